// Single-precision instantiation of the typed operations.
#include "ops_impl.hpp"
template struct sc::Ops<float>;
