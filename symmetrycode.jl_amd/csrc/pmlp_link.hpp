// Internal interface between the solver and the pointwise MLP closure
// (csrc/pmlp.hip): the solver attaches a network as closure kind 7 and calls
// closure_tau on its own stream.
#pragma once
#include <hip/hip_runtime.h>

struct sc_pmlp;

namespace pmlplink {
int grad_channels(const sc_pmlp* c);
int dim(const sc_pmlp* c);
bool uses_redelta(const sc_pmlp* c);
// tau (packed physical, component-major) = Δ² |∇ū|² mlp(∇ū / (|∇ū| + eps), z)
// from the physical velocity gradient GG (D² components, nonsym order); with a
// Re_Δ input z = (log(Δ² sqrt(mean |∇ū|²) / visc) - mu) / sigma of this field
template <typename T>
void closure_tau(sc_pmlp* c, const T* GG, long np, double delta, double visc, double mu, double sigma, T* tau,
                 hipStream_t s);
}  // namespace pmlplink
