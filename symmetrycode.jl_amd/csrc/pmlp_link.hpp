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
int arch(const sc_pmlp* c);  // 0 mlp, 1 TBNN
// tau (packed physical, component-major) from the physical velocity gradient
// GG (D² components, nonsym order): fullchain's Δ² |∇ū|² mlp(∇ū / (|∇ū| +
// eps), z) (averaged over the cube's symmetry group when sym != 0), or the
// TBNN's Δ² |∇ū|² Σ_b w_b(invariants, z) B_b; with a Re_Δ input z = (log(Δ²
// sqrt(mean |∇ū|²) / visc) - mu) / sigma of this field
template <typename T>
void closure_tau(sc_pmlp* c, const T* GG, long np, double delta, double visc, double mu, double sigma, int sym, T* tau,
                 hipStream_t s);
}  // namespace pmlplink
