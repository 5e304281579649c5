// Typed operations of the library (instantiated for float and double in
// lib_f32.hip / lib_f64.hip).  All launch on ctx.stream; pointers are device
// pointers in the device layout.
#pragma once
#include "ctx.hpp"

namespace sc {

template <typename T>
struct Ops {
  using CT = C<T>;
  static bool supported(int n);
  static void init_tables(Ctx& c);
  // 1D passes and 3D transforms
  static void c2r(Ctx& c, T* phys, const CT* spec, int ncomp, double scale, bool mask23);
  static void r2c(Ctx& c, CT* spec, const T* phys, int ncomp, double scale);
  // spectral kernels
  static void twothirds(Ctx& c, CT* u, int ncomp);
  static void project(Ctx& c, CT* u);
  static void viscosity(Ctx& c, CT* sig, const CT* u, double nu);
  static void tensordivergence(Ctx& c, CT* du, const CT* sig);
  static void vectorgradient(Ctx& c, CT* G, const CT* u);
  static void strainrate(Ctx& c, CT* S, const CT* u);
  static void vectordivergence(Ctx& c, CT* d, const CT* u);
  static void gaussian(Ctx& c, CT* u, int ncomp, double delta);
  static void axpy(Ctx& c, CT* y, const CT* x, double a, int ncomp);
  static void spec_tracefree(Ctx& c, CT* tau);
  // the RHS
  static void nonlinearity(Ctx& c, CT* sig, const CT* u, unsigned long long* vmax);
  static void rhs(Ctx& c, CT* du, const CT* u, double nu);
  static void wray3(Ctx& c, CT* u, double dt, double nu);
  static void rk4(Ctx& c, CT* u, double dt, double nu);
  static double max_velocity(Ctx& c, const CT* u);
  static void max_velocity_into(Ctx& c, const CT* u);
  static void rhs_fused(Ctx& c, CT* du, const CT* u, double nu, unsigned long long* vmax, bool lin = true);
  static double wray3_dev(Ctx& c, CT* u, double dt, double nu, double cfl, int les_closure, double p1, double p2);
  static double rk4_dev(Ctx& c, CT* u, double dt, double nu, double cfl, int les_closure = -1, double p1 = 0,
                        double p2 = 0);
  static double read_dt(Ctx& c);
  static void rk4_steps(Ctx& c, CT* u, int nsteps, double cfl, double nu, double dtmax, double* dts);
  // adjoint
  static void lincomb(Ctx& c, CT* dst, const CT* x, double a, const CT* y, double b, int ncomp);
  // closure < 0: the DNS right-hand side; 0 or 4: les_rhs with that closure
  static void rhs_vjp(Ctx& c, CT* out, const CT* u, const CT* w, double nu, int closure = -1);
  static void rhs_vjp_kept(Ctx& c, CT* out, const CT* u, const CT* w, double nu, int closure, int kept);
  static void rk4_vjp(Ctx& c, CT* out, const CT* u0, const CT* ubar, double dt, double nu, int closure = -1);
  static void wray3_vjp(Ctx& c, CT* out, const CT* u0, const CT* ubar, double dt, double nu, int closure = -1);
  static void stage_rhs(Ctx& c, CT* du, const CT* u, double nu, int closure);
  // reductions (return to host)
  static double sum_modes(Ctx& c, const CT* u, int ncomp, int kind, double nu);
  static void sum_modes_noalloc(Ctx& c, const CT* u, int ncomp, int kind, double nu, double* scr, double* out);
  static void dot_modes(Ctx& c, const CT* u, const CT* v, int ncomp, double* re, double* im);
  static void shell_energy(Ctx& c, const CT* u, int ncomp, int nshell, double* host_out);
  static void shell_scale(Ctx& c, CT* u, int ncomp, int nshell, const double* host_factor);
  // initial conditions
  static void randn(Ctx& c, CT* u, int ncomp, unsigned long long seed);
  static void symmetrize(Ctx& c, CT* u, int ncomp);
  static void taylorgreen(Ctx& c, CT* u, double V0, int doproject);
  // filtering between two contexts (coarse <- fine)
  static void cutoff(Ctx& cb, Ctx& cf, CT* ub, const CT* u, int ncomp);
  static void inverse_cutoff(Ctx& cb, Ctx& cf, CT* u, const CT* ub, int ncomp);
  // closures (physical packed tensors)
  static void clark(Ctx& c, T* tau, const T* GG, double delta);
  static void smagorinsky(Ctx& c, T* tau, const T* GG, double cs, double delta);
  static void verstappen(Ctx& c, T* tau, const T* GG, double cq, double delta);
  static void shell_transfer(Ctx& c, const CT* u, const T* tau, int nshell, double* host_out);
  static void sfs_dissipation(Ctx& c, T* eps, const T* tau, const T* GG);
  static void sample_moments(Ctx& c, const T* x, long n, double* out);
  static double sample_mean(Ctx& c, const T* x, long n);
  static void apriori(Ctx& c, const T* model, const T* ref, long n, double* out);
  static void bardina(Ctx& c, T* tau, const CT* u, double s, double delta);
  static void gradient_phys(Ctx& c, T* GG, const CT* u);
  static double dynsmag(Ctx& c, T* tau, const CT* u, double delta);
  static void les_rhs(Ctx& c, CT* du, const CT* u, double nu, int closure, double p1, double p2);
  static void cnn_tau(Ctx& c, T* tau, const CT* u);
  static const double* band_coeff(Ctx& c, const CT* u);
  static void phys_tracefree(Ctx& c, T* tau);
  static void les_wray3(Ctx& c, CT* u, double dt, double nu, int closure, double p1, double p2);
  static void launch_phys(Ctx& c, T* y, const T* x, T a, T b, long n);
  static void smag_tensor(Ctx& c, T* m, const T* S, double delta);
};

}  // namespace sc
