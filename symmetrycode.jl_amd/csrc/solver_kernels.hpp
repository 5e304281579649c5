// Pointwise spectral/physical kernels and FP64 reductions.
// Each spectral kernel runs one thread per entry of the padded device array
// (id = x + nhp*(y + n*z)) and skips the pad columns x >= nh.
#pragma once
#include "common.hpp"

namespace sc {

struct SpecGeom {
  int D, n, nh, nhp, nz;
  long s_comp;   // spectral component stride (elements)
  long total;    // nhp * n * nz
  double kf;     // (pi / l) * 2, the reference's wavenumber factor
};

struct Mode {
  int x, y, z;     // indices
  int kx, ky, kz;  // integer wavenumbers
  long off;
};

DEV bool mode_of(long id, const SpecGeom& g, Mode& m) {
  if (id >= g.total) return false;
  m.x = int(id % g.nhp);
  if (m.x >= g.nh) return false;
  long r = id / g.nhp;
  m.y = int(r % g.n);
  m.z = int(r / g.n);
  m.kx = m.x;
  m.ky = freq(m.y, g.n);
  m.kz = g.D == 3 ? freq(m.z, g.n) : 0;
  m.off = id;
  return true;
}

DEV long gid() { return long(blockIdx.x) * blockDim.x + threadIdx.x; }

template <typename T> DEV C<T> czero() { return mk<T>(T(0), T(0)); }

// i * k * a for real k
template <typename T> DEV C<T> ik(T k, C<T> a) { return mk<T>(-k * a.y, k * a.x); }

DEV bool in23(const Mode& m, int kcut) {
  bool nonzero = (m.kx | m.ky | m.kz) != 0;
  return nonzero && abs(m.kx) <= kcut && abs(m.ky) <= kcut && abs(m.kz) <= kcut;
}

// twothirds! (src/solver.jl:371-389)
template <typename T>
__global__ void k_twothirds(C<T>* u, SpecGeom g, int ncomp) {
  Mode m;
  if (!mode_of(gid(), g, m)) return;
  if (in23(m, g.n / 3)) return;
  for (int c = 0; c < ncomp; ++c) u[c * g.s_comp + m.off] = czero<T>();
}

// project! (src/solver.jl:350-368); the k-division is done in FP64.
template <typename T, int D>
DEV void project_mode(const Mode& m, const SpecGeom& g, C<T>* uv) {
  double k[3] = {g.kf * m.kx, g.kf * m.ky, g.kf * m.kz};
  double pr = 0, pi = 0, k2 = 0;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    pr += k[c] * double(uv[c].x);
    pi += k[c] * double(uv[c].y);
    k2 += k[c] * k[c];
  }
  if (k2 == 0.0) return;  // constant mode left intact
  pr /= k2;
  pi /= k2;
#pragma unroll
  for (int c = 0; c < D; ++c) uv[c] = mk<T>(T(double(uv[c].x) - k[c] * pr), T(double(uv[c].y) - k[c] * pi));
}

template <typename T, int D>
__global__ void k_project(C<T>* u, SpecGeom g) {
  Mode m;
  if (!mode_of(gid(), g, m)) return;
  C<T> uv[D];
#pragma unroll
  for (int c = 0; c < D; ++c) uv[c] = u[c * g.s_comp + m.off];
  project_mode<T, D>(m, g, uv);
#pragma unroll
  for (int c = 0; c < D; ++c) u[c * g.s_comp + m.off] = uv[c];
}

template <typename T> DEV void kvec(const Mode& m, const SpecGeom& g, T* k) {
  k[0] = T(g.kf) * T(m.kx);
  k[1] = T(g.kf) * T(m.ky);
  k[2] = T(g.kf) * T(m.kz);
}

// viscosity! (src/solver.jl:458-476)
template <typename T, int D>
DEV void viscosity_mode(const T* k, const C<T>* u, T nu, C<T>* s) {
  if constexpr (D == 2) {
    s[0] = sub<T>(s[0], ik<T>(2 * nu * k[0], u[0]));
    s[1] = sub<T>(s[1], ik<T>(2 * nu * k[1], u[1]));
    s[2] = sub<T>(s[2], scl<T>(ik<T>(T(1), add<T>(scl<T>(u[0], k[1]), scl<T>(u[1], k[0]))), nu));
  } else {
    s[0] = sub<T>(s[0], ik<T>(2 * nu * k[0], u[0]));
    s[1] = sub<T>(s[1], ik<T>(2 * nu * k[1], u[1]));
    s[2] = sub<T>(s[2], ik<T>(2 * nu * k[2], u[2]));
    s[3] = sub<T>(s[3], scl<T>(ik<T>(T(1), add<T>(scl<T>(u[0], k[1]), scl<T>(u[1], k[0]))), nu));
    s[4] = sub<T>(s[4], scl<T>(ik<T>(T(1), add<T>(scl<T>(u[1], k[2]), scl<T>(u[2], k[1]))), nu));
    s[5] = sub<T>(s[5], scl<T>(ik<T>(T(1), add<T>(scl<T>(u[2], k[0]), scl<T>(u[0], k[2]))), nu));
  }
}

// tensordivergence! (src/solver.jl:509-521): du_i = -i k_j sigma_ij
template <typename T, int D>
DEV void divergence_mode(const T* k, const C<T>* s, C<T>* du) {
  if constexpr (D == 2) {
    du[0] = mk<T>(T(0), T(0));
    du[0] = sub<T>(sub<T>(du[0], ik<T>(k[0], s[0])), ik<T>(k[1], s[2]));
    du[1] = sub<T>(sub<T>(czero<T>(), ik<T>(k[0], s[2])), ik<T>(k[1], s[1]));
  } else {
    du[0] = sub<T>(sub<T>(sub<T>(czero<T>(), ik<T>(k[0], s[0])), ik<T>(k[1], s[3])), ik<T>(k[2], s[5]));
    du[1] = sub<T>(sub<T>(sub<T>(czero<T>(), ik<T>(k[0], s[3])), ik<T>(k[1], s[1])), ik<T>(k[2], s[4]));
    du[2] = sub<T>(sub<T>(sub<T>(czero<T>(), ik<T>(k[0], s[5])), ik<T>(k[1], s[4])), ik<T>(k[2], s[2]));
  }
}

template <typename T, int D>
__global__ void k_viscosity(C<T>* sig, const C<T>* u, double nu, SpecGeom g) {
  constexpr int NT = D == 3 ? 6 : 3;
  Mode m;
  if (!mode_of(gid(), g, m)) return;
  T k[3];
  kvec<T>(m, g, k);
  C<T> uv[D], s[NT];
#pragma unroll
  for (int c = 0; c < D; ++c) uv[c] = u[c * g.s_comp + m.off];
#pragma unroll
  for (int c = 0; c < NT; ++c) s[c] = sig[c * g.s_comp + m.off];
  viscosity_mode<T, D>(k, uv, T(nu), s);
#pragma unroll
  for (int c = 0; c < NT; ++c) sig[c * g.s_comp + m.off] = s[c];
}

template <typename T, int D>
__global__ void k_tensordivergence(C<T>* du, const C<T>* sig, SpecGeom g) {
  constexpr int NT = D == 3 ? 6 : 3;
  Mode m;
  if (!mode_of(gid(), g, m)) return;
  T k[3];
  kvec<T>(m, g, k);
  C<T> s[NT], d[D];
#pragma unroll
  for (int c = 0; c < NT; ++c) s[c] = sig[c * g.s_comp + m.off];
  divergence_mode<T, D>(k, s, d);
#pragma unroll
  for (int c = 0; c < D; ++c) du[c * g.s_comp + m.off] = d[c];
}

// Fused viscosity! + tensordivergence! (+ optional linear band forcing
// du += A u on 0 < |k|^2 <= kband2, A read from device memory).
template <typename T, int D>
__global__ void k_div_visc(C<T>* du, const C<T>* sig, const C<T>* u, double nu, SpecGeom g,
                           const double* forceA, int kband2) {
  constexpr int NT = D == 3 ? 6 : 3;
  Mode m;
  if (!mode_of(gid(), g, m)) return;
  T k[3];
  kvec<T>(m, g, k);
  C<T> uv[D], s[NT], d[D];
#pragma unroll
  for (int c = 0; c < D; ++c) uv[c] = u[c * g.s_comp + m.off];
#pragma unroll
  for (int c = 0; c < NT; ++c) s[c] = sig[c * g.s_comp + m.off];
  viscosity_mode<T, D>(k, uv, T(nu), s);
  divergence_mode<T, D>(k, s, d);
  if (forceA != nullptr) {
    int k2 = m.kx * m.kx + m.ky * m.ky + m.kz * m.kz;
    if (k2 > 0 && k2 <= kband2) {
      T A = T(*forceA);
#pragma unroll
      for (int c = 0; c < D; ++c) d[c] = add<T>(d[c], scl<T>(uv[c], A));
    }
  }
#pragma unroll
  for (int c = 0; c < D; ++c) du[c * g.s_comp + m.off] = d[c];
}

// vectorgradient! (src/solver.jl:711-733); output order is the reference's
// tensorfield_nonsym order xx, yx, zx, xy, yy, zy, xz, yz, zz where
// G.ij = i k_j u_i (2D: xx, yx, xy, yy).
template <typename T, int D>
__global__ void k_vectorgradient(C<T>* G, const C<T>* u, SpecGeom g) {
  Mode m;
  if (!mode_of(gid(), g, m)) return;
  T k[3];
  kvec<T>(m, g, k);
  C<T> uv[D];
#pragma unroll
  for (int c = 0; c < D; ++c) uv[c] = u[c * g.s_comp + m.off];
  // component index of G.ij in the nonsym order is i + D*j
#pragma unroll
  for (int j = 0; j < D; ++j)
#pragma unroll
    for (int i = 0; i < D; ++i) G[(i + D * j) * g.s_comp + m.off] = ik<T>(k[j], uv[i]);
}

// strainrate! (src/solver.jl:776-794)
template <typename T, int D>
__global__ void k_strainrate(C<T>* S, const C<T>* u, SpecGeom g) {
  Mode m;
  if (!mode_of(gid(), g, m)) return;
  T k[3];
  kvec<T>(m, g, k);
  C<T> uv[D];
#pragma unroll
  for (int c = 0; c < D; ++c) uv[c] = u[c * g.s_comp + m.off];
  const long o = m.off, s = g.s_comp;
  if constexpr (D == 2) {
    S[0 * s + o] = ik<T>(k[0], uv[0]);
    S[1 * s + o] = ik<T>(k[1], uv[1]);
    S[2 * s + o] = scl<T>(ik<T>(T(1), add<T>(scl<T>(uv[0], k[1]), scl<T>(uv[1], k[0]))), T(0.5));
  } else {
    S[0 * s + o] = ik<T>(k[0], uv[0]);
    S[1 * s + o] = ik<T>(k[1], uv[1]);
    S[2 * s + o] = ik<T>(k[2], uv[2]);
    S[3 * s + o] = scl<T>(ik<T>(T(1), add<T>(scl<T>(uv[0], k[1]), scl<T>(uv[1], k[0]))), T(0.5));
    S[4 * s + o] = scl<T>(ik<T>(T(1), add<T>(scl<T>(uv[1], k[2]), scl<T>(uv[2], k[1]))), T(0.5));
    S[5 * s + o] = scl<T>(ik<T>(T(1), add<T>(scl<T>(uv[2], k[0]), scl<T>(uv[0], k[2]))), T(0.5));
  }
}

// vectordivergence! (src/solver.jl:484-493)
template <typename T, int D>
__global__ void k_vectordivergence(C<T>* dv, const C<T>* u, SpecGeom g) {
  Mode m;
  if (!mode_of(gid(), g, m)) return;
  T k[3];
  kvec<T>(m, g, k);
  C<T> r = czero<T>();
#pragma unroll
  for (int c = 0; c < D; ++c) r = add<T>(r, ik<T>(k[c], u[c * g.s_comp + m.off]));
  dv[m.off] = r;
}

// gaussianfilter! (src/filtering.jl:44-67): exp(-Delta^2 k^2 / 24) except the
// forced shells |k_int|^2 < 9.
template <typename T>
__global__ void k_gaussian(C<T>* u, double delta, SpecGeom g, int ncomp) {
  Mode m;
  if (!mode_of(gid(), g, m)) return;
  double kx = g.kf * m.kx, ky = g.kf * m.ky, kz = g.kf * m.kz;
  double k2 = kx * kx + ky * ky + kz * kz;
  int k2i = m.kx * m.kx + m.ky * m.ky + m.kz * m.kz;
  T w = k2i < 9 ? T(1) : T(exp(-delta * delta * k2 / 24));
  for (int c = 0; c < ncomp; ++c) u[c * g.s_comp + m.off] = scl<T>(u[c * g.s_comp + m.off], w);
}

// cutoff! (src/filtering.jl:28-34): coarse mode (x, y, z) <- fine mode
// (x, y', z') with negative wavenumbers counted from the end.
template <typename T>
__global__ void k_cutoff(C<T>* ub, const C<T>* u, SpecGeom gb, SpecGeom gf, int ncomp) {
  Mode m;
  if (!mode_of(gid(), gb, m)) return;
  int half = gb.n / 2;
  int fy = m.y >= half ? m.y + gf.n - gb.n : m.y;
  int fz = (gb.D == 3 && m.z >= half) ? m.z + gf.n - gb.n : m.z;
  long fo = m.x + long(gf.nhp) * (fy + long(gf.n) * fz);
  for (int c = 0; c < ncomp; ++c) ub[c * gb.s_comp + m.off] = u[c * gf.s_comp + fo];
}

// inverse_cutoff! (src/filtering.jl:36-42): scatter coarse into fine
template <typename T>
__global__ void k_inverse_cutoff(C<T>* u, const C<T>* ub, SpecGeom gb, SpecGeom gf, int ncomp) {
  Mode m;
  if (!mode_of(gid(), gb, m)) return;
  int half = gb.n / 2;
  int fy = m.y >= half ? m.y + gf.n - gb.n : m.y;
  int fz = (gb.D == 3 && m.z >= half) ? m.z + gf.n - gb.n : m.z;
  long fo = m.x + long(gf.nhp) * (fy + long(gf.n) * fz);
  for (int c = 0; c < ncomp; ++c) u[c * gf.s_comp + fo] = ub[c * gb.s_comp + m.off];
}

// ------------------------------------------------------------- RK updates
// wray3! stage update (src/solver.jl:997-1008):
//   u = P(src + a dt du);  if (write_ustart) ustart = src + b dt du
// src is u itself on the first stage (the reference copies u to ustart).
// dt is read from device memory (set by the host or by the fused CFL step);
// the coefficients are a[i] and b[i] rounded to T, as in the reference.
template <typename T, int D>
__global__ void k_wray3_update(C<T>* u, C<T>* ustart, const C<T>* src, const C<T>* du, const double* dtp, T a,
                               T b, int write_ustart, SpecGeom g) {
  Mode m;
  if (!mode_of(gid(), g, m)) return;
  const double dt = *dtp;
  const T adt = T(double(a) * dt), bdt = T(double(b) * dt);
  C<T> s[D], d[D], v[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    s[c] = src[c * g.s_comp + m.off];
    d[c] = du[c * g.s_comp + m.off];
    v[c] = add<T>(s[c], scl<T>(d[c], adt));
  }
  project_mode<T, D>(m, g, v);
#pragma unroll
  for (int c = 0; c < D; ++c) u[c * g.s_comp + m.off] = v[c];
  if (write_ustart) {
#pragma unroll
    for (int c = 0; c < D; ++c) ustart[c * g.s_comp + m.off] = add<T>(s[c], scl<T>(d[c], bdt));
  }
}

// RK4 stage update.  stage 0: acc = k;  1,2: acc += 2k;  3: acc += k.
// stages 0-2 write ust = P(u0 + frac dt k); stage 3 writes u0 = P(u0 + dt/6 acc).
template <typename T, int D>
__global__ void k_rk4_update(C<T>* u0, C<T>* ust, C<T>* acc, const C<T>* k, const double* dtp, double frac, int stage,
                             SpecGeom g) {
  Mode m;
  if (!mode_of(gid(), g, m)) return;
  const T cdt = T(*dtp * frac);
  C<T> v[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const long o = c * g.s_comp + m.off;
    C<T> kk = k[o];
    C<T> a;
    if (stage == 0) a = kk;
    else if (stage == 3) a = add<T>(acc[o], kk);
    else a = add<T>(acc[o], scl<T>(kk, T(2)));
    if (stage < 3) acc[o] = a;
    C<T> b = u0[o];
    v[c] = add<T>(b, scl<T>(stage < 3 ? kk : a, cdt));
  }
  project_mode<T, D>(m, g, v);
  C<T>* dst = stage < 3 ? ust : u0;
#pragma unroll
  for (int c = 0; c < D; ++c) dst[c * g.s_comp + m.off] = v[c];
}

// RK4 stage update whose k arrives without its linear terms (the divergence
// of k_z_fwd_rhs<.., false>): k = k_div - nu |k|^2 us (+ A_f us on the band
// 0 < |k|^2 <= kband2, fA non-null), us the stage input (u0 at stage 0, ust
// after; read before ust is written, so the update runs in place).  Then as
// k_rk4_update.
template <typename T, int D>
__global__ void k_rk4_update_lin(C<T>* u0, C<T>* ust, C<T>* acc, const C<T>* k, const double* dtp, double frac,
                                 int stage, SpecGeom g, double nu, const double* fA, int kband2) {
  Mode m;
  if (!mode_of(gid(), g, m)) return;
  const T cdt = T(*dtp * frac);
  const T kf = T(g.kf);
  const T kx = kf * T(m.kx), ky = kf * T(m.ky), kz = kf * T(m.kz);
  const T nk2 = T(nu) * (kx * kx + ky * ky + kz * kz);
  const int k2i = m.kx * m.kx + m.ky * m.ky + m.kz * m.kz;
  const T fa = (fA && k2i > 0 && k2i <= kband2) ? T(*fA) : T(0);
  C<T> v[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const long o = c * g.s_comp + m.off;
    const C<T> b = u0[o];
    C<T> uu = b;
    if (stage > 0) uu = ust[o];
    C<T> kk = k[o];
    kk = mk<T>(kk.x - nk2 * uu.x, kk.y - nk2 * uu.y);
    if (fa != T(0)) kk = add<T>(kk, scl<T>(uu, fa));
    C<T> a;
    if (stage == 0) a = kk;
    else if (stage == 3) a = add<T>(acc[o], kk);
    else a = add<T>(acc[o], scl<T>(kk, T(2)));
    if (stage < 3) acc[o] = a;
    v[c] = add<T>(b, scl<T>(stage < 3 ? kk : a, cdt));
  }
  project_mode<T, D>(m, g, v);
  C<T>* dst = stage < 3 ? ust : u0;
#pragma unroll
  for (int c = 0; c < D; ++c) dst[c * g.s_comp + m.off] = v[c];
}

// y = a*x + y over ncomp spectral components
template <typename T>
__global__ void k_axpy(C<T>* y, const C<T>* x, T a, SpecGeom g, int ncomp) {
  Mode m;
  if (!mode_of(gid(), g, m)) return;
  for (int c = 0; c < ncomp; ++c) {
    long o = c * g.s_comp + m.off;
    y[o] = add<T>(y[o], scl<T>(x[o], a));
  }
}

// dst = a*x + b*y over ncomp spectral components (dst may alias x or y)
template <typename T>
__global__ void k_lincomb(C<T>* dst, const C<T>* x, T a, const C<T>* y, T b, SpecGeom g, int ncomp) {
  Mode m;
  if (!mode_of(gid(), g, m)) return;
  for (int c = 0; c < ncomp; ++c) {
    long o = c * g.s_comp + m.off;
    dst[o] = add<T>(scl<T>(x[o], a), scl<T>(y[o], b));
  }
}

// --------------------------------------------------------------- adjoint
// Physical part of the RHS transpose: a_m = sum_j g_mj v_j + sum_i g_im v_i
// with g the nonsym physical gradient of the cotangent (index i + D*j holds
// d w_i / d x_j) and v the truncated velocity.  a overwrites v.
template <typename T, int D>
__global__ void k_adj_products(T* v, const T* gr, long np) {
  long id = gid();
  if (id >= np) return;
  T vv[D], a[D];
#pragma unroll
  for (int c = 0; c < D; ++c) vv[c] = v[c * np + id];
#pragma unroll
  for (int m = 0; m < D; ++m) {
    T s = 0;
#pragma unroll
    for (int j = 0; j < D; ++j) s += gr[(m + D * j) * np + id] * vv[j];
#pragma unroll
    for (int i = 0; i < D; ++i) s += gr[(i + D * m) * np + id] * vv[i];
    a[m] = s;
  }
#pragma unroll
  for (int c = 0; c < D; ++c) v[c * np + id] = a[c];
}

// out += -nu (|k|^2 w + k (k.w)) + A B w - coefB B u, the viscous part of
// the transpose (self-adjoint) and the band-forcing part; bandp = {A, coefB}
// or null when no forcing.
template <typename T, int D>
__global__ void k_adj_visc_force(C<T>* out, const C<T>* w, const C<T>* u, double nu, SpecGeom g,
                                 const double* bandp, int kband2) {
  Mode m;
  if (!mode_of(gid(), g, m)) return;
  double k[3] = {g.kf * m.kx, g.kf * m.ky, g.kf * m.kz};
  double k2 = 0, kwr = 0, kwi = 0;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    C<T> ww = w[c * g.s_comp + m.off];
    k2 += k[c] * k[c];
    kwr += k[c] * double(ww.x);
    kwi += k[c] * double(ww.y);
  }
  const int k2i = m.kx * m.kx + m.ky * m.ky + m.kz * m.kz;
  const bool inband = bandp != nullptr && k2i > 0 && k2i <= kband2;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const long o = c * g.s_comp + m.off;
    C<T> ww = w[o];
    double rx = -nu * (k2 * double(ww.x) + k[c] * kwr);
    double ry = -nu * (k2 * double(ww.y) + k[c] * kwi);
    if (inband) {
      C<T> uu = u[o];
      rx += bandp[0] * double(ww.x) - bandp[1] * double(uu.x);
      ry += bandp[0] * double(ww.y) - bandp[1] * double(uu.y);
    }
    C<T> r = out[o];
    out[o] = mk<T>(T(double(r.x) + rx), T(double(r.y) + ry));
  }
}

// Band sums for the forcing transpose: S = sum_band w|u|^2, c = sum_band w Re(conj(u) w)
// -> bandp = {A, 2 A c / S} with A = injection / S
template <typename T>
__global__ void k_band_adj(const C<T>* u, const C<T>* wv, int ncomp, SpecGeom g, int kband, double injection,
                           double* bandp) {
  __shared__ double s1[256], s2[256];
  double a = 0.0, b = 0.0;
  const int kb2 = kband * kband;
  const int w1 = 2 * kband + 1;
  const int nz = g.D == 3 ? w1 : 1;
  const int total = (kband + 1) * w1 * nz;
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    int kx = i % (kband + 1);
    int ky = (i / (kband + 1)) % w1 - kband;
    int kz = g.D == 3 ? (i / ((kband + 1) * w1)) - kband : 0;
    int k2 = kx * kx + ky * ky + kz * kz;
    if (k2 == 0 || k2 > kb2) continue;
    int y = ky < 0 ? ky + g.n : ky;
    int z = kz < 0 ? kz + g.n : kz;
    long off = kx + long(g.nhp) * (y + long(g.n) * z);
    double wt = (kx >= 1 && kx <= g.nh - 2) ? 2.0 : 1.0;
    for (int c = 0; c < ncomp; ++c) {
      C<T> uu = u[c * g.s_comp + off], ww = wv[c * g.s_comp + off];
      a += wt * (double(uu.x) * uu.x + double(uu.y) * uu.y);
      b += wt * (double(uu.x) * ww.x + double(uu.y) * ww.y);
    }
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      s1[threadIdx.x] += s1[threadIdx.x + s];
      s2[threadIdx.x] += s2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double A = s1[0] > 0 ? injection / s1[0] : 0.0;
    bandp[0] = A;
    bandp[1] = s1[0] > 0 ? 2 * A * s2[0] / s1[0] : 0.0;
  }
}

// ------------------------------------------------------------- reductions
// Block-partial FP64 sums of an RFFT-weighted mode function; the final sum
// over blocks is done by k_finish_sum (fixed order: deterministic).
constexpr int RB = 256;

// Scratch of the no-allocation reduction (sc_energy_noalloc), in doubles:
// the partials of one chunk of at most NOALLOC_PARTIALS blocks, the sums of
// at most NOALLOC_CHUNKS chunks, and the total.  The largest field it takes
// has NOALLOC_PARTIALS * NOALLOC_CHUNKS * RB entries per component (1.7e10).
constexpr long NOALLOC_PARTIALS = 65536;
constexpr long NOALLOC_CHUNKS = 1024;
constexpr long NOALLOC_DOUBLES = NOALLOC_PARTIALS + NOALLOC_CHUNKS + 1;

template <class F>
DEV void block_sum_store(double v, double* partial) {
  __shared__ double sm[RB];
  sm[threadIdx.x] = v;
  __syncthreads();
  for (int s = RB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sm[0];
}

DEV double rfft_w(const Mode& m, const SpecGeom& g) { return (m.x >= 1 && m.x <= g.nh - 2) ? 2.0 : 1.0; }

// kind 0: sum_c w |u_c|^2          (getenergy / energy)
// kind 1: nu * sum_c w k^2 |u_c|^2 (get_dissipation!)
// kind 2: nu * sum_c w k^4 |u_c|^2 (get_palinstrophy!)
// kind 3: sum_c |u_c|^2 over the stored half spectrum, unweighted (LinearAlgebra.norm^2)
// (base: index of the block's first mode, for launches over part of the field)
template <typename T>
__global__ void k_sum_modes(const C<T>* u, int ncomp, SpecGeom g, int kind, double nu, double* partial, long base) {
  Mode m;
  double acc = 0.0;
  if (mode_of(base + gid(), g, m)) {
    double a = 0;
    for (int c = 0; c < ncomp; ++c) {
      C<T> v = u[c * g.s_comp + m.off];
      a += double(v.x) * double(v.x) + double(v.y) * double(v.y);
    }
    double w = kind == 3 ? 1.0 : rfft_w(m, g);
    if (kind == 4 && !in23(m, g.n / 3)) a = 0;  // filter_reynolds: gradient inside the 2/3 box
    if (kind == 1 || kind == 2 || kind == 4) {
      double kx = g.kf * m.kx, ky = g.kf * m.ky, kz = g.kf * m.kz;
      double k2 = kx * kx + ky * ky + kz * kz;
      a *= nu * (kind == 2 ? k2 * k2 : k2);
    }
    acc = w * a;
  }
  block_sum_store<int>(acc, partial);
}

// RFFT-weighted sum_c conj(u_c) v_c (spectraldot, src/solver.jl:708): (re, im)
template <typename T>
__global__ void k_dot_modes(const C<T>* u, const C<T>* v, int ncomp, SpecGeom g, double* partial) {
  Mode m;
  double re = 0.0, im = 0.0;
  if (mode_of(gid(), g, m)) {
    double w = rfft_w(m, g);
    for (int c = 0; c < ncomp; ++c) {
      C<T> a = u[c * g.s_comp + m.off], b = v[c * g.s_comp + m.off];
      re += w * (double(a.x) * b.x + double(a.y) * b.y);
      im += w * (double(a.x) * b.y - double(a.y) * b.x);
    }
  }
  block_sum_store<int>(re, partial);
  __syncthreads();
  block_sum_store<int>(im, partial + gridDim.x);
}

static __global__ void k_finish_sum(const double* partial, long n, double* out) {
  __shared__ double sm[1024];
  double a = 0.0;
  for (long i = threadIdx.x; i < n; i += blockDim.x) a += partial[i];
  sm[threadIdx.x] = a;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sm[0];
}

// integer shell index floor(sqrt(k2))
DEV int isqrt_shell(int k2) {
  int s = int(sqrt(double(k2)));
  while ((s + 1) * (s + 1) <= k2) ++s;
  while (s * s > k2) --s;
  return s;
}

// Per-shell RFFT-weighted energy 1/2 sum_c |u_c|^2, shells 0..nshell-1
// (shell i: i^2 <= k2int < (i+1)^2).  FP64 atomics into `out` (zeroed).
template <typename T>
__global__ void k_shell_energy(const C<T>* u, int ncomp, SpecGeom g, int nshell, double* out) {
  extern __shared__ double hist[];
  for (int i = threadIdx.x; i < nshell; i += blockDim.x) hist[i] = 0.0;
  __syncthreads();
  Mode m;
  if (mode_of(gid(), g, m)) {
    int k2 = m.kx * m.kx + m.ky * m.ky + m.kz * m.kz;
    int s = isqrt_shell(k2);
    if (s < nshell) {
      double a = 0;
      for (int c = 0; c < ncomp; ++c) {
        C<T> v = u[c * g.s_comp + m.off];
        a += double(v.x) * double(v.x) + double(v.y) * double(v.y);
      }
      atomicAdd(&hist[s], 0.5 * rfft_w(m, g) * a);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nshell; i += blockDim.x)
    if (hist[i] != 0.0) atomicAdd(&out[i], hist[i]);
}

// Shell-resolved SFS transfer of solve_les's reduce! (src/les.jl:125-139):
// per mode clo_i = -i k_j tau_ij (tensordivergence!) of the packed spectral
// stress th, T = w Re(sum_i conj(u_i) clo_i) with the RFFT weight w,
// histogrammed into shells 0..nshell-1 (out zeroed, FP64 atomics)
template <typename T, int D>
__global__ void k_shell_transfer(const C<T>* u, const C<T>* th, SpecGeom g, int nshell, double* out) {
  constexpr int NT = D == 3 ? 6 : 3;
  extern __shared__ double hist[];
  for (int i = threadIdx.x; i < nshell; i += blockDim.x) hist[i] = 0.0;
  __syncthreads();
  Mode m;
  if (mode_of(gid(), g, m)) {
    int k2 = m.kx * m.kx + m.ky * m.ky + m.kz * m.kz;
    int sh = isqrt_shell(k2);
    if (sh < nshell) {
      T k[3];
      kvec<T>(m, g, k);
      C<T> s[NT], clo[D];
#pragma unroll
      for (int c = 0; c < NT; ++c) s[c] = th[c * g.s_comp + m.off];
      divergence_mode<T, D>(k, s, clo);
      double a = 0;
#pragma unroll
      for (int i = 0; i < D; ++i) {
        const C<T> v = u[i * g.s_comp + m.off];
        a += double(v.x) * double(clo[i].x) + double(v.y) * double(clo[i].y);
      }
      atomicAdd(&hist[sh], rfft_w(m, g) * a);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nshell; i += blockDim.x)
    if (hist[i] != 0.0) atomicAdd(&out[i], hist[i]);
}

// u *= factor[shell] for shells < nshell (others untouched)
template <typename T>
__global__ void k_shell_scale(C<T>* u, int ncomp, SpecGeom g, int nshell, const double* factor) {
  Mode m;
  if (!mode_of(gid(), g, m)) return;
  int k2 = m.kx * m.kx + m.ky * m.ky + m.kz * m.kz;
  int s = isqrt_shell(k2);
  if (s >= nshell) return;
  T f = T(factor[s]);
  for (int c = 0; c < ncomp; ++c) u[c * g.s_comp + m.off] = scl<T>(u[c * g.s_comp + m.off], f);
}

// Linear band forcing coefficient A = injection / sum_{0<|k|^2<=kb2} w |u|^2,
// one block over the (small) band.
template <typename T>
__global__ void k_band_coeff(const C<T>* u, int ncomp, SpecGeom g, int kband, double injection, double* A,
                             int tx = 0, long lp = 0) {
  __shared__ double sm[256];
  double acc = 0.0;
  const int kb2 = kband * kband;
  const int w1 = 2 * kband + 1;
  const int nz = g.D == 3 ? w1 : 1;
  const int total = (kband + 1) * w1 * nz;
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    int kx = i % (kband + 1);
    int ky = (i / (kband + 1)) % w1 - kband;
    int kz = g.D == 3 ? (i / ((kband + 1) * w1)) - kband : 0;
    int k2 = kx * kx + ky * ky + kz * kz;
    if (k2 == 0 || k2 > kb2) continue;
    int y = ky < 0 ? ky + g.n : ky;
    int z = kz < 0 ? kz + g.n : kz;
    // tx > 0: u in z-pencil tiles of tx x-values, lines lp elements apart
    // (component stride g.s_comp)
    long off = tx > 0 ? (long(kx / tx) * g.n + y) * lp + long(z) * tx + kx % tx : kx + long(g.nhp) * (y + long(g.n) * z);
    double w = (kx >= 1 && kx <= g.nh - 2) ? 2.0 : 1.0;
    for (int c = 0; c < ncomp; ++c) {
      C<T> v = u[c * g.s_comp + off];
      acc += w * (double(v.x) * v.x + double(v.y) * v.y);
    }
  }
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *A = sm[0] > 0 ? injection / sm[0] : 0.0;
}

// ------------------------------------------------------------ initial data
DEV unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Counter-based complex normals (re, im ~ N(0, 1/2)) keyed by the reference
// linear index x + nh*(y + n*z) and the component.
template <typename T>
__global__ void k_randn(C<T>* u, int ncomp, SpecGeom g, unsigned long long seed) {
  Mode m;
  if (!mode_of(gid(), g, m)) return;
  const unsigned long long G = 0x9E3779B97F4A7C15ull;
  unsigned long long idx = (unsigned long long)m.x + (unsigned long long)g.nh * ((unsigned long long)m.y + (unsigned long long)g.n * m.z);
  for (int c = 0; c < ncomp; ++c) {
    unsigned long long key = mix64(seed * G + (unsigned long long)(c + 1));
    unsigned long long b0 = mix64(key + (2ull * idx + 1ull) * G);
    unsigned long long b1 = mix64(key + (2ull * idx + 2ull) * G);
    double u1 = double((b0 >> 11) + 1ull) * 0x1.0p-53;
    double u2 = double(b1 >> 11) * 0x1.0p-53;
    double r = sqrt(-2.0 * log(u1));
    double th = 2.0 * 3.14159265358979323846 * u2;
    double h = sqrt(0.5);
    u[c * g.s_comp + m.off] = mk<T>(T(r * cos(th) * h), T(r * sin(th) * h));
  }
}

// symmetrize! (src/solver.jl:398-430) on the kx = 0 and kx = n/2 planes.
// Every source entry (ix, piy, piz) is itself never written, so the kernel
// is race free.
template <typename T>
__global__ void k_symmetrize(C<T>* u, int ncomp, SpecGeom g) {
  long id = gid();
  const int n = g.n, kmax = n / 2;
  long plane = long(n) * g.nz;
  if (id >= 2 * plane) return;
  int ix = id < plane ? 0 : kmax;
  long r = id % plane;
  int iy = int(r % n), iz = int(r / n);
  bool iy_self = iy == 0 || iy == kmax;
  for (int c = 0; c < ncomp; ++c) {
    C<T>* uc = u + c * g.s_comp;
    long off = ix + long(g.nhp) * (iy + long(n) * iz);
    if (g.D == 2) {
      if (iy_self) uc[off].y = T(0);
      else if (iy > kmax) uc[off] = conjc<T>(uc[ix + long(g.nhp) * (n - iy)]);
      continue;
    }
    bool iz_self = iz == 0 || iz == kmax;
    int piy = iy_self ? iy : n - iy;
    int piz = iz_self ? iz : n - iz;
    if (iy_self && iz_self) uc[off].y = T(0);
    else if (iy > piy || (iy == piy && iz > piz))
      uc[off] = conjc<T>(uc[ix + long(g.nhp) * (piy + long(n) * piz)]);
  }
}

// taylorgreen (src/solver.jl:537-570): physical field on cell centres.
template <typename T>
__global__ void k_taylorgreen_phys(T* v, int comp, int D, int n, double l, double V0) {
  long id = gid();
  long tot = D == 3 ? long(n) * n * n : long(n) * n;
  if (id >= tot) return;
  int ix = int(id % n), iy = int((id / n) % n), iz = int(id / (long(n) * n));
  double h = l / n;
  double x = h / 2 + ix * h, y = h / 2 + iy * h, z = h / 2 + iz * h;
  // range(h/2, l-h/2, n) is exact at the end points; interior points agree to rounding
  double sx = sinpi(2 * x / l), cx = cospi(2 * x / l);
  double sy = sinpi(2 * y / l), cy = cospi(2 * y / l);
  double r;
  if (D == 2) r = comp == 0 ? -V0 * sx * cy : V0 * cx * sy;
  else {
    double cz = cospi(2 * z / l);
    r = comp == 0 ? V0 * sx * cy * cz : (comp == 1 ? -V0 * cx * sy * cz : 0.0);
  }
  v[id] = T(r);
}

// ----------------------------------------------------------------- closures
// Physical tensor fields: component stride sp.  G uses the nonsym order
// (index i + D*j holds dG_i/dx_j); tau is packed xx, yy, zz, xy, yz, zx.
template <typename T, int D>
DEV void load_grad(const T* GG, long sp, long id, T (&G)[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) G[i][j] = T(0);
#pragma unroll
  for (int j = 0; j < D; ++j)
#pragma unroll
    for (int i = 0; i < D; ++i) G[i][j] = GG[(i + D * j) * sp + id];
}

template <typename T, int D>
DEV void store_sym(T* tau, long sp, long id, const T (&t)[3][3]) {
  if constexpr (D == 2) {
    tau[0 * sp + id] = t[0][0];
    tau[1 * sp + id] = t[1][1];
    tau[2 * sp + id] = t[0][1];
  } else {
    tau[0 * sp + id] = t[0][0];
    tau[1 * sp + id] = t[1][1];
    tau[2 * sp + id] = t[2][2];
    tau[3 * sp + id] = t[0][1];
    tau[4 * sp + id] = t[1][2];
    tau[5 * sp + id] = t[2][0];
  }
}

// clark_kernel! (src/closures.jl:37-42): tau = Delta^2/12 G G^T
template <typename T, int D>
__global__ void k_clark(T* tau, const T* GG, double delta, long np) {
  long id = gid();
  if (id >= np) return;
  T G[3][3], t[3][3];
  load_grad<T, D>(GG, np, id, G);
  const T c = T(delta * delta / 12);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      T s = 0;
#pragma unroll
      for (int k = 0; k < D; ++k) s += G[i][k] * G[j][k];
      t[i][j] = c * s;
    }
  store_sym<T, D>(tau, np, id, t);
}

// smagorinsky_kernel! (src/closures.jl:44-51)
template <typename T, int D>
__global__ void k_smagorinsky(T* tau, const T* GG, double cs, double delta, long np) {
  long id = gid();
  if (id >= np) return;
  T G[3][3], S[3][3];
  load_grad<T, D>(GG, np, id, G);
  T ss = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      S[i][j] = (G[i][j] + G[j][i]) / 2;
      ss += S[i][j] * S[i][j];
    }
  T nu = T(cs * cs * delta * delta) * sqrt(2 * ss);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) S[i][j] = -2 * nu * S[i][j];
  store_sym<T, D>(tau, np, id, S);
}

// verstappen_kernel! (src/closures.jl:53-62), the Q-R model (3D):
// q = tr(S S) / 2, r = tr(S S S) / 3, nu = C^2 Delta^2 |r| / q, tau = -2 nu S
template <typename T>
__global__ void k_verstappen(T* tau, const T* GG, double cq, double delta, long np) {
  long id = gid();
  if (id >= np) return;
  T G[3][3], S[3][3], SS[3][3];
  load_grad<T, 3>(GG, np, id, G);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) S[i][j] = (G[i][j] + G[j][i]) / 2;
  T q = 0, r = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      SS[i][j] = S[i][0] * S[0][j] + S[i][1] * S[1][j] + S[i][2] * S[2][j];
      q += S[i][j] * S[j][i];
    }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) r += SS[i][j] * S[j][i];
  q /= 2;
  r /= 3;
  const T nu = T(cq * cq * delta * delta) * fabs(r) / q;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) S[i][j] = -2 * nu * S[i][j];
  store_sym<T, 3>(tau, np, id, S);
}

// pointwise SFS dissipation eps = -tau:S (contract_dissipation,
// src/solver.jl:746-759; strain_from_gradient, :742-744) from the packed
// stress and the physical velocity gradient
template <typename T, int D>
__global__ void k_sfs_dissipation(T* eps, const T* tau, const T* GG, long np) {
  long id = gid();
  if (id >= np) return;
  T G[3][3];
  load_grad<T, D>(GG, np, id, G);
  auto S = [&](int i, int j) { return (G[i][j] + G[j][i]) / 2; };
  T d;
  if constexpr (D == 2) {
    d = tau[id] * S(0, 0) + tau[np + id] * S(1, 1) + 2 * tau[2 * np + id] * S(0, 1);
  } else {
    d = tau[id] * S(0, 0) + tau[np + id] * S(1, 1) + tau[2 * np + id] * S(2, 2) +
        2 * (tau[3 * np + id] * S(0, 1) + tau[4 * np + id] * S(1, 2) + tau[5 * np + id] * S(2, 0));
  }
  eps[id] = -d;
}

// FP64 block partials of sample statistics over x[0, n): kind 0 the sum;
// kind 1 sum (x - m)^2, sum (x - m)^3 and the count of x < 0 (three partial
// arrays of gridDim.x); kind 2 for the pair (a, b): sums of a, b, a^2
// (relative to ma), b^2 (relative to mb), (a - ma)(b - mb), (b - a)^2
template <typename T>
__global__ void k_sample_stats(const T* x, const T* y, long n, int kind, double ma, double mb, double* partial) {
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (long i = gid(); i < n; i += long(gridDim.x) * blockDim.x) {
    const double a = double(x[i]);
    if (kind == 0) {
      v[0] += a;
    } else if (kind == 1) {
      const double e = a - ma;
      v[0] += e * e;
      v[1] += e * e * e;
      v[2] += a < 0 ? 1.0 : 0.0;
    } else {
      const double b = double(y[i]);
      const double ea = a - ma, eb = b - mb;
      v[0] += a;
      v[1] += b;
      v[2] += ea * ea;
      v[3] += eb * eb;
      v[4] += ea * eb;
      v[5] += (b - a) * (b - a);
    }
  }
  const int nv = kind == 0 ? 1 : (kind == 1 ? 3 : 6);
  for (int q = 0; q < nv; ++q) {
    block_sum_store<int>(v[q], partial + q * gridDim.x);
    __syncthreads();
  }
}

// ---------------------------------------------------------- radix selection
// Order statistics of floating-point samples without sorting: the value of
// rank k is found digit by digit (8 bits at a time, most significant first)
// on order-preserving unsigned keys (negative values bit-inverted, the sign
// bit set on the others), each pass counting the digit over the samples whose
// higher digits match the prefix found so far.  Two ranks are selected in the
// same passes (the two middle elements of a median).
DEV unsigned long long okey(float x) {
  const unsigned b = __float_as_uint(x);
  return (b >> 31) ? ~b : (b | 0x80000000u);
}
DEV unsigned long long okey(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | (1ull << 63));
}
struct RadixSel {
  unsigned long long prefix, mask;  // digits found so far and their bit mask
  long long k;                      // rank among the samples matching the prefix
};
static __global__ void k_radix_init(RadixSel* sel, long long r0, long long r1, unsigned long long* hist) {
  if (threadIdx.x == 0) {
    sel[0] = RadixSel{0, 0, r0};
    sel[1] = RadixSel{0, 0, r1};
  }
  for (int b = threadIdx.x; b < 512; b += blockDim.x) hist[b] = 0;
}
// hist[256 j + d]: samples matching selection j's prefix whose digit at
// `shift` is d
template <typename T>
__global__ void __launch_bounds__(256) k_radix_hist(const T* x, long n, const RadixSel* sel, int shift,
                                                    unsigned long long* hist) {
  __shared__ unsigned h[512];
  for (int b = threadIdx.x; b < 512; b += 256) h[b] = 0;
  __syncthreads();
  const RadixSel s0 = sel[0], s1 = sel[1];
  for (long i = gid(); i < n; i += long(gridDim.x) * blockDim.x) {
    const unsigned long long key = okey(x[i]);
    const unsigned d = unsigned(key >> shift) & 255u;
    if ((key & s0.mask) == s0.prefix) atomicAdd(&h[d], 1u);
    if ((key & s1.mask) == s1.prefix) atomicAdd(&h[256 + d], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < 512; b += 256)
    if (h[b]) atomicAdd(&hist[b], (unsigned long long)h[b]);
}
// the digit holding each selection's rank extends its prefix; the
// histogram is cleared for the next pass
static __global__ void k_radix_pick(RadixSel* sel, int shift, unsigned long long* hist) {
  if (threadIdx.x < 2) {
    RadixSel s = sel[threadIdx.x];
    const unsigned long long* h = hist + 256 * threadIdx.x;
    unsigned long long below = 0;
    for (int b = 0; b < 256; ++b) {
      if ((unsigned long long)s.k < below + h[b]) {
        s.prefix |= (unsigned long long)b << shift;
        s.mask |= 255ull << shift;
        s.k -= (long long)below;
        break;
      }
      below += h[b];
    }
    sel[threadIdx.x] = s;
  }
  __syncthreads();
  for (int b = threadIdx.x; b < 512; b += blockDim.x) hist[b] = 0;
}

// smagorinsky_tensor! (src/closures.jl:117-150), optional output scale c
template <typename T, int D>
__global__ void k_smag_tensor(T* m, const T* S, double delta, long np) {
  long id = gid();
  if (id >= np) return;
  constexpr int NT = D == 3 ? 6 : 3;
  T s[NT];
#pragma unroll
  for (int c = 0; c < NT; ++c) s[c] = S[c * np + id];
  T S2;
  if constexpr (D == 2) S2 = s[0] * s[0] + s[1] * s[1] + 2 * s[2] * s[2];
  else S2 = s[0] * s[0] + s[1] * s[1] + s[2] * s[2] + 2 * s[3] * s[3] + 2 * s[4] * s[4] + 2 * s[5] * s[5];
  T nu = T(delta * delta) * sqrt(2 * S2);
#pragma unroll
  for (int c = 0; c < NT; ++c) m[c * np + id] = -2 * nu * s[c];
}

// The dynamic model's spectral steps.  dst = 2/3 G(src) (- minus when minus
// is non-null): the Gaussian filter of width delta with the forced-shell
// carve-out, then the 2/3 mask (twothirds! after gaussianfilter!, one pass);
// dst may alias src or minus.
template <typename T>
__global__ void k_gauss23(C<T>* dst, const C<T>* src, const C<T>* minus, double delta, SpecGeom g, int ncomp) {
  Mode m;
  if (!mode_of(gid(), g, m)) return;
  const bool keep = in23(m, g.n / 3);
  const double kx = g.kf * m.kx, ky = g.kf * m.ky, kz = g.kf * m.kz;
  const int k2i = m.kx * m.kx + m.ky * m.ky + m.kz * m.kz;
  const T w = k2i < 9 ? T(1) : T(exp(-delta * delta * (kx * kx + ky * ky + kz * kz) / 24));
  for (int c = 0; c < ncomp; ++c) {
    const long o = c * g.s_comp + m.off;
    C<T> v = keep ? scl<T>(src[o], w) : czero<T>();
    if (minus) v = sub<T>(v, minus[o]);
    dst[o] = v;
  }
}

// smagorinsky_ml_mm! (src/closures.jl:152-179) for the dynamic model with M
// formed on the fly: M = Mt - m2, Mt the test-filtered m1 and m2 =
// smagorinsky_tensor(St, dd) of the combined-filter strain St; FP64 block
// sums of M:L_dev (partial[b]) and M:M (partial[nblocks + b]).
template <typename T, int D>
__global__ void k_ml_mm_fused(const T* Mt, const T* St, const T* L, double dd, long np, double* partial) {
  constexpr int NT = D == 3 ? 6 : 3;
  const long id = gid();
  double ml = 0, mm = 0;
  if (id < np) {
    T s[NT], l[NT], M[NT];
#pragma unroll
    for (int c = 0; c < NT; ++c) {
      s[c] = St[c * np + id];
      l[c] = L[c * np + id];
    }
    T S2;
    if constexpr (D == 2) S2 = s[0] * s[0] + s[1] * s[1] + 2 * s[2] * s[2];
    else S2 = s[0] * s[0] + s[1] * s[1] + s[2] * s[2] + 2 * s[3] * s[3] + 2 * s[4] * s[4] + 2 * s[5] * s[5];
    const T nu = T(dd * dd) * sqrt(2 * S2);
#pragma unroll
    for (int c = 0; c < NT; ++c) M[c] = Mt[c * np + id] - (-2 * nu * s[c]);
    if constexpr (D == 2) {
      const T tr = (l[0] + l[1]) / 2;
      l[0] -= tr;
      l[1] -= tr;
      ml = double(M[0] * l[0] + M[1] * l[1] + 2 * M[2] * l[2]);
      mm = double(M[0] * M[0] + M[1] * M[1] + 2 * M[2] * M[2]);
    } else {
      const T tr = (l[0] + l[1] + l[2]) / 3;
      l[0] -= tr;
      l[1] -= tr;
      l[2] -= tr;
      ml = double(M[0] * l[0] + M[1] * l[1] + M[2] * l[2] + 2 * M[3] * l[3] + 2 * M[4] * l[4] + 2 * M[5] * l[5]);
      mm = double(M[0] * M[0] + M[1] * M[1] + M[2] * M[2] + 2 * M[3] * M[3] + 2 * M[4] * M[4] + 2 * M[5] * M[5]);
    }
  }
  block_sum_store<int>(ml, partial);
  __syncthreads();
  block_sum_store<int>(mm, partial + gridDim.x);
}

// the Lilly coefficient c = max(-<ML>/<MM>, 0), 0 when <MM> = 0, from the
// two sums (sums[0] = <ML>, sums[1] = <MM>), kept on the device
static __global__ void k_dyn_coef(const double* sums, double* coef) {
  *coef = sums[1] == 0.0 ? 0.0 : fmax(-sums[0] / sums[1], 0.0);
}

// y = c scale 2/3(x) (accumulate: y += ...), c read from the device: the
// dynamic model's stress c to_spec(dealias_phys(m1)) from x = rfft(m1)
template <typename T>
__global__ void k_coef23(C<T>* y, const C<T>* x, const double* coef, double scale, int accumulate, SpecGeom g,
                         int ncomp) {
  Mode m;
  if (!mode_of(gid(), g, m)) return;
  const bool keep = in23(m, g.n / 3);
  const T a = T(*coef * scale);
  for (int c = 0; c < ncomp; ++c) {
    const long o = c * g.s_comp + m.off;
    const C<T> v = keep ? scl<T>(x[o], a) : czero<T>();
    y[o] = accumulate ? add<T>(y[o], v) : v;
  }
}

// physical helpers
template <typename T>
__global__ void k_phys_axpby(T* y, const T* x, T a, T b, long n) {  // y = a*x + b*y
  long id = gid();
  if (id < n) y[id] = a * x[id] + b * y[id];
}
template <typename T>
__global__ void k_phys_scale(T* y, double a, long n) {
  long id = gid();
  if (id < n) y[id] = T(a) * y[id];
}
// make_tracefree! on D diagonal comps of a packed physical tensor
template <typename T>
__global__ void k_phys_tracefree(T* tau, int D, long np) {
  long id = gid();
  if (id >= np) return;
  if (D == 2) {
    T tr = (tau[id] + tau[np + id]) / 2;
    tau[id] -= tr; tau[np + id] -= tr;
  } else {
    T tr = (tau[id] + tau[np + id] + tau[2 * np + id]) / 3;
    tau[id] -= tr; tau[np + id] -= tr; tau[2 * np + id] -= tr;
  }
}
template <typename T>
__global__ void k_spec_tracefree(C<T>* tau, SpecGeom g) {
  Mode m;
  if (!mode_of(gid(), g, m)) return;
  const long s = g.s_comp, o = m.off;
  if (g.D == 2) {
    C<T> t2 = add<T>(tau[o], tau[s + o]);
    C<T> tr = mk<T>(t2.x / T(2), t2.y / T(2));
    tau[o] = sub<T>(tau[o], tr); tau[s + o] = sub<T>(tau[s + o], tr);
  } else {
    C<T> t3 = add<T>(add<T>(tau[o], tau[s + o]), tau[2 * s + o]);
    C<T> tr = mk<T>(t3.x / T(3), t3.y / T(3));
    tau[o] = sub<T>(tau[o], tr); tau[s + o] = sub<T>(tau[s + o], tr); tau[2 * s + o] = sub<T>(tau[2 * s + o], tr);
  }
}

}  // namespace sc
