// Host-side launchers of the typed operations (included once per precision).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#include "ops.hpp"
#include "fft_passes.hpp"
#include "rhs_passes.hpp"
#include "cnn_link.hpp"
#include "pmlp_link.hpp"

namespace sc {

#define SC_FOR_N(X) X(4) X(6) X(8) X(12) X(16) X(24) X(30) X(32) X(40) X(48) X(64) X(96) X(128) X(256) X(512) X(810)
#define SC_FOR_M(X) X(2) X(3) X(4) X(6) X(8) X(12) X(15) X(16) X(20) X(24) X(32) X(48) X(64) X(128) X(256) X(405)

static inline unsigned nblk(long total, int bs = 256) { return unsigned((total + bs - 1) / bs); }

template <typename T> static int tile_x_rt(int N) {
  switch (N) {
#define X(NN) case NN: return tile_x<T, NN>();
    SC_FOR_N(X)
#undef X
  }
  throw Error(2, "unsupported transform length " + std::to_string(N));
}
static int tpp_rt(int N) {
  switch (N) {
#define X(NN) case NN: return Plan<NN>::tpp;
    SC_FOR_N(X)
    X(2) X(3) X(15) X(20) X(405)
#undef X
  }
  throw Error(2, "unsupported transform length " + std::to_string(N));
}
static int rows_rt(int M) {
  switch (M) {
#define X(MM) case MM: return rows_per_block<MM>();
    SC_FOR_M(X)
#undef X
  }
  throw Error(2, "unsupported half length " + std::to_string(M));
}

template <typename T>
bool Ops<T>::supported(int n) {
  switch (n) {
#define X(NN) case NN: return true;
    SC_FOR_N(X)
#undef X
  }
  return false;
}

// geometry of a strided complex pass along axis (1 = y, 2 = z) of length n,
// input and output in the x-contiguous spectral layout
template <typename T>
static StridedGeom axis_geom(const Ctx& c, int axis, int xs, bool prune_other, int kin, bool mask23, double scale) {
  const int n = c.n;
  const int TX = tile_x_rt<T>(n);
  StridedGeom g;
  g.xs = xs;
  g.ntile = (xs + TX - 1) / TX;
  const int kc = n / 3;
  const int nother_full = axis == 1 ? c.nz : n;
  if (prune_other && axis == 2) {
    g.nother = 2 * kc + 1;
    g.other_lo = kc + 1;
    g.other_gap = n - 2 * kc - 1;
  } else {
    g.nother = nother_full;
    g.other_lo = nother_full;
    g.other_gap = 0;
  }
  g.n_other = axis == 1 ? c.nz : n;
  g.s_axis = axis == 1 ? long(c.nhp) : long(c.nhp) * n;
  g.s_other = axis == 1 ? long(c.nhp) * n : long(c.nhp);
  g.s_comp = c.spec_elems;
  g.s_tile = TX;
  g.o_axis = g.s_axis;
  g.o_other = g.s_other;
  g.o_comp = g.s_comp;
  g.o_tile = g.s_tile;
  g.kin = kin;
  g.mask23 = mask23 ? 1 : 0;
  g.kcut = kc;
  g.scale = scale;
  return g;
}

template <typename T>
static void c2c_launch(Ctx& c, const StridedGeom& g, int axis, int sign, const C<T>* in, C<T>* out, int ncomp) {
  const int n = c.n;
  dim3 grid(unsigned(g.ntile) * unsigned(g.nother), 1);
  const C<T>* tw = reinterpret_cast<const C<T>*>(c.tw_n);
  static const char* names[2][2] = {{"c2c_y_fwd", "c2c_y_inv"}, {"c2c_z_fwd", "c2c_z_inv"}};
  c.prof_begin(names[axis - 1][sign > 0]);
  switch (n) {
#define X(NN)                                                                                        \
  case NN:                                                                                           \
    if (sign < 0)                                                                                    \
      hipLaunchKernelGGL((k_c2c_strided<T, NN, -1>), grid, dim3(tile_x<T, NN>() * Plan<NN>::tpp), 0, \
                         c.stream, in, out, tw, g, ncomp);                                           \
    else                                                                                             \
      hipLaunchKernelGGL((k_c2c_strided<T, NN, 1>), grid, dim3(tile_x<T, NN>() * Plan<NN>::tpp), 0,  \
                         c.stream, in, out, tw, g, ncomp);                                           \
    break;
    SC_FOR_N(X)
#undef X
    default: throw Error(2, "unsupported grid size " + std::to_string(n));
  }
  c.prof_end();
  SC_HIP(hipGetLastError());
}

// strided complex pass along axis (1 = y, 2 = z) of length n
template <typename T>
static void c2c_axis(Ctx& c, int axis, int sign, const C<T>* in, C<T>* out, int ncomp, int xs, bool prune_other,
                     int kin, bool mask23, double scale) {
  c2c_launch<T>(c, axis_geom<T>(c, axis, xs, prune_other, kin, mask23, scale), axis, sign, in, out, ncomp);
}

template <typename T>
static RowGeom row_geom(const Ctx& c, int kin, double scale) {
  RowGeom g;
  g.nrows = long(c.n) * c.nz;
  g.nhp = c.nhp;
  g.n = c.n;
  g.kin = kin;
  g.s_comp_spec = c.spec_elems;
  g.s_comp_phys = c.phys_elems;
  g.scale = scale;
  return g;
}

template <typename T>
static void x_c2r(Ctx& c, T* out, const C<T>* in, int ncomp, int kin, double scale) {
  const int M = c.n / 2;
  RowGeom g = row_geom<T>(c, kin, scale);
  const int TY = rows_rt(M);
  dim3 grid(unsigned((g.nrows + TY - 1) / TY), unsigned(ncomp));
  auto twm = reinterpret_cast<const C<T>*>(c.tw_m);
  auto twn = reinterpret_cast<const C<T>*>(c.tw_n);
  c.prof_begin("x_c2r");
  switch (M) {
#define X(MM)                                                                                           \
  case MM:                                                                                              \
    hipLaunchKernelGGL((k_x_c2r<T, MM>), grid, dim3(rows_per_block<MM>() * Plan<MM>::tpp), 0, c.stream, \
                       in, out, twm, twn, g);                                                           \
    break;
    SC_FOR_M(X)
#undef X
    default: throw Error(2, "unsupported grid size");
  }
  c.prof_end();
  SC_HIP(hipGetLastError());
}

template <typename T>
static void x_r2c(Ctx& c, C<T>* out, const T* in, int ncomp, double scale) {
  const int M = c.n / 2;
  RowGeom g = row_geom<T>(c, M, scale);
  const int TY = rows_rt(M);
  dim3 grid(unsigned((g.nrows + TY - 1) / TY), unsigned(ncomp));
  auto twm = reinterpret_cast<const C<T>*>(c.tw_m);
  auto twn = reinterpret_cast<const C<T>*>(c.tw_n);
  c.prof_begin("x_r2c");
  switch (M) {
#define X(MM)                                                                                           \
  case MM:                                                                                              \
    hipLaunchKernelGGL((k_x_r2c<T, MM>), grid, dim3(rows_per_block<MM>() * Plan<MM>::tpp), 0, c.stream, \
                       in, out, twm, twn, g);                                                           \
    break;
    SC_FOR_M(X)
#undef X
    default: throw Error(2, "unsupported grid size");
  }
  c.prof_end();
  SC_HIP(hipGetLastError());
}

#ifndef SC_XPROD_PREFETCH
#define SC_XPROD_PREFETCH 1
#endif
#ifndef SC_XPROD_WAVE
#define SC_XPROD_WAVE 1
#endif
template <typename T>
static void x_products(Ctx& c, C<T>* out, const C<T>* in, int kin, double scale, unsigned long long* vmax) {
  const int M = c.n / 2;
  RowGeom g = row_geom<T>(c, kin, scale);
  auto twm = reinterpret_cast<const C<T>*>(c.tw_m);
  auto twn = reinterpret_cast<const C<T>*>(c.tw_n);
  c.prof_begin("x_products");
  if (SC_XPROD_PREFETCH && M == 405 && g.kin < xp_entries<T>()) {
    // persistent: two blocks of SC_XP_WAVES wavefronts per compute unit
    constexpr int NW = SC_XP_WAVES;
    const long blocks = std::min<long>((g.nrows + NW - 1) / NW, 2L * c.ncu);
    if (c.D == 3)
      hipLaunchKernelGGL((k_x_products_p<T, 405, 3>), dim3(unsigned(blocks)), dim3(64 * NW), 0, c.stream, in, out,
                         twm, twn, g, vmax);
    else
      hipLaunchKernelGGL((k_x_products_p<T, 405, 2>), dim3(unsigned(blocks)), dim3(64 * NW), 0, c.stream, in, out,
                         twm, twn, g, vmax);
    c.prof_end();
    SC_HIP(hipGetLastError());
    return;
  }
  if (SC_XPROD_WAVE && tpp_rt(M) <= 64) {
    constexpr int NW = SC_XW_WAVES;
    switch (M) {
#define X(MM)                                                                                            \
  case MM:                                                                                               \
    if constexpr (Plan<MM>::tpp <= 64) {                                                                 \
      constexpr long RB = long(NW) * xw_rows_per_wave<MM>();  /* rows per block */                       \
      const dim3 grid(unsigned((g.nrows + RB - 1) / RB), 1);                                             \
      if (c.D == 3)                                                                                      \
        hipLaunchKernelGGL((k_x_products_w<T, MM, 3>), grid, dim3(64 * NW), 0, c.stream, in, out, twm, twn, g, vmax); \
      else                                                                                               \
        hipLaunchKernelGGL((k_x_products_w<T, MM, 2>), grid, dim3(64 * NW), 0, c.stream, in, out, twm, twn, g, vmax); \
    }                                                                                                    \
    break;
      SC_FOR_M(X)
#undef X
      default: throw Error(2, "unsupported grid size");
    }
  } else {
    const int TY = rows_rt(M);
    dim3 grid(unsigned((g.nrows + TY - 1) / TY), 1);
    switch (M) {
#define X(MM)                                                                                                 \
  case MM:                                                                                                    \
    if (c.D == 3)                                                                                             \
      hipLaunchKernelGGL((k_x_products<T, MM, 3>), grid, dim3(rows_per_block<MM>() * Plan<MM>::tpp), 0,       \
                         c.stream, in, out, twm, twn, g, vmax);                                               \
    else                                                                                                      \
      hipLaunchKernelGGL((k_x_products<T, MM, 2>), grid, dim3(rows_per_block<MM>() * Plan<MM>::tpp), 0,       \
                         c.stream, in, out, twm, twn, g, vmax);                                               \
    break;
      SC_FOR_M(X)
#undef X
      default: throw Error(2, "unsupported grid size");
    }
  }
  c.prof_end();
  SC_HIP(hipGetLastError());
}

// geometry of a full (unpruned) strided pass along axis (1 = y, 2 = z)
template <typename T>
static StridedGeom full_geom(const Ctx& c, int axis) {
  const int n = c.n;
  const int TX = tile_x_rt<T>(n);
  StridedGeom g;
  g.xs = c.nh;
  g.ntile = (c.nh + TX - 1) / TX;
  g.nother = axis == 1 ? c.nz : n;
  g.other_lo = g.nother;
  g.other_gap = 0;
  g.n_other = axis == 1 ? c.nz : n;
  g.s_axis = axis == 1 ? long(c.nhp) : long(c.nhp) * n;
  g.s_other = axis == 1 ? long(c.nhp) * n : long(c.nhp);
  g.s_comp = c.spec_elems;
  g.s_tile = TX;
  g.o_axis = g.s_axis;
  g.o_other = g.s_other;
  g.o_comp = g.s_comp;
  g.o_tile = g.s_tile;
  g.kin = n;
  g.mask23 = 0;
  g.kcut = n / 3;
  g.scale = 1.0;
  return g;
}

// component stride of the tiled buffers between strided passes (x tiles of
// TX, each holding n x n pencil positions)
template <typename T>
static long tile_line_rt(const Ctx& c) {
  return long(c.n) * tile_x_rt<T>(c.n) + SC_TILE_PAD;
}
template <typename T>
static long tiled_slot(const Ctx& c) {
  const int TX = tile_x_rt<T>(c.n);
  return long((c.nh + TX - 1) / TX) * c.n * tile_line_rt<T>(c);
}

template <typename T>
static void y_fwd_div(Ctx& c, const C<T>* buf, C<T>* outT) {
  StridedGeom g = full_geom<T>(c, 1);
  dim3 grid(unsigned(g.ntile) * unsigned(g.nother));
  const C<T>* tw = reinterpret_cast<const C<T>*>(c.tw_n);
  const double kf = c.geom().kf;
  c.prof_begin("y_fwd_div");
  switch (c.n) {
#define X(NN)                                                                                          \
  case NN:                                                                                             \
    hipLaunchKernelGGL((k_y_fwd_div<T, NN>), grid, dim3(wide_threads<T, NN>()), 0, c.stream, buf, outT, \
                       tiled_slot<T>(c), tw, g, kf);                                                  \
    break;
    SC_FOR_N(X)
#undef X
    default: throw Error(2, "unsupported grid size");
  }
  c.prof_end();
  SC_HIP(hipGetLastError());
}

template <typename T>
static void fwd_rhs_last(Ctx& c, const C<T>* buf, const C<T>* u, C<T>* du, RhsParams rp, bool lin = true) {
  const int axis = c.D == 3 ? 2 : 1;
  StridedGeom g = full_geom<T>(c, axis);
  dim3 grid(unsigned(g.ntile) * unsigned(g.nother));
  const C<T>* tw = reinterpret_cast<const C<T>*>(c.tw_n);
  c.prof_begin(c.D == 3 ? (lin ? "z_fwd_rhs" : "z_fwd_div") : "y_fwd_rhs2d");
  switch (c.n) {
#define X(NN)                                                                                                  \
  case NN:                                                                                                     \
    if (c.D == 3 && lin)                                                                                       \
      hipLaunchKernelGGL((k_z_fwd_rhs<T, NN, true>), grid, dim3(wide_threads<T, NN>()), 0, c.stream, buf,         \
                         tiled_slot<T>(c), u, du, tw, g, rp);                                                  \
    else if (c.D == 3)                                                                                         \
      hipLaunchKernelGGL((k_z_fwd_rhs<T, NN, false>), grid, dim3(wide_threads<T, NN>()), 0, c.stream, buf,        \
                         tiled_slot<T>(c), u, du, tw, g, rp);                                                  \
    else                                                                                                       \
      hipLaunchKernelGGL((k_y_fwd_rhs2d<T, NN>), grid, dim3(wide_threads<T, NN>()), 0, c.stream, buf, u, du, tw, g, rp); \
    break;
    SC_FOR_N(X)
#undef X
    default: throw Error(2, "unsupported grid size");
  }
  c.prof_end();
  SC_HIP(hipGetLastError());
}

// ----------------------------------------------------------------- tables
template <typename T>
void Ops<T>::init_tables(Ctx& c) {
  const int n = c.n, M = n / 2;
  c.state_elems = c.D == 3 ? std::max(c.spec_elems, tiled_slot<T>(c)) : c.spec_elems;
  std::vector<C<T>> hn(n), hm(M);
  for (int k = 0; k < n; ++k) {
    long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
    hn[k] = mk<T>(T(std::cos(a)), T(std::sin(a)));
  }
  for (int k = 0; k < M; ++k) {
    long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)M;
    hm[k] = mk<T>(T(std::cos(a)), T(std::sin(a)));
  }
  // tw_n carries four zero entries after its n twiddles (a zero source for
  // direct-to-LDS loads)
  hn.resize(n + 4, mk<T>(T(0), T(0)));
  SC_HIP(hipMalloc(&c.tw_n, (n + 4) * sizeof(C<T>)));
  SC_HIP(hipMalloc(&c.tw_m, (M > 0 ? M : 1) * sizeof(C<T>)));
  SC_HIP(hipMemcpy(c.tw_n, hn.data(), (n + 4) * sizeof(C<T>), hipMemcpyHostToDevice));
  SC_HIP(hipMemcpy(c.tw_m, hm.data(), M * sizeof(C<T>), hipMemcpyHostToDevice));
}

// ------------------------------------------------------------- transforms
// spectral -> physical: z (3D) and y complex inverse passes into scratch,
// then the x C2R.  mask23 applies twothirds! on the fly (inputs outside the
// 2/3 box are never read).
template <typename T>
void Ops<T>::c2r(Ctx& c, T* phys, const CT* spec, int ncomp, double scale, bool mask23) {
  CT* s = reinterpret_cast<CT*>(c.scratch(ncomp));
  const int kc = c.n / 3;
  const int kin = mask23 ? kc : c.n;
  const int xs = mask23 ? kc + 1 : c.nh;
  if (c.D == 3) {
    c2c_axis<T>(c, 2, +1, spec, s, ncomp, xs, mask23, kin, mask23, 1.0);
    c2c_axis<T>(c, 1, +1, s, s, ncomp, xs, false, kin, false, 1.0);
  } else {
    c2c_axis<T>(c, 1, +1, spec, s, ncomp, xs, false, kin, mask23, 1.0);
  }
  x_c2r<T>(c, phys, s, ncomp, mask23 ? kc : c.nh, scale);
}

template <typename T>
void Ops<T>::r2c(Ctx& c, CT* spec, const T* phys, int ncomp, double scale) {
  x_r2c<T>(c, spec, phys, ncomp, scale);
  c2c_axis<T>(c, 1, -1, spec, spec, ncomp, c.nh, false, c.n, false, 1.0);
  if (c.D == 3) c2c_axis<T>(c, 2, -1, spec, spec, ncomp, c.nh, false, c.n, false, 1.0);
}

// ----------------------------------------------------------- spectral ops
#define SPEC_LAUNCH(kern, ...)                                                             \
  do {                                                                                     \
    SpecGeom g_ = c.geom();                                                                \
    c.prof_begin(#kern);                                                                   \
    hipLaunchKernelGGL(kern, dim3(nblk(g_.total)), dim3(256), 0, c.stream, __VA_ARGS__);  \
    c.prof_end();                                                                          \
    SC_HIP(hipGetLastError());                                                             \
  } while (0)

template <typename T> void Ops<T>::twothirds(Ctx& c, CT* u, int ncomp) { SPEC_LAUNCH(k_twothirds<T>, u, g_, ncomp); }
template <typename T> void Ops<T>::project(Ctx& c, CT* u) {
  if (c.D == 3) SPEC_LAUNCH((k_project<T, 3>), u, g_);
  else SPEC_LAUNCH((k_project<T, 2>), u, g_);
}
template <typename T> void Ops<T>::viscosity(Ctx& c, CT* sig, const CT* u, double nu) {
  if (c.D == 3) SPEC_LAUNCH((k_viscosity<T, 3>), sig, u, nu, g_);
  else SPEC_LAUNCH((k_viscosity<T, 2>), sig, u, nu, g_);
}
template <typename T> void Ops<T>::tensordivergence(Ctx& c, CT* du, const CT* sig) {
  if (c.D == 3) SPEC_LAUNCH((k_tensordivergence<T, 3>), du, sig, g_);
  else SPEC_LAUNCH((k_tensordivergence<T, 2>), du, sig, g_);
}
template <typename T> void Ops<T>::vectorgradient(Ctx& c, CT* G, const CT* u) {
  if (c.D == 3) SPEC_LAUNCH((k_vectorgradient<T, 3>), G, u, g_);
  else SPEC_LAUNCH((k_vectorgradient<T, 2>), G, u, g_);
}
template <typename T> void Ops<T>::strainrate(Ctx& c, CT* S, const CT* u) {
  if (c.D == 3) SPEC_LAUNCH((k_strainrate<T, 3>), S, u, g_);
  else SPEC_LAUNCH((k_strainrate<T, 2>), S, u, g_);
}
template <typename T> void Ops<T>::vectordivergence(Ctx& c, CT* d, const CT* u) {
  if (c.D == 3) SPEC_LAUNCH((k_vectordivergence<T, 3>), d, u, g_);
  else SPEC_LAUNCH((k_vectordivergence<T, 2>), d, u, g_);
}
template <typename T> void Ops<T>::gaussian(Ctx& c, CT* u, int ncomp, double delta) {
  SPEC_LAUNCH(k_gaussian<T>, u, delta, g_, ncomp);
}
template <typename T> void Ops<T>::axpy(Ctx& c, CT* y, const CT* x, double a, int ncomp) {
  SPEC_LAUNCH(k_axpy<T>, y, x, T(a), g_, ncomp);
}
template <typename T> void Ops<T>::spec_tracefree(Ctx& c, CT* tau) { SPEC_LAUNCH(k_spec_tracefree<T>, tau, g_); }

// ---------------------------------------------------------------- the RHS
// nonlinearity! (src/solver.jl:434-456) as five passes: inverse z (pruned to
// the 2/3 box), inverse y, fused x C2R/products/R2C, forward y, forward z.
template <typename T>
void Ops<T>::nonlinearity(Ctx& c, CT* sig, const CT* u, unsigned long long* vmax) {
  CT* A = reinterpret_cast<CT*>(c.ensure_spec(c.bufA, c.D));
  const int kc = c.n / 3;
  const int nt = c.tdim();
  if (c.D == 3) {
    c2c_axis<T>(c, 2, +1, u, A, 3, kc + 1, true, kc, true, 1.0);
    c2c_axis<T>(c, 1, +1, A, A, 3, kc + 1, false, kc, false, 1.0);
  } else {
    c2c_axis<T>(c, 1, +1, u, A, 2, kc + 1, false, kc, true, 1.0);
  }
  const double fac = 1.0 / std::pow(double(c.n), c.D);
  x_products<T>(c, sig, A, kc, fac, vmax);
  c2c_axis<T>(c, 1, -1, sig, sig, nt, c.nh, false, c.n, false, 1.0);
  if (c.D == 3) c2c_axis<T>(c, 2, -1, sig, sig, nt, c.nh, false, c.n, false, 1.0);
}

// convectiondiffusion! (src/solver.jl:949-953) plus the optional band forcing
template <typename T>
void Ops<T>::rhs(Ctx& c, CT* du, const CT* u, double nu) {
  CT* B = reinterpret_cast<CT*>(c.ensure_spec(c.bufB, c.tdim()));
  nonlinearity(c, B, u, nullptr);
  const double* fA = band_coeff(c, u);
  if (c.D == 3) SPEC_LAUNCH((k_div_visc<T, 3>), du, B, u, nu, g_, fA, c.kband * c.kband);
  else SPEC_LAUNCH((k_div_visc<T, 2>), du, B, u, nu, g_, fA, c.kband * c.kband);
}

// device coefficient A of the linear band forcing (null when forcing is off)
template <typename T>
const double* Ops<T>::band_coeff(Ctx& c, const CT* u) {
  if (c.force_kind != 1) return nullptr;
  if (!c.forceA) SC_HIP(hipMalloc(&c.forceA, sizeof(double)));
  hipLaunchKernelGGL(k_band_coeff<T>, dim3(1), dim3(256), 0, c.stream, u, c.D, c.geom(), c.kband, c.injection,
                     c.forceA, 0);
  SC_HIP(hipGetLastError());
  return c.forceA;
}

// timing scope of the placement choices below: profiling off (their launches
// are not the caller's), two events; restored / destroyed on every exit
struct TuneScope {
  Ctx& c;
  bool prof;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  explicit TuneScope(Ctx& c_) : c(c_), prof(c_.prof_on) {
    SC_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
      (void)hipEventDestroy(e0);
      throw Error(3, "hipEventCreate failed");
    }
    c.prof_on = false;
  }
  ~TuneScope() {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c.prof_on = prof;
  }
  // ms between the two events once e1 has completed
  float elapsed() {
    SC_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    SC_HIP(hipEventElapsedTime(&ms, e0, e1));
    return ms;
  }
};

// Passes that write a large buffer run at a rate that depends on where its
// pages landed, and an allocation keeps its rate.  On MI355X at 810^3 FP64
// the forward y-pass ran 11.5-11.9 ms per launch into some allocations of the
// z-pencil-tiled buffer and 13.5-14.1 ms into others, whatever its input
// (the inverse z-pass, which writes the same buffer, 2.3 against 2.8 ms),
// and the inverse y-pass 2.9 or 3.1 ms by its output buffer.  So at the
// first RHS whose buffers are large, launch(candidate) is timed (after one
// untimed launch) into the buffer the context has and into up to max_extra
// fresh allocations of the same size (as many as fit with 8 GB to spare),
// and the fastest is kept; the others are freed.  Each launch writes all of
// its candidate, so the kept buffer holds the launch's output, and the
// results are the same whichever is kept.  Returns false, having launched
// nothing, when the buffer is small or no candidate fits.
// The first allocation of a large 3D grid's pipeline buffers (one spectral
// component of 512 MiB or more) goes through Ctx::scattered_allocs; allocs()
// allocates whatever of those buffers does not exist yet.
template <typename T, class F>
static void scatter_first(Ctx& c, F&& allocs) {
  const size_t comp = size_t(c.state_or_spec()) * c.celem();
  if (!SC_SCATTER || c.scattered || c.D != 3 || comp < (size_t(512) << 20)) {
    allocs();
    return;
  }
  c.scattered = true;
  // bufA (3), bufB (6), bufT (6 tiled slots), du, ust, acc (3 each)
  const size_t need = size_t(24) * comp;
  c.scattered_allocs(need, allocs);
}

template <class Launch>
static bool choose_placement(Ctx& c, void*& buf, size_t bytes, int max_extra, Launch launch) {
  if (bytes < (size_t(1) << 30) || !SC_TUNE_TILED) return false;
  size_t freeb = 0, totalb = 0;
  SC_HIP(hipMemGetInfo(&freeb, &totalb));
  const size_t spare = size_t(8) << 30;
  const int extra = int(std::min<size_t>(size_t(max_extra),
                                         freeb > spare ? (freeb - spare) / (bytes + (size_t(5) << 20)) : 0));
  if (extra == 0) return false;
  std::vector<void*> cand{buf};
  size_t ib = 0;
  try {
    for (int i = 0; i < extra; ++i) {
      void* q = nullptr;
      c.dmalloc(&q, bytes);
      cand.push_back(q);
    }
    TuneScope ts(c);
    float best = 0.f;
    for (size_t i = 0; i < cand.size(); ++i) {
      launch(cand[i]);  // first touch
      SC_HIP(hipEventRecord(ts.e0, c.stream));
      launch(cand[i]);
      SC_HIP(hipEventRecord(ts.e1, c.stream));
      const float ms = ts.elapsed();
      if (i == 0 || ms < best) {
        best = ms;
        ib = i;
      }
    }
  } catch (...) {
    for (size_t i = 1; i < cand.size(); ++i) c.dfree(cand[i]);
    throw;
  }
  buf = cand[ib];
  for (size_t i = 0; i < cand.size(); ++i)
    if (i != ib) c.dfree(cand[i]);
  return true;
}

// Right-hand side of the time steppers: the five-pass pipeline with the
// divergence, viscous term and forcing fused into the forward passes.  The
// gradient part -nu k (k.u) of the viscous term is omitted (removed by the
// projection that follows every stage).  With vmax non-null the stage-input
// velocity maximum is accumulated for the CFL step.  lin = false (3D): du is
// the divergence alone; the caller adds -nu |k|^2 u and the band forcing
// (whose coefficient is left in c.forceA).
template <typename T>
void Ops<T>::rhs_fused(Ctx& c, CT* du, const CT* u, double nu, unsigned long long* vmax, bool lin) {
  if (c.D == 3)
    scatter_first<T>(c, [&] {
      c.ensure_spec(c.bufA, c.D);
      c.ensure_spec(c.bufB, c.tdim());
      if (!c.bufT) c.dmalloc(&c.bufT, size_t(6) * tiled_slot<T>(c) * sizeof(CT));
    });
  CT* A = reinterpret_cast<CT*>(c.ensure_spec(c.bufA, c.D));
  CT* B = reinterpret_cast<CT*>(c.ensure_spec(c.bufB, c.tdim()));
  const int kc = c.n / 3;
  StridedGeom gz{};
  if (c.D == 3) {
    // the inverse z-pass writes y-pencil tiles, element (x, y, z) of a
    // component at ((tile * n + z) * n + y) * TX + x % TX (in the buffer the
    // forward passes use later): both its writes and the y-pass's reads then
    // stay within few address-translation pages
    if (!c.bufT) c.dmalloc(&c.bufT, size_t(6) * tiled_slot<T>(c) * sizeof(CT));
    CT* AT = reinterpret_cast<CT*>(c.bufT);
    const long n = c.n, TX = tile_x_rt<T>(c.n), ts = tiled_slot<T>(c), LP = tile_line_rt<T>(c);
    gz = axis_geom<T>(c, 2, kc + 1, true, kc, true, 1.0);
    gz.o_tile = n * LP;
    gz.o_axis = LP;
    gz.o_other = TX;
    gz.o_comp = ts;
    gz.pair = 1;
    c2c_launch<T>(c, gz, 2, +1, u, AT, 3);
    StridedGeom gy = axis_geom<T>(c, 1, kc + 1, false, kc, false, 1.0);
    gy.s_tile = n * LP;
    gy.s_axis = TX;
    gy.s_other = LP;
    gy.s_comp = ts;
    auto inv_y = [&](void* dst) { c2c_launch<T>(c, gy, 1, +1, AT, reinterpret_cast<CT*>(dst), 3); };
    bool done = false;
    if (!c.inv_tuned) {
      c.inv_tuned = true;
      done = choose_placement(c, c.bufA, size_t(c.D) * c.spec_elems * c.celem(), 3, inv_y);
      A = reinterpret_cast<CT*>(c.bufA);  // if chosen, it holds the inverse y-pass's output
    }
    if (!done) inv_y(A);
  } else {
    c2c_axis<T>(c, 1, +1, u, A, 2, kc + 1, false, kc, true, 1.0);
  }
  x_products<T>(c, B, A, kc, 1.0 / std::pow(double(c.n), c.D), vmax);
  RhsParams rp;
  rp.kf = c.geom().kf;
  rp.nu = nu;
  rp.forceA = nullptr;
  rp.kband2 = c.kband * c.kband;
  if (c.force_kind == 1) {
    if (!c.forceA) SC_HIP(hipMalloc(&c.forceA, sizeof(double)));
    c.prof_begin("band_coeff");
    hipLaunchKernelGGL(k_band_coeff<T>, dim3(1), dim3(256), 0, c.stream, u, c.D, c.geom(), c.kband, c.injection,
                       c.forceA, 0);
    c.prof_end();
    SC_HIP(hipGetLastError());
    rp.forceA = c.forceA;
  }
  if (c.D == 3) {
    if (!c.tiled_tuned) {
      c.tiled_tuned = true;
      choose_placement(c, c.bufT, size_t(6) * tiled_slot<T>(c) * sizeof(CT), 5, [&](void* dst) {
        c2c_launch<T>(c, gz, 2, +1, u, reinterpret_cast<CT*>(dst), 3);
        y_fwd_div<T>(c, B, reinterpret_cast<CT*>(dst));
      });
    }
    CT* BT = reinterpret_cast<CT*>(c.bufT);
    y_fwd_div<T>(c, B, BT);
    fwd_rhs_last<T>(c, BT, u, du, rp, lin);
  } else {
    fwd_rhs_last<T>(c, B, u, du, rp, true);
  }
}

template <typename T>
static double* dt_slot(Ctx& c) {
  if (!c.dtdev) SC_HIP(hipMalloc(&c.dtdev, 2 * sizeof(double)));
  return c.dtdev;
}

// CFL time step computed on the device from the maximum collected in vmax
template <typename T>
static void cfl_dt(Ctx& c, double cfl, double nu, double dtmax) {
  hipLaunchKernelGGL(k_cfl_dt, dim3(1), dim3(1), 0, c.stream, c.vmax, cfl, c.l / c.n, c.D, nu, dtmax, dt_slot<T>(c));
  SC_HIP(hipGetLastError());
}

template <typename T>
static void zero_vmax(Ctx& c) {
  if (!c.vmax) SC_HIP(hipMalloc(&c.vmax, sizeof(unsigned long long)));
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, c.stream, c.vmax, 0ull);
  SC_HIP(hipGetLastError());
}

// wray3! (src/solver.jl:968-1012).  cfl > 0: the step is
// min(cfl * propose_timestep(u), dtmax) with the velocity maximum taken from
// the first stage's x-pass (the same C2R of the truncated u).
template <typename T>
double Ops<T>::wray3_dev(Ctx& c, CT* u, double dt, double nu, double cfl, int les_closure, double p1, double p2) {
  CT* us = reinterpret_cast<CT*>(c.ensure_spec(c.ustart, c.D));
  CT* du = reinterpret_cast<CT*>(c.ensure_spec(c.du, c.D));
  double* dtp = dt_slot<T>(c);
  const T a[3] = {T(8.0 / 15.0), T(5.0 / 12.0), T(3.0 / 4.0)};
  const T b[2] = {T(1.0 / 4.0), T(0.0)};
  if (cfl > 0) zero_vmax<T>(c);
  else hipLaunchKernelGGL(k_set_f64, dim3(1), dim3(1), 0, c.stream, dtp, dt);
  for (int i = 0; i < 3; ++i) {
    if (les_closure >= 0) {
      if (cfl > 0 && i == 0) max_velocity_into(c, u);
      les_rhs(c, du, u, nu, les_closure, p1, p2);
    } else {
      rhs_fused(c, du, u, nu, (cfl > 0 && i == 0) ? c.vmax : nullptr);
    }
    if (cfl > 0 && i == 0) cfl_dt<T>(c, cfl, nu, dt);
    const CT* src = i == 0 ? u : us;
    const T bi = i < 2 ? b[i] : T(0);
    int wr = (i < 2 && b[i] != T(0)) ? 1 : 0;
    if (c.D == 3) SPEC_LAUNCH((k_wray3_update<T, 3>), u, us, src, du, (const double*)dtp, a[i], bi, wr, g_);
    else SPEC_LAUNCH((k_wray3_update<T, 2>), u, us, src, du, (const double*)dtp, a[i], bi, wr, g_);
  }
  return read_dt(c);
}

// classic RK4 with a projection of every stage value
// The RK4 update's rate likewise depends on the pages of the two buffers it
// both reads and writes, the stage input ust and the accumulator acc (810^3
// FP64: 11.5-11.8 ms per launch on average with some allocations, 12.5-13.0
// ms with others; replacing either buffer alone by a fresh allocation was
// enough, the RHS buffer makes no difference).  At the first large 3D RK4
// step a stage-1 update (which writes only ust and acc, both rewritten by
// stage 0 before they are read) is timed for every pairing of the context's
// ust or one of up to three fresh allocations with its acc or one of up to
// three fresh ones, and the fastest pairing is kept.
template <typename T>
static void tune_state(Ctx& c, const C<T>* u, const C<T>* du) {
  c.state_tuned = true;
  const size_t bytes = size_t(c.D) * c.state_or_spec() * c.celem();
  if (bytes < (size_t(1) << 30) || !SC_TUNE_TILED) return;
  size_t freeb = 0, totalb = 0;
  SC_HIP(hipMemGetInfo(&freeb, &totalb));
  const size_t spare = size_t(8) << 30, one = bytes + (size_t(5) << 20);
  const int extra = int(std::min<size_t>(3, freeb > spare ? (freeb - spare) / (2 * one) : 0));
  if (extra == 0) return;
  std::vector<void*> us{c.ust}, ac{c.acc};
  size_t bi = 0, bj = 0;
  try {
    for (int i = 0; i < extra; ++i) {
      void *a = nullptr, *b = nullptr;
      c.dmalloc(&a, bytes);
      us.push_back(a);
      c.dmalloc(&b, bytes);
      ac.push_back(b);
    }
    TuneScope ts(c);
    const SpecGeom g = c.geom();
    const double* dtp = c.dtdev;  // its value does not matter: the outputs are overwritten
    float best = 0.f;
    bool first = true;
    for (size_t i = 0; i < us.size(); ++i)
      for (size_t j = 0; j < ac.size(); ++j) {
        C<T>* s_ = reinterpret_cast<C<T>*>(us[i]);
        C<T>* a_ = reinterpret_cast<C<T>*>(ac[j]);
        for (int rep = 0; rep < 2; ++rep) {  // first touch, then timed
          if (rep == 1) SC_HIP(hipEventRecord(ts.e0, c.stream));
          hipLaunchKernelGGL((k_rk4_update_lin<T, 3>), dim3(nblk(g.total)), dim3(256), 0, c.stream,
                             const_cast<C<T>*>(u), s_, a_, du, dtp, 0.5, 1, g, 0.0, (const double*)nullptr, 0);
          SC_HIP(hipGetLastError());
        }
        SC_HIP(hipEventRecord(ts.e1, c.stream));
        const float ms = ts.elapsed();
        if (first || ms < best) {
          best = ms;
          bi = i;
          bj = j;
          first = false;
        }
      }
  } catch (...) {
    for (size_t i = 1; i < us.size(); ++i) c.dfree(us[i]);
    for (size_t j = 1; j < ac.size(); ++j) c.dfree(ac[j]);
    throw;
  }
  c.ust = us[bi];
  c.acc = ac[bj];
  for (size_t i = 0; i < us.size(); ++i)
    if (i != bi) c.dfree(us[i]);
  for (size_t j = 0; j < ac.size(); ++j)
    if (j != bj) c.dfree(ac[j]);
}

template <typename T>
double Ops<T>::rk4_dev(Ctx& c, CT* u, double dt, double nu, double cfl, int les_closure, double p1, double p2) {
  c.ensure_spec(c.ust, c.D);
  c.ensure_spec(c.acc, c.D);
  CT* du = reinterpret_cast<CT*>(c.ensure_spec(c.du, c.D));
  double* dtp = dt_slot<T>(c);
  if (c.D == 3 && les_closure < 0 && !c.state_tuned) tune_state<T>(c, u, du);
  CT* ust = reinterpret_cast<CT*>(c.ust);
  CT* acc = reinterpret_cast<CT*>(c.acc);
  const double frac[4] = {0.5, 0.5, 1.0, 1.0 / 6.0};
  if (cfl > 0) zero_vmax<T>(c);
  else hipLaunchKernelGGL(k_set_f64, dim3(1), dim3(1), 0, c.stream, dtp, dt);
  for (int s = 0; s < 4; ++s) {
    if (les_closure >= 0) {
      if (cfl > 0 && s == 0) max_velocity_into(c, u);
      les_rhs(c, du, s == 0 ? u : ust, nu, les_closure, p1, p2);
    } else {
      // 3D: the viscous and forcing terms join k in the contiguous update pass
      rhs_fused(c, du, s == 0 ? u : ust, nu, (cfl > 0 && s == 0) ? c.vmax : nullptr, c.D != 3);
    }
    if (cfl > 0 && s == 0) cfl_dt<T>(c, cfl, nu, dt);
    if (c.D == 3 && les_closure < 0)
      SPEC_LAUNCH((k_rk4_update_lin<T, 3>), u, ust, acc, du, (const double*)dtp, frac[s], s, g_, nu,
                  (const double*)(c.force_kind == 1 ? c.forceA : nullptr), c.kband * c.kband);
    else if (c.D == 3) SPEC_LAUNCH((k_rk4_update<T, 3>), u, ust, acc, du, (const double*)dtp, frac[s], s, g_);
    else SPEC_LAUNCH((k_rk4_update<T, 2>), u, ust, acc, du, (const double*)dtp, frac[s], s, g_);
  }
  return read_dt(c);
}

// nsteps RK4 steps with the Runge-Kutta state held in z-pencil tiles (3D):
// u is converted to the tiled layout once on entry and back on exit, and each
// stage runs the inverse z-pass (reading the tiled stage input), the inverse
// y-pass, the x-pass, the forward y-pass and k_z_fwd_rk4, the last forward
// pass fused with the viscous and forcing terms, the RK4 accumulation and the
// projection.  The stage arithmetic is that of rk4_dev.  cfl > 0: every step
// takes min(cfl * propose_timestep, dtmax) from its first stage's velocity
// maximum; otherwise dt = dtmax.  dts (host, nullable) receives the steps.
template <typename T>
static void ztile_convert(Ctx& c, const C<T>* src, C<T>* dst, int ncomp, bool to_tiles) {
  const int n = c.n, TX = tile_x_rt<T>(n);
  const int ntile = (c.nh + TX - 1) / TX;
  const unsigned nblocks = unsigned(ntile) * unsigned((n + 7) / 8) * unsigned((n + 31) / 32);
  const long ts = tiled_slot<T>(c);
  c.prof_begin(to_tiles ? "ztile_in" : "ztile_out");
#define SC_CONV(TXV)                                                                                          \
  if (TX == TXV) {                                                                                            \
    if (to_tiles)                                                                                             \
      hipLaunchKernelGGL((k_ztile_convert<T, TXV, true>), dim3(nblocks, unsigned(ncomp)), dim3(256), 0, c.stream, \
                         src, dst, n, c.nhp, c.spec_elems, ts, tile_line_rt<T>(c));                          \
    else                                                                                                      \
      hipLaunchKernelGGL((k_ztile_convert<T, TXV, false>), dim3(nblocks, unsigned(ncomp)), dim3(256), 0,       \
                         c.stream, src, dst, n, c.nhp, c.spec_elems, ts, tile_line_rt<T>(c));                \
  }
  SC_CONV(2) SC_CONV(4) SC_CONV(8) SC_CONV(16)
#undef SC_CONV
  c.prof_end();
  SC_HIP(hipGetLastError());
}

template <typename T>
static void z_fwd_rk4(Ctx& c, const C<T>* BT, C<T>* U, C<T>* US, C<T>* AC, RhsParams rp, const double* dtp,
                      double frac, int stage) {
  StridedGeom g = full_geom<T>(c, 2);
  dim3 grid(unsigned(g.ntile) * unsigned(g.nother));
  const C<T>* tw = reinterpret_cast<const C<T>*>(c.tw_n);
  c.prof_begin("z_fwd_rk4");
  switch (c.n) {
#define X(NN)                                                                                                  \
  case NN:                                                                                                     \
    if constexpr (z_rk4_fits<T, NN>())                                                                         \
      hipLaunchKernelGGL((k_z_fwd_rk4<T, NN>), grid, dim3(wide_threads<T, NN>()), 0, c.stream, BT,             \
                         tiled_slot<T>(c), U, US, AC, tw, g, rp, dtp, frac, stage);                             \
    else                                                                                                       \
      throw Error(2, "no fused RK4 stage at this size");                                                       \
    break;
    SC_FOR_N(X)
#undef X
    default: throw Error(2, "unsupported grid size");
  }
  c.prof_end();
  SC_HIP(hipGetLastError());
}

template <typename T>
static bool z_rk4_fits_rt(int n) {
  switch (n) {
#define X(NN) case NN: return z_rk4_fits<T, NN>();
    SC_FOR_N(X)
#undef X
  }
  return false;
}

template <typename T>
void Ops<T>::rk4_steps(Ctx& c, CT* u, int nsteps, double cfl, double nu, double dtmax, double* dts) {
  if (c.D != 3 || !z_rk4_fits_rt<T>(c.n)) {
    for (int i = 0; i < nsteps; ++i) {
      const double d = rk4_dev(c, u, dtmax, nu, cfl);
      if (dts) dts[i] = d;
    }
    return;
  }
  const int n = c.n, kc = n / 3;
  const long TX = tile_x_rt<T>(n), ts = tiled_slot<T>(c), LP = tile_line_rt<T>(c);
  auto allocs = [&] {
    c.ensure_spec(c.bufA, 3);
    c.ensure_spec(c.bufB, 6);
    if (!c.bufT) c.dmalloc(&c.bufT, size_t(6) * ts * sizeof(CT));
    c.ensure_spec(c.du, 3);
    c.ensure_spec(c.ust, 3);
    c.ensure_spec(c.acc, 3);
  };
  scatter_first<T>(c, allocs);
  CT* A = reinterpret_cast<CT*>(c.bufA);
  CT* B = reinterpret_cast<CT*>(c.bufB);
  CT* U = reinterpret_cast<CT*>(c.du);   // u0, tiled
  CT* US = reinterpret_cast<CT*>(c.ust); // stage value, tiled
  CT* AC = reinterpret_cast<CT*>(c.acc); // accumulator, tiled
  double* dtp = dt_slot<T>(c);
  // the placement choice of the pipeline's buffers happens at the first RHS
  // (its output goes to the stage buffer, which stage 0 overwrites)
  if (!c.inv_tuned || !c.tiled_tuned) {
    rhs_fused(c, US, u, nu, nullptr, false);
    A = reinterpret_cast<CT*>(c.bufA);
  }
  if (c.dtlog_n < nsteps) {
    if (c.dtlog) SC_HIP(hipFree(c.dtlog));
    c.dtlog = nullptr;
    SC_HIP(hipMalloc(&c.dtlog, size_t(nsteps) * sizeof(double)));
    c.dtlog_n = nsteps;
  }
  double* dlog = c.dtlog;
  ztile_convert<T>(c, u, U, 3, true);
  StridedGeom gz = axis_geom<T>(c, 2, kc + 1, true, kc, true, 1.0);
  gz.s_tile = n * LP;
  gz.s_axis = TX;
  gz.s_other = LP;
  gz.s_comp = ts;
  gz.o_tile = n * LP;
  gz.o_axis = LP;
  gz.o_other = TX;
  gz.o_comp = ts;
  gz.pair = 1;
  StridedGeom gy = axis_geom<T>(c, 1, kc + 1, false, kc, false, 1.0);
  gy.s_tile = n * LP;
  gy.s_axis = TX;
  gy.s_other = LP;
  gy.s_comp = ts;
  SpecGeom gt = c.geom();
  gt.s_comp = ts;
  RhsParams rp;
  rp.kf = c.geom().kf;
  rp.nu = nu;
  rp.forceA = nullptr;
  rp.kband2 = c.kband * c.kband;
  if (c.force_kind == 1) {
    if (!c.forceA) SC_HIP(hipMalloc(&c.forceA, sizeof(double)));
    rp.forceA = c.forceA;
  }
  CT* AT = reinterpret_cast<CT*>(c.bufT);
  const double frac[4] = {0.5, 0.5, 1.0, 1.0 / 6.0};
  const double fac = 1.0 / std::pow(double(n), 3);
  for (int step = 0; step < nsteps; ++step) {
    if (cfl > 0) zero_vmax<T>(c);
    else hipLaunchKernelGGL(k_set_f64, dim3(1), dim3(1), 0, c.stream, dtp, dtmax);
    for (int s = 0; s < 4; ++s) {
      const CT* sin = s == 0 ? U : US;
      c2c_launch<T>(c, gz, 2, +1, sin, AT, 3);
      c2c_launch<T>(c, gy, 1, +1, AT, A, 3);
      x_products<T>(c, B, A, kc, fac, (cfl > 0 && s == 0) ? c.vmax : nullptr);
      if (cfl > 0 && s == 0) cfl_dt<T>(c, cfl, nu, dtmax);
      if (c.force_kind == 1) {
        c.prof_begin("band_coeff");
        hipLaunchKernelGGL(k_band_coeff<T>, dim3(1), dim3(256), 0, c.stream, sin, 3, gt, c.kband, c.injection,
                           c.forceA, int(TX), LP);
        c.prof_end();
        SC_HIP(hipGetLastError());
      }
      y_fwd_div<T>(c, B, AT);
      z_fwd_rk4<T>(c, AT, U, US, AC, rp, dtp, frac[s], s);
    }
    hipLaunchKernelGGL(k_copy_f64, dim3(1), dim3(1), 0, c.stream, dlog + step, (const double*)dtp);
    SC_HIP(hipGetLastError());
  }
  ztile_convert<T>(c, U, u, 3, false);
  if (dts && nsteps > 0) {
    double* h = c.hostbuf(nsteps);
    SC_HIP(hipMemcpyAsync(h, dlog, size_t(nsteps) * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SC_HIP(hipStreamSynchronize(c.stream));
    std::memcpy(dts, h, size_t(nsteps) * sizeof(double));
  }
}

template <typename T>
double Ops<T>::read_dt(Ctx& c) {
  double* h = c.hostbuf(1);
  SC_HIP(hipMemcpyAsync(h, c.dtdev, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  SC_HIP(hipStreamSynchronize(c.stream));
  return h[0];
}

template <typename T> void Ops<T>::wray3(Ctx& c, CT* u, double dt, double nu) { wray3_dev(c, u, dt, nu, 0.0, -1, 0, 0); }
template <typename T> void Ops<T>::rk4(Ctx& c, CT* u, double dt, double nu) { rk4_dev(c, u, dt, nu, 0.0); }

// velocity maximum of C2R(2/3 u) accumulated into c.vmax (no reset)
template <typename T>
void Ops<T>::max_velocity_into(Ctx& c, const CT* u) {
  CT* A = reinterpret_cast<CT*>(c.ensure_spec(c.bufA, c.D));
  CT* B = reinterpret_cast<CT*>(c.ensure_spec(c.bufB, c.tdim()));
  const int kc = c.n / 3;
  if (c.D == 3) {
    c2c_axis<T>(c, 2, +1, u, A, 3, kc + 1, true, kc, true, 1.0);
    c2c_axis<T>(c, 1, +1, A, A, 3, kc + 1, false, kc, false, 1.0);
  } else {
    c2c_axis<T>(c, 1, +1, u, A, 2, kc + 1, false, kc, true, 1.0);
  }
  x_products<T>(c, B, A, kc, 1.0, c.vmax);
}

// max |C2R(2/3 u)| (propose_timestep, src/solver.jl:931-947)
template <typename T>
double Ops<T>::max_velocity(Ctx& c, const CT* u) {
  zero_vmax<T>(c);
  max_velocity_into(c, u);
  unsigned long long bits = 0;
  SC_HIP(hipMemcpyAsync(&bits, c.vmax, sizeof(bits), hipMemcpyDeviceToHost, c.stream));
  SC_HIP(hipStreamSynchronize(c.stream));
  double v;
  std::memcpy(&v, &bits, sizeof(v));
  return v;
}

// ------------------------------------------------------------- reductions
template <typename T>
double Ops<T>::sum_modes(Ctx& c, const CT* u, int ncomp, int kind, double nu) {
  SpecGeom g = c.geom();
  unsigned nb = nblk(g.total);
  double* part = c.dscratch(long(nb) + 1);
  hipLaunchKernelGGL(k_sum_modes<T>, dim3(nb), dim3(RB), 0, c.stream, u, ncomp, g, kind, nu, part, 0L);
  SC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(1024), 0, c.stream, part, long(nb), part + nb);
  double* h = c.hostbuf(1);
  SC_HIP(hipMemcpyAsync(h, part + nb, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  SC_HIP(hipStreamSynchronize(c.stream));
  return h[0];
}

// sum_modes through the caller's scratch of NOALLOC_DOUBLES doubles, with no
// allocation: k_sum_modes runs in chunks of at most NOALLOC_PARTIALS blocks,
// each chunk's partials are summed into its slot before the next chunk
// reuses them, then the slots are summed; the total goes to the host double
// *out.  Every scratch entry read is written earlier in the same call.
template <typename T>
void Ops<T>::sum_modes_noalloc(Ctx& c, const CT* u, int ncomp, int kind, double nu, double* scr, double* out) {
  const SpecGeom g = c.geom();
  const long nb = (g.total + RB - 1) / RB;
  const long nch = (nb + NOALLOC_PARTIALS - 1) / NOALLOC_PARTIALS;
  if (nch > NOALLOC_CHUNKS)
    throw Error(1, "field too large for the no-allocation reduction (at most " +
                       std::to_string(NOALLOC_PARTIALS * NOALLOC_CHUNKS * RB) + " entries per component)");
  double* sums = scr + NOALLOC_PARTIALS;
  double* total = sums + NOALLOC_CHUNKS;
  for (long ch = 0; ch < nch; ++ch) {
    const long b0 = ch * NOALLOC_PARTIALS;
    const long blocks = std::min(NOALLOC_PARTIALS, nb - b0);
    hipLaunchKernelGGL(k_sum_modes<T>, dim3(unsigned(blocks)), dim3(RB), 0, c.stream, u, ncomp, g, kind, nu, scr,
                       b0 * RB);
    hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(1024), 0, c.stream, (const double*)scr, blocks, sums + ch);
  }
  hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(1024), 0, c.stream, (const double*)sums, nch, total);
  SC_HIP(hipGetLastError());
  SC_HIP(hipMemcpyAsync(out, total, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  SC_HIP(hipStreamSynchronize(c.stream));
}

template <typename T>
void Ops<T>::dot_modes(Ctx& c, const CT* u, const CT* v, int ncomp, double* re, double* im) {
  SpecGeom g = c.geom();
  unsigned nb = nblk(g.total);
  double* part = c.dscratch(2 * long(nb) + 2);
  hipLaunchKernelGGL(k_dot_modes<T>, dim3(nb), dim3(RB), 0, c.stream, u, v, ncomp, g, part);
  SC_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(1024), 0, c.stream, part, long(nb), part + 2 * nb);
  hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(1024), 0, c.stream, part + nb, long(nb), part + 2 * nb + 1);
  double* h = c.hostbuf(2);
  SC_HIP(hipMemcpyAsync(h, part + 2 * nb, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  SC_HIP(hipStreamSynchronize(c.stream));
  *re = h[0];
  *im = h[1];
}

template <typename T>
void Ops<T>::shell_energy(Ctx& c, const CT* u, int ncomp, int nshell, double* host_out) {
  SpecGeom g = c.geom();
  double* d = c.dscratch(nshell);
  SC_HIP(hipMemsetAsync(d, 0, nshell * sizeof(double), c.stream));
  hipLaunchKernelGGL(k_shell_energy<T>, dim3(nblk(g.total)), dim3(256), nshell * sizeof(double), c.stream, u, ncomp,
                     g, nshell, d);
  SC_HIP(hipGetLastError());
  double* h = c.hostbuf(nshell);
  SC_HIP(hipMemcpyAsync(h, d, nshell * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  SC_HIP(hipStreamSynchronize(c.stream));
  std::memcpy(host_out, h, nshell * sizeof(double));
}

// solve_les's shell-resolved transfer (src/les.jl:125-139) of the physical
// packed stress tau against the spectral u: shells 0..nshell-1 into host_out
template <typename T>
void Ops<T>::shell_transfer(Ctx& c, const CT* u, const T* tau, int nshell, double* host_out) {
  const int nt = c.tdim();
  CT* th = reinterpret_cast<CT*>(c.work(2, size_t(nt) * c.spec_elems * sizeof(CT)));
  r2c(c, th, tau, nt, 1.0 / std::pow(double(c.n), c.D));
  twothirds(c, th, nt);
  SpecGeom g = c.geom();
  double* d = c.dscratch(nshell);
  SC_HIP(hipMemsetAsync(d, 0, nshell * sizeof(double), c.stream));
  if (c.D == 3)
    hipLaunchKernelGGL((k_shell_transfer<T, 3>), dim3(nblk(g.total)), dim3(256), nshell * sizeof(double), c.stream, u,
                       th, g, nshell, d);
  else
    hipLaunchKernelGGL((k_shell_transfer<T, 2>), dim3(nblk(g.total)), dim3(256), nshell * sizeof(double), c.stream, u,
                       th, g, nshell, d);
  SC_HIP(hipGetLastError());
  double* h = c.hostbuf(nshell);
  SC_HIP(hipMemcpyAsync(h, d, nshell * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  SC_HIP(hipStreamSynchronize(c.stream));
  std::memcpy(host_out, h, nshell * sizeof(double));
}

template <typename T>
void Ops<T>::sfs_dissipation(Ctx& c, T* eps, const T* tau, const T* GG) {
  if (c.D == 3) hipLaunchKernelGGL((k_sfs_dissipation<T, 3>), dim3(nblk(c.phys_elems)), dim3(256), 0, c.stream, eps, tau, GG, c.phys_elems);
  else hipLaunchKernelGGL((k_sfs_dissipation<T, 2>), dim3(nblk(c.phys_elems)), dim3(256), 0, c.stream, eps, tau, GG, c.phys_elems);
  SC_HIP(hipGetLastError());
}

// FP64 sums of k_sample_stats (nv values) into host out
template <typename T>
static void sample_sums(Ctx& c, const T* x, const T* y, long n, int kind, double ma, double mb, int nv, double* out) {
  const unsigned nb = unsigned(std::min<long>(1024, std::max<long>(1, (n + RB - 1) / RB)));
  double* part = c.dscratch(long(nv) * (nb + 1));
  hipLaunchKernelGGL(k_sample_stats<T>, dim3(nb), dim3(RB), 0, c.stream, x, y, n, kind, ma, mb, part);
  SC_HIP(hipGetLastError());
  for (int q = 0; q < nv; ++q)
    hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(1024), 0, c.stream, part + long(q) * nb, long(nb),
                       part + long(nv) * nb + q);
  SC_HIP(hipGetLastError());
  double* h = c.hostbuf(nv);
  SC_HIP(hipMemcpyAsync(h, part + long(nv) * nb, nv * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  SC_HIP(hipStreamSynchronize(c.stream));
  std::memcpy(out, h, nv * sizeof(double));
}

template <typename T>
double Ops<T>::sample_mean(Ctx& c, const T* x, long n) {
  double sum = 0;
  sample_sums<T>(c, x, nullptr, n, 0, 0, 0, 1, &sum);
  return sum / double(n);
}

// moments (src/analysis.jl:9-18) and the backscatter fraction
// (:132) of the device samples x[0, n): mean, median, std (n - 1), skewness
// (third standardised moment, 0 when std is 0), fraction of x < 0.  The
// median's two middle order statistics come from a device radix selection
// (k_radix_hist / k_radix_pick: 8-bit digits of order-preserving keys).
template <typename T>
void Ops<T>::sample_moments(Ctx& c, const T* x, long n, double* out) {
  if (n < 2) throw Error(1, "moments need at least two samples");
  double sum = 0, cm[3];
  sample_sums<T>(c, x, nullptr, n, 0, 0, 0, 1, &sum);
  const double m = sum / double(n);
  sample_sums<T>(c, x, nullptr, n, 1, m, 0, 3, cm);
  const double sd = std::sqrt(cm[0] / double(n - 1));
  auto* sel = reinterpret_cast<RadixSel*>(c.work(7, 2 * sizeof(RadixSel) + 512 * sizeof(unsigned long long)));
  auto* hist = reinterpret_cast<unsigned long long*>(sel + 2);
  hipLaunchKernelGGL(k_radix_init, dim3(1), dim3(256), 0, c.stream, sel, (long long)((n - 1) / 2),
                     (long long)(n / 2), hist);
  const unsigned nb = unsigned(std::min<long>(1024, (n + 255) / 256));
  for (int shift = 8 * int(sizeof(T)) - 8; shift >= 0; shift -= 8) {
    hipLaunchKernelGGL(k_radix_hist<T>, dim3(nb), dim3(256), 0, c.stream, x, n, sel, shift, hist);
    hipLaunchKernelGGL(k_radix_pick, dim3(1), dim3(256), 0, c.stream, sel, shift, hist);
  }
  SC_HIP(hipGetLastError());
  static_assert(2 * sizeof(RadixSel) <= 6 * sizeof(double), "host staging");
  RadixSel* hs = reinterpret_cast<RadixSel*>(c.hostbuf(6));
  SC_HIP(hipMemcpyAsync(hs, sel, 2 * sizeof(RadixSel), hipMemcpyDeviceToHost, c.stream));
  SC_HIP(hipStreamSynchronize(c.stream));
  // the keys back to values (the inverse of okey)
  auto value = [](unsigned long long key) -> double {
    if constexpr (sizeof(T) == 8) {
      const unsigned long long b = (key >> 63) ? (key & ~(1ull << 63)) : ~key;
      double v;
      std::memcpy(&v, &b, sizeof v);
      return v;
    } else {
      const unsigned k32 = unsigned(key);
      const unsigned b = (k32 >> 31) ? (k32 & 0x7fffffffu) : ~k32;
      float v;
      std::memcpy(&v, &b, sizeof v);
      return double(v);
    }
  };
  const double mid0 = value(hs[0].prefix), mid1 = value(hs[1].prefix);
  out[0] = m;
  out[1] = n % 2 ? mid0 : (mid0 + mid1) / 2;
  out[2] = sd;
  out[3] = sd == 0.0 ? 0.0 : cm[1] / double(n) / (sd * sd * sd);
  out[4] = cm[2] / double(n);
}

// relerr and crosscor of compute_sfs_stats (src/analysis.jl:116-121) over
// the n stacked entries of model (b) and reference (a)
template <typename T>
void Ops<T>::apriori(Ctx& c, const T* model, const T* ref, long n, double* out) {
  if (n < 1) throw Error(1, "empty tensors");
  double s1[6], s2[6];
  sample_sums<T>(c, ref, model, n, 2, 0, 0, 6, s1);
  const double ma = s1[0] / double(n), mb = s1[1] / double(n);
  sample_sums<T>(c, ref, model, n, 2, ma, mb, 6, s2);
  out[0] = std::sqrt(s2[5]) / std::sqrt(s1[2]);
  out[1] = s2[4] / std::sqrt(s2[3] * s2[2]);
}

template <typename T>
void Ops<T>::shell_scale(Ctx& c, CT* u, int ncomp, int nshell, const double* host_factor) {
  double* d = c.dscratch(nshell);
  double* h = c.hostbuf(nshell);
  SC_HIP(hipStreamSynchronize(c.stream));
  std::memcpy(h, host_factor, nshell * sizeof(double));
  SC_HIP(hipMemcpyAsync(d, h, nshell * sizeof(double), hipMemcpyHostToDevice, c.stream));
  SPEC_LAUNCH(k_shell_scale<T>, u, ncomp, g_, nshell, (const double*)d);
  SC_HIP(hipStreamSynchronize(c.stream));
}

// --------------------------------------------------------- initial fields
template <typename T> void Ops<T>::randn(Ctx& c, CT* u, int ncomp, unsigned long long seed) {
  SPEC_LAUNCH(k_randn<T>, u, ncomp, g_, seed);
}
template <typename T> void Ops<T>::symmetrize(Ctx& c, CT* u, int ncomp) {
  long tot = 2L * c.n * c.nz;
  hipLaunchKernelGGL(k_symmetrize<T>, dim3(nblk(tot)), dim3(256), 0, c.stream, u, ncomp, c.geom());
  SC_HIP(hipGetLastError());
}
template <typename T> void Ops<T>::taylorgreen(Ctx& c, CT* u, double V0, int doproject) {
  T* v = reinterpret_cast<T*>(c.scratch(1));  // one spectral scalar holds one physical scalar
  for (int comp = 0; comp < c.D; ++comp) {
    hipLaunchKernelGGL(k_taylorgreen_phys<T>, dim3(nblk(c.phys_elems)), dim3(256), 0, c.stream, v, comp, c.D, c.n,
                       c.l, V0);
    SC_HIP(hipGetLastError());
    r2c(c, u + comp * c.spec_elems, v, 1, 1.0 / std::pow(double(c.n), c.D));
  }
  if (doproject) project(c, u);
}

// ------------------------------------------------------------- filtering
template <typename T> void Ops<T>::cutoff(Ctx& cb, Ctx& cf, CT* ub, const CT* u, int ncomp) {
  SpecGeom gb = cb.geom(), gf = cf.geom();
  hipLaunchKernelGGL(k_cutoff<T>, dim3(nblk(gb.total)), dim3(256), 0, cb.stream, ub, u, gb, gf, ncomp);
  SC_HIP(hipGetLastError());
}
template <typename T> void Ops<T>::inverse_cutoff(Ctx& cb, Ctx& cf, CT* u, const CT* ub, int ncomp) {
  SpecGeom gb = cb.geom(), gf = cf.geom();
  hipLaunchKernelGGL(k_inverse_cutoff<T>, dim3(nblk(gb.total)), dim3(256), 0, cb.stream, u, ub, gb, gf, ncomp);
  SC_HIP(hipGetLastError());
}

// -------------------------------------------------------------- closures
template <typename T> void Ops<T>::clark(Ctx& c, T* tau, const T* GG, double delta) {
  if (c.D == 3) hipLaunchKernelGGL((k_clark<T, 3>), dim3(nblk(c.phys_elems)), dim3(256), 0, c.stream, tau, GG, delta, c.phys_elems);
  else hipLaunchKernelGGL((k_clark<T, 2>), dim3(nblk(c.phys_elems)), dim3(256), 0, c.stream, tau, GG, delta, c.phys_elems);
  SC_HIP(hipGetLastError());
}
template <typename T> void Ops<T>::smagorinsky(Ctx& c, T* tau, const T* GG, double cs, double delta) {
  if (c.D == 3) hipLaunchKernelGGL((k_smagorinsky<T, 3>), dim3(nblk(c.phys_elems)), dim3(256), 0, c.stream, tau, GG, cs, delta, c.phys_elems);
  else hipLaunchKernelGGL((k_smagorinsky<T, 2>), dim3(nblk(c.phys_elems)), dim3(256), 0, c.stream, tau, GG, cs, delta, c.phys_elems);
  SC_HIP(hipGetLastError());
}
template <typename T> void Ops<T>::verstappen(Ctx& c, T* tau, const T* GG, double cq, double delta) {
  if (c.D != 3) throw Error(1, "the Q-R model is only defined in 3D");
  hipLaunchKernelGGL(k_verstappen<T>, dim3(nblk(c.phys_elems)), dim3(256), 0, c.stream, tau, GG, cq, delta, c.phys_elems);
  SC_HIP(hipGetLastError());
}
template <typename T> void Ops<T>::phys_tracefree(Ctx& c, T* tau) {
  hipLaunchKernelGGL(k_phys_tracefree<T>, dim3(nblk(c.phys_elems)), dim3(256), 0, c.stream, tau, c.D, c.phys_elems);
  SC_HIP(hipGetLastError());
}

// physical velocity gradient GG (D*D comps): vectorgradient! + 2/3 + C2R
// (src/les.jl:11-16)
template <typename T> void Ops<T>::gradient_phys(Ctx& c, T* GG, const CT* u) {
  const int ng = c.D * c.D;
  CT* G = reinterpret_cast<CT*>(c.scratch(2 * ng)) + long(ng) * c.spec_elems;
  vectorgradient(c, G, u);
  // the twothirds mask is applied while reading the spectral gradient
  c2r(c, GG, G, ng, 1.0, true);
}

template <typename T>
static void dynsmag_add_spec(Ctx& c, C<T>* sig, const C<T>* u, double delta);

// les! (src/les.jl:4-33) for the built-in closures:
// closure 1 = Clark(Delta=p1), 2 = Smagorinsky(C_S=p1, Delta=p2),
// 3 = dynamic Smagorinsky(Delta=p1), 4 = the attached learned closure,
// 5 = Verstappen Q-R(C=p1, Delta=p2), 6 = Bardina(s=p1, Delta=p2), 7 = the
// attached pointwise MLP, 0 = none.
// The band forcing is added when it is enabled on the context.
template <typename T>
void Ops<T>::les_rhs(Ctx& c, CT* du, const CT* u, double nu, int closure, double p1, double p2) {
  const int nt = c.tdim();
  const int ng = c.D * c.D;
  if (closure < 0 || closure > 7) throw Error(1, "unknown closure kind");
  CT* B = reinterpret_cast<CT*>(c.ensure_spec(c.bufB, nt));
  nonlinearity(c, B, u, nullptr);
  if (closure == 3) {
    dynsmag_add_spec<T>(c, B, u, p1);
  } else if (closure != 0) {
    T* tau = reinterpret_cast<T*>(c.work(0, size_t(nt) * c.phys_elems * sizeof(T)));
    if (closure == 4) {
      cnn_tau(c, tau, u);
    } else if (closure == 6) {
      bardina(c, tau, u, p1, p2);
    } else if (closure == 7) {
      if (!c.pmlp) throw Error(1, "closure 7 needs a network attached with sc_set_pmlp_closure");
      T* GG = reinterpret_cast<T*>(c.work(1, size_t(ng) * c.phys_elems * sizeof(T)));
      gradient_phys(c, GG, u);
      pmlplink::closure_tau<T>(c.pmlp, GG, c.phys_elems, c.pmlp_delta, c.pmlp_visc, c.pmlp_mu, c.pmlp_sigma, c.pmlp_sym, tau,
                               c.stream);
    } else {
      T* GG = reinterpret_cast<T*>(c.work(1, size_t(ng) * c.phys_elems * sizeof(T)));
      gradient_phys(c, GG, u);
      if (closure == 1) clark(c, tau, GG, p1);
      else if (closure == 2) smagorinsky(c, tau, GG, p1, p2);
      else verstappen(c, tau, GG, p1, p2);
    }
    // sigma += to_spec!(tau) per component (the closure stress is not truncated)
    CT* th = reinterpret_cast<CT*>(c.work(2, size_t(nt) * c.spec_elems * sizeof(CT)));
    r2c(c, th, tau, nt, 1.0 / std::pow(double(c.n), c.D));
    axpy(c, B, th, 1.0, nt);
  }
  // viscosity! + tensordivergence! (+ forcing) in one pass
  const double* fA = band_coeff(c, u);
  if (c.D == 3) SPEC_LAUNCH((k_div_visc<T, 3>), du, B, u, nu, g_, fA, c.kband * c.kband);
  else SPEC_LAUNCH((k_div_visc<T, 2>), du, B, u, nu, g_, fA, c.kband * c.kband);
}

namespace {
// runs the network on the solver's stream for the lifetime of the object
struct OnStream {
  ::sc_cnn* net;
  hipStream_t old;
  OnStream(::sc_cnn* n, hipStream_t s) : net(n), old(cnnlink::set_stream(n, s)) {}
  ~OnStream() { cnnlink::set_stream(net, old); }
};
}  // namespace

// learned closure: tau = s_out * net(C2R(2/3 u) / s_in, extra), the six packed
// components xx, yy, zz, xy, yz, zx in physical space
template <typename T>
void Ops<T>::cnn_tau(Ctx& c, T* tau, const CT* u) {
  if (!c.cnn) throw Error(1, "closure 4 needs a network attached with sc_set_cnn_closure");
  T* v = reinterpret_cast<T*>(c.work(1, size_t(c.D) * c.phys_elems * sizeof(T)));
  c2r(c, v, u, c.D, 1.0, true);
  OnStream on(c.cnn, c.stream);
  const cnnlink::Bufs b = cnnlink::buffers(c.cnn);
  void* x = cnnlink::slot_x(c.cnn, c.cnn_slot);
  cnnlink::pack_velocity<T>(c.cnn, v, c.phys_elems, 1.0 / c.cnn_sin, c.cnn_extra, x);
  cnnlink::forward_slot(c.cnn, c.cnn_slot, b.y);
  cnnlink::unpack_tau<T>(c.cnn, b.y, c.cnn_sout, tau, c.phys_elems);
}

// create_bardina's model (src/closures.jl:68-115): tau = tracefree(
// G~ * (u_i u_j) - (G~ * u_i)(G~ * u_j)), test filter G~ Gaussian of width
// s Delta, each filtered field 2/3-truncated, products of the truncated
// velocities (nonlinearity!)
template <typename T>
void Ops<T>::bardina(Ctx& c, T* tau, const CT* u, double s, double delta) {
  const int D = c.D, nt = c.tdim();
  const long sp = c.spec_elems, np = c.phys_elems;
  const double dt = s * delta;
  CT* UT = reinterpret_cast<CT*>(c.work(3, size_t(D) * sp * sizeof(CT)));
  CT* SIG = reinterpret_cast<CT*>(c.work(4, size_t(nt) * sp * sizeof(CT)));
  CT* SIGT = reinterpret_cast<CT*>(c.work(5, size_t(nt) * sp * sizeof(CT)));
  T* P = reinterpret_cast<T*>(c.work(8, size_t(nt) * np * sizeof(T)));
  SC_HIP(hipMemcpyAsync(UT, u, size_t(D) * sp * sizeof(CT), hipMemcpyDeviceToDevice, c.stream));
  gaussian(c, UT, D, dt);
  twothirds(c, UT, D);
  nonlinearity(c, SIG, u, nullptr);
  nonlinearity(c, SIGT, UT, nullptr);
  gaussian(c, SIG, nt, dt);
  twothirds(c, SIG, nt);
  c2r(c, tau, SIG, nt, 1.0, false);
  c2r(c, P, SIGT, nt, 1.0, false);
  launch_phys(c, tau, P, T(-1), T(1), long(nt) * np);
  phys_tracefree(c, tau);
}

// create_dynamic_smagorinsky's model (src/closures.jl:188-271), with the
// test-filtered strain computed from the truncated strain (the reference
// reads a buffer its own C2R destroyed; see the oracle), arranged so that
// nothing is transformed twice and nothing returns to the host:
//   u~ = 2/3 G~ u, s~ = nonlinearity(u~); L = to_phys(2/3 G~ s - s~) in one
//   C2R (to_phys is linear); S = to_phys(2/3 S^), S~ = to_phys(2/3 G~ S^);
//   m1 = smagorinsky_tensor(S, Delta) and M1 = rfft(m1).  The dealiased m1 is
//   irfft(2/3 M1), so its test filter is irfft(2/3 G~ M1) and the stress's
//   own spectrum is c 2/3 M1 / n^D: no second forward transform.
//   <M:L_dev> and <M:M> with M = that test filter - smagorinsky_tensor(S~,
//   Delta^) formed inside the sums; c = max(-<ML>/<MM>, 0) on the device.
// sig: nonlinearity(u) when the caller has it (les!), else null.  Returns
// M1 = rfft(m1) (unnormalised) and the device address of c.
template <typename T>
static void dynsmag_core(Ctx& c, const C<T>* u, const C<T>* sig, double delta, C<T>*& M1, double*& coef_dev) {
  using CT = C<T>;
  const int D = c.D, nt = c.tdim();
  const long sp = c.spec_elems, np = c.phys_elems;
  const double dt = 2 * delta;
  const double dd = std::sqrt(dt * dt + delta * delta);
  const double inv = 1.0 / std::pow(double(c.n), D);
  CT* UT = reinterpret_cast<CT*>(c.work(3, size_t(D) * sp * sizeof(CT)));
  CT* SIGT = reinterpret_cast<CT*>(c.work(5, size_t(nt) * sp * sizeof(CT)));  // s~, then L^
  CT* SH = reinterpret_cast<CT*>(c.work(6, size_t(nt) * sp * sizeof(CT)));    // S^, then 2/3 G~ M1
  T* L = reinterpret_cast<T*>(c.work(7, size_t(nt) * np * sizeof(T)));
  T* S = reinterpret_cast<T*>(c.work(9, size_t(nt) * np * sizeof(T)));        // S, then m1
  T* St = reinterpret_cast<T*>(c.work(10, size_t(nt) * np * sizeof(T)));
  M1 = reinterpret_cast<CT*>(c.work(11, size_t(nt) * sp * sizeof(CT)));
  T* Mt = reinterpret_cast<T*>(c.work(12, size_t(nt) * np * sizeof(T)));
  if (!sig) {
    CT* SIG = reinterpret_cast<CT*>(c.work(4, size_t(nt) * sp * sizeof(CT)));
    Ops<T>::nonlinearity(c, SIG, u, nullptr);
    sig = SIG;
  }
  // Germano L
  SPEC_LAUNCH(k_gauss23<T>, UT, u, (const CT*)nullptr, dt, g_, D);
  Ops<T>::nonlinearity(c, SIGT, UT, nullptr);
  SPEC_LAUNCH(k_gauss23<T>, SIGT, sig, (const CT*)SIGT, dt, g_, nt);
  Ops<T>::c2r(c, L, SIGT, nt, 1.0, false);
  // strain at the grid and combined filter levels (the 2/3 mask on load)
  Ops<T>::strainrate(c, SH, u);
  Ops<T>::c2r(c, S, SH, nt, 1.0, true);
  Ops<T>::gaussian(c, SH, nt, dt);
  Ops<T>::c2r(c, St, SH, nt, 1.0, true);
  // m1 and its spectrum; the test-filtered dealiased m1
  Ops<T>::smag_tensor(c, S, S, delta);
  Ops<T>::r2c(c, M1, S, nt, 1.0);
  SPEC_LAUNCH(k_gauss23<T>, SH, M1, (const CT*)nullptr, dt, g_, nt);
  Ops<T>::c2r(c, Mt, SH, nt, inv, true);
  // Lilly average on the device
  const unsigned nb = nblk(np);
  double* part = c.dscratch(2 * long(nb) + 3);
  if (D == 3) hipLaunchKernelGGL((k_ml_mm_fused<T, 3>), dim3(nb), dim3(RB), 0, c.stream, Mt, St, L, dd, np, part);
  else hipLaunchKernelGGL((k_ml_mm_fused<T, 2>), dim3(nb), dim3(RB), 0, c.stream, Mt, St, L, dd, np, part);
  hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(1024), 0, c.stream, part, long(nb), part + 2 * nb);
  hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(1024), 0, c.stream, part + nb, long(nb), part + 2 * nb + 1);
  hipLaunchKernelGGL(k_dyn_coef, dim3(1), dim3(1), 0, c.stream, (const double*)(part + 2 * nb), part + 2 * nb + 2);
  SC_HIP(hipGetLastError());
  coef_dev = part + 2 * nb + 2;
}

// the model's physical stress c dealias(m1); returns c
template <typename T>
double Ops<T>::dynsmag(Ctx& c, T* tau, const CT* u, double delta) {
  const int nt = c.tdim();
  const double inv = 1.0 / std::pow(double(c.n), c.D);
  CT* M1 = nullptr;
  double* cd = nullptr;
  dynsmag_core<T>(c, u, nullptr, delta, M1, cd);
  CT* X = reinterpret_cast<CT*>(c.work(6, size_t(nt) * c.spec_elems * sizeof(CT)));
  SPEC_LAUNCH(k_coef23<T>, X, (const CT*)M1, (const double*)cd, inv, 0, g_, nt);
  c2r(c, tau, X, nt, 1.0, true);
  double* h = c.hostbuf(1);
  SC_HIP(hipMemcpyAsync(h, cd, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  SC_HIP(hipStreamSynchronize(c.stream));
  return h[0];
}

// les!'s stress with the dynamic model: sig += c to_spec(dealias(m1)), sig
// holding nonlinearity(u) on entry (shared with the model's Germano L)
template <typename T>
static void dynsmag_add_spec(Ctx& c, C<T>* sig, const C<T>* u, double delta) {
  const double inv = 1.0 / std::pow(double(c.n), c.D);
  C<T>* M1 = nullptr;
  double* cd = nullptr;
  dynsmag_core<T>(c, u, sig, delta, M1, cd);
  SPEC_LAUNCH(k_coef23<T>, sig, (const C<T>*)M1, (const double*)cd, inv, 1, g_, c.tdim());
}

template <typename T>
void Ops<T>::launch_phys(Ctx& c, T* y, const T* x, T a, T b, long n) {
  hipLaunchKernelGGL(k_phys_axpby<T>, dim3(nblk(n)), dim3(256), 0, c.stream, y, x, a, b, n);
  SC_HIP(hipGetLastError());
}

template <typename T>
void Ops<T>::smag_tensor(Ctx& c, T* m, const T* S, double delta) {
  if (c.D == 3) hipLaunchKernelGGL((k_smag_tensor<T, 3>), dim3(nblk(c.phys_elems)), dim3(256), 0, c.stream, m, S, delta, c.phys_elems);
  else hipLaunchKernelGGL((k_smag_tensor<T, 2>), dim3(nblk(c.phys_elems)), dim3(256), 0, c.stream, m, S, delta, c.phys_elems);
  SC_HIP(hipGetLastError());
}

// wray3!(les!, u, dt, grid, cache; model, visc)
template <typename T>
void Ops<T>::les_wray3(Ctx& c, CT* u, double dt, double nu, int closure, double p1, double p2) {
  wray3_dev(c, u, dt, nu, 0.0, closure, p1, p2);
}

// ------------------------------------------------------------------ adjoint
template <typename T>
void Ops<T>::lincomb(Ctx& c, CT* dst, const CT* x, double a, const CT* y, double b, int ncomp) {
  SPEC_LAUNCH(k_lincomb<T>, dst, x, T(a), y, T(b), g_, ncomp);
}

// J(u)^T w of the DNS right-hand side (convectiondiffusion! plus the band
// forcing when enabled) under the RFFT-weighted inner product:
//   g_ij = C2R(i k_j w_i), v = C2R(2/3 u), a_m = sum_j g_mj v_j + sum_i g_im v_i,
//   out = 2/3 R2C(a)/n^D - nu (|k|^2 w + k (k.w)) + forcing transpose.
template <typename T>
void Ops<T>::rhs_vjp(Ctx& c, CT* out, const CT* u, const CT* w, double nu, int closure) {
  rhs_vjp_kept(c, out, u, w, nu, closure, -1);
}

// rhs_vjp; with kept >= 0 the network's forward on u is already in activation
// slot `kept` (cnn_tau on u with cnn_slot = kept, nothing evaluated in that
// slot since), and the backward starts from it
template <typename T>
void Ops<T>::rhs_vjp_kept(Ctx& c, CT* out, const CT* u, const CT* w, double nu, int closure, int kept) {
  const int D = c.D, ng = D * D;
  const long np = c.phys_elems;
  if (closure > 0 && closure != 4) throw Error(1, "the transpose exists for closure kinds 0 and 4");
  if (closure == 4 && !c.cnn) throw Error(1, "closure 4 needs a network attached with sc_set_cnn_closure");
  T* v = reinterpret_cast<T*>(c.work(13, size_t(D) * np * sizeof(T)));
  T* gr = reinterpret_cast<T*>(c.work(14, size_t(ng) * np * sizeof(T)));
  CT* G = reinterpret_cast<CT*>(c.work(15, size_t(ng) * c.spec_elems * sizeof(CT)));
  c2r(c, v, u, D, 1.0, true);
  vectorgradient(c, G, w);
  c2r(c, gr, G, ng, 1.0, false);
  // learned closure: recompute the network on v, then push the stress
  // cotangent tau_bar = sym(g) through it (parameter gradients accumulate).
  // Cotangents under sums over grid points: y_bar = s_out tau_bar / n^D, and
  // the velocity part a += n^D x_bar / s_in.
  const double nD = std::pow(double(c.n), D);
  cnnlink::Bufs cb{};
  if (closure == 4) {
    OnStream on(c.cnn, c.stream);
    cb = cnnlink::buffers(c.cnn);
    const int s = kept >= 0 ? kept : 0;
    if (kept < 0) {
      cnnlink::pack_velocity<T>(c.cnn, v, np, 1.0 / c.cnn_sin, c.cnn_extra, cnnlink::slot_x(c.cnn, 0));
      cnnlink::forward_slot(c.cnn, 0, cb.y);
    }
    cnnlink::pack_tau_cotangent<T>(c.cnn, gr, np, c.cnn_sout / nD, cb.dy);
    cnnlink::backward_slot(c.cnn, s, cb.dy, cb.dx);
  }
  if (D == 3) hipLaunchKernelGGL((k_adj_products<T, 3>), dim3(nblk(np)), dim3(256), 0, c.stream, v, gr, np);
  else hipLaunchKernelGGL((k_adj_products<T, 2>), dim3(nblk(np)), dim3(256), 0, c.stream, v, gr, np);
  SC_HIP(hipGetLastError());
  if (closure == 4) {
    OnStream on(c.cnn, c.stream);
    cnnlink::add_velocity_cotangent<T>(c.cnn, cb.dx, nD / c.cnn_sin, v, np);
  }
  r2c(c, out, v, D, 1.0 / std::pow(double(c.n), D));
  twothirds(c, out, D);
  const double* bandp = nullptr;
  if (c.force_kind == 1) {
    if (!c.bandadj) SC_HIP(hipMalloc(&c.bandadj, 2 * sizeof(double)));
    hipLaunchKernelGGL(k_band_adj<T>, dim3(1), dim3(256), 0, c.stream, u, w, D, c.geom(), c.kband, c.injection,
                       c.bandadj);
    SC_HIP(hipGetLastError());
    bandp = c.bandadj;
  }
  if (D == 3) SPEC_LAUNCH((k_adj_visc_force<T, 3>), out, w, u, nu, g_, bandp, c.kband * c.kband);
  else SPEC_LAUNCH((k_adj_visc_force<T, 2>), out, w, u, nu, g_, bandp, c.kband * c.kband);
}

// Transposed Jacobian of one RK4 step (fixed dt) applied to ubar; the stage
// values are recomputed from the checkpoint u0.
template <typename T>
void Ops<T>::stage_rhs(Ctx& c, CT* du, const CT* u, double nu, int closure) {
  if (closure < 0) rhs_fused(c, du, u, nu, nullptr);
  else les_rhs(c, du, u, nu, closure, 0.0, 0.0);
}

template <typename T>
void Ops<T>::rk4_vjp(Ctx& c, CT* out, const CT* u0, const CT* ubar, double dt, double nu, int closure) {
  const int D = c.D;
  const size_t fb = size_t(D) * c.spec_elems * sizeof(CT);
  CT* k = reinterpret_cast<CT*>(c.adj(0, fb));
  CT* U1 = reinterpret_cast<CT*>(c.adj(1, fb));
  CT* U2 = reinterpret_cast<CT*>(c.adj(2, fb));
  CT* U3 = reinterpret_cast<CT*>(c.adj(3, fb));
  CT* Sb = reinterpret_cast<CT*>(c.adj(4, fb));
  CT* KB = reinterpret_cast<CT*>(c.adj(5, fb));
  CT* Tt = reinterpret_cast<CT*>(c.adj(6, fb));
  CT* acc = reinterpret_cast<CT*>(c.adj(7, fb));
  // forward recompute of the stage values.  With the learned closure and
  // four activation slots, the network's forwards on u0, U1, U2 stay in
  // slots 1, 2, 3 and the transposes of those stages start from them
  const bool keep = closure == 4 && c.cnn && cnnlink::slots(c.cnn) >= 4;
  auto stage = [&](CT* kk, const CT* uu, int slot) {
    if (keep) c.cnn_slot = slot;
    try {
      stage_rhs(c, kk, uu, nu, closure);
    } catch (...) {
      c.cnn_slot = 0;
      throw;
    }
    c.cnn_slot = 0;
  };
  stage(k, u0, 1);
  lincomb(c, U1, u0, 1.0, k, dt / 2, D);
  project(c, U1);
  stage(k, U1, 2);
  lincomb(c, U2, u0, 1.0, k, dt / 2, D);
  project(c, U2);
  stage(k, U2, 3);
  lincomb(c, U3, u0, 1.0, k, dt, D);
  project(c, U3);
  // reverse sweep
  SC_HIP(hipMemcpyAsync(Sb, ubar, fb, hipMemcpyDeviceToDevice, c.stream));
  project(c, Sb);
  SC_HIP(hipMemcpyAsync(acc, Sb, fb, hipMemcpyDeviceToDevice, c.stream));
  lincomb(c, KB, Sb, dt / 6, Sb, 0.0, D);
  rhs_vjp(c, Tt, U3, KB, nu, closure);
  project(c, Tt);
  axpy(c, acc, Tt, 1.0, D);
  lincomb(c, KB, Sb, dt / 3, Tt, dt, D);
  rhs_vjp_kept(c, Tt, U2, KB, nu, closure, keep ? 3 : -1);
  project(c, Tt);
  axpy(c, acc, Tt, 1.0, D);
  lincomb(c, KB, Sb, dt / 3, Tt, dt / 2, D);
  rhs_vjp_kept(c, Tt, U1, KB, nu, closure, keep ? 2 : -1);
  project(c, Tt);
  axpy(c, acc, Tt, 1.0, D);
  lincomb(c, KB, Sb, dt / 6, Tt, dt / 2, D);
  rhs_vjp_kept(c, Tt, u0, KB, nu, closure, keep ? 1 : -1);
  lincomb(c, out, acc, 1.0, Tt, 1.0, D);
}

// Transposed Jacobian of one Wray3 step (fixed dt) applied to ubar.
template <typename T>
void Ops<T>::wray3_vjp(Ctx& c, CT* out, const CT* u0, const CT* ubar, double dt, double nu, int closure) {
  const int D = c.D;
  const size_t fb = size_t(D) * c.spec_elems * sizeof(CT);
  const double a1 = double(T(8.0 / 15.0)), a2 = double(T(5.0 / 12.0)), a3 = double(T(3.0 / 4.0));
  const double b1 = double(T(1.0 / 4.0));
  CT* k = reinterpret_cast<CT*>(c.adj(0, fb));
  CT* U1 = reinterpret_cast<CT*>(c.adj(1, fb));
  CT* S1 = reinterpret_cast<CT*>(c.adj(2, fb));
  CT* U2 = reinterpret_cast<CT*>(c.adj(3, fb));
  CT* Q = reinterpret_cast<CT*>(c.adj(4, fb));
  CT* S1b = reinterpret_cast<CT*>(c.adj(5, fb));
  CT* Tt = reinterpret_cast<CT*>(c.adj(6, fb));
  CT* KB = reinterpret_cast<CT*>(c.adj(7, fb));
  stage_rhs(c, k, u0, nu, closure);
  lincomb(c, U1, u0, 1.0, k, a1 * dt, D);
  project(c, U1);
  lincomb(c, S1, u0, 1.0, k, b1 * dt, D);
  stage_rhs(c, k, U1, nu, closure);
  lincomb(c, U2, S1, 1.0, k, a2 * dt, D);
  project(c, U2);
  // reverse
  SC_HIP(hipMemcpyAsync(Q, ubar, fb, hipMemcpyDeviceToDevice, c.stream));
  project(c, Q);                                   // q3
  SC_HIP(hipMemcpyAsync(S1b, Q, fb, hipMemcpyDeviceToDevice, c.stream));
  lincomb(c, KB, Q, a3 * dt, Q, 0.0, D);
  rhs_vjp(c, Tt, U2, KB, nu, closure);                      // u2bar
  project(c, Tt);                                  // q2
  axpy(c, S1b, Tt, 1.0, D);
  lincomb(c, KB, Tt, a2 * dt, Tt, 0.0, D);
  rhs_vjp(c, Q, U1, KB, nu, closure);                       // u1bar
  project(c, Q);                                   // q1
  lincomb(c, KB, Q, a1 * dt, S1b, b1 * dt, D);     // k1bar
  rhs_vjp(c, Tt, u0, KB, nu, closure);
  lincomb(c, out, Q, 1.0, S1b, 1.0, D);
  axpy(c, out, Tt, 1.0, D);
}

}  // namespace sc
