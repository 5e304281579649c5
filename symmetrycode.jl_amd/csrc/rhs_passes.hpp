// Fused forward passes of the DNS right-hand side (time-stepper path).
//
// After the fused x-pass the six stress components sigma_ij(kx, y, z) are
// transformed along y and z.  The divergence du_i = -i k_j sigma_ij is split
// over the two passes: the y-pass forms A_i = kx sigma_ix + ky sigma_iy as
// soon as ky is known, the z-pass adds kz sigma_iz, the viscous term
// -nu |k|^2 u_i and the band forcing, and writes du directly.  The viscous
// term of viscosity! also carries -nu k_i (k.u), a gradient that the
// projection of every Runge-Kutta stage removes exactly; the stepper path
// therefore omits it (the stand-alone convectiondiffusion keeps it).
#pragma once
#include "fft_passes.hpp"

namespace sc {

// threads per pencil giving one butterfly per thread in the largest stage
#ifndef SC_WIDE
#define SC_WIDE 1
#endif
template <int N> constexpr int tpp_wide() {
  if (!SC_WIDE) return Plan<N>::tpp;
  int m = 1;
  for (int s = 0; s < Plan<N>::ns; ++s) {
    int b = N / Plan<N>::r[s];
    m = b > m ? b : m;
  }
  return m;
}
template <typename T, int N> constexpr int wide_threads() { return tile_x<T, N>() * tpp_wide<N>(); }

struct RhsParams {
  double kf;             // (pi / l) * 2
  double nu;
  const double* forceA;  // band forcing coefficient (device) or null
  int kband2;
};

// Copy an n-entry twiddle table into LDS (all threads participate).
template <typename T>
DEV void stage_twiddles(C<T>* dst, const C<T>* __restrict__ src, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
}

// Output slots of the y-pass: 0 A_x, 1 A_y, 2 s_zz, 3 A_z, 4 s_yz, 5 s_zx.
// Input slots (packed stress): 0 xx, 1 yy, 2 zz, 3 xy, 4 yz, 5 zx.  The
// output goes to a z-pencil-tiled buffer, element (x, y, z) of a slot at
// ((tile * n + y) * n + z) * TX + x % TX with tile = x / TX: the z-pass then
// reads each of its pencil tiles as one contiguous run, instead of TX
// elements every 16 (n/2+1) n bytes, which misses the address-translation
// cache on almost every access.  The global loads of the next
// component are issued before the current component's FFT (software
// pipelining); twiddles are read from LDS so that waiting for them never
// waits for the prefetch.
template <typename T, int N>
__global__ void SC_LB_WIDE((wide_threads<T, N>()))
k_y_fwd_div(const C<T>* buf, C<T>* outT, long s_slot, const C<T>* __restrict__ tw, StridedGeom g, double kf) {
  constexpr int TX = tile_x<T, N>();
  constexpr int TPP = tpp_wide<N>();
  using PE = Pencil<T, N, -1, TPP>;
  constexpr int E = PE::EL;
  __shared__ C<T> lds[TX * N];
  __shared__ C<T> twl[N];
  stage_twiddles<T>(twl, tw, N);
  const int nb = g.ntile * g.nother;
  const int b = xcd_remap(blockIdx.x, nb);
  const int tile = b % g.ntile;
  const int oi = b / g.ntile;  // z index
  const int t = threadIdx.x;
  const int xl = t % TX, p = t / TX;
  const int x = tile * TX + xl;
  const bool xok = x < g.xs;
  const C<T>* base = buf + long(oi) * g.s_other + x;
  C<T>* tbase = outT + (long(tile) * N * N + oi) * TX + xl;  // (tile, y = 0, z = oi)
  const T kx = T(kf) * T(x);
  C<T> v[PE::MAXE];
  C<T> nx[PE::E0];
  C<T> acc[E];
  auto fetch = [&](int comp) {
    const C<T>* src = base + long(comp) * g.s_comp;
    PE::load(nx, p, [&](int j) { return xok ? src[long(j) * g.s_axis] : mk<T>(T(0), T(0)); });
  };
  // run the FFT of the prefetched component, prefetching `next` (-1: none)
  auto fft = [&](int next) {
#pragma unroll
    for (int e = 0; e < PE::E0; ++e) v[e] = nx[e];
    if (next >= 0) fetch(next);
    PE::run(v, p, lds + xl, TX, twl);
  };
  auto put = [&](int slot, const C<T>* val) {
    if (!xok) return;
    C<T>* dst = tbase + long(slot) * s_slot;
    PE::for_out(p, [&](int e, int k) { dst[long(k) * (N * TX)] = val[e]; });
  };
  auto kyv = [&](int k) { return T(kf) * T(freq(k, N)); };
  fetch(0);
  // xx -> A_x = kx s_xx
  fft(3);
  PE::for_out(p, [&](int e, int) { acc[e] = scl<T>(v[e], kx); });
  // xy -> A_x += ky s_xy (write slot 0); A_y = kx s_xy
  fft(1);
  PE::for_out(p, [&](int e, int k) { acc[e] = add<T>(acc[e], scl<T>(v[e], kyv(k))); });
  put(0, acc);
  PE::for_out(p, [&](int e, int) { acc[e] = scl<T>(v[e], kx); });
  // yy -> A_y += ky s_yy (slot 1)
  fft(2);
  PE::for_out(p, [&](int e, int k) { acc[e] = add<T>(acc[e], scl<T>(v[e], kyv(k))); });
  put(1, acc);
  // zz -> slot 2
  fft(5);
  put(2, v);
  // zx -> slot 5; A_z = kx s_zx
  fft(4);
  put(5, v);
  PE::for_out(p, [&](int e, int) { acc[e] = scl<T>(v[e], kx); });
  // yz -> slot 4; A_z += ky s_yz (slot 3)
  fft(-1);
  put(4, v);
  PE::for_out(p, [&](int e, int k) { acc[e] = add<T>(acc[e], scl<T>(v[e], kyv(k))); });
  put(3, acc);
}

// Final z-pass: du_i = -i (A_i + kz s_iz) - nu |k|^2 u_i (+ A_f u_i on the band).
// With LIN false only the divergence -i (A_i + kz s_iz) is written and u is
// not read: the RK4 update adds the viscous and forcing terms in its
// contiguous pass (k_rk4_update_lin), where u costs no strided access.  The
// six slots come from the y-pass's z-pencil-tiled buffer (slot stride s_slot).
template <typename T, int N, bool LIN>
__global__ void SC_LB_WIDE((wide_threads<T, N>()))
k_z_fwd_rhs(const C<T>* bufT, long s_slot, const C<T>* __restrict__ u, C<T>* du, const C<T>* __restrict__ tw,
            StridedGeom g, RhsParams rp) {
  constexpr int TX = tile_x<T, N>();
  constexpr int TPP = tpp_wide<N>();
  using PE = Pencil<T, N, -1, TPP>;
  constexpr int E = PE::EL;
  __shared__ C<T> lds[TX * N];
  __shared__ C<T> twl[N];
  __shared__ C<T> accl[TX * N];  // kz * FFT(s_iz), natural order (each thread reads back its own)
  stage_twiddles<T>(twl, tw, N);
  const int nb = g.ntile * g.nother;
  const int b = xcd_remap(blockIdx.x, nb);
  const int tile = b % g.ntile;
  const int oi = b / g.ntile;  // y index
  const int t = threadIdx.x;
  const int xl = t % TX, p = t / TX;
  const int x = tile * TX + xl;
  const bool xok = x < g.xs;
  const long off = long(oi) * g.s_other + x;
  const T kf = T(rp.kf);
  const int kyi = freq(oi, g.n_other);
  const T kx = kf * T(x), ky = kf * T(kyi);
  const T nu = T(rp.nu);
  T fA = T(0);
  if (rp.forceA) fA = T(*rp.forceA);
  C<T> v[PE::MAXE];
  C<T> nx[PE::E0];
  const C<T>* tsrc = bufT + (long(tile) * N + oi) * N * TX + xl;  // (tile, y = oi, z = 0)
  auto fetch = [&](int slot) {
    const C<T>* src = tsrc + long(slot) * s_slot;
    PE::load(nx, p, [&](int j) { return xok ? src[j * TX] : mk<T>(T(0), T(0)); });
  };
  auto fft = [&](int next) {
#pragma unroll
    for (int e = 0; e < PE::E0; ++e) v[e] = nx[e];
    if (next >= 0) fetch(next);
    PE::run(v, p, lds + xl, TX, twl);
  };
  constexpr int SA[3] = {0, 1, 3};  // A_x, A_y, A_z
  constexpr int SZ[3] = {5, 4, 2};  // s_zx, s_yz, s_zz
  fetch(SZ[0]);
#pragma unroll 1
  for (int i = 0; i < 3; ++i) {
    fft(SA[i]);
    PE::for_out(p, [&](int e, int k) { accl[k * TX + xl] = scl<T>(v[e], kf * T(freq(k, N))); });
    fft(i < 2 ? SZ[i + 1] : -1);
    if (xok) {
      const C<T>* ui = u + long(i) * g.s_comp + off;
      C<T>* di = du + long(i) * g.s_comp + off;
      PE::for_out(p, [&](int e, int k) {
        const int kzi = freq(k, N);
        const T kz = kf * T(kzi);
        C<T> a = add<T>(accl[k * TX + xl], v[e]);
        C<T> r = mk<T>(a.y, -a.x);  // -i a
        if constexpr (LIN) {
          C<T> uu = ui[long(k) * g.s_axis];
          const T k2 = kx * kx + ky * ky + kz * kz;
          r = mk<T>(r.x - nu * k2 * uu.x, r.y - nu * k2 * uu.y);  // - nu k^2 u
          if (rp.forceA) {
            int k2i = x * x + kyi * kyi + kzi * kzi;
            if (k2i > 0 && k2i <= rp.kband2) r = add<T>(r, scl<T>(uu, fA));
          }
        }
        di[long(k) * g.s_axis] = r;
      });
    }
  }
}

// 2D final y-pass over the packed stress (xx, yy, xy).
template <typename T, int N>
__global__ void SC_LB_WIDE((wide_threads<T, N>()))
k_y_fwd_rhs2d(const C<T>* buf, const C<T>* __restrict__ u, C<T>* du, const C<T>* __restrict__ tw, StridedGeom g,
              RhsParams rp) {
  constexpr int TX = tile_x<T, N>();
  constexpr int TPP = tpp_wide<N>();
  using PE = Pencil<T, N, -1, TPP>;
  constexpr int E = PE::PI::E_LAST;
  __shared__ C<T> lds[TX * N];
  __shared__ C<T> twl[N];
  stage_twiddles<T>(twl, tw, N);
  const int b = xcd_remap(blockIdx.x, g.ntile);
  const int t = threadIdx.x;
  const int xl = t % TX, p = t / TX;
  const int x = b * TX + xl;
  const bool xok = x < g.xs;
  const T kf = T(rp.kf);
  const T kx = kf * T(x);
  const T nu = T(rp.nu);
  T fA = T(0);
  if (rp.forceA) fA = T(*rp.forceA);
  C<T> v[PE::MAXE];
  C<T> sxy[E];
  auto fft = [&](int slot) {
    const C<T>* src = buf + long(slot) * g.s_comp + x;
    PE::load(v, p, [&](int j) { return xok ? src[long(j) * g.s_axis] : mk<T>(T(0), T(0)); });
    PE::run(v, p, lds + xl, TX, twl);
  };
  fft(2);
  PE::for_out(p, [&](int e, int) { sxy[e] = v[e]; });
#pragma unroll 1
  for (int i = 0; i < 2; ++i) {
    fft(i);
    if (xok) {
      const C<T>* ui = u + long(i) * g.s_comp + x;
      C<T>* di = du + long(i) * g.s_comp + x;
      PE::for_out(p, [&](int e, int k) {
        const int kyi = freq(k, N);
        const T ky = kf * T(kyi);
        // i = 0: kx s_xx + ky s_xy ; i = 1: kx s_xy + ky s_yy
        C<T> a = i == 0 ? add<T>(scl<T>(v[e], kx), scl<T>(sxy[e], ky)) : add<T>(scl<T>(sxy[e], kx), scl<T>(v[e], ky));
        C<T> uu = ui[long(k) * g.s_axis];
        const T k2 = kx * kx + ky * ky;
        C<T> r = mk<T>(a.y - nu * k2 * uu.x, -a.x - nu * k2 * uu.y);
        if (rp.forceA) {
          int k2i = x * x + kyi * kyi;
          if (k2i > 0 && k2i <= rp.kband2) r = add<T>(r, scl<T>(uu, fA));
        }
        di[long(k) * g.s_axis] = r;
      });
    }
  }
}

// --------------------------------------------------------------- CFL dt
static __global__ void k_cfl_dt(const unsigned long long* vmax_bits, double cfl, double h, int D, double nu, double dtmax,
                         double* dt) {
  double vmax = __longlong_as_double((long long)*vmax_bits);
  double dconv = h / vmax;
  double ddiff = h * h / D / 2 / nu;
  double d = cfl * fmin(dconv, ddiff);
  *dt = fmin(d, dtmax);
}
static __global__ void k_set_f64(double* p, double v) { *p = v; }
static __global__ void k_set_u64(unsigned long long* p, unsigned long long v) { *p = v; }

}  // namespace sc
