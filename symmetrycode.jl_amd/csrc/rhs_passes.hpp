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
#include <type_traits>
#include "fft_passes.hpp"
#include "solver_kernels.hpp"

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
  int tile, oi;  // oi: z index; z pairs side by side (the output is z-tiled)
  block_tile_other(xcd_remap(blockIdx.x, nb), g.ntile, g.nother, 1, tile, oi);
  const int t = threadIdx.x;
  const int xl = t % TX, p = t / TX;
  const int x = tile * TX + xl;
  const bool xok = x < g.xs;
  const C<T>* base = buf + long(oi) * g.s_other + x;
  constexpr long LP = tile_line<T, N>();
  C<T>* tbase = outT + long(tile) * N * LP + long(oi) * TX + xl;  // (tile, y = 0, z = oi)
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
    PE::for_out(p, [&](int e, int k) { dst[long(k) * LP] = val[e]; });
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
  const C<T>* tsrc = bufT + (long(tile) * N + oi) * tile_line<T, N>() + xl;  // (tile, y = oi, z = 0)
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

// LDS of k_z_fwd_rk4 (the transform exchange, half a twiddle table and two
// waiting components): the fused stage exists where it fits one block
template <typename T, int N> constexpr bool z_rk4_fits() {
  return (3L * tile_x<T, N>() * N + N / 2) * long(sizeof(C<T>)) <= 163840;
}

// RK4 stage as the last forward pass, for a Runge-Kutta state kept in
// z-pencil tiles like the y-pass output (element (x, y, z) of a component at
// ((x/TX n + y) n + z) TX + x%TX, component stride s_slot, so one block's
// share of every state component is one contiguous run).  For each component
// i the divergence -i (A_i + kz s_iz) of the six transformed slots gives
// k_i = div_i - nu |k|^2 us_i (+ A_f us_i on the band 0 < |k|^2 <= kband2),
// us the stage input (u0 at stage 0, ust after), and then the RK4 update with
// acc holding the running new state: stage 0 acc = u0 + dt/6 k, stages 1-2
// acc += dt/3 k, the stage value v = u0 + c dt k; stage 3's value is the new
// state acc + dt/6 k, so the last stage reads no u0 (three streams instead of
// four).  The first two components' stage values wait in LDS for the third
// (each thread reads back its own entries); then the three are projected with
// the FP64 |k|^2 division (acc carries the unprojected k, and P(u0 + c dt k)
// = u0 + c dt P(k) for the projected u0) and written to ust (stage 3: u0).  kz s_iz of components 0 and 1 is parked in
// the LDS slot that the component's stage value overwrites afterwards; the
// third component is peeled off the loop and holds it in registers, which the
// absent prefetch leaves free.  LDS: the transform exchange, half a twiddle
// table and the two waiting components (158 KB at n = 810 FP64).  Each state
// element is read and written by the one thread that owns its (x, z)
// position, so the update runs in place.
template <typename T, int N>
__global__ void SC_LB_WIDE((wide_threads<T, N>()))
k_z_fwd_rk4(const C<T>* bufT, long s_slot, C<T>* u0, C<T>* ust, C<T>* acc, const C<T>* __restrict__ tw,
            StridedGeom g, RhsParams rp, const double* dtp, double frac, int stage) {
  constexpr int TX = tile_x<T, N>();
  constexpr int TPP = tpp_wide<N>();
  using PE = Pencil<T, N, -1, TPP>;
  constexpr int E = PE::EL;
  static_assert(z_rk4_fits<T, N>(), "k_z_fwd_rk4: LDS");
  __shared__ C<T> lds[TX * N];
  __shared__ C<T> twh[N / 2];
  __shared__ C<T> vw[2][TX * N];  // kz s_iz, then the stage values of components 0 and 1 (natural order)
  for (int i = threadIdx.x; i < N / 2; i += blockDim.x) twh[i] = tw[i];
  __syncthreads();
  const HalfTwiddles<T, N> twl{twh};
  const int nb = g.ntile * g.nother;
  const int b = xcd_remap(blockIdx.x, nb);
  const int tile = b % g.ntile;
  const int oi = b / g.ntile;  // y index
  const int t = threadIdx.x;
  const int xl = t % TX, p = t / TX;
  const int x = tile * TX + xl;
  const bool xok = x < g.xs;
  const long toff = (long(tile) * N + oi) * tile_line<T, N>() + xl;  // (tile, y = oi, z = 0) of a slot or component
  const T kf = T(rp.kf);
  const int kyi = freq(oi, g.n_other);
  const T kx = kf * T(x), ky = kf * T(kyi);
  const T nu = T(rp.nu);
  const T cdt = T(*dtp * frac);
  const T wdt = T(*dtp * (stage == 0 || stage == 3 ? 1.0 / 6.0 : 1.0 / 3.0));  // RK4 weight x dt
  T fA = T(0);
  if (rp.forceA) fA = T(*rp.forceA);
  SpecGeom pg{};
  pg.kf = rp.kf;
  C<T> v[PE::MAXE];
  C<T> nx[PE::E0];
  C<T> kzs[E];  // kz * FFT(s_zz) of the third component
  const C<T>* tsrc = bufT + toff;
  auto fetch = [&](int slot) __attribute__((always_inline)) {
    const C<T>* src = tsrc + long(slot) * s_slot;
    PE::load(nx, p, [&](int j) { return xok ? src[j * TX] : mk<T>(T(0), T(0)); });
  };
  auto fft = [&](int next) __attribute__((always_inline)) {
#pragma unroll
    for (int e = 0; e < PE::E0; ++e) v[e] = nx[e];
    if (next >= 0) fetch(next);
    PE::run(v, p, lds + xl, TX, twl);
  };
  // slots of A_i (A_x, A_y, A_z) and of s_iz (s_zx, s_yz, s_zz)
  auto SA = [](int i) { return i < 2 ? i : 3; };
  auto SZ = [](int i) { return i < 2 ? 5 - i : 2; };
  // the two transforms of component i; v holds FFT(A_i) afterwards
  auto transforms = [&](int i, auto last) __attribute__((always_inline)) {
    fft(SA(i));
    if constexpr (decltype(last)::value)
      PE::for_out(p, [&](int e, int k) { kzs[e] = scl<T>(v[e], kf * T(freq(k, N))); });
    else
      PE::for_out(p, [&](int e, int k) { vw[i][k * TX + xl] = scl<T>(v[e], kf * T(freq(k, N))); });
    fft(decltype(last)::value ? -1 : SZ(i + 1));
  };
  // the update of component i; sink(k, v) receives its stage value at z index k
  auto update = [&](int i, auto last, auto&& sink) __attribute__((always_inline)) {
    const long ci = long(i) * s_slot + toff;
    PE::for_out(p, [&](int e, int k) {
      const int kzi = freq(k, N);
      const T kz = kf * T(kzi);
      C<T> a;
      if constexpr (decltype(last)::value) a = add<T>(kzs[e], v[e]);
      else a = add<T>(vw[i][k * TX + xl], v[e]);
      C<T> kk = mk<T>(a.y, -a.x);  // -i a
      const long o = ci + long(k) * TX;
      const C<T> uu = stage == 0 ? u0[o] : ust[o];
      const T k2 = kx * kx + ky * ky + kz * kz;
      kk = mk<T>(kk.x - nu * k2 * uu.x, kk.y - nu * k2 * uu.y);
      if (rp.forceA) {
        const int k2i = x * x + kyi * kyi + kzi * kzi;
        if (k2i > 0 && k2i <= rp.kband2) kk = add<T>(kk, scl<T>(uu, fA));
      }
      // separate branches: with the accumulator's load and store in one
      // statement the compiler may issue the load at stage 0 too
      if (stage == 3) {
        sink(k, add<T>(acc[o], scl<T>(kk, wdt)));
      } else if (stage == 0) {
        acc[o] = add<T>(uu, scl<T>(kk, wdt));
        sink(k, add<T>(uu, scl<T>(kk, cdt)));
      } else {
        const C<T> b0 = u0[o];
        acc[o] = add<T>(acc[o], scl<T>(kk, wdt));
        sink(k, add<T>(b0, scl<T>(kk, cdt)));
      }
    });
  };
  fetch(SZ(0));
#pragma unroll 1
  for (int i = 0; i < 2; ++i) {
    transforms(i, std::false_type{});
    if (xok) update(i, std::false_type{}, [&](int k, C<T> vv) { vw[i][k * TX + xl] = vv; });
  }
  transforms(2, std::true_type{});
  if (xok)
    update(2, std::true_type{}, [&](int k, C<T> vv) {
      Mode m;
      m.kx = x;
      m.ky = kyi;
      m.kz = freq(k, N);
      C<T> w[3] = {vw[0][k * TX + xl], vw[1][k * TX + xl], vv};
      project_mode<T, 3>(m, pg, w);
      C<T>* dst = (stage < 3 ? ust : u0) + toff + long(k) * TX;
#pragma unroll
      for (int c = 0; c < 3; ++c) dst[long(c) * s_slot] = w[c];
    });
}

// Layout conversion of ncomp spectral components between the x-contiguous
// layout (element (x, y, z) at (z n + y) nhp + x, component stride s_spec)
// and z-pencil tiles ((x/TX n + y) lp + z TX + x%TX, lp = n TX + padding,
// component stride ts).
// A block moves a tile of TX x by 8 y by 32 z through LDS, so the spectral
// side is accessed in runs of TX elements and the tiled side in runs of
// 32 TX.  Grid: (ntile * ceil(n/8) * ceil(n/32), ncomp).
template <typename T, int TX, bool TO_TILES>
__global__ void __launch_bounds__(256) k_ztile_convert(const C<T>* __restrict__ src, C<T>* dst, int n, int nhp,
                                                       long s_spec, long ts, long lp) {
  constexpr int YB = 8, ZB = 32;
  constexpr int NE = YB * ZB * TX;
  __shared__ C<T> buf[YB][ZB][TX + 1];
  const int nyb = (n + YB - 1) / YB, nzb = (n + ZB - 1) / ZB;
  long bid = blockIdx.x;
  const int zb = int(bid % nzb);
  bid /= nzb;
  const int yb = int(bid % nyb);
  const int tile = int(bid / nyb);
  const long x0 = long(tile) * TX;
  const C<T>* sp = src + long(blockIdx.y) * (TO_TILES ? s_spec : ts);
  C<T>* dp = dst + long(blockIdx.y) * (TO_TILES ? ts : s_spec);
  // spectral order: xl fastest, then y, then z; tiled order: xl, z, y
  auto spec_at = [&](int e, int& yy, int& zz, long& off) {
    const int xl = e % TX;
    yy = (e / TX) % YB;
    zz = e / (TX * YB);
    off = (long(zb * ZB + zz) * n + (yb * YB + yy)) * nhp + x0 + xl;
  };
  auto tile_at = [&](int e, int& yy, int& zz, long& off) {
    const int xl = e % TX;
    zz = (e / TX) % ZB;
    yy = e / (TX * ZB);
    off = (long(tile) * n + (yb * YB + yy)) * lp + long(zb * ZB + zz) * TX + xl;
  };
  for (int e = threadIdx.x; e < NE; e += 256) {
    int yy, zz;
    long off;
    if (TO_TILES) spec_at(e, yy, zz, off);
    else tile_at(e, yy, zz, off);
    if (yb * YB + yy < n && zb * ZB + zz < n) buf[yy][zz][e % TX] = sp[off];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < NE; e += 256) {
    int yy, zz;
    long off;
    if (TO_TILES) tile_at(e, yy, zz, off);
    else spec_at(e, yy, zz, off);
    if (yb * YB + yy < n && zb * ZB + zz < n) dp[off] = buf[yy][zz][e % TX];
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
static __global__ void k_copy_f64(double* dst, const double* src) { *dst = *src; }
static __global__ void k_set_u64(unsigned long long* p, unsigned long long v) { *p = v; }

}  // namespace sc
