// FFT pass kernels over the device spectral layout.
//
// Device layout of one spectral scalar: C<T>[nz][n][nhp] (x fastest), with
// nhp = nh = n/2+1 rounded up to a 128-byte multiple so every x-row starts on
// a cache line.  nz = n in 3D and 1 in 2D.  Physical scalars are T[nz][n][n]
// (the reference layout).  Multi-component fields store components back to
// back at a fixed component stride.
#pragma once
#include "fft_engine.hpp"

namespace sc {

// Blocks are dispatched round-robin over the 8 XCDs.  Remap so that
// consecutive logical tiles (which share 128-byte lines and L2 working sets)
// land on the same XCD at about the same time.
DEV int xcd_remap(int b, int nb) {
  int q = nb >> 3;
  if (b >= (q << 3)) return b;
  return (b & 7) * q + (b >> 3);
}

// Logical block b -> (tile, other index o).  pair = 0: tile fastest.
// pair = 1: adjacent other indices side by side (b = 2 (o/2 ntile + tile) +
// o%2; an odd last index runs alone at the end).  A pass whose output is
// tiled with the other index next to the x tile (element (.., o, x%TX), a
// 64-byte run per o) then has the two blocks that write the halves of one
// 128-byte line running together, so the line is completed in L2 instead of
// reaching HBM as two partial writes far apart in time.
#ifndef SC_PAIR_OTHER
#define SC_PAIR_OTHER 1
#endif
DEV void block_tile_other(int b, int ntile, int nother, int pair, int& tile, int& o) {
  if (!(SC_PAIR_OTHER && pair)) {
    tile = b % ntile;
    o = b / ntile;
    return;
  }
  const int full = (nother / 2) * 2 * ntile;
  if (b < full) {
    tile = (b >> 1) % ntile;
    o = 2 * ((b >> 1) / ntile) + (b & 1);
  } else {
    tile = b - full;
    o = nother - 1;
  }
}

// pencils per block along x for the strided passes: a 128-byte row segment,
// halved until the LDS tile fits in 64 KiB
#ifndef SC_TILE_BYTES
#define SC_TILE_BYTES 128
#endif
template <typename T, int N> constexpr int tile_x() {
  int tx = SC_TILE_BYTES / int(sizeof(C<T>));
  while (tx > 1 && tx * N * int(sizeof(C<T>)) > 65536) tx /= 2;
  return tx;
}
// elements from one line of the tiled layouts (the n pencil positions of a
// tile of TX x-values) to the next: n TX plus SC_TILE_PAD elements
#ifndef SC_TILE_PAD
#define SC_TILE_PAD 0
#endif
template <typename T, int N> constexpr long tile_line() { return long(N) * tile_x<T, N>() + SC_TILE_PAD; }

// ------------------------------------------------------------------
// Strided complex pass along y or z.  Each block transforms TX consecutive-x
// pencils at one "other" index (z for a y-pass, y for a z-pass).
struct StridedGeom {
  int xs;             // process x in [0, xs)
  int ntile;          // ceil(xs / TX)
  int nother;         // number of other-axis indices processed
  int other_lo;       // other index o maps to o if o < other_lo, else o + other_gap
  int other_gap;
  int n_other;        // length of the other axis (for its frequency)
  long s_axis;        // element stride along the transform axis
  long s_other;       // element stride along the other axis
  long s_comp;        // component stride
  long s_tile;        // stride of an x tile (TX for the x-contiguous layout)
  // output strides (k_c2c_strided; equal to the input's unless the output
  // goes to another layout)
  long o_axis, o_other, o_comp, o_tile;
  int kin;            // read only axis positions with |freq| <= kin
  int mask23;         // apply the 2/3 box mask (and zero mode) on load
  int kcut;           // n / 3, for mask23
  double scale;       // output scale
  int pair = 0;       // block order of block_tile_other (1: the output is tiled along the other axis)
};

// minimum waves per SIMD requested from the register allocator (0 = none)
#ifndef SC_WPE_C2C
#define SC_WPE_C2C 0
#endif
#ifndef SC_WPE_X
#define SC_WPE_X 0
#endif
#ifndef SC_WPE_WIDE
#define SC_WPE_WIDE 0
#endif
#if SC_WPE_C2C
#define SC_LB_C2C(n) __launch_bounds__((n), SC_WPE_C2C)
#else
#define SC_LB_C2C(n) __launch_bounds__((n))
#endif
#if SC_WPE_X
#define SC_LB_X(n) __launch_bounds__((n), SC_WPE_X)
#else
#define SC_LB_X(n) __launch_bounds__((n))
#endif
#if SC_WPE_WIDE
#define SC_LB_WIDE(n) __launch_bounds__((n), SC_WPE_WIDE)
#else
#define SC_LB_WIDE(n) __launch_bounds__((n))
#endif

// All components of a tile are handled by one block: the global loads of
// component c+1 are issued before the FFT of component c, and twiddles are
// read from LDS, so the memory pipeline stays busy during the transforms.
template <typename T, int N, int SGN>
__global__ void SC_LB_C2C((tile_x<T, N>() * Plan<N>::tpp))
k_c2c_strided(const C<T>* __restrict__ in, C<T>* out, const C<T>* __restrict__ tw, StridedGeom g, int ncomp) {
  constexpr int TX = tile_x<T, N>();
  using PE = Pencil<T, N, SGN>;
  __shared__ C<T> lds[TX * N];
  __shared__ C<T> twl[N];
  for (int i = threadIdx.x; i < N; i += blockDim.x) twl[i] = tw[i];
  __syncthreads();
  const int nb = g.ntile * g.nother;
  int tile, o;
  block_tile_other(xcd_remap(blockIdx.x, nb), g.ntile, g.nother, g.pair, tile, o);
  const int oi = o < g.other_lo ? o : o + g.other_gap;
  const int t = threadIdx.x;
  const int xl = t % TX, p = t / TX;
  const int x = tile * TX + xl;
  const bool xok = x < g.xs;
  const long base = long(oi) * g.s_other + tile * g.s_tile + xl;
  const long obase = long(oi) * g.o_other + tile * g.o_tile + xl;
  bool xmask = true;
  bool zero_o = false;
  if (g.mask23) {
    int ko = freq(oi, g.n_other);
    xmask = (x <= g.kcut) && (ko <= g.kcut && ko >= -g.kcut);
    zero_o = (x == 0) && (ko == 0);
  }
  C<T> v[PE::MAXE];
  C<T> nx[PE::E0];
  auto fetch = [&](int c) {
    const C<T>* src = in + long(c) * g.s_comp + base;
    PE::load(nx, p, [&](int j) {
      int kj = freq(j, N);
      bool ok = xok && xmask && kj <= g.kin && kj >= -g.kin;
      if (zero_o && j == 0) ok = false;
      return ok ? src[long(j) * g.s_axis] : mk<T>(T(0), T(0));
    });
  };
  const T sc_ = T(g.scale);
  fetch(0);
#pragma unroll 1
  for (int c = 0; c < ncomp; ++c) {
#pragma unroll
    for (int e = 0; e < PE::E0; ++e) v[e] = nx[e];
    if (c + 1 < ncomp) fetch(c + 1);
    PE::run(v, p, lds + xl, TX, twl);
    if (xok) {
      C<T>* dst = out + long(c) * g.o_comp + obase;
      PE::store(v, p, [&](int k, C<T> val) { dst[long(k) * g.o_axis] = scl<T>(val, sc_); });
    }
  }
}

// ------------------------------------------------------------------
// x axis: real rows of length n = 2M <-> Hermitian half rows of nh = M+1,
// through one complex FFT of length M on z[m] = x[2m] + i x[2m+1].
struct RowGeom {
  long nrows;      // rows processed (row r: spectral offset r*nhp, physical r*n)
  int nhp;         // spectral row pitch
  int n;           // physical row length / pitch
  int kin;         // spectral inputs read only for kx <= kin
  long s_comp_spec, s_comp_phys;
  double scale;
};

constexpr int pow2_ceil(int v) {
  int p = 1;
  while (p < v) p *= 2;
  return p;
}

#ifndef SC_ROWS_THREADS
#define SC_ROWS_THREADS 512
#endif
template <int M> constexpr int rows_per_block() {
  int tpp = Plan<M>::tpp;
  int r = SC_ROWS_THREADS / tpp;
  return r < 1 ? 1 : r;
}

// Build the length-M FFT input of a C2R row from the half spectrum X[0..M]:
// Zin[k] = (X[k] + conj X[M-k]) + i e^{+2 pi i k/n} (X[k] - conj X[M-k]).
// The imaginary parts of X[0] and X[M] are discarded (the convention of every
// multi-dimensional C2R for its halved axis).
template <typename T, int M, class Get>
DEV C<T> c2r_pre(int k, const C<T>* __restrict__ twn, Get&& get) {
  C<T> a = get(k);
  C<T> b = get(M - k);
  if (k == 0) { a.y = T(0); b.y = T(0); }
  b = conjc<T>(b);
  C<T> w = conjc<T>(twn[k]);  // exp(+2 pi i k / n)
  C<T> s = add<T>(a, b), d = sub<T>(a, b);
  return add<T>(s, muli<1, T>(cmul<T>(w, d)));
}

// R2C post-processing: X[k] = 1/2 (Z[k] + conj Z[M-k]) - i/2 W^k (Z[k] - conj Z[M-k]),
// W = exp(-2 pi i / n), Z[M] = Z[0].
template <typename T, int M>
DEV C<T> r2c_post(int k, const C<T>* zl, const C<T>* __restrict__ twn) {
  C<T> a = zl[k == M ? 0 : k];
  C<T> b = conjc<T>(zl[k == 0 ? 0 : M - k]);
  C<T> s = add<T>(a, b), d = sub<T>(a, b);
  C<T> wd = cmul<T>(twn[k], d);
  C<T> r = sub<T>(s, muli<1, T>(wd));
  return scl<T>(r, T(0.5));
}

// r2c_post for one output from its two entries: X = 1/2 (za + conj zb) -
// i/2 w (za - conj zb), za = Z[k], zb = Z[M - k] (Z[M] = Z[0]), w = W^k
template <typename T>
DEV C<T> r2c_pair(C<T> a, C<T> zb, C<T> w) {
  C<T> b = conjc<T>(zb);
  C<T> s = add<T>(a, b), d = sub<T>(a, b);
  C<T> wd = cmul<T>(w, d);
  C<T> r = sub<T>(s, muli<1, T>(wd));
  return scl<T>(r, T(0.5));
}

template <typename T, int M>
__global__ void __launch_bounds__((rows_per_block<M>() * Plan<M>::tpp))
k_x_c2r(const C<T>* __restrict__ in, T* out, const C<T>* __restrict__ twm,
        const C<T>* __restrict__ twn, RowGeom g) {
  constexpr int TY = rows_per_block<M>();
  using PE = Pencil<T, M, +1>;
  __shared__ C<T> lds[TY * M];
  const int t = threadIdx.x;
  const int ry = t / PE::TPP, p = t % PE::TPP;
  const long row = long(blockIdx.x) * TY + ry;
  const bool rok = (ry < TY) && row < g.nrows;
  const C<T>* src = in + long(blockIdx.y) * g.s_comp_spec + row * g.nhp;
  T* dst = out + long(blockIdx.y) * g.s_comp_phys + row * g.n;
  C<T> v[PE::MAXE];
  PE::load(v, p, [&](int k) {
    if (!rok) return mk<T>(T(0), T(0));
    return c2r_pre<T, M>(k, twn, [&](int kk) { return kk <= g.kin ? src[kk] : mk<T>(T(0), T(0)); });
  });
  PE::run(v, p, lds + ry * M, 1, twm);
  if (rok) {
    const T s = T(g.scale);
    typedef C<T> C2;
    PE::store(v, p, [&](int m, C2 z) { reinterpret_cast<C2*>(dst)[m] = scl<T>(z, s); });
  }
}

template <typename T, int M>
__global__ void __launch_bounds__((rows_per_block<M>() * Plan<M>::tpp))
k_x_r2c(const T* __restrict__ in, C<T>* out, const C<T>* __restrict__ twm,
        const C<T>* __restrict__ twn, RowGeom g) {
  constexpr int TY = rows_per_block<M>();
  using PE = Pencil<T, M, -1>;
  __shared__ C<T> lds[TY * M];
  const int t = threadIdx.x;
  const int ry = t / PE::TPP, p = t % PE::TPP;
  const long row = long(blockIdx.x) * TY + ry;
  const bool rok = (ry < TY) && row < g.nrows;
  const T* src = in + long(blockIdx.y) * g.s_comp_phys + row * g.n;
  C<T>* dst = out + long(blockIdx.y) * g.s_comp_spec + row * g.nhp;
  C<T> v[PE::MAXE];
  PE::load(v, p, [&](int m) {
    return rok ? reinterpret_cast<const C<T>*>(src)[m] : mk<T>(T(0), T(0));
  });
  PE::run(v, p, lds + ry * M, 1, twm);
  C<T>* zl = lds + ry * M;
  PE::out_to_lds(v, p, zl, 1);
  __syncthreads();
  if (rok) {
    const T s = T(g.scale);
    for (int k = p; k <= M; k += PE::TPP) dst[k] = scl<T>(r2c_post<T, M>(k, zl, twn), s);
  }
}

// ------------------------------------------------------------------
// Fused x-pass of the nonlinear term: per (y,z) row, C2R of the D velocity
// components (inputs band-limited to kx <= kin), the D(D+1)/2 pointwise
// products v_i v_j, and R2C of each product (times scale).  Also the
// row-wise max of |v| for the CFL time step.
template <typename T, int M, int D>
__global__ void SC_LB_X((rows_per_block<M>() * Plan<M>::tpp))
k_x_products(const C<T>* __restrict__ in, C<T>* out, const C<T>* __restrict__ twm,
             const C<T>* __restrict__ twn, RowGeom g, unsigned long long* vmax_bits) {
  constexpr int TY = rows_per_block<M>();
  constexpr int NP = D == 3 ? 6 : 3;
  using PI = Pencil<T, M, +1>;
  using PF = Pencil<T, M, -1>;
  constexpr int E = PlanInfo<M>::E_LAST;
  __shared__ C<T> lds[TY * M];
  __shared__ double red[TY * Plan<M>::tpp];
  const int t = threadIdx.x;
  const int ry = t / PI::TPP, p = t % PI::TPP;
  const long row = long(blockIdx.x) * TY + ry;
  const bool rok = (ry < TY) && row < g.nrows;
  C<T>* ll = lds + ry * M;
  C<T> u[D][E];
  C<T> v[PlanInfo<M>::MAXE];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const C<T>* src = in + long(c) * g.s_comp_spec + row * g.nhp;
    PI::load(v, p, [&](int k) {
      if (!rok) return mk<T>(T(0), T(0));
      return c2r_pre<T, M>(k, twn, [&](int kk) { return kk <= g.kin ? src[kk] : mk<T>(T(0), T(0)); });
    });
    PI::run(v, p, ll, 1, twm);
#pragma unroll
    for (int e = 0; e < E; ++e) u[c][e] = v[e];
  }
  // CFL: max |v| over the row's points
  double vm = 0.0;
  PI::for_out(p, [&](int e, int) {
    double a = 0, b = 0;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      a += double(u[c][e].x) * double(u[c][e].x);
      b += double(u[c][e].y) * double(u[c][e].y);
    }
    vm = fmax(vm, fmax(a, b));
  });
  if (vmax_bits != nullptr) {
    red[t] = rok ? vm : 0.0;
    __syncthreads();
    constexpr int NTH = TY * Plan<M>::tpp;
    constexpr int P2 = pow2_ceil(NTH);
    for (int s = P2 / 2; s > 0; s >>= 1) {
      if (t < s && t + s < NTH) red[t] = fmax(red[t], red[t + s]);
      __syncthreads();
    }
    if (t == 0) atomicMax(vmax_bits, (unsigned long long)__double_as_longlong(sqrt(red[0])));
  }
  constexpr int PA[6] = {0, 1, 2, 0, 1, 2};
  constexpr int PB[6] = {0, 1, 2, 1, 2, 0};
  constexpr int QA[3] = {0, 1, 0};
  constexpr int QB[3] = {0, 1, 1};
  const T s = T(g.scale);
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const int i = D == 3 ? PA[q] : QA[q];
    const int j = D == 3 ? PB[q] : QB[q];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = mk<T>(u[i][e].x * u[j][e].x, u[i][e].y * u[j][e].y);
    PI::out_to_in(v, p, ll, 1);
    PF::run(v, p, ll, 1, twm);
    PF::out_to_lds(v, p, ll, 1);
    __syncthreads();
    if (rok) {
      C<T>* dst = out + long(q) * g.s_comp_spec + row * g.nhp;
      for (int k = p; k <= M; k += PI::TPP) dst[k] = scl<T>(r2c_post<T, M>(k, ll, twn), s);
    }
  }
}

// ------------------------------------------------------------------
// The same fused x-pass with each row held by one wavefront: every LDS
// exchange is wavefront-local (no block barrier after the twiddle staging),
// so the wavefronts of a CU progress independently and hide each other's
// latency.  When the pencil's thread count divides 64, a wavefront holds
// 64 / TPP rows side by side (lane = row slot * TPP + p): all its lanes
// transform (at M = 64, 128, 256 one row would leave 56, 48 or 32 of them
// idle); otherwise it holds one row, and lanes beyond TPP stay idle in the
// transforms and share the R2C post-processing.
#ifndef SC_XW_WAVES
#define SC_XW_WAVES 2
#endif
template <int M> constexpr int xw_rows_per_wave() {
  return 64 % Plan<M>::tpp == 0 ? 64 / Plan<M>::tpp : 1;
}
template <typename T, int M, int D>
__global__ void __launch_bounds__(64 * SC_XW_WAVES)
k_x_products_w(const C<T>* __restrict__ in, C<T>* out, const C<T>* __restrict__ twm,
               const C<T>* __restrict__ twn, RowGeom g, unsigned long long* vmax_bits) {
  constexpr int TPP = Plan<M>::tpp;
  static_assert(TPP <= 64, "one wavefront per row");
  constexpr int NW = SC_XW_WAVES;
  constexpr int RPW = xw_rows_per_wave<M>();
  constexpr int LPR = 64 / RPW;  // lanes per row slot
  constexpr int NP = D == 3 ? 6 : 3;
  using PI = Pencil<T, M, +1, TPP, true>;
  using PF = Pencil<T, M, -1, TPP, true>;
  // forward transforms with the radices reversed read the products in the
  // registers where the inverse transforms left the velocities: no LDS
  // reordering between the pointwise product and the forward FFT
  using PFR = Pencil<T, M, -1, TPP, true, PlanRev<M>>;
  constexpr bool DIRECT = M / PlanInfo<M>::R(PlanInfo<M>::NS - 1) == TPP;
  constexpr int E = PlanInfo<M>::E_LAST;
  __shared__ C<T> lds[NW * RPW * M];
  __shared__ C<T> twm_l[M];
  __shared__ C<T> twn_l[M + 1];
  __shared__ double red[NW];
  for (int i = threadIdx.x; i < M; i += blockDim.x) twm_l[i] = twm[i];
  for (int i = threadIdx.x; i <= M; i += blockDim.x) twn_l[i] = twn[i];
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int slot = lane / LPR, q = lane % LPR;  // row slot, lane within it
  const bool act = q < TPP;
  const int p = act ? q : 0;
  const long row = (long(blockIdx.x) * NW + w) * RPW + slot;
  const bool rok = row < g.nrows;
  C<T>* ll = lds + (w * RPW + slot) * M;
  C<T> u[D][E];
  C<T> v[PlanInfo<M>::MAXE > PFR::MAXE ? PlanInfo<M>::MAXE : PFR::MAXE];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const C<T>* src = in + long(c) * g.s_comp_spec + row * g.nhp;
    PI::load(v, p, [&](int k) {
      if (!rok) return mk<T>(T(0), T(0));
      return c2r_pre<T, M>(k, twn_l, [&](int kk) { return kk <= g.kin ? src[kk] : mk<T>(T(0), T(0)); });
    });
    PI::run(v, p, ll, 1, twm_l);
#pragma unroll
    for (int e = 0; e < E; ++e) u[c][e] = v[e];
  }
  if (vmax_bits != nullptr) {
    double vm = 0.0;
    PI::for_out(p, [&](int e, int) {
      double a = 0, b = 0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        a += double(u[c][e].x) * double(u[c][e].x);
        b += double(u[c][e].y) * double(u[c][e].y);
      }
      vm = fmax(vm, fmax(a, b));
    });
    if (!rok) vm = 0.0;
    for (int o = 32; o > 0; o >>= 1) vm = fmax(vm, __shfl_xor(vm, o));
    if (lane == 0) red[w] = vm;
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = red[0];
      for (int i = 1; i < NW; ++i) m = fmax(m, red[i]);
      atomicMax(vmax_bits, (unsigned long long)__double_as_longlong(sqrt(m)));
    }
  }
  constexpr int PA[6] = {0, 1, 2, 0, 1, 2};
  constexpr int PB[6] = {0, 1, 2, 1, 2, 0};
  constexpr int QA[3] = {0, 1, 0};
  constexpr int QB[3] = {0, 1, 1};
  const T s = T(g.scale);
#pragma unroll
  for (int qq = 0; qq < NP; ++qq) {
    const int i = D == 3 ? PA[qq] : QA[qq];
    const int j = D == 3 ? PB[qq] : QB[qq];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = mk<T>(u[i][e].x * u[j][e].x, u[i][e].y * u[j][e].y);
    if constexpr (DIRECT) {
      static_assert(PFR::E0 == E, "reversed plan reads the inverse's output slots");
      PFR::run(v, p, ll, 1, twm_l);
      PFR::out_to_lds(v, p, ll, 1);
    } else {
      PI::out_to_in(v, p, ll, 1);
      PF::run(v, p, ll, 1, twm_l);
      PF::out_to_lds(v, p, ll, 1);
    }
    xsync<true>();
    if (rok) {
      C<T>* dst = out + long(qq) * g.s_comp_spec + row * g.nhp;
      for (int k = q; k <= M; k += LPR) dst[k] = scl<T>(r2c_post<T, M>(k, ll, twn_l), s);
    }
  }
}

// Persistent form of k_x_products_w for rows whose inputs are band-limited to
// kx <= kin < xp_entries<T>(): wavefront gw walks rows gw, gw + W, ... (W
// waves in the grid).  Each velocity row is staged in LDS with 16-byte direct
// global->LDS loads issued one transform ahead of its use (the next component
// while the current one is transformed, the next row's first component while
// the current row's products are transformed), so no register holds an
// in-flight load and the loads' latency overlaps the FFTs.  The staging row is
// zero beyond kin (entries above kin are loaded from a zero source, those past
// xp_entries<T>() are cleared once; FP32 loads carry two entries, so entry
// kin + 1 may hold the row's own value and is masked where it is read), so
// the C2R pre-processing reads it without bounds tests.
#ifndef SC_XP_WAVES
#define SC_XP_WAVES 4
#endif
// entries per 16-byte load, chunks of 64 loads per staged row, entries staged
template <typename T> constexpr int xp_per_load() { return 16 / int(sizeof(C<T>)); }
template <typename T> constexpr int xp_chunks() { return sizeof(T) == 8 ? 5 : 3; }
template <typename T> constexpr int xp_entries() { return 64 * xp_chunks<T>() * xp_per_load<T>(); }

// stg[k] = k <= kin ? src[k] : 0 for k < xp_entries<T>() (FP32: the pair
// holding kin is loaded whole); lane l of chunk j fills the 16 bytes at
// stg + per-load (64 j + l) (a direct load writes the wave's base + 16 l)
template <typename T>
DEV void xp_stage(const C<T>* src, const C<T>* zero, C<T>* stg, int kin, int lane) {
  constexpr int PL = xp_per_load<T>();
  // the wave's earlier LDS reads of stg have returned (lgkmcnt(0))
  __builtin_amdgcn_s_waitcnt(0xC07F);
#pragma unroll
  for (int j = 0; j < xp_chunks<T>(); ++j) {
    const int k = (j * 64 + lane) * PL;
    const C<T>* a = k <= kin ? src + k : zero;
    __builtin_amdgcn_global_load_lds((const void*)a, (__attribute__((address_space(3))) void*)(stg + j * 64 * PL), 16,
                                     0, 0);
  }
}

template <typename T, int M, int D>
__global__ void __launch_bounds__(64 * SC_XP_WAVES, 8 / SC_XP_WAVES)
k_x_products_p(const C<T>* __restrict__ in, C<T>* out, const C<T>* __restrict__ twm, const C<T>* __restrict__ twn,
               RowGeom g, unsigned long long* vmax_bits) {
  constexpr int TPP = Plan<M>::tpp;
  static_assert(TPP <= 64, "one wavefront per row");
  static_assert(xp_entries<T>() <= (M + 1 + 7) / 8 * 8, "staged chunks fit the staging row");
  constexpr int NW = SC_XP_WAVES;
  constexpr int NP = D == 3 ? 6 : 3;
  constexpr int XS = (M + 1 + 7) / 8 * 8;  // staging row: entries 0..M
  using PI = Pencil<T, M, +1, TPP, true>;
  using PF = Pencil<T, M, -1, TPP, true>;
  using PFR = Pencil<T, M, -1, TPP, true, PlanRev<M>>;
  constexpr bool DIRECT = M / PlanInfo<M>::R(PlanInfo<M>::NS - 1) == TPP;
  constexpr int E = PlanInfo<M>::E_LAST;
  __shared__ C<T> lds[NW * M];
  __shared__ C<T> stg[NW * XS];
  __shared__ C<T> twm_l[M];
  __shared__ C<T> twn_l[M + 1];
  for (int i = threadIdx.x; i < M; i += blockDim.x) twm_l[i] = twm[i];
  for (int i = threadIdx.x; i <= M; i += blockDim.x) twn_l[i] = twn[i];
  for (int i = threadIdx.x; i < NW * XS; i += blockDim.x) stg[i] = mk<T>(T(0), T(0));
  __syncthreads();
  const C<T>* zero = twn + 2 * M;
  const int w = threadIdx.x >> 6, lane0 = threadIdx.x & 63;
  const int p0 = lane0 < TPP ? lane0 : 0;
  C<T>* ll = lds + w * M;
  C<T>* sw = stg + w * XS;
  const long W = long(gridDim.x) * NW;
  const int kin = g.kin;
  long row = long(blockIdx.x) * NW + w;
  if (row < g.nrows) xp_stage<T>(in + row * g.nhp, zero, sw, kin, lane0);
  double vm = 0.0;
  constexpr int PA[6] = {0, 1, 2, 0, 1, 2};
  constexpr int PB[6] = {0, 1, 2, 1, 2, 0};
  constexpr int QA[3] = {0, 1, 0};
  constexpr int QB[3] = {0, 1, 1};
  const T s = T(g.scale);
  for (; row < g.nrows; row += W) {
    // opaque copies of the lane indices: every per-lane address is formed
    // inside the loop instead of being kept live across it
    int p = p0, lane = lane0;
    asm volatile("" : "+v"(p), "+v"(lane));
    // only the TPP lanes of the pencil transform; the others would repeat
    // lane 0's LDS writes, which the LDS serialises as bank conflicts
    const bool act = lane < TPP;
    C<T> u[D][E];
    C<T> v[PlanInfo<M>::MAXE > PFR::MAXE ? PlanInfo<M>::MAXE : PFR::MAXE];
    const long nxt = row + W;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      if (act)
        PI::load(v, p, [&](int k) {
          return c2r_pre<T, M>(k, twn_l, [&](int kk) {
            if constexpr (xp_per_load<T>() > 1) return kk <= kin ? sw[kk] : mk<T>(T(0), T(0));
            else return sw[kk];
          });
        });
      // the staging row is free: the next component, or the next row's first
      // (after the last row, this row's again: unused, but every wave issues
      // the same loads, so the waits can count them)
      if (c + 1 < D) xp_stage<T>(in + long(c + 1) * g.s_comp_spec + row * g.nhp, zero, sw, kin, lane);
      else xp_stage<T>(in + (nxt < g.nrows ? nxt : row) * g.nhp, zero, sw, kin, lane);
      if (act) {
        PI::run(v, p, ll, 1, twm_l);
#pragma unroll
        for (int e = 0; e < E; ++e) u[c][e] = v[e];
      }
    }
    if (vmax_bits != nullptr && act) {
      PI::for_out(p, [&](int e, int) {
        double a = 0, b = 0;
#pragma unroll
        for (int c = 0; c < D; ++c) {
          a += double(u[c][e].x) * double(u[c][e].x);
          b += double(u[c][e].y) * double(u[c][e].y);
        }
        vm = fmax(vm, fmax(a, b));
      });
    }
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      const int i = D == 3 ? PA[q] : QA[q];
      const int j = D == 3 ? PB[q] : QB[q];
      if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = mk<T>(u[i][e].x * u[j][e].x, u[i][e].y * u[j][e].y);
        if constexpr (DIRECT) {
          PFR::run(v, p, ll, 1, twm_l);
          PFR::out_to_lds(v, p, ll, 1);
        } else {
          PI::out_to_in(v, p, ll, 1);
          PF::run(v, p, ll, 1, twm_l);
          PF::out_to_lds(v, p, ll, 1);
        }
      }
      xsync<true>();
      C<T>* dst = out + long(q) * g.s_comp_spec + row * g.nhp;
      // X[k] and X[M - k] from the same two entries Z[k], Z[M - k] (Z[M] =
      // Z[0]); lanes past k = M / 2 store that pair again (same addresses,
      // same values), so every store is issued and the waits can count them.
      // Not unrolled: the unrolled form needs more registers than the
      // wavefront has next to the three live velocities (FP64: 11 spilled)
#pragma unroll 1
      for (int t = 0; t < (M / 2 + 64) / 64; ++t) {
        const int k = min(lane + 64 * t, M / 2);
        const C<T> za = ll[k], zb = ll[k == 0 ? 0 : M - k];
        dst[k] = scl<T>(r2c_pair<T>(za, zb, twn_l[k]), s);
        dst[M - k] = scl<T>(r2c_pair<T>(zb, za, twn_l[M - k]), s);
      }
    }
  }
  if (vmax_bits != nullptr) {
    for (int o = 32; o > 0; o >>= 1) vm = fmax(vm, __shfl_xor(vm, o));
    if (lane0 == 0 && vm > 0.0) atomicMax(vmax_bits, (unsigned long long)__double_as_longlong(sqrt(vm)));
  }
}

}  // namespace sc
