// Learned closure: 3x3x3 circular-padding convolutional network on BF16 MFMA
// (v_mfma_f32_16x16x32_bf16) with FP32 accumulation and FP32 master weights.
//
// Architecture: cin -> W -> W -> W -> cout, GELU (tanh form) after the three
// hidden layers, no bias on the last layer (the reference's pointwise nets,
// src/nets.jl:247-298, use the same bias/activation pattern with 1x1 kernels).
//
// Layouts (all tensors channels-last, one batch of B cubes of side n):
//   activations  bf16 [B][n][n][n][C]   C padded to a multiple of 32
//   weights      fp32 master [27][Cout_p][Cin_p] (padding entries stay zero: their
//                inputs / output gradients are zero); bf16 forward copy [27][Cout_p][Cin_p];
//                bf16 transposed copy [27][Cin_p][Cout_p] with the taps mirrored
//                (backward data = convolution with flipped taps)
//   tap t = (dz+1)*9 + (dy+1)*3 + (dx+1): output voxel v reads input v + (dz,dy,dx)
//
// Convolution = implicit GEMM: M = voxels (64 per block, 16 per wavefront),
// N = output channels, K = 27 taps x input channels.  The 9 (dz, dy) input
// rows of a 64-voxel output segment (with one halo voxel each side) are
// staged in LDS; weights stream from L2.
#include <type_traits>
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "../../include/symcode.h"
#include "cnn_link.hpp"
#include "optim.hpp"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

namespace cnn {

#define CNN_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// periodic index for v in [-n, 2n)
__device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// tanh-form GELU, 0.5 z (1 + tanh(u)) with u = sqrt(2/pi) (z + 0.044715 z^3),
// evaluated as z * sigmoid(2u) (the same function) with the hardware exp2 and
// reciprocal: relative error ~1e-7, far below the bf16 storage it feeds
__device__ __forceinline__ float sig2u(float z) {
  const float c2 = -2.0f * 0.7978845608028654f * 1.4426950408889634f;  // -2 sqrt(2/pi) log2(e)
  const float u = z * __builtin_fmaf(0.044715f * z, z, 1.0f);
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(c2 * u));
}
__device__ __forceinline__ float gelu(float z) { return z * sig2u(z); }
// d/dz: s + 2 z s (1 - s) sqrt(2/pi) (1 + 3 * 0.044715 z^2), s = sigmoid(2u)
__device__ __forceinline__ float gelu_grad(float z) {
  const float s = sig2u(z);
  return s + 2.0f * 0.7978845608028654f * z * s * (1.0f - s) * __builtin_fmaf(0.134145f * z, z, 1.0f);
}

enum Mode { HIDDEN = 0, LAST = 1, BWD_HIDDEN = 2, BWD_INPUT = 3 };

constexpr int VPB = 64;   // output voxels per x segment (n must be a multiple)
#ifndef SC_CNN_PREFETCH
#define SC_CNN_PREFETCH 0
#endif
#ifndef SC_CNN_TROWS
#define SC_CNN_TROWS 4
#endif
// output y rows per block: 4 rows (4 wavefronts, 57-61 KB of LDS for a
// 64-channel input plane) let two blocks share a compute unit, so one
// block's plane fetch, plane commits and epilogue overlap the other's MFMAs
// (with 8 rows, 95-102 KB, a unit holds one block and those stalls are
// exposed: the 16-column last layer took 2.99 against 2.78 ms with 8 rows,
// for all its smaller y halo)
constexpr int conv_trows(int) { return SC_CNN_TROWS; }
// y rows per wavefront.  RPW = 2 halves the weight-fragment traffic per flop
// but leaves one wavefront per SIMD (the LDS tile allows one block per CU):
// batch-16 x 128^3 training steps took 87.8 ms against 80.3 ms with RPW = 1
#ifndef SC_CNN_RPW
#define SC_CNN_RPW 1
#endif
constexpr int RPW = SC_CNN_RPW;
constexpr int conv_threads(int coutt) { return 64 * (conv_trows(coutt) / RPW); }  // threads per block
// Staged input rows, chunk-major: for each (plane, row, 8-channel chunk c8)
// the row's voxels are 16-byte entries, voxel x at position x ^ (c8 & 7) of
// a run of XWP = 80 entries (1280 bytes, a multiple of 256).  Measured on
// this chip with a one-wavefront probe and the bank-conflict counter,
// ds_read_b128 is conflict-free when the four 16-lane groups of the wave
// start at the same offset modulo 256 bytes (aligned or not, contiguous or
// strided within a group), and conflicts 2-way when they do not (groups
// 1168 or 272 bytes apart).  An A-fragment read (group lk = chunk c8 of 16
// consecutive voxels) covers the 16 slots of a 256-byte window once (the
// XOR permutes within aligned 8-entry blocks) at congruent offsets, and the
// staging stores (8 consecutive lanes = the 8 chunks of one voxel) land in 8
// distinct 16-byte slots modulo 128 bytes, as ds_write_b128 needs.
constexpr int XWP = 80;

constexpr int kpad(int cin) { return (27 * cin + 31) / 32 * 32; }

#ifndef SC_CNN_SPLIT
#define SC_CNN_SPLIT 1  // two rows x half the output tiles per wavefront (see k_conv)
#endif

// One convolution layer (or its data gradient) as an implicit GEMM with the
// reduction index k = t * CIN + c (tap-major; narrow layers pack several taps
// into one 32-wide MFMA step).  Block: 4 consecutive y rows of one z plane,
// 64 x voxels each, RPW rows per wavefront; a wavefront holds its
// (64 RPW)-voxel x COUTT output tile in 4 RPW x (COUTT/16) MFMA accumulators,
// so every B fragment feeds 4 RPW MFMAs and every A fragment COUTT/16 (with
// RPW = 2 the block's weight-fragment traffic through L1 is half that of one
// row per wavefront, for the same LDS tile).  Inputs are staged in LDS: one input
// plane (10 rows x 66 voxels) at a time for CIN = 64, all three planes for the
// narrow layers.  Weights are pre-arranged per (32-wide k step, 16 output
// channels) as the 64 lanes' B fragments (1 KB contiguous, see
// k_pack_weights), read through L1, two k steps ahead.  The next input plane
// is fetched into registers while the current one's MFMAs run.
//  HIDDEN:     z = conv + bias; out_z = z, out_a = gelu(z)            (bf16)
//  LAST:       out_f = conv, with tracefree: channels 0-2 (xx, yy, zz)
//              minus their mean                                       (fp32)
//  BWD_HIDDEN: out_a = conv * gelu'(zprev)                            (bf16)
//  BWD_INPUT:  out_f = conv (first COUTS channels)                    (fp32)
template <int CIN, int COUTT, int COUTS, int MODE>
__global__ void __launch_bounds__(conv_threads(COUTT), 8 / (conv_threads(COUTT) / 64)) k_conv(const bf16* __restrict__ in, const bf16* __restrict__ W,
                                                 const float* __restrict__ bias, bf16* out_a, bf16* out_z,
                                                 float* out_f, const bf16* __restrict__ zprev, int n, long ntiles,
                                                 int tracefree) {
  constexpr int TROWS = conv_trows(COUTT), CONV_THREADS = conv_threads(COUTT);
  constexpr int RW = TROWS + 2, XW = VPB + 2;
  static_assert(XWP >= (XW | 7) + 1 && XWP % 16 == 0, "staged row width");
  constexpr bool BYPLANE = CIN >= 32;
  constexpr int NPL = BYPLANE ? 1 : 3;
  constexpr int NKC = kpad(CIN) / 32;
  // with 4 output tiles (and one row per wavefront otherwise) a wavefront
  // takes two rows and half the output tiles: the same MFMAs per k step,
  // half the weight-fragment traffic through L1 (8 wavefronts of a block
  // request the same fragments), twice the A reads from LDS
  constexpr bool SPLIT = SC_CNN_SPLIT && COUTT == 64 && RPW == 1;
  // PAIR (the last layer with at most 8 output channels): the 16 rows of an
  // output tile hold 8 channels of two consecutive output rows; a k step
  // pairs one staged input row with the taps dy (rows 0-7) and dy - 1 (rows
  // 8-15) that map it onto those two rows, so one plane takes 4 input rows x
  // 3 dx x 2 channel halves = 24 k steps for two rows (36 for two rows with
  // 8 of 16 MFMA rows used).  Wavefront w takes rows 2 (w / 2), + 1 and the
  // x half w % 2 of the block's 64 voxels (two M tiles)
  constexpr bool PAIR = MODE == LAST && COUTS == 8 && CIN == 64 && RPW == 1;
  constexpr int NT = SPLIT ? 2 : COUTT / 16;   // output tiles per wavefront
  constexpr int NTA = COUTT / 16;              // output tiles in the weight layout
  constexpr int CH8 = CIN / 8;
  constexpr int MT = SPLIT ? 8 : PAIR ? 2 : 4 * RPW;  // M tiles of 16 voxels: 4 per row
  constexpr int NPH = BYPLANE ? 3 : 1;
  constexpr int NCH = NPL * RW * XW * CH8;
  constexpr int NPT = (NCH + CONV_THREADS - 1) / CONV_THREADS;
  // HIDDEN and BWD_HIDDEN with SPLIT: the epilogue passes the block's
  // outputs through the same LDS (two row pairs x 128 voxels x 64 channels)
  constexpr bool LINES = (MODE == HIDDEN || MODE == BWD_HIDDEN) && SPLIT;
  constexpr int WIN = NPL * RW * CH8 * XWP * 8;
  constexpr int EPI = LINES ? 2 * 128 * 64 : 0;
  __shared__ __attribute__((aligned(16))) bf16 win[WIN > EPI ? WIN : EPI];
  const int segs = n / VPB, ytiles = n / TROWS;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int li = l & 15, lk = l >> 4;
  const int wr = (SPLIT || PAIR) ? 2 * (w / 2) : w * RPW;  // the wavefront's first row
  const int wn = SPLIT ? 2 * (w % 2) : 0;                  // and first output tile
  const int wx = PAIR ? 32 * (w % 2) : 0;                  // and first voxel of its x range
  struct Tile { long b; int z, y0, x0; };
  auto decode = [&](long t) {
    Tile T;
    T.x0 = int(t % segs) * VPB;
    t /= segs;
    T.y0 = int(t % ytiles) * TROWS;
    t /= ytiles;
    T.z = int(t % n);
    T.b = t / n;
    return T;
  };
  floatx4 acc[MT][NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = floatx4{0, 0, 0, 0};
  // A fragment of M tile mt (row wr + mt / 4, x 16 (mt % 4) + li) at k step
  // kc: lane (li, lk) reads chunk lk of the k step's 32 channels; the
  // swizzle of x only touches its low three bits, so the M tiles' offsets
  // stay constants
  const bf16* lane_a = &win[lk * XWP * 8];
  auto mt_off = [](int mt) { return ((mt / 4) * CH8 * XWP + 16 * (mt % 4)) * 8; };
  auto load_a = [&](int kc, bf16x8* a) {
    if constexpr (PAIR) {  // k step within the plane: (staged row offset ro, dx, channel half)
      const int ks = kc % 24, ro = ks / 6, dx = (ks / 2) % 3 - 1, cb = (ks % 2) * 32;
      const int c8 = cb / 8 + lk;
      const bf16* base = lane_a + (((wr + ro) * CH8 + cb / 8) * XWP + ((1 + dx + li) ^ (c8 & 7)) + wx) * 8;
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) a[mt] = *reinterpret_cast<const bf16x8*>(base + mt_off(mt));
      return;
    }
    if constexpr (CIN >= 32) {  // one tap per k step: the row offsets are uniform
      const int t = (kc * 32) / CIN, cb = (kc * 32) % CIN;
      const int dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
      const int c8 = cb / 8 + lk;
      const bf16* base = lane_a + (((wr + 1 + dy) * CH8 + cb / 8) * XWP + ((1 + dx + li) ^ (c8 & 7))) * 8;
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) a[mt] = *reinterpret_cast<const bf16x8*>(base + mt_off(mt));
      return;
    }
    const int k0 = kc * 32 + 8 * lk;
    const int t = k0 / CIN, c0 = k0 % CIN;
    if (t < 27) {
      const int pl = BYPLANE ? 0 : t / 9;
      const int dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
      const int c8 = c0 / 8;
      const bf16* base = &win[(((pl * RW + wr + 1 + dy) * CH8 + c8) * XWP + ((1 + li + dx) ^ (c8 & 7))) * 8];
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) a[mt] = *reinterpret_cast<const bf16x8*>(base + mt_off(mt));
    } else {
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int e = 0; e < 8; ++e) a[mt][e] = bf16(0.f);
    }
  };
  auto load_b = [&](int kc, bf16x8* bv) {
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
      bv[nt] = *reinterpret_cast<const bf16x8*>(&W[(long(kc * NTA + wn + nt) * 64 + l) * 8]);
  };
  // staging of (tile T, phase PH): all of a phase's 16-byte chunks are fetched
  // into registers at once, and written to LDS between two barriers
  u32x4 pre[NPT];
#define CONV_FETCH(T, PH)                                                                                   \
  _Pragma("clang loop unroll(full)") for (int j = 0; j < NPT; ++j) {                                        \
    const int i = min(int(threadIdx.x) + CONV_THREADS * j, NCH - 1);                                        \
    const int c8 = i % CH8;                                                                                 \
    int r = i / CH8;                                                                                        \
    const int xi = r % XW;                                                                                  \
    r /= XW;                                                                                                \
    const int row = r % RW, pl = r / RW;                                                                    \
    const int zs = wrap((T).z + (BYPLANE ? (PH) : pl) - 1, n), ys = wrap((T).y0 - 1 + row, n);              \
    const int xx = wrap((T).x0 - 1 + xi, n);                                                                \
    const unsigned vox = ((unsigned((T).b) * n + zs) * n + ys) * n + xx;                                   \
    pre[j] = reinterpret_cast<const u32x4*>(in)[vox * CH8 + c8];                                            \
  }
#define CONV_COMMIT()                                                                                       \
  _Pragma("clang loop unroll(full)") for (int j = 0; j < NPT; ++j) {                                        \
    const int i = threadIdx.x + CONV_THREADS * j;                                                           \
    if (i < NCH)                                                                                            \
      *reinterpret_cast<u32x4*>(&win[(((i / CH8) / XW * CH8 + i % CH8) * XWP + ((i / CH8) % XW ^ (i % CH8 & 7))) * 8]) = pre[j]; \
  }
  if (long(blockIdx.x) >= ntiles) return;
  const Tile cur = decode(blockIdx.x);
  CONV_FETCH(cur, 0)
  CONV_COMMIT()
  __syncthreads();
#pragma unroll 1
  for (int ph = 0; ph < NPH; ++ph) {
    // prefetch this tile's next input plane while this one is used (without
    // it, the plane is fetched at the commit, and the other block on the
    // compute unit covers the wait)
    if (SC_CNN_PREFETCH && ph + 1 < NPH) {
      CONV_FETCH(cur, ph + 1)
    }
    constexpr int NS = PAIR ? 24 : BYPLANE ? NKC / 3 : NKC;  // k steps of the phase
    const int kc0 = ph * NS;
    // software pipeline over the phase's k steps, fully unrolled (the step
    // count is a compile-time constant, so the code is straight-line): B
    // fragments are requested four steps ahead, A fragments one step ahead.
    // With the loads under run-time conditions the compiler either waits
    // with vmcnt(0) (draining the fragments requested for later steps too)
    // or sinks a load into the block that uses it
    bf16x8 af[2][MT], bf[4][NT];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (j < NS) load_b(kc0 + j, bf[j]);
    load_a(kc0, af[0]);
#pragma unroll
    for (int kk = 0; kk < NS; ++kk) {
      // the scheduler keeps each step's loads in that step (left to itself
      // it moves every load next to its first use, which undoes the pipeline)
      __builtin_amdgcn_sched_barrier(0);
      if (kk + 1 < NS) load_a(kc0 + kk + 1, af[(kk + 1) & 1]);
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf[kk & 3][nt], af[kk & 1][mt], acc[mt][nt], 0, 0, 0);
      if (kk + 4 < NS) load_b(kc0 + kk + 4, bf[kk & 3]);
    }
    if (ph + 1 < NPH) {
      if (!SC_CNN_PREFETCH) {
        CONV_FETCH(cur, ph + 1)
      }
      __syncthreads();
      CONV_COMMIT()
      __syncthreads();
    }
  }
  if constexpr (LINES) {
    // The block's 4 rows x 64 voxels x 64 channels go through LDS so that
    // every store of the epilogue is a whole 128-byte voxel line (8 lanes x
    // 16 bytes); from the MFMA layout each lane would store 8 bytes, a
    // quarter line per 16 lanes.  HIDDEN: z, then gelu(z).  BWD_HIDDEN:
    // conv * gelu'(zprev), zprev read in the MFMA layout (loading its lines
    // through LDS as well, after the last k step, was slower: the barrier
    // then waits for the loads).  Row pair wr / 2: voxel vloc = 64 (row - wr) + x of 128, a
    // line of 8 chunks of 8 channels, chunk q at position q ^ (vloc & 7)
    static_assert(TROWS == 4 && NT == 2 && MT == 8 && COUTS == 64, "block epilogue shape");
    uint2* ew = reinterpret_cast<uint2*>(win) + (w / 2) * 128 * 16;
    uint4* el = reinterpret_cast<uint4*>(win) + (w / 2) * 128 * 8;
    float bco[NT][4] = {};
    if constexpr (MODE == HIDDEN) {
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int r = 0; r < 4; ++r) bco[nt][r] = bias[(wn + nt) * 16 + 4 * lk + r];
    }
    // this lane's 8 bytes of (M tile mt, output tile nt)
    auto slot = [&](int mt, int nt) __attribute__((always_inline)) {
      const int vloc = 64 * (mt / 4) + 16 * (mt % 4) + li;
      const int q = ((wn + nt) * 2 + lk / 2) ^ (vloc & 7);
      return vloc * 16 + q * 2 + (lk & 1);
    };
    // the lane's first voxel in the MFMA layout (row wr, x li)
    const unsigned v0w = ((unsigned(cur.b) * n + cur.z) * n + cur.y0 + wr) * unsigned(n) + cur.x0 + li;
    auto put = [&](int act) __attribute__((always_inline)) {  // 0: z, 1: gelu(z), 2: conv * gelu'(zprev)
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          const int e = slot(mt, nt);
          bf16x4 o, zp;
          if (act == 2) {
            const unsigned v = v0w + (mt / 4) * unsigned(n) + 16 * (mt % 4);
            zp = *reinterpret_cast<const bf16x4*>(&zprev[v * 64u + (wn + nt) * 16 + 4 * lk]);
          }
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float d = acc[mt][nt][r];
            const float z = d + bco[nt][r];
            o[r] = bf16(act == 2 ? d * gelu_grad(float(zp[r])) : act == 1 ? gelu(z) : z);
          }
          uint2 u;
          __builtin_memcpy(&u, &o, 8);
          ew[e] = u;
        }
    };
    // wavefront w % 2 of the pair moves the lines of row wr + w % 2, 8 voxels per instruction
    const int rr = w % 2, p = l % 8;
    const unsigned v0 = ((unsigned(cur.b) * n + cur.z) * n + cur.y0 + wr + rr) * unsigned(n) + cur.x0;
    auto drain = [&](bf16* out) __attribute__((always_inline)) {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int x = 8 * j + l / 8, vloc = 64 * rr + x;
        *reinterpret_cast<uint4*>(&out[(v0 + x) * 64u + p * 8]) = el[vloc * 8 + (p ^ (vloc & 7))];
      }
    };
    if constexpr (MODE == HIDDEN) {
      __syncthreads();  // every wavefront is done with the staged input
      put(0);
      __syncthreads();
      drain(out_z);
      __syncthreads();
      put(1);
      __syncthreads();
      drain(out_a);
    } else {
      __syncthreads();  // every wavefront is done with the staged input
      put(2);
      __syncthreads();
      drain(out_a);
    }
    return;
  }
  // epilogue: with the weights as the MFMA's A operand, lane (li, lk) holds
  // D[co = 16 nt + 4 lk + r][voxel = 16 mt + li]: four consecutive channels of
  // one voxel, stored (and, for the data gradient, the pre-activation read)
  // as one 8- or 16-byte access
  // (PAIR: rows 0-7 of the tile are output row wr, rows 8-15 row wr + 1)
  const unsigned vrow0 = ((unsigned(cur.b) * n + cur.z) * n + cur.y0 + wr + (PAIR ? lk / 2 : 0)) * unsigned(n) +
                         cur.x0 + wx + li;
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    const int co0 = PAIR ? 4 * (lk % 2) : (wn + nt) * 16 + 4 * lk;
    if (co0 >= COUTS) continue;
    float bco[4] = {0.f, 0.f, 0.f, 0.f};
    if constexpr (MODE == HIDDEN) {
#pragma unroll
      for (int r = 0; r < 4; ++r) bco[r] = bias[co0 + r];
    }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      // voxel and element offsets fit in 32 bits (at most 2^31 elements per tensor)
      const unsigned v = vrow0 + (PAIR ? 0u : (mt / 4) * unsigned(n)) + 16 * (mt % 4);
      const floatx4 d = acc[mt][nt];
      if constexpr (MODE == HIDDEN) {
        bf16x4 zz, aa;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          zz[r] = bf16(d[r] + bco[r]);
          aa[r] = bf16(gelu(d[r] + bco[r]));
        }
        *reinterpret_cast<bf16x4*>(&out_z[v * COUTS + co0]) = zz;
        *reinterpret_cast<bf16x4*>(&out_a[v * COUTS + co0]) = aa;
      } else if constexpr (MODE == BWD_HIDDEN) {
        const bf16x4 zp = *reinterpret_cast<const bf16x4*>(&zprev[v * COUTS + co0]);
        bf16x4 o;
#pragma unroll
        for (int r = 0; r < 4; ++r) o[r] = bf16(d[r] * gelu_grad(float(zp[r])));
        *reinterpret_cast<bf16x4*>(&out_a[v * COUTS + co0]) = o;
      } else {
        floatx4 o = d;
        if constexpr (MODE == LAST) {
          // trace removal: lane lk = 0 of tile nt = 0 holds xx, yy, zz of its voxel
          if (tracefree && co0 == 0) {
            const float t = (o[0] + o[1] + o[2]) / 3.f;
            o[0] -= t;
            o[1] -= t;
            o[2] -= t;
          }
        }
        *reinterpret_cast<floatx4*>(&out_f[v * COUTS + co0]) = o;
      }
    }
  }
#undef CONV_FETCH
#undef CONV_COMMIT
}

// 8 x 8 transpose of 16-bit elements held as 8 rows of 4 dwords:
// out[c] (dwords 0..3 = columns 2j, 2j+1 of row pair) = column c of in
__device__ __forceinline__ void transpose8x8(const uint32_t in[8][4], uint32_t out[8][4]) {
#pragma unroll
  for (int c = 0; c < 8; ++c)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t a = in[2 * j][c >> 1], bb = in[2 * j + 1][c >> 1];
      // v_perm_b32 selector: bytes of {bb, a}; low half of a|b or high half
      out[c][j] = (c & 1) ? __builtin_amdgcn_perm(bb, a, 0x07060302u) : __builtin_amdgcn_perm(bb, a, 0x05040100u);
    }
}

// Weight gradient of a 64 -> 64 layer:
//   dW[t][co][ci] += sum_v dout[v][co] * in[v + (dz,dy,dx)][ci]
// as GEMMs with M = co, N = ci, K = the voxels of an x row.  Persistent
// blocks: each plane group of three blocks (dz = -1, 0, 1) takes the (batch,
// z) planes g, g + ngroups, ... whole, walking y with a sliding window of
// input rows, so each block keeps all nine (dy, dx) taps of its dz in
// registers (36 accumulator tiles per wavefront: all taps of a block in one
// block would not fit).  Block = 8 wavefronts: wavefronts 0-3 compute
// (wavefront w owns input channels 16 w .. 16 w + 15 and all output
// channels), wavefronts 4-7 stage rows.  Staging is software-pipelined:
// the producers load row sets PD rows ahead into registers and, while row y
// is computed, write the set for row y + 1 (dout row y + 1, input row y + 2)
// into the LDS slots not in use (one barrier per row).
//
// LDS layout: a row of a 16-channel tile is n / 8 blocks of 8 voxels x 16
// channels; a block is 16 chunks of 16 bytes, chunk = one channel's 8
// voxels, at chunk index swz(channel, block) = channel ^ ((4 (channel >> 3))
// ^ block) & 7); tiles start 16 bytes apart modulo 256.  An MFMA operand read
// (16 channels of one block per 16 lanes) is then one 256-byte block, which
// ds_read_b128 serves without bank conflicts, and the staging stores (8
// lanes = 8 channel blocks of one voxel block, or 8 voxel blocks of one
// channel block) hit 8 distinct 16-byte slots modulo 128 bytes, which
// ds_write_b128 needs.  Input rows carry one halo block on each side; the x
// shifts dx = -1, +1 are formed in registers from a block's chunk and one
// dword of each neighbouring block: the producers also write every block's
// first and last dword per channel into edge arrays xE[slot][tile][first /
// last][block][channel], where a wavefront's 64 lanes (16 channels x 4
// consecutive blocks) read 256 contiguous bytes (4-byte reads, conflict-free)
// instead of the neighbours' whole chunks.
// Block sums go to part[g][t][co][ci] (k_reduce_partials adds them up); with
// db set, the dz = 0 blocks also add the bias gradient sum_v dout[v][co].
__device__ __forceinline__ int wg_swz(int ch, int blk) { return ch ^ (((4 * (ch >> 3)) ^ blk) & 7); }

template <int CIN, int COUT, int PD>
__global__ void __launch_bounds__(512, 1) k_conv_wgrad(const bf16* __restrict__ in, const bf16* __restrict__ dout,
                                                       float* part, float* db, int n, long nplanes) {
  constexpr int NBM = 16;                      // x blocks per row (n <= 128)
  constexpr int MT = COUT / 16;                // output-channel tiles, all per compute wavefront
  constexpr int NW = CIN / 16;                 // 16-channel input tiles, one per compute wavefront
  static_assert(CIN == 64 && COUT == 64, "the 64 x 64 layer (the narrow layers use k_conv_wgrad3)");
  // tile strides in 16-byte chunks (+1: tiles start 16 bytes apart)
  constexpr int DTS = NBM * 16 + 1, XTS = (NBM + 2) * 16 + 1;
  // one staging job (8 voxels x 8 channels) per producer thread and row
  // (n <= 128): the dout row's blocks, then the input row's; the halos are
  // written by the jobs of x blocks 0 and n / 8 - 1
  static_assert(16 * (COUT / 8) + 16 * (CIN / 8) <= 256, "wgrad staging jobs");
  static_assert(64 % PD == 0, "row sets in flight divide n");
  __shared__ uint4 dT[2 * MT * DTS];  // dout rows y, y + 1
  __shared__ uint4 xT[4 * NW * XTS];  // input rows y - 1 .. y + 2
  constexpr int EBS = (NBM + 2) * 16;  // dwords per edge array of a row tile (blocks x channels)
  __shared__ uint32_t xE[4 * NW * 2 * EBS];  // [slot][tile][0: first, 1: last dword][block][channel]
  // block -> (plane group pg, dz): the three dz blocks of a plane group are
  // consecutive multiples of 8 apart in the launch order, so they run on one
  // XCD (workgroups are dealt to the 8 XCDs round-robin) and share its L2
  const int ngroups = int(gridDim.x / 3);
  const int xcd = int(blockIdx.x % 8), kx = int(blockIdx.x / 8);
  const int pg = xcd * (ngroups / 8) + kx / 3;
  const int dz = kx % 3 - 1;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;  // w: wave-uniform
  const int li = l & 15, lk = l >> 4;
  const int wc = w;  // compute wavefront's input tile
  const bool producer = w >= 4;
  const int pt = int(threadIdx.x) - 256;  // producer thread index
  floatx4 acc[3][3][MT];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int m = 0; m < MT; ++m) acc[a][d][m] = floatx4{0, 0, 0, 0};
  float bs[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // bias partial sums of this producer's dout channel block
  const bool dobias = db != nullptr && dz == 0;
  const int nb8 = n / 8;
  const int nd_row = nb8 * (COUT / 8);
  const int nx_row = nb8 * (CIN / 8);
  uint32_t v[PD][8][4];  // blocks loaded by this producer thread, not yet staged
  // this producer thread's staging job: a dout block (pt < nd_row) or an
  // input block, channel block fastest (8 lanes read one voxel's 128 bytes);
  // threads without a job re-read input block 0.  The global loads are issued
  // by every producer thread on every row, so the compiler can count the
  // loads in flight exactly (waiting for one row set while the PD - 1 later
  // sets stay in flight)
  const bool jd = pt < nd_row, jx = !jd && pt < nd_row + nx_row;
  const int jj = jx ? pt - nd_row : 0;
  const int cb = jd ? pt % (COUT / 8) : jj % (CIN / 8);
  const int xb = jd ? pt / (COUT / 8) : jj / (CIN / 8);
  // global loads of this thread's block of dout row drow or input row xrow into v[p]
  auto load = [&](int p, long drow, long xrow) __attribute__((always_inline)) {
    const bf16* src = jd ? &dout[(drow * n + xb * 8) * COUT + cb * 8] : &in[(xrow * n + xb * 8) * CIN + cb * 8];
    const int stride = jd ? COUT : CIN;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const uint4 q = *reinterpret_cast<const uint4*>(src + e * stride);
      v[p][e][0] = q.x; v[p][e][1] = q.y; v[p][e][2] = q.z; v[p][e][3] = q.w;
    }
  };
  // transposed LDS stores of v[p]: a dout block (if dvalid) into dT slot ds,
  // an input block (if xvalid) into xT slot xs and, for x blocks 0 and
  // n / 8 - 1, into the periodic halo block on the other side
  auto store = [&](int p, bool dvalid, int ds, bool xvalid, int xs) __attribute__((always_inline)) {
    uint32_t o[8][4];
    transpose8x8(v[p], o);
    const int ch0 = (cb % 2) * 8;  // channel of o[0] within its 16-channel tile
    if (jd && dvalid) {
      uint4* tile = &dT[(ds * MT + cb / 2) * DTS + xb * 16];
#pragma unroll
      for (int c = 0; c < 8; ++c) tile[wg_swz(ch0 + c, xb)] = uint4{o[c][0], o[c][1], o[c][2], o[c][3]};
      if (dobias) {  // o[c] holds channel cb * 8 + c at the block's 8 voxels, two per dword
        typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
        const bf16x2_t one2 = {__bf16(1.0f), __bf16(1.0f)};
#pragma unroll
        for (int c = 0; c < 8; ++c)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            bf16x2_t pr;
            __builtin_memcpy(&pr, &o[c][j], 4);
            bs[c] = __builtin_amdgcn_fdot2_f32_bf16(pr, one2, bs[c], false);  // exact products, fp32 sums
          }
      }
    } else if (jx && xvalid) {
      uint4* tile = &xT[(xs * NW + cb / 2) * XTS];
      const int bk = xb + 1;  // LDS block of x block xb (block 0 is the left halo)
#pragma unroll
      for (int c = 0; c < 8; ++c) tile[bk * 16 + wg_swz(ch0 + c, bk)] = uint4{o[c][0], o[c][1], o[c][2], o[c][3]};
      // edge dwords: the first dword of blocks 2 .. n / 8 + 1 and the last
      // of blocks 0 .. n / 8 - 1 are read (halos: x block 0 as block
      // n / 8 + 1, x block n / 8 - 1 as block 0)
      uint32_t* ef = &xE[(xs * NW + cb / 2) * 2 * EBS];
      uint32_t* el = ef + EBS;
      const int fb = xb == 0 ? nb8 + 1 : bk, lb = xb == nb8 - 1 ? 0 : bk;
      *reinterpret_cast<uint4*>(&ef[fb * 16 + ch0]) = uint4{o[0][0], o[1][0], o[2][0], o[3][0]};
      *reinterpret_cast<uint4*>(&ef[fb * 16 + ch0 + 4]) = uint4{o[4][0], o[5][0], o[6][0], o[7][0]};
      *reinterpret_cast<uint4*>(&el[lb * 16 + ch0]) = uint4{o[0][3], o[1][3], o[2][3], o[3][3]};
      *reinterpret_cast<uint4*>(&el[lb * 16 + ch0 + 4]) = uint4{o[4][3], o[5][3], o[6][3], o[7][3]};
      if (xb == 0 || xb == nb8 - 1) {
        const int hb = xb == 0 ? nb8 + 1 : 0;  // x + n or x - n
#pragma unroll
        for (int c = 0; c < 8; ++c) tile[hb * 16 + wg_swz(ch0 + c, hb)] = uint4{o[c][0], o[c][1], o[c][2], o[c][3]};
      }
    }
  };
  // one row of MFMAs: dout row y against input rows y - 1, y, y + 1, in
  // steps of (32 voxels, one input row).  The LDS reads of the next step are
  // issued before the MFMAs of the current one (two register sets).  A lane
  // reads x block bx = k0 / 8 + lk.
  struct BOps {
    u32x4 c;          // input block bx + 1 (x ..)
    uint32_t pw, nx;  // last dword of block bx (x - 2, x - 1), first of block bx + 2 (x + 8, x + 9)
  };
  const u32x4* dTv = reinterpret_cast<const u32x4*>(dT);
  const u32x4* xTv = reinterpret_cast<const u32x4*>(xT);
  auto chunk = [&](int blk) __attribute__((always_inline)) { return blk * 16 + wg_swz(li, blk); };
  auto read_a = [&](u32x4* av, int drow0, int k0) __attribute__((always_inline)) {
#pragma unroll
    for (int m = 0; m < MT; ++m) av[m] = dTv[drow0 + m * DTS + chunk(k0 / 8 + lk)];
  };
  auto read_b = [&](BOps& o, int slot, int k0) __attribute__((always_inline)) {
    const u32x4* xr = &xTv[(slot * NW + wc) * XTS];
    const uint32_t* er = &xE[(slot * NW + wc) * 2 * EBS];
    const int bx = k0 / 8 + lk;
    o.c = xr[chunk(bx + 1)];
    o.pw = er[EBS + bx * 16 + li];
    o.nx = er[(bx + 2) * 16 + li];
  };
  auto mfma_row = [&](const u32x4* am, const BOps& o, floatx4 (&accr)[3][MT]) __attribute__((always_inline)) {
    const u32x4 L0 = o.c;
    u32x4 S[3];
    // dx = -1: x0 - 1 .. x0 + 6;  dx = 0: x0 .. x0 + 7;  dx = +1: x0 + 1 .. x0 + 8
    S[0] = u32x4{__builtin_amdgcn_alignbyte(L0.x, o.pw, 2), __builtin_amdgcn_alignbyte(L0.y, L0.x, 2),
                 __builtin_amdgcn_alignbyte(L0.z, L0.y, 2), __builtin_amdgcn_alignbyte(L0.w, L0.z, 2)};
    S[1] = L0;
    S[2] = u32x4{__builtin_amdgcn_alignbyte(L0.y, L0.x, 2), __builtin_amdgcn_alignbyte(L0.z, L0.y, 2),
                 __builtin_amdgcn_alignbyte(L0.w, L0.z, 2), __builtin_amdgcn_alignbyte(o.nx, L0.w, 2)};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      bf16x8 bv;
      __builtin_memcpy(&bv, &S[d], 16);
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        bf16x8 av;
        __builtin_memcpy(&av, &am[m], 16);
        accr[d][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, accr[d][m], 0, 0, 0);
      }
    }
  };
  // A row is 2 n / 32 sections of 12 MFMAs (x step q = s / 3, input row
  // s % 3).  Each section issues the reads of the section two ahead (B set
  // (s + 2) % 3; the A set of the next x step one or two sections ahead)
  // before its MFMAs (sched_barrier / sched_group_barrier: DS reads, then
  // MFMAs; the scheduler would otherwise move the reads next to their first
  // use), so a read has two sections of MFMAs to arrive.  The reads after
  // the last step repeat the last step's, in range.
#define WG_ORDER(NREADS)                                                         \
  __builtin_amdgcn_sched_group_barrier(0x100, NREADS, 0); /* DS reads */    \
  __builtin_amdgcn_sched_group_barrier(0x008, 3 * MT, 0); /* MFMAs */
  auto compute_row = [&](int y) __attribute__((always_inline)) {
    const int sl[3] = {y & 3, (y + 1) & 3, (y + 2) & 3};  // input rows y - 1, y, y + 1
    const int drow0 = (y & 1) * MT * DTS;
    u32x4 a0[MT], a1[MT];
    BOps b0, b1, b2;
    read_a(a0, drow0, 0);
    read_b(b0, sl[0], 0);
    read_b(b1, sl[1], 0);
#pragma unroll 1
    for (int k0 = 0; k0 < n; k0 += 64) {  // n is a multiple of 64: two x steps per iteration
      const int kn = min(k0 + 64, n - 32);
      __builtin_amdgcn_sched_barrier(0);
      read_b(b2, sl[2], k0);
      mfma_row(a0, b0, acc[0]);
      WG_ORDER(3)
      __builtin_amdgcn_sched_barrier(0);
      read_a(a1, drow0, k0 + 32);
      read_b(b0, sl[0], k0 + 32);
      mfma_row(a0, b1, acc[1]);
      WG_ORDER(MT + 3)
      __builtin_amdgcn_sched_barrier(0);
      read_b(b1, sl[1], k0 + 32);
      mfma_row(a0, b2, acc[2]);
      WG_ORDER(3)
      __builtin_amdgcn_sched_barrier(0);
      read_b(b2, sl[2], k0 + 32);
      mfma_row(a1, b0, acc[0]);
      WG_ORDER(3)
      __builtin_amdgcn_sched_barrier(0);
      read_a(a0, drow0, kn);
      read_b(b0, sl[0], kn);
      mfma_row(a1, b1, acc[1]);
      WG_ORDER(MT + 3)
      __builtin_amdgcn_sched_barrier(0);
      read_b(b1, sl[1], kn);
      mfma_row(a1, b2, acc[2]);
      WG_ORDER(3)
    }
  };
#undef WG_ORDER
  // Row set s = dout row s + 1 and input row s + 2; it is loaded into
  // v[s % PD] during row s - PD (or before the plane's first row) and staged
  // during row s.  Slot of input row r (r in [-1, n]): (r + 1) & 3.
  // Producers and compute wavefronts run separate loops (the role is
  // wave-uniform) with the same barriers: two per plane and one per row.
  if (producer) {
#pragma unroll 1
    for (long q = pg; q < nplanes; q += ngroups) {
      // the plane group's three blocks read the same dout plane q at the same time
      const int z = int(q % n);
      const long b = q / n;
      const long pin = b * n + wrap(z + dz, n);  // input plane
      const long pout = q;                       // dout plane
      __syncthreads();
      load(0, pout * n + 0, pin * n + (n - 1));  // dout row 0; input row -1
      store(0, true, 0, true, 0);
      load(0, pout * n + 0, pin * n + 0);  // input row 0
      store(0, false, 0, true, 1);
      load(0, pout * n + 0, pin * n + (1 % n));  // input row 1
      store(0, false, 0, true, 2);
#pragma unroll
      for (int s = 0; s < PD; ++s) load(s, pout * n + min(s + 1, n - 1), pin * n + (s + 2) % n);
      __syncthreads();
#pragma unroll 1
      for (int y0 = 0; y0 < n; y0 += PD) {
#pragma unroll
        for (int p = 0; p < PD; ++p) {
          const int y = y0 + p;
          // set y: dout row y + 1, input row y + 2.  Set n - 1 (dout row n
          // does not exist) is loaded like the others and its input block is
          // written to the free slots, which the next plane overwrites.
          store(p, y + 1 < n, (y + 1) & 1, true, (y + 3) & 3);
          load(p, pout * n + min(y + PD + 1, n - 1), pin * n + (y + PD + 2) % n);
          __syncthreads();
        }
      }
    }
  } else {
#pragma unroll 1
    for (long q = pg; q < nplanes; q += ngroups) {
      __syncthreads();
      __syncthreads();
#pragma unroll 1
      for (int y = 0; y < n; ++y) {
        compute_row(y);
        __syncthreads();
      }
    }
  }
  if (dobias) {
    // a producer's dout block is the same on every row: reduce the 8-channel
    // sums over the x blocks
    __syncthreads();
    float* red = reinterpret_cast<float*>(xT);  // >= 8 KB for every CIN
    if (producer) {
#pragma unroll
      for (int c = 0; c < 8; ++c) red[pt * 8 + c] = pt < nd_row ? bs[c] : 0.f;
    }
    __syncthreads();
    if (threadIdx.x < COUT) {
      const int cbt = threadIdx.x / 8, c = threadIdx.x % 8;
      float t = 0.f;
      for (int x8 = 0; x8 < nb8; ++x8) t += red[(x8 * (COUT / 8) + cbt) * 8 + c];
      atomicAdd(&db[threadIdx.x], t);
    }
  }
  if (producer) return;
  // D[co = 16 m + 4 lk + r][ci = 16 wc + li]
  const int ci = 16 * wc + li;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int t = (dz + 1) * 9 + a * 3 + d;
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int co = m * 16 + 4 * lk + r;
          part[((long(pg) * 27 + t) * COUT + co) * CIN + ci] = acc[a][d][m][r];
        }
    }
}

// Weight gradient of the two narrow layers with all 27 taps in one block, so
// each staged row reaches one compute unit once (the three-block form above
// moves every row into three compute units, which bounds these layers by
// the bytes delivered per unit, not by MFMAs).  Block p takes the (batch, z)
// planes p, p + gridDim.x, ... whole.  Staging (wavefronts NCWV .. NCWV + 3),
// the staged rows' chunk layout and the producers' row pipeline are those of
// k_conv_wgrad; the x neighbours are read as whole chunks (edge arrays as
// there did not change these kernels' time).
//  INPACK (CIN <= 8, COUT = 64): the input planes z - 1, z, z + 1 are staged
//    and (tap, input channel) pairs go on M: compute wavefront w owns the
//    taps s = 2 w, 2 w + 1 of s = 3 (dz + 1) + dy + 1 (lanes 0-7 and 8-15,
//    one channel each), reads their x blocks once and forms the dx = -1, 0, 1
//    fragments, each against the 4 output tiles, and tap s = 8 (lanes 0-7;
//    lanes 8-15 repeat it and their rows are dropped) against output tile w:
//    15 MFMAs per 32 voxels.
//    C4 (at most 4 real input channels): 16-row tiles of 4 taps x 4
//    channels; compute wavefront w owns the taps s = 4 (w & 1) .. + 3 (lane
//    group li / 4 one tap, lane li % 4 its channel) against output tiles
//    2 (w / 2), + 1, and tap 8 against output tile w: 9 MFMAs per 32 voxels.
//    The row slot and plane strides are 64 and 192 bytes modulo 256, so the
//    four taps a 16-lane group reads lie in four distinct 64-byte quarters.
//  otherwise (CIN = 64, COUT = 16 or 8): the input plane z and the dout planes
//    z + 1, z, z - 1 (dz = -1, 0, 1) are staged; compute wavefront w owns
//    input tile w and all 27 taps.  COUT = 16: 27 MFMAs per 32 voxels.
//    COUT = 8 (the 6-channel output layer): the dout planes dz = -1, 0 share
//    one 16-row tile (channels 0-7 and 8-15) and dz = 1 has a tile of its
//    own with channels 8-15 zero, so 18 MFMAs per 32 voxels cover all taps.
template <int CIN, int COUT, int PD, bool C4 = false>
__global__ void __launch_bounds__(512, 1)
    k_conv_wgrad3(const bf16* __restrict__ in, const bf16* __restrict__ dout, float* part, float* db, int n,
                  long nplanes) {
  constexpr bool INPACK = CIN <= 8;
  constexpr bool PACKO = !INPACK && COUT == 8;  // two dout planes per 16-row tile
  static_assert(!C4 || INPACK, "C4: the input layer");
  constexpr int NCWV = 4;  // compute wavefronts
  constexpr int NBM = 16;               // x blocks per row (n <= 128)
  constexpr int CP = CIN < 16 ? 16 : CIN;
  constexpr int MT = (COUT + 15) / 16;  // output tiles (COUT = 8: one tile, channels 8-15 zero)
  constexpr int NW = CP / 16;
  constexpr int NPI = INPACK ? 3 : 1;  // input planes per row step
  constexpr int NPO = INPACK ? 1 : 3;  // dout planes per row step
  constexpr int NTD = PACKO ? 2 : NPO;  // dout tiles per row slot and output tile
  constexpr int NA = INPACK ? 0 : NTD;  // A operands (dout tiles) per x step, CIN = 64
  // tile strides in 16-byte chunks.  CIN = 64: input tiles start 16 bytes
  // apart (the staging stores of 8 channel blocks hit distinct 16-byte slots
  // modulo 128 bytes).  INPACK: the row slots and planes are 128 bytes apart
  // modulo 256, so the two taps a 16-lane group reads (plane + row offset of
  // opposite parity) fill the two halves of one 256-byte window
  constexpr int DTS = NBM * 16 + 1;
  constexpr int XTS = C4 ? (NBM + 2) * 16 + 4 : INPACK ? (NBM + 2) * 16 + 8 : (NBM + 2) * 16 + 1;
  constexpr int XPS = C4 ? 4 * XTS + 12 : INPACK ? 4 * XTS + 8 : 4 * NW * XTS;  // input plane stride
  static_assert(NPO * 16 * (COUT / 8) + NPI * 16 * (CIN / 8) <= 256, "staging jobs");
  static_assert(INPACK ? (MT == 4 && NW == 1) : (MT == 1 && NW == 4), "wgrad3 shapes");
  static_assert(COUT % 8 == 0, "dout blocks of 8 channels");
  static_assert(64 % PD == 0, "row sets in flight divide n");
  __shared__ uint4 dT[NTD * 2 * MT * DTS];  // [dout tile][row slot y & 1][output tile]
  __shared__ uint4 xT[NPI * XPS];           // [input plane][row slot (r + 1) & 3][tile]
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;  // w: wave-uniform
  const int li = l & 15, lk = l >> 4;
  const bool producer = w >= NCWV;
  const int pt = int(threadIdx.x) - 64 * NCWV;  // producer thread index
  constexpr int NACC = C4 ? 9 : INPACK ? 3 * MT + 3 : 9 * NA;
  floatx4 acc[NACC];
#pragma unroll
  for (int i = 0; i < NACC; ++i) acc[i] = floatx4{0, 0, 0, 0};
  if (CIN < CP) {  // the zero-padded channels are never written: clear all input slots once
    for (int i = threadIdx.x; i < NPI * XPS; i += 64 * (NCWV + 4)) xT[i] = uint4{0, 0, 0, 0};
  }
  if (COUT < 16 * MT) {  // likewise the zero-padded output channels
    for (int i = threadIdx.x; i < NTD * 2 * MT * DTS; i += 64 * (NCWV + 4)) dT[i] = uint4{0, 0, 0, 0};
  }
  float bs[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int nb8 = n / 8;
  const int ndp = nb8 * (COUT / 8), nxp = nb8 * (CIN / 8);  // jobs per dout / input plane row
  const int nd_row = NPO * ndp, nx_row = NPI * nxp;
  const bool jd = pt < nd_row, jx = !jd && pt < nd_row + nx_row;
  const int jj = jd ? pt : (jx ? pt - nd_row : 0);
  const int jpl = jd ? jj / ndp : jj / nxp;          // the job's plane index (po or pi)
  const int jr = jd ? jj % ndp : jj % nxp;
  const int cb = jd ? jr % (COUT / 8) : jr % (CIN / 8);
  const int xb = jd ? jr / (COUT / 8) : jr / (CIN / 8);
  // the bias gradient: every dout row once (INPACK: the one dout plane;
  // otherwise the dz = 0 plane)
  const bool dobias = db != nullptr && (INPACK || jpl == 1);
  uint32_t v[PD][8][4];
  long myplane = 0;  // this job's plane, set per block plane
  auto load = [&](int p, int yd, int yx) __attribute__((always_inline)) {
    const bf16* src = jd ? &dout[((myplane * n + yd) * n + xb * 8) * COUT + cb * 8]
                         : &in[((myplane * n + yx) * n + xb * 8) * CIN + cb * 8];
    const int stride = jd ? COUT : CIN;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const uint4 q = *reinterpret_cast<const uint4*>(src + e * stride);
      v[p][e][0] = q.x; v[p][e][1] = q.y; v[p][e][2] = q.z; v[p][e][3] = q.w;
    }
  };
  auto store = [&](int p, bool dvalid, int ds, bool xvalid, int xs) __attribute__((always_inline)) {
    uint32_t o[8][4];
    transpose8x8(v[p], o);
    if (jd && dvalid) {
      // PACKO: dout plane jpl -> tile jpl / 2, channels 8 (jpl % 2) .. + 7
      const int td = PACKO ? jpl >> 1 : jpl;
      const int ch0 = PACKO ? (jpl & 1) * 8 : (cb % 2) * 8;
      uint4* tile = &dT[((td * 2 + ds) * MT + cb / 2) * DTS + xb * 16];
#pragma unroll
      for (int c = 0; c < 8; ++c) tile[wg_swz(ch0 + c, xb)] = uint4{o[c][0], o[c][1], o[c][2], o[c][3]};
      if (dobias) {
        typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
        const bf16x2_t one2 = {__bf16(1.0f), __bf16(1.0f)};
#pragma unroll
        for (int c = 0; c < 8; ++c)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            bf16x2_t pr;
            __builtin_memcpy(&pr, &o[c][j], 4);
            bs[c] = __builtin_amdgcn_fdot2_f32_bf16(pr, one2, bs[c], false);
          }
      }
    } else if (jx && xvalid) {
      const int ch0 = (cb % 2) * 8;
      uint4* tile = &xT[jpl * XPS + (xs * NW + cb / 2) * XTS];
      const int bk = xb + 1;
#pragma unroll
      for (int c = 0; c < 8; ++c) tile[bk * 16 + wg_swz(ch0 + c, bk)] = uint4{o[c][0], o[c][1], o[c][2], o[c][3]};
      if (xb == 0 || xb == nb8 - 1) {
        const int hb = xb == 0 ? nb8 + 1 : 0;
#pragma unroll
        for (int c = 0; c < 8; ++c) tile[hb * 16 + wg_swz(ch0 + c, hb)] = uint4{o[c][0], o[c][1], o[c][2], o[c][3]};
      }
    }
  };
  const u32x4* dTv = reinterpret_cast<const u32x4*>(dT);
  const u32x4* xTv = reinterpret_cast<const u32x4*>(xT);
  auto mfma = [&](const u32x4& a, const u32x4& b, floatx4& d) __attribute__((always_inline)) {
    bf16x8 av, bv;
    __builtin_memcpy(&av, &a, 16);
    __builtin_memcpy(&bv, &b, 16);
    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, d, 0, 0, 0);
  };
  // the x shifts of a staged row: from the chunks of x blocks bx - 1, bx,
  // bx + 1 (X[0..2]; LDS blocks bx, bx + 1, bx + 2, block 0 the left halo)
  auto chunk = [&](int blk, int ch) __attribute__((always_inline)) { return blk * 16 + wg_swz(ch, blk); };
  auto read_x = [&](u32x4* X, const u32x4* xr, int k0, int ch) __attribute__((always_inline)) {
    const int bx = k0 / 8 + lk;
#pragma unroll
    for (int i = 0; i < 3; ++i) X[i] = xr[chunk(bx + i, ch)];
  };
  auto shifts = [&](const u32x4* X, u32x4* S) __attribute__((always_inline)) {
    const u32x4 P = X[0], C = X[1], N = X[2];
    asm volatile("" : : "v"(P), "v"(N));  // keeps the reads 16 bytes wide (conflict-free); no instruction
    S[0] = u32x4{__builtin_amdgcn_alignbyte(C.x, P.w, 2), __builtin_amdgcn_alignbyte(C.y, C.x, 2),
                 __builtin_amdgcn_alignbyte(C.z, C.y, 2), __builtin_amdgcn_alignbyte(C.w, C.z, 2)};
    S[1] = C;
    S[2] = u32x4{__builtin_amdgcn_alignbyte(C.y, C.x, 2), __builtin_amdgcn_alignbyte(C.z, C.y, 2),
                 __builtin_amdgcn_alignbyte(C.w, C.z, 2), __builtin_amdgcn_alignbyte(N.x, C.w, 2)};
  };
  // this lane's rows: INPACK: tap s = min(2 w + li / 8, 8) (plane dz + 1 = s / 3,
  // row y - 1 + s % 3), channel li % 8; otherwise the input tile w rows
  const int sl_tap = C4 ? 4 * (w & 1) + li / 4 : INPACK ? 2 * w + li / 8 : 0;
  // one row: the x steps in pairs (n is a multiple of 64), every operand
  // chunk read one step ahead, each step's reads kept at its top
  // (sched_barrier) so that its MFMAs wait only for reads issued a step
  // earlier; the read after the last step repeats the last step's (in range)
  auto compute_row = [&](int y) __attribute__((always_inline)) {
    const u32x4* dr = &dTv[(y & 1) * MT * DTS];  // dout row y (slot y & 1), output tile 0
    if constexpr (C4) {
      const int ci = li & 3, m0 = 2 * (w >> 1);
      const u32x4* xr = &xTv[(sl_tap / 3) * XPS + ((y + sl_tap % 3) & 3) * XTS];
      const u32x4* xr8 = &xTv[2 * XPS + ((y + 2) & 3) * XTS];  // tap 8: plane z + 1, row y + 1
      struct Ops {
        u32x4 B[2];  // output tiles m0, m0 + 1 (tile w is B[w & 1])
        u32x4 X[3], X8[3];
      };
      auto read_ops = [&](Ops& o, int k0) __attribute__((always_inline)) {
        const int c = chunk(k0 / 8 + lk, li);
        o.B[0] = dr[m0 * DTS + c];
        o.B[1] = dr[(m0 + 1) * DTS + c];
        read_x(o.X, xr, k0, ci);
        read_x(o.X8, xr8, k0, ci);
      };
      // M8 = w & 1 as a constant: tile w is B[M8] (a run-time index would
      // move the operand sets to memory)
      auto step = [&](const Ops& o, auto M8) __attribute__((always_inline)) {
        u32x4 S[3], S8[3];
        shifts(o.X, S);
        shifts(o.X8, S8);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
#pragma unroll
          for (int m = 0; m < 2; ++m) mfma(S[d], o.B[m], acc[d * 2 + m]);
          mfma(S8[d], o.B[decltype(M8)::value], acc[6 + d]);  // tap 8 against output tile w
        }
      };
      auto run = [&](auto M8) __attribute__((always_inline)) {
        Ops o0, o1;
        read_ops(o0, 0);
#pragma unroll 1
        for (int k0 = 0; k0 < n; k0 += 64) {
          __builtin_amdgcn_sched_barrier(0);
          read_ops(o1, k0 + 32);
          step(o0, M8);
          __builtin_amdgcn_sched_barrier(0);
          read_ops(o0, min(k0 + 64, n - 32));
          step(o1, M8);
        }
      };
      if (w & 1)
        run(std::integral_constant<int, 1>{});
      else
        run(std::integral_constant<int, 0>{});
    } else if constexpr (INPACK) {
      const int ci = li & 7;
      const u32x4* xr = &xTv[(sl_tap / 3) * XPS + ((y + sl_tap % 3) & 3) * XTS];
      const u32x4* xr8 = &xTv[2 * XPS + ((y + 2) & 3) * XTS];  // tap 8: plane z + 1, row y + 1
      struct Ops {
        u32x4 B[MT + 1];  // output tiles 0 .. MT - 1, then tile w (tap 8)
        u32x4 X[3], X8[3];
      };
      auto read_ops = [&](Ops& o, int k0) __attribute__((always_inline)) {
        const int c = chunk(k0 / 8 + lk, li);
#pragma unroll
        for (int m = 0; m < MT; ++m) o.B[m] = dr[m * DTS + c];
        o.B[MT] = dr[w * DTS + c];
        read_x(o.X, xr, k0, ci);
        read_x(o.X8, xr8, k0, ci);
      };
      auto step = [&](const Ops& o) __attribute__((always_inline)) {
        u32x4 S[3], S8[3];
        shifts(o.X, S);
        shifts(o.X8, S8);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
#pragma unroll
          for (int m = 0; m < MT; ++m) mfma(S[d], o.B[m], acc[d * MT + m]);
          mfma(S8[d], o.B[MT], acc[3 * MT + d]);  // tap 8 against output tile w
        }
      };
      Ops o0, o1;
      read_ops(o0, 0);
#pragma unroll 1
      for (int k0 = 0; k0 < n; k0 += 64) {
        __builtin_amdgcn_sched_barrier(0);
        read_ops(o1, k0 + 32);
        step(o0);
        __builtin_amdgcn_sched_barrier(0);
        read_ops(o0, min(k0 + 64, n - 32));
        step(o1);
      }
    } else {
      const u32x4* xr[3];  // input rows y - 1, y, y + 1 of tile w
#pragma unroll
      for (int a = 0; a < 3; ++a) xr[a] = &xTv[(((y + a) & 3) * NW + w) * XTS];
      // PACKO: tile 0 = dout planes dz = -1 (rows 0-7), 0 (rows 8-15);
      // tile 1 = dz = 1 (rows 0-7).  Otherwise tile pp = plane pp
      struct Ops {
        u32x4 A[NA];
        u32x4 X[3][3];
      };
      auto read_ops = [&](Ops& o, int k0) __attribute__((always_inline)) {
        const int c = chunk(k0 / 8 + lk, li);
#pragma unroll
        for (int pp = 0; pp < NA; ++pp) o.A[pp] = dr[pp * 2 * DTS + c];
#pragma unroll
        for (int a = 0; a < 3; ++a) read_x(o.X[a], xr[a], k0, li);
      };
      auto step = [&](const Ops& o) __attribute__((always_inline)) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          u32x4 S[3];
          shifts(o.X[a], S);
#pragma unroll
          for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int pp = 0; pp < NA; ++pp) mfma(o.A[pp], S[d], acc[pp * 9 + a * 3 + d]);
        }
      };
      Ops o0, o1;
      read_ops(o0, 0);
#pragma unroll 1
      for (int k0 = 0; k0 < n; k0 += 64) {
        __builtin_amdgcn_sched_barrier(0);
        read_ops(o1, k0 + 32);
        step(o0);
        __builtin_amdgcn_sched_barrier(0);
        read_ops(o0, min(k0 + 64, n - 32));
        step(o1);
      }
    }
  };
  if (producer) {
#pragma unroll 1
    for (long q = blockIdx.x; q < nplanes; q += gridDim.x) {
      const int z = int(q % n);
      const long b = q / n;
      // INPACK: dout plane q, input planes z - 1, z, z + 1; otherwise input
      // plane q, dout planes z + 1, z, z - 1 (dz = -1, 0, 1: dout plane z - dz)
      myplane = jd ? (INPACK ? q : b * n + wrap(z + 1 - jpl, n)) : (INPACK ? b * n + wrap(z + jpl - 1, n) : q);
      __syncthreads();
      load(0, 0, n - 1);  // dout row 0; input row -1
      store(0, true, 0, true, 0);
      load(0, 0, 0);  // input row 0
      store(0, false, 0, true, 1);
      load(0, 0, 1 % n);  // input row 1
      store(0, false, 0, true, 2);
#pragma unroll
      for (int s2 = 0; s2 < PD; ++s2) load(s2, min(s2 + 1, n - 1), (s2 + 2) % n);
      __syncthreads();
#pragma unroll 1
      for (int y0 = 0; y0 < n; y0 += PD) {
#pragma unroll
        for (int p = 0; p < PD; ++p) {
          const int y = y0 + p;
          store(p, y + 1 < n, (y + 1) & 1, true, (y + 3) & 3);
          load(p, min(y + PD + 1, n - 1), (y + PD + 2) % n);
          __syncthreads();
        }
      }
    }
  } else {
#pragma unroll 1
    for (long q = blockIdx.x; q < nplanes; q += gridDim.x) {
      __syncthreads();
      __syncthreads();
#pragma unroll 1
      for (int y = 0; y < n; ++y) {
        compute_row(y);
        __syncthreads();
      }
    }
  }
  if (db != nullptr) {
    __syncthreads();
    float* red = reinterpret_cast<float*>(xT);  // >= 8 KB
    if (producer) {
#pragma unroll
      for (int c = 0; c < 8; ++c) red[pt * 8 + c] = (jd && dobias) ? bs[c] : 0.f;
    }
    __syncthreads();
    if (threadIdx.x < COUT) {
      const int cbt = threadIdx.x / 8, c = threadIdx.x % 8;
      const int p0 = INPACK ? 0 : ndp;  // the plane whose rows carry the bias sums
      float t = 0.f;
      for (int x8 = 0; x8 < nb8; ++x8) t += red[(p0 + x8 * (COUT / 8) + cbt) * 8 + c];
      atomicAdd(&db[threadIdx.x], t);
    }
  }
  if (producer) return;
  const long pbase = long(blockIdx.x) * 27;
  if constexpr (C4) {
    // D[row = 4 lk + r][co = 16 (m0 + m) + li]: tap 4 (w & 1) + lk, channel r;
    // tap 8 in lane group lk = 0.  Channels 4-7 of every tap get zeros.
    const int m0 = 2 * (w >> 1);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        const int t = (4 * (w & 1) + lk) * 3 + d;  // (dz + 1) * 9 + (dy + 1) * 3 + dx + 1
        float* dst = &part[((pbase + t) * COUT + 16 * (m0 + m) + li) * CIN];
        *reinterpret_cast<floatx4*>(dst) = acc[d * 2 + m];
        *reinterpret_cast<floatx4*>(dst + 4) = floatx4{0, 0, 0, 0};
      }
      if (lk == 0) {
        float* dst = &part[((pbase + 24 + d) * COUT + 16 * w + li) * CIN];
        *reinterpret_cast<floatx4*>(dst) = acc[6 + d];
        *reinterpret_cast<floatx4*>(dst + 4) = floatx4{0, 0, 0, 0};
      }
    }
  } else if constexpr (INPACK) {
    // D[row = 4 lk + r][co = 16 m + li]; row -> (tap 2 w + row / 8, channel row % 8)
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 4 * lk + r, s = 2 * w + row / 8, ci = row % 8;
          const int t = s * 3 + d;  // (dz + 1) * 9 + (dy + 1) * 3 + dx + 1
          const int co = 16 * m + li;
          part[((pbase + t) * COUT + co) * CIN + ci] = acc[d * MT + m][r];
        }
    // tap 8: rows 0-7 (channels), output tile w
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 4 * lk + r;
        if (row >= 8) continue;
        const int co = 16 * w + li;
        part[((pbase + 24 + d) * COUT + co) * CIN + row] = acc[3 * MT + d][r];
      }
  } else {
    // D[row = 4 lk + r][ci = 16 w + li]; PACKO: row -> (dout plane 2 pp + row / 8,
    // channel row % 8), otherwise (plane pp, channel row)
    const int ci = 16 * w + li;
#pragma unroll
    for (int pp = 0; pp < NA; ++pp) {
      const int po = PACKO ? 2 * pp + lk / 2 : pp;
      if (po > 2 || (!PACKO && 4 * lk >= COUT)) continue;
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int co = PACKO ? (4 * lk + r) % 8 : 4 * lk + r;
            const int t = po * 9 + a * 3 + d;
            part[((pbase + t) * COUT + co) * CIN + ci] = acc[pp * 9 + a * 3 + d][r];
          }
    }
  }
}

__global__ void k_reduce_partials(const float* part, float* out, long nparts, long len) {
  long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= len) return;
  float s = 0.f;
  for (long p = 0; p < nparts; ++p) s += part[p * len + i];
  out[i] += s;
}


// MSE loss over the first `nout` channels of y (fp32, channel pitch cp):
// loss partials, and the gradient dy = 2 (y - t) / (nvox * nout) as bf16 with
// channel pitch cp (padding channels zero).
__global__ void k_mse(const float* y, const float* tgt, bf16* dy, float* dyf, double* part, long nvox, int nout,
                      int cp, double inv) {
  __shared__ double sm[256];
  long v = long(blockIdx.x) * blockDim.x + threadIdx.x;
  double acc = 0;
  const bool vec = cp == 8 && nout == 6 && reinterpret_cast<uintptr_t>(y) % 16 == 0 &&
                   reinterpret_cast<uintptr_t>(tgt) % 8 == 0 && reinterpret_cast<uintptr_t>(dy) % 16 == 0 &&
                   reinterpret_cast<uintptr_t>(dyf) % 16 == 0;
  if (v < nvox && vec) {
    // the network's 6-channel output: a voxel's record in 16-, 8- and
    // 16-byte accesses (per-channel accesses would touch 4 or 2 bytes of a
    // 32-, 24- or 16-byte record per lane); the same arithmetic in the same order
    const float4 y0 = reinterpret_cast<const float4*>(y + v * 8)[0];
    const float4 y1 = reinterpret_cast<const float4*>(y + v * 8)[1];
    const float2* tv = reinterpret_cast<const float2*>(tgt + v * 6);
    const float2 t0 = tv[0], t1 = tv[1], t2 = tv[2];
    const float e[6] = {y0.x - t0.x, y0.y - t0.y, y0.z - t1.x, y0.w - t1.y, y1.x - t2.x, y1.y - t2.y};
    bf16x8 o;
    float g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      acc += double(e[c]) * e[c];
      g[c] = float(2.0 * inv) * e[c];
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) o[c] = bf16(g[c]);
    *reinterpret_cast<bf16x8*>(dy + v * 8) = o;
    if (dyf) {
      reinterpret_cast<float4*>(dyf + v * 8)[0] = float4{g[0], g[1], g[2], g[3]};
      reinterpret_cast<float4*>(dyf + v * 8)[1] = float4{g[4], g[5], g[6], g[7]};
    }
  } else if (v < nvox) {
    for (int c = 0; c < cp; ++c) {
      float g = 0.f;
      if (c < nout) {
        float e = y[v * cp + c] - tgt[v * nout + c];
        acc += double(e) * e;
        g = float(2.0 * inv) * e;
      }
      dy[v * cp + c] = bf16(g);
      if (dyf) dyf[v * cp + c] = g;
    }
  }
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// Adam on FP32 master weights; refreshes the two BF16 copies.
// sum of squares of x[0, len) in block partials (FP64)
__global__ void k_sumsq(const float* x, long len, double* part) {
  __shared__ double sm[256];
  // 16-byte loads into four independent sums where x is 16-byte aligned
  const long gid = long(blockIdx.x) * blockDim.x + threadIdx.x, stride = long(gridDim.x) * blockDim.x;
  const long n4 = reinterpret_cast<uintptr_t>(x) % 16 == 0 ? len / 4 : 0;
  double a[4] = {0, 0, 0, 0};
  for (long i = gid; i < n4; i += stride) {
    const float4 q = reinterpret_cast<const float4*>(x)[i];
    a[0] += double(q.x) * q.x;
    a[1] += double(q.y) * q.y;
    a[2] += double(q.z) * q.z;
    a[3] += double(q.w) * q.w;
  }
  for (long i = 4 * n4 + gid; i < len; i += stride) {
    const double e = x[i];
    a[0] += e * e;
  }
  double acc = (a[0] + a[1]) + (a[2] + a[3]);
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// trace removal on the last layer's weight gradient: for every (tap, input
// channel), rows co = 0, 1, 2 (xx, yy, zz) lose their mean.  The map is the
// projection the forward pass applies to the output, and it is its own
// transpose, so this turns the gradient of the raw weights' output into the
// gradient through the projected output.  Idempotent: earlier (projected)
// accumulations are unchanged.
__global__ void k_tracefree_rows(float* g, int cop, int cip) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 27 * cip) return;
  const int t = i / cip, ci = i % cip;
  float* r = g + long(t) * cop * cip + ci;
  const float m = (r[0] + r[cip] + r[2 * cip]) / 3.f;
  r[0] -= m;
  r[cip] -= m;
  r[2 * cip] -= m;
}

// master [27][cop][cip] -> bf16 B fragments of the forward convolution,
// W[co][k = t * cip + ci] (k padded to kpad(cip), co to a multiple of 16), and
// of the backward-data convolution, W'[ci][k = t' * cop + co] = W[26 - t'][co][ci]
// (taps mirrored).  Fragment layout: for k step kc and channel tile nt, lane
// L = 16 j + i holds entries (row 16 nt + i, k = 32 kc + 8 j .. + 7), stored at
// ((kc * NT + nt) * 64 + L) * 8; padding entries are zero.  With tracefree
// (last layer only) the backward-data weights are those of the projected
// output: rows co = 0, 1, 2 minus their mean, so the data gradient is the
// raw weights' transpose applied to the projected cotangent.
__device__ __forceinline__ long frag_index(int row, int k, int nrows) {
  const int kc = k >> 5, j = (k >> 3) & 3, e = k & 7, nt = row >> 4, i = row & 15;
  return ((long(kc) * (nrows / 16) + nt) * 64 + 16 * j + i) * 8 + e;
}
// With paired (the last layer, cop = 8, cip = 64; see PAIR in k_conv) the
// forward fragments take the paired layout: k step kc = 24 pz + (ro 3 + dxi)
// 2 + ch (plane pz = dz + 1, staged row offset ro, dxi = dx + 1, channel half
// ch), row r = 8 h + co holding the tap (dz, dy = ro - 1 - h, dx) of output
// channel co, zero where dy leaves [-1, 1].
__device__ __forceinline__ int pair_kpad() { return 72 * 32; }
__global__ void k_pack_weights(const float* w, bf16* wf, bf16* wb, int cop, int cip, int tracefree, int paired) {
  const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  const int kpf = paired ? pair_kpad() : kpad(cip), kpb = kpad(cop);
  const int rf = paired ? 16 : (cop + 15) / 16 * 16, rb = (cip + 15) / 16 * 16;
  const long nf = long(rf) * kpf, nb = long(rb) * kpb;
  if (i < nf && paired) {
    const int r = int(i / kpf), k = int(i % kpf);
    const int kc = k / 32, ci = (kc % 2) * 32 + k % 32;
    const int pz = kc / 24, ro = (kc % 24) / 6, dxi = (kc / 2) % 3;
    const int h = r / 8, co = r % 8, dy = ro - 1 - h;
    const int t = pz * 9 + (dy + 1) * 3 + dxi;
    wf[frag_index(r, k, rf)] = bf16(dy >= -1 && dy <= 1 && co < cop ? w[(long(t) * cop + co) * cip + ci] : 0.f);
  } else if (i < nf) {
    const int co = int(i / kpf), k = int(i % kpf);
    const int t = k / cip, ci = k % cip;
    wf[frag_index(co, k, rf)] = bf16(t < 27 && co < cop ? w[(long(t) * cop + co) * cip + ci] : 0.f);
  } else if (i < nf + nb) {
    const long j = i - nf;
    const int ci = int(j / kpb), k = int(j % kpb);
    const int tp = k / cop, co = k % cop;
    float val = 0.f;
    if (tp < 27 && ci < cip) {
      const float* r = w + long(26 - tp) * cop * cip + ci;
      val = r[long(co) * cip];
      if (tracefree && co < 3) val -= (r[0] + r[cip] + r[2 * cip]) / 3.f;
    }
    wb[frag_index(ci, k, rb)] = bf16(val);
  }
}

// pack physical fields (component-major, ncomp x nvox, T) into channels-last
// bf16 [nvox][cp]: channels 0..ncomp-1 = field * scale, channel ncomp = extra
// (constant, e.g. the Re_Delta feature) when has_extra, rest zero
template <typename T>
__global__ void k_pack_input(const T* f, long nvox, int ncomp, double scale, int has_extra, float extra, bf16* out,
                             int cp, long cstride) {
  long v = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= nvox) return;
  for (int c = 0; c < cp; ++c) {
    float val = 0.f;
    if (c < ncomp) val = float(double(f[c * cstride + v]) * scale);
    else if (c == ncomp && has_extra) val = extra;
    out[v * cp + c] = bf16(val);
  }
}

// pack ncomp component-major fields into channels-last fp32 [nvox][ncomp] * scale
template <typename T>
__global__ void k_pack_target(const T* f, long nvox, int ncomp, double scale, float* out, long cstride) {
  long v = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= nvox) return;
  for (int c = 0; c < ncomp; ++c) out[v * ncomp + c] = float(double(f[c * cstride + v]) * scale);
}

// unpack channels-last fp32 [nvox][cp] -> component-major T (first ncomp)
template <typename T>
__global__ void k_unpack(const float* y, long nvox, int ncomp, int cp, double scale, T* out, long stride) {
  long v = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= nvox) return;
  for (int c = 0; c < ncomp; ++c) out[c * stride + v] = T(double(y[v * cp + c]) * scale);
}

// cotangent of the six packed stress channels (xx, yy, zz, xy, yz, zx) from
// the nine physical components g (index i + 3 j): a diagonal entry takes g_ii,
// an off-diagonal entry appears twice in the divergence and takes g_ij + g_ji
template <typename T>
__global__ void k_pack_tau_cot(const T* g, long nvox, long stride, double scale, bf16* dy, int cp) {
  long v = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= nvox) return;
  auto G = [&](int i, int j) { return double(g[long(i + 3 * j) * stride + v]); };
  const double t[6] = {G(0, 0), G(1, 1), G(2, 2), G(0, 1) + G(1, 0), G(1, 2) + G(2, 1), G(2, 0) + G(0, 2)};
  for (int c = 0; c < cp; ++c) dy[v * cp + c] = bf16(c < 6 ? float(scale * t[c]) : 0.f);
}

// a_m += scale * dx[v][m] for the three velocity channels
template <typename T>
__global__ void k_add_vel_cot(const float* dx, long nvox, int cp, double scale, T* a, long stride) {
  long v = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= nvox) return;
  for (int m = 0; m < 3; ++m) a[m * stride + v] += T(scale * double(dx[v * cp + m]));
}

}  // namespace cnn

// ------------------------------------------------------------------ host side
struct sc_cnn {
  int n = 0, bmax = 0, cin = 0, width = 0, cout = 0;
  int cin_p = 8, cout_p = 16;
  hipStream_t stream = nullptr;
  bool own_stream = false;  // created here (destroyed by sc_cnn_destroy) or the caller's
  std::string err;
  // master weights / grads / Adam state, one flat buffer each; offsets per layer
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  long wlen = 0;
  long woff[4] = {}, wsz[4] = {}, boff[3] = {};
  bf16 *wf[4] = {}, *wb[4] = {};
  // activations for the largest batch: z1..z3, a1..a3 and their gradients
  bf16 *z[3] = {}, *a[3] = {}, *dz[3] = {};
  // one-cube input (one per activation slot) / output / cotangent buffers of
  // the LES closure (allocated on first use by the solver)
  bf16 *a0 = nullptr, *dyb = nullptr;
  float *y = nullptr, *dx = nullptr, *part = nullptr;
  long part_len = 0;
  double* lossp = nullptr;
  double* hostd = nullptr;
  double* gnorm = nullptr;  // per-array gradient norm partials, then the squared norms (NORM_DOUBLES)
  int adam_t = 0;
  bool tracefree = false;  // output channels 0-2 (xx, yy, zz) lose their mean (cout 6)
  int ncu = 256;  // compute units of the device (persistent wgrad grid)
  // per-launch HIP-event timing (sc_cnn_profile): name, events, executed flops
  struct Prof { const char* name; hipEvent_t a, b; double flop; };
  bool prof_on = false;
  std::vector<Prof> prof;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  hipEvent_t next_event() {
    if (ev_used == ev_pool.size()) {
      hipEvent_t e;
      CNN_HIP(hipEventCreate(&e));
      ev_pool.push_back(e);
    }
    return ev_pool[ev_used++];
  }
  void pbeg(const char* name, double flop) {
    if (!prof_on) return;
    Prof r{name, next_event(), next_event(), flop};
    CNN_HIP(hipEventRecord(r.a, stream));
    prof.push_back(r);
  }
  void pend() {
    if (prof_on) CNN_HIP(hipEventRecord(prof.back().b, stream));
  }
};

namespace {
using namespace cnn;
using namespace optim;
thread_local std::string g_cnn_create_err;

int layer_cin(const sc_cnn* c, int l) { return l == 0 ? c->cin_p : c->width; }
int layer_cout(const sc_cnn* c, int l) { return l == 3 ? c->cout_p : c->width; }
int real_cin(const sc_cnn* c, int l) { return l == 0 ? c->cin : c->width; }
int real_cout(const sc_cnn* c, int l) { return l == 3 ? c->cout : c->width; }

template <class F>
int cguard(sc_cnn* c, F&& f) {
  if (!c) return 1;
  (void)hipGetLastError();  // drop an error state left by unrelated HIP calls of the caller
  try {
    f();
    CNN_HIP(hipStreamSynchronize(c->stream));
    return 0;
  } catch (std::exception& e) {
    c->err = e.what();
    return 3;
  }
}

void launch_conv(sc_cnn* c, int cinp, int coutp, int mode, const bf16* in, const bf16* W, const float* bias,
                 bf16* oa, bf16* oz, float* of, const bf16* zp, int batch) {
  const long rows = long(batch) * c->n * c->n;

  static const char* names[4] = {"conv_fwd", "conv_last", "conv_bwd_data", "conv_bwd_input"};
  const bool pair = mode == LAST && coutp == 8 && cinp == 64;
  const double flop = pair ? double(rows) * c->n * 16 * 72 * 32  // half the rows, 72 k steps of 32
                           : 2.0 * double(rows) * c->n * kpad(cinp) * ((coutp + 15) / 16 * 16);
#define CONV(CI, COT, COS, MO)                                                                             \
  if (cinp == CI && coutp == COS && mode == MO) {                                                         \
    c->pbeg(names[MO], flop);                                                                             \
    const long ntiles = long(batch) * c->n * (c->n / conv_trows(COT)) * (c->n / VPB);                    \
    hipLaunchKernelGGL((k_conv<CI, COT, COS, MO>), dim3(unsigned(ntiles)), dim3(conv_threads(COT)), 0, c->stream, in, W, \
                       bias, oa, oz, of, zp, c->n, ntiles, c->tracefree ? 1 : 0);                        \
    CNN_HIP(hipGetLastError());                                                                           \
    c->pend();                                                                                            \
    return;                                                                                               \
  }
  CONV(8, 64, 64, HIDDEN) CONV(64, 64, 64, HIDDEN) CONV(64, 16, 16, LAST) CONV(64, 16, 8, LAST)
  CONV(16, 64, 64, BWD_HIDDEN) CONV(8, 64, 64, BWD_HIDDEN) CONV(64, 64, 64, BWD_HIDDEN) CONV(64, 16, 8, BWD_INPUT)
#undef CONV
  throw std::runtime_error("unsupported conv shape");
}

// row sets in flight per producer thread (registers: 32 VGPRs each)
#ifndef SC_WG_PD
#define SC_WG_PD 2
#endif
#ifndef SC_WG3_PD
#define SC_WG3_PD 2
#endif


void launch_wgrad(sc_cnn* c, int cinp, int coutp, const bf16* in, const bf16* dout, float* grad, float* db,
                  int batch) {
  const long nplanes = long(batch) * c->n;
  // the narrow layers take all 27 taps per block (one block per CU); the
  // 64 x 64 layer uses plane groups of three blocks: a multiple of 8 (one XCD
  // holds a group's three blocks), at most one block per CU
  const bool narrow = !(cinp >= 64 && coutp >= 64);
  const long G = narrow ? std::min<long>(nplanes, std::max(1, c->ncu)) : 8 * std::max(1, c->ncu / 24);
  const long len = 27L * coutp * cinp;
  if (c->part_len < G * len) {
    if (c->part) CNN_HIP(hipFree(c->part));
    CNN_HIP(hipMalloc(&c->part, G * len * sizeof(float)));
    c->part_len = G * len;
  }
  c->pbeg("conv_wgrad", 2.0 * double(nplanes) * c->n * c->n * 27.0 * ((cinp + 15) / 16 * 16) * coutp);
  if (cinp == 64 && coutp == 64) {
    hipLaunchKernelGGL((k_conv_wgrad<64, 64, SC_WG_PD>), dim3(unsigned(3 * G)), dim3(512), 0, c->stream, in, dout,
                       c->part, db, c->n, nplanes);
  } else if (cinp == 8 && coutp == 64 && c->cin <= 4) {
    hipLaunchKernelGGL((k_conv_wgrad3<8, 64, SC_WG3_PD, true>), dim3(unsigned(G)), dim3(64 * 8), 0, c->stream, in,
                       dout, c->part, db, c->n, nplanes);
  } else if (cinp == 8 && coutp == 64) {
    hipLaunchKernelGGL((k_conv_wgrad3<8, 64, SC_WG3_PD>), dim3(unsigned(G)), dim3(64 * 8), 0, c->stream, in, dout,
                       c->part, db, c->n, nplanes);
  } else if (cinp == 64 && coutp == 16) {
    hipLaunchKernelGGL((k_conv_wgrad3<64, 16, SC_WG3_PD>), dim3(unsigned(G)), dim3(64 * 8), 0, c->stream, in, dout,
                       c->part, db, c->n, nplanes);
  } else if (cinp == 64 && coutp == 8) {
    hipLaunchKernelGGL((k_conv_wgrad3<64, 8, SC_WG3_PD>), dim3(unsigned(G)), dim3(64 * 8), 0, c->stream, in, dout,
                       c->part, db, c->n, nplanes);
  } else {
    throw std::runtime_error("unsupported wgrad shape");
  }
  CNN_HIP(hipGetLastError());
  c->pend();
  c->pbeg("wgrad_reduce", 0);
  hipLaunchKernelGGL(k_reduce_partials, dim3(unsigned((len + 255) / 256)), dim3(256), 0, c->stream, c->part, grad, G,
                     len);
  CNN_HIP(hipGetLastError());
  c->pend();
}

// the last layer's forward fragments take the paired layout (k_conv PAIR)
// when it has at most 8 output channels
bool paired_last(const sc_cnn* c, int l) { return l == 3 && c->cout_p == 8 && layer_cin(c, l) == 64; }
long wf_len(const sc_cnn* c, int l) {
  return paired_last(c, l) ? 16L * 72 * 32 : long((layer_cout(c, l) + 15) / 16 * 16) * kpad(layer_cin(c, l));
}
long wb_len(const sc_cnn* c, int l) { return long((layer_cin(c, l) + 15) / 16 * 16) * kpad(layer_cout(c, l)); }

void pack_all(sc_cnn* c) {
  for (int l = 0; l < 4; ++l) {
    const long len = wf_len(c, l) + wb_len(c, l);
    hipLaunchKernelGGL(k_pack_weights, dim3(unsigned((len + 255) / 256)), dim3(256), 0, c->stream, c->w + c->woff[l],
                       c->wf[l], c->wb[l], layer_cout(c, l), layer_cin(c, l), (l == 3 && c->tracefree) ? 1 : 0,
                       paired_last(c, l) ? 1 : 0);
    CNN_HIP(hipGetLastError());
  }
}
}  // namespace

// ----------------------------------------------- interface used by the solver
namespace cnnlink {
using namespace cnn;

// forward / backward over `batch` samples whose activations start at sample
// slot s0 of the batch buffers
static void forward_at(sc_cnn* c, const void* x, float* y, int batch, int s0) {
  if (batch < 1 || s0 < 0 || s0 + batch > c->bmax) throw std::runtime_error("batch out of range");
  const long off = long(s0) * c->n * c->n * c->n * c->width;
  const bf16* in = (const bf16*)x;
  for (int l = 0; l < 3; ++l) {
    launch_conv(c, layer_cin(c, l), c->width, HIDDEN, in, c->wf[l], c->w + c->boff[l], c->a[l] + off, c->z[l] + off,
                nullptr, nullptr, batch);
    in = c->a[l] + off;
  }
  launch_conv(c, c->width, c->cout_p, LAST, in, c->wf[3], nullptr, nullptr, nullptr, y, nullptr, batch);
}

static void backward_at(sc_cnn* c, const void* x, const void* dy, float* dx, int batch, int s0) {
  if (batch < 1 || s0 < 0 || s0 + batch > c->bmax) throw std::runtime_error("batch out of range");
  const long off = long(s0) * c->n * c->n * c->n * c->width;
  const bf16* dY = (const bf16*)dy;
  // last layer: dW4 (projected when the output is trace-free), then the data
  // gradient into dz[2] with gelu'(z3)
  launch_wgrad(c, c->width, c->cout_p, c->a[2] + off, dY, c->g + c->woff[3], nullptr, batch);
  if (c->tracefree) {
    hipLaunchKernelGGL(k_tracefree_rows, dim3(unsigned((27 * c->width + 255) / 256)), dim3(256), 0, c->stream,
                       c->g + c->woff[3], c->cout_p, c->width);
    CNN_HIP(hipGetLastError());
  }
  launch_conv(c, c->cout_p, c->width, BWD_HIDDEN, dY, c->wb[3], nullptr, c->dz[2], nullptr, nullptr,
              c->z[2] + off, batch);
  for (int l = 2; l >= 0; --l) {
    const bf16* inl = l == 0 ? (const bf16*)x : c->a[l - 1] + off;
    launch_wgrad(c, layer_cin(c, l), c->width, inl, c->dz[l], c->g + c->woff[l], c->g + c->boff[l], batch);
    if (l > 0)
      launch_conv(c, c->width, c->width, BWD_HIDDEN, c->dz[l], c->wb[l], nullptr, c->dz[l - 1], nullptr, nullptr,
                  c->z[l - 1] + off, batch);
    else if (dx)
      launch_conv(c, c->width, c->cin_p, BWD_INPUT, c->dz[0], c->wb[0], nullptr, nullptr, nullptr, dx, nullptr, batch);
  }
}

void forward(sc_cnn* c, const void* x, float* y, int batch) { forward_at(c, x, y, batch, 0); }

void backward(sc_cnn* c, const void* x, const void* dy, float* dx, int batch) { backward_at(c, x, dy, dx, batch, 0); }

int slots(sc_cnn* c) { return std::min(4, c->bmax); }

void* slot_x(sc_cnn* c, int s) {
  if (s < 0 || s >= slots(c)) throw std::runtime_error("activation slot out of range");
  buffers(c);
  return c->a0 + long(s) * c->n * c->n * c->n * c->cin_p;
}

void forward_slot(sc_cnn* c, int s, float* y) { forward_at(c, slot_x(c, s), y, 1, s); }

void backward_slot(sc_cnn* c, int s, const void* dy, float* dx) { backward_at(c, slot_x(c, s), dy, dx, 1, s); }

int grid_n(const sc_cnn* c) { return c->n; }
int in_channels(const sc_cnn* c) { return c->cin; }
int out_channels(const sc_cnn* c) { return c->cout; }

Bufs buffers(sc_cnn* c) {
  const long nv = long(c->n) * c->n * c->n;
  if (!c->a0) {
    CNN_HIP(hipMalloc(&c->a0, nv * c->cin_p * 2 * slots(c)));  // one input per activation slot
    CNN_HIP(hipMalloc(&c->y, nv * c->cout_p * 4));
    CNN_HIP(hipMalloc(&c->dyb, nv * c->cout_p * 2));
    CNN_HIP(hipMalloc(&c->dx, nv * c->cin_p * 4));
  }
  return Bufs{c->a0, c->y, c->dyb, c->dx};
}

hipStream_t set_stream(sc_cnn* c, hipStream_t s) {
  hipStream_t old = c->stream;
  c->stream = s;
  return old;
}

static unsigned nb1(long nv) { return unsigned((nv + 255) / 256); }

template <typename T>
void pack_velocity(sc_cnn* c, const T* v, long stride, double scale, double extra, void* x) {
  const long nv = long(c->n) * c->n * c->n;
  const int has_extra = c->cin > 3 ? 1 : 0;
  hipLaunchKernelGGL(k_pack_input<T>, dim3(nb1(nv)), dim3(256), 0, c->stream, v, nv, 3, scale, has_extra,
                     float(extra), (bf16*)x, c->cin_p, stride);
  CNN_HIP(hipGetLastError());
}

template <typename T>
void unpack_tau(sc_cnn* c, const float* y, double scale, T* tau, long stride) {
  const long nv = long(c->n) * c->n * c->n;
  hipLaunchKernelGGL(k_unpack<T>, dim3(nb1(nv)), dim3(256), 0, c->stream, y, nv, 6, c->cout_p, scale, tau, stride);
  CNN_HIP(hipGetLastError());
}

template <typename T>
void pack_tau_cotangent(sc_cnn* c, const T* g9, long stride, double scale, void* dy) {
  const long nv = long(c->n) * c->n * c->n;
  hipLaunchKernelGGL(k_pack_tau_cot<T>, dim3(nb1(nv)), dim3(256), 0, c->stream, g9, nv, stride, scale, (bf16*)dy,
                     c->cout_p);
  CNN_HIP(hipGetLastError());
}

template <typename T>
void add_velocity_cotangent(sc_cnn* c, const float* dx, double scale, T* a, long stride) {
  const long nv = long(c->n) * c->n * c->n;
  hipLaunchKernelGGL(k_add_vel_cot<T>, dim3(nb1(nv)), dim3(256), 0, c->stream, dx, nv, c->cin_p, scale, a, stride);
  CNN_HIP(hipGetLastError());
}

template void pack_velocity<float>(sc_cnn*, const float*, long, double, double, void*);
template void pack_velocity<double>(sc_cnn*, const double*, long, double, double, void*);
template void unpack_tau<float>(sc_cnn*, const float*, double, float*, long);
template void unpack_tau<double>(sc_cnn*, const float*, double, double*, long);
template void pack_tau_cotangent<float>(sc_cnn*, const float*, long, double, void*);
template void pack_tau_cotangent<double>(sc_cnn*, const double*, long, double, void*);
template void add_velocity_cotangent<float>(sc_cnn*, const float*, double, float*, long);
template void add_velocity_cotangent<double>(sc_cnn*, const float*, double, double*, long);

}  // namespace cnnlink

using namespace optim;

extern "C" {

int sc_cnn_create(int n, int batch_max, int cin, int width, int cout, uint64_t seed, sc_cnn** out) {
  if (!out) return 1;
  *out = nullptr;
  if (n % VPB != 0 || n > 128 || batch_max < 1 || cin < 1 || cin > 8 || width != 64 || cout < 1 || cout > 16)
    return 1;
  g_cnn_create_err.clear();
  (void)hipGetLastError();
  sc_cnn* c = new sc_cnn();
  c->n = n; c->bmax = batch_max; c->cin = cin; c->width = width; c->cout = cout;
  // output channel pitch: 8 when it suffices, so the loss gradient and the
  // last layer's data and weight gradients carry 8 channels, not 16 (the
  // forward's MFMA tile stays 16 wide and stores 8)
  c->cout_p = cout <= 8 ? 8 : 16;
  try {
    CNN_HIP(hipStreamCreate(&c->stream));
    c->own_stream = true;
    {
      int dev = 0;
      CNN_HIP(hipGetDevice(&dev));
      CNN_HIP(hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, dev));
    }
    long off = 0;
    for (int l = 0; l < 4; ++l) {
      c->woff[l] = off;
      c->wsz[l] = 27L * layer_cout(c, l) * layer_cin(c, l);
      off += c->wsz[l];
    }
    for (int l = 0; l < 3; ++l) { c->boff[l] = off; off += width; }
    c->wlen = off;
    // Glorot-uniform-style init from a splitmix64 stream (fan_in + fan_out = 27 (cin + cout))
    std::vector<float> h(c->wlen, 0.f);
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 12345;
    auto rnd = [&]() {
      s += 0x9E3779B97F4A7C15ull;
      uint64_t zz = s;
      zz = (zz ^ (zz >> 30)) * 0xBF58476D1CE4E5B9ull;
      zz = (zz ^ (zz >> 27)) * 0x94D049BB133111EBull;
      zz ^= zz >> 31;
      return double(zz >> 11) * 0x1.0p-53;
    };
    for (int l = 0; l < 4; ++l) {
      double lim = std::sqrt(6.0 / (27.0 * (real_cin(c, l) + real_cout(c, l))));
      const int cop = layer_cout(c, l), cip = layer_cin(c, l);
      for (int t = 0; t < 27; ++t)
        for (int co = 0; co < real_cout(c, l); ++co)
          for (int ci = 0; ci < real_cin(c, l); ++ci)
            h[c->woff[l] + (long(t) * cop + co) * cip + ci] = float((2 * rnd() - 1) * lim);
    }
    CNN_HIP(hipMalloc(&c->w, c->wlen * 4));
    CNN_HIP(hipMalloc(&c->g, c->wlen * 4));
    CNN_HIP(hipMalloc(&c->m, c->wlen * 4));
    CNN_HIP(hipMalloc(&c->v, c->wlen * 4));
    CNN_HIP(hipMemcpy(c->w, h.data(), c->wlen * 4, hipMemcpyHostToDevice));
    CNN_HIP(hipMemset(c->g, 0, c->wlen * 4));
    CNN_HIP(hipMemset(c->m, 0, c->wlen * 4));
    CNN_HIP(hipMemset(c->v, 0, c->wlen * 4));
    for (int l = 0; l < 4; ++l) {
      CNN_HIP(hipMalloc(&c->wf[l], wf_len(c, l) * 2));
      CNN_HIP(hipMalloc(&c->wb[l], wb_len(c, l) * 2));
    }
    const long nv = long(batch_max) * n * n * n;
    for (int l = 0; l < 3; ++l) {
      CNN_HIP(hipMalloc(&c->z[l], nv * width * 2));
      CNN_HIP(hipMalloc(&c->a[l], nv * width * 2));
      CNN_HIP(hipMalloc(&c->dz[l], nv * width * 2));
    }
    CNN_HIP(hipMalloc(&c->lossp, ((nv + 255) / 256) * sizeof(double) + 16));
    CNN_HIP(hipMalloc(&c->gnorm, NORM_DOUBLES * sizeof(double)));
    CNN_HIP(hipHostMalloc(&c->hostd, 64));
    pack_all(c);
    CNN_HIP(hipStreamSynchronize(c->stream));
  } catch (std::exception& e) {
    g_cnn_create_err = e.what();
    sc_cnn_destroy(c);
    return 3;
  }
  *out = c;
  return 0;
}

int sc_cnn_destroy(sc_cnn* c) {
  if (!c) return 0;
  (void)hipStreamSynchronize(c->stream);
  void* ptrs[] = {c->w, c->g, c->m, c->v, c->a0, c->y, c->dyb, c->dx, c->part, c->lossp, c->gnorm};
  for (void* p : ptrs) if (p) (void)hipFree(p);
  for (int l = 0; l < 4; ++l) { if (c->wf[l]) (void)hipFree(c->wf[l]); if (c->wb[l]) (void)hipFree(c->wb[l]); }
  for (int l = 0; l < 3; ++l) {
    if (c->z[l]) (void)hipFree(c->z[l]);
    if (c->a[l]) (void)hipFree(c->a[l]);
    if (c->dz[l]) (void)hipFree(c->dz[l]);
  }
  if (c->hostd) (void)hipHostFree(c->hostd);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->own_stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

const char* sc_cnn_last_error(const sc_cnn* c) { return c ? c->err.c_str() : g_cnn_create_err.c_str(); }

int sc_cnn_info(const sc_cnn* c, int64_t* info) {
  if (!c || !info) return 1;
  info[0] = c->n; info[1] = c->bmax; info[2] = c->cin; info[3] = c->width; info[4] = c->cout;
  info[5] = c->cin_p; info[6] = c->cout_p; info[7] = c->wlen;
  return 0;
}

// FP32 master parameters: all conv weights ([27][cout_p][cin_p] per layer,
// padding entries zero) then the three hidden biases
int sc_cnn_get_params(sc_cnn* c, float* host) {
  return cguard(c, [&] { CNN_HIP(hipMemcpyAsync(host, c->w, c->wlen * 4, hipMemcpyDeviceToHost, c->stream)); });
}
int sc_cnn_set_params(sc_cnn* c, const float* host) {
  return cguard(c, [&] {
    CNN_HIP(hipMemcpyAsync(c->w, host, c->wlen * 4, hipMemcpyHostToDevice, c->stream));
    pack_all(c);
  });
}
int sc_cnn_get_grads(sc_cnn* c, float* host) {
  return cguard(c, [&] { CNN_HIP(hipMemcpyAsync(host, c->g, c->wlen * 4, hipMemcpyDeviceToHost, c->stream)); });
}
int sc_cnn_zero_grad(sc_cnn* c) {
  return cguard(c, [&] { CNN_HIP(hipMemsetAsync(c->g, 0, c->wlen * 4, c->stream)); });
}

// input: device bf16 [batch][n^3][cin_p] channels-last (see sc_cnn_pack_input)
// output: device fp32 [batch][n^3][cout_p]
int sc_cnn_forward(sc_cnn* c, const void* x, void* y, int batch) {
  return cguard(c, [&] { cnnlink::forward(c, x, (float*)y, batch); });
}

// backward of the last sc_cnn_forward: dy device bf16 [batch][n^3][cout_p];
// accumulates parameter gradients; dx (nullable) receives fp32 [batch][n^3][cin_p]
int sc_cnn_backward(sc_cnn* c, const void* x, const void* dy, void* dx, int batch) {
  return cguard(c, [&] { cnnlink::backward(c, x, dy, (float*)dx, batch); });
}

// Loss of y (fp32 [batch][n^3][cout_p]) against tgt (fp32 [batch][n^3][cout])
// over the first cout channels; writes dy = dL/dy (bf16 [batch][n^3][cout_p])
// and returns the loss.  kind 0: mean squared error; kind 1: the relative
// error sum|y - t|^2 / (sum|t|^2 + eps_f32) of the whole batch.
int sc_cnn_loss(sc_cnn* c, const void* y, const void* tgt, void* dy, int batch, int kind, double* loss) {
  return cguard(c, [&] {
    if (kind != 0 && kind != 1) throw std::runtime_error("loss kind must be 0 (mse) or 1 (relative)");
    if (batch < 1 || batch > c->bmax) throw std::runtime_error("batch out of range");
    const long nv = long(batch) * c->n * c->n * c->n;
    const unsigned nb = unsigned((nv + 255) / 256);
    double inv = 1.0 / (double(nv) * c->cout);
    if (kind == 1) {
      // sum of squares of the target first (it fixes the gradient's scale)
      const unsigned ns = std::min(nb, 1024u);
      hipLaunchKernelGGL(k_sumsq, dim3(ns), dim3(256), 0, c->stream, (const float*)tgt, nv * c->cout, c->lossp);
      CNN_HIP(hipGetLastError());
      std::vector<double> h(ns);
      CNN_HIP(hipMemcpyAsync(h.data(), c->lossp, ns * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      CNN_HIP(hipStreamSynchronize(c->stream));
      double s = 0;
      for (double v : h) s += v;
      inv = 1.0 / (s + double(FLT_EPSILON));
    }
    c->pbeg("loss", 0);
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, c->stream, (const float*)y, (const float*)tgt, (bf16*)dy, nullptr,
                       c->lossp, nv, c->cout, c->cout_p, inv);
    CNN_HIP(hipGetLastError());
    c->pend();
    std::vector<double> h(nb);
    CNN_HIP(hipMemcpyAsync(h.data(), c->lossp, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CNN_HIP(hipStreamSynchronize(c->stream));
    double s = 0;
    for (double v : h) s += v;
    if (loss) *loss = s * inv;
  });
}

int sc_cnn_mse(sc_cnn* c, const void* y, const void* tgt, void* dy, int batch, double* loss) {
  return sc_cnn_loss(c, y, tgt, dy, batch, 0, loss);
}

// One optimiser step on the accumulated gradient: per-array ClipNorm(clip)
// (clip <= 0: none), then AdamW (weight_decay 0: Adam); the gradient is kept.
int sc_cnn_adamw(sc_cnn* c, double lr, double b1, double b2, double eps, double weight_decay, double clip) {
  return cguard(c, [&] {
    c->adam_t += 1;
    const float c1 = float(1.0 - std::pow(b1, c->adam_t)), c2 = float(1.0 - std::pow(b2, c->adam_t));
    Arrays A;
    for (int l = 0; l < 4; ++l) A.b[l] = c->woff[l];
    for (int l = 0; l < 3; ++l) A.b[4 + l] = c->boff[l];
    A.b[7] = c->wlen;
    A.n = 7;
    double* nrm2 = c->gnorm + NORM_OFFSET;
    if (clip > 0) {
      hipLaunchKernelGGL(k_array_sumsq, dim3(NORM_BLOCKS, A.n), dim3(256), 0, c->stream, c->g, A, c->gnorm);
      CNN_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_finish_norms, dim3(1), dim3(64), 0, c->stream, c->gnorm, nrm2, A.n);
      CNN_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_adamw, dim3(unsigned((c->wlen + 255) / 256)), dim3(256), 0, c->stream, c->w, c->g, c->m, c->v,
                       c->wlen, float(lr), float(b1), float(b2), float(eps), float(weight_decay), c1, c2, nrm2, A,
                       clip);
    CNN_HIP(hipGetLastError());
    pack_all(c);
  });
}

int sc_cnn_adam(sc_cnn* c, double lr, double b1, double b2, double eps) {
  return sc_cnn_adamw(c, lr, b1, b2, eps, 0.0, 0.0);
}

// Trace removal on the output (needs cout 6: xx, yy, zz, xy, yz, zx)
int sc_cnn_set_tracefree(sc_cnn* c, int on) {
  return cguard(c, [&] {
    if (on && c->cout != 6) throw std::runtime_error("trace removal needs cout = 6 (xx, yy, zz, xy, yz, zx)");
    c->tracefree = on != 0;
    pack_all(c);
  });
}

// pack ncomp physical fields (component stride cstride elements, one cube per
// batch entry at batch stride bstride) into the channels-last bf16 input
int sc_cnn_pack_input(sc_cnn* c, const void* f, int fp64, int ncomp, long cstride, long bstride, int batch,
                      double scale, int has_extra, double extra, void* x) {
  return cguard(c, [&] {
    const long nv = long(c->n) * c->n * c->n;
    for (int b = 0; b < batch; ++b) {
      bf16* out = (bf16*)x + long(b) * nv * c->cin_p;
      if (fp64)
        hipLaunchKernelGGL(k_pack_input<double>, dim3(unsigned((nv + 255) / 256)), dim3(256), 0, c->stream,
                           (const double*)f + b * bstride, nv, ncomp, scale, has_extra, float(extra), out, c->cin_p,
                           cstride);
      else
        hipLaunchKernelGGL(k_pack_input<float>, dim3(unsigned((nv + 255) / 256)), dim3(256), 0, c->stream,
                           (const float*)f + b * bstride, nv, ncomp, scale, has_extra, float(extra), out, c->cin_p,
                           cstride);
      CNN_HIP(hipGetLastError());
    }
  });
}

int sc_cnn_pack_target(sc_cnn* c, const void* f, int fp64, long cstride, long bstride, int batch, double scale,
                       void* out) {
  return cguard(c, [&] {
    const long nv = long(c->n) * c->n * c->n;
    for (int b = 0; b < batch; ++b) {
      float* o = (float*)out + long(b) * nv * c->cout;
      if (fp64)
        hipLaunchKernelGGL(k_pack_target<double>, dim3(unsigned((nv + 255) / 256)), dim3(256), 0, c->stream,
                           (const double*)f + b * bstride, nv, c->cout, scale, o, cstride);
      else
        hipLaunchKernelGGL(k_pack_target<float>, dim3(unsigned((nv + 255) / 256)), dim3(256), 0, c->stream,
                           (const float*)f + b * bstride, nv, c->cout, scale, o, cstride);
      CNN_HIP(hipGetLastError());
    }
  });
}

int sc_cnn_profile(sc_cnn* c, int enable) {
  return cguard(c, [&] {
    c->prof.clear();
    c->ev_used = 0;
    c->prof_on = enable != 0;
  });
}

// one line per launch name: "name count total_ms executed_flop"
int sc_cnn_profile_report(sc_cnn* c, char* buf, int buflen) {
  return cguard(c, [&] {
    if (!buf || buflen < 1) throw std::runtime_error("buffer");
    CNN_HIP(hipStreamSynchronize(c->stream));
    std::vector<std::string> names;
    std::vector<double> tot, flop;
    std::vector<long> cnt;
    for (auto& r : c->prof) {
      float ms = 0;
      CNN_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      size_t i = 0;
      while (i < names.size() && names[i] != r.name) ++i;
      if (i == names.size()) { names.push_back(r.name); tot.push_back(0); cnt.push_back(0); flop.push_back(0); }
      tot[i] += ms;
      cnt[i] += 1;
      flop[i] += r.flop;
    }
    std::string out;
    char line[256];
    for (size_t i = 0; i < names.size(); ++i) {
      std::snprintf(line, sizeof line, "%s %ld %.6f %.6e\n", names[i].c_str(), cnt[i], tot[i], flop[i]);
      out += line;
    }
    std::strncpy(buf, out.c_str(), size_t(buflen) - 1);
    buf[buflen - 1] = 0;
  });
}

int sc_cnn_set_stream(sc_cnn* c, void* s) {
  return cguard(c, [&] {
    CNN_HIP(hipStreamSynchronize(c->stream));
    if (c->own_stream) CNN_HIP(hipStreamDestroy(c->stream));
    c->stream = (hipStream_t)s;
    c->own_stream = false;
  });
}

}  // extern "C"
