// Pointwise MLP closure: the reference's `mlp` learned closure and its
// `fullchain` wrapper (src/nets.jl:37-81,247-334; src/data.jl:689-741).
//
//   x = [∇ū / (|∇ū| + eps) , z]          (9 or 4 gradient channels, z = the
//                                          standardised log Re_Δ when used)
//   h1 = gelu(W1ᵀ x + b1), h2 = gelu(W2ᵀ h1 + b2), h3 = gelu(W3ᵀ h2 + b3)
//   y = symm(W4ᵀ h3)                      (trace removal of xx, yy[, zz])
//   τ = Δ² |∇ū|² y                        (packed xx, yy, zz, xy, yz, zx)
//
// with |∇ū|² = Σ_ij G_ij² pointwise and eps = eps(Float64).  Training uses the
// dataloader's pairs (x, y = τ / (Δ² |∇ū|² + eps)) in FP32.  Layout: voxel
// matrices are channels-fastest, [nvox][c] (= column-major c x nvox); weights
// are Lux's (c_in, c_out) column-major (c_in fastest), so the flat parameter
// vector [W1, b1, W2, b2, W3, b3, W4] is Lux's own parameter order.  Training
// runs every layer as a GEMM over the voxels (k_dense: bias + GELU fused into
// the forward's store, GELU' into the data gradient's, the weight gradient
// split over voxel chunks and summed in order); the FP64 inference of the
// closure runs the fused MFMA kernels below, or k_dense per layer for
// networks too wide for them.  The output trace removal, the feature
// normalisation, the loss and the optimiser are this file's kernels too.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>
#include "../../include/symcode.h"
#include "pmlp_link.hpp"
#include "optim.hpp"

namespace pmlp {

#define PM_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace {

// tanh-form GELU (NNlib's gelu) and its derivative
template <typename E>
__device__ __forceinline__ E gelu(E z) {
  const E u = E(0.7978845608028654) * (z + E(0.044715) * z * z * z);
  return E(0.5) * z * (E(1) + tanh(u));
}
__device__ __forceinline__ float gelu_grad(float z) {
  const float u = 0.7978845608028654f * (z + 0.044715f * z * z * z);
  const float t = tanhf(u);
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * z * z);
}

__device__ __forceinline__ long gid() { return long(blockIdx.x) * blockDim.x + threadIdx.x; }

// create_dataloader's normalised pair of one voxel set (src/data.jl:729-741):
// x[v][c] = G_c / (sqrt(A2) + eps), x[v][ng] = z (with has_re), and (tau
// non-null) y[v][c] = tau_c / (Δ² A2 + eps); A2 = Σ_c G_c², eps = eps(T), all
// in FP64, stored in the FP32 training precision
template <typename T>
__global__ void k_features(const T* GG, long np, int ng, double d2, int has_re, float z, const T* tau, int nt,
                           float* x, int nin, float* y) {
  const long v = gid();
  if (v >= np) return;
  const double eps = double(std::numeric_limits<T>::epsilon());
  double a2 = 0;
  for (int c = 0; c < ng; ++c) {
    const double g = double(GG[c * np + v]);
    a2 += g * g;
  }
  const double den = sqrt(a2) + eps;
  for (int c = 0; c < ng; ++c) x[v * nin + c] = float(double(GG[c * np + v]) / den);
  if (has_re) x[v * nin + ng] = z;
  if (tau) {
    const double r = d2 * a2 + eps;
    for (int c = 0; c < nt; ++c) y[v * nt + c] = float(double(tau[c * np + v]) / r);
  }
}

// fullchain's input (src/nets.jl:317-325) for voxels [v0, v0 + nv) of the
// field, FP64: x[i][c] = G_c / (sqrt(A2) + eps(T)), x[i][ng] = *z when z
template <typename T>
__global__ void k_features_inf(const T* GG, long np, long v0, long nv, int ng, const double* z, double* x, int nin) {
  const long i = gid();
  if (i >= nv) return;
  const long v = v0 + i;
  double a2 = 0;
  for (int c = 0; c < ng; ++c) {
    const double g = double(GG[c * np + v]);
    a2 += g * g;
  }
  const double den = sqrt(a2) + double(std::numeric_limits<T>::epsilon());
  for (int c = 0; c < ng; ++c) x[i * nin + c] = double(GG[c * np + v]) / den;
  if (z) x[i * nin + ng] = *z;
}

// block partial sums of A2 = Σ_c G_c² over the field (FP64)
template <typename T>
__global__ void k_a2_partials(const T* GG, long np, int ng, double* part) {
  __shared__ double sm[256];
  double acc = 0;
  for (long v = gid(); v < np; v += long(gridDim.x) * blockDim.x)
    for (int c = 0; c < ng; ++c) {
      const double g = double(GG[c * np + v]);
      acc += g * g;
    }
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// the live standardised log Re_Δ of fullchain: Re_Δ = Δ² sqrt(mean A2) / ν,
// z = (log Re_Δ - μ) / σ, rounded to the field precision (append_redelta,
// src/data.jl:689-694)
// one wavefront: lane l adds partials l, l + 64, ..., then an xor butterfly
// (every lane ends with the same sum, a + b = b + a exactly)
__global__ void k_redelta_z(const double* part, int nblk, long np, double d2, double visc, double mu, double sigma,
                            int fp32, double* z) {
  if (blockIdx.x != 0 || threadIdx.x >= 64) return;
  const int l = threadIdx.x;
  double s = 0;
  for (int i = l; i < nblk; i += 64) s += part[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (l != 0) return;
  const double re = d2 * sqrt(s / double(np)) / visc;
  const double zz = (log(re) - mu) / sigma;
  *z = fp32 ? double(float(zz)) : zz;
}

__global__ void k_widen(const float* p, double* q, long n) {
  const long i = gid();
  if (i < n) q[i] = double(p[i]);
}


// ------------------------------------------ fused FP64 inference on MFMA
// The four layers of the network on v_mfma_f64_16x16x4f64 (D[i][j] +=
// A[i][k] B[k][j]; lane l holds A[l%16][l/16], B[l/16][l%16] and
// D[l/16 + 4r][l%16], r = 0..3), 16 input rows per row tile: i runs over 16
// output channels, j over the rows, k over 4 input channels.  Tile m of a
// layer's result leaves lane (c = l/16, v = l%16) holding channels
// 16m + c + 4r of row v, which is the B operand of the next layer's K step
// (tile m, r) when that step's k index c stands for channel 16m + c + 4r; the
// weights W[co][ci] (the A operand) are read from LDS in that order.  The
// activations stay in registers from the input row to the output, so only
// the inputs, the outputs and the weights (once per block) move through
// memory; every width is at most FT tiles of 16 channels.
typedef double d4_t __attribute__((ext_vector_type(4)));
constexpr int FT = 4;  // channel tiles per layer (widths up to 64)
#ifndef SC_PMLP_FR
#define SC_PMLP_FR 2
#endif
constexpr int FR = SC_PMLP_FR;  // row tiles per wavefront pass

// padded network in LDS: layer l's weights as rows[l] (input channels padded
// to 4) rows of 16 TOM doubles (output channel co at column co, TOM = the
// widest layer in tiles of 16) at wofs[l], the biases of the hidden layers
// (16 to[l]) at bofs[l]; padding is zero, so padded channels stay exactly
// zero through the layers (gelu(0) = 0).  The fixed row length makes every
// fragment's LDS offset a constant.
struct FusedNet {
  int cin[4], cout[4], ti[4], to[4], rows[4], wofs[4], bofs[3];
  long wsrc[4], bsrc[3];  // offsets of W (co x ci, ci fastest) and b in pd
};

// the tanh-form GELU with 0.5 (1 + tanh(u)) = 1 / (1 + exp(-2u)): the same
// function as gelu() above, one exp instead of a tanh; 1 / (1 + e) by the
// hardware reciprocal and two Newton steps (within an ulp or two of the
// division), the exponent capped so that 1 + e stays finite
__device__ __forceinline__ double gelu_exp(double z) {
  const double u = 0.7978845608028654 * (z + 0.044715 * z * z * z);
  const double d = 1.0 + exp(fmin(-2.0 * u, 708.0));
  double r = __builtin_amdgcn_rcp(d);
  r = fma(fma(-d, r, 1.0), r, r);
  r = fma(fma(-d, r, 1.0), r, r);
  return z * r;
}

// orders a wavefront's LDS stores before its later loads of the same words
// by other lanes (and the loads before later stores)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// acc[rt][m] = sum over the input channels of W[16m + v][ci] in[rt][ci-tile];
// wv = this lane's (c, v) entry of the layer's first LDS row
template <int TOM>
__device__ __forceinline__ void fused_layer(const double* wv, int ti, int cin, int to, const d4_t (&in)[FR][TOM],
                                            d4_t (&acc)[FR][TOM]) {
#pragma unroll
  for (int m = 0; m < TOM; ++m)
#pragma unroll
    for (int rt = 0; rt < FR; ++rt) acc[rt][m] = d4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < TOM; ++t) {
    if (t >= ti) break;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      // channels 16t + 4r + (0..3) of this K step; none is live past cin
      if (16 * t + 4 * r >= cin) break;
#pragma unroll
      for (int m = 0; m < TOM; ++m) {
        if (m >= to) break;
        const double a = wv[(16 * t + 4 * r) * 16 * TOM + 16 * m];
#pragma unroll
        for (int rt = 0; rt < FR; ++rt)
          acc[rt][m] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, in[rt][t][r], acc[rt][m], 0, 0, 0);
      }
    }
  }
}

// h = gelu(h + b) on the live channels (co < cout); bc = this lane's b[c].
// Channel groups 16m + 4r + (0..3) wholly past cout hold exact zeros
// (zero weights and bias) and gelu(0) = 0, so they are skipped.
template <int TOM>
__device__ __forceinline__ void fused_bias_gelu(d4_t (&h)[FR][TOM], const double* bc, int cout) {
#pragma unroll
  for (int m = 0; m < TOM; ++m) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (16 * m + 4 * r >= cout) break;
      const double bb = bc[16 * m + 4 * r];
#pragma unroll
      for (int rt = 0; rt < FR; ++rt) h[rt][m][r] = gelu_exp(h[rt][m][r] + bb);
    }
  }
}

// the padded network from pd into LDS (the whole block)
template <int TOM>
__device__ __forceinline__ void stage_net(double* wl, const double* __restrict__ pd, const FusedNet& f) {
  constexpr int COLS = 16 * TOM;
  for (int l = 0; l < 4; ++l) {
    const int ci = f.cin[l], co = f.cout[l];
    for (int i = threadIdx.x; i < f.rows[l] * COLS; i += blockDim.x) {
      const int a = i / COLS, b = i - a * COLS;
      double w = 0.0;
      if (a < ci && b < co) w = pd[f.wsrc[l] + a + long(ci) * b];
      wl[f.wofs[l] + i] = w;
    }
    if (l < 3)
      for (int i = threadIdx.x; i < 16 * f.to[l]; i += blockDim.x) {
        double b = 0.0;
        if (i < co) b = pd[f.bsrc[l] + i];
        wl[f.bofs[l] + i] = b;
      }
  }
}

// y (nrows x cout[3]) = net(x) for x (nrows rows of leading dimension ldx,
// the first cin[0] columns read); LDSD doubles of LDS hold the padded network,
// TOM tiles of 16 channels the widest layer.  Grid-stride over passes of FR
// row tiles per wavefront.
template <int TOM, int LDSD>
__global__ void __launch_bounds__(256) k_mlp_fused(const double* __restrict__ pd, FusedNet f,
                                                   const double* __restrict__ x, int ldx, long nrows,
                                                   double* __restrict__ y) {
  __shared__ double wl[LDSD];
  constexpr int COLS = 16 * TOM;
  stage_net<TOM>(wl, pd, f);
  __syncthreads();
  const int lane = threadIdx.x & 63, c = lane >> 4, v = lane & 15;
  const int nw = blockDim.x >> 6;
  const long npass = (nrows + 16 * FR - 1) / (16 * FR);
  const int co3 = f.cout[3];
  const double* wlv = wl + c * COLS + v;
  for (long p = long(blockIdx.x) * nw + (threadIdx.x >> 6); p < npass; p += long(gridDim.x) * nw) {
    d4_t A[FR][TOM], B[FR][TOM];
#pragma unroll
    for (int rt = 0; rt < FR; ++rt) {
      const long row = (p * FR + rt) * 16 + v;
      const double* xr = x + row * ldx + c;
#pragma unroll
      for (int t = 0; t < TOM; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          double e = 0.0;
          if (16 * t + c + 4 * r < f.cin[0] && row < nrows) e = xr[16 * t + 4 * r];
          A[rt][t][r] = e;
        }
    }
    fused_layer<TOM>(wlv + f.wofs[0], f.ti[0], f.cin[0], f.to[0], A, B);
    fused_bias_gelu<TOM>(B, wl + f.bofs[0] + c, f.cout[0]);
    fused_layer<TOM>(wlv + f.wofs[1], f.ti[1], f.cin[1], f.to[1], B, A);
    fused_bias_gelu<TOM>(A, wl + f.bofs[1] + c, f.cout[1]);
    fused_layer<TOM>(wlv + f.wofs[2], f.ti[2], f.cin[2], f.to[2], A, B);
    fused_bias_gelu<TOM>(B, wl + f.bofs[2] + c, f.cout[2]);
    fused_layer<TOM>(wlv + f.wofs[3], f.ti[3], f.cin[3], f.to[3], B, A);
#pragma unroll
    for (int rt = 0; rt < FR; ++rt) {
      const long row = (p * FR + rt) * 16 + v;
      if (row >= nrows) continue;
#pragma unroll
      for (int m = 0; m < TOM; ++m) {
        if (m >= f.to[3]) break;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int co = 16 * m + c + 4 * r;
          if (co < co3) y[row * co3 + co] = A[rt][m][r];
        }
      }
    }
  }
}


// ------------------------------------------------------- dense layers
// C[n ldc + m] = epilogue(sum_k A(m, k) B(k, n)) for m < M, n < N, over a
// k range: a register-blocked GEMM on the VALU, one 64 x 64 tile of C per
// block of 256 threads (4 x 4 per thread), the A and B tiles of 16 k staged
// in LDS.  The layer shapes are thin (M, K at most a few hundred channels,
// N the voxels), so the layers are bound by their activations' HBM traffic
// and the epilogue that a separate pass would add; fusing the bias and the
// activation into the store removes that pass.
//   A(m, k) = A[m lda + k] (AT false) or A[k lda + m] (AT true)
//   B(k, n) = B[n ldb + k] (BT false) or B[k ldb + n] (BT true)
// Grid: (N tiles, M tiles, k splits).  Split over k: block z takes k in
// [z kc, min(K, (z + 1) kc)) and writes the partial tile at C + z N ldc
// (DN_PLAIN only).
enum DenseMode { DN_PLAIN = 0, DN_BIAS_GELU_TRAIN = 1, DN_BIAS_GELU = 2, DN_GELU_BACK = 3 };
//   DN_BIAS_GELU_TRAIN: z = sum + b[m] to C, gelu(z) to C2
//   DN_BIAS_GELU:       gelu(sum + b[m]) to C
//   DN_GELU_BACK:       sum * gelu'(zp[n ldc + m]) to C
constexpr int DN_BM = 64, DN_BN = 64, DN_BK = 16;
template <typename E, int MODE, bool AT, bool BT>
__global__ void __launch_bounds__(256) k_dense(const E* __restrict__ A, long lda, const E* __restrict__ B, long ldb,
                                              E* C, long ldc, int M, long N, long K, long kc,
                                              const E* __restrict__ bias, E* C2, const E* __restrict__ zp) {
  __shared__ __attribute__((aligned(16))) E As[DN_BK][DN_BM + 4];
  __shared__ __attribute__((aligned(16))) E Bs[DN_BK][DN_BN + 4];
  const int t = threadIdx.x;
  const int tm = t % 16, tn = t / 16;  // consecutive threads on consecutive m: coalesced stores
  const int m0 = blockIdx.y * DN_BM;
  const long n0 = long(blockIdx.x) * DN_BN;
  const long k0 = long(blockIdx.z) * kc, k1 = K < k0 + kc ? K : k0 + kc;
  E acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = E(0);
  for (long kb = k0; kb < k1; kb += DN_BK) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < DN_BM * DN_BK / 256; ++r) {
      const int e = t + 256 * r;
      // the index along the operand's contiguous direction varies fastest
      const int kk = AT ? e / DN_BM : e % DN_BK, mm = AT ? e % DN_BM : e / DN_BK;
      const long k = kb + kk;
      const int m = m0 + mm;
      As[kk][mm] = (k < k1 && m < M) ? (AT ? A[k * lda + m] : A[long(m) * lda + k]) : E(0);
    }
#pragma unroll
    for (int r = 0; r < DN_BN * DN_BK / 256; ++r) {
      const int e = t + 256 * r;
      const int kk = BT ? e / DN_BN : e % DN_BK, nn = BT ? e % DN_BN : e / DN_BK;
      const long k = kb + kk, n = n0 + nn;
      Bs[kk][nn] = (k < k1 && n < N) ? (BT ? B[k * ldb + n] : B[n * ldb + k]) : E(0);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < DN_BK; ++kk) {
      E a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][4 * tm + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][4 * tn + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
    }
  }
  E* Cz = C + long(blockIdx.z) * N * ldc;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const long n = n0 + 4 * tn + j;
    if (n >= N) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = m0 + 4 * tm + i;
      if (m >= M) continue;
      const long o = n * ldc + m;
      if constexpr (MODE == DN_PLAIN) {
        Cz[o] = acc[i][j];
      } else if constexpr (MODE == DN_BIAS_GELU_TRAIN) {
        const E zz = acc[i][j] + bias[m];
        C[o] = zz;
        C2[o] = gelu(zz);
      } else if constexpr (MODE == DN_BIAS_GELU) {
        C[o] = gelu(acc[i][j] + bias[m]);
      } else {
        C[o] = acc[i][j] * gelu_grad(zp[o]);
      }
    }
  }
}

// FP64 dense layer of the inference on v_mfma_f64_16x16x4f64 (lane l holds
// A[l%16][l/16], B[l/16][l%16], D[l/16 + 4r][l%16]): rows of the product are
// voxels, columns output channels, so each D row is 16 consecutive channels
// of one voxel (128-byte stores).  out[v ldo + o] = act(sum_k x[v ldx + k]
// W[o ldw + k] + b[o]) for v < nv, o < co, k < ci; act = gelu on hidden
// layers (GELU) or the identity.  Block: 128 voxels x 32 channels, 4
// wavefronts of 32 voxels x 32 channels (2 x 2 MFMA tiles); 32-channel k
// slices of the voxels and the weights staged in LDS, the next slice loaded
// into registers while the current one is multiplied.  Grid: (channel
// blocks, voxel blocks), so the blocks sharing a voxel tile run side by side
// and read it from L2 once.
constexpr int DM_BV = 128, DM_BO = 32, DM_BK = 32, DM_PAD = DM_BK + 1;
template <bool GELU>
__global__ void __launch_bounds__(256) k_dense_mfma64(const double* __restrict__ x, long ldx,
                                                     const double* __restrict__ W, long ldw,
                                                     const double* __restrict__ bias, double* out, long ldo,
                                                     long nv, int co, int ci) {
  __shared__ double Xs[DM_BV][DM_PAD];
  __shared__ double Ws[DM_BO][DM_PAD];
  constexpr int NX = DM_BV * DM_BK / 256, NW = DM_BO * DM_BK / 256;
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const int o0 = blockIdx.x * DM_BO;
  const long v0 = long(blockIdx.y) * DM_BV;
  const int wv = 32 * w;  // the wavefront's first voxel within the block
  // staging of the k slice from k0: entry e = t + 256 r, k fastest
  double px[NX], pw[NW];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int r = 0; r < NX; ++r) {
      const int e = t + 256 * r, kk = e % DM_BK;
      const long v = v0 + e / DM_BK;
      px[r] = (v < nv && k0 + kk < ci) ? x[v * ldx + k0 + kk] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < NW; ++r) {
      const int e = t + 256 * r, kk = e % DM_BK, o = o0 + e / DM_BK;
      pw[r] = (o < co && k0 + kk < ci) ? W[long(o) * ldw + k0 + kk] : 0.0;
    }
  };
  d4_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = d4_t{0.0, 0.0, 0.0, 0.0};
  fetch(0);
  for (int k0 = 0; k0 < ci; k0 += DM_BK) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NX; ++r) Xs[(t + 256 * r) / DM_BK][(t + 256 * r) % DM_BK] = px[r];
#pragma unroll
    for (int r = 0; r < NW; ++r) Ws[(t + 256 * r) / DM_BK][(t + 256 * r) % DM_BK] = pw[r];
    __syncthreads();
    if (k0 + DM_BK < ci) fetch(k0 + DM_BK);
#pragma unroll
    for (int ks = 0; ks < DM_BK; ks += 4) {
      double av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) av[a] = Xs[wv + 16 * a + (l & 15)][ks + (l >> 4)];
#pragma unroll
      for (int b = 0; b < 2; ++b) bv[b] = Ws[16 * b + (l & 15)][ks + (l >> 4)];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
  }
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int o = o0 + 16 * b + (l & 15);
    if (o >= co) continue;
    const double bo = GELU ? bias[o] : 0.0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long v = v0 + wv + 16 * a + (l >> 4) + 4 * r;
        if (v >= nv) continue;
        const double s = acc[a][b][r];
        out[v * ldo + o] = GELU ? gelu(s + bo) : s;
      }
  }
}

// out[j] += sum over z of part[z len + j], z in order (deterministic)
__global__ void k_dense_reduce(const float* part, long nz, long len, float* out) {
  const long j = gid();
  if (j >= len) return;
  float s = 0.f;
  for (long z = 0; z < nz; ++z) s += part[z * len + j];
  out[j] += s;
}

// part[z co + o] = sum of d[v co + o] over the voxels v of chunk z
__global__ void k_colsum_partials(const float* d, int co, long nv, long chunk, float* part) {
  const long v0 = long(blockIdx.x) * chunk, v1 = nv < v0 + chunk ? nv : v0 + chunk;
  for (int o = threadIdx.x; o < co; o += blockDim.x) {
    float s = 0.f;
    for (long v = v0; v < v1; ++v) s += d[v * co + o];
    part[long(blockIdx.x) * co + o] = s;
  }
}

// trace removal of packed (xx, yy[, zz], ...) rows, in place: the `symm`
// layer of `mlp` (src/nets.jl:275-293); the map is its own transpose
template <typename E>
__global__ void k_tracefree(E* y, long nv, int nt, int D) {
  const long v = gid();
  if (v >= nv) return;
  E* r = y + v * nt;
  if (D == 3) {
    const E t = (r[0] + r[1] + r[2]) / E(3);
    r[0] -= t;
    r[1] -= t;
    r[2] -= t;
  } else {
    const E t = (r[0] + r[1]) / E(2);
    r[0] -= t;
    r[1] -= t;
  }
}

// sum of squares of x[0, n) in block partials (FP64)
__global__ void k_sumsq(const float* x, long n, double* part) {
  __shared__ double sm[256];
  double acc = 0;
  for (long i = gid(); i < n; i += long(gridDim.x) * blockDim.x) acc += double(x[i]) * double(x[i]);
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// squared error partials and dy = scale * (y - t), both [nv][nt]
__global__ void k_loss(const float* y, const float* t, float* dy, long n, double scale, double* part) {
  __shared__ double sm[256];
  double acc = 0;
  for (long i = gid(); i < n; i += long(gridDim.x) * blockDim.x) {
    const float e = y[i] - t[i];
    acc += double(e) * double(e);
    dy[i] = float(scale * double(e));
  }
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// fullchain's output scaling for voxels [v0, v0 + nv): tau_c = y_c (Δ² A2)
// in the solver's precision, A2 recomputed from GG
template <typename T>
__global__ void k_tau_out(const double* y, const T* GG, long np, long v0, long nv, int ng, int nt, double d2, T* tau) {
  const long i = gid();
  if (i >= nv) return;
  const long v = v0 + i;
  double a2 = 0;
  for (int c = 0; c < ng; ++c) {
    const double g = double(GG[c * np + v]);
    a2 += g * g;
  }
  const double s = d2 * a2;
  for (int c = 0; c < nt; ++c) tau[c * np + v] = T(y[i * nt + c] * s);
}

// ------------------------------------------------ the cube's symmetry group
// group_stuff (src/symmetry.jl:42-67): element e = (permutation e % NPERM,
// signs e / NPERM), in the reference's order; (R A Rᵀ)_ab = s_a s_b A_{p_a p_b}
__constant__ int c_perm3[6][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}, {1, 0, 2}, {0, 2, 1}};
__constant__ int c_sign3[8][3] = {{1, 1, 1},  {-1, 1, 1}, {1, -1, 1}, {1, 1, -1},
                                  {-1, -1, 1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, -1}};
__constant__ int c_perm2[2][2] = {{0, 1}, {1, 0}};
__constant__ int c_sign2[4][2] = {{1, 1}, {-1, 1}, {1, -1}, {-1, -1}};

__device__ __forceinline__ void group_elem(int D, int e, int* p, int* sg) {
  if (D == 3) {
    const int pi = e % 6, si = e / 6;
    for (int a = 0; a < 3; ++a) {
      p[a] = c_perm3[pi][a];
      sg[a] = c_sign3[si][a];
    }
  } else {
    const int pi = e % 2, si = e / 2;
    for (int a = 0; a < 2; ++a) {
      p[a] = c_perm2[pi][a];
      sg[a] = c_sign2[si][a];
    }
  }
}
// packed index of the symmetric component (a, b): xx, yy, zz, xy, yz, zx (3D)
// or xx, yy, xy (2D)
__device__ __forceinline__ int packed_index(int D, int a, int b) {
  if (a == b) return a;
  if (D == 2) return 2;
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  return lo == 0 ? (hi == 1 ? 3 : 5) : 4;
}

// symmetrize_pointwise's inputs (src/nets.jl:350-389) for voxels [v0, v0 + nv)
// and every group element: row e nv + i holds fullchain's input for R_e A R_eᵀ
// of voxel v0 + i (its norm taken over the rotated components, in order)
template <typename T>
__global__ void k_features_sym(const T* GG, long np, long v0, long nv, int D, int G, const double* z, double* x,
                               int nin) {
  const long id = gid();
  if (id >= nv * G) return;
  const int e = int(id / nv);
  const long v = v0 + (id - long(e) * nv);
  int p[3], sg[3];
  group_elem(D, e, p, sg);
  double r[9];
  for (int b = 0; b < D; ++b)
    for (int a = 0; a < D; ++a) r[a + D * b] = double(T(sg[a] * sg[b]) * GG[long(p[a] + D * p[b]) * np + v]);
  const int ng = D * D;
  double a2 = 0;
  for (int k = 0; k < ng; ++k) a2 += r[k] * r[k];
  const double den = sqrt(a2) + double(std::numeric_limits<T>::epsilon());
  for (int k = 0; k < ng; ++k) x[id * nin + k] = r[k] / den;
  if (z) x[id * nin + ng] = *z;
}

// the group average of the conjugated-back outputs, scaled to tau:
// tau_ab = Δ² A2 / |G| Σ_e qs_a qs_b y_e[q_a q_b], (q, qs) the inverse element
template <typename T>
__global__ void k_sym_out(const double* y, const T* GG, long np, long v0, long nv, int D, int nt, int G, double d2,
                          T* tau) {
  const long i = gid();
  if (i >= nv) return;
  const long v = v0 + i;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const int ca[6] = {0, 1, 2, 0, 1, 2}, cb[6] = {0, 1, 2, 1, 2, 0};
  const int da[3] = {0, 1, 0}, db[3] = {0, 1, 1};
  for (int e = 0; e < G; ++e) {
    int p[3], sg[3], q[3], qs[3];
    group_elem(D, e, p, sg);
    for (int a = 0; a < D; ++a) q[p[a]] = a;
    for (int a = 0; a < D; ++a) qs[a] = sg[q[a]];
    const double* row = y + (long(e) * nv + i) * nt;
    for (int t = 0; t < nt; ++t) {
      const int a = D == 3 ? ca[t] : da[t], b = D == 3 ? cb[t] : db[t];
      acc[t] += double(qs[a] * qs[b]) * row[packed_index(D, q[a], q[b])];
    }
  }
  double a2 = 0;
  for (int c = 0; c < D * D; ++c) {
    const double g = double(GG[c * np + v]);
    a2 += g * g;
  }
  const double sc = d2 * a2;
  for (int t = 0; t < nt; ++t) tau[t * np + v] = T(acc[t] / double(G) * sc);
}

// fullchain's closure (the mlp, src/nets.jl:316-331) and its group average
// (symmetrize_pointwise, :350-389) in one kernel, the work of k_features_inf
// / k_features_sym, k_mlp_fused, k_tracefree and k_tau_out / k_sym_out: a
// wavefront takes RW = 16 FR voxels through each of the G group elements
// (G = 1: the identity alone).  Per element it forms the network input of
// R_e A R_eᵀ from the voxels' gradients staged in LDS (gb, component-major),
// runs the four layers, stages the outputs (yb), removes their trace and
// adds them, conjugated back, to the voxel's sums; then
// tau_ab = Δ² A2 / G Σ_e qs_a qs_b y_e[q_a q_b] in the field's precision.
// The norm of the rotated gradient is taken in the unrotated order (the same
// squares; its rounding may differ from a rotated-order sum in the last bit).
// Each wavefront owns RW (ng + nt) doubles of LDS at ybase.
template <int TOM, int LDSD, typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_mlp_closure(const double* __restrict__ pd, FusedNet f, int ybase,
                                                     const T* __restrict__ GG, long np, int D, int G,
                                                     const double* __restrict__ zp, double d2, T* __restrict__ tau) {
  __shared__ double wl[LDSD];
  constexpr int RW = 16 * FR;
  constexpr int NSET = 64 / RW;            // output sets per row
  constexpr int NACC = (6 + NSET - 1) / NSET;
  stage_net<TOM>(wl, pd, f);
  __syncthreads();
  constexpr int COLS = 16 * TOM;
  const int lane = threadIdx.x & 63, c = lane >> 4, v = lane & 15;
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int ng = D * D, nt = f.cout[3], nin = f.cin[0];
  double* gb = wl + ybase + wave * RW * (ng + nt);
  double* yb = gb + RW * ng;
  const double z = zp ? *zp : 0.0;
  const double eps = double(std::numeric_limits<T>::epsilon());
  const double* wlv = wl + c * COLS + v;
  const int ro = lane % RW, set = lane / RW;  // this lane's output row and set
  const long npass = (np + RW - 1) / RW;
  for (long p = long(blockIdx.x) * nw + wave; p < npass; p += long(gridDim.x) * nw) {
    const long vb = p * RW;
    for (int i = lane; i < RW * ng; i += 64) {
      const int k = i / RW, row = i - k * RW;
      double gv = 0.0;
      if (vb + row < np) gv = double(GG[long(k) * np + vb + row]);
      gb[i] = gv;
    }
    wave_sync();
    double den[FR];
#pragma unroll
    for (int rt = 0; rt < FR; ++rt) {
      double a2 = 0.0;
      for (int k = 0; k < ng; ++k) {
        const double gv = gb[k * RW + rt * 16 + v];
        a2 += gv * gv;
      }
      den[rt] = sqrt(a2) + eps;
    }
    double acc[NACC];
#pragma unroll
    for (int j = 0; j < NACC; ++j) acc[j] = 0.0;
    for (int e = 0; e < G; ++e) {
      int pe[3] = {0, 1, 2}, se[3] = {1, 1, 1};
      group_elem(D, e, pe, se);
      d4_t A[FR][TOM], B[FR][TOM];
#pragma unroll
      for (int rt = 0; rt < FR; ++rt)
#pragma unroll
        for (int t = 0; t < TOM; ++t)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            // input channel ci of row rt*16 + v: (R_e A R_eᵀ)_ab = s_a s_b A_{p_a p_b},
            // component a + D b; then the Re_Δ input
            const int ci = 16 * t + c + 4 * r;
            double xv = 0.0;
            if (ci < ng) {
              const int a = ci % D, b = ci / D;
              const int pa = a == 0 ? pe[0] : (a == 1 ? pe[1] : pe[2]);
              const int pb = b == 0 ? pe[0] : (b == 1 ? pe[1] : pe[2]);
              const int sa = a == 0 ? se[0] : (a == 1 ? se[1] : se[2]);
              const int sb = b == 0 ? se[0] : (b == 1 ? se[1] : se[2]);
              xv = double(sa * sb) * gb[(pa + D * pb) * RW + rt * 16 + v] / den[rt];
            } else if (ci < nin) {
              xv = z;
            }
            A[rt][t][r] = xv;
          }
      fused_layer<TOM>(wlv + f.wofs[0], f.ti[0], f.cin[0], f.to[0], A, B);
      fused_bias_gelu<TOM>(B, wl + f.bofs[0] + c, f.cout[0]);
      fused_layer<TOM>(wlv + f.wofs[1], f.ti[1], f.cin[1], f.to[1], B, A);
      fused_bias_gelu<TOM>(A, wl + f.bofs[1] + c, f.cout[1]);
      fused_layer<TOM>(wlv + f.wofs[2], f.ti[2], f.cin[2], f.to[2], A, B);
      fused_bias_gelu<TOM>(B, wl + f.bofs[2] + c, f.cout[2]);
      fused_layer<TOM>(wlv + f.wofs[3], f.ti[3], f.cin[3], f.to[3], B, A);
      // the output tile: channel c + 4r of row rt*16 + v (nt <= 6)
#pragma unroll
      for (int rt = 0; rt < FR; ++rt)
#pragma unroll
        for (int r = 0; r < 2; ++r)
          if (c + 4 * r < nt) yb[(c + 4 * r) * RW + rt * 16 + v] = A[rt][0][r];
      wave_sync();
      // trace removal (k_tracefree), then the inverse element's conjugation
      // q[pe[a]] = a, qs[a] = se[q[a]], by selects (no register-indexed stores)
      int q[3], qs[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        q[k] = pe[0] == k ? 0 : (pe[1] == k ? 1 : 2);
        qs[k] = q[k] == 0 ? se[0] : (q[k] == 1 ? se[1] : se[2]);
      }
      const double tr = D == 3 ? (yb[ro] + yb[RW + ro] + yb[2 * RW + ro]) / 3.0 : (yb[ro] + yb[RW + ro]) / 2.0;
#pragma unroll
      for (int j = 0; j < NACC; ++j) {
        const int t = set + NSET * j;
        if (t >= nt) break;
        // component (a, b) of output t: xx, yy, zz, xy, yz, zx (3D); xx, yy, xy (2D)
        const int a = D == 3 ? (t < 3 ? t : t - 3) : (t < 2 ? t : 0);
        const int b = D == 3 ? (t < 3 ? t : (t == 5 ? 0 : t - 2)) : (t < 2 ? t : 1);
        const int qa = a == 0 ? q[0] : (a == 1 ? q[1] : q[2]);
        const int qb = b == 0 ? q[0] : (b == 1 ? q[1] : q[2]);
        const int sa = a == 0 ? qs[0] : (a == 1 ? qs[1] : qs[2]);
        const int sb = b == 0 ? qs[0] : (b == 1 ? qs[1] : qs[2]);
        const int k = packed_index(D, qa, qb);
        double yv = yb[k * RW + ro];
        if (k < D) yv -= tr;
        acc[j] += double(sa * sb) * yv;
      }
      wave_sync();
    }
    const long vox = vb + ro;
    if (vox < np) {
      double a2 = 0.0;
      for (int k = 0; k < ng; ++k) {
        const double gv = gb[k * RW + ro];
        a2 += gv * gv;
      }
      const double sc = d2 * a2;
#pragma unroll
      for (int j = 0; j < NACC; ++j) {
        const int t = set + NSET * j;
        if (t >= nt) break;
        tau[long(t) * np + vox] = T(acc[j] / double(G) * sc);
      }
    }
    wave_sync();
  }
}

// ------------------------------------------------------------------ TBNN
// tb_kernel! (src/nets.jl:455-500) on one voxel's gradient G (nonsym order,
// G_ij at i + D j): A = G / (|G| + eps), S, R its symmetric and antisymmetric
// parts; inv[ni] the invariants, bas[t + nt b] the deviatoric basis tensors
// (packed); returns |G|² (FP64)
struct Mat3 {
  double m[3][3];
};
__device__ __forceinline__ Mat3 mmul(const Mat3& A, const Mat3& B) {
  Mat3 C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C.m[i][j] = A.m[i][0] * B.m[0][j] + A.m[i][1] * B.m[1][j] + A.m[i][2] * B.m[2][j];
  return C;
}
__device__ __forceinline__ Mat3 msub(const Mat3& A, const Mat3& B) {
  Mat3 C;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C.m[i][j] = A.m[i][j] - B.m[i][j];
  return C;
}
__device__ __forceinline__ double mtr(const Mat3& A) { return A.m[0][0] + A.m[1][1] + A.m[2][2]; }
// deviator, packed xx yy zz xy yz zx at out[t * nt]
__device__ __forceinline__ void put_dev3(const Mat3& B, double* out, int stride) {
  const double t = mtr(B) / 3;
  out[0 * stride] = B.m[0][0] - t;
  out[1 * stride] = B.m[1][1] - t;
  out[2 * stride] = B.m[2][2] - t;
  out[3 * stride] = B.m[0][1];
  out[4 * stride] = B.m[1][2];
  out[5 * stride] = B.m[2][0];
}

template <typename T>
__device__ __forceinline__ double tb_voxel(int D, const T* GG, long np, long v, double* inv, double* bas) {
  const int ng = D * D;
  double g[9];
  double a2 = 0;
  for (int c = 0; c < ng; ++c) {
    g[c] = double(GG[c * np + v]);
    a2 += g[c] * g[c];
  }
  const double den = sqrt(a2) + double(std::numeric_limits<T>::epsilon());
  if (D == 3) {
    Mat3 S, R;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const double aij = g[i + 3 * j] / den, aji = g[j + 3 * i] / den;
        S.m[i][j] = (aij + aji) / 2;
        R.m[i][j] = (aij - aji) / 2;
      }
    const Mat3 SS = mmul(S, S), RR = mmul(R, R), SR = mmul(S, R), RS = mmul(R, S);
    inv[0] = mtr(SS);
    inv[1] = mtr(RR);
    inv[2] = mtr(mmul(SS, S));
    inv[3] = mtr(mmul(SR, R));
    inv[4] = mtr(mmul(mmul(SS, R), R));
    put_dev3(S, bas + 0 * 6, 1);
    put_dev3(SS, bas + 1 * 6, 1);
    put_dev3(RR, bas + 2 * 6, 1);
    put_dev3(msub(SR, RS), bas + 3 * 6, 1);
    const Mat3 RSR = mmul(RS, R);
    put_dev3(RSR, bas + 4 * 6, 1);
    put_dev3(msub(mmul(RS, S), mmul(SS, R)), bas + 5 * 6, 1);
    put_dev3(msub(mmul(RSR, R), mmul(mmul(RR, S), R)), bas + 6 * 6, 1);
  } else {
    // 2D: A = [xx xy; yx yy] with G nonsym order xx, yx, xy, yy
    const double axx = g[0] / den, ayx = g[1] / den, axy = g[2] / den, ayy = g[3] / den;
    const double s00 = axx, s11 = ayy, s01 = (axy + ayx) / 2, r01 = (axy - ayx) / 2;
    // S = [s00 s01; s01 s11], R = [0 r01; -r01 0]
    inv[0] = s00 * s00 + s01 * s01 + s01 * s01 + s11 * s11;  // tr(S S)
    inv[1] = -r01 * r01 - r01 * r01;                          // tr(R R)
    const double t = (s00 + s11) / 2;
    bas[0] = s00 - t;
    bas[1] = s11 - t;
    bas[2] = s01;
    // S R - R S
    const double c00 = s01 * (-r01) - r01 * s01;
    const double c01 = s00 * r01 - r01 * s11;
    const double c11 = s01 * r01 - (-r01) * s01;
    const double t2 = (c00 + c11) / 2;
    bas[3] = c00 - t2;
    bas[4] = c11 - t2;
    bas[5] = c01;
  }
  return a2;
}

// create_dataloader_tbnn's pair (src/data.jl:759-805): x[v] = [invariants,
// z, basis (t + nt b)], y = tau / (Δ² A2 + eps), FP32
template <typename T>
__global__ void k_tb_features(const T* GG, long np, int D, double d2, int has_re, float z, const T* tau, int nt,
                              float* x, int ctot, int ni, int nbas, float* y) {
  const long v = gid();
  if (v >= np) return;
  double inv[5], bas[42];
  const double a2 = tb_voxel<T>(D, GG, np, v, inv, bas);
  float* r = x + v * ctot;
  for (int k = 0; k < ni; ++k) r[k] = float(inv[k]);
  const int o = ni + (has_re ? 1 : 0);
  if (has_re) r[ni] = z;
  for (int k = 0; k < nt * nbas; ++k) r[o + k] = float(bas[k]);
  if (tau) {
    const double den = d2 * a2 + double(std::numeric_limits<T>::epsilon());
    for (int c = 0; c < nt; ++c) y[v * nt + c] = float(double(tau[c * np + v]) / den);
  }
}

// the TBNN model's input (tbnn, src/nets.jl:558-583) for voxels [v0, v0 + nv), FP64
template <typename T>
__global__ void k_tb_features_inf(const T* GG, long np, long v0, long nv, int D, const double* z, double* x, int ctot,
                                  int ni, int nt, int nbas) {
  const long i = gid();
  if (i >= nv) return;
  double inv[5], bas[42];
  tb_voxel<T>(D, GG, np, v0 + i, inv, bas);
  double* r = x + i * ctot;
  for (int k = 0; k < ni; ++k) r[k] = inv[k];
  const int o = ni + (z ? 1 : 0);
  if (z) r[ni] = *z;
  for (int k = 0; k < nt * nbas; ++k) r[o + k] = bas[k];
}

// basis contraction y[v][t] = Σ_b w[v][b] x[v][boff + t + nt b]
template <typename E>
__global__ void k_tb_contract(const E* w, const E* x, long nv, int ctot, int boff, int nt, int nbas, E* y) {
  const long v = gid();
  if (v >= nv) return;
  const E* bx = x + v * ctot + boff;
  for (int t = 0; t < nt; ++t) {
    E m = E(0);
    for (int b = 0; b < nbas; ++b) m += bx[t + nt * b] * w[v * nbas + b];
    y[v * nt + t] = m;
  }
}

// its transpose: dw[v][b] = Σ_t dy[v][t] x[v][boff + t + nt b]
__global__ void k_tb_dw(const float* dy, const float* x, long nv, int ctot, int boff, int nt, int nbas, float* dw) {
  const long v = gid();
  if (v >= nv) return;
  const float* bx = x + v * ctot + boff;
  for (int b = 0; b < nbas; ++b) {
    float a = 0.f;
    for (int t = 0; t < nt; ++t) a += dy[v * nt + t] * bx[t + nt * b];
    dw[v * nbas + b] = a;
  }
}

// tbnn's closure (src/nets.jl:558-583) in one kernel, the work of
// k_tb_features_inf, k_mlp_fused, k_tb_contract and k_tau_out: per pass a
// wavefront's lanes 0..RW-1 stage the invariants [and z] of RW = 16 FR voxels
// in LDS (xb, channel-major), the four layers give the basis coefficients
// (staged in yb), and each lane recomputes its voxel's deviatoric basis and
// contracts it: tau_t = Δ² A2 Σ_b w_b B_b[t].  DD is the grid's dimension
// (2: 2 invariants, 2 basis tensors, 3 components; 3: 5, 7, 6).  Each
// wavefront owns RW (cin[0] + nbas) doubles of LDS at ybase.
template <int TOM, int LDSD, int DD, typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_tb_closure(const double* __restrict__ pd, FusedNet f, int ybase, const T* __restrict__ GG, long np,
             const double* __restrict__ zp, double d2, T* __restrict__ tau) {
  __shared__ double wl[LDSD];
  constexpr int RW = 16 * FR;
  constexpr int NSET = 64 / RW;
  constexpr int NI = DD == 3 ? 5 : 2, NT = DD == 3 ? 6 : 3, NB = DD == 3 ? 7 : 2;
  constexpr int NACC = (NT + NSET - 1) / NSET;
  stage_net<TOM>(wl, pd, f);
  __syncthreads();
  constexpr int COLS = 16 * TOM;
  const int lane = threadIdx.x & 63, c = lane >> 4, v = lane & 15;
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int nin = f.cin[0];
  double* xb = wl + ybase + wave * RW * (nin + NB);
  double* yb = xb + RW * nin;
  const double* wlv = wl + c * COLS + v;
  const int ro = lane % RW, set = lane / RW;
  const long npass = (np + RW - 1) / RW;
  for (long p = long(blockIdx.x) * nw + wave; p < npass; p += long(gridDim.x) * nw) {
    const long vb = p * RW;
    if (lane < RW) {
      double inv[5] = {0, 0, 0, 0, 0}, bas[42];
      if (vb + lane < np) tb_voxel<T>(DD, GG, np, vb + lane, inv, bas);
#pragma unroll
      for (int k = 0; k < NI; ++k) xb[k * RW + lane] = inv[k];
      if (nin > NI) xb[NI * RW + lane] = *zp;
    }
    wave_sync();
    d4_t A[FR][TOM], B[FR][TOM];
#pragma unroll
    for (int rt = 0; rt < FR; ++rt)
#pragma unroll
      for (int t = 0; t < TOM; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int ci = 16 * t + c + 4 * r;
          double xv = 0.0;
          if (ci < nin) xv = xb[ci * RW + rt * 16 + v];
          A[rt][t][r] = xv;
        }
    fused_layer<TOM>(wlv + f.wofs[0], f.ti[0], f.cin[0], f.to[0], A, B);
    fused_bias_gelu<TOM>(B, wl + f.bofs[0] + c, f.cout[0]);
    fused_layer<TOM>(wlv + f.wofs[1], f.ti[1], f.cin[1], f.to[1], B, A);
    fused_bias_gelu<TOM>(A, wl + f.bofs[1] + c, f.cout[1]);
    fused_layer<TOM>(wlv + f.wofs[2], f.ti[2], f.cin[2], f.to[2], A, B);
    fused_bias_gelu<TOM>(B, wl + f.bofs[2] + c, f.cout[2]);
    fused_layer<TOM>(wlv + f.wofs[3], f.ti[3], f.cin[3], f.to[3], B, A);
    // the coefficients: channel c + 4r of row rt*16 + v (NB <= 7)
#pragma unroll
    for (int rt = 0; rt < FR; ++rt)
#pragma unroll
      for (int r = 0; r < 2; ++r)
        if (c + 4 * r < NB) yb[(c + 4 * r) * RW + rt * 16 + v] = A[rt][0][r];
    wave_sync();
    const long vox = vb + ro;
    if (vox < np) {
      double inv[5], bas[42];
      const double sc = d2 * tb_voxel<T>(DD, GG, np, vox, inv, bas);
      double w[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) w[b] = yb[b * RW + ro];
#pragma unroll
      for (int j = 0; j < NACC; ++j) {
        // component t = set + NSET j, selected among the compile-time
        // candidates so that bas stays in registers
        double m[NSET];
#pragma unroll
        for (int s = 0; s < NSET; ++s) {
          const int t = s + NSET * j;
          double a = 0.0;
          if (t < NT) {
#pragma unroll
            for (int b = 0; b < NB; ++b) a += bas[t + NT * b] * w[b];
          }
          m[s] = a;
        }
        double mm = m[0];
#pragma unroll
        for (int s = 1; s < NSET; ++s)
          if (set == s) mm = m[s];
        const int t = set + NSET * j;
        if (t < NT) tau[long(t) * np + vox] = T(mm * sc);
      }
    }
    wave_sync();
  }
}

// ------------------------------------------------------ equivariant_net
// The synthesis of equivariant_net's projected weights (src/nets.jl:102-173,
// with get_weight_synthesis, src/symmetry.jl:156-197) from the compact
// trainables, and its transpose for the gradient.  Tables gt: perm, sign,
// pinv (inverse permutation) [nreg][3], relidx[m][n] = index of g_n^-1 g_m,
// ninv[m][k] = the n with relidx[m][n] = k.  Network weights are Lux's
// (c_in, c_out) column-major; compact arrays: lift (nten, c1), mid (nreg,
// c_out, c_in), sink (nten, c3), all column-major.
struct EqTab {
  const int *perm, *sign, *pinv, *relidx, *ninv;
};
__device__ __forceinline__ EqTab eq_tab(const int* gt, int nreg) {
  return {gt, gt + 3 * nreg, gt + 6 * nreg, gt + 9 * nreg, gt + 9 * nreg + nreg * nreg};
}

// lift: W[mu][h + nreg j] = alpha s_x s_y w[p_x + D p_y][j] (mu = x + D y),
// W[nten][h + nreg j] = w_re[j]; b[h + nreg j] = bias[j]
template <typename E>
__global__ void k_eq_lift(const float* w, const float* wre, const float* bias, E* W, E* b, int D, int nreg, int nin,
                          int c1, E alpha, const int* gt) {
  const long id = gid();
  const int nten = D * D;
  const long ncol = long(nreg) * c1;
  if (id < ncol) b[id] = E(bias[id / nreg]);
  if (id >= ncol * nin) return;
  const EqTab t = eq_tab(gt, nreg);
  const int mu = int(id % nin);
  const long k = id / nin;
  const int h = int(k % nreg), j = int(k / nreg);
  if (mu == nten) {
    W[id] = E(wre[j]);
    return;
  }
  const int x = mu % D, y = mu / D;
  const int nu = t.perm[3 * h + x] + D * t.perm[3 * h + y];
  W[id] = (alpha * E(t.sign[3 * h + x] * t.sign[3 * h + y])) * E(w[nu + nten * j]);
}

// mid: W[n + nreg ci][m + nreg co] = alpha w[relidx[m][n]][co][ci]; b[m + nreg co] = bias[co]
template <typename E>
__global__ void k_eq_mid(const float* w, const float* bias, E* W, E* b, int nreg, int cin, int cout, E alpha,
                         const int* gt) {
  const long id = gid();
  const long nrow = long(nreg) * cin, ncol = long(nreg) * cout;
  if (id < ncol) b[id] = E(bias[id / nreg]);
  if (id >= nrow * ncol) return;
  const EqTab t = eq_tab(gt, nreg);
  const long r = id % nrow, k = id / nrow;
  const int n = int(r % nreg), ci = int(r / nreg);
  const int m = int(k % nreg), co = int(k / nreg);
  W[id] = alpha * E(w[t.relidx[m * nreg + n] + long(nreg) * (co + long(cout) * ci)]);
}

// sink: W[h + nreg ci][mu] = alpha s_x s_y w[p_x + D p_y][ci]
template <typename E>
__global__ void k_eq_sink(const float* w, E* W, int D, int nreg, int c3, E alpha, const int* gt) {
  const long id = gid();
  const int nten = D * D;
  const long nrow = long(nreg) * c3;
  if (id >= nrow * nten) return;
  const EqTab t = eq_tab(gt, nreg);
  const long r = id % nrow;
  const int mu = int(id / nrow);
  const int h = int(r % nreg), ci = int(r / nreg);
  const int x = mu % D, y = mu / D;
  const int nu = t.perm[3 * h + x] + D * t.perm[3 * h + y];
  W[id] = (alpha * E(t.sign[3 * h + x] * t.sign[3 * h + y])) * E(w[nu + nten * ci]);
}

// transposes (gradients of the compact arrays, accumulated)
__global__ void k_eq_lift_T(const float* GW, const float* Gb, float* gw, float* gre, float* gb, int D, int nreg,
                            int nin, int c1, float alpha, const int* gt) {
  const long id = gid();
  const int nten = D * D;
  const EqTab t = eq_tab(gt, nreg);
  if (id < long(nten) * c1) {
    const int nu = int(id % nten), j = int(id / nten);
    const int i = nu % D, jj = nu / D;
    float acc = 0.f;
    for (int h = 0; h < nreg; ++h) {
      const int x = t.pinv[3 * h + i], y = t.pinv[3 * h + jj];
      acc += float(t.sign[3 * h + x] * t.sign[3 * h + y]) * GW[(x + D * y) + long(nin) * (h + long(nreg) * j)];
    }
    gw[id] += alpha * acc;
  } else if (id < long(nten) * c1 + c1) {
    const int j = int(id - long(nten) * c1);
    float ab = 0.f, ar = 0.f;
    for (int h = 0; h < nreg; ++h) {
      ab += Gb[h + long(nreg) * j];
      if (gre) ar += GW[nten + long(nin) * (h + long(nreg) * j)];
    }
    gb[j] += ab;
    if (gre) gre[j] += ar;
  }
}

__global__ void k_eq_mid_T(const float* GW, const float* Gb, float* gw, float* gb, int nreg, int cin, int cout,
                           float alpha, const int* gt) {
  const long id = gid();
  const EqTab t = eq_tab(gt, nreg);
  const long nw = long(nreg) * cout * cin;
  const long nrow = long(nreg) * cin;
  if (id < nw) {
    const int kk = int(id % nreg);
    const long rest = id / nreg;
    const int co = int(rest % cout), ci = int(rest / cout);
    float acc = 0.f;
    for (int m = 0; m < nreg; ++m) {
      const int n = t.ninv[m * nreg + kk];
      acc += GW[(n + long(nreg) * ci) + nrow * (m + long(nreg) * co)];
    }
    gw[id] += alpha * acc;
  } else if (id < nw + cout) {
    const int co = int(id - nw);
    float acc = 0.f;
    for (int m = 0; m < nreg; ++m) acc += Gb[m + long(nreg) * co];
    gb[co] += acc;
  }
}

__global__ void k_eq_sink_T(const float* GW, float* gw, int D, int nreg, int c3, float alpha, const int* gt) {
  const long id = gid();
  const int nten = D * D;
  if (id >= long(nten) * c3) return;
  const EqTab t = eq_tab(gt, nreg);
  const int nu = int(id % nten), ci = int(id / nten);
  const int i = nu % D, jj = nu / D;
  const long nrow = long(nreg) * c3;
  float acc = 0.f;
  for (int h = 0; h < nreg; ++h) {
    const int x = t.pinv[3 * h + i], y = t.pinv[3 * h + jj];
    acc += float(t.sign[3 * h + x] * t.sign[3 * h + y]) * GW[(h + long(nreg) * ci) + nrow * (x + D * y)];
  }
  gw[id] += alpha * acc;
}

// equivariant_net's `symm` layer: the D^2 sink outputs (mu = x + D y) made
// symmetric and trace-free, packed (src/nets.jl:196-215)
template <typename E>
__global__ void k_symm_sink(const E* s9, E* y, long nv, int D) {
  const long v = gid();
  if (v >= nv) return;
  const E* q = s9 + v * D * D;
  if (D == 3) {
    const E xx = q[0], yy = q[4], zz = q[8];
    const E xy = (q[1] + q[3]) / E(2), yz = (q[5] + q[7]) / E(2), zx = (q[2] + q[6]) / E(2);
    const E t = (xx + yy + zz) / E(3);
    E* o = y + v * 6;
    o[0] = xx - t; o[1] = yy - t; o[2] = zz - t; o[3] = xy; o[4] = yz; o[5] = zx;
  } else {
    const E xx = q[0], yy = q[3], xy = (q[1] + q[2]) / E(2);
    const E t = (xx + yy) / E(2);
    E* o = y + v * 3;
    o[0] = xx - t; o[1] = yy - t; o[2] = xy;
  }
}
__global__ void k_symm_sink_T(const float* dy, float* d9, long nv, int D) {
  const long v = gid();
  if (v >= nv) return;
  float* q = d9 + v * D * D;
  if (D == 3) {
    const float* g = dy + v * 6;
    const float mm = (g[0] + g[1] + g[2]) / 3.f;
    q[0] = g[0] - mm; q[4] = g[1] - mm; q[8] = g[2] - mm;
    q[1] = q[3] = g[3] / 2.f;
    q[5] = q[7] = g[4] / 2.f;
    q[2] = q[6] = g[5] / 2.f;
  } else {
    const float* g = dy + v * 3;
    const float mm = (g[0] + g[1]) / 2.f;
    q[0] = g[0] - mm; q[3] = g[1] - mm;
    q[1] = q[2] = g[2] / 2.f;
  }
}

unsigned nb(long n) { return unsigned((n + 255) / 256); }

}  // namespace
}  // namespace pmlp

struct sc_pmlp {
  int D = 3, nin = 9, nt = 6, ng = 9, w[3] = {0, 0, 0};
  // arch 0: mlp (output nt, trace removal); 1: TBNN (inputs ni invariants
  // [+ z], output the nbas basis coefficients, rows of the input matrix carry
  // the basis tensors after the network inputs: ctot = nin + nt nbas)
  int arch = 0, ni = 0, nbas = 0, ctot = 9;
  // arch 2: equivariant_net, the mlp over |G| = nreg copies of each hidden
  // channel whose weights are synthesised from compact trainable parameters
  // (p, g, m, v hold the compact vector; pf, gf the synthesised network)
  int nreg = 1, nten = 0;
  int wf[3] = {0, 0, 0};  // network widths (w times nreg for arch 2)
  bool redelta = false;
  long nvmax = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  long plen = 0;
  long cb[9] = {};  // bounds of the trainable arrays (ncarr of them)
  int ncarr = 7;
  // network parameters: W1 b1 W2 b2 W3 b3 W4 at fo[] (flen in all); the
  // trainable vector itself except for arch 2
  float *pf = nullptr, *gf = nullptr;
  long fo[8] = {};
  long flen = 0;
  int* gt = nullptr;  // arch 2 group tables (device): perm, sign, pinv [nreg][3], relidx, ninv [nreg][nreg]
  // activations of the last forward (z pre-activation, a = gelu(z)) per
  // hidden layer, two gradient scratch buffers, the parameter gradients'
  // partial sums (dpart_len floats)
  float *z[3] = {}, *a[3] = {}, *d0 = nullptr, *d1 = nullptr, *dpart = nullptr;
  long dpart_len = 0;
  // FP64 inference (the closure): widened parameters, the live Re_Δ input,
  // input / ping-pong / output buffers for chunks of up to icap voxels
  double *pd = nullptr, *zdev = nullptr, *xd = nullptr, *hA = nullptr, *hB = nullptr, *yd = nullptr;
  long icap = 0;
  int ncu = 0;  // compute units of the device (the fused kernel's grid)
  double *part = nullptr, *gnorm = nullptr;
  int adam_t = 0;
  long last_nv = 0;
  int nout() const { return arch == 1 ? nbas : (arch == 2 ? nten : nt); }
  int maxw() const { return std::max({wf[0], wf[1], wf[2], nin, nt, nout()}); }
  int cin(int l) const { return l == 0 ? nin : wf[l - 1]; }
  int cout(int l) const { return l == 3 ? nout() : wf[l]; }
  int ldin(int l) const { return l == 0 ? ctot : cin(l); }
  // the synthesis matrices' scale 1 / sqrt|G|, stored in FP32 as the
  // reference's training precision keeps them
  float alpha() const { return float(1.0 / std::sqrt(double(nreg))); }
};

namespace {
using namespace pmlp;
using namespace optim;
thread_local std::string g_pmlp_err;

template <class F>
int pguard(sc_pmlp* c, F&& f) {
  if (!c) return 1;
  (void)hipGetLastError();
  try {
    f();
    PM_HIP(hipStreamSynchronize(c->stream));
    return 0;
  } catch (std::exception& e) {
    c->err = e.what();
    return 3;
  }
}

// equivariant_net: network parameters dst (fo layout) from the trainables
template <typename E>
void synthesise(sc_pmlp* c, E* dst, E alpha) {
  const int D = c->D, nreg = c->nreg;
  const float* p = c->p;
  const long* cb = c->cb;
  const int re = c->redelta ? 1 : 0;
  const long n0 = long(c->nin) * c->wf[0];
  hipLaunchKernelGGL(k_eq_lift<E>, dim3(nb(n0)), dim3(256), 0, c->stream, p + cb[0], re ? p + cb[2] : nullptr,
                     p + cb[1], dst + c->fo[0], dst + c->fo[1], D, nreg, c->nin, c->w[0], alpha, (const int*)c->gt);
  for (int l = 1; l < 3; ++l) {
    const long nw = long(c->wf[l - 1]) * c->wf[l];
    const long wo = cb[2 + re + 2 * (l - 1)], bo = wo + long(nreg) * c->w[l] * c->w[l - 1];
    hipLaunchKernelGGL(k_eq_mid<E>, dim3(nb(nw)), dim3(256), 0, c->stream, p + wo, p + bo, dst + c->fo[2 * l],
                       dst + c->fo[2 * l + 1], nreg, c->w[l - 1], c->w[l], alpha, (const int*)c->gt);
  }
  const long n3 = long(c->wf[2]) * c->nten;
  hipLaunchKernelGGL(k_eq_sink<E>, dim3(nb(n3)), dim3(256), 0, c->stream, p + cb[c->ncarr - 1], dst + c->fo[6], D,
                     nreg, c->w[2], alpha, (const int*)c->gt);
  PM_HIP(hipGetLastError());
}

// equivariant_net: trainable gradient g += Sᵀ gf
void synthesise_T(sc_pmlp* c) {
  const int D = c->D, nreg = c->nreg;
  float* g = c->g;
  const long* cb = c->cb;
  const int re = c->redelta ? 1 : 0;
  const float al = c->alpha();
  const long n0 = long(c->nten) * c->w[0] + c->w[0];
  hipLaunchKernelGGL(k_eq_lift_T, dim3(nb(n0)), dim3(256), 0, c->stream, (const float*)c->gf + c->fo[0],
                     (const float*)c->gf + c->fo[1], g + cb[0], re ? g + cb[2] : nullptr, g + cb[1], D, nreg, c->nin,
                     c->w[0], al, (const int*)c->gt);
  for (int l = 1; l < 3; ++l) {
    const long nw = long(nreg) * c->w[l] * c->w[l - 1];
    const long wo = cb[2 + re + 2 * (l - 1)];
    hipLaunchKernelGGL(k_eq_mid_T, dim3(nb(nw + c->w[l])), dim3(256), 0, c->stream,
                       (const float*)c->gf + c->fo[2 * l], (const float*)c->gf + c->fo[2 * l + 1], g + wo,
                       g + wo + nw, nreg, c->w[l - 1], c->w[l], al, (const int*)c->gt);
  }
  const long n3 = long(c->nten) * c->w[2];
  hipLaunchKernelGGL(k_eq_sink_T, dim3(nb(n3)), dim3(256), 0, c->stream, (const float*)c->gf + c->fo[6],
                     g + cb[c->ncarr - 1], D, nreg, c->w[2], al, (const int*)c->gt);
  PM_HIP(hipGetLastError());
}

// C (ldc x N, M rows used) = A B through k_dense, k split over nz blocks
template <typename E, int MODE, bool AT, bool BT>
void dense(hipStream_t s, const E* A, long lda, const E* B, long ldb, E* C, long ldc, int M, long N, long K,
           long nz = 1, const E* bias = nullptr, E* C2 = nullptr, const E* zp = nullptr) {
  const long kc = (K + nz - 1) / nz;
  const dim3 grid(unsigned((N + DN_BN - 1) / DN_BN), unsigned((M + DN_BM - 1) / DN_BM), unsigned(nz));
  hipLaunchKernelGGL((k_dense<E, MODE, AT, BT>), grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, M, N, K, kc, bias,
                     C2, zp);
  PM_HIP(hipGetLastError());
}

// layer l of the training network on nv voxels: Z (co x nv) = W (co x ci) X
// (ci x nv, leading dimension ldin(l)); hidden layers add the bias and
// store z in Z and gelu(z) in act
void layer_fwd(sc_pmlp* c, int l, const float* X, float* Z, float* act, long nv) {
  const int ci = c->cin(l), co = c->cout(l);
  if (l < 3)
    dense<float, DN_BIAS_GELU_TRAIN, false, false>(c->stream, c->pf + c->fo[2 * l], ci, X, c->ldin(l), Z, co, co, nv,
                                                  ci, 1, c->pf + c->fo[2 * l + 1], act);
  else
    dense<float, DN_PLAIN, false, false>(c->stream, c->pf + c->fo[2 * l], ci, X, c->ldin(l), Z, co, co, nv, ci);
}

// partial-sum scratch of the parameter gradients (floats)
float* grad_scratch(sc_pmlp* c, long n) {
  if (c->dpart_len < n) {
    if (c->dpart) PM_HIP(hipFree(c->dpart));
    c->dpart = nullptr;
    PM_HIP(hipMalloc(&c->dpart, n * sizeof(float)));
    c->dpart_len = n;
  }
  return c->dpart;
}

void forward_impl(sc_pmlp* c, const float* x, float* y, long nv) {
  if (nv < 1 || nv > c->nvmax) throw std::runtime_error("voxel count out of range");
  if (c->arch == 2) synthesise<float>(c, c->pf, float(c->alpha()));
  const float* in = x;
  for (int l = 0; l < 3; ++l) {
    layer_fwd(c, l, in, c->z[l], c->a[l], nv);
    in = c->a[l];
  }
  if (c->arch == 2) {
    // the sink's D^2 outputs, then the symmetrising trace removal
    layer_fwd(c, 3, in, c->d1, nullptr, nv);
    hipLaunchKernelGGL(k_symm_sink<float>, dim3(nb(nv)), dim3(256), 0, c->stream, (const float*)c->d1, y, nv, c->D);
  } else if (c->arch == 1) {
    // basis coefficients, then the contraction with the rows' basis tensors
    layer_fwd(c, 3, in, c->d1, nullptr, nv);
    hipLaunchKernelGGL(k_tb_contract<float>, dim3(nb(nv)), dim3(256), 0, c->stream, (const float*)c->d1, x, nv, c->ctot,
                       c->nin, c->nt, c->nbas, y);
  } else {
    layer_fwd(c, 3, in, y, nullptr, nv);
    hipLaunchKernelGGL(k_tracefree<float>, dim3(nb(nv)), dim3(256), 0, c->stream, y, nv, c->nt, c->D);
  }
  PM_HIP(hipGetLastError());
  c->last_nv = nv;
}

// accumulate parameter gradients of the last forward for dy = dL/dy
void backward_impl(sc_pmlp* c, const float* x, const float* dy, long nv) {
  if (nv != c->last_nv) throw std::runtime_error("backward needs the voxel count of the last forward");
  float* d = c->d0;
  if (c->arch == 2) {
    hipLaunchKernelGGL(k_symm_sink_T, dim3(nb(nv)), dim3(256), 0, c->stream, dy, d, nv, c->D);
    PM_HIP(hipMemsetAsync(c->gf, 0, c->flen * sizeof(float), c->stream));
  } else if (c->arch == 1) {
    // dL/dw_b = Σ_t dy_t basis_tb
    hipLaunchKernelGGL(k_tb_dw, dim3(nb(nv)), dim3(256), 0, c->stream, dy, x, nv, c->ctot, c->nin, c->nt, c->nbas, d);
  } else {
    // dL/d(raw) = P(dy) (the trace removal is self-adjoint)
    PM_HIP(hipMemcpyAsync(d, dy, size_t(nv) * c->nt * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(k_tracefree<float>, dim3(nb(nv)), dim3(256), 0, c->stream, d, nv, c->nt, c->D);
  }
  PM_HIP(hipGetLastError());
  for (int l = 3; l >= 0; --l) {
    const int ci = c->cin(l), co = c->cout(l);
    const float* A = l == 0 ? x : c->a[l - 1];
    // dW (co x ci, stored [o][i]) += sum over the voxels of dZ Aᵀ: the voxels
    // split over nz blocks per 64 x 64 tile, the partials added in order
    const long tiles = long((ci + DN_BM - 1) / DN_BM) * ((co + DN_BN - 1) / DN_BN);
    const long nz = std::max(1L, std::min((nv + 255) / 256, 1024 / tiles));
    float* part = grad_scratch(c, nz * long(ci) * co + nz * co);
    dense<float, DN_PLAIN, true, true>(c->stream, A, c->ldin(l), d, co, part, ci, ci, co, nv, nz);
    hipLaunchKernelGGL(k_dense_reduce, dim3(nb(long(ci) * co)), dim3(256), 0, c->stream, (const float*)part, nz,
                       long(ci) * co, c->gf + c->fo[2 * l]);
    if (l < 3) {
      // db (co) += sum over the voxels of dZ
      float* pb = part + nz * long(ci) * co;
      const long chunk = (nv + nz - 1) / nz;
      hipLaunchKernelGGL(k_colsum_partials, dim3(unsigned((nv + chunk - 1) / chunk)), dim3(256), 0, c->stream,
                         (const float*)d, co, nv, chunk, pb);
      hipLaunchKernelGGL(k_dense_reduce, dim3(nb(co)), dim3(256), 0, c->stream, (const float*)pb,
                         (nv + chunk - 1) / chunk, long(co), c->gf + c->fo[2 * l + 1]);
    }
    PM_HIP(hipGetLastError());
    if (l > 0) {
      // dZ_{l-1} (ci x nv) = (Wᵀ dZ) * gelu'(z_{l-1})
      float* dn = d == c->d0 ? c->d1 : c->d0;
      dense<float, DN_GELU_BACK, true, false>(c->stream, c->pf + c->fo[2 * l], ci, d, co, dn, ci, ci, nv, co, 1,
                                              nullptr, nullptr, c->z[l - 1]);
      d = dn;
    }
  }
  if (c->arch == 2) synthesise_T(c);  // the trainable gradient += Sᵀ (network gradient)
}

double sum_partials(sc_pmlp* c, unsigned nblk) {
  std::vector<double> h(nblk);
  PM_HIP(hipMemcpyAsync(h.data(), c->part, nblk * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PM_HIP(hipStreamSynchronize(c->stream));
  double s = 0;
  for (double v : h) s += v;
  return s;
}

}  // namespace

// ------------------------------------------------------ solver-side interface
namespace pmlplink {
int grad_channels(const sc_pmlp* c) { return c->ng; }
int dim(const sc_pmlp* c) { return c->D; }
bool uses_redelta(const sc_pmlp* c) { return c->redelta; }
int arch(const sc_pmlp* c) { return c->arch; }

// the fused kernel's layout of this network (tom = the widest layer in
// tiles of 16 channels, lds = doubles of LDS); false when a layer is wider
// than FT tiles or the padded network exceeds 16384 doubles (128 KB)
bool fused_plan(const sc_pmlp* c, FusedNet& f, int& tom, int& lds) {
  tom = 1;
  for (int l = 0; l < 4; ++l) {
    f.cin[l] = c->cin(l);
    f.cout[l] = c->cout(l);
    f.ti[l] = (f.cin[l] + 15) / 16;
    f.to[l] = (f.cout[l] + 15) / 16;
    tom = std::max({tom, f.ti[l], f.to[l]});
    f.rows[l] = (f.cin[l] + 3) / 4 * 4;
    f.wsrc[l] = c->fo[2 * l];
  }
  if (tom > FT) return false;
  int off = 0;
  for (int l = 0; l < 4; ++l) {
    f.wofs[l] = off;
    off += f.rows[l] * 16 * tom;
  }
  for (int l = 0; l < 3; ++l) {
    f.bofs[l] = off;
    off += 16 * f.to[l];
    f.bsrc[l] = c->fo[2 * l + 1];
  }
  lds = off;
  return off <= 16384;
}

template <int TOM>
void launch_fused(sc_pmlp* c, const FusedNet& f, int lds, const double* x, long nv, double* out) {
  const long blocks = (nv + 64 * FR - 1) / (64 * FR);  // 4 wavefronts per block
  // LDS classes of 32, 72 and 128 KB: 4, 2 and 1 blocks per compute unit
  if (lds <= 4096)
    hipLaunchKernelGGL((k_mlp_fused<TOM, 4096>), dim3(unsigned(std::min(blocks, 4L * c->ncu))), dim3(256), 0,
                       c->stream, (const double*)c->pd, f, x, c->ctot, nv, out);
  else if (lds <= 9216)
    hipLaunchKernelGGL((k_mlp_fused<TOM, 9216>), dim3(unsigned(std::min(blocks, 2L * c->ncu))), dim3(256), 0,
                       c->stream, (const double*)c->pd, f, x, c->ctot, nv, out);
  else
    hipLaunchKernelGGL((k_mlp_fused<TOM, 16384>), dim3(unsigned(std::min(blocks, 1L * c->ncu))), dim3(256), 0,
                       c->stream, (const double*)c->pd, f, x, c->ctot, nv, out);
  PM_HIP(hipGetLastError());
}

// k_mlp_closure over the field: need = doubles of LDS (network and the
// wavefronts' staging at ybase)
template <int TOM, typename T>
void launch_closure_t(sc_pmlp* c, const FusedNet& f, int need, int ybase, const T* GG, long np, int G,
                      const double* zp, double d2, T* tau) {
  const long blocks = (np + 64 * FR - 1) / (64 * FR);  // 4 wavefronts of 16 FR voxels per block
  if (need <= 4096)
    hipLaunchKernelGGL((k_mlp_closure<TOM, 4096, T>), dim3(unsigned(std::min(blocks, 4L * c->ncu))), dim3(256), 0,
                       c->stream, (const double*)c->pd, f, ybase, GG, np, c->D, G, zp, d2, tau);
  else if (need <= 9216)
    hipLaunchKernelGGL((k_mlp_closure<TOM, 9216, T>), dim3(unsigned(std::min(blocks, 2L * c->ncu))), dim3(256), 0,
                       c->stream, (const double*)c->pd, f, ybase, GG, np, c->D, G, zp, d2, tau);
  else
    hipLaunchKernelGGL((k_mlp_closure<TOM, 16384, T>), dim3(unsigned(std::min(blocks, 1L * c->ncu))), dim3(256), 0,
                       c->stream, (const double*)c->pd, f, ybase, GG, np, c->D, G, zp, d2, tau);
  PM_HIP(hipGetLastError());
}
template <typename T>
void launch_closure(sc_pmlp* c, const FusedNet& f, int tom, int need, int ybase, const T* GG, long np, int G,
                    const double* zp, double d2, T* tau) {
  if (tom == 1) launch_closure_t<1, T>(c, f, need, ybase, GG, np, G, zp, d2, tau);
  else if (tom == 2) launch_closure_t<2, T>(c, f, need, ybase, GG, np, G, zp, d2, tau);
  else if (tom == 3) launch_closure_t<3, T>(c, f, need, ybase, GG, np, G, zp, d2, tau);
  else launch_closure_t<4, T>(c, f, need, ybase, GG, np, G, zp, d2, tau);
}

// k_tb_closure over the field (need = doubles of LDS; at most 9216)
template <int TOM, int DD, typename T>
void launch_tb_closure_t(sc_pmlp* c, const FusedNet& f, int need, int ybase, const T* GG, long np,
                         const double* zp, double d2, T* tau) {
  const long blocks = (np + 64 * FR - 1) / (64 * FR);
  if (need <= 4096)
    hipLaunchKernelGGL((k_tb_closure<TOM, 4096, DD, T>), dim3(unsigned(std::min(blocks, 4L * c->ncu))), dim3(256),
                       0, c->stream, (const double*)c->pd, f, ybase, GG, np, zp, d2, tau);
  else
    hipLaunchKernelGGL((k_tb_closure<TOM, 9216, DD, T>), dim3(unsigned(std::min(blocks, 2L * c->ncu))), dim3(256),
                       0, c->stream, (const double*)c->pd, f, ybase, GG, np, zp, d2, tau);
  PM_HIP(hipGetLastError());
}
template <typename T>
void launch_tb_closure(sc_pmlp* c, const FusedNet& f, int tom, int need, int ybase, const T* GG, long np,
                       const double* zp, double d2, T* tau) {
  if (c->D == 3) {
    if (tom == 1) launch_tb_closure_t<1, 3, T>(c, f, need, ybase, GG, np, zp, d2, tau);
    else if (tom == 2) launch_tb_closure_t<2, 3, T>(c, f, need, ybase, GG, np, zp, d2, tau);
    else launch_tb_closure_t<3, 3, T>(c, f, need, ybase, GG, np, zp, d2, tau);
  } else {
    if (tom == 1) launch_tb_closure_t<1, 2, T>(c, f, need, ybase, GG, np, zp, d2, tau);
    else if (tom == 2) launch_tb_closure_t<2, 2, T>(c, f, need, ybase, GG, np, zp, d2, tau);
    else launch_tb_closure_t<3, 2, T>(c, f, need, ybase, GG, np, zp, d2, tau);
  }
}

// one chunk of FP64 inference on nv input rows (leading dimension ctot):
// y (nt x nv) = symm(mlp(x)) or, for the TBNN, the basis contraction of the
// coefficients
void infer_chunk(sc_pmlp* c, const double* x, double* y, long nv) {
  const double* in = x;
  double* out = c->hA;
  FusedNet f;
  int tom = 0, lds = 0;
  if (fused_plan(c, f, tom, lds)) {
    // the whole network in one kernel: the last layer writes y (mlp) or the
    // head's input (hB)
    if (!c->ncu) {
      int dev = 0;
      PM_HIP(hipGetDevice(&dev));
      PM_HIP(hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, dev));
    }
    out = c->arch == 0 ? y : c->hB;
    if (tom == 1) launch_fused<1>(c, f, lds, x, nv, out);
    else if (tom == 2) launch_fused<2>(c, f, lds, x, nv, out);
    else if (tom == 3) launch_fused<3>(c, f, lds, x, nv, out);
    else launch_fused<4>(c, f, lds, x, nv, out);
  } else {
    // wider networks: one GEMM per layer, the activations through HBM
    for (int l = 0; l < 4; ++l) {
      const int ci = c->cin(l), co = c->cout(l);
      double* dst = l == 3 && c->arch == 0 ? y : out;  // the others end in a head
      const double* W = c->pd + c->fo[2 * l];
      const dim3 grid(unsigned((co + DM_BO - 1) / DM_BO), unsigned((nv + DM_BV - 1) / DM_BV));
      const long ldx = l == 0 ? c->ctot : ci;
      if (l < 3) {
        hipLaunchKernelGGL(k_dense_mfma64<true>, grid, dim3(256), 0, c->stream, in, ldx, W, long(ci),
                           (const double*)c->pd + c->fo[2 * l + 1], dst, long(co), nv, co, ci);
        PM_HIP(hipGetLastError());
        in = dst;
        out = dst == c->hA ? c->hB : c->hA;
      } else {
        hipLaunchKernelGGL(k_dense_mfma64<false>, grid, dim3(256), 0, c->stream, in, ldx, W, long(ci),
                           (const double*)nullptr, dst, long(co), nv, co, ci);
        PM_HIP(hipGetLastError());
      }
    }
  }
  if (c->arch == 2)
    hipLaunchKernelGGL(k_symm_sink<double>, dim3(nb(nv)), dim3(256), 0, c->stream, (const double*)out, y, nv, c->D);
  else if (c->arch == 1)
    hipLaunchKernelGGL(k_tb_contract<double>, dim3(nb(nv)), dim3(256), 0, c->stream, (const double*)out, x, nv,
                       c->ctot, c->nin, c->nt, c->nbas, y);
  else
    hipLaunchKernelGGL(k_tracefree<double>, dim3(nb(nv)), dim3(256), 0, c->stream, y, nv, c->nt, c->D);
  PM_HIP(hipGetLastError());
}

// fullchain's model (src/nets.jl:316-331), or tbnn's (:558-583), on the
// solver's stream s: the network runs in FP64 (the reference widens the
// trained parameters with `f64` before inference), over chunks of at most
// 2^20 input rows; sym != 0 averages the mlp over the cube's symmetry group
// (symmetrize_pointwise, src/nets.jl:350-389), |G| rows per voxel
template <typename T>
void closure_tau(sc_pmlp* c, const T* GG, long np, double delta, double visc, double mu, double sigma, int sym,
                 T* tau, hipStream_t s) {
  constexpr long CHUNK = 1L << 20, CHUNK_ELEMS = 1L << 27;
  if (sym && c->arch != 0) throw std::runtime_error("the group average applies to the mlp closure");
  const int G = sym ? (c->D == 3 ? 48 : 8) : 1;
  hipStream_t old = c->stream;
  c->stream = s;
  try {
    // the mlp whose padded network and per-wavefront staging fit in LDS runs
    // as one fused kernel over the field; otherwise chunks of features, the
    // network and the output scaling
    FusedNet f;
    int tom = 0, lds = 0;
    const int ng = c->D * c->D;
    const bool planned = (c->arch == 0 || c->arch == 1) && fused_plan(c, f, tom, lds);
    const int need = c->arch == 0 ? lds + 4 * 16 * FR * (ng + c->nt) : lds + 4 * 16 * FR * (c->nin + c->nbas);
    const bool fused = planned && (c->arch == 0 ? need <= 16384 : (tom <= 3 && need <= 9216));
    // voxels per chunk: at most 2^20 input rows and 2^27 activations
    const long vc = std::max(1L, std::min({np, CHUNK / G, CHUNK_ELEMS / (long(c->maxw()) * G)}));
    const long cap = vc * G;                  // input rows per chunk
    if (!fused && cap > c->icap) {
      for (double** q : {&c->xd, &c->hA, &c->hB, &c->yd})
        if (*q) {
          PM_HIP(hipFree(*q));
          *q = nullptr;
        }
      c->icap = 0;
      PM_HIP(hipMalloc(&c->xd, size_t(cap) * c->ctot * sizeof(double)));
      PM_HIP(hipMalloc(&c->hA, size_t(cap) * c->maxw() * sizeof(double)));
      PM_HIP(hipMalloc(&c->hB, size_t(cap) * c->maxw() * sizeof(double)));
      PM_HIP(hipMalloc(&c->yd, size_t(cap) * c->nt * sizeof(double)));
      c->icap = cap;
    }
    if (!c->pd) {
      PM_HIP(hipMalloc(&c->pd, c->flen * sizeof(double)));
      PM_HIP(hipMalloc(&c->zdev, sizeof(double)));
    }
    if (c->arch == 2) {
      // the synthesis in FP64 from the FP32 trainables: alpha is the FP32
      // 1/sqrt|G| of the reference's synthesis matrices
      synthesise<double>(c, c->pd, double(c->alpha()));
    } else {
      hipLaunchKernelGGL(k_widen, dim3(nb(c->plen)), dim3(256), 0, s, c->p, c->pd, c->plen);
      PM_HIP(hipGetLastError());
    }
    const double d2 = delta * delta;
    if (c->redelta) {
      const unsigned nblk = std::min(nb(np), 1024u);
      hipLaunchKernelGGL(k_a2_partials<T>, dim3(nblk), dim3(256), 0, s, GG, np, c->ng, c->part);
      PM_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_redelta_z, dim3(1), dim3(64), 0, s, (const double*)c->part, int(nblk), np, d2, visc, mu,
                         sigma, sizeof(T) == 4 ? 1 : 0, c->zdev);
      PM_HIP(hipGetLastError());
    }
    const double* zp = c->redelta ? (const double*)c->zdev : nullptr;
    if (fused) {
      if (!c->ncu) {
        int dev = 0;
        PM_HIP(hipGetDevice(&dev));
        PM_HIP(hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, dev));
      }
      if (c->arch == 0)
        launch_closure<T>(c, f, tom, need, lds, GG, np, G, zp, d2, tau);
      else
        launch_tb_closure<T>(c, f, tom, need, lds, GG, np, zp, d2, tau);
    } else {
      for (long v0 = 0; v0 < np; v0 += vc) {
        const long nv = std::min(vc, np - v0);
        if (c->arch == 1)
          hipLaunchKernelGGL(k_tb_features_inf<T>, dim3(nb(nv)), dim3(256), 0, s, GG, np, v0, nv, c->D, zp, c->xd,
                             c->ctot, c->ni, c->nt, c->nbas);
        else if (sym)
          hipLaunchKernelGGL(k_features_sym<T>, dim3(nb(nv * G)), dim3(256), 0, s, GG, np, v0, nv, c->D, G, zp, c->xd,
                             c->nin);
        else
          hipLaunchKernelGGL(k_features_inf<T>, dim3(nb(nv)), dim3(256), 0, s, GG, np, v0, nv, c->ng, zp, c->xd,
                             c->nin);
        PM_HIP(hipGetLastError());
        infer_chunk(c, c->xd, c->yd, nv * G);
        if (sym)
          hipLaunchKernelGGL(k_sym_out<T>, dim3(nb(nv)), dim3(256), 0, s, (const double*)c->yd, GG, np, v0, nv, c->D,
                             c->nt, G, d2, tau);
        else
          hipLaunchKernelGGL(k_tau_out<T>, dim3(nb(nv)), dim3(256), 0, s, (const double*)c->yd, GG, np, v0, nv, c->ng,
                             c->nt, d2, tau);
        PM_HIP(hipGetLastError());
      }
    }
  } catch (...) {
    c->stream = old;
    throw;
  }
  c->stream = old;
}
template void closure_tau<float>(sc_pmlp*, const float*, long, double, double, double, double, int, float*,
                                 hipStream_t);
template void closure_tau<double>(sc_pmlp*, const double*, long, double, double, double, double, int, double*,
                                  hipStream_t);
}  // namespace pmlplink

extern "C" {

namespace {
// group_stuff's tables (src/symmetry.jl:42-67) for the synthesis kernels:
// perm, sign, pinv [nreg][3], relidx[m][n] = index of g_n^-1 g_m (the Cayley
// gather of get_weight_synthesis), ninv[m][k] = the n with relidx[m][n] = k
std::vector<int> group_tables(int D) {
  const int np = D == 3 ? 6 : 2, ns = D == 3 ? 8 : 4, nreg = np * ns;
  const int P3[6][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}, {1, 0, 2}, {0, 2, 1}};
  const int S3[8][3] = {{1, 1, 1}, {-1, 1, 1}, {1, -1, 1}, {1, 1, -1}, {-1, -1, 1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, -1}};
  const int P2[2][3] = {{0, 1, 2}, {1, 0, 2}};
  const int S2[4][3] = {{1, 1, 1}, {-1, 1, 1}, {1, -1, 1}, {-1, -1, 1}};
  std::vector<int> t(9 * nreg + 2 * nreg * nreg, 0);
  int* perm = t.data();
  int* sign = perm + 3 * nreg;
  int* pinv = perm + 6 * nreg;
  int* relidx = perm + 9 * nreg;
  int* ninv = relidx + nreg * nreg;
  std::vector<std::array<int, 9>> mats(nreg);
  for (int e = 0; e < nreg; ++e) {
    const int pi = e % np, si = e / np;
    for (int a = 0; a < 3; ++a) {
      perm[3 * e + a] = D == 3 ? P3[pi][a] : P2[pi][a];
      sign[3 * e + a] = D == 3 ? S3[si][a] : S2[si][a];
    }
    for (int a = 0; a < D; ++a) pinv[3 * e + perm[3 * e + a]] = a;
    std::array<int, 9> M{};
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) M[3 * i + j] = sign[3 * e + i] * (perm[3 * e + i] == j ? 1 : 0);
    mats[e] = M;
  }
  auto find = [&](const std::array<int, 9>& M) {
    for (int e = 0; e < nreg; ++e)
      if (mats[e] == M) return e;
    throw std::runtime_error("group tables: product outside the group");
  };
  auto mul = [&](int a, int b) {
    std::array<int, 9> C{};
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j)
        for (int k = 0; k < D; ++k) C[3 * i + j] += mats[a][3 * i + k] * mats[b][3 * k + j];
    return find(C);
  };
  std::vector<int> inv(nreg);
  for (int e = 0; e < nreg; ++e)
    for (int f = 0; f < nreg; ++f)
      if (mul(e, f) == 0) inv[e] = f;
  for (int m = 0; m < nreg; ++m)
    for (int n = 0; n < nreg; ++n) {
      const int k = mul(inv[n], m);
      relidx[m * nreg + n] = k;
      ninv[m * nreg + k] = n;
    }
  return t;
}

int create_impl(int arch, int D, int use_redelta, int c1, int c2, int c3, uint64_t seed, long nvox_max,
                sc_pmlp** out) {
  if (!out) return 1;
  *out = nullptr;
  // every activation and input matrix (nvox_max x width) stays below 2^31 entries
  const long nreg = arch == 2 ? (D == 3 ? 48 : 8) : 1;
  if ((D != 2 && D != 3) || c1 < 1 || c2 < 1 || c3 < 1 || c1 > 1024 || c2 > 1024 || c3 > 1024 || nvox_max < 1 ||
      nvox_max * std::max({nreg * c1, nreg * c2, nreg * c3, 48L}) >= (1L << 31))
    return 1;
  g_pmlp_err.clear();
  (void)hipGetLastError();
  sc_pmlp* c = new sc_pmlp();
  c->arch = arch;
  c->D = D;
  c->ng = D * D;
  c->nt = D == 3 ? 6 : 3;
  c->redelta = use_redelta != 0;
  if (arch == 1) {
    // invariants and basis of tb_kernel! (src/nets.jl:416-453)
    c->ni = D == 3 ? 5 : 2;
    c->nbas = D == 3 ? 7 : 2;
    c->nin = c->ni + (c->redelta ? 1 : 0);
    c->ctot = c->nin + c->nt * c->nbas;
  } else {
    c->nin = c->ng + (c->redelta ? 1 : 0);
    c->ctot = c->nin;
    if (arch == 2) {
      c->nreg = int(nreg);
      c->nten = c->ng;
    }
  }
  c->w[0] = c1; c->w[1] = c2; c->w[2] = c3;
  for (int l = 0; l < 3; ++l) c->wf[l] = c->w[l] * c->nreg;
  c->nvmax = nvox_max;
  try {
    PM_HIP(hipStreamCreate(&c->stream));
    c->own_stream = true;
    long o = 0;
    for (int l = 0; l < 4; ++l) {
      c->fo[2 * l] = o;
      o += long(c->cin(l)) * c->cout(l);
      if (l < 3) {
        c->fo[2 * l + 1] = o;
        o += c->cout(l);
      }
    }
    c->fo[7] = o;
    c->flen = o;
    // trainable arrays with their glorot_uniform bounds (Lux's fans of each
    // array; biases zero)
    std::vector<double> lims;
    if (arch == 2) {
      // lift (weight, bias[, weight_re]), mid_1 and mid_2 (weight, bias), sink
      const int re = c->redelta ? 1 : 0, nten = c->nten;
      std::vector<std::pair<long, double>> arrs = {{long(nten) * c1, std::sqrt(6.0 / (nten + c1))}, {c1, 0.0}};
      if (re) arrs.push_back({c1, std::sqrt(6.0 / (1 + c1))});
      for (int l = 1; l < 3; ++l) {
        arrs.push_back({nreg * c->w[l] * c->w[l - 1], std::sqrt(6.0 / (nreg * (c->w[l] + c->w[l - 1])))});
        arrs.push_back({c->w[l], 0.0});
      }
      arrs.push_back({long(nten) * c3, std::sqrt(6.0 / (nten + c3))});
      long q = 0;
      for (size_t i = 0; i < arrs.size(); ++i) {
        c->cb[i] = q;
        q += arrs[i].first;
        lims.push_back(arrs[i].second);
      }
      c->ncarr = int(arrs.size());
      c->cb[c->ncarr] = q;
      c->plen = q;
    } else {
      for (int i = 0; i < 8; ++i) c->cb[i] = c->fo[i];
      for (int l = 0; l < 4; ++l) {
        lims.push_back(std::sqrt(6.0 / (c->cin(l) + c->cout(l))));
        if (l < 3) lims.push_back(0.0);
      }
      c->ncarr = 7;
      c->plen = o;
    }
    o = c->plen;
    // uniform(-lim, lim) from a splitmix64 stream
    std::vector<float> h(o, 0.f);
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 777;
    auto rnd = [&]() {
      s += 0x9E3779B97F4A7C15ull;
      uint64_t zz = s;
      zz = (zz ^ (zz >> 30)) * 0xBF58476D1CE4E5B9ull;
      zz = (zz ^ (zz >> 27)) * 0x94D049BB133111EBull;
      zz ^= zz >> 31;
      return double(zz >> 11) * 0x1.0p-53;
    };
    for (int a = 0; a < c->ncarr; ++a)
      if (lims[a] > 0)
        for (long i = c->cb[a]; i < c->cb[a + 1]; ++i) h[i] = float((2 * rnd() - 1) * lims[a]);
    for (float** q : {&c->p, &c->g, &c->m, &c->v}) PM_HIP(hipMalloc(q, o * sizeof(float)));
    PM_HIP(hipMemcpy(c->p, h.data(), o * sizeof(float), hipMemcpyHostToDevice));
    for (float* q : {c->g, c->m, c->v}) PM_HIP(hipMemset(q, 0, o * sizeof(float)));
    if (arch == 2) {
      PM_HIP(hipMalloc(&c->pf, c->flen * sizeof(float)));
      PM_HIP(hipMalloc(&c->gf, c->flen * sizeof(float)));
      const std::vector<int> t = group_tables(D);
      PM_HIP(hipMalloc(&c->gt, t.size() * sizeof(int)));
      PM_HIP(hipMemcpy(c->gt, t.data(), t.size() * sizeof(int), hipMemcpyHostToDevice));
    } else {
      c->pf = c->p;
      c->gf = c->g;
    }
    const size_t nv = size_t(nvox_max);
    for (int l = 0; l < 3; ++l) {
      PM_HIP(hipMalloc(&c->z[l], nv * c->wf[l] * sizeof(float)));
      PM_HIP(hipMalloc(&c->a[l], nv * c->wf[l] * sizeof(float)));
    }
    PM_HIP(hipMalloc(&c->d0, nv * c->maxw() * sizeof(float)));
    PM_HIP(hipMalloc(&c->d1, nv * c->maxw() * sizeof(float)));
    PM_HIP(hipMalloc(&c->part, 1024 * sizeof(double)));
    PM_HIP(hipMalloc(&c->gnorm, NORM_DOUBLES * sizeof(double)));
  } catch (std::exception& e) {
    g_pmlp_err = e.what();
    sc_pmlp_destroy(c);
    return 3;
  }
  *out = c;
  return 0;
}
}  // namespace

int sc_pmlp_create(int D, int use_redelta, int c1, int c2, int c3, uint64_t seed, long nvox_max, sc_pmlp** out) {
  return create_impl(0, D, use_redelta, c1, c2, c3, seed, nvox_max, out);
}

int sc_tbnn_create(int D, int use_redelta, int c1, int c2, int c3, uint64_t seed, long nvox_max, sc_pmlp** out) {
  return create_impl(1, D, use_redelta, c1, c2, c3, seed, nvox_max, out);
}

int sc_equi_create(int D, int use_redelta, int c1, int c2, int c3, uint64_t seed, long nvox_max, sc_pmlp** out) {
  return create_impl(2, D, use_redelta, c1, c2, c3, seed, nvox_max, out);
}

int sc_pmlp_destroy(sc_pmlp* c) {
  if (!c) return 0;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void* ptrs[] = {c->p, c->g, c->m, c->v, c->d0, c->d1, c->dpart, c->pd, c->zdev, c->xd, c->hA, c->hB, c->yd, c->part, c->gnorm,
                  c->gt};
  for (void* q : ptrs) if (q) (void)hipFree(q);
  if (c->arch == 2) {
    if (c->pf) (void)hipFree(c->pf);
    if (c->gf) (void)hipFree(c->gf);
  }
  for (int l = 0; l < 3; ++l) {
    if (c->z[l]) (void)hipFree(c->z[l]);
    if (c->a[l]) (void)hipFree(c->a[l]);
  }
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

const char* sc_pmlp_last_error(const sc_pmlp* c) { return c ? c->err.c_str() : g_pmlp_err.c_str(); }

int sc_pmlp_info(const sc_pmlp* c, int64_t* info) {
  if (!c || !info) return 1;
  info[0] = c->D; info[1] = c->nin; info[2] = c->w[0]; info[3] = c->w[1]; info[4] = c->w[2];
  info[5] = c->nt; info[6] = c->plen; info[7] = c->nvmax; info[8] = c->arch; info[9] = c->ctot;
  info[10] = c->nreg; info[11] = c->flen;
  return 0;
}

int sc_pmlp_set_stream(sc_pmlp* c, void* s) {
  return pguard(c, [&] {
    PM_HIP(hipStreamSynchronize(c->stream));
    if (c->own_stream) PM_HIP(hipStreamDestroy(c->stream));
    c->stream = (hipStream_t)s;
    c->own_stream = false;
  });
}

int sc_pmlp_get_params(sc_pmlp* c, float* host) {
  return pguard(c, [&] { PM_HIP(hipMemcpyAsync(host, c->p, c->plen * 4, hipMemcpyDeviceToHost, c->stream)); });
}
int sc_pmlp_set_params(sc_pmlp* c, const float* host) {
  return pguard(c, [&] { PM_HIP(hipMemcpyAsync(c->p, host, c->plen * 4, hipMemcpyHostToDevice, c->stream)); });
}
int sc_pmlp_get_grads(sc_pmlp* c, float* host) {
  return pguard(c, [&] { PM_HIP(hipMemcpyAsync(host, c->g, c->plen * 4, hipMemcpyDeviceToHost, c->stream)); });
}
int sc_pmlp_zero_grad(sc_pmlp* c) {
  return pguard(c, [&] { PM_HIP(hipMemsetAsync(c->g, 0, c->plen * 4, c->stream)); });
}

int sc_pmlp_features(sc_pmlp* c, const void* GG, const void* tau, int fp64, long np, double delta, double z, void* x,
                     void* y) {
  return pguard(c, [&] {
    if (np < 1 || !x || (tau && !y)) throw std::runtime_error("arguments");
    const double d2 = delta * delta;
    const int re = c->redelta ? 1 : 0;
    if (c->arch == 1) {
      if (fp64)
        hipLaunchKernelGGL(k_tb_features<double>, dim3(nb(np)), dim3(256), 0, c->stream, (const double*)GG, np, c->D,
                           d2, re, float(z), (const double*)tau, c->nt, (float*)x, c->ctot, c->ni, c->nbas,
                           (float*)y);
      else
        hipLaunchKernelGGL(k_tb_features<float>, dim3(nb(np)), dim3(256), 0, c->stream, (const float*)GG, np, c->D,
                           d2, re, float(z), (const float*)tau, c->nt, (float*)x, c->ctot, c->ni, c->nbas,
                           (float*)y);
    } else if (fp64) {
      hipLaunchKernelGGL(k_features<double>, dim3(nb(np)), dim3(256), 0, c->stream, (const double*)GG, np, c->ng, d2,
                         re, float(z), (const double*)tau, c->nt, (float*)x, c->nin, (float*)y);
    } else {
      hipLaunchKernelGGL(k_features<float>, dim3(nb(np)), dim3(256), 0, c->stream, (const float*)GG, np, c->ng, d2,
                         re, float(z), (const float*)tau, c->nt, (float*)x, c->nin, (float*)y);
    }
    PM_HIP(hipGetLastError());
  });
}

int sc_pmlp_forward(sc_pmlp* c, const void* x, void* y, long nvox) {
  return pguard(c, [&] { forward_impl(c, (const float*)x, (float*)y, nvox); });
}

int sc_pmlp_backward(sc_pmlp* c, const void* x, const void* dy, long nvox) {
  return pguard(c, [&] { backward_impl(c, (const float*)x, (const float*)dy, nvox); });
}

int sc_pmlp_loss(sc_pmlp* c, const void* y, const void* target, void* dy, long nvox, int kind, double* loss) {
  return pguard(c, [&] {
    if (kind != 0 && kind != 1) throw std::runtime_error("loss kind must be 0 (mse) or 1 (relative)");
    const long n = nvox * c->nt;
    const unsigned nblk = std::min(nb(n), 1024u);
    double inv = 1.0 / double(n);
    if (kind == 1) {
      hipLaunchKernelGGL(k_sumsq, dim3(nblk), dim3(256), 0, c->stream, (const float*)target, n, c->part);
      PM_HIP(hipGetLastError());
      inv = 1.0 / (sum_partials(c, nblk) + double(FLT_EPSILON));
    }
    hipLaunchKernelGGL(k_loss, dim3(nblk), dim3(256), 0, c->stream, (const float*)y, (const float*)target, (float*)dy, n,
                       2.0 * inv, c->part);
    PM_HIP(hipGetLastError());
    const double s = sum_partials(c, nblk);
    if (loss) *loss = s * inv;
  });
}

int sc_pmlp_adamw(sc_pmlp* c, double lr, double b1, double b2, double eps, double weight_decay, double clip) {
  return pguard(c, [&] {
    c->adam_t += 1;
    const float c1 = float(1.0 - std::pow(b1, c->adam_t)), c2 = float(1.0 - std::pow(b2, c->adam_t));
    Arrays A;
    A.n = c->ncarr;
    for (int i = 0; i <= A.n; ++i) A.b[i] = c->cb[i];
    double* nrm2 = c->gnorm + NORM_OFFSET;
    if (clip > 0) {
      hipLaunchKernelGGL(k_array_sumsq, dim3(NORM_BLOCKS, A.n), dim3(256), 0, c->stream, c->g, A, c->gnorm);
      PM_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_finish_norms, dim3(1), dim3(64), 0, c->stream, c->gnorm, nrm2, A.n);
      PM_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_adamw, dim3(unsigned((c->plen + 255) / 256)), dim3(256), 0, c->stream, c->p, c->g, c->m, c->v,
                       c->plen, float(lr), float(b1), float(b2), float(eps), float(weight_decay), c1, c2, nrm2, A,
                       clip);
    PM_HIP(hipGetLastError());
  });
}

}  // extern "C"
