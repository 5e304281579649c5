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
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#include "../../include/symcode.h"
#include "cnn_link.hpp"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

namespace cnn {

#define CNN_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

__device__ __forceinline__ float gelu(float z) {
  const float c = 0.7978845608028654f;  // sqrt(2/pi)
  float t = tanhf(c * (z + 0.044715f * z * z * z));
  return 0.5f * z * (1.0f + t);
}
__device__ __forceinline__ float gelu_grad(float z) {
  const float c = 0.7978845608028654f;
  float u = c * (z + 0.044715f * z * z * z);
  float t = tanhf(u);
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * z * z);
}

enum Mode { HIDDEN = 0, LAST = 1, BWD_HIDDEN = 2, BWD_INPUT = 3 };

constexpr int VPB = 64;   // output voxels per block (one x segment)
constexpr int PADC = 8;   // LDS channel padding (16 bytes) against bank conflicts

// One convolution layer (or its data gradient).  n must be a multiple of 64.
//  HIDDEN:     z = conv + bias; out_z = z, out_a = gelu(z)            (bf16)
//  LAST:       out_f = conv                                           (fp32)
//  BWD_HIDDEN: out_a = conv * gelu'(zprev)                            (bf16)
//  BWD_INPUT:  out_f = conv                                           (fp32)
template <int CIN, int COUT, int MODE>
__global__ void __launch_bounds__(256) k_conv(const bf16* __restrict__ in, const bf16* __restrict__ W,
                                              const float* __restrict__ bias, bf16* out_a, bf16* out_z,
                                              float* out_f, const bf16* __restrict__ zprev, int n) {
  constexpr int CL = CIN + PADC;
  __shared__ __attribute__((aligned(16))) bf16 win[9 * (VPB + 2) * CL];
  const int segs = n / VPB;
  const long row = blockIdx.x / segs;  // (b*n + z)*n + y
  const int x0 = (blockIdx.x % segs) * VPB;
  const int y = int(row % n);
  const long bz = row / n;
  const int z = int(bz % n);
  const long b = bz / n;
  // stage the 9 neighbour rows, 16 bytes (8 channels) per load
  constexpr int CH8 = CIN / 8;
  for (int i = threadIdx.x; i < 9 * (VPB + 2) * CH8; i += blockDim.x) {
    const int c8 = i % CH8;
    const int xx = (i / CH8) % (VPB + 2);
    const int rr = i / (CH8 * (VPB + 2));
    const int dz = rr / 3 - 1, dy = rr % 3 - 1;
    const int zs = (z + dz + n) % n, ys = (y + dy + n) % n, xs = (x0 + xx - 1 + n) % n;
    const long src = (((b * n + zs) * n + ys) * n + xs) * CIN + c8 * 8;
    *reinterpret_cast<bf16x8*>(&win[(rr * (VPB + 2) + xx) * CL + c8 * 8]) =
        *reinterpret_cast<const bf16x8*>(&in[src]);
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int li = l & 15, lk = l >> 4;
  floatx4 acc[COUT / 16];
#pragma unroll
  for (int nb = 0; nb < COUT / 16; ++nb) acc[nb] = floatx4{0, 0, 0, 0};
#pragma unroll 1
  for (int t = 0; t < 27; ++t) {
    const int rr = t / 3, dx = t % 3 - 1;
    const bf16* arow = &win[(rr * (VPB + 2) + 1 + w * 16 + li + dx) * CL + 8 * lk];
    const bf16* wt = W + long(t) * COUT * CIN + 8 * lk;
#pragma unroll
    for (int kc = 0; kc < CIN / 32; ++kc) {
      bf16x8 a = *reinterpret_cast<const bf16x8*>(arow + kc * 32);
#pragma unroll
      for (int nb = 0; nb < COUT / 16; ++nb) {
        bf16x8 bv = *reinterpret_cast<const bf16x8*>(wt + long(nb * 16 + li) * CIN + kc * 32);
        acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[nb], 0, 0, 0);
      }
    }
  }
  // epilogue: lane holds D[voxel = 4*lk + r][co = li + 16 nb]
  const long vbase = row * n + x0 + w * 16;
#pragma unroll
  for (int nb = 0; nb < COUT / 16; ++nb) {
    const int co = nb * 16 + li;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const long v = vbase + 4 * lk + r;
      float d = acc[nb][r];
      if constexpr (MODE == HIDDEN) {
        d += bias[co];
        out_z[v * COUT + co] = bf16(d);
        out_a[v * COUT + co] = bf16(gelu(d));
      } else if constexpr (MODE == LAST) {
        out_f[v * COUT + co] = d;
      } else if constexpr (MODE == BWD_HIDDEN) {
        out_a[v * COUT + co] = bf16(d * gelu_grad(float(zprev[v * COUT + co])));
      } else {
        out_f[v * COUT + co] = d;
      }
    }
  }
}

// Weight gradient of one layer for the three taps sharing (dz, dy) = blockIdx.y:
//   dW[t][co][ci] += sum_v dout[v][co] * in[v + (dz,dy,dx)][ci]
// over the rows [blockIdx.x * rpb, +rpb).  Partial sums go to
// part[blockIdx.x][t][co][ci]; k_reduce_partials adds them up.
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_conv_wgrad(const bf16* __restrict__ in, const bf16* __restrict__ dout,
                                                    float* part, int n, long nrows, int rpb) {
  constexpr int XL = 136;  // 128 voxels + halo, padded (x contiguous in LDS)
  __shared__ __attribute__((aligned(16))) bf16 dT[COUT * XL];
  __shared__ __attribute__((aligned(16))) bf16 iT[CIN * XL];
  const int rr = blockIdx.y;
  const int dz = rr / 3 - 1, dy = rr % 3 - 1;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int li = l & 15, lk = l >> 4;
  // wave w owns output-channel blocks cb = w, w+4, ... and all input blocks
  constexpr int NCB = COUT / 16, NIB = CIN / 16;
  constexpr int MYCB = (NCB + 3) / 4;
  floatx4 acc[MYCB][NIB][3];
#pragma unroll
  for (int a = 0; a < MYCB; ++a)
#pragma unroll
    for (int i = 0; i < NIB; ++i)
#pragma unroll
      for (int d = 0; d < 3; ++d) acc[a][i][d] = floatx4{0, 0, 0, 0};
  const long r0 = long(blockIdx.x) * rpb;
#pragma unroll 1
  for (long row = r0; row < r0 + rpb && row < nrows; ++row) {
    const int y = int(row % n);
    const long bz = row / n;
    const int z = int(bz % n);
    const long b = bz / n;
    const long srow = ((b * n + (z + dz + n) % n) * n + (y + dy + n) % n) * n;
    __syncthreads();
    // stage dout[row][x][co] transposed to dT[co][x], and the shifted input
    // row in[srow][x-1 .. x+n][ci] transposed to iT[ci][x]
    for (int i = threadIdx.x; i < n * (COUT / 8); i += blockDim.x) {
      const int c8 = i % (COUT / 8), x = i / (COUT / 8);
      bf16x8 v8 = *reinterpret_cast<const bf16x8*>(&dout[(row * n + x) * COUT + c8 * 8]);
#pragma unroll
      for (int e = 0; e < 8; ++e) dT[(c8 * 8 + e) * XL + x] = v8[e];
    }
    for (int i = threadIdx.x; i < (n + 2) * (CIN / 8); i += blockDim.x) {
      const int c8 = i % (CIN / 8), xx = i / (CIN / 8);
      const int xs = (xx - 1 + n) % n;
      bf16x8 v8 = *reinterpret_cast<const bf16x8*>(&in[(srow + xs) * CIN + c8 * 8]);
#pragma unroll
      for (int e = 0; e < 8; ++e) iT[(c8 * 8 + e) * XL + xx] = v8[e];
    }
    __syncthreads();
    // K = voxels x in [0, n), 32 per MFMA
#pragma unroll 1
    for (int k0 = 0; k0 < n; k0 += 32) {
#pragma unroll
      for (int a = 0; a < MYCB; ++a) {
        const int cb = w + 4 * a;
        if (cb >= NCB) continue;
        bf16x8 av;
#pragma unroll
        for (int e = 0; e < 8; ++e) av[e] = dT[(cb * 16 + li) * XL + k0 + 8 * lk + e];
#pragma unroll
        for (int ib = 0; ib < NIB; ++ib) {
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            // output voxel x reads input x + dx -> iT column x + dx + 1
            bf16x8 bv;
#pragma unroll
            for (int e = 0; e < 8; ++e) bv[e] = iT[(ib * 16 + li) * XL + k0 + 8 * lk + e + d];
            acc[a][ib][d] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[a][ib][d], 0, 0, 0);
          }
        }
      }
    }
  }
  // D[i = co in block][j = ci in block]: lane holds rows 4*lk + r, column li
#pragma unroll
  for (int a = 0; a < MYCB; ++a) {
    const int cb = w + 4 * a;
    if (cb >= NCB) continue;
#pragma unroll
    for (int ib = 0; ib < NIB; ++ib)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int t = rr * 3 + d;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int co = cb * 16 + 4 * lk + r, ci = ib * 16 + li;
          part[((long(blockIdx.x) * 27 + t) * COUT + co) * CIN + ci] = acc[a][ib][d][r];
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

// bias gradient: db[c] += sum_v d[v][c] (block partials then atomic add)
template <int C>
__global__ void k_bias_grad(const bf16* d, float* db, long nvox) {
  __shared__ float sm[256];
  const int c = threadIdx.x % C;
  const int lane = threadIdx.x / C;
  const int lanes = 256 / C;
  float s = 0.f;
  for (long v = long(blockIdx.x) * lanes + lane; v < nvox; v += long(gridDim.x) * lanes) s += float(d[v * C + c]);
  sm[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < C) {
    float t = 0.f;
    for (int i = 0; i < lanes; ++i) t += sm[i * C + threadIdx.x];
    atomicAdd(&db[threadIdx.x], t);
  }
}

// MSE loss over the first `nout` channels of y (fp32, channel pitch cp):
// loss partials, and the gradient dy = 2 (y - t) / (nvox * nout) as bf16 with
// channel pitch cp (padding channels zero).
__global__ void k_mse(const float* y, const float* tgt, bf16* dy, float* dyf, double* part, long nvox, int nout,
                      int cp, double inv) {
  __shared__ double sm[256];
  long v = long(blockIdx.x) * blockDim.x + threadIdx.x;
  double acc = 0;
  if (v < nvox) {
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
__global__ void k_adam(float* w, const float* g, float* m, float* v, long len, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= len) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1 - b1) * gi;
  float vi = b2 * v[i] + (1 - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// master [27][cout][cin] -> forward bf16 [27][cop][cip] and mirrored
// transposed bf16 [27][cip][cop] (tap t -> 26 - t)
__global__ void k_pack_weights(const float* w, bf16* wf, bf16* wb, int cout, int cin, int cop, int cip) {
  long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  long len = 27L * cop * cip;
  if (i >= len) return;
  int ci = int(i % cip), co = int((i / cip) % cop), t = int(i / (long(cip) * cop));
  float val = (co < cout && ci < cin) ? w[(long(t) * cout + co) * cin + ci] : 0.f;
  wf[i] = bf16(val);
  wb[(long(26 - t) * cip + ci) * cop + co] = bf16(val);
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
  int cin_p = 32, cout_p = 32;
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
  // one-cube input / output / cotangent buffers of the LES closure (allocated
  // on first use by the solver)
  bf16 *a0 = nullptr, *dyb = nullptr;
  float *y = nullptr, *dx = nullptr, *part = nullptr;
  long part_len = 0;
  double* lossp = nullptr;
  double* hostd = nullptr;
  int adam_t = 0;
};

namespace {
using namespace cnn;
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
  dim3 grid(unsigned(rows * (c->n / VPB)));
#define CONV(CI, CO, MO)                                                                                   \
  if (cinp == CI && coutp == CO && mode == MO) {                                                          \
    hipLaunchKernelGGL((k_conv<CI, CO, MO>), grid, dim3(256), 0, c->stream, in, W, bias, oa, oz, of, zp, c->n); \
    CNN_HIP(hipGetLastError());                                                                           \
    return;                                                                                               \
  }
  CONV(32, 64, HIDDEN) CONV(64, 64, HIDDEN) CONV(64, 32, LAST)
  CONV(32, 64, BWD_HIDDEN) CONV(64, 64, BWD_HIDDEN) CONV(64, 32, BWD_INPUT)
#undef CONV
  throw std::runtime_error("unsupported conv shape");
}

void launch_wgrad(sc_cnn* c, int cinp, int coutp, const bf16* in, const bf16* dout, float* grad, int batch) {
  const long rows = long(batch) * c->n * c->n;
  const int rpb = 64;
  const long nchunks = (rows + rpb - 1) / rpb;
  const long len = 27L * coutp * cinp;
  if (c->part_len < nchunks * len) {
    if (c->part) CNN_HIP(hipFree(c->part));
    CNN_HIP(hipMalloc(&c->part, nchunks * len * sizeof(float)));
    c->part_len = nchunks * len;
  }
  dim3 grid(unsigned(nchunks), 9);
#define WG(CI, CO)                                                                                              \
  if (cinp == CI && coutp == CO) {                                                                               \
    hipLaunchKernelGGL((k_conv_wgrad<CI, CO>), grid, dim3(256), 0, c->stream, in, dout, c->part, c->n, rows, rpb); \
    CNN_HIP(hipGetLastError());                                                                                  \
  } else
  WG(32, 64) WG(64, 64) WG(64, 32) { throw std::runtime_error("unsupported wgrad shape"); }
#undef WG
  hipLaunchKernelGGL(k_reduce_partials, dim3(unsigned((len + 255) / 256)), dim3(256), 0, c->stream, c->part, grad,
                     nchunks, len);
  CNN_HIP(hipGetLastError());
}

void pack_all(sc_cnn* c) {
  for (int l = 0; l < 4; ++l) {
    long len = 27L * layer_cout(c, l) * layer_cin(c, l);
    hipLaunchKernelGGL(k_pack_weights, dim3(unsigned((len + 255) / 256)), dim3(256), 0, c->stream, c->w + c->woff[l],
                       c->wf[l], c->wb[l], layer_cout(c, l), layer_cin(c, l), layer_cout(c, l), layer_cin(c, l));
    CNN_HIP(hipGetLastError());
  }
}
}  // namespace

// ----------------------------------------------- interface used by the solver
namespace cnnlink {
using namespace cnn;

void forward(sc_cnn* c, const void* x, float* y, int batch) {
  if (batch < 1 || batch > c->bmax) throw std::runtime_error("batch out of range");
  const bf16* in = (const bf16*)x;
  for (int l = 0; l < 3; ++l) {
    launch_conv(c, layer_cin(c, l), c->width, HIDDEN, in, c->wf[l], c->w + c->boff[l], c->a[l], c->z[l], nullptr,
                nullptr, batch);
    in = c->a[l];
  }
  launch_conv(c, c->width, c->cout_p, LAST, in, c->wf[3], nullptr, nullptr, nullptr, y, nullptr, batch);
}

void backward(sc_cnn* c, const void* x, const void* dy, float* dx, int batch) {
  if (batch < 1 || batch > c->bmax) throw std::runtime_error("batch out of range");
  const long nv = long(batch) * c->n * c->n * c->n;
  const bf16* dY = (const bf16*)dy;
  // last layer: dW4, then the data gradient into dz[2] with gelu'(z3)
  launch_wgrad(c, c->width, c->cout_p, c->a[2], dY, c->g + c->woff[3], batch);
  launch_conv(c, c->cout_p, c->width, BWD_HIDDEN, dY, c->wb[3], nullptr, c->dz[2], nullptr, nullptr, c->z[2], batch);
  for (int l = 2; l >= 0; --l) {
    const bf16* inl = l == 0 ? (const bf16*)x : c->a[l - 1];
    launch_wgrad(c, layer_cin(c, l), c->width, inl, c->dz[l], c->g + c->woff[l], batch);
    hipLaunchKernelGGL(k_bias_grad<64>, dim3(1024), dim3(256), 0, c->stream, c->dz[l], c->g + c->boff[l], nv);
    CNN_HIP(hipGetLastError());
    if (l > 0)
      launch_conv(c, c->width, c->width, BWD_HIDDEN, c->dz[l], c->wb[l], nullptr, c->dz[l - 1], nullptr, nullptr,
                  c->z[l - 1], batch);
    else if (dx)
      launch_conv(c, c->width, c->cin_p, BWD_INPUT, c->dz[0], c->wb[0], nullptr, nullptr, nullptr, dx, nullptr, batch);
  }
}

int grid_n(const sc_cnn* c) { return c->n; }
int in_channels(const sc_cnn* c) { return c->cin; }
int out_channels(const sc_cnn* c) { return c->cout; }

Bufs buffers(sc_cnn* c) {
  const long nv = long(c->n) * c->n * c->n;
  if (!c->a0) {
    CNN_HIP(hipMalloc(&c->a0, nv * c->cin_p * 2));
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

extern "C" {

int sc_cnn_create(int n, int batch_max, int cin, int width, int cout, uint64_t seed, sc_cnn** out) {
  if (!out) return 1;
  *out = nullptr;
  if (n % VPB != 0 || n > 128 || batch_max < 1 || cin < 1 || cin > 32 || width != 64 || cout < 1 || cout > 32)
    return 1;
  g_cnn_create_err.clear();
  (void)hipGetLastError();
  sc_cnn* c = new sc_cnn();
  c->n = n; c->bmax = batch_max; c->cin = cin; c->width = width; c->cout = cout;
  try {
    CNN_HIP(hipStreamCreate(&c->stream));
    c->own_stream = true;
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
      long len = 27L * layer_cout(c, l) * layer_cin(c, l);
      CNN_HIP(hipMalloc(&c->wf[l], len * 2));
      CNN_HIP(hipMalloc(&c->wb[l], len * 2));
    }
    const long nv = long(batch_max) * n * n * n;
    for (int l = 0; l < 3; ++l) {
      CNN_HIP(hipMalloc(&c->z[l], nv * width * 2));
      CNN_HIP(hipMalloc(&c->a[l], nv * width * 2));
      CNN_HIP(hipMalloc(&c->dz[l], nv * width * 2));
    }
    CNN_HIP(hipMalloc(&c->lossp, ((nv + 255) / 256) * sizeof(double) + 16));
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
  void* ptrs[] = {c->w, c->g, c->m, c->v, c->a0, c->y, c->dyb, c->dx, c->part, c->lossp};
  for (void* p : ptrs) if (p) (void)hipFree(p);
  for (int l = 0; l < 4; ++l) { if (c->wf[l]) (void)hipFree(c->wf[l]); if (c->wb[l]) (void)hipFree(c->wb[l]); }
  for (int l = 0; l < 3; ++l) {
    if (c->z[l]) (void)hipFree(c->z[l]);
    if (c->a[l]) (void)hipFree(c->a[l]);
    if (c->dz[l]) (void)hipFree(c->dz[l]);
  }
  if (c->hostd) (void)hipHostFree(c->hostd);
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

// MSE loss of y (fp32 [batch][n^3][cout_p]) against tgt (fp32 [batch][n^3][cout]);
// writes dy (bf16 [batch][n^3][cout_p]) and returns the loss
int sc_cnn_mse(sc_cnn* c, const void* y, const void* tgt, void* dy, int batch, double* loss) {
  return cguard(c, [&] {
    const long nv = long(batch) * c->n * c->n * c->n;
    const unsigned nb = unsigned((nv + 255) / 256);
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, c->stream, (const float*)y, (const float*)tgt, (bf16*)dy, nullptr,
                       c->lossp, nv, c->cout, c->cout_p, 1.0 / (double(nv) * c->cout));
    CNN_HIP(hipGetLastError());
    std::vector<double> h(nb);
    CNN_HIP(hipMemcpyAsync(h.data(), c->lossp, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CNN_HIP(hipStreamSynchronize(c->stream));
    double s = 0;
    for (double v : h) s += v;
    if (loss) *loss = s / (double(nv) * c->cout);
  });
}

int sc_cnn_adam(sc_cnn* c, double lr, double b1, double b2, double eps) {
  return cguard(c, [&] {
    c->adam_t += 1;
    const float c1 = float(1.0 - std::pow(b1, c->adam_t)), c2 = float(1.0 - std::pow(b2, c->adam_t));
    hipLaunchKernelGGL(k_adam, dim3(unsigned((c->wlen + 255) / 256)), dim3(256), 0, c->stream, c->w, c->g, c->m, c->v,
                       c->wlen, float(lr), float(b1), float(b2), float(eps), c1, c2);
    CNN_HIP(hipGetLastError());
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

int sc_cnn_set_stream(sc_cnn* c, void* s) {
  return cguard(c, [&] {
    CNN_HIP(hipStreamSynchronize(c->stream));
    if (c->own_stream) CNN_HIP(hipStreamDestroy(c->stream));
    c->stream = (hipStream_t)s;
    c->own_stream = false;
  });
}

}  // extern "C"
