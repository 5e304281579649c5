// Optimiser kernels shared by the learned closures: per-array ClipNorm and
// AdamW on a flat FP32 parameter vector of at most eight arrays (the rules of
// OptimiserChain(ClipNorm, AdamW), src/training.jl:50, Optimisers.jl 0.4).
// Internal linkage: every source that includes this gets its own copy.
#pragma once
#include <hip/hip_runtime.h>

namespace optim {
namespace {

// Boundaries of the n (at most MAX_ARRAYS) parameter arrays in the flat
// parameter vector: array a is [b[a], b[a + 1])
constexpr int MAX_ARRAYS = 8;
struct Arrays {
  long b[MAX_ARRAYS + 1];
  int n;
};

// norm workspace (doubles): partials [MAX_ARRAYS][NORM_BLOCKS], then the
// squared norms at NORM_OFFSET
constexpr int NORM_BLOCKS = 32;
constexpr int NORM_OFFSET = MAX_ARRAYS * NORM_BLOCKS;
constexpr int NORM_DOUBLES = NORM_OFFSET + MAX_ARRAYS;

// squared L2 norm of each parameter array's gradient: block (j, a) sums a
// strided share of array a into part[a][j]; k_finish_norms adds the shares
__global__ void k_array_sumsq(const float* g, Arrays A, double* part) {
  __shared__ double sm[256];
  const int a = blockIdx.y;
  double acc = 0;
  for (long i = A.b[a] + long(blockIdx.x) * blockDim.x + threadIdx.x; i < A.b[a + 1];
       i += long(gridDim.x) * blockDim.x) {
    const double e = g[i];
    acc += e * e;
  }
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] += sm[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[a * NORM_BLOCKS + blockIdx.x] = sm[0];
}
__global__ void k_finish_norms(const double* part, double* nrm2, int n) {
  const int a = threadIdx.x;
  if (a >= n) return;
  double s = 0;
  for (int j = 0; j < NORM_BLOCKS; ++j) s += part[a * NORM_BLOCKS + j];
  nrm2[a] = s;
}

// ClipNorm(clip) per parameter array (gradient scaled by min(1, clip / |g_a|);
// clip <= 0 disables it) followed by AdamW, its weight decay scaled by the
// learning rate: w -= lr (mhat / (sqrt(vhat) + eps) + wd w)
__global__ void k_adamw(float* w, const float* g, float* m, float* v, long len, float lr, float b1, float b2,
                        float eps, float wd, float c1, float c2, const double* nrm2, Arrays A, double clip) {
  long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= len) return;
  float gi = g[i];
  if (clip > 0) {
    int a = 0;
    while (a < A.n - 1 && i >= A.b[a + 1]) ++a;
    const double nrm = sqrt(nrm2[a]);
    if (nrm > clip) gi *= float(clip / nrm);
  }
  const float wi = w[i];
  float mi = b1 * m[i] + (1 - b1) * gi;
  float vi = b2 * v[i] + (1 - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] = wi - lr * ((mi / c1) / (sqrtf(vi / c2) + eps) + wd * wi);
}


}  // namespace
}  // namespace optim
