// LDS bank-conflict probe for gfx950 (not part of the library).  One
// wavefront repeats a ds_read_b128 at the per-lane byte offset
//   off(lane) = gbase[lane / 16] + lbase + (lane % 16) * stride
// for the patterns listed in main(), one launch each, so that the
// SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE counters of each dispatch give the
// conflict rate of one pattern.  Build and run (from the repository root):
//   hipcc --offload-arch=gfx950 -O3 -Xclang -target-feature -Xclang -packed-fp32-ops -o lds_bank_probe symmetrycode.jl_amd/probes/lds_bank_probe.hip
//   rocprofv3 --pmc SQ_LDS_BANK_CONFLICT SQ_LDS_IDX_ACTIVE -d <dir> -o run -f csv -- ./lds_bank_probe
// Every offset stays inside the 32 KB LDS array (masked to 16 KB, plus a
// second copy 16 KB above it).
#include <hip/hip_runtime.h>
#include <cstdio>
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(64) k_probe(unsigned* out, int g0, int g1, int g2, int g3, int lbase, int stride, int reps) {
  __shared__ __attribute__((aligned(1024))) unsigned char lds[32768];
  for (int i = threadIdx.x; i < 32768 / 4; i += 64) reinterpret_cast<unsigned*>(lds)[i] = i;
  __syncthreads();
  const int li = threadIdx.x & 15, lk = threadIdx.x >> 4;
  const int gb = lk == 0 ? g0 : lk == 1 ? g1 : lk == 2 ? g2 : g3;
  int off = gb + lbase + li * stride;
  off &= 16383 & ~15;
  u32x4 acc = {0, 0, 0, 0};
  for (int r = 0; r < reps; ++r) {
    u32x4 v = *reinterpret_cast<const volatile u32x4*>(lds + off);
    acc ^= v;
    off ^= (r & 1) << 14;  // alternate between two copies 16 KB apart (same bank pattern)
  }
  out[threadIdx.x] = acc[0] ^ acc[1] ^ acc[2] ^ acc[3];
}
int main() {
  unsigned* d;
  hipMalloc(&d, 64 * 4);
  struct P { const char* name; int g[4]; int lbase, stride; };
  P ps[] = {
      {"aligned256_groups1024", {0, 1024, 2048, 3072}, 0, 16},
      {"mis16_groups1024", {0, 1024, 2048, 3072}, 16, 16},
      {"mis32_groups1024", {0, 1024, 2048, 3072}, 32, 16},
      {"mis128_groups1024", {0, 1024, 2048, 3072}, 128, 16},
      {"stride144", {0, 2304, 4608, 6912}, 0, 144},
      {"groups_1168apart", {0, 1168, 2336, 3504}, 0, 16},
      {"groups_1168apart_mis16", {0, 1168, 2336, 3504}, 16, 16},
      {"groups_1168apart_mis32", {0, 1168, 2336, 3504}, 32, 16},
      {"groups_same_block", {0, 0, 0, 0}, 0, 16},
      {"groups_256apart_aligned", {0, 256, 512, 768}, 0, 16},
      {"groups_272apart", {0, 272, 544, 816}, 0, 16},
  };
  for (auto& p : ps) {
    printf("probe %s\n", p.name);
    fflush(stdout);
    hipLaunchKernelGGL(k_probe, dim3(1), dim3(64), 0, 0, d, p.g[0], p.g[1], p.g[2], p.g[3], p.lbase, p.stride, 4096);
    hipDeviceSynchronize();
  }
  unsigned h[64];
  hipMemcpy(h, d, sizeof h, hipMemcpyDeviceToHost);
  printf("done %u\n", h[0]);
  return 0;
}
