// Shared device/host helpers for the symcode HIP library (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

#define DEV __device__ __forceinline__
#define HD __host__ __device__ __forceinline__

namespace sc {

template <typename T> struct cvec;
template <> struct cvec<float> { using type = float2; };
template <> struct cvec<double> { using type = double2; };
template <typename T> using C = typename cvec<T>::type;

template <typename T> HD C<T> mk(T re, T im) { C<T> r; r.x = re; r.y = im; return r; }
template <typename T> HD C<T> add(C<T> a, C<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> HD C<T> sub(C<T> a, C<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> HD C<T> scl(C<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
template <typename T> HD C<T> cmul(C<T> a, C<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> HD C<T> conjc(C<T> a) { return mk<T>(a.x, -a.y); }
// multiply by +i (s=+1) or -i (s=-1)
template <int S, typename T> HD C<T> muli(C<T> a) {
  if constexpr (S > 0) return mk<T>(-a.y, a.x);
  else return mk<T>(a.y, -a.x);
}

// Integer frequency of 0-based index j on an axis of length n: j < n/2 -> j,
// else j - n (the Nyquist index n/2 maps to -n/2).
HD int freq(int j, int n) { return j < ((n + 1) >> 1) ? j : j - n; }

// ------------------------------------------------------------------
// Compile-time cos/sin(2*pi*m/N) for codelet twiddles (Taylor series after
// reduction to [0, pi/4]); accurate to ~1 ulp for the small N used here.
constexpr double cx_pi = 3.14159265358979323846264338327950288;
constexpr double cx_taylor_sin(double x) {
  double t = x, s = x;
  for (int k = 1; k < 14; ++k) { t *= -x * x / ((2 * k) * (2 * k + 1)); s += t; }
  return s;
}
constexpr double cx_taylor_cos(double x) {
  double t = 1, s = 1;
  for (int k = 1; k < 14; ++k) { t *= -x * x / ((2 * k - 1) * (2 * k)); s += t; }
  return s;
}
// angle = 2*pi*m/N, exact octant reduction on the rational m/N
constexpr void cx_sincos(long m, long N, double& s, double& c) {
  m %= N; if (m < 0) m += N;
  // work in units of N/8 : a = 8m/N in [0, 8)
  long m8 = 8 * m;
  long oct = m8 / N;  // octant 0..7
  long rem = m8 - oct * N;  // remainder, angle within octant = (pi/4)*rem/N
  double th = (cx_pi / 4) * double(rem) / double(N);
  double sr = cx_taylor_sin(th), cr = cx_taylor_cos(th);
  double thc = (cx_pi / 4) * double(N - rem) / double(N);
  double sc_ = cx_taylor_sin(thc), cc = cx_taylor_cos(thc);
  switch (oct) {
    case 0: s = sr; c = cr; break;
    case 1: s = cc; c = sc_; break;  // pi/2 - thc
    case 2: s = cr; c = -sr; break;
    case 3: s = sc_; c = -cc; break;
    case 4: s = -sr; c = -cr; break;
    case 5: s = -cc; c = -sc_; break;
    case 6: s = -cr; c = sr; break;
    default: s = -sc_; c = cc; break;
  }
}
constexpr double cx_cos(long m, long N) { double s = 0, c = 0; cx_sincos(m, N, s, c); return c; }
constexpr double cx_sin(long m, long N) { double s = 0, c = 0; cx_sincos(m, N, s, c); return s; }

}  // namespace sc
