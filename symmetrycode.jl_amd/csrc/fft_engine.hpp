// Mixed-radix (2/3/5) Stockham FFT engine for one pencil held by a group of
// threads.  Butterfly data lives in registers; LDS is used only to exchange
// data between stages.  Every size has a compile-time plan, so all butterfly
// indexing is constant-folded and the register arrays never spill.
//
// Stage formulation (Stockham autosort, decimation in time): with Ns the
// product of the radices already applied and R the current radix, butterfly
// i in [0, N/R) reads x[i + r N/R], multiplies by w^(r (i mod Ns)) with
// w = exp(sign 2 pi i / (Ns R)), performs an R-point DFT and writes
// y[(i / Ns) Ns R + (i mod Ns) + r Ns].  After all stages y is in natural
// order.
#pragma once
#include "common.hpp"

namespace sc {

// ------------------------------------------------------------ codelets
// In-place R-point DFT X_k = sum_j x_j w^(jk), w = exp(S 2 pi i / R).
template <int R, int S, typename T> struct Dft;

template <int S, typename T> struct Dft<1, S, T> {
  static DEV void run(C<T>*) {}
};
template <int S, typename T> struct Dft<2, S, T> {
  static DEV void run(C<T>* v) {
    C<T> a = v[0], b = v[1];
    v[0] = add<T>(a, b);
    v[1] = sub<T>(a, b);
  }
};
template <int S, typename T> struct Dft<3, S, T> {
  static DEV void run(C<T>* v) {
    constexpr T h = T(0.5);
    constexpr T s3 = T(0.866025403784438646763723170752936183);
    C<T> a = v[0], b = v[1], c = v[2];
    C<T> t1 = add<T>(b, c);
    C<T> t2 = sub<T>(a, scl<T>(t1, h));
    C<T> t3 = scl<T>(sub<T>(b, c), s3);
    C<T> it3 = muli<S, T>(t3);
    v[0] = add<T>(a, t1);
    v[1] = add<T>(t2, it3);
    v[2] = sub<T>(t2, it3);
  }
};
template <int S, typename T> struct Dft<4, S, T> {
  static DEV void run(C<T>* v) {
    C<T> a = v[0], b = v[1], c = v[2], d = v[3];
    C<T> s0 = add<T>(a, c), d0 = sub<T>(a, c);
    C<T> s1 = add<T>(b, d), d1 = muli<S, T>(sub<T>(b, d));
    v[0] = add<T>(s0, s1);
    v[2] = sub<T>(s0, s1);
    v[1] = add<T>(d0, d1);
    v[3] = sub<T>(d0, d1);
  }
};
template <int S, typename T> struct Dft<5, S, T> {
  static DEV void run(C<T>* v) {
    constexpr T c1 = T(0.309016994374947424102293417182819059);   // cos(2pi/5)
    constexpr T c2 = T(-0.809016994374947424102293417182819059);  // cos(4pi/5)
    constexpr T s1 = T(0.951056516295153572116439333379382143);   // sin(2pi/5)
    constexpr T s2 = T(0.587785252292473129168705954639072769);   // sin(4pi/5)
    C<T> a = v[0];
    C<T> t1 = add<T>(v[1], v[4]), t2 = add<T>(v[2], v[3]);
    C<T> t3 = sub<T>(v[1], v[4]), t4 = sub<T>(v[2], v[3]);
    C<T> r1 = mk<T>(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
    C<T> r2 = mk<T>(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
    C<T> i1 = muli<S, T>(mk<T>(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
    C<T> i2 = muli<S, T>(mk<T>(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
    v[0] = add<T>(a, add<T>(t1, t2));
    v[1] = add<T>(r1, i1);
    v[4] = sub<T>(r1, i1);
    v[2] = add<T>(r2, i2);
    v[3] = sub<T>(r2, i2);
  }
};

// multiply by the constant exp(S 2 pi i m / N), skipping trivial cases
template <int m, int N, int S, typename T> DEV C<T> twc(C<T> a) {
  constexpr int mm = ((m % N) + N) % N;
  if constexpr (mm == 0) return a;
  else if constexpr (4 * mm == N) return muli<S, T>(a);
  else if constexpr (2 * mm == N) return mk<T>(-a.x, -a.y);
  else if constexpr (4 * mm == 3 * N) return muli<-S, T>(a);
  else {
    constexpr T c = T(cx_cos(mm, N));
    constexpr T s = T(S * cx_sin(mm, N));
    return mk<T>(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

// Composite codelet R = P*Q (input n = Q n1 + n2, output k = k1 + P k2).
template <int P, int Q, int S, typename T> struct DftPQ {
  template <int n2> static DEV void col(C<T>* v, C<T>* y) {
    C<T> t[P];
#pragma unroll
    for (int n1 = 0; n1 < P; ++n1) t[n1] = v[Q * n1 + n2];
    Dft<P, S, T>::run(t);
    tw_col<n2, 0>(t, y);
  }
  template <int n2, int k1> static DEV void tw_col(C<T>* t, C<T>* y) {
    if constexpr (k1 < P) {
      y[n2 * P + k1] = twc<n2 * k1, P * Q, S, T>(t[k1]);
      tw_col<n2, k1 + 1>(t, y);
    }
  }
  template <int n2> static DEV void cols(C<T>* v, C<T>* y) {
    if constexpr (n2 < Q) { col<n2>(v, y); cols<n2 + 1>(v, y); }
  }
  static DEV void run(C<T>* v) {
    C<T> y[P * Q];
    cols<0>(v, y);
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
      C<T> t[Q];
#pragma unroll
      for (int n2 = 0; n2 < Q; ++n2) t[n2] = y[n2 * P + k1];
      Dft<Q, S, T>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < Q; ++k2) v[k1 + P * k2] = t[k2];
    }
  }
};
template <int S, typename T> struct Dft<6, S, T> : DftPQ<2, 3, S, T> {};
template <int S, typename T> struct Dft<8, S, T> : DftPQ<2, 4, S, T> {};
template <int S, typename T> struct Dft<9, S, T> : DftPQ<3, 3, S, T> {};
template <int S, typename T> struct Dft<10, S, T> : DftPQ<2, 5, S, T> {};
template <int S, typename T> struct Dft<12, S, T> : DftPQ<3, 4, S, T> {};
template <int S, typename T> struct Dft<15, S, T> : DftPQ<3, 5, S, T> {};
template <int S, typename T> struct Dft<16, S, T> : DftPQ<4, 4, S, T> {};

// ------------------------------------------------------------ plans
// radices and threads-per-pencil (TPP) for each supported length
template <int N> struct Plan;
#ifndef SC_TPP810
#define SC_TPP810 90
#endif
#ifndef SC_TPP405
#define SC_TPP405 45
#endif
#define SC_PLAN(N_, TPP_, ...)                                   \
  template <> struct Plan<N_> {                                  \
    static constexpr int tpp = TPP_;                             \
    static constexpr int r[] = {__VA_ARGS__};                    \
    static constexpr int ns = sizeof(r) / sizeof(int);           \
    static constexpr int rad(int s) { return r[s]; }             \
  };
SC_PLAN(2, 1, 2)
SC_PLAN(3, 1, 3)
SC_PLAN(4, 1, 4)
SC_PLAN(6, 1, 6)
SC_PLAN(8, 1, 8)
SC_PLAN(12, 1, 12)
SC_PLAN(15, 1, 15)
SC_PLAN(16, 4, 4, 4)
SC_PLAN(20, 4, 5, 4)
SC_PLAN(24, 4, 6, 4)
SC_PLAN(30, 6, 5, 6)
SC_PLAN(32, 8, 4, 8)
SC_PLAN(40, 8, 5, 8)
SC_PLAN(48, 8, 6, 8)
SC_PLAN(64, 8, 8, 8)
SC_PLAN(96, 16, 6, 4, 4)
SC_PLAN(128, 16, 8, 4, 4)
SC_PLAN(256, 32, 4, 8, 8)
SC_PLAN(405, SC_TPP405, 5, 9, 9)
SC_PLAN(512, 64, 8, 8, 8)
SC_PLAN(810, SC_TPP810, 10, 9, 9)
#undef SC_PLAN

// the same radices in reverse order: a transform with this plan reads its
// stage-0 input in the layout a Plan<N> transform leaves after its last stage
// when N / (last radix) == TPP (both are positions p + r N / R)
template <int N> struct PlanRev {
  static constexpr int tpp = Plan<N>::tpp;
  static constexpr int ns = Plan<N>::ns;
  static constexpr int rad(int s) { return Plan<N>::r[ns - 1 - s]; }
};

template <int N, int TPP_ = Plan<N>::tpp, class PL = Plan<N>> struct PlanInfo {
  using P = PL;
  static constexpr int TPP = TPP_;
  static constexpr int NS = P::ns;
  static constexpr int R(int s) { return P::rad(s); }
  static constexpr int Ns(int s) { int p = 1; for (int i = 0; i < s; ++i) p *= P::rad(i); return p; }
  static constexpr int NB(int s) { return (N / R(s) + TPP - 1) / TPP; }
  static constexpr int maxE() {
    int m = 0;
    for (int s = 0; s < NS; ++s) m = (NB(s) * R(s) > m) ? NB(s) * R(s) : m;
    return m;
  }
  static constexpr int MAXE = maxE();
  static constexpr int E_LAST = NB(NS - 1) * R(NS - 1);
  static constexpr bool check() { return Ns(NS) == N; }
  static_assert(check(), "plan radices must multiply to N");
};

// Pencil engine: TPP threads (index p) cooperatively transform one pencil of
// length N.  LDS pencil storage is element i at lds[i * LS] (LS = pencil
// interleave stride).
// Barrier between the LDS write and read of an exchange: the whole block, or
// (WAVE) only the calling wavefront when one wavefront owns the pencil.  LDS
// instructions of one wavefront execute in order, so a wavefront-scope fence
// that keeps the compiler from reordering them is sufficient there.
// Twiddles exp(-2 pi i k / N), k < N, from a table of the first N/2
// (w^(k + N/2) = -w^k, N even): half the LDS of a full table.
template <typename T, int N> struct HalfTwiddles {
  const C<T>* t;
  DEV C<T> operator[](int k) const {
    if (k < N / 2) return t[k];
    const C<T> w = t[k - N / 2];
    return mk<T>(-w.x, -w.y);
  }
};

template <bool WAVE> DEV void xsync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

template <typename T, int N, int SGN, int TPP_ = Plan<N>::tpp, bool WAVE = false, class PL = Plan<N>> struct Pencil {
  using PI = PlanInfo<N, TPP_, PL>;
  static constexpr int TPP = PI::TPP;
  static constexpr int NS = PI::NS;
  static constexpr int MAXE = PI::MAXE;
  static constexpr int E0 = PI::NB(0) * PI::R(0);   // stage-0 input slots
  static constexpr int EL = PI::E_LAST;             // last-stage output slots

  // positions read by stage s, slot (m, r)
  template <int s> static DEV int in_pos(int p, int m, int r) {
    constexpr int R = PI::R(s);
    return p + TPP * m + r * (N / R);
  }
  template <int s> static DEV bool valid(int p, int m) {
    constexpr int R = PI::R(s);
    constexpr int NB = PI::NB(s);
    if constexpr (NB * TPP == N / R) return true;
    else return p + TPP * m < N / R;
  }
  // positions written by stage s, slot (m, r)
  template <int s> static DEV int out_pos(int p, int m, int r) {
    constexpr int R = PI::R(s);
    constexpr int Ns = PI::Ns(s);
    int i = p + TPP * m;
    return (i / Ns) * Ns * R + (i % Ns) + r * Ns;
  }

  // load inputs of stage 0 through f(pos) -> C<T>
  template <class F> static DEV void load(C<T>* v, int p, F&& f) {
    constexpr int R = PI::R(0);
    constexpr int NB = PI::NB(0);
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      if (valid<0>(p, m)) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[m * R + r] = f(in_pos<0>(p, m, r));
      }
    }
  }
  // hand the outputs of the last stage to f(pos, value)
  template <class F> static DEV void store(const C<T>* v, int p, F&& f) {
    constexpr int s = NS - 1;
    constexpr int R = PI::R(s);
    constexpr int NB = PI::NB(s);
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      if (valid<s>(p, m)) {
#pragma unroll
        for (int r = 0; r < R; ++r) f(out_pos<s>(p, m, r), v[m * R + r]);
      }
    }
  }
  // the same, for a register array produced by the last stage
  template <class F> static DEV void for_out(int p, F&& f) {
    constexpr int s = NS - 1;
    constexpr int R = PI::R(s);
    constexpr int NB = PI::NB(s);
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      if (valid<s>(p, m)) {
#pragma unroll
        for (int r = 0; r < R; ++r) f(m * R + r, out_pos<s>(p, m, r));
      }
    }
  }

  // tw: the forward twiddles exp(-2 pi i k / N), k < N, as a pointer or an
  // accessor with operator[] (HalfTwiddles)
  template <int s, class TW> static DEV void compute(C<T>* v, int p, const TW& tw) {
    constexpr int R = PI::R(s);
    constexpr int Ns = PI::Ns(s);
    constexpr int NB = PI::NB(s);
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      if (valid<s>(p, m)) {
        C<T>* w = v + m * R;
        if constexpr (Ns > 1) {
          int j = (p + TPP * m) % Ns;
          constexpr int step = N / (Ns * R);
#pragma unroll
          for (int r = 1; r < R; ++r) {
            C<T> t = tw[r * j * step];
            if constexpr (SGN > 0) t.y = -t.y;
            w[r] = cmul<T>(w[r], t);
          }
        }
        Dft<R, SGN, T>::run(w);
      }
    }
  }
  template <int s> static DEV void lds_write(const C<T>* v, int p, C<T>* lds, int LS) {
    constexpr int R = PI::R(s);
    constexpr int NB = PI::NB(s);
#pragma unroll
    for (int m = 0; m < NB; ++m)
      if (valid<s>(p, m)) {
#pragma unroll
        for (int r = 0; r < R; ++r) lds[out_pos<s>(p, m, r) * LS] = v[m * R + r];
      }
  }
  template <int s> static DEV void lds_read(C<T>* v, int p, const C<T>* lds, int LS) {
    constexpr int R = PI::R(s);
    constexpr int NB = PI::NB(s);
#pragma unroll
    for (int m = 0; m < NB; ++m)
      if (valid<s>(p, m)) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[m * R + r] = lds[in_pos<s>(p, m, r) * LS];
      }
  }

  template <int s, class TW> static DEV void run_from(C<T>* v, int p, C<T>* lds, int LS, const TW& tw) {
    compute<s>(v, p, tw);
    if constexpr (s + 1 < NS) {
      xsync<WAVE>();  // previous users of lds are done
      lds_write<s>(v, p, lds, LS);
      xsync<WAVE>();
      lds_read<s + 1>(v, p, lds, LS);
      run_from<s + 1>(v, p, lds, LS, tw);
    }
  }
  // Run all stages.  All threads of the block must call this (barriers).
  template <class TW> static DEV void run(C<T>* v, int p, C<T>* lds, int LS, const TW& tw) {
    run_from<0>(v, p, lds, LS, tw);
  }

  // Move the final-stage outputs (natural order) to the stage-0 input
  // layout of another transform of the same length, through LDS.
  static DEV void out_to_in(C<T>* v, int p, C<T>* lds, int LS) {
    xsync<WAVE>();
    lds_write<NS - 1>(v, p, lds, LS);
    xsync<WAVE>();
    lds_read<0>(v, p, lds, LS);
  }
  // Write final-stage outputs into LDS in natural order (caller syncs after)
  static DEV void out_to_lds(const C<T>* v, int p, C<T>* lds, int LS) {
    xsync<WAVE>();
    lds_write<NS - 1>(v, p, lds, LS);
  }
};

}  // namespace sc
