// Register-resident mixed-radix Stockham FFT for gfx950 (see LinePlan below): a team of
// T threads per line, E elements per thread, butterflies in registers and LDS only for the
// exchanges between stages. Zero inputs and unused outputs known at compile time are
// pruned by constant propagation (the register DFTs are fully unrolled).
//
// Register DFTs are recursive decimation-in-time codelets; their twiddles are constants
// computed at compile time (exact octant reduction plus Taylor series in double). The
// runtime stage twiddles come from a table tw[m] = exp(-2 pi i m / L) built in FP64 on the
// host.
#pragma once

#include <type_traits>

#include "common.h"

namespace sopht {
namespace rfft {

// ------------------------------------------------------------------ compile-time loops
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// ------------------------------------------------------------------ constexpr twiddles
struct CD {
  double re, im;
};

constexpr double taylor_sin(double x) {  // |x| <= pi/4
  double x2 = x * x, term = x, sum = x;
  for (int k = 1; k < 14; ++k) {
    term *= -x2 / double((2 * k) * (2 * k + 1));
    sum += term;
  }
  return sum;
}

constexpr double taylor_cos(double x) {  // |x| <= pi/4
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int k = 1; k < 14; ++k) {
    term *= -x2 / double((2 * k - 1) * (2 * k));
    sum += term;
  }
  return sum;
}

// exp(-2 pi i m / n), reduced exactly to an angle phi <= pi/4 plus a quarter-turn
constexpr CD twiddle(long m, long n) {
  m %= n;
  if (m < 0) m += n;
  const long q = (4 * m) / n, r = (4 * m) % n;
  const double half_pi = 1.57079632679489661923132169163975144;
  double c = 0, s = 0;  // cos, sin of the in-quadrant angle
  if (2 * r <= n) {
    const double phi = half_pi * (double(r) / double(n));
    c = taylor_cos(phi);
    s = taylor_sin(phi);
  } else {
    const double phi = half_pi * (double(n - r) / double(n));
    c = taylor_sin(phi);
    s = taylor_cos(phi);
  }
  double ct = c, st = s;
  switch (q & 3) {
    case 1: ct = -s; st = c; break;
    case 2: ct = -c; st = -s; break;
    case 3: ct = s; st = -c; break;
    default: break;
  }
  return CD{ct, -st};
}

// x * w_N^E (forward) or x * conj(w_N^E) (inverse); exact special cases avoid work
template <typename T, int N, int E, bool INV>
__device__ __forceinline__ Cplx<T> twmul(Cplx<T> x) {
  constexpr int e = ((E % N) + N) % N;
  if constexpr (e == 0) {
    return x;
  } else if constexpr (4 * e == N) {  // w = -i
    return INV ? Cplx<T>{-x.im, x.re} : Cplx<T>{x.im, -x.re};
  } else if constexpr (2 * e == N) {  // w = -1
    return Cplx<T>{-x.re, -x.im};
  } else if constexpr (4 * e == 3 * N) {  // w = +i
    return INV ? Cplx<T>{x.im, -x.re} : Cplx<T>{-x.im, x.re};
  } else {
    constexpr CD w = twiddle(e, N);
    const T wr = T(w.re), wi = T(INV ? -w.im : w.im);
    return Cplx<T>{x.re * wr - x.im * wi, x.re * wi + x.im * wr};
  }
}

// ------------------------------------------------------------------ register codelets
template <typename T, bool INV>
__device__ __forceinline__ void dft2(Cplx<T>* v) {
  const Cplx<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft3(Cplx<T>* v) {
  const T h = T(0.86602540378443864676372317075294);  // sqrt(3)/2
  const Cplx<T> t1 = cadd(v[1], v[2]);
  const Cplx<T> t2 = {v[0].re - T(0.5) * t1.re, v[0].im - T(0.5) * t1.im};
  const Cplx<T> d = csub(v[1], v[2]);
  const Cplx<T> t3 = mul_mi<T, INV>(Cplx<T>{h * d.re, h * d.im});
  v[0] = cadd(v[0], t1);
  v[1] = cadd(t2, t3);
  v[2] = csub(t2, t3);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft4(Cplx<T>* v) {
  const Cplx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  const Cplx<T> t2 = cadd(v[1], v[3]), t3 = mul_mi<T, INV>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

// DFT of length N in registers (unnormalised; INV uses the conjugate roots)
template <typename T, int N, bool INV>
__device__ __forceinline__ void rdft(Cplx<T>* v) {
  if constexpr (N == 1) {
  } else if constexpr (N == 2) {
    dft2<T, INV>(v);
  } else if constexpr (N == 3) {
    dft3<T, INV>(v);
  } else if constexpr (N == 4) {
    dft4<T, INV>(v);
  } else {
    // decimation in time: N = R * M, X[k + M q] = sum_r w_N^{r k} w_R^{r q} S_r[k]
    constexpr int R = (N % 4 == 0) ? 4 : ((N % 2 == 0) ? 2 : 3);
    constexpr int M = N / R;
    Cplx<T> s[R][M];
    static_for<0, R>([&](auto r) {
      static_for<0, M>([&](auto m) { s[r][m] = v[decltype(r)::value + R * decltype(m)::value]; });
    });
    static_for<0, R>([&](auto r) { rdft<T, M, INV>(s[decltype(r)::value]); });
    static_for<0, M>([&](auto k) {
      constexpr int kk = decltype(k)::value;
      Cplx<T> t[R];
      static_for<0, R>([&](auto r) {
        constexpr int rr = decltype(r)::value;
        t[rr] = twmul<T, N, rr * kk, INV>(s[rr][kk]);
      });
      if constexpr (R == 2) dft2<T, INV>(t);
      else if constexpr (R == 3) dft3<T, INV>(t);
      else dft4<T, INV>(t);
      static_for<0, R>([&](auto q) { v[kk + M * decltype(q)::value] = t[decltype(q)::value]; });
    });
  }
}

// DFT of length N whose inputs v[N/2 .. N) are zero: X[2m] = DFT_{N/2}(v)[m] and
// X[2m+1] = DFT_{N/2}(v[r] w_N^r)[m]. (Constant propagation cannot drop the zeros itself:
// x + 0.0 is not x for x = -0.0.)
template <typename T, int N, bool INV>
__device__ __forceinline__ void rdft_half_in(Cplx<T>* v) {
  static_assert(N % 2 == 0, "half-zero codelet needs an even length");
  constexpr int M = N / 2;
  Cplx<T> a[M], b[M];
  static_for<0, M>([&](auto m_) {
    constexpr int m = decltype(m_)::value;
    a[m] = v[m];
    b[m] = twmul<T, N, m, INV>(v[m]);
  });
  rdft<T, M, INV>(a);
  rdft<T, M, INV>(b);
  static_for<0, M>([&](auto m_) {
    constexpr int m = decltype(m_)::value;
    v[2 * m] = a[m];
    v[2 * m + 1] = b[m];
  });
}

// ------------------------------------------------------------------ line plans
// A line of length L is transformed by a team of T = L / E threads; thread t holds the E
// elements x[t + T e] (e < E). Stage s of the Stockham sequence has radix R_s, with R_s | E:
// each thread computes E / R_s butterflies j = t + p T, reading j + r L/R_s and writing
// (j / Ns) Ns R_s + j % Ns + r Ns (Ns = product of the earlier radices), with twiddles
// w_{Ns R_s}^{r (j % Ns)}. The first stage's inputs and the last stage's outputs are both
// x[t + T e] held by thread t, so a forward transform, a pointwise multiply and an
// inverse transform need no reordering, and the zero-padded half of a line (t + T e >= L/2,
// i.e. e >= E/2) is known at compile time. Between stages the elements are exchanged
// through LDS (one write, then one read after a wavefront-scope sync). T divides 64.
template <int... Rs>
struct Radices {};

template <int L>
struct LinePlan;
#define SOPHT_LINE_PLAN(L_, E_, ...)                                        \
  template <>                                                               \
  struct LinePlan<L_> {                                                     \
    static constexpr int E = E_, T = L_ / E_;                               \
    using R = Radices<__VA_ARGS__>;                                         \
  };
SOPHT_LINE_PLAN(8, 8, 8)
SOPHT_LINE_PLAN(12, 12, 12)
SOPHT_LINE_PLAN(16, 8, 8, 2)
SOPHT_LINE_PLAN(24, 12, 2, 12)
SOPHT_LINE_PLAN(32, 8, 8, 4)
SOPHT_LINE_PLAN(48, 12, 4, 12)
SOPHT_LINE_PLAN(64, 8, 8, 8)
SOPHT_LINE_PLAN(96, 12, 4, 4, 6)
SOPHT_LINE_PLAN(128, 8, 8, 4, 4)
SOPHT_LINE_PLAN(192, 12, 4, 4, 12)
SOPHT_LINE_PLAN(256, 8, 8, 8, 4)
SOPHT_LINE_PLAN(384, 12, 4, 4, 2, 12)
#ifndef SOPHT_PLAN512
SOPHT_LINE_PLAN(512, 8, 8, 8, 8)
#elif SOPHT_PLAN512 == 16
SOPHT_LINE_PLAN(512, 16, 16, 16, 2)
#else
SOPHT_LINE_PLAN(512, 32, 32, 16)
#endif
SOPHT_LINE_PLAN(768, 12, 4, 4, 4, 12)
SOPHT_LINE_PLAN(1024, 16, 16, 8, 8)
#undef SOPHT_LINE_PLAN

constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v / 2); }

#ifndef SOPHT_TWIDDLE_PRODUCTS
#define SOPHT_TWIDDLE_PRODUCTS 1
#endif
constexpr bool kTwiddleProducts = SOPHT_TWIDDLE_PRODUCTS;

// LDS layout of the exchange after a stage of radix R and write stride NS: element p of the
// team's line lives at p ^ (((p >> A) & (2^MB - 1)) << B). The stage writes
// (j / NS) NS R + j % NS + r NS (stride R or NS between lanes) and the next stage reads
// t + T e (unit stride); the XOR moves the lane bits that would collide onto the bank bits the
// other pattern leaves free, so on gfx950 (64-bit accesses served 32 lanes at a time by 64
// four-byte banks) both patterns are conflict-free for every plan from 128 points on
// (checked exhaustively on the host). The map is an involution on [0, L). Shorter lines keep
// the identity.
template <int L, int NS, int R>
struct Swizzle {
  static constexpr int T = LinePlan<L>::T;
  static constexpr bool kOn = L >= 128;
  static constexpr int A = T >= 32 ? 5 : 4;
  static constexpr int MB = NS == 1 ? ilog2(R) : (ilog2(R) < 3 ? ilog2(R) : 3);
  static constexpr int B = NS == 1 ? 0 : 2;
  static_assert(!kOn || (1 << ilog2(R)) == R, "swizzled exchanges need a power-of-two radix");
  static_assert(!kOn || (B + MB <= A && L % (1 << (A + MB)) == 0), "swizzle must stay inside the line");
  static __device__ __forceinline__ int map(int p) {
    if constexpr (kOn) return p ^ (((p >> A) & ((1 << MB) - 1)) << B);
    else return p;
  }
};

// LDS elements reserved per line. Teams of 16 threads share a 32-lane half-wave, so their
// regions start 16 slots apart (mod 32). Column tiles stage WC = 8 columns with lanes
// (column, row); a stride of 4 (mod 32) between column regions keeps those accesses
// conflict-free too.
template <int L, bool COLUMN>
constexpr int region_stride() {
  if (LinePlan<L>::T < 32) return L + 16;
  return COLUMN ? L + 4 : L;
}

// All lanes of a team sit in one wavefront (T divides 64), so the exchanges need only
// wavefront-scope ordering: LDS operations of one wavefront execute in order, and the
// fences keep the compiler from moving accesses across the exchange point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Twiddle tables copied to LDS are stored padded, element m at m + m / 32: a stage reads the
// table at stride r L / (NS R) across its lanes, and without the padding those reads fall on
// 2 to 8 lanes per bank (the stage with NS = 8 of a 512-point line); with it, on one for
// every plan of 128, 256, 512 and 1024 points (two for 384 and 768).
__host__ __device__ constexpr int tw_slot(int m) { return m + (m >> 5); }
template <int L>
constexpr int tw_lds_size() { return L + L / 32 + 1; }

// HALF_IN (first stage only): the line's upper half x[L/2 .. L) is zero, i.e. the inputs
// r >= R/2 of every butterfly. PADTW: tw is a padded LDS table (tw_slot).
template <typename T, int L, int R, int NS, bool INV, bool HALF_IN, bool PADTW = false>
__device__ __forceinline__ void stockham_stage(Cplx<T>* v, Cplx<T>* xb, const Cplx<T>* tw,
                                               int t) {
  constexpr int E = LinePlan<L>::E, TT = LinePlan<L>::T;
  constexpr int PER = E / R;
  static_assert(E % R == 0, "radix must divide the elements per thread");
  // butterfly p takes v[p + PER r] (= x[j + r L/R], j = t + p T)
  static_for<0, PER>([&](auto p_) {
    constexpr int p = decltype(p_)::value;
    Cplx<T> u[R];
    static_for<0, R>([&](auto r_) { u[decltype(r_)::value] = v[p + PER * decltype(r_)::value]; });
    if constexpr (NS > 1) {
      // w^{r k}, w = exp(-2 pi i / (NS R)): the powers r = 2^b come from the table, the others
      // are products of two of them (fewer vector-memory instructions per butterfly)
      const int k = (t + p * TT) % NS;
      constexpr int step = L / (NS * R);
      Cplx<T> wr[R];
      static_for<1, R>([&](auto r_) {
        constexpr int r = decltype(r_)::value;
        constexpr int hb = kTwiddleProducts ? 1 << ilog2(r) : r;
        if constexpr (hb == r) wr[r] = tw[PADTW ? tw_slot(r * k * step) : r * k * step];
        else wr[r] = cmul(wr[hb], wr[r - hb]);
      });
      static_for<1, R>([&](auto r_) {
        constexpr int r = decltype(r_)::value;
        u[r] = INV ? cmulc(u[r], wr[r]) : cmul(u[r], wr[r]);
      });
    }
    if constexpr (HALF_IN) rdft_half_in<T, R, INV>(u);
    else rdft<T, R, INV>(u);
    static_for<0, R>([&](auto r_) { v[p + PER * decltype(r_)::value] = u[decltype(r_)::value]; });
  });
  if constexpr (NS * R < L) {
    // exchange: outputs of butterfly j go to (j/NS) NS R + j%NS + r NS; the next stage reads
    // x[t + T e] again
    static_for<0, PER>([&](auto p_) {
      constexpr int p = decltype(p_)::value;
      const int j = t + p * TT;
      const int base = (j / NS) * NS * R + j % NS;
      static_for<0, R>([&](auto r_) {
        constexpr int r = decltype(r_)::value;
        xb[Swizzle<L, NS, R>::map(base + r * NS)] = v[p + PER * r];
      });
    });
    wave_sync();
    static_for<0, E>([&](auto e_) {
      constexpr int e = decltype(e_)::value;
      v[e] = xb[Swizzle<L, NS, R>::map(t + TT * e)];
    });
    wave_sync();
  }
}

// Twiddles of butterfly input r of a stage with write stride NS and radix R, for butterfly k of
// its group (k = j % NS): w^{r k} with w = exp(-2 pi i / (NS R)), computed as stockham_stage
// does (r = 2^b from the table, the other powers as products of two of them).
template <typename T, int L, int R, int NS>
__device__ __forceinline__ void stage_twiddles(Cplx<T>* wr, const Cplx<T>* tw, int k) {
  constexpr int step = L / (NS * R);
  static_for<1, R>([&](auto r_) {
    constexpr int r = decltype(r_)::value;
    constexpr int hb = kTwiddleProducts ? 1 << ilog2(r) : r;
    if constexpr (hb == r) wr[r] = tw[r * k * step];
    else wr[r] = cmul(wr[hb], wr[r - hb]);
  });
}

// A stage of radix R and write stride NS whose twiddles are given: wr[p][r] for butterfly p of
// the thread (j = t + p T, from stage_twiddles with k = j % NS). Without an exchange after it
// (NS R = L, the last stage) the outputs stay in v; otherwise the exchange writes and the
// read-back x[t + T e] for the next stage follow, as in stockham_stage.
template <typename T, int L, int R, int NS, bool INV>
__device__ __forceinline__ void stockham_stage_given(Cplx<T>* v, Cplx<T>* xb,
                                                     const Cplx<T> (*wr)[R], int t) {
  constexpr int E = LinePlan<L>::E, TT = LinePlan<L>::T;
  constexpr int PER = E / R;
  static_assert(E % R == 0 && NS > 1, "a stage with twiddles");
  static_for<0, PER>([&](auto p_) {
    constexpr int p = decltype(p_)::value;
    Cplx<T> u[R];
    static_for<0, R>([&](auto r_) { u[decltype(r_)::value] = v[p + PER * decltype(r_)::value]; });
    static_for<1, R>([&](auto r_) {
      constexpr int r = decltype(r_)::value;
      u[r] = INV ? cmulc(u[r], wr[p][r]) : cmul(u[r], wr[p][r]);
    });
    rdft<T, R, INV>(u);
    static_for<0, R>([&](auto r_) { v[p + PER * decltype(r_)::value] = u[decltype(r_)::value]; });
  });
  if constexpr (NS * R < L) {
    static_for<0, PER>([&](auto p_) {
      constexpr int p = decltype(p_)::value;
      const int j = t + p * TT;
      const int base = (j / NS) * NS * R + j % NS;
      static_for<0, R>([&](auto r_) {
        constexpr int r = decltype(r_)::value;
        xb[Swizzle<L, NS, R>::map(base + r * NS)] = v[p + PER * r];
      });
    });
    wave_sync();
    static_for<0, E>([&](auto e_) {
      constexpr int e = decltype(e_)::value;
      v[e] = xb[Swizzle<L, NS, R>::map(t + TT * e)];
    });
    wave_sync();
  }
}

template <typename T, int L, bool INV, bool HALF_IN, bool PADTW, int NS, int R0, int... Rest>
__device__ __forceinline__ void stockham_run(Cplx<T>* v, Cplx<T>* xb, const Cplx<T>* tw, int t) {
  stockham_stage<T, L, R0, NS, INV, HALF_IN && NS == 1, PADTW>(v, xb, tw, t);
  if constexpr (sizeof...(Rest) > 0)
    stockham_run<T, L, INV, HALF_IN, PADTW, NS * R0, Rest...>(v, xb, tw, t);
}

template <typename T, int L, bool INV, bool HALF_IN, bool PADTW, int... Rs>
__device__ __forceinline__ void stockham_plan(Cplx<T>* v, Cplx<T>* xb, const Cplx<T>* tw, int t,
                                              Radices<Rs...>) {
  stockham_run<T, L, INV, HALF_IN, PADTW, 1, Rs...>(v, xb, tw, t);
}

// In-place transform of a line held as v[e] = x[t + T e] by thread t of its team; on return
// v[e] = X[t + T e] (forward: exp(-), inverse: exp(+), unnormalised). `region` (L elements
// of LDS) belongs to the team alone; its contents are undefined on return. HALF_IN: the
// caller guarantees x[L/2 .. L) = 0 (v[e] = 0 for e >= E/2), which prunes the first stage.
// PADTW: tw is a padded LDS table (tw_slot).
template <typename T, int L, bool INV, bool HALF_IN = false, bool PADTW = false>
__device__ __forceinline__ void team_fft(Cplx<T>* v, Cplx<T>* region, const Cplx<T>* tw, int t) {
  static_assert(64 % LinePlan<L>::T == 0, "a team must sit inside one wavefront");
  stockham_plan<T, L, INV, HALF_IN, PADTW>(v, region, tw, t, typename LinePlan<L>::R{});
}

// Copy the twiddle table tw[0, L) into the padded LDS table twl (threads i0, i0 + step, ...).
template <typename T, int L>
__device__ __forceinline__ void load_tw_lds(Cplx<T>* twl, const Cplx<T>* tw, int i0, int step) {
  for (int i = i0; i < L; i += step) twl[tw_slot(i)] = tw[i];
}

}  // namespace rfft
}  // namespace sopht
