// Mixed-radix (2, 3, 4, 8) Stockham FFT on lines held in LDS, for gfx950.
//
// A "line" of length L lives in LDS at buf[e * pitch] (e = 0..L-1). A team of TL threads
// (t = 0..TL-1) transforms one line; a workgroup holds several lines side by side and every
// thread of the workgroup calls fft_lds with uniform control flow (the stages contain
// workgroup barriers). Each stage is done in place: a thread loads all butterflies it owns
// into registers, the workgroup synchronises, then every thread stores its results.
//
// Stockham autosort: stage s with radix R and Ns = product of earlier radices maps
// butterfly j (0 <= j < L/R) to inputs j + r L/R and outputs (j/Ns) Ns R + (j%Ns) + r Ns,
// with twiddles w_{Ns R}^{r (j%Ns)}; input and output are both in natural order.
// Twiddles come from an LDS table tw[m] = exp(-2 pi i m / L), m < L, built in FP64 on the
// host; the inverse transform uses their conjugates and is unnormalised.
#pragma once

#include "common.h"

namespace sopht {
namespace fft {

// ------------------------------------------------------------------ register codelets
template <typename T, bool INV>
__device__ __forceinline__ void dft2(Cplx<T>* v) {
  const Cplx<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft3(Cplx<T>* v) {
  // X1 = x0 - (x1+x2)/2 -+ i (sqrt3/2)(x1-x2), X2 = conjugate pairing
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

template <typename T, bool INV>
__device__ __forceinline__ void dft8(Cplx<T>* v) {
  // radix-2 DIT over two radix-4 halves
  Cplx<T> e[4] = {v[0], v[2], v[4], v[6]};
  Cplx<T> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<T, INV>(e);
  dft4<T, INV>(o);
  const T r = T(0.70710678118654752440084436210485);  // 1/sqrt(2)
  // o[k] *= w8^{+-k}
  Cplx<T> o1, o2, o3;
  if (!INV) {
    o1 = {r * (o[1].re + o[1].im), r * (o[1].im - o[1].re)};
    o2 = {o[2].im, -o[2].re};
    o3 = {r * (o[3].im - o[3].re), -r * (o[3].re + o[3].im)};
  } else {
    o1 = {r * (o[1].re - o[1].im), r * (o[1].re + o[1].im)};
    o2 = {-o[2].im, o[2].re};
    o3 = {-r * (o[3].re + o[3].im), r * (o[3].re - o[3].im)};
  }
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void dft(Cplx<T>* v) {
  if constexpr (R == 2) dft2<T, INV>(v);
  else if constexpr (R == 3) dft3<T, INV>(v);
  else if constexpr (R == 4) dft4<T, INV>(v);
  else if constexpr (R == 8) dft8<T, INV>(v);
  else static_assert(R == 2, "unsupported radix");
}

// ------------------------------------------------------------------ compile-time plans
template <int... Rs>
struct Radices {};

template <int L>
struct Plan;  // radices in stage order
template <> struct Plan<8> { using R = Radices<8>; };
template <> struct Plan<12> { using R = Radices<4, 3>; };
template <> struct Plan<16> { using R = Radices<4, 4>; };
template <> struct Plan<24> { using R = Radices<8, 3>; };
template <> struct Plan<32> { using R = Radices<4, 8>; };
template <> struct Plan<48> { using R = Radices<4, 4, 3>; };
template <> struct Plan<64> { using R = Radices<8, 8>; };
template <> struct Plan<96> { using R = Radices<8, 4, 3>; };
template <> struct Plan<128> { using R = Radices<8, 4, 4>; };
template <> struct Plan<192> { using R = Radices<8, 8, 3>; };
template <> struct Plan<256> { using R = Radices<8, 8, 4>; };
template <> struct Plan<384> { using R = Radices<8, 4, 4, 3>; };
template <> struct Plan<512> { using R = Radices<8, 8, 8>; };
template <> struct Plan<768> { using R = Radices<8, 8, 4, 3>; };
template <> struct Plan<1024> { using R = Radices<8, 8, 4, 4>; };

constexpr int cgcd(int a, int b) { return b == 0 ? a : cgcd(b, a % b); }

template <int L, typename RS>
struct TeamSize;
template <int L, int... Rs>
struct TeamSize<L, Radices<Rs...>> {
  static constexpr int g = [] {
    int r = 0;
    ((r = cgcd(r, L / Rs)), ...);
    return r;
  }();
};

// Threads per line: the largest power of two dividing every stage's butterfly count,
// capped at 64 (one wavefront per line at most).
template <int L>
constexpr int team_size() {
  int g = TeamSize<L, typename Plan<L>::R>::g;
  int t = 1;
  while (t * 2 <= 64 && g % (t * 2) == 0) t *= 2;
  return t;
}

template <int L>
constexpr bool supported_length() {
  return L == 8 || L == 12 || L == 16 || L == 24 || L == 32 || L == 48 || L == 64 || L == 96 ||
         L == 128 || L == 192 || L == 256 || L == 384 || L == 512 || L == 768 || L == 1024;
}

// One Stockham stage.
template <typename T, int L, int R, int NS, int TL, bool INV>
__device__ __forceinline__ void stage(Cplx<T>* buf, int pitch, int t, const Cplx<T>* tw) {
  constexpr int NB = L / R;       // butterflies in the line
  constexpr int PER = NB / TL;    // butterflies per thread
  static_assert(NB % TL == 0, "team size must divide the butterfly count");
  Cplx<T> v[PER][R];
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int j = t + p * TL;
    const int k = j % NS;
#pragma unroll
    for (int r = 0; r < R; ++r) v[p][r] = buf[(j + r * NB) * pitch];
    if constexpr (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const Cplx<T> w = tw[r * k * (L / (NS * R))];
        v[p][r] = INV ? cmulc(v[p][r], w) : cmul(v[p][r], w);
      }
    }
    dft<T, R, INV>(v[p]);
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int j = t + p * TL;
    const int k = j % NS;
    const int base = (j / NS) * NS * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) buf[(base + r * NS) * pitch] = v[p][r];
  }
  __syncthreads();
}

template <typename T, int L, int TL, bool INV, int NS, int R0, int... Rest>
__device__ __forceinline__ void run_stages(Cplx<T>* buf, int pitch, int t, const Cplx<T>* tw) {
  stage<T, L, R0, NS, TL, INV>(buf, pitch, t, tw);
  if constexpr (sizeof...(Rest) > 0) run_stages<T, L, TL, INV, NS * R0, Rest...>(buf, pitch, t, tw);
}

template <typename T, int L, int TL, bool INV, int... Rs>
__device__ __forceinline__ void run_plan(Cplx<T>* buf, int pitch, int t, const Cplx<T>* tw,
                                         Radices<Rs...>) {
  run_stages<T, L, TL, INV, 1, Rs...>(buf, pitch, t, tw);
}

// Transform one line in place. All threads of the workgroup must call this together
// (barriers inside); the caller synchronises before (data loaded) and the routine ends
// with a barrier (results visible).
template <typename T, int L, bool INV>
__device__ __forceinline__ void fft_lds(Cplx<T>* line, int pitch, int t, const Cplx<T>* tw) {
  constexpr int TL = team_size<L>();
  run_plan<T, L, TL, INV>(line, pitch, t, tw, typename Plan<L>::R{});
}

}  // namespace fft
}  // namespace sopht
