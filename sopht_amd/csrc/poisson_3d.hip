// Free-space Poisson solve -lap(psi) = omega on an (nz, ny, nx) grid by Hockney-Eastwood
// domain doubling (UnboundedPoissonSolverPYFFTW3D.py:9-172), as pruned FFT passes on gfx950.
//
// The doubled domain is never materialised. Per component c (all components of a vector
// solve go through each pass in one launch):
//   K1 xr2c : rows (c, z<nz, y<ny): real row (zero-padded to 2nx) -> nx+1 complex, two rows
//             packed per complex FFT and separated through the conjugate pairing.
//   K2 yfwd : columns (c, z<nz, kx): y<ny inputs, zero-padded FFT of length 2ny -> all ky.
//   K3 zconv: pencils (c, ky, kx): z<nz inputs, FFT 2nz, times G^ (real, even: read from the
//             (nz+1)(ny+1)(nx+1) octant), inverse FFT, keep z<nz.
//   K4 yinv : columns: inverse FFT over all ky, keep y<ny.
//   K5 xc2r : rows: Hermitian half-spectra of two rows -> one complex inverse FFT, keep x<nx.
// All of them work in the spectral slab H[c][z<nz][ky<2ny][kx<=nx] (complex, real_t).
// G^ is the DFT of the even-reflected Green's function G = 1/(4 pi r), G(0) = 1/(4 pi dx)
// (:56-83); being real and even per axis it is a separable DCT-I of the octant, evaluated
// here in FP64 on the device, then scaled by dx^3 / (8 nz ny nx) (the dx^3 of :47-49 and the
// 1/(2N)^3 backward normalisation of pyfftw) and stored in real_t.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <new>

#include "common.h"
#include "fft_reg.h"

struct sopht_poisson3d {
  int dtype;
  int64_t nz, ny, nx;
  double x_range, dx;
  void* slab;       // complex slab, 3 components
  void* ghat;       // real octant
  void* tw[3];      // twiddles for lengths 2nx, 2ny, 2nz (x, y, z)
  int64_t slab_bytes, ghat_bytes;
};

namespace sopht {
namespace {

using rfft::ColSlot;
using rfft::LinePlan;
using rfft::RowSlot;
using rfft::team_fft;
using rfft::static_for;

// Dynamic LDS per workgroup stays within 64 KiB, so two 1024-thread workgroups fit a CU.
constexpr int kLdsBudget = 64 * 1024;

// Threads per workgroup: 1024 (a 128-VGPR budget) for 8-element FP32 lines; 512 (256
// VGPRs) for the 12/24-element plans and FP64, which need more registers.
template <typename T, int L>
constexpr int max_threads() {
  return (LinePlan<L>::E > 8 || sizeof(T) == 8) ? 512 : 1024;
}

// Column tiles (y and z passes): WC neighbouring kx columns per workgroup, T threads per
// column.
template <typename T, int L>
constexpr int cols_per_block() {
  constexpr int sc = int(sizeof(Cplx<T>));
  int wc = max_threads<T, L>() / LinePlan<L>::T;
  const int cap = kLdsBudget / (L * sc);
  if (wc > cap) wc = cap;
  return wc < 1 ? 1 : wc;
}

// Row tiles (x passes): LPB row pairs per workgroup.
template <typename T, int L>
constexpr int rows_per_block() {
  constexpr int sc = int(sizeof(Cplx<T>));
  int lpb = max_threads<T, L>() / LinePlan<L>::T;
  const int cap = kLdsBudget / (RowSlot<L>::kStride * sc);
  if (lpb > cap) lpb = cap;
  return lpb < 1 ? 1 : lpb;
}

template <typename T, int L>
constexpr size_t row_lds_bytes() {
  return size_t(rows_per_block<T, L>()) * RowSlot<L>::kStride * sizeof(Cplx<T>);
}

template <typename T, int L>
constexpr size_t col_lds_bytes() {
  return size_t(L) * cols_per_block<T, L>() * sizeof(Cplx<T>);
}

struct SlabGeom {
  int64_t nz, ny, nx;  // physical grid
  int ncomp;
  // H index of (c, z, ky, kx)
  __device__ __forceinline__ int64_t h(int c, int64_t z, int64_t ky, int64_t kx) const {
    return ((int64_t(c) * nz + z) * (2 * ny) + ky) * (nx + 1) + kx;
  }
  __device__ __forceinline__ int64_t row_h(int64_t r) const {  // H row of input row r
    const int c = int(r / (nz * ny));
    const int64_t rem = r % (nz * ny);
    return h(c, rem / ny, rem % ny, 0);
  }
};

struct CompPtrs {
  const void* p[3];
  int64_t sz, sy, sx;
};
struct CompPtrsW {
  void* p[3];
  int64_t sz, sy, sx;
};

// K1: x real-to-complex of zero-padded rows, two rows (ra, rb) per complex FFT of length
// L = 2nx; the pair is separated with A = (Z[k] + conj Z[L-k])/2, B = (Z[k] - conj Z[L-k])/(2i).
template <typename T, int L>
__global__ void __launch_bounds__((max_threads<T, L>())) k1_xr2c(SlabGeom g, CompPtrs in, Cplx<T>* H,
                                                const Cplx<T>* tw) {
  constexpr int E = LinePlan<L>::E, TT = LinePlan<L>::T;
  constexpr int LPB = rows_per_block<T, L>();
  extern __shared__ unsigned char smem_raw[];
  Cplx<T>* xb = reinterpret_cast<Cplx<T>*>(smem_raw);
  const int t = threadIdx.x % TT;
  const int line = threadIdx.x / TT;
  const int64_t rows = int64_t(g.ncomp) * g.nz * g.ny;
  const int64_t q = int64_t(blockIdx.x) * LPB + line;
  const int64_t ra = 2 * q, rb = 2 * q + 1;
  auto row_ptr = [&](int64_t r) -> const T* {
    const int c = int(r / (g.nz * g.ny));
    const int64_t rem = r % (g.nz * g.ny);
    return static_cast<const T*>(in.p[c]) + (rem / g.ny) * in.sz + (rem % g.ny) * in.sy;
  };
  const T* pa = ra < rows ? row_ptr(ra) : nullptr;
  const T* pb = rb < rows ? row_ptr(rb) : nullptr;
  Cplx<T> v[E];
  static_for<0, E>([&](auto e_) {
    constexpr int e = decltype(e_)::value;
    if constexpr (e < E / 2) {  // x = t + T e < nx
      const int64_t x = t + TT * e;
      v[e] = Cplx<T>{pa ? pa[x * in.sx] : T(0), pb ? pb[x * in.sx] : T(0)};
    } else {
      v[e] = Cplx<T>{T(0), T(0)};
    }
  });
  RowSlot<L> slot{line};
  team_fft<T, L, false>(v, xb, tw, t, slot);
  static_for<0, E>([&](auto e_) {
    constexpr int e = decltype(e_)::value;
    xb[slot(t + TT * e)] = v[e];
  });
  __syncthreads();
  Cplx<T>* ha = ra < rows ? H + g.row_h(ra) : nullptr;
  Cplx<T>* hb = rb < rows ? H + g.row_h(rb) : nullptr;
  for (int k = t; k <= g.nx; k += TT) {
    const Cplx<T> zk = xb[slot(k)];
    const Cplx<T> zm = xb[slot((L - k) % L)];
    if (ha) ha[k] = Cplx<T>{T(0.5) * (zk.re + zm.re), T(0.5) * (zk.im - zm.im)};
    if (hb) hb[k] = Cplx<T>{T(0.5) * (zk.im + zm.im), T(0.5) * (zm.re - zk.re)};
  }
}

// K2 / K4: FFT along y of columns (c, z, kx). Forward: inputs y < ny (zero padding), all
// 2ny outputs stored. Inverse: all 2ny inputs, outputs y < ny stored.
template <typename T, int L, bool INV>
__global__ void __launch_bounds__((max_threads<T, L>())) k_ycols(SlabGeom g, Cplx<T>* H, const Cplx<T>* tw) {
  constexpr int E = LinePlan<L>::E, TT = LinePlan<L>::T;
  constexpr int WC = cols_per_block<T, L>();
  extern __shared__ unsigned char smem_raw[];
  Cplx<T>* xb = reinterpret_cast<Cplx<T>*>(smem_raw);
  const int w = threadIdx.x % WC;
  const int t = threadIdx.x / WC;
  const int64_t ncols = g.nx + 1;
  const int64_t tiles = (ncols + WC - 1) / WC;
  const int64_t tile = blockIdx.x % tiles;
  const int64_t cz = blockIdx.x / tiles;  // c * nz + z
  const int c = int(cz / g.nz);
  const int64_t z = cz % g.nz;
  const int64_t kx = tile * WC + w;
  const bool live = kx < ncols;
  Cplx<T>* col = H + g.h(c, z, 0, live ? kx : 0);
  const int64_t ys = ncols;  // stride between consecutive y (or ky)
  Cplx<T> v[E];
  static_for<0, E>([&](auto e_) {
    constexpr int e = decltype(e_)::value;
    if constexpr (INV || e < E / 2) v[e] = col[(t + TT * e) * ys];
    else v[e] = Cplx<T>{T(0), T(0)};
  });
  team_fft<T, L, INV>(v, xb, tw, t, ColSlot<WC>{w});
  if (live)
    static_for<0, (INV ? E / 2 : E)>([&](auto e_) {
      constexpr int e = decltype(e_)::value;
      col[(t + TT * e) * ys] = v[e];
    });
}

// K3: z pencils (c, ky, kx): zero-padded forward FFT, times G^ (real, even: read from the
// octant), inverse FFT, crop to z < nz.
template <typename T, int L>
__global__ void __launch_bounds__((max_threads<T, L>())) k3_zconv(SlabGeom g, Cplx<T>* H, const T* ghat,
                                                 const Cplx<T>* tw) {
  constexpr int E = LinePlan<L>::E, TT = LinePlan<L>::T;
  constexpr int WC = cols_per_block<T, L>();
  extern __shared__ unsigned char smem_raw[];
  Cplx<T>* xb = reinterpret_cast<Cplx<T>*>(smem_raw);
  const int w = threadIdx.x % WC;
  const int t = threadIdx.x / WC;
  const int64_t ncols = g.nx + 1;
  const int64_t tiles = (ncols + WC - 1) / WC;
  const int64_t tile = blockIdx.x % tiles;
  const int64_t cky = blockIdx.x / tiles;  // c * 2ny + ky
  const int c = int(cky / (2 * g.ny));
  const int64_t ky = cky % (2 * g.ny);
  const int64_t kx = tile * WC + w;
  const bool live = kx < ncols;
  const int64_t kxs = live ? kx : 0;
  Cplx<T>* col = H + g.h(c, 0, ky, kxs);
  const int64_t zs = 2 * g.ny * ncols;  // stride between consecutive z
  Cplx<T> v[E];
  static_for<0, E>([&](auto e_) {
    constexpr int e = decltype(e_)::value;
    if constexpr (e < E / 2) v[e] = col[(t + TT * e) * zs];
    else v[e] = Cplx<T>{T(0), T(0)};
  });
  // G^ at this thread's frequencies kz = t + T e, loaded ahead of the transform
  T gv[E];
  {
    const int64_t kyo = ky <= g.ny ? ky : 2 * g.ny - ky;
    const T* gcol = ghat + kyo * ncols + kxs;
    const int64_t gz = (g.ny + 1) * ncols;
    static_for<0, E>([&](auto e_) {
      constexpr int e = decltype(e_)::value;
      const int kz = t + TT * e;
      gv[e] = gcol[(kz <= L / 2 ? kz : L - kz) * gz];
    });
  }
  ColSlot<WC> slot{w};
  team_fft<T, L, false>(v, xb, tw, t, slot);
  static_for<0, E>([&](auto e_) {
    constexpr int e = decltype(e_)::value;
    v[e] = Cplx<T>{v[e].re * gv[e], v[e].im * gv[e]};
  });
  team_fft<T, L, true>(v, xb, tw, t, slot);
  if (live)
    static_for<0, E / 2>([&](auto e_) {
      constexpr int e = decltype(e_)::value;
      col[(t + TT * e) * zs] = v[e];
    });
}

// K5: x complex-to-real of two rows per complex inverse FFT, crop to x < nx. The Hermitian
// half spectra A, B (kx <= nx) of rows ra, rb become Z[k] = A[k] + i B[k],
// Z[L-k] = conj(A[k]) + i conj(B[k]); imaginary parts at kx = 0 and nx are dropped
// (irfft semantics).
template <typename T, int L>
__global__ void __launch_bounds__((max_threads<T, L>())) k5_xc2r(SlabGeom g, const Cplx<T>* H, CompPtrsW out,
                                                const Cplx<T>* tw) {
  constexpr int E = LinePlan<L>::E, TT = LinePlan<L>::T;
  constexpr int LPB = rows_per_block<T, L>();
  extern __shared__ unsigned char smem_raw[];
  Cplx<T>* xb = reinterpret_cast<Cplx<T>*>(smem_raw);
  const int t = threadIdx.x % TT;
  const int line = threadIdx.x / TT;
  const int64_t rows = int64_t(g.ncomp) * g.nz * g.ny;
  const int64_t q = int64_t(blockIdx.x) * LPB + line;
  const int64_t ra = 2 * q, rb = 2 * q + 1;
  const int64_t nx = g.nx;
  const Cplx<T>* ha = ra < rows ? H + g.row_h(ra) : nullptr;
  const Cplx<T>* hb = rb < rows ? H + g.row_h(rb) : nullptr;
  RowSlot<L> slot{line};
  for (int k = t; k <= nx; k += TT) {
    Cplx<T> a = {T(0), T(0)}, bb = {T(0), T(0)};
    if (ha) a = ha[k];
    if (hb) bb = hb[k];
    if (k == 0 || k == nx) {
      a.im = T(0);
      bb.im = T(0);
    }
    xb[slot(k)] = Cplx<T>{a.re - bb.im, a.im + bb.re};
    if (k > 0 && k < nx) xb[slot(L - k)] = Cplx<T>{a.re + bb.im, bb.re - a.im};
  }
  __syncthreads();
  Cplx<T> v[E];
  static_for<0, E>([&](auto e_) {
    constexpr int e = decltype(e_)::value;
    v[e] = xb[slot(t + TT * e)];
  });
  __syncthreads();  // exchange area reuse
  team_fft<T, L, true>(v, xb, tw, t, slot);
  auto row_ptr = [&](int64_t r) -> T* {
    const int c = int(r / (g.nz * g.ny));
    const int64_t rem = r % (g.nz * g.ny);
    return static_cast<T*>(out.p[c]) + (rem / g.ny) * out.sz + (rem % g.ny) * out.sy;
  };
  T* pa = ra < rows ? row_ptr(ra) : nullptr;
  T* pb = rb < rows ? row_ptr(rb) : nullptr;
  static_for<0, E / 2>([&](auto e_) {
    constexpr int e = decltype(e_)::value;
    const int64_t x = t + TT * e;
    if (pa) pa[x * out.sx] = v[e].re;
    if (pb) pb[x * out.sx] = v[e].im;
  });
}

// ------------------------------------------------------------------ Green's function octant
// g[z][y][x] = 1/(4 pi r dx) on integer offsets, g[0] = 1/(4 pi dx), in FP64.
__global__ void k_green_octant(double* g, int64_t nz1, int64_t ny1, int64_t nx1, double dx) {
  const int64_t n = nz1 * ny1 * nx1;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t x = i % nx1, y = (i / nx1) % ny1, z = i / (nx1 * ny1);
    const double r = sqrt(double(x * x + y * y + z * z)) * dx;
    g[i] = (i == 0) ? 1.0 / (4.0 * M_PI * dx) : 1.0 / (4.0 * M_PI * r);
  }
}

// DCT-I along one axis of the octant (length n+1 per line, doubled period 2n):
// out[k] = g[0] + (-1)^k g[n] + 2 sum_{j=1}^{n-1} g[j] cos(pi j k / n), k = 0..n.
// `stride` is the element stride of the axis, `lines` enumerates all other indices through
// (outer, inner) with inner extent `inner` (= stride).
__global__ void k_dct1_axis(const double* in, double* out, int64_t n, int64_t stride,
                            int64_t inner, int64_t outer, const double* cos_tab) {
  const int64_t total = outer * (n + 1) * inner;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < total;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t in_i = i % inner;
    const int64_t k = (i / inner) % (n + 1);
    const int64_t o = i / (inner * (n + 1));
    const double* line = in + o * (n + 1) * inner + in_i;
    double acc = line[0] + ((k & 1) ? -line[n * stride] : line[n * stride]);
    double s = 0.0;
    for (int64_t j = 1; j < n; ++j) s += line[j * stride] * cos_tab[(j * k) % (2 * n)];
    out[i] = acc + 2.0 * s;
  }
}

template <typename T>
__global__ void k_scale_convert(const double* in, T* out, int64_t n, double s) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    out[i] = T(in[i] * s);
}

// ------------------------------------------------------------------ dispatch on length
#define SOPHT_FFT_LENGTHS(X) \
  X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024)

bool length_supported(int64_t L) {
#define X(n) if (L == n) return true;
  SOPHT_FFT_LENGTHS(X)
#undef X
  return false;
}

template <typename T>
int launch_k1(int64_t L, SlabGeom g, CompPtrs in, Cplx<T>* H, const Cplx<T>* tw, hipStream_t s) {
  const int64_t rows = int64_t(g.ncomp) * g.nz * g.ny;
  const int64_t pairs = (rows + 1) / 2;
#define X(n)                                                                          \
  if (L == n) {                                                                       \
    constexpr int lpb = rows_per_block<T, n>();                                       \
    const int threads = lpb * LinePlan<n>::T;                                         \
    k1_xr2c<T, n><<<ceil_div(pairs, lpb), threads, row_lds_bytes<T, n>(), s>>>(g, in, H, tw); \
    SOPHT_CHECK_LAUNCH();                                                             \
    return SOPHT_OK;                                                                  \
  }
  SOPHT_FFT_LENGTHS(X)
#undef X
  return SOPHT_ERR_UNSUPPORTED;
}

template <typename T, bool INV>
int launch_ycols(int64_t L, SlabGeom g, Cplx<T>* H, const Cplx<T>* tw, hipStream_t s) {
#define X(n)                                                                          \
  if (L == n) {                                                                       \
    constexpr int wc = cols_per_block<T, n>();                                        \
    const int threads = wc * LinePlan<n>::T;                                          \
    const int64_t tiles = (g.nx + 1 + wc - 1) / wc;                                   \
    k_ycols<T, n, INV><<<tiles * g.ncomp * g.nz, threads, col_lds_bytes<T, n>(), s>>>(g, H, tw); \
    SOPHT_CHECK_LAUNCH();                                                             \
    return SOPHT_OK;                                                                  \
  }
  SOPHT_FFT_LENGTHS(X)
#undef X
  return SOPHT_ERR_UNSUPPORTED;
}

template <typename T>
int launch_k3(int64_t L, SlabGeom g, Cplx<T>* H, const T* ghat, const Cplx<T>* tw,
              hipStream_t s) {
#define X(n)                                                                          \
  if (L == n) {                                                                       \
    constexpr int wc = cols_per_block<T, n>();                                        \
    const int threads = wc * LinePlan<n>::T;                                          \
    const int64_t tiles = (g.nx + 1 + wc - 1) / wc;                                   \
    k3_zconv<T, n><<<tiles * g.ncomp * 2 * g.ny, threads, col_lds_bytes<T, n>(), s>>>(g, H, ghat, tw); \
    SOPHT_CHECK_LAUNCH();                                                             \
    return SOPHT_OK;                                                                  \
  }
  SOPHT_FFT_LENGTHS(X)
#undef X
  return SOPHT_ERR_UNSUPPORTED;
}

template <typename T>
int launch_k5(int64_t L, SlabGeom g, const Cplx<T>* H, CompPtrsW out, const Cplx<T>* tw,
              hipStream_t s) {
  const int64_t rows = int64_t(g.ncomp) * g.nz * g.ny;
  const int64_t pairs = (rows + 1) / 2;
#define X(n)                                                                          \
  if (L == n) {                                                                       \
    constexpr int lpb = rows_per_block<T, n>();                                       \
    const int threads = lpb * LinePlan<n>::T;                                         \
    k5_xc2r<T, n><<<ceil_div(pairs, lpb), threads, row_lds_bytes<T, n>(), s>>>(g, H, out, tw); \
    SOPHT_CHECK_LAUNCH();                                                             \
    return SOPHT_OK;                                                                  \
  }
  SOPHT_FFT_LENGTHS(X)
#undef X
  return SOPHT_ERR_UNSUPPORTED;
}

template <typename T>
int solve_impl(sopht_poisson3d* P, int ncomp, void* const sol[3], const int64_t* ss,
               const void* const rhs[3], const int64_t* rs, hipStream_t s,
               StageTimer* timer = nullptr, int first_stage = 0) {
  SlabGeom g{P->nz, P->ny, P->nx, ncomp};
  CompPtrs in;
  CompPtrsW out;
  for (int c = 0; c < 3; ++c) {
    in.p[c] = c < ncomp ? rhs[c] : nullptr;
    out.p[c] = c < ncomp ? sol[c] : nullptr;
  }
  in.sz = rs[0]; in.sy = rs[1]; in.sx = rs[2];
  out.sz = ss[0]; out.sy = ss[1]; out.sx = ss[2];
  Cplx<T>* H = static_cast<Cplx<T>*>(P->slab);
  const Cplx<T>* twx = static_cast<const Cplx<T>*>(P->tw[0]);
  const Cplx<T>* twy = static_cast<const Cplx<T>*>(P->tw[1]);
  const Cplx<T>* twz = static_cast<const Cplx<T>*>(P->tw[2]);
  int rc;
  auto mark = [&](int i) {
    if (timer) timer->mark(first_stage + i, s);
  };
  if ((rc = launch_k1<T>(2 * P->nx, g, in, H, twx, s))) return rc;
  mark(0);
  if ((rc = launch_ycols<T, false>(2 * P->ny, g, H, twy, s))) return rc;
  mark(1);
  if ((rc = launch_k3<T>(2 * P->nz, g, H, static_cast<const T*>(P->ghat), twz, s))) return rc;
  mark(2);
  if ((rc = launch_ycols<T, true>(2 * P->ny, g, H, twy, s))) return rc;
  mark(3);
  rc = launch_k5<T>(2 * P->nx, g, H, out, twx, s);
  mark(4);
  return rc;
}

// exp(-2 pi i m / L) in FP64 with exact argument reduction to the first octant.
void twiddle_fp64(int64_t L, int64_t m, double& re, double& im) {
  m %= L;
  // angle = 2 pi m / L; use symmetry to keep the argument small
  const long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)L;
  re = double(cosl(a));
  im = -double(sinl(a));
}

template <typename T>
int upload_twiddles(int64_t L, void** dst, hipStream_t s) {
  Cplx<T>* host = static_cast<Cplx<T>*>(malloc(sizeof(Cplx<T>) * L));
  if (!host) return SOPHT_ERR_INVALID;
  for (int64_t m = 0; m < L; ++m) {
    double re, im;
    twiddle_fp64(L, m, re, im);
    host[m] = Cplx<T>{T(re), T(im)};
  }
  hipError_t e = hipMalloc(dst, sizeof(Cplx<T>) * L);
  if (e == hipSuccess) e = hipMemcpyAsync(*dst, host, sizeof(Cplx<T>) * L, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  free(host);
  SOPHT_CHECK_HIP(e);
  return SOPHT_OK;
}

int build_ghat(sopht_poisson3d* P, hipStream_t s) {
  const int64_t nz1 = P->nz + 1, ny1 = P->ny + 1, nx1 = P->nx + 1;
  const int64_t n = nz1 * ny1 * nx1;
  double *a = nullptr, *b = nullptr, *ct = nullptr;
  SOPHT_CHECK_HIP(hipMalloc(&a, n * sizeof(double)));
  SOPHT_CHECK_HIP(hipMalloc(&b, n * sizeof(double)));
  const int64_t nmax = std::max(P->nx, std::max(P->ny, P->nz));
  SOPHT_CHECK_HIP(hipMalloc(&ct, 2 * nmax * sizeof(double)));
  double* host = static_cast<double*>(malloc(2 * nmax * sizeof(double)));
  const int blocks = std::min<int64_t>(65536, (n + 255) / 256);
  k_green_octant<<<blocks, 256, 0, s>>>(a, nz1, ny1, nx1, P->dx);
  SOPHT_CHECK_LAUNCH();
  // axis passes: x (stride 1), y (stride nx1), z (stride nx1*ny1)
  struct Ax { int64_t n, stride, inner, outer; } axes[3] = {
      {P->nx, 1, 1, nz1 * ny1},
      {P->ny, nx1, nx1, nz1},
      {P->nz, nx1 * ny1, nx1 * ny1, 1}};
  double* src = a;
  double* dst = b;
  for (const Ax& ax : axes) {
    for (int64_t m = 0; m < 2 * ax.n; ++m)
      host[m] = double(cosl(3.14159265358979323846264338327950288L * (long double)m / (long double)ax.n));
    SOPHT_CHECK_HIP(hipMemcpyAsync(ct, host, 2 * ax.n * sizeof(double), hipMemcpyHostToDevice, s));
    SOPHT_CHECK_HIP(hipStreamSynchronize(s));
    k_dct1_axis<<<blocks, 256, 0, s>>>(src, dst, ax.n, ax.stride, ax.inner, ax.outer, ct);
    SOPHT_CHECK_LAUNCH();
    std::swap(src, dst);
  }
  const double scale = P->dx * P->dx * P->dx / (8.0 * double(P->nz) * double(P->ny) * double(P->nx));
  if (P->dtype == SOPHT_F32)
    k_scale_convert<float><<<blocks, 256, 0, s>>>(src, static_cast<float*>(P->ghat), n, scale);
  else
    k_scale_convert<double><<<blocks, 256, 0, s>>>(src, static_cast<double*>(P->ghat), n, scale);
  SOPHT_CHECK_LAUNCH();
  SOPHT_CHECK_HIP(hipStreamSynchronize(s));
  free(host);
  SOPHT_CHECK_HIP(hipFree(a));
  SOPHT_CHECK_HIP(hipFree(b));
  SOPHT_CHECK_HIP(hipFree(ct));
  return SOPHT_OK;
}

}  // namespace

int poisson_vector_solve_timed(sopht_poisson3d* P, void* const sol[3], const int64_t* ss,
                               const void* const rhs[3], const int64_t* rs, hipStream_t s,
                               StageTimer* timer, int first_stage) {
  return P->dtype == SOPHT_F32 ? solve_impl<float>(P, 3, sol, ss, rhs, rs, s, timer, first_stage)
                               : solve_impl<double>(P, 3, sol, ss, rhs, rs, s, timer, first_stage);
}

}  // namespace sopht

using namespace sopht;

extern "C" {

int sopht_poisson3d_create(int dtype, int64_t nz, int64_t ny, int64_t nx, double x_range,
                           void* stream, sopht_poisson3d** out) {
  SOPHT_REQUIRE(out != nullptr, "plan pointer is null");
  *out = nullptr;
  SOPHT_REQUIRE(dtype == SOPHT_F32 || dtype == SOPHT_F64, "dtype must be SOPHT_F32 or SOPHT_F64");
  SOPHT_REQUIRE(nz > 0 && ny > 0 && nx > 0 && x_range > 0, "grid sizes and x_range must be positive");
  if (!length_supported(2 * nx) || !length_supported(2 * ny) || !length_supported(2 * nz)) {
    set_last_error("unsupported Poisson grid size: 2*n must be one of 8, 12, 16, 24, 32, 48, 64, "
                   "96, 128, 192, 256, 384, 512, 768, 1024", __FILE__, __LINE__);
    return SOPHT_ERR_UNSUPPORTED;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  sopht_poisson3d* P = new (std::nothrow) sopht_poisson3d();
  SOPHT_REQUIRE(P != nullptr, "out of host memory");
  P->dtype = dtype;
  P->nz = nz; P->ny = ny; P->nx = nx;
  P->x_range = x_range;
  P->dx = x_range / double(nx);
  if (dtype == SOPHT_F32) P->dx = double(float(P->dx));  // the reference's real_t(dx)
  const size_t cs = dtype == SOPHT_F32 ? sizeof(Cplx<float>) : sizeof(Cplx<double>);
  const size_t rsz = dtype == SOPHT_F32 ? sizeof(float) : sizeof(double);
  P->slab_bytes = int64_t(3) * nz * 2 * ny * (nx + 1) * int64_t(cs);
  P->ghat_bytes = (nz + 1) * (ny + 1) * (nx + 1) * int64_t(rsz);
  int rc = SOPHT_OK;
  auto fail = [&](int code) {
    sopht_poisson3d_destroy(P);
    return code;
  };
  if (hipMalloc(&P->slab, P->slab_bytes) != hipSuccess) {
    set_last_error("hipMalloc of the Poisson slab failed", __FILE__, __LINE__);
    return fail(SOPHT_ERR_HIP);
  }
  if (hipMalloc(&P->ghat, P->ghat_bytes) != hipSuccess) {
    set_last_error("hipMalloc of the Green's function failed", __FILE__, __LINE__);
    return fail(SOPHT_ERR_HIP);
  }
  const int64_t Ls[3] = {2 * nx, 2 * ny, 2 * nz};
  for (int a = 0; a < 3; ++a) {
    rc = dtype == SOPHT_F32 ? upload_twiddles<float>(Ls[a], &P->tw[a], s)
                            : upload_twiddles<double>(Ls[a], &P->tw[a], s);
    if (rc) return fail(rc);
  }
  if ((rc = build_ghat(P, s))) return fail(rc);
  *out = P;
  return SOPHT_OK;
}

int sopht_poisson3d_destroy(sopht_poisson3d* P) {
  if (!P) return SOPHT_OK;
  if (P->slab) (void)hipFree(P->slab);
  if (P->ghat) (void)hipFree(P->ghat);
  for (int a = 0; a < 3; ++a)
    if (P->tw[a]) (void)hipFree(P->tw[a]);
  delete P;
  return SOPHT_OK;
}

int64_t sopht_poisson3d_scratch_bytes(const sopht_poisson3d* P) {
  return P ? P->slab_bytes + P->ghat_bytes : 0;
}

int sopht_poisson3d_solve(sopht_poisson3d* P, void* sol, const int64_t* ss, const void* rhs,
                          const int64_t* rs, void* stream) {
  SOPHT_REQUIRE(P != nullptr, "plan is null");
  void* const so[3] = {sol, nullptr, nullptr};
  const void* const rh[3] = {rhs, nullptr, nullptr};
  hipStream_t s = static_cast<hipStream_t>(stream);
  return P->dtype == SOPHT_F32 ? solve_impl<float>(P, 1, so, ss, rh, rs, s)
                               : solve_impl<double>(P, 1, so, ss, rh, rs, s);
}

int sopht_poisson3d_vector_solve(sopht_poisson3d* P, void* const sol[3], const int64_t* ss,
                                 const void* const rhs[3], const int64_t* rs, void* stream) {
  SOPHT_REQUIRE(P != nullptr, "plan is null");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return P->dtype == SOPHT_F32 ? solve_impl<float>(P, 3, sol, ss, rhs, rs, s)
                               : solve_impl<double>(P, 3, sol, ss, rhs, rs, s);
}

}  // extern "C"
