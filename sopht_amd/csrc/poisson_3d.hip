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
  int64_t gz;       // G^ line stride (elements): nz + 1 rounded up to a 128-byte line
  int slab_comps;   // components the slab holds (3, or 3 per instance of a batched class)
  int64_t slab_bytes, ghat_bytes;
  int passes;       // 3: x and y passes fused (three-pass schedule), or 5
  int zjobs;        // three-pass z pass: jobs per workgroup, chosen at plan creation
};

namespace sopht {
namespace {

// A streaming store of one complex value: the next pass reads it back from HBM, so it need
// not stay in the cache hierarchy.
template <typename T>
__device__ __forceinline__ void store_stream(Cplx<T>* p, Cplx<T> v) {
  typedef T pair __attribute__((ext_vector_type(2)));
  pair w;
  w.x = v.re;
  w.y = v.im;
  __builtin_nontemporal_store(w, reinterpret_cast<pair*>(p));
}

using rfft::LinePlan;
using rfft::region_stride;
using rfft::static_for;
using rfft::team_fft;

// Slab rows (kx = 0..nx) are padded to a whole number of 128-byte lines, so every column
// tile of the y and z passes reads and writes whole cache lines.
template <typename T>
constexpr int kLineCplx = 128 / int(sizeof(Cplx<T>));
#ifndef SOPHT_COL_HALF
#define SOPHT_COL_HALF 1
#endif
#ifndef SOPHT_ROW_PAD
#define SOPHT_ROW_PAD 1
#endif
// Column tiles are half a 128-byte line when SOPHT_COL_HALF is set; the two halves of a line
// then go to workgroups b and b + 8, which the dispatcher sends to the same XCD one after the
// other, so the second half is read from that XCD's L2.
constexpr int kColHalf = SOPHT_COL_HALF;
constexpr int64_t kRowPad = SOPHT_ROW_PAD;
// Columns per slab tile of the three-pass schedule. Measured at 256^3 FP32 (pass A + z pass +
// pass C): 16 columns 1.21 ms, 8 columns 1.20 ms (the z pass faster, passes A and C slower),
// 4 columns 2.2 ms (pass A's writes become 32-byte pieces of lines from four planes, which
// other workgroups write).
#ifndef SOPHT_SLAB_TILE
#define SOPHT_SLAB_TILE 16
#endif
constexpr int kSlabTile = SOPHT_SLAB_TILE;
static_assert(16 % kSlabTile == 0, "slab tiles divide a 128-byte line");

// (row, first kx) of the column tile of workgroup b; false for the padding workgroups.
__device__ __forceinline__ bool col_tile(int64_t b, int64_t ncols, int64_t nrows, int wc,
                                         int64_t& row, int64_t& kx0) {
  if (kColHalf) {
    const int64_t lines = (ncols + 2 * wc - 1) / (2 * wc);
    const int64_t line = (b / 16) * 8 + b % 8;
    const int64_t half = (b / 8) % 2;
    row = line / lines;
    kx0 = ((line % lines) * 2 + half) * wc;
    return row < nrows;
  }
  const int64_t tiles = (ncols + wc - 1) / wc;
  row = b / tiles;
  kx0 = (b % tiles) * wc;
  return row < nrows;
}

// The same with a line split in PARTS tiles: the parts of a line go to workgroups
// b + 8 p (p < PARTS), which the dispatcher sends to one XCD.
template <int PARTS>
__device__ __forceinline__ bool col_tile_parts(int64_t b, int64_t ncols, int64_t nrows, int wc,
                                               int64_t& row, int64_t& kx0) {
  const int64_t lines = (ncols + PARTS * wc - 1) / (PARTS * wc);
  const int64_t line = (b / (8 * PARTS)) * 8 + b % 8;
  const int64_t part = (b / 8) % PARTS;
  row = line / lines;
  kx0 = ((line % lines) * PARTS + part) * wc;
  return row < nrows;
}

inline int64_t col_tile_parts_blocks(int parts, int64_t ncols, int64_t nrows, int wc) {
  const int64_t lines = (ncols + parts * wc - 1) / (parts * wc);
  return 8 * parts * ((lines * nrows + 7) / 8);
}

inline int64_t col_tile_blocks(int64_t ncols, int64_t nrows, int wc) {
  if (kColHalf) {
    const int64_t lines = (ncols + 2 * wc - 1) / (2 * wc);
    return 16 * ((lines * nrows + 7) / 8);
  }
  return (ncols + wc - 1) / wc * nrows;
}

inline int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

struct SlabGeom {
  int64_t nz, ny, nx;  // physical grid
  int64_t rs;          // row stride (complex elements), >= nx + 1
  int64_t gz;          // G^ line stride (elements), >= nz + 1
  int ncomp;
  // H index of (c, z, ky, kx)
  __device__ __forceinline__ int64_t h(int c, int64_t z, int64_t ky, int64_t kx) const {
    return ((int64_t(c) * nz + z) * (2 * ny) + ky) * rs + kx;
  }
  // the same as a 32-bit offset (plan creation checks that the slab has < 2^31 elements):
  // loads then take a uniform base plus one VGPR of offset
  __device__ __forceinline__ int hi(int c, int z, int ky, int kx) const {
    return ((c * int(nz) + z) * (2 * int(ny)) + ky) * int(rs) + kx;
  }
  // The three-pass schedule keeps the slab in column tiles of kSlabTile columns with z inside
  // a tile: (c, ky, kx / kSlabTile, z, kx % kSlabTile). The z pass then reads and writes each
  // tile's z line as one contiguous run; with 16 columns (one 128-byte line per plane) the x/y
  // passes still move whole lines (rs is a multiple of 16 there).
  // (Unsigned arithmetic: the divisions by the power-of-two tile width are shifts.) Moving by
  // dr rows adds dr rs nz; moving by a whole number of tiles, dk columns, adds dk nz.
  __device__ __forceinline__ int ti(int c, int z, int ky, int kx) const {
    const unsigned k = unsigned(kx);
    return int((((unsigned(c) * 2u * unsigned(ny) + unsigned(ky)) * (unsigned(rs) / kSlabTile) + k / kSlabTile) *
                    unsigned(nz) + unsigned(z)) * kSlabTile + k % kSlabTile);
  }
  __device__ __forceinline__ int64_t row_h(int64_t r) const {  // H row of input row r
    const int c = int(r / (nz * ny));
    const int64_t rem = r % (nz * ny);
    return h(c, rem / ny, rem % ny, 0);
  }
};

// Component fields of a solve: up to three pointers by value, or (batched solves over many
// instances of one grid size) a device table of one pointer per component.
struct CompPtrs {
  const void* p[3];
  const void* const* tab;
  int64_t sz, sy, sx;
  __device__ __forceinline__ const void* get(int c) const { return tab ? tab[c] : p[c]; }
};
struct CompPtrsW {
  void* p[3];
  void* const* tab;
  int64_t sz, sy, sx;
  __device__ __forceinline__ void* get(int c) const { return tab ? tab[c] : p[c]; }
};

// Row passes (x): one team of T threads per row pair, RowTile::NTEAM teams per workgroup.
template <typename T, int L>
struct RowTile {
  static constexpr int E = LinePlan<L>::E, TT = LinePlan<L>::T;
  static constexpr int NTEAM = TT >= 256 ? 1 : 256 / TT;
  static constexpr int NT = NTEAM * TT;
  static constexpr int S = region_stride<L, false>();
  static constexpr size_t lds = size_t(NTEAM) * S * sizeof(Cplx<T>);
};

// Column passes (y, z): WC neighbouring kx columns per workgroup (one 128-byte line of a
// slab row), one team per column. Loads and stores run with kx fastest (lanes (w, row));
// the transform runs one team per column; LDS region w (S elements) holds column w.
template <typename T, int L>
struct ColTile {
  static constexpr int E = LinePlan<L>::E, TT = LinePlan<L>::T;
  static constexpr int WC = kColHalf ? kLineCplx<T> / 2 : kLineCplx<T>;
  static constexpr int NT = WC * TT;
  static constexpr int S = region_stride<L, true>();
  static constexpr size_t lds = size_t(WC) * S * sizeof(Cplx<T>);
};

// K1: x real-to-complex of zero-padded rows, two rows (ra, rb) per complex FFT of length
// L = 2nx; the pair is separated with A = (Z[k] + conj Z[L-k])/2, B = (Z[k] - conj Z[L-k])/(2i).
template <typename T, int L>
__global__ void __launch_bounds__((RowTile<T, L>::NT)) k1_xr2c(SlabGeom g, CompPtrs in, Cplx<T>* H,
                                                             const Cplx<T>* tw) {
  using RT = RowTile<T, L>;
  constexpr int E = RT::E, TT = RT::TT;
  extern __shared__ unsigned char smem_raw[];
  Cplx<T>* xb = reinterpret_cast<Cplx<T>*>(smem_raw);
  const int t = threadIdx.x % TT;
  const int team = threadIdx.x / TT;
  const int64_t rows = int64_t(g.ncomp) * g.nz * g.ny;
  const int64_t q = int64_t(blockIdx.x) * RT::NTEAM + team;
  const int64_t ra = 2 * q, rb = 2 * q + 1;
  auto row_ptr = [&](int64_t r) -> const T* {
    const int c = int(r / (g.nz * g.ny));
    const int64_t rem = r % (g.nz * g.ny);
    return static_cast<const T*>(in.get(c)) + (rem / g.ny) * in.sz + (rem % g.ny) * in.sy;
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
  Cplx<T>* const line = xb + team * RT::S;
  // the twiddle table in LDS, shared by the workgroup's teams
  Cplx<T>* const twl = xb + RT::NTEAM * RT::S;
  rfft::load_tw_lds<T, L>(twl, tw, threadIdx.x, RT::NT);
  __syncthreads();
  team_fft<T, L, false, true, true>(v, line, twl, t);
  static_for<0, E>([&](auto e_) {
    constexpr int e = decltype(e_)::value;
    line[t + TT * e] = v[e];
  });
  rfft::wave_sync();
  Cplx<T>* ha = ra < rows ? H + g.row_h(ra) : nullptr;
  Cplx<T>* hb = rb < rows ? H + g.row_h(rb) : nullptr;
  for (int k = t; k <= g.nx; k += TT) {
    const Cplx<T> zk = line[k];
    const Cplx<T> zm = line[(L - k) % L];
    if (ha) store_stream(ha + k, Cplx<T>{T(0.5) * (zk.re + zm.re), T(0.5) * (zk.im - zm.im)});
    if (hb) store_stream(hb + k, Cplx<T>{T(0.5) * (zk.im + zm.im), T(0.5) * (zm.re - zk.re)});
  }
}

// K2 / K4: FFT along y of columns (c, z, kx). Forward: inputs y < ny (zero padding), all
// 2ny outputs stored. Inverse: all 2ny inputs, outputs y < ny stored.
template <typename T, int L, bool INV>
__global__ void __launch_bounds__((ColTile<T, L>::NT)) k_ycols(SlabGeom g, Cplx<T>* H, const Cplx<T>* tw) {
  using CT = ColTile<T, L>;
  constexpr int E = CT::E, TT = CT::TT, WC = CT::WC, S = CT::S;
  constexpr int KIN = INV ? E : E / 2, KOUT = INV ? E / 2 : E;  // rows = TT * K
  extern __shared__ unsigned char smem_raw[];
  Cplx<T>* xb = reinterpret_cast<Cplx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int lw = tid % WC, lr = tid / WC;  // load / store role
  const int team = tid / TT, t = tid % TT;  // transform role
  const int64_t ncols = g.nx + 1;
  int64_t cz, kx0;  // cz = c * nz + z
  if (!col_tile(blockIdx.x, ncols, int64_t(g.ncomp) * g.nz, WC, cz, kx0)) return;
  const int c = int(cz / g.nz);
  const int64_t z = cz % g.nz;
  const int64_t kx = kx0 + lw;
  const bool live = kx < ncols;
  // a wavefront whose teams all have kx >= ncols skips its transforms
  const bool wave_live = kx0 + (tid / 64) * (64 / TT) < ncols;
  const int col = g.hi(c, int(z), 0, live ? int(kx) : 0);
  const int ys = int(g.rs);  // stride between consecutive y (or ky)
  // the twiddle table in LDS (one load per thread), read by every stage of the 8 transforms
  Cplx<T>* const twl = xb + WC * S;
  rfft::load_tw_lds<T, L>(twl, tw, tid, CT::NT);
  static_for<0, KIN>([&](auto k_) {
    constexpr int k = decltype(k_)::value;
    const int y = lr + TT * k;
    xb[lw * S + y] = live ? H[col + y * ys] : Cplx<T>{T(0), T(0)};
  });
  __syncthreads();
  Cplx<T>* const line = xb + team * S;
  Cplx<T> v[E];
  static_for<0, E>([&](auto e_) {
    constexpr int e = decltype(e_)::value;
    if constexpr (e < KIN) v[e] = line[t + TT * e];
    else v[e] = Cplx<T>{T(0), T(0)};
  });
  rfft::wave_sync();
  if (wave_live) team_fft<T, L, INV, !INV, true>(v, line, twl, t);
  static_for<0, KOUT>([&](auto e_) {
    constexpr int e = decltype(e_)::value;
    line[t + TT * e] = v[e];
  });
  __syncthreads();
  if (live)
    static_for<0, KOUT>([&](auto k_) {
      constexpr int k = decltype(k_)::value;
      const int y = lr + TT * k;
      if constexpr (INV)
        H[col + y * ys] = xb[lw * S + y];
      else
        store_stream(H + col + y * ys, xb[lw * S + y]);
    });
}

// Waves per SIMD the z pass is compiled for (its VGPR budget: 128 in FP32; a budget of 80,
// for three 512-thread workgroups per CU, spills).
template <typename T>
constexpr int kZconvWaves = sizeof(T) == 4 ? 4 : 2;
// more elements per thread need more registers: twelve (L = 384, 768) one wavefront less per
// SIMD, sixteen (L = 1024) two less
template <typename T, int L>
constexpr int kZconvWavesFor = sizeof(T) != 4 ? kZconvWaves<T>
                               : (LinePlan<L>::E >= 16 ? 2 : (LinePlan<L>::E > 8 ? 3 : kZconvWaves<T>));
// pencil tiles per z-pass workgroup at most (all of them for a 3-component solve)
constexpr int kZconvJobs = 6;

// Three-stage line plans: the twiddles of a thread's butterflies in the second and last stages
// depend on the thread only, so a z-pass workgroup computes them once for all its pencils.
template <typename R>
struct Plan3 {
  static constexpr bool ok = false;
  static constexpr int r0 = 0, r1 = 0, r2 = 0;
};
template <int A, int B, int C>
struct Plan3<rfft::Radices<A, B, C>> {
  static constexpr bool ok = true;
  static constexpr int r0 = A, r1 = B, r2 = C;
};
template <int L>
constexpr bool last_twiddles_hoisted() {  // the E = 8 plans (L = 128, 256, 512): 24-28 VGPRs
  using P = Plan3<typename LinePlan<L>::R>;
  return P::ok && LinePlan<L>::E == 8;
}
constexpr int at_least_1(int v) { return v > 0 ? v : 1; }

// K3: z pencils (c, ky, kx): zero-padded forward FFT, times G^ (real, even), inverse FFT,
// crop to z < nz. A workgroup owns one kx tile of one |ky| = kyo and runs the pencils of
// every component at ky = kyo and ky = 2ny - kyo (the same G^ line) one after another: G^
// stays in registers, and the next pencil tile's loads are in flight during the transforms.
// G^ is stored [kyo][kx][kz], kz fastest, so a team reads its line contiguously.
template <typename T, int L>
__global__ void __launch_bounds__((ColTile<T, L>::NT)) __attribute__((amdgpu_waves_per_eu(kZconvWavesFor<T, L>))) k3_zconv(SlabGeom g, Cplx<T>* H, const T* ghat,
                                                              const Cplx<T>* tw) {
  using CT = ColTile<T, L>;
  constexpr int E = CT::E, TT = CT::TT, WC = CT::WC, S = CT::S, K = E / 2;
  extern __shared__ unsigned char smem_raw[];
  Cplx<T>* xb = reinterpret_cast<Cplx<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const int lw = tid % WC, lr = tid / WC;
  const int team = tid / TT, t = tid % TT;
  const int64_t ncols = g.nx + 1;
  int64_t kyo, kx0;  // kyo = 0..ny
  if (!col_tile(blockIdx.x, ncols, g.ny + 1, WC, kyo, kx0)) return;
  // jobs (component, ky or 2ny - ky) of this workgroup: chunk blockIdx.y of kZconvJobs
  const int alljobs = (kyo == 0 || kyo == g.ny) ? g.ncomp : 2 * g.ncomp;
  const int jbeg = int(blockIdx.y) * kZconvJobs;
  if (jbeg >= alljobs) return;
  const int njobs = alljobs - jbeg < kZconvJobs ? alljobs - jbeg : kZconvJobs;
  const int64_t kx = kx0 + lw;
  const bool live = kx < ncols;
  const bool wave_live = kx0 + (tid / 64) * (64 / TT) < ncols;
  T gv[E];
  {
    const int64_t kxt = kx0 + team;
    const T* gq = ghat + int((kyo * ncols + (kxt < ncols ? kxt : 0)) * g.gz);
    static_for<0, E>([&](auto e_) {
      constexpr int e = decltype(e_)::value;
      const int kz = t + TT * e;
      gv[e] = gq[kz <= L / 2 ? kz : L - kz];
    });
  }
  const int zs = 2 * int(g.ny) * int(g.rs);  // stride between consecutive z
  // the twiddles of this thread's butterflies in the second and last stages, the same for
  // every pencil of the workgroup
  using P3 = Plan3<typename LinePlan<L>::R>;
  constexpr int R1 = at_least_1(P3::r1), R2 = at_least_1(P3::r2);
  constexpr int PER2 = at_least_1(E / R1), PER3 = at_least_1(E / R2);
  Cplx<T> w2[PER2][R1], w3[PER3][R2];
  if constexpr (last_twiddles_hoisted<L>()) {
    static_for<0, PER2>([&](auto p_) {
      constexpr int p = decltype(p_)::value;
      rfft::stage_twiddles<T, L, R1, P3::r0>(w2[p], tw, (t + p * TT) % P3::r0);
    });
    static_for<0, PER3>([&](auto p_) {
      constexpr int p = decltype(p_)::value;
      rfft::stage_twiddles<T, L, R2, P3::r0 * R1>(w3[p], tw, (t + p * TT) % (P3::r0 * R1));
    });
  }
  auto column = [&](int jl) -> int {
    const int j = jbeg + jl;
    const int c = j % g.ncomp;
    const int ky = j < g.ncomp ? int(kyo) : 2 * int(g.ny) - int(kyo);
    return g.hi(c, 0, ky, live ? int(kx) : 0);
  };
  Cplx<T> pf[K];
  auto load = [&](int j) {
    const int col = column(j);
    static_for<0, K>([&](auto k_) {
      constexpr int k = decltype(k_)::value;
      pf[k] = live ? H[col + (lr + TT * k) * zs] : Cplx<T>{T(0), T(0)};
    });
  };
  Cplx<T>* const line = xb + team * S;
  load(0);
  for (int j = 0; j < njobs; ++j) {
    static_for<0, K>([&](auto k_) {
      constexpr int k = decltype(k_)::value;
      const int z = lr + TT * k;
      xb[lw * S + z] = pf[k];
    });
    __syncthreads();
    Cplx<T> v[E];
    static_for<0, E>([&](auto e_) {
      constexpr int e = decltype(e_)::value;
      if constexpr (e < K) v[e] = line[t + TT * e];
      else v[e] = Cplx<T>{T(0), T(0)};
    });
    rfft::wave_sync();
    if (j + 1 < njobs) load(j + 1);
    if (wave_live) {
      if constexpr (last_twiddles_hoisted<L>()) {
        constexpr int R0 = P3::r0, NS3 = R0 * R1;
        rfft::stockham_stage<T, L, R0, 1, false, true>(v, line, tw, t);
        rfft::stockham_stage_given<T, L, R1, R0, false>(v, line, w2, t);
        rfft::stockham_stage_given<T, L, R2, NS3, false>(v, line, w3, t);
        static_for<0, E>([&](auto e_) {
          constexpr int e = decltype(e_)::value;
          v[e] = Cplx<T>{v[e].re * gv[e], v[e].im * gv[e]};
        });
        rfft::stockham_stage<T, L, R0, 1, true, false>(v, line, tw, t);
        rfft::stockham_stage_given<T, L, R1, R0, true>(v, line, w2, t);
        rfft::stockham_stage_given<T, L, R2, NS3, true>(v, line, w3, t);
      } else {
        team_fft<T, L, false, true>(v, line, tw, t);
        static_for<0, E>([&](auto e_) {
          constexpr int e = decltype(e_)::value;
          v[e] = Cplx<T>{v[e].re * gv[e], v[e].im * gv[e]};
        });
        team_fft<T, L, true>(v, line, tw, t);
      }
    }
    static_for<0, K>([&](auto e_) {
      constexpr int e = decltype(e_)::value;
      line[t + TT * e] = v[e];
    });
    __syncthreads();
    if (live) {
      const int col = column(j);
      static_for<0, K>([&](auto k_) {
        constexpr int k = decltype(k_)::value;
        const int z = lr + TT * k;
        store_stream(H + col + z * zs, xb[lw * S + z]);
      });
    }
    __syncthreads();
  }
}

// K5: x complex-to-real of two rows per complex inverse FFT, crop to x < nx. The Hermitian
// half spectra A, B (kx <= nx) of rows ra, rb become Z[k] = A[k] + i B[k],
// Z[L-k] = conj(A[k]) + i conj(B[k]); imaginary parts at kx = 0 and nx are dropped
// (irfft semantics).
template <typename T, int L, bool ACC>
__global__ void __launch_bounds__((RowTile<T, L>::NT)) k5_xc2r(SlabGeom g, const Cplx<T>* H, CompPtrsW out,
                                                             const Cplx<T>* tw) {
  using RT = RowTile<T, L>;
  constexpr int E = RT::E, TT = RT::TT;
  extern __shared__ unsigned char smem_raw[];
  Cplx<T>* xb = reinterpret_cast<Cplx<T>*>(smem_raw);
  const int t = threadIdx.x % TT;
  const int team = threadIdx.x / TT;
  const int64_t rows = int64_t(g.ncomp) * g.nz * g.ny;
  const int64_t q = int64_t(blockIdx.x) * RT::NTEAM + team;
  const int64_t ra = 2 * q, rb = 2 * q + 1;
  const int64_t nx = g.nx;
  const Cplx<T>* ha = ra < rows ? H + g.row_h(ra) : nullptr;
  const Cplx<T>* hb = rb < rows ? H + g.row_h(rb) : nullptr;
  Cplx<T>* const line = xb + team * RT::S;
  // the twiddle table in LDS, shared by the workgroup's teams (one barrier, below)
  Cplx<T>* const twl = xb + RT::NTEAM * RT::S;
  rfft::load_tw_lds<T, L>(twl, tw, threadIdx.x, RT::NT);
  for (int k = t; k <= nx; k += TT) {
    Cplx<T> a = {T(0), T(0)}, bb = {T(0), T(0)};
    if (ha) a = ha[k];
    if (hb) bb = hb[k];
    if (k == 0 || k == nx) {
      a.im = T(0);
      bb.im = T(0);
    }
    line[k] = Cplx<T>{a.re - bb.im, a.im + bb.re};
    if (k > 0 && k < nx) line[L - k] = Cplx<T>{a.re + bb.im, bb.re - a.im};
  }
  rfft::wave_sync();
  Cplx<T> v[E];
  static_for<0, E>([&](auto e_) {
    constexpr int e = decltype(e_)::value;
    v[e] = line[t + TT * e];
  });
  __syncthreads();  // the twiddle table is complete
  team_fft<T, L, true, false, true>(v, line, twl, t);
  auto row_ptr = [&](int64_t r) -> T* {
    const int c = int(r / (g.nz * g.ny));
    const int64_t rem = r % (g.nz * g.ny);
    return static_cast<T*>(out.get(c)) + (rem / g.ny) * out.sz + (rem % g.ny) * out.sy;
  };
  T* pa = ra < rows ? row_ptr(ra) : nullptr;
  T* pb = rb < rows ? row_ptr(rb) : nullptr;
  static_for<0, E / 2>([&](auto e_) {
    constexpr int e = decltype(e_)::value;
    const int64_t x = t + TT * e;
    if constexpr (ACC) {  // solution += result (the adjoint adds G psi_bar to omega_bar)
      if (pa) pa[x * out.sx] += v[e].re;
      if (pb) pb[x * out.sx] += v[e].im;
    } else {
      if (pa) __builtin_nontemporal_store(v[e].re, pa + x * out.sx);
      if (pb) __builtin_nontemporal_store(v[e].im, pb + x * out.sx);
    }
  });
}

// ------------------------------------------------------------------ three-pass schedule
// The y transform (length LY = 2 ny = 4 L2) is split by decimation in time, y = 4 y2 + y1 and
// ky = k2 + L2 k1:
//   Y[k2 + L2 k1] = sum_{y1<4} w4^{y1 k1} * [ w_LY^{y1 k2} * sum_{y2} x[4 y2 + y1] w_L2^{y2 k2} ]
// and its inverse the same way with conjugate roots. Pass A (ka_xy_fwd) runs per (c, z, y1):
// the x transforms of the rows y = 4 y2 + y1, then the inner L2-point sums over y2 and the
// twiddle, stored in slab row y1 L2 + k2. The z pass (k3_zconv4) applies the 4-point sums
// over y1 before its convolution and their inverse after it. Pass C (kc_yx_inv) undoes the
// twiddle and the inner sums and runs the inverse x transforms. So the x-pass outputs
// (K1 -> K2) and the inverse y-pass outputs (K4 -> K5) of the five-pass schedule never reach
// HBM: 24 words per cell fewer.

// Elements of a thread's half spectrum per chunk of the y sums: the most, up to 64 / TX, that
// divide EX / 2 and keep the chunk's column teams and staging rows within the workgroup.
constexpr int xy_chunk_elems(int EX, int TX, int T2, int NT, int L2) {
  int best = 1;
  for (int k = 1; k == 1 || k * TX <= 64; ++k) {
    const int cw = k * TX;
    if ((EX / 2) % k == 0 && cw * T2 <= NT && NT % cw == 0 && L2 % (NT / cw) == 0) best = k;
  }
  return best;
}

// A workgroup of pass A or C holds the ny / 4 rows of its (c, z, y1) as row pairs (two real
// rows per complex x transform), PPT pairs per team of TX threads. The y sums run on chunks of
// CW = KC TX columns (kx = CW ch + TX kk + t, element KC ch + kk of every x thread; the last
// chunk is kx = nx alone), one team of T2 threads per column, staged through LDS. Lines of
// fewer than 64 x threads take KC > 1, so a chunk still spans a wavefront's width of columns
// and the workgroup passes fewer chunk barriers.
template <typename T, int LX, int LY>
struct XYPass {
  static constexpr int EX = LinePlan<LX>::E, TX = LinePlan<LX>::T;
  static constexpr int L2 = LY / 4;
  static constexpr int E2 = LinePlan<L2>::E, T2 = LinePlan<L2>::T;
  static constexpr int NY2 = L2 / 2;      // rows of a workgroup (y2 < ny / 4)
  static constexpr int NPAIR = NY2 / 2;
  static constexpr int NTEAM = NPAIR * TX <= 1024 ? NPAIR : 1024 / TX;
  static constexpr int PPT = NPAIR / NTEAM;  // row pairs per x team
  static constexpr int NT = NTEAM * TX;
  static constexpr int KC = xy_chunk_elems(EX, TX, T2, NT, L2);
  static constexpr int CW = KC * TX;         // columns per chunk: kx = CW ch + kk TX + t, kk < KC
  static constexpr int NCH = EX / 2 / KC + 1;  // the last chunk is kx = nx alone
  static constexpr int SX = region_stride<LX, false>();
  static constexpr int S2 = region_stride<L2, true>();
  // staging tile [row][P]: the column teams' accesses (T2 rows of 32 / T2 columns per half
  // wavefront) fall in distinct banks when P = 32 / T2 (mod 32)
  static constexpr int P = (CW + 31) / 32 * 32 + 32 / T2;
  static constexpr int TW2 = rfft::tw_lds_size<L2>();                // padded w_L2 table
  static constexpr int XPH = NTEAM * SX + rfft::tw_lds_size<LX>();  // x regions, x table
  static constexpr int CREG = CW * S2;         // column regions
  static constexpr int STG = L2 * P;           // staging tile
  static constexpr int UNION = XPH > CREG ? (XPH > STG ? XPH : STG) : (CREG > STG ? CREG : STG);
  static constexpr size_t lds = size_t(TW2 + L2 + UNION) * sizeof(Cplx<T>);
  // the staging tile moves to and from HBM as rows of CW columns: thread tid takes column
  // tid % CW of the rows tid / CW + q NT / CW
  static constexpr int RSTEP = NT / CW, NQ = L2 / RSTEP;
  static_assert(LY % 8 == 0 && NPAIR % NTEAM == 0 && NTEAM >= T2 && T2 <= 32, "three-pass geometry");
  static_assert(NT % CW == 0 && L2 % RSTEP == 0 && CW * T2 <= NT, "whole staging rows per thread");
};

// Pass A: x real-to-complex of the rows y = 4 y2 + y1 of (c, z) (two rows per complex FFT,
// separated as in k1_xr2c), then per column kx the zero-padded L2-point FFT over y2 times
// w_LY^{y1 k2}, stored in slab row y1 L2 + k2.
template <typename T, int LX, int LY>
__global__ void __launch_bounds__((XYPass<T, LX, LY>::NT)) ka_xy_fwd(SlabGeom g, CompPtrs in, Cplx<T>* H,
                                                                    const Cplx<T>* twx, const Cplx<T>* twy) {
  using F = XYPass<T, LX, LY>;
  constexpr int EX = F::EX, TX = F::TX, L2 = F::L2, E2 = F::E2, T2 = F::T2, PPT = F::PPT, P = F::P;
  constexpr int CW = F::CW, NT = F::NT;
  extern __shared__ unsigned char smem_raw[];
  Cplx<T>* const tw2 = reinterpret_cast<Cplx<T>*>(smem_raw);  // w_L2^m = w_LY^{4m}, padded
  Cplx<T>* const twr = tw2 + F::TW2;                           // w_LY^{y1 k2}
  Cplx<T>* const U = twr + L2;
  const int tid = threadIdx.x;
  const int y1 = int(blockIdx.x & 3);
  const int cz = int(blockIdx.x >> 2);
  const int c = cz / int(g.nz), z = cz % int(g.nz);
  const int team = tid / TX, t = tid % TX;
  for (int i = tid; i < L2; i += NT) {
    tw2[rfft::tw_slot(i)] = twy[4 * i];  // padded (w_L2^m = w_LY^{4m})
    twr[i] = twy[(y1 * i) % LY];
  }
  Cplx<T>* const twxl = U + F::NTEAM * F::SX;
  rfft::load_tw_lds<T, LX>(twxl, twx, tid, NT);
  const T* base = static_cast<const T*>(in.get(c)) + int64_t(z) * in.sz;
  Cplx<T> v[PPT][EX];
  static_for<0, PPT>([&](auto p_) {
    constexpr int p = decltype(p_)::value;
    const int y2 = 2 * (team * PPT + p);
    const T* ra = base + int64_t(4 * y2 + y1) * in.sy + int64_t(t) * in.sx;
    const T* rb = ra + 4 * in.sy;
    static_for<0, EX>([&](auto e_) {
      constexpr int e = decltype(e_)::value;
      if constexpr (e < EX / 2) {
        v[p][e] = Cplx<T>{ra[int64_t(TX * e) * in.sx], rb[int64_t(TX * e) * in.sx]};
      } else {
        v[p][e] = Cplx<T>{T(0), T(0)};
      }
    });
  });
  __syncthreads();  // tables complete
  Cplx<T>* const xline = U + team * F::SX;
  // half spectra kx = TX e + t (e <= EX / 2; e = EX / 2 is kx = nx, lane 0 only)
  Cplx<T> A[PPT][EX / 2 + 1], B[PPT][EX / 2 + 1];
  const int mirror_lane = (TX - t) % TX;
  static_for<0, PPT>([&](auto p_) {
    constexpr int p = decltype(p_)::value;
    rfft::team_fft<T, LX, false, true, true>(v[p], xline, twxl, t);
    static_for<0, EX / 2 + 1>([&](auto e_) {
      constexpr int e = decltype(e_)::value;
      // Z[LX - k], k = TX e + t: lane TX - t, element EX - 1 - e (t > 0); lane 0: EX - e
      const Cplx<T> zk = v[p][e];
      Cplx<T> zm;
      zm.re = __shfl(v[p][EX - 1 - e].re, mirror_lane, TX);
      zm.im = __shfl(v[p][EX - 1 - e].im, mirror_lane, TX);
      if (t == 0) zm = v[p][(EX - e) % EX];
      A[p][e] = Cplx<T>{T(0.5) * (zk.re + zm.re), T(0.5) * (zk.im - zm.im)};
      B[p][e] = Cplx<T>{T(0.5) * (zk.im + zm.im), T(0.5) * (zm.re - zk.re)};
    });
  });
  __syncthreads();  // x regions free: the chunk phase reuses the area
  const int col = tid / T2, t2 = tid % T2;
  const bool col_role = tid < CW * T2;
  Cplx<T>* const creg = U + col * F::S2;
  // staging rows to HBM: column sc of rows sr + q RSTEP
  const int sc = tid % CW, sr = tid / CW;
  static_for<0, F::NCH>([&](auto ch_) {
    constexpr int ch = decltype(ch_)::value;
    constexpr bool last = ch == F::NCH - 1;  // kx = nx alone
    if constexpr (last) {
      if (t == 0) {
        static_for<0, PPT>([&](auto p_) {
          constexpr int p = decltype(p_)::value;
          const int y2 = 2 * (team * PPT + p);
          U[y2 * P] = A[p][EX / 2];
          U[(y2 + 1) * P] = B[p][EX / 2];
        });
      }
    } else {
      static_for<0, F::KC>([&](auto k_) {
        constexpr int kk = decltype(k_)::value;
        static_for<0, PPT>([&](auto p_) {
          constexpr int p = decltype(p_)::value;
          const int y2 = 2 * (team * PPT + p);
          U[y2 * P + kk * TX + t] = A[p][ch * F::KC + kk];
          U[(y2 + 1) * P + kk * TX + t] = B[p][ch * F::KC + kk];
        });
      });
    }
    __syncthreads();
    const bool live = col_role && (!last || col == 0);
    Cplx<T> w[E2];
    if (live) {
      static_for<0, E2>([&](auto e_) {
        constexpr int e = decltype(e_)::value;
        if constexpr (e < E2 / 2) w[e] = U[(t2 + T2 * e) * P + col];
        else w[e] = Cplx<T>{T(0), T(0)};
      });
    }
    __syncthreads();
    if (live) {
      rfft::team_fft<T, L2, false, true, true>(w, creg, tw2, t2);
      static_for<0, E2>([&](auto e_) {
        constexpr int e = decltype(e_)::value;
        w[e] = cmul(w[e], twr[t2 + T2 * e]);
      });
    }
    __syncthreads();
    if (live) {
      static_for<0, E2>([&](auto e_) {
        constexpr int e = decltype(e_)::value;
        U[(t2 + T2 * e) * P + col] = w[e];
      });
    }
    __syncthreads();
    if (!last || sc == 0) {
      static_for<0, F::NQ>([&](auto q_) {
        constexpr int q = decltype(q_)::value;
        store_stream(H + g.ti(c, z, y1 * L2 + sr + q * F::RSTEP, CW * ch + sc), U[(sr + q * F::RSTEP) * P + sc]);
      });
    }
    __syncthreads();
  });
}

// Pass C: per column kx, w_LY^{-y1 k2} times the slab rows y1 L2 + k2, the inverse L2-point
// FFT over k2 (outputs y2 < ny / 4), then the Hermitian x half spectra of each row pair to one
// complex inverse FFT (as in k5_xc2r), rows y = 4 y2 + y1 stored (ACC: added). The next
// chunk's rows are loaded while the current one is transformed.
template <typename T, int LX, int LY, bool ACC>
__global__ void __launch_bounds__((XYPass<T, LX, LY>::NT)) kc_yx_inv(SlabGeom g, const Cplx<T>* H, CompPtrsW out,
                                                                    const Cplx<T>* twx, const Cplx<T>* twy) {
  using F = XYPass<T, LX, LY>;
  constexpr int EX = F::EX, TX = F::TX, L2 = F::L2, E2 = F::E2, T2 = F::T2, PPT = F::PPT, P = F::P;
  constexpr int CW = F::CW, NT = F::NT;
  extern __shared__ unsigned char smem_raw[];
  Cplx<T>* const tw2 = reinterpret_cast<Cplx<T>*>(smem_raw);
  Cplx<T>* const twr = tw2 + F::TW2;
  Cplx<T>* const U = twr + L2;
  const int tid = threadIdx.x;
  const int y1 = int(blockIdx.x & 3);
  const int cz = int(blockIdx.x >> 2);
  const int c = cz / int(g.nz), z = cz % int(g.nz);
  const int team = tid / TX, t = tid % TX;
  for (int i = tid; i < L2; i += NT) {
    tw2[rfft::tw_slot(i)] = twy[4 * i];  // padded (w_L2^m = w_LY^{4m})
    twr[i] = twy[(y1 * i) % LY];
  }
  const int col = tid / T2, t2 = tid % T2;
  const bool col_role = tid < CW * T2;
  Cplx<T>* const creg = U + col * F::S2;
  const int mirror_lane = (TX - t) % TX;
  // the rows of a chunk: thread tid holds elements tid + NT q of the chunk's L2 x ncol block
  // (addresses formed per load: a hoisted row pointer costs the registers of the prefetch)
  Cplx<T> pre[F::NQ];
  const int hsrc = g.ti(c, z, y1 * L2 + tid / CW, tid % CW);
  const int rowel = int(g.rs) * int(g.nz);  // elements per slab row
  auto load_chunk = [&](auto ch_) {
    constexpr int ch = decltype(ch_)::value;
    constexpr int ncol = ch == F::NCH - 1 ? 1 : CW;
    static_for<0, F::NQ>([&](auto q_) {
      constexpr int q = decltype(q_)::value;
      const int i = tid + q * NT;
      if constexpr (ncol == CW && CW % kSlabTile == 0) {
        // row i / CW = tid / CW + q NT / CW, column CW ch + tid % CW
        pre[q] = H[hsrc + q * (NT / CW) * rowel + CW * ch * int(g.nz)];
      } else {
        if (i < L2 * ncol) pre[q] = H[g.ti(c, z, y1 * L2 + i / ncol, CW * ch + i % ncol)];
      }
    });
  };
  load_chunk(std::integral_constant<int, 0>{});
  Cplx<T> v[PPT][EX];
  static_for<0, F::NCH>([&](auto ch_) {
    constexpr int ch = decltype(ch_)::value;
    constexpr bool last = ch == F::NCH - 1;
    constexpr int ncol = last ? 1 : CW;
    static_for<0, F::NQ>([&](auto q_) {
      constexpr int q = decltype(q_)::value;
      const int i = tid + q * NT;
      if (i < L2 * ncol) U[(i / ncol) * P + i % ncol] = pre[q];
    });
    if constexpr (ch + 1 < F::NCH) load_chunk(std::integral_constant<int, ch + 1>{});
    __syncthreads();  // (first chunk: the tables too)
    const bool live = col_role && (!last || col == 0);
    Cplx<T> w[E2];
    if (live) {
      static_for<0, E2>([&](auto e_) {
        constexpr int e = decltype(e_)::value;
        w[e] = cmulc(U[(t2 + T2 * e) * P + col], twr[t2 + T2 * e]);
      });
    }
    __syncthreads();
    if (live) rfft::team_fft<T, L2, true, false, true>(w, creg, tw2, t2);
    __syncthreads();
    if (live) {
      static_for<0, E2 / 2>([&](auto e_) {
        constexpr int e = decltype(e_)::value;
        U[(t2 + T2 * e) * P + col] = w[e];
      });
    }
    __syncthreads();
    static_for<0, (last ? 1 : F::KC)>([&](auto k_) {
      constexpr int kk = decltype(k_)::value;
      constexpr int e = last ? EX / 2 : ch * F::KC + kk;  // the element of kx = TX e + t
      static_for<0, PPT>([&](auto p_) {
        constexpr int p = decltype(p_)::value;
        const int y2 = 2 * (team * PPT + p);
        Cplx<T> a = U[y2 * P + kk * TX + t], b = U[(y2 + 1) * P + kk * TX + t];
        if ((e == 0 && t == 0) || last) {  // kx = 0 and kx = nx: imaginary parts dropped
          a.im = T(0);
          b.im = T(0);
        }
        const Cplx<T> zk = {a.re - b.im, a.im + b.re};  // Z[k] = A + i B
        const Cplx<T> zm = {a.re + b.im, b.re - a.im};  // Z[LX - k] = conj(A) + i conj(B)
        if constexpr (last) {
          if (t == 0) v[p][EX / 2] = zk;
        } else {
          Cplx<T> r;
          r.re = __shfl(zm.re, mirror_lane, TX);
          r.im = __shfl(zm.im, mirror_lane, TX);
          v[p][e] = zk;
          // Z[LX - k'] of lane TX - t (k' = TX e + TX - t) is element EX - 1 - e of lane t;
          // lane 0 keeps its own mirror, element EX - e
          if (t != 0) v[p][EX - 1 - e] = r;
          if constexpr (e > 0) {
            if (t == 0) v[p][EX - e] = zm;
          }
        }
      });
    });
    __syncthreads();
  });
  Cplx<T>* const twxl = U + F::NTEAM * F::SX;
  rfft::load_tw_lds<T, LX>(twxl, twx, tid, NT);
  __syncthreads();
  Cplx<T>* const xline = U + team * F::SX;
  T* const base = static_cast<T*>(out.get(c)) + int64_t(z) * out.sz;
  static_for<0, PPT>([&](auto p_) {
    constexpr int p = decltype(p_)::value;
    rfft::team_fft<T, LX, true, false, true>(v[p], xline, twxl, t);
    const int y2 = 2 * (team * PPT + p);
    T* const ra = base + int64_t(4 * y2 + y1) * out.sy + int64_t(t) * out.sx;
    T* const rb = ra + 4 * out.sy;
    static_for<0, EX / 2>([&](auto e_) {
      constexpr int e = decltype(e_)::value;
      const int64_t x = int64_t(TX * e) * out.sx;
      if constexpr (ACC) {
        ra[x] += v[p][e].re;
        rb[x] += v[p][e].im;
      } else {
        __builtin_nontemporal_store(v[p][e].re, ra + x);
        __builtin_nontemporal_store(v[p][e].im, rb + x);
      }
    });
  });
}

// Build-time choices of the three-pass z pass (measured: ordinary stores beat streaming ones
// there, quarter-line tiles beat half-line ones, six jobs per workgroup beat two or three).
#ifndef SOPHT_ZROW_STREAM
#define SOPHT_ZROW_STREAM 0
#endif
#ifndef SOPHT_ZROW_PARTS
#define SOPHT_ZROW_PARTS 4
#endif
#ifndef SOPHT_ZROW_JOBS
#define SOPHT_ZROW_JOBS 6
#endif
#ifndef SOPHT_ZROW_SLOTS
#define SOPHT_ZROW_SLOTS 2
#endif
#ifndef SOPHT_ZROW_PREFETCH
#define SOPHT_ZROW_PREFETCH 1
#endif
constexpr bool kZrowStream = SOPHT_ZROW_STREAM;
constexpr bool kZrowPrefetch = SOPHT_ZROW_PREFETCH;
constexpr int kZrowJobs = SOPHT_ZROW_JOBS;

// The z pass of the three-pass schedule. A workgroup owns one kx tile (WC columns) of the row
// class k2 (slab rows y1 L2 + k2, y1 < 4) and of its mirror class L2 - k2. Per job (component,
// class) the four rows are loaded, the 4-point DFT over y1 gives the pencils ky = k2 + L2 k1,
// each pencil is convolved along z as in k3_zconv, and the inverse 4-point DFT returns the rows.
// The workgroup has two teams per column and transforms two pencils at a time (two LDS tiles),
// so a thread holds E / 4 values of each row: the next job's rows are loaded during the
// transforms within 128 VGPRs. The G^ lines of the class, |k2 + L2 k1| folded to <= ny for
// k1 < 4, are staged in LDS once per workgroup. Slot s of a column takes pencil k1 = s + 2r in
// round r of a job, and so G^ line s + 2r; in a mirror job it takes pencil 3 - s - 2r, whose
// ky' = L2 - k2 + L2 (3 - s - 2r) has the line of 2 ny - ky' = k2 + L2 (s + 2r): the same.
template <typename T, int L>
struct ZRowTile {
  static constexpr int E = LinePlan<L>::E, TT = LinePlan<L>::T;
  // a quarter of a 128-byte line per workgroup (two workgroups per CU); the four quarters of a
  // line go to workgroups b, b + 8, b + 16, b + 24 (the same XCD), so three come from its L2
  static constexpr int PARTS = SOPHT_ZROW_PARTS;
  static constexpr int SLOTS = SOPHT_ZROW_SLOTS;  // pencils per column transformed at once
  static constexpr int WC = kLineCplx<T> / PARTS;
  static constexpr int NT = SLOTS * WC * TT;
  // tiles [slot][column]: the load role's lanes (column, z) hit distinct banks when the
  // column stride is 8 (mod 32) slots
  static constexpr int S = TT < 32 ? L + 16 : (WC == 4 ? L + 8 : L + 4);
  static constexpr int KL = E / (2 * SLOTS);  // z values per row per thread (z = lr + SLOTS TT k)
  static constexpr int GL = L / 2 + 1;
  static constexpr size_t lds = size_t(SLOTS * WC * S) * sizeof(Cplx<T>) + size_t(4 * WC * GL) * sizeof(T);
  static_assert(E % (2 * SLOTS) == 0 && 4 % SLOTS == 0, "the z rows split among the slots");
};

template <typename T, int L>
__global__ void __launch_bounds__((ZRowTile<T, L>::NT)) __attribute__((amdgpu_waves_per_eu(kZconvWavesFor<T, L>))) k3_zconv4(SlabGeom g, Cplx<T>* H, const T* ghat,
                                                                  const Cplx<T>* tw, int jobs) {
  using ZT = ZRowTile<T, L>;
  constexpr int E = ZT::E, TT = ZT::TT, WC = ZT::WC, S = ZT::S, KL = ZT::KL, GL = ZT::GL;
  constexpr int SL = ZT::SLOTS, ZSTEP = SL * TT;
  extern __shared__ unsigned char smem_raw[];
  Cplx<T>* xb = reinterpret_cast<Cplx<T>*>(smem_raw);  // tiles [slot][column][z]
  T* const gl = reinterpret_cast<T*>(xb + SL * WC * S);  // G^ lines [k1][column][kz]
  const int tid = threadIdx.x;
  const int lw = tid % WC, lr = tid / WC;  // load role: column, first z
  const int team = tid / TT, t = tid % TT;  // transform role: team = slot * WC + column
  const int tcol = team % WC, slot = team / WC;
  const int64_t ncols = g.nx + 1;
  const int L2 = int(g.ny) / 2;
  int64_t k2, kx0;
  if (!col_tile_parts<ZT::PARTS>(blockIdx.x, ncols, L2 / 2 + 1, WC, k2, kx0)) return;
  const bool self = k2 == 0 || 2 * k2 == L2;
  const int alljobs = self ? g.ncomp : 2 * g.ncomp;
  const int jbeg = int(blockIdx.y) * jobs;
  if (jbeg >= alljobs) return;
  const int njobs = alljobs - jbeg < jobs ? alljobs - jbeg : jobs;
  const int64_t kx = kx0 + lw;
  const bool live = kx < ncols;
  const int first_col = 64 / TT >= WC ? 0 : ((tid / 64) * (64 / TT)) % WC;
  const bool wave_live = kx0 + first_col < ncols;
  {
    // every load in flight before the first LDS write (a loop would wait on each in turn)
    constexpr int NG = 4 * WC * GL, NIT = (NG + ZT::NT - 1) / ZT::NT;
    T gt[NIT];
    static_for<0, NIT>([&](auto i_) {
      constexpr int it = decltype(i_)::value;
      const int i = tid + it * ZT::NT;
      const int kz = i % GL, rest = i / GL;
      const int w = rest % WC, k1 = (rest / WC) & 3;
      const int ky = int(k2) + L2 * k1;
      const int kyo = ky <= int(g.ny) ? ky : 2 * int(g.ny) - ky;
      const int kxt = kx0 + w < ncols ? int(kx0) + w : 0;
      gt[it] = (it + 1 < NIT || i < NG) ? ghat[(kyo * int(ncols) + kxt) * int(g.gz) + kz] : T(0);
    });
    static_for<0, NIT>([&](auto i_) {
      constexpr int it = decltype(i_)::value;
      const int i = tid + it * ZT::NT;
      if (it + 1 < NIT || i < NG) gl[i] = gt[it];
    });
  }
  constexpr int zs = kSlabTile;  // consecutive z of a column tile (ti layout)
  using P3 = Plan3<typename LinePlan<L>::R>;
  constexpr int R1 = at_least_1(P3::r1), R2 = at_least_1(P3::r2);
  constexpr int PER2 = at_least_1(E / R1), PER3 = at_least_1(E / R2);
  Cplx<T> w2[PER2][R1], w3[PER3][R2];
  if constexpr (last_twiddles_hoisted<L>()) {
    static_for<0, PER2>([&](auto p_) {
      constexpr int p = decltype(p_)::value;
      rfft::stage_twiddles<T, L, R1, P3::r0>(w2[p], tw, (t + p * TT) % P3::r0);
    });
    static_for<0, PER3>([&](auto p_) {
      constexpr int p = decltype(p_)::value;
      rfft::stage_twiddles<T, L, R2, P3::r0 * R1>(w3[p], tw, (t + p * TT) % (P3::r0 * R1));
    });
  }
  auto row_col = [&](int jl, int y1) -> int {
    const int j = jbeg + jl;
    const int cc = j % g.ncomp;
    const int cls = j < g.ncomp ? int(k2) : L2 - int(k2);
    return g.ti(cc, 0, y1 * L2 + cls, live ? int(kx) : 0);
  };
  Cplx<T> pf[4][KL];
  auto load = [&](int j) {
    static_for<0, 4>([&](auto y_) {
      constexpr int y1 = decltype(y_)::value;
      const int colp = row_col(j, y1);
      static_for<0, KL>([&](auto k_) {
        constexpr int k = decltype(k_)::value;
        pf[y1][k] = live ? H[colp + (lr + ZSTEP * k) * zs] : Cplx<T>{T(0), T(0)};
      });
    });
  };
  Cplx<T>* const line = xb + team * S;
  if constexpr (kZrowPrefetch) load(0);
  __syncthreads();  // G^ lines staged
  for (int j = 0; j < njobs; ++j) {
    const bool mirror = jbeg + j >= g.ncomp;
    if constexpr (!kZrowPrefetch) load(j);
    Cplx<T> y[4][KL];
    static_for<0, KL>([&](auto k_) {
      constexpr int k = decltype(k_)::value;
      Cplx<T> u[4] = {pf[0][k], pf[1][k], pf[2][k], pf[3][k]};
      rfft::dft4<T, false>(u);
      static_for<0, 4>([&](auto q_) { y[decltype(q_)::value][k] = u[decltype(q_)::value]; });
    });
    if constexpr (kZrowPrefetch) {
      if (j + 1 < njobs) load(j + 1);
    }
    static_for<0, 4 / SL>([&](auto r_) {
      constexpr int r = decltype(r_)::value;
      // slot sl takes pencil sl + SL r (a mirror job: 3 - sl - SL r)
      static_for<0, SL>([&](auto s_) {
        constexpr int sl = decltype(s_)::value;
        static_for<0, KL>([&](auto k_) {
          constexpr int k = decltype(k_)::value;
          // selects on values (a select of two array elements' addresses would leave y in
          // scratch memory)
          const Cplx<T> a = y[3 - sl - SL * r][k], b = y[sl + SL * r][k];
          xb[(sl * WC + lw) * S + lr + ZSTEP * k] = Cplx<T>{mirror ? a.re : b.re, mirror ? a.im : b.im};
        });
      });
      __syncthreads();
      Cplx<T> v[E];
      static_for<0, E>([&](auto e_) {
        constexpr int e = decltype(e_)::value;
        if constexpr (e < E / 2) v[e] = line[t + TT * e];
        else v[e] = Cplx<T>{T(0), T(0)};
      });
      rfft::wave_sync();
      if (wave_live) {
        const T* const gq = gl + ((slot + SL * r) * WC + tcol) * GL;
        if constexpr (last_twiddles_hoisted<L>()) {
          constexpr int R0 = P3::r0, NS3 = R0 * R1;
          rfft::stockham_stage<T, L, R0, 1, false, true>(v, line, tw, t);
          rfft::stockham_stage_given<T, L, R1, R0, false>(v, line, w2, t);
          rfft::stockham_stage_given<T, L, R2, NS3, false>(v, line, w3, t);
        } else {
          rfft::team_fft<T, L, false, true>(v, line, tw, t);
        }
        static_for<0, E>([&](auto e_) {
          constexpr int e = decltype(e_)::value;
          const int kz = t + TT * e;
          const T gv = gq[kz <= L / 2 ? kz : L - kz];
          v[e] = Cplx<T>{v[e].re * gv, v[e].im * gv};
        });
        if constexpr (last_twiddles_hoisted<L>()) {
          constexpr int R0 = P3::r0, NS3 = R0 * R1;
          rfft::stockham_stage<T, L, R0, 1, true, false>(v, line, tw, t);
          rfft::stockham_stage_given<T, L, R1, R0, true>(v, line, w2, t);
          rfft::stockham_stage_given<T, L, R2, NS3, true>(v, line, w3, t);
        } else {
          rfft::team_fft<T, L, true>(v, line, tw, t);
        }
      }
      static_for<0, E / 2>([&](auto e_) {
        constexpr int e = decltype(e_)::value;
        line[t + TT * e] = v[e];
      });
      __syncthreads();
      static_for<0, SL>([&](auto s_) {
        constexpr int sl = decltype(s_)::value;
        static_for<0, KL>([&](auto k_) {
          constexpr int k = decltype(k_)::value;
          const Cplx<T> val = xb[(sl * WC + lw) * S + lr + ZSTEP * k];
          Cplx<T>& a = y[3 - sl - SL * r][k];
          Cplx<T>& b = y[sl + SL * r][k];
          a = Cplx<T>{mirror ? val.re : a.re, mirror ? val.im : a.im};
          b = Cplx<T>{mirror ? b.re : val.re, mirror ? b.im : val.im};
        });
      });
      __syncthreads();
    });
    if (live) {
      static_for<0, KL>([&](auto k_) {
        constexpr int k = decltype(k_)::value;
        Cplx<T> u[4] = {y[0][k], y[1][k], y[2][k], y[3][k]};
        rfft::dft4<T, true>(u);
        static_for<0, 4>([&](auto q_) {
          constexpr int y1 = decltype(q_)::value;
          if constexpr (kZrowStream) store_stream(H + row_col(j, y1) + (lr + ZSTEP * k) * zs, u[y1]);
          else H[row_col(j, y1) + (lr + ZSTEP * k) * zs] = u[y1];
        });
      });
    }
  }
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

// scaled, converted to real_t and stored [y][x][z] (z fastest, line stride gz)
template <typename T>
__global__ void k_ghat_store(const double* in, T* out, int64_t nz1, int64_t ny1, int64_t nx1,
                             int64_t gz, double s) {
  const int64_t n = nz1 * ny1 * nx1;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t z = i % nz1, yx = i / nz1;  // output order: consecutive threads, consecutive z
    const int64_t x = yx % nx1, y = yx / nx1;
    out[yx * gz + z] = T(in[(z * ny1 + y) * nx1 + x] * s);
  }
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
    using RT = RowTile<T, n>;                                                         \
    k1_xr2c<T, n><<<ceil_div(pairs, RT::NTEAM), RT::NT, RT::lds + rfft::tw_lds_size<n>() * sizeof(Cplx<T>), s>>>(g, in, H, tw); \
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
    using CT = ColTile<T, n>;                                                         \
    const int64_t nb = col_tile_blocks(g.nx + 1, int64_t(g.ncomp) * g.nz, CT::WC);    \
    k_ycols<T, n, INV><<<nb, CT::NT, CT::lds + rfft::tw_lds_size<n>() * sizeof(Cplx<T>), s>>>(g, H, tw);   \
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
    using CT = ColTile<T, n>;                                                         \
    const int64_t nb = col_tile_blocks(g.nx + 1, g.ny + 1, CT::WC);                   \
    const unsigned nch = unsigned((2 * g.ncomp + kZconvJobs - 1) / kZconvJobs);        \
    k3_zconv<T, n><<<dim3(unsigned(nb), nch), CT::NT, CT::lds, s>>>(g, H, ghat, tw);   \
    SOPHT_CHECK_LAUNCH();                                                             \
    return SOPHT_OK;                                                                  \
  }
  SOPHT_FFT_LENGTHS(X)
#undef X
  return SOPHT_ERR_UNSUPPORTED;
}

template <typename T>
int launch_k5(int64_t L, SlabGeom g, const Cplx<T>* H, CompPtrsW out, const Cplx<T>* tw,
              hipStream_t s, bool acc) {
  const int64_t rows = int64_t(g.ncomp) * g.nz * g.ny;
  const int64_t pairs = (rows + 1) / 2;
#define X(n)                                                                          \
  if (L == n) {                                                                       \
    using RT = RowTile<T, n>;                                                         \
    if (acc)                                                                          \
      k5_xc2r<T, n, true><<<ceil_div(pairs, RT::NTEAM), RT::NT, RT::lds + rfft::tw_lds_size<n>() * sizeof(Cplx<T>), s>>>(g, H, out, tw); \
    else                                                                              \
      k5_xc2r<T, n, false><<<ceil_div(pairs, RT::NTEAM), RT::NT, RT::lds + rfft::tw_lds_size<n>() * sizeof(Cplx<T>), s>>>(g, H, out, tw); \
    SOPHT_CHECK_LAUNCH();                                                             \
    return SOPHT_OK;                                                                  \
  }
  SOPHT_FFT_LENGTHS(X)
#undef X
  return SOPHT_ERR_UNSUPPORTED;
}

// Line lengths 2 nx = 2 ny of the three-pass schedule, and 2 nz (longer z lines leave the
// z pass without registers for its next rows).
#define SOPHT_XY_LENGTHS(X) X(96) X(128) X(192) X(256) X(384) X(512)
#define SOPHT_ZROW_LENGTHS(X) \
  X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

bool zrow_length_supported(int64_t L) {
#define X(n) if (L == n) return true;
  SOPHT_ZROW_LENGTHS(X)
#undef X
  return false;
}

bool xy_length_supported(int64_t L) {
#define X(n) if (L == n) return true;
  SOPHT_XY_LENGTHS(X)
#undef X
  return false;
}

// Workgroups above 64 KB of LDS declare it once per kernel.
template <typename K>
int allow_lds(K* kernel, size_t bytes) {
  if (bytes <= 65536) return SOPHT_OK;
  SOPHT_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
  return SOPHT_OK;
}

template <typename T>
int launch_xy_fwd(int64_t L, SlabGeom g, CompPtrs in, Cplx<T>* H, const Cplx<T>* twx,
                  const Cplx<T>* twy, hipStream_t s) {
  const unsigned blocks = unsigned(int64_t(g.ncomp) * g.nz * 4);
#define X(n)                                                                          \
  if (L == n) {                                                                       \
    using F = XYPass<T, n, n>;                                                        \
    static const int lds_ok = allow_lds(ka_xy_fwd<T, n, n>, F::lds);                  \
    if (lds_ok) return lds_ok;                                                        \
    ka_xy_fwd<T, n, n><<<blocks, F::NT, F::lds, s>>>(g, in, H, twx, twy);             \
    SOPHT_CHECK_LAUNCH();                                                             \
    return SOPHT_OK;                                                                  \
  }
  SOPHT_XY_LENGTHS(X)
#undef X
  return SOPHT_ERR_UNSUPPORTED;
}

template <typename T>
int launch_yx_inv(int64_t L, SlabGeom g, const Cplx<T>* H, CompPtrsW out, const Cplx<T>* twx,
                  const Cplx<T>* twy, hipStream_t s, bool acc) {
  const unsigned blocks = unsigned(int64_t(g.ncomp) * g.nz * 4);
#define X(n)                                                                          \
  if (L == n) {                                                                       \
    using F = XYPass<T, n, n>;                                                        \
    if (acc) {                                                                        \
      static const int lds_ok = allow_lds(kc_yx_inv<T, n, n, true>, F::lds);          \
      if (lds_ok) return lds_ok;                                                      \
      kc_yx_inv<T, n, n, true><<<blocks, F::NT, F::lds, s>>>(g, H, out, twx, twy);    \
    } else {                                                                          \
      static const int lds_ok = allow_lds(kc_yx_inv<T, n, n, false>, F::lds);         \
      if (lds_ok) return lds_ok;                                                      \
      kc_yx_inv<T, n, n, false><<<blocks, F::NT, F::lds, s>>>(g, H, out, twx, twy);   \
    }                                                                                 \
    SOPHT_CHECK_LAUNCH();                                                             \
    return SOPHT_OK;                                                                  \
  }
  SOPHT_XY_LENGTHS(X)
#undef X
  return SOPHT_ERR_UNSUPPORTED;
}

template <typename T>
int launch_k3_rows(int64_t L, SlabGeom g, Cplx<T>* H, const T* ghat, const Cplx<T>* tw,
                   hipStream_t s, int jobs) {
#define X(n)                                                                          \
  if (L == n) {                                                                       \
    using ZT = ZRowTile<T, n>;                                                        \
    const int64_t nb = col_tile_parts_blocks(ZT::PARTS, g.nx + 1, g.ny / 4 + 1, ZT::WC); \
    const unsigned nch = unsigned((2 * g.ncomp + jobs - 1) / jobs);                    \
    static const int lds_ok = allow_lds(k3_zconv4<T, n>, ZT::lds);                    \
    if (lds_ok) return lds_ok;                                                        \
    k3_zconv4<T, n><<<dim3(unsigned(nb), nch), ZT::NT, ZT::lds, s>>>(g, H, ghat, tw, jobs); \
    SOPHT_CHECK_LAUNCH();                                                             \
    return SOPHT_OK;                                                                  \
  }
  SOPHT_ZROW_LENGTHS(X)
#undef X
  return SOPHT_ERR_UNSUPPORTED;
}

template <typename T>
SlabGeom slab_geom(const sopht_poisson3d* P, int ncomp) {
  return SlabGeom{P->nz, P->ny, P->nx, kRowPad ? round_up(P->nx + 1, kLineCplx<T>) : P->nx + 1,
                  P->gz, ncomp};
}

// ncomp <= 3 components by value (sol, rhs), or any number from device pointer tables
// (sol_tab, rhs_tab; the slab must hold ncomp components).
template <typename T>
int solve_impl(sopht_poisson3d* P, int ncomp, void* const sol[3], const int64_t* ss,
               const void* const rhs[3], const int64_t* rs, hipStream_t s,
               StageTimer* timer = nullptr, int first_stage = 0, void* const* sol_tab = nullptr,
               const void* const* rhs_tab = nullptr, bool accumulate = false) {
  if (ncomp > P->slab_comps) {
    set_last_error("more components than the Poisson plan's slab holds", __FILE__, __LINE__);
    return SOPHT_ERR_INVALID;
  }
  const SlabGeom g = slab_geom<T>(P, ncomp);
  CompPtrs in;
  CompPtrsW out;
  in.tab = rhs_tab;
  out.tab = sol_tab;
  for (int c = 0; c < 3; ++c) {
    in.p[c] = (!rhs_tab && c < ncomp) ? rhs[c] : nullptr;
    out.p[c] = (!sol_tab && c < ncomp) ? sol[c] : nullptr;
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
  if constexpr (sizeof(T) == 4) {  // the three-pass kernels are built for FP32
    if (P->passes == 3) {
      if ((rc = launch_xy_fwd<T>(2 * P->nx, g, in, H, twx, twy, s))) return rc;
      mark(0);
      if ((rc = launch_k3_rows<T>(2 * P->nz, g, H, static_cast<const T*>(P->ghat), twz, s, P->zjobs)))
        return rc;
      mark(1);
      rc = launch_yx_inv<T>(2 * P->nx, g, H, out, twx, twy, s, accumulate);
      mark(2);
      return rc;
    }
  }
  if ((rc = launch_k1<T>(2 * P->nx, g, in, H, twx, s))) return rc;
  mark(0);
  if ((rc = launch_ycols<T, false>(2 * P->ny, g, H, twy, s))) return rc;
  mark(1);
  if ((rc = launch_k3<T>(2 * P->nz, g, H, static_cast<const T*>(P->ghat), twz, s))) return rc;
  mark(2);
  if ((rc = launch_ycols<T, true>(2 * P->ny, g, H, twy, s))) return rc;
  mark(3);
  rc = launch_k5<T>(2 * P->nx, g, H, out, twx, s, accumulate);
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
    k_ghat_store<float><<<blocks, 256, 0, s>>>(src, static_cast<float*>(P->ghat), nz1, ny1, nx1, P->gz, scale);
  else
    k_ghat_store<double><<<blocks, 256, 0, s>>>(src, static_cast<double*>(P->ghat), nz1, ny1, nx1, P->gz, scale);
  SOPHT_CHECK_LAUNCH();
  SOPHT_CHECK_HIP(hipStreamSynchronize(s));
  free(host);
  SOPHT_CHECK_HIP(hipFree(a));
  SOPHT_CHECK_HIP(hipFree(b));
  SOPHT_CHECK_HIP(hipFree(ct));
  return SOPHT_OK;
}

// Jobs per workgroup of the three-pass z pass: the fewer, the more workgroups (and the more
// G^ staging). Each candidate is timed on the plan's own slab (zeroed) and stream, and the
// fastest kept; SOPHT_POISSON_ZJOBS in the environment fixes the choice instead. Results do
// not depend on it: a pencil's arithmetic is the same in any workgroup.
int tune_zpass(sopht_poisson3d* P, hipStream_t s) {
  P->zjobs = kZrowJobs;
  if (P->passes != 3 || P->dtype != SOPHT_F32) return SOPHT_OK;
  if (const char* forced = std::getenv("SOPHT_POISSON_ZJOBS")) {
    const int j = std::atoi(forced);
    if (j >= 1) P->zjobs = j;
    return SOPHT_OK;
  }
  const SlabGeom g = slab_geom<float>(P, P->slab_comps);
  Cplx<float>* H = static_cast<Cplx<float>*>(P->slab);
  const float* gh = static_cast<const float*>(P->ghat);
  const Cplx<float>* twz = static_cast<const Cplx<float>*>(P->tw[2]);
  SOPHT_CHECK_HIP(hipMemsetAsync(P->slab, 0, size_t(P->slab_bytes), s));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  SOPHT_CHECK_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    set_last_error("hipEventCreate failed", __FILE__, __LINE__);
    return SOPHT_ERR_HIP;
  }
  int rc = SOPHT_OK;
  float best = -1.f;
  for (const int jobs : {1, 2, 3, 6}) {
    if (jobs > 2 * g.ncomp) continue;
    if ((rc = launch_k3_rows<float>(2 * P->nz, g, H, gh, twz, s, jobs))) break;  // warm-up
    if (hipEventRecord(e0, s) != hipSuccess) { rc = SOPHT_ERR_HIP; break; }
    for (int r = 0; r < 3 && rc == SOPHT_OK; ++r) rc = launch_k3_rows<float>(2 * P->nz, g, H, gh, twz, s, jobs);
    if (rc) break;
    float ms = 0.f;
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      rc = SOPHT_ERR_HIP;
      break;
    }
    if (best < 0.f || ms < best) {
      best = ms;
      P->zjobs = jobs;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc == SOPHT_ERR_HIP) set_last_error("timing the z pass at plan creation failed", __FILE__, __LINE__);
  return rc;
}

}  // namespace

int poisson_stage_count(const sopht_poisson3d* P) { return P->passes; }

int poisson_vector_solve_timed(sopht_poisson3d* P, void* const sol[3], const int64_t* ss,
                               const void* const rhs[3], const int64_t* rs, hipStream_t s,
                               StageTimer* timer, int first_stage) {
  return P->dtype == SOPHT_F32 ? solve_impl<float>(P, 3, sol, ss, rhs, rs, s, timer, first_stage)
                               : solve_impl<double>(P, 3, sol, ss, rhs, rs, s, timer, first_stage);
}

int poisson_vector_solve_accumulate(sopht_poisson3d* P, void* const sol[3], const int64_t* ss,
                                    const void* const rhs[3], const int64_t* rs, hipStream_t s) {
  return P->dtype == SOPHT_F32
             ? solve_impl<float>(P, 3, sol, ss, rhs, rs, s, nullptr, 0, nullptr, nullptr, true)
             : solve_impl<double>(P, 3, sol, ss, rhs, rs, s, nullptr, 0, nullptr, nullptr, true);
}

int poisson_table_solve(sopht_poisson3d* P, int ncomp, void* const* sol_tab, const int64_t* ss,
                        const void* const* rhs_tab, const int64_t* rs, hipStream_t s) {
  return P->dtype == SOPHT_F32
             ? solve_impl<float>(P, ncomp, nullptr, ss, nullptr, rs, s, nullptr, 0, sol_tab, rhs_tab)
             : solve_impl<double>(P, ncomp, nullptr, ss, nullptr, rs, s, nullptr, 0, sol_tab, rhs_tab);
}

}  // namespace sopht

using namespace sopht;

extern "C" {

int sopht_poisson3d_create(int dtype, int64_t nz, int64_t ny, int64_t nx, double x_range,
                           void* stream, sopht_poisson3d** out) {
  return sopht::poisson_create(dtype, nz, ny, nx, x_range, 3, stream, out);
}

}  // extern "C"

namespace sopht {

int poisson_create(int dtype, int64_t nz, int64_t ny, int64_t nx, double x_range, int comps,
                   void* stream, sopht_poisson3d** out) {
  SOPHT_REQUIRE(out != nullptr, "plan pointer is null");
  SOPHT_REQUIRE(comps >= 1, "a Poisson plan holds at least one component");
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
  const int64_t rs = kRowPad ? round_up(nx + 1, int64_t(128 / cs)) : nx + 1;
  P->gz = kRowPad ? round_up(nz + 1, int64_t(128 / rsz)) : nz + 1;
  P->slab_comps = comps;
  P->slab_bytes = int64_t(comps) * nz * 2 * ny * rs * int64_t(cs);
  if (int64_t(comps) * nz * 2 * ny * rs >= (int64_t(1) << 31)) {
    delete P;
    set_last_error("Poisson grid too large: the spectral slab must have < 2^31 elements", __FILE__, __LINE__);
    return SOPHT_ERR_UNSUPPORTED;
  }
  P->ghat_bytes = (ny + 1) * (nx + 1) * P->gz * int64_t(rsz);
  // the three-pass schedule where its kernels are built (FP32, nx = ny, 2 nx in
  // SOPHT_XY_LENGTHS, 2 nz in SOPHT_ZROW_LENGTHS); SOPHT_POISSON_PASSES=5 in the environment
  // keeps the five-pass one
  const char* forced = std::getenv("SOPHT_POISSON_PASSES");
  const bool five = forced != nullptr && std::atoi(forced) == 5;
  P->passes = (!five && kRowPad && dtype == SOPHT_F32 && nx == ny && xy_length_supported(2 * nx) &&
               zrow_length_supported(2 * nz)) ? 3 : 5;
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
  if ((rc = tune_zpass(P, s))) return fail(rc);
  *out = P;
  return SOPHT_OK;
}

}  // namespace sopht

extern "C" {

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

int sopht_poisson3d_schedule(const sopht_poisson3d* P, int* passes, int* zpass_jobs) {
  SOPHT_REQUIRE(P != nullptr, "plan is null");
  if (passes) *passes = P->passes;
  if (zpass_jobs) *zpass_jobs = P->passes == 3 ? P->zjobs : 0;
  return SOPHT_OK;
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
