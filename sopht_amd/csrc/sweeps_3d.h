// Device code of the fused flow-step sweeps (S1, SC, S2) and the CFL reduction, shared by
// the single-simulator step (flow_step_3d.hip) and the batched step (batch_3d.hip). Each sweep
// is a __device__ body over one work item of one instance plus thin kernels that map the
// workgroup index to (instance, item).
#pragma once

#include "common.h"

namespace sopht {
namespace {

struct Grid {
  int64_t nz, ny, nx;
  __device__ __forceinline__ int64_t at(int64_t z, int64_t y, int64_t x) const {
    return (z * ny + y) * nx + x;
  }
  __device__ __forceinline__ bool interior(int64_t z, int64_t y, int64_t x) const {
    return z > 0 && y > 0 && x > 0 && z < nz - 1 && y < ny - 1 && x < nx - 1;
  }
};

// z planes per S1 work item (two extra planes of B are recomputed per item)
constexpr int kS1Chunk = 16;

// S1: omega2 = omega1 + pB curl_h(u x omega1), omega1 = omega + pF curl_h(F), with neither
// omega1 nor B stored in HBM. A workgroup owns a TX x TY tile of (x, y) columns over a chunk
// of kS1Chunk planes and marches up z. Iteration i
//   (1) produces omega1 and B on plane i over the tile plus its 1-cell edge ring (without
//       corners); curl_h(F) reads F's neighbours through the caches; B goes to a 3-plane LDS
//       ring, the tile's own omega1(i), B(i), B(i-1) stay in registers;
//   (2) after the one barrier, consumes plane i-1: its curl_h(B) takes the in-plane
//       neighbours from LDS and B(i), B(i-2) from registers.
// The ring slot written at iteration i was last read before the previous iteration's barrier,
// so one barrier per plane is enough. The small LDS footprint leaves room for several
// workgroups per CU, whose loads overlap each other's barriers.
template <typename T>
struct S1Tile {
  static constexpr int TX = 64, TY = sizeof(T) == 4 ? 8 : 4, NT = TX * TY;
  static constexpr int GX = TX + 2, GY = TY + 2, GP = GX * GY;  // B box (1-cell ring)
  static constexpr int NEDGE = 2 * (TX + TY);                    // edge-ring items (no corners)
};

// true when any bit of v is set (-0.0 and NaN count)
__device__ __forceinline__ bool nonzero_bits(float v) { return __float_as_uint(v) != 0u; }
__device__ __forceinline__ bool nonzero_bits(double v) { return __double_as_longlong(v) != 0ll; }

template <typename T>
__device__ __forceinline__ void s1_body(Grid g, const T* w, const T* u, const T* F, T* w2out, T pF,
                                        T pB, int with_forcing, int zchunk, int tiles_x,
                                        int tiles_y, int64_t item, T* w1out = nullptr,
                                        int* fflags = nullptr) {
  using S = S1Tile<T>;
  constexpr int TX = S::TX, TY = S::TY, GX = S::GX, GP = S::GP;
  __shared__ T bs[3][3][GP];
  const int64_t tx_i = item % tiles_x;
  const int64_t ty_i = (item / tiles_x) % tiles_y;
  const int64_t zc_i = item / (int64_t(tiles_x) * tiles_y);
  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  // 32-bit offsets (the host checks 3 nz ny nx < 2^31): one VGPR per address
  const int x0 = int(tx_i) * TX, y0 = int(ty_i) * TY;
  const int z0 = int(zc_i) * zchunk;
  const int z1 = z0 + zchunk < g.nz ? z0 + zchunk : int(g.nz);
  const int n = int(g.nz * g.ny * g.nx), sz = int(g.ny * g.nx), sy = int(g.nx);
  const bool tv = x0 + tx < g.nx && y0 + ty < g.ny;
  // edge-ring item: x = -1 / TX columns, then y = -1 / TY rows
  int ex = 0, ey = 0;
  if (tid < 2 * TY) {
    ex = (tid & 1) ? TX : -1;
    ey = tid >> 1;
  } else if (tid < S::NEDGE) {
    ex = (tid - 2 * TY) >> 1;
    ey = ((tid - 2 * TY) & 1) ? TY : -1;
  }
  const bool ev = tid < S::NEDGE && x0 + ex >= 0 && x0 + ex < g.nx && y0 + ey >= 0 && y0 + ey < g.ny;

  // omega1 and B = u x omega1 at (x0 + xr, y0 + yr, q); cf: curl_h F already formed (tile
  // cells), or null (edge-ring cells read F's neighbours)
  auto produce = [&](int xr, int yr, int q, T b[3], T a[3], const T* cf) {
    const int x = x0 + xr, y = y0 + yr;
    const int j = q * sz + y * sy + x;
    a[0] = w[j];
    a[1] = w[n + j];
    a[2] = w[2 * n + j];
    const T u0 = u[j], u1 = u[n + j], u2 = u[2 * n + j];
    if (with_forcing && g.interior(q, y, x)) {
      T c0, c1, c2;
      if (cf) {
        c0 = cf[0];
        c1 = cf[1];
        c2 = cf[2];
      } else {
        const T* F0 = F;
        const T* F1 = F + n;
        const T* F2 = F + 2 * n;
        c0 = F2[j + sy] - F2[j - sy] - F1[j + sz] + F1[j - sz];
        c1 = F0[j + sz] - F0[j - sz] - F2[j + 1] + F2[j - 1];
        c2 = F1[j + 1] - F1[j - 1] - F0[j + sy] + F0[j - sy];
      }
      a[0] = a[0] + pF * c0;
      a[1] = a[1] + pF * c1;
      a[2] = a[2] + pF * c2;
    }
    // elementwise_ops_3d.py:427-447: B_i = u_j w_k - w_j u_k
    b[0] = u1 * a[2] - a[1] * u2;
    b[1] = u2 * a[0] - a[2] * u0;
    b[2] = u0 * a[1] - a[0] * u1;
  };

  T bm2[3] = {T(0), T(0), T(0)}, bm1[3] = {T(0), T(0), T(0)}, a1prev[3] = {T(0), T(0), T(0)};
  const int gt = (ty + 1) * GX + tx + 1;
  // curl_h F of the tile cells: F0 and F1 of the lane's own column ride along z in registers,
  // the x neighbours of F1 and F2 come from the adjacent lanes of the row (a tile row is one
  // wavefront), the y neighbours are row loads
  static_assert(TX == 64, "a tile row must be one wavefront");
  const int xt = x0 + tx, yt = y0 + ty;
  const int colj = yt * sy + xt;
  auto fown = [&](const T* f, int q) -> T {
    return (with_forcing && tv && q >= 0 && q < g.nz) ? f[q * sz + colj] : T(0);
  };
  T f0m = fown(F, z0 - 2), f0c = fown(F, z0 - 1), f1m = fown(F + n, z0 - 2), f1c = fown(F + n, z0 - 1);
  bool fset = false;  // some bit of F set in this lane's column (the tile's SC zeroes F then)
  for (int i = z0 - 1; i <= z1; ++i) {
    T b0[3] = {T(0), T(0), T(0)}, a1[3] = {T(0), T(0), T(0)};
    const T f0p = fown(F, i + 1), f1p = fown(F + n, i + 1);
    const T f2c = fown(F + 2 * n, i);
    fset = fset || nonzero_bits(f0p) || nonzero_bits(f1p) || nonzero_bits(f2c);
    T f2r = __shfl_down(f2c, 1, 64), f2l = __shfl_up(f2c, 1, 64);
    T f1r = __shfl_down(f1c, 1, 64), f1l = __shfl_up(f1c, 1, 64);
    T cf[3] = {T(0), T(0), T(0)};
    if (with_forcing && tv && i >= 0 && i < g.nz && g.interior(i, yt, xt)) {
      const int j = i * sz + colj;  // interior: every neighbour below is inside the grid
      if (tx == 0) {
        f2l = F[2 * n + j - 1];
        f1l = F[n + j - 1];
      }
      if (tx == TX - 1) {
        f2r = F[2 * n + j + 1];
        f1r = F[n + j + 1];
      }
      cf[0] = F[2 * n + j + sy] - F[2 * n + j - sy] - f1p + f1m;
      cf[1] = f0p - f0m - f2r + f2l;
      cf[2] = f1r - f1l - F[j + sy] + F[j - sy];
    }
    f0m = f0c;
    f0c = f0p;
    f1m = f1c;
    f1c = f1p;
    if (i >= 0 && i < g.nz) {
      T(*Bq)[GP] = bs[i % 3];
      if (tv) {
        produce(tx, ty, i, b0, a1, cf);
        Bq[0][gt] = b0[0];
        Bq[1][gt] = b0[1];
        Bq[2][gt] = b0[2];
      }
      if (ev) {
        T eb[3], ea[3];
        produce(ex, ey, i, eb, ea, nullptr);
        const int ge = (ey + 1) * GX + ex + 1;
        Bq[0][ge] = eb[0];
        Bq[1][ge] = eb[1];
        Bq[2][ge] = eb[2];
      }
    }
    __syncthreads();
    const int r = i - 1;  // consumed plane: B(r-1) = bm2, B(r+1) = b0, omega1(r) = a1prev
    if (r >= z0 && tv) {
      const int x = x0 + tx, y = y0 + ty;
      T o0 = a1prev[0], o1 = a1prev[1], o2 = a1prev[2];
      if (g.interior(r, y, x)) {
        const T(*Bc)[GP] = bs[r % 3];
        const T c0 = Bc[2][gt + GX] - Bc[2][gt - GX] - b0[1] + bm2[1];
        const T c1 = b0[0] - bm2[0] - Bc[2][gt + 1] + Bc[2][gt - 1];
        const T c2 = Bc[1][gt + 1] - Bc[1][gt - 1] - Bc[0][gt + GX] + Bc[0][gt - GX];
        o0 = o0 + pB * c0;
        o1 = o1 + pB * c1;
        o2 = o2 + pB * c2;
      }
      const int j = r * sz + y * sy + x;
      __builtin_nontemporal_store(o0, w2out + j);
      __builtin_nontemporal_store(o1, w2out + n + j);
      __builtin_nontemporal_store(o2, w2out + 2 * n + j);
      if (w1out) {  // omega1 as well (trajectories recorded for the adjoint)
        w1out[j] = a1prev[0];
        w1out[n + j] = a1prev[1];
        w1out[2 * n + j] = a1prev[2];
      }
    }
    for (int c = 0; c < 3; ++c) {
      bm2[c] = bm1[c];
      bm1[c] = b0[c];
      a1prev[c] = a1[c];
    }
  }
  if (fflags) {  // per tile (logical item index): does the tile's F need zeroing
    const int any = __syncthreads_or(fset ? 1 : 0);
    if (tid == 0) fflags[item] = any;
  }
}

// Workgroups are dealt to the 8 XCDs in turn; with the remap, XCD k takes a contiguous run of
// work items (neighbouring y tiles share halo rows in its L2).
__device__ __forceinline__ int64_t xcd_item(int remap) {
  int64_t item = blockIdx.x;
  if (remap) item = (item % 8) * (gridDim.x / 8) + item / 8;
  return item;
}

template <typename T>
__global__ void __launch_bounds__(S1Tile<T>::NT) __attribute__((amdgpu_waves_per_eu(6)))
k_s1(Grid g, const T* w, const T* u, const T* F, T* w2out, T pF, T pB, int with_forcing,
     int zchunk, int tiles_x, int tiles_y, int remap, T* w1out, int* fflags) {
  s1_body<T>(g, w, u, F, w2out, pF, pB, with_forcing, zchunk, tiles_x, tiles_y, xcd_item(remap),
             w1out, fflags);
}

struct Pen {
  int w;
  // per-axis factors: index 0..w-1 front, then w back entries for n-w..n-1
  double f[3][16];
};

// Column-marching sweeps (SC, S2): a workgroup owns a 64-cell-wide tile of rows over
// kMarchChunk planes. Each thread holds VEC neighbouring cells of one row (one 16-byte
// access per row instead of VEC scalar ones) and walks up z with the z neighbours in
// registers; x neighbours across threads come from the adjacent lane of the row's lane
// group (tile edges load them). Work items are dealt so that each XCD takes a contiguous run
// (y neighbours share rows in that XCD's L2). Offsets are 32-bit (3 nz ny nx < 2^31 is
// checked at plan creation). VEC = 1 covers row lengths that are not a multiple of the
// vector width and penalty zones wider than a vector.
constexpr int MX = 64, kMarchThreads = 256, kMarchChunk = 16;

template <typename T, int VEC>
struct alignas(sizeof(T) * VEC) VecT {
  T v[VEC];
};

template <typename T, int VEC>
__device__ __forceinline__ VecT<T, VEC> vload(const T* __restrict__ p, int off) {
  return *reinterpret_cast<const VecT<T, VEC>*>(p + off);
}
template <typename T, int VEC>
__device__ __forceinline__ void vstore(T* __restrict__ p, int off, const VecT<T, VEC>& v) {
  *reinterpret_cast<VecT<T, VEC>*>(p + off) = v;
}
// A streaming (non-temporal) store: for outputs the next kernel reads back from HBM.
template <typename T, int VEC>
__device__ __forceinline__ void vstore_stream(T* __restrict__ p, int off, const VecT<T, VEC>& v) {
  if constexpr (VEC == 1) {
    __builtin_nontemporal_store(v.v[0], p + off);
  } else {
    typedef T vec __attribute__((ext_vector_type(VEC)));
    vec w;
#pragma unroll
    for (int k = 0; k < VEC; ++k) w[k] = v.v[k];
    __builtin_nontemporal_store(w, reinterpret_cast<vec*>(p + off));
  }
}
template <typename T, int VEC>
__device__ __forceinline__ VecT<T, VEC> vzero() {
  VecT<T, VEC> r;
  for (int k = 0; k < VEC; ++k) r.v[k] = T(0);
  return r;
}

template <int VEC>
struct MarchGeom {
  static constexpr int MXT = MX / VEC;              // threads per row
  static constexpr int MYV = kMarchThreads / MXT;   // rows per tile
};

struct MarchTile {
  int tx, x0, y, z0, z1;
};

template <int VEC>
__device__ __forceinline__ MarchTile march_tile(const Grid& g, int tiles_x, int tiles_y,
                                                int64_t item, int chunk = kMarchChunk) {
  using G = MarchGeom<VEC>;
  MarchTile m;
  m.tx = int(threadIdx.x % G::MXT);
  m.x0 = int(item % tiles_x) * MX + m.tx * VEC;
  m.y = int((item / tiles_x) % tiles_y) * G::MYV + int(threadIdx.x / G::MXT);
  m.z0 = int(item / (int64_t(tiles_x) * tiles_y)) * chunk;
  m.z1 = m.z0 + chunk < int(g.nz) ? m.z0 + chunk : int(g.nz);
  return m;
}

struct MarchLaunch {
  unsigned blocks;
  int tiles_x, tiles_y, remap, vec, chunk;
};

// chunk = z planes per work item. The default keeps kMarchChunk; fill = true halves it (down
// to 2) while the launch has fewer than about 1024 workgroups, so small grids still put four
// wavefronts on every SIMD (the extra z-window planes of a shorter chunk come from L2).
inline MarchLaunch march_launch(int64_t nz, int64_t ny, int64_t nx, int vec, bool fill = false) {
  MarchLaunch l;
  l.vec = vec;
  l.tiles_x = int(ceil_div(nx, MX));
  l.tiles_y = int(ceil_div(ny, kMarchThreads / (MX / vec)));
  l.chunk = kMarchChunk;
  if (fill)
    while (l.chunk > 2 && int64_t(l.tiles_x) * l.tiles_y * ceil_div(nz, l.chunk) < 1024) l.chunk /= 2;
  const int64_t items = int64_t(l.tiles_x) * l.tiles_y * ceil_div(nz, l.chunk);
  l.blocks = unsigned(items);
  l.remap = items % 8 == 0 ? 1 : 0;
  return l;
}

// x neighbours of a row vector spread over the MXT lanes of a row: the element left of
// v[0] and right of v[VEC-1]. Every lane of the row must call this (lane shuffles).
template <typename T, int VEC>
__device__ __forceinline__ void x_nbrs(const VecT<T, VEC>& v, const T* __restrict__ row, int x0,
                                       int nx, int tx, bool act, T& left, T& right) {
  constexpr int MXT = MarchGeom<VEC>::MXT;
  if constexpr (VEC == 1) {
    // scalar path: a lane's column may be a clamped one (penalty zone), so the adjacent
    // lane does not necessarily hold the adjacent column
    left = (act && x0 > 0) ? row[x0 - 1] : T(0);
    right = (act && x0 + 1 < nx) ? row[x0 + 1] : T(0);
  } else {
    left = __shfl_up(v.v[VEC - 1], 1, MXT);
    right = __shfl_down(v.v[0], 1, MXT);
    if (tx == 0) left = (act && x0 > 0) ? row[x0 - 1] : T(0);
    if (tx == MXT - 1) right = (act && x0 + VEC < nx) ? row[x0 + VEC] : T(0);
  }
}

// S1 with 16-byte accesses (rows whose length is a multiple of VEC). The scalar sweep above
// issues one load instruction per array, plane and cell. Here a thread holds VEC neighbouring
// cells of one row, so each array and plane costs one 16-byte load per VEC cells (3.4x fewer
// load instructions; the kernel is then bound by HBM traffic). The tile is TX = 64 or 128
// cells by TY rows and marches a chunk of z planes as s1_body does. 128-cell rows of 8 (FP32)
// fetch fewer x edge lines per cell than 64-cell rows of 16: at 256^3 FP32 0.210 against
// 0.225 ms, 1.19x against 1.54x its algorithmic bytes.
// Per plane i:
//   - the tile threads form omega1 and B on their cells; the x neighbours of F come from the
//     adjacent lanes of the row, its y neighbours are row loads, and F0 and F1 ride along z in
//     registers;
//   - wavefront 0 forms B on the y ring rows (y0 - 1 and y0 + TY);
//   - wavefront 1 forms B on the x edge columns (x0 - 1 and x0 + TX).
// LDS keeps three planes of B: the B0 and B2 rows of the tile and ring (y derivatives), the B1
// rows of the tile, and the B1 and B2 edge columns (x derivatives at the tile's first and last
// lanes). After the one barrier per plane, plane i - 1 is consumed: its x derivatives from the
// adjacent lanes, its y and z derivatives from LDS. From one plane to the next, registers keep
// omega1 of the consumed plane and F0 and F1 of the thread's cells (their z window). The
// F0/F2 rows also going through LDS (for their y neighbours) cut the sweep's HBM fetch from
// 500 to 435 MB at 256^3 FP32 but not its time (0.229 against 0.224 ms), and were not kept.
// Every expression keeps the operand order of s1_body, so both sweeps give the same bits.
template <typename T, int TXW = 64, int TYW = sizeof(T) == 4 ? 16 : 8>
struct S1VTile {
  static constexpr int VEC = 16 / int(sizeof(T));
  static constexpr int TX = TXW, MXT = TX / VEC;           // cells and lanes per row
  static constexpr int TY = TYW, NT = MXT * TY;
  static constexpr int RW = TY + 2, ROWS = RW * TX;        // a component's row box (ring rows included)
  static constexpr int B1OFF = 2 * ROWS, EOFF = B1OFF + TY * TX;
  static constexpr int PL = EOFF + 4 * TY;  // B0 and B2 row boxes, B1 tile rows, B1 and B2 edge columns
  static_assert(NT >= 64 + 2 * TY && 2 * MXT <= 64 && TX % MX == 0,
                "ring rows in wavefront 0, edge columns in wavefront 1");
};

// Where a flow step's S1 tiles lie, for the F-zeroing flags it leaves per tile (SC reads them).
struct S1Map {
  int tiles_x, tiles_y, rows, chunk;  // tiles of `cols` cells by `rows` rows and `chunk` planes
  int cols = 64;
};

template <typename T>
__device__ __forceinline__ bool any_bits(const VecT<T, 16 / int(sizeof(T))>& v) {
  bool r = false;
  for (int k = 0; k < 16 / int(sizeof(T)); ++k) r = r || nonzero_bits(v.v[k]);
  return r;
}

template <typename T, int TXW = 64, int TYW = sizeof(T) == 4 ? 16 : 8>
__device__ __forceinline__ void s1v_body(Grid g, const T* __restrict__ w, const T* __restrict__ u,
                                         const T* __restrict__ F, T* __restrict__ w2out, T pF, T pB,
                                         int with_forcing, int zchunk, int tiles_x, int tiles_y,
                                         int64_t item, T* __restrict__ w1out = nullptr,
                                         int* fflags = nullptr) {
  using S = S1VTile<T, TXW, TYW>;
  constexpr int VEC = S::VEC, TX = S::TX, MXT = S::MXT, TY = S::TY, ROWS = S::ROWS;
  constexpr int B1OFF = S::B1OFF, EOFF = S::EOFF;
  using V = VecT<T, VEC>;
  __shared__ __align__(16) T bs[3][S::PL];
  const int tid = threadIdx.x;
  const int nz = int(g.nz), ny = int(g.ny), nx = int(g.nx);
  const int n = nz * ny * nx, sz = ny * nx, sy = nx;
  const int x0 = int(item % tiles_x) * TX;
  const int y0 = int((item / tiles_x) % tiles_y) * TY;
  const int z0 = int(item / (int64_t(tiles_x) * tiles_y)) * zchunk;
  const int z1 = z0 + zchunk < nz ? z0 + zchunk : nz;
  const bool fw = with_forcing != 0;
  const T* const F0 = F;
  const T* const F1 = F + n;
  const T* const F2 = F + 2 * n;
  // tile role: cells xs .. xs + VEC - 1 of row yt
  const int tx = tid % MXT, ty = tid / MXT;
  const int xs = x0 + tx * VEC, yt = y0 + ty;
  const bool tv = xs < nx && yt < ny;
  const int colj = yt * sy + xs;
  const int lrow = (ty + 1) * TX + tx * VEC;  // the cells' slot in a row box
  const int l1 = B1OFF + ty * TX + tx * VEC;   // and in the B1 rows
  // ring role (wavefront 0): cells xs .. of row yr
  const bool ring = tid < 2 * MXT;
  const int yr = tid < MXT ? y0 - 1 : y0 + TY;
  const bool rv = ring && xs < nx && yr >= 0 && yr < ny;
  const int rrow = (tid < MXT ? 0 : TY + 1) * TX + tx * VEC;
  // edge role (wavefront 1): cell (xe, ye)
  const int e = tid - 64;
  const bool edge = e >= 0 && e < 2 * TY;
  const int side = e >= TY ? 1 : 0, ey = y0 + (e - side * TY);
  const int xe = side ? x0 + TX : x0 - 1;
  const bool ev = edge && xe >= 0 && xe < nx && ey < ny;

  // Accesses at a uniform base (a component's start) plus a 32-bit byte offset, so an address
  // is one VGPR over an SGPR pair (the host keeps n sizeof(T) below 2^32).
  auto at = [](const T* p, int off) -> const T* {
    return reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + uint32_t(off) * uint32_t(sizeof(T)));
  };
  // one vector-typed load (a struct load is split per element and a conditionally used
  // element sunk into a branch, which breaks the 16-byte access apart)
  typedef T vecT __attribute__((ext_vector_type(VEC)));
  auto vl = [&](const T* p, int off) -> V {
    const vecT x = *reinterpret_cast<const vecT*>(at(p, off));
    V r;
    for (int k = 0; k < VEC; ++k) r.v[k] = x[k];
    return r;
  };
  auto sl = [&](const T* p, int off) -> T { return *at(p, off); };
  auto ld = [&](const T* p, int off, bool ok) -> V { return ok ? vl(p, off) : vzero<T, VEC>(); };
  // x neighbours of a row vector at element offset `off` (adjacent lanes; the tile's first and
  // last lanes load); every lane of the row takes part
  auto xn = [&](const V& v, const T* p, int off, bool act, T& l, T& r) {
    l = __shfl_up(v.v[VEC - 1], 1, MXT);
    r = __shfl_down(v.v[0], 1, MXT);
    if (tx == 0) l = (act && xs > 0) ? sl(p, off - 1) : T(0);
    if (tx == MXT - 1) r = (act && xs + VEC < nx) ? sl(p, off + VEC) : T(0);
  };
  auto lds_ld = [&](const T* p) -> V {
    const vecT x = *reinterpret_cast<const vecT*>(p);
    V r;
    for (int k = 0; k < VEC; ++k) r.v[k] = x[k];
    return r;
  };
  auto lds_st = [&](T* p, const V& v) {
    vecT x;
    for (int k = 0; k < VEC; ++k) x[k] = v.v[k];
    *reinterpret_cast<vecT*>(p) = x;
  };
  auto vst = [&](T* p, int off, const V& v, bool stream) {
    T* q = const_cast<T*>(at(p, off));
    vecT x;
    for (int k = 0; k < VEC; ++k) x[k] = v.v[k];
    if (stream) __builtin_nontemporal_store(x, reinterpret_cast<vecT*>(q));
    else *reinterpret_cast<vecT*>(q) = x;
  };
  auto fz = [&](const T* f, int q) -> V { return ld(f, q * sz + colj, fw && tv && q >= 0 && q < nz); };
  // F0 and F1 of the thread's cells ride along z (m: plane i - 1, c: plane i)
  V f0m = fz(F0, z0 - 2), f0c = fz(F0, z0 - 1), f1m = fz(F1, z0 - 2), f1c = fz(F1, z0 - 1);
  // omega1 of the consumed plane
  V a1p[3] = {vzero<T, VEC>(), vzero<T, VEC>(), vzero<T, VEC>()};
  bool fset = false;  // some bit of F set in this thread's cells (the tile's SC zeroes F then)
#pragma unroll 1
  for (int i = z0 - 1; i <= z1; ++i) {
    const bool pin = i >= 0 && i < nz;
    const bool qin = i > 0 && i < nz - 1;
    // F0 and F1 of plane i + 1 and F2 of plane i (the flag looks at these, as s1_body does)
    const V f0p = fz(F0, i + 1), f1p = fz(F1, i + 1), f2c = fz(F2, i);
    if (fw) fset = fset || any_bits<T>(f0p) || any_bits<T>(f1p) || any_bits<T>(f2c);
    // x neighbours of F1 and F2 on plane i (every lane of the row takes part in the shuffles)
    const bool fact = fw && tv && pin;
    T f1l, f1r, f2l, f2r;
    xn(f1c, F1, i * sz + colj, fact, f1l, f1r);
    xn(f2c, F2, i * sz + colj, fact, f2l, f2r);
    T* const Bq = bs[(i + 3) % 3];
    if (ring) {
      // B0 and B2 of the ring row yr on plane i
      const bool ok = rv && pin;
      const int j = i * sz + yr * sy + xs;
      const V r1 = ld(F1, j, fw && ok), r2 = ld(F2, j, fw && ok);
      T l1, q1, l2, q2;
      xn(r1, F1, j, fw && ok, l1, q1);
      xn(r2, F2, j, fw && ok, l2, q2);
      if (ok) {
        V a0 = vl(w, j), a1 = vl(w + n, j), a2 = vl(w + 2 * n, j);
        const V u0 = vl(u, j), u1 = vl(u + n, j), u2 = vl(u + 2 * n, j);
        if (fw && qin && yr > 0 && yr < ny - 1) {
          const V f2yp = vl(F2, j + sy), f2ym = vl(F2, j - sy);
          const V f0yp = vl(F0, j + sy), f0ym = vl(F0, j - sy);
          const V f1zp = vl(F1, j + sz), f1zm = vl(F1, j - sz);
          const V f0zp = vl(F0, j + sz), f0zm = vl(F0, j - sz);
          for (int k = 0; k < VEC; ++k) {
            const int x = xs + k;
            const bool inner = x > 0 && x < nx - 1;
            const T f2xl = k > 0 ? r2.v[k - 1] : l2, f2xr = k < VEC - 1 ? r2.v[k + 1] : q2;
            const T f1xl = k > 0 ? r1.v[k - 1] : l1, f1xr = k < VEC - 1 ? r1.v[k + 1] : q1;
            const T c0 = f2yp.v[k] - f2ym.v[k] - f1zp.v[k] + f1zm.v[k];
            const T c1 = f0zp.v[k] - f0zm.v[k] - f2xr + f2xl;
            const T c2 = f1xr - f1xl - f0yp.v[k] + f0ym.v[k];
            const T o0 = a0.v[k] + pF * c0, o1 = a1.v[k] + pF * c1, o2 = a2.v[k] + pF * c2;
            a0.v[k] = inner ? o0 : a0.v[k];
            a1.v[k] = inner ? o1 : a1.v[k];
            a2.v[k] = inner ? o2 : a2.v[k];
          }
        }
        V rb0, rb2;
        for (int k = 0; k < VEC; ++k) {
          rb0.v[k] = u1.v[k] * a2.v[k] - a1.v[k] * u2.v[k];
          rb2.v[k] = u0.v[k] * a1.v[k] - a0.v[k] * u1.v[k];
        }
        lds_st(Bq + rrow, rb0);
        lds_st(Bq + ROWS + rrow, rb2);
      }
    }
    V an[3] = {vzero<T, VEC>(), vzero<T, VEC>(), vzero<T, VEC>()};
    if (tv && pin) {
      const int j = i * sz + colj;
      an[0] = vl(w, j);
      an[1] = vl(w + n, j);
      an[2] = vl(w + 2 * n, j);
      const V u0 = vl(u, j), u1 = vl(u + n, j), u2 = vl(u + 2 * n, j);
      if (fw && qin && yt > 0 && yt < ny - 1) {
        const V f2yp = vl(F2, j + sy), f2ym = vl(F2, j - sy);
        const V f0yp = vl(F0, j + sy), f0ym = vl(F0, j - sy);
        // every element is formed and the border cells keep omega by a select, so the
        // loads stay whole 16-byte ones
        for (int k = 0; k < VEC; ++k) {
          const int x = xs + k;
          const bool inner = x > 0 && x < nx - 1;
          const T f2xl = k > 0 ? f2c.v[k - 1] : f2l, f2xr = k < VEC - 1 ? f2c.v[k + 1] : f2r;
          const T f1xl = k > 0 ? f1c.v[k - 1] : f1l, f1xr = k < VEC - 1 ? f1c.v[k + 1] : f1r;
          const T c0 = f2yp.v[k] - f2ym.v[k] - f1p.v[k] + f1m.v[k];
          const T c1 = f0p.v[k] - f0m.v[k] - f2xr + f2xl;
          const T c2 = f1xr - f1xl - f0yp.v[k] + f0ym.v[k];
          const T o0 = an[0].v[k] + pF * c0, o1 = an[1].v[k] + pF * c1, o2 = an[2].v[k] + pF * c2;
          an[0].v[k] = inner ? o0 : an[0].v[k];
          an[1].v[k] = inner ? o1 : an[1].v[k];
          an[2].v[k] = inner ? o2 : an[2].v[k];
        }
      }
      // elementwise_ops_3d.py:427-447: B_i = u_j w_k - w_j u_k
      V b0n, b1n, b2n;
      for (int k = 0; k < VEC; ++k) {
        b0n.v[k] = u1.v[k] * an[2].v[k] - an[1].v[k] * u2.v[k];
        b1n.v[k] = u2.v[k] * an[0].v[k] - an[2].v[k] * u0.v[k];
        b2n.v[k] = u0.v[k] * an[1].v[k] - an[0].v[k] * u1.v[k];
      }
      lds_st(Bq + lrow, b0n);
      lds_st(Bq + ROWS + lrow, b2n);
      lds_st(Bq + l1, b1n);
    }
    if (ev && pin) {
      // B1 and B2 of the edge cell (xe, ey) on plane i
      const int j = i * sz + ey * sy + xe;
      T a0 = sl(w, j), a1 = sl(w + n, j), a2 = sl(w + 2 * n, j);
      const T u0 = sl(u, j), u1 = sl(u + n, j), u2 = sl(u + 2 * n, j);
      if (fw && g.interior(i, ey, xe)) {
        const T c0 = sl(F2, j + sy) - sl(F2, j - sy) - sl(F1, j + sz) + sl(F1, j - sz);
        const T c1 = sl(F0, j + sz) - sl(F0, j - sz) - sl(F2, j + 1) + sl(F2, j - 1);
        const T c2 = sl(F1, j + 1) - sl(F1, j - 1) - sl(F0, j + sy) + sl(F0, j - sy);
        a0 = a0 + pF * c0;
        a1 = a1 + pF * c1;
        a2 = a2 + pF * c2;
      }
      Bq[EOFF + e] = u2 * a0 - a2 * u0;
      Bq[EOFF + 2 * TY + e] = u0 * a1 - a0 * u1;
    }
    f0m = f0c;
    f0c = f0p;
    f1m = f1c;
    f1c = f1p;
    __syncthreads();
    const int r = i - 1;  // consumed plane (omega1 in a1p)
    const T* const Bc = bs[(r + 3) % 3];
    const T* const Bm = bs[(r + 2) % 3];  // plane r - 1
    // B1 and B2 of the thread's cells on plane r and their x neighbours (every lane shuffles)
    const V b1c = lds_ld(Bc + l1);
    const V b2c = lds_ld(Bc + ROWS + lrow);
    T b1l = __shfl_up(b1c.v[VEC - 1], 1, MXT), b1r = __shfl_down(b1c.v[0], 1, MXT);
    T b2l = __shfl_up(b2c.v[VEC - 1], 1, MXT), b2r = __shfl_down(b2c.v[0], 1, MXT);
    if (r >= z0 && tv) {
      if (tx == 0) {
        b1l = Bc[EOFF + ty];
        b2l = Bc[EOFF + 2 * TY + ty];
      }
      if (tx == MXT - 1) {
        b1r = Bc[EOFF + TY + ty];
        b2r = Bc[EOFF + 3 * TY + ty];
      }
      V o[3] = {a1p[0], a1p[1], a1p[2]};
      if (r > 0 && r < nz - 1 && yt > 0 && yt < ny - 1) {
        const V b0yp = lds_ld(Bc + lrow + TX);
        const V b0ym = lds_ld(Bc + lrow - TX);
        const V b2yp = lds_ld(Bc + ROWS + lrow + TX);
        const V b2ym = lds_ld(Bc + ROWS + lrow - TX);
        const V b0zp = lds_ld(Bq + lrow), b0zm = lds_ld(Bm + lrow);
        const V b1zp = lds_ld(Bq + l1), b1zm = lds_ld(Bm + l1);
        for (int k = 0; k < VEC; ++k) {
          const int x = xs + k;
          const bool inner = x > 0 && x < nx - 1;
          const T b2xl = k > 0 ? b2c.v[k - 1] : b2l, b2xr = k < VEC - 1 ? b2c.v[k + 1] : b2r;
          const T b1xl = k > 0 ? b1c.v[k - 1] : b1l, b1xr = k < VEC - 1 ? b1c.v[k + 1] : b1r;
          const T c0 = b2yp.v[k] - b2ym.v[k] - b1zp.v[k] + b1zm.v[k];
          const T c1 = b0zp.v[k] - b0zm.v[k] - b2xr + b2xl;
          const T c2 = b1xr - b1xl - b0yp.v[k] + b0ym.v[k];
          const T q0 = o[0].v[k] + pB * c0, q1 = o[1].v[k] + pB * c1, q2 = o[2].v[k] + pB * c2;
          o[0].v[k] = inner ? q0 : o[0].v[k];
          o[1].v[k] = inner ? q1 : o[1].v[k];
          o[2].v[k] = inner ? q2 : o[2].v[k];
        }
      }
      const int j = r * sz + colj;
      vst(w2out, j, o[0], true);
      vst(w2out + n, j, o[1], true);
      vst(w2out + 2 * n, j, o[2], true);
      if (w1out) {  // omega1 as well (trajectories recorded for the adjoint)
        vst(w1out, j, a1p[0], false);
        vst(w1out + n, j, a1p[1], false);
        vst(w1out + 2 * n, j, a1p[2], false);
      }
    }
    a1p[0] = an[0];
    a1p[1] = an[1];
    a1p[2] = an[2];
  }
  if (fflags) {  // per tile (logical item index): does the tile's F need zeroing
    const int any = __syncthreads_or(fset ? 1 : 0);
    if (tid == 0) fflags[item] = any;
  }
}

template <typename T, int TXW = 64, int TYW = sizeof(T) == 4 ? 16 : 8>
__global__ void __launch_bounds__((S1VTile<T, TXW, TYW>::NT)) __attribute__((amdgpu_waves_per_eu(2)))
k_s1v(Grid g, const T* w, const T* u, const T* F, T* w2out, T pF, T pB, int with_forcing,
      int zchunk, int tiles_x, int tiles_y, int remap, T* w1out, int* fflags) {
  s1v_body<T, TXW, TYW>(g, w, u, F, w2out, pF, pB, with_forcing, zchunk, tiles_x, tiles_y, xcd_item(remap),
              w1out, fflags);
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// SC: diffusion of omega2 at the clamped index, penalised with the per-axis sine factors
// (penalise_field_boundary_3d.py:191-208, x then y then z), into omega; F = 0. A thread's
// row is the clamped row cy; the x clamp stays inside the thread's vector (VEC >= width).
template <typename T, int VEC>
__device__ __forceinline__ void sc_body(Grid g, const T* __restrict__ w2, T* __restrict__ w,
                                        T* __restrict__ F, T pD, const Pen& pen,
                                        int zero_forcing, int tiles_x, int tiles_y,
                                        int64_t item, int chunk = kMarchChunk,
                                        const int* fflags = nullptr,
                                        S1Map s1 = S1Map{0, 0, 1, kS1Chunk}) {
  using V = VecT<T, VEC>;
  const MarchTile m = march_tile<VEC>(g, tiles_x, tiles_y, item, chunk);
  if (zero_forcing && fflags && m.x0 < int(g.nx) && m.y < int(g.ny)) {
    // the S1 tile holding these cells saw no bit of F set: F is zero there already
    const int t = ((m.z0 / s1.chunk) * s1.tiles_y + m.y / s1.rows) * s1.tiles_x + m.x0 / s1.cols;
    if (fflags[t] == 0) zero_forcing = 0;
  }
  const int nz = int(g.nz), ny = int(g.ny), nx = int(g.nx);
  const int n = nz * ny * nx, sz = ny * nx, sy = nx;
  const bool act = m.x0 < nx && m.y < ny;  // whole vectors (nx % VEC == 0)
  const int W = pen.w;
  const int y = act ? m.y : 0;
  const int cy = W > 0 ? clampi(y, W - 1, ny - W) : y;
  T fy = T(1);
  if (W > 0) {
    if (y < W) fy = T(pen.f[1][y]);
    else if (y >= ny - W) fy = T(pen.f[1][W + y - (ny - W)]);
  }
  // VEC == 1: the thread's source cell is the clamped one; VEC > 1: its own vector, with the
  // clamp resolved inside it
  const int x0 = act ? m.x0 : 0;
  const int xs0 = (VEC == 1 && W > 0) ? clampi(x0, W - 1, nx - W) : x0;
  int src[VEC];
  T fx[VEC];
  for (int k = 0; k < VEC; ++k) {
    const int x = x0 + k;
    const int cx = W > 0 ? clampi(x, W - 1, nx - W) : x;
    src[k] = cx - xs0;
    fx[k] = T(1);
    if (W > 0) {
      if (x < W) fx[k] = T(pen.f[0][x]);
      else if (x >= nx - W) fx[k] = T(pen.f[0][W + x - (nx - W)]);
    }
  }
  const bool yin = cy > 0 && cy < ny - 1;
  const int row = cy * sy + xs0;
  // diffused vector of plane q from its three planes (a: q-1, q, q+1 of row cy)
  auto diffuse = [&](const T* __restrict__ f, int q, const V& am, const V& a0, const V& ap, V& out) {
    const int i = q * sz + row;
    const bool qin = act && yin && q > 0 && q < nz - 1;
    const V yp = qin ? vload<T, VEC>(f, i + sy) : vzero<T, VEC>();
    const V ym = qin ? vload<T, VEC>(f, i - sy) : vzero<T, VEC>();
    T left, right;
    x_nbrs<T, VEC>(a0, f + q * sz + cy * sy, xs0, nx, m.tx, act, left, right);
    for (int k = 0; k < VEC; ++k) {
      const int x = xs0 + k;
      const T xl = k > 0 ? a0.v[k - 1] : left;
      const T xr = k < VEC - 1 ? a0.v[k + 1] : right;
      const bool inner = qin && x > 0 && x < nx - 1;
      out.v[k] = inner ? a0.v[k] + pD * (ap.v[k] + am.v[k] + yp.v[k] + ym.v[k] + xr + xl -
                                         T(6) * a0.v[k])
                       : a0.v[k];
    }
  };
  V am[3], a0[3], ap[3];
  for (int c = 0; c < 3; ++c) {
    const T* f = w2 + c * n;
    am[c] = (act && m.z0 - 1 >= 0) ? vload<T, VEC>(f, (m.z0 - 1) * sz + row) : vzero<T, VEC>();
    a0[c] = act ? vload<T, VEC>(f, m.z0 * sz + row) : vzero<T, VEC>();
  }
#pragma unroll 2
  for (int z = m.z0; z < m.z1; ++z) {
    int cz = z;
    T fz = T(1);
    if (W > 0) {
      cz = clampi(z, W - 1, nz - W);
      if (z < W) fz = T(pen.f[2][z]);
      else if (z >= nz - W) fz = T(pen.f[2][W + z - (nz - W)]);
    }
    V out[3];
    for (int c = 0; c < 3; ++c) {
      const T* f = w2 + c * n;
      ap[c] = (act && z + 1 < nz) ? vload<T, VEC>(f, (z + 1) * sz + row) : vzero<T, VEC>();
      if (cz == z) {  // the column's own window
        diffuse(f, z, am[c], a0[c], ap[c], out[c]);
      } else {        // z penalty zone (uniform for the workgroup): the planes around cz
        const V bm = act ? vload<T, VEC>(f, (cz - 1) * sz + row) : vzero<T, VEC>();
        const V b0 = act ? vload<T, VEC>(f, cz * sz + row) : vzero<T, VEC>();
        const V bp = act ? vload<T, VEC>(f, (cz + 1) * sz + row) : vzero<T, VEC>();
        diffuse(f, cz, bm, b0, bp, out[c]);
      }
    }
    if (act) {
      const int o = z * sz + m.y * sy + m.x0;
      for (int c = 0; c < 3; ++c) {
        V r;
        for (int k = 0; k < VEC; ++k) {
          T v = out[c].v[0];  // out[c].v[src[k]] as selects (a dynamic index would spill)
          for (int j = 1; j < VEC; ++j) v = src[k] == j ? out[c].v[j] : v;
          const bool pen_cell = fx[k] != T(1) || fy != T(1) || fz != T(1) || src[k] != k ||
                                cy != m.y || cz != z || xs0 != x0;
          if (pen_cell) v = ((v * fx[k]) * fy) * fz;
          r.v[k] = v;
        }
        vstore_stream<T, VEC>(w + c * n, o, r);
        if (zero_forcing) vstore<T, VEC>(F + c * n, o, vzero<T, VEC>());
      }
    }
    for (int c = 0; c < 3; ++c) {
      am[c] = a0[c];
      a0[c] = ap[c];
    }
  }
}

template <typename T, int VEC>
__global__ void __launch_bounds__(kMarchThreads) k_sc(Grid g, const T* __restrict__ w2,
                                                      T* __restrict__ w, T* __restrict__ F, T pD,
                                                      Pen pen, int zero_forcing, int tiles_x,
                                                      int tiles_y, int remap, int chunk,
                                                      const int* fflags, S1Map s1) {
  sc_body<T, VEC>(g, w2, w, F, pD, pen, zero_forcing, tiles_x, tiles_y, xcd_item(remap), chunk,
                  fflags, s1);
}

// S2: u = p curl_h(psi) (interior), 0 on the border, + U_inf.
template <typename T, int VEC>
__device__ __forceinline__ void s2_body(Grid g, const T* __restrict__ psi, T* __restrict__ u, T p,
                                        T U0, T U1, T U2, int tiles_x, int tiles_y,
                                        int64_t item, int chunk = kMarchChunk,
                                        T* umax = nullptr) {
  using V = VecT<T, VEC>;
  const MarchTile m = march_tile<VEC>(g, tiles_x, tiles_y, item, chunk);
  T vmax = T(0);  // max over this thread's cells of |u0| + |u1| + |u2| (when umax is given)
  const int nz = int(g.nz), ny = int(g.ny), nx = int(g.nx);
  const int n = nz * ny * nx, sz = ny * nx, sy = nx;
  const bool act = m.x0 < nx && m.y < ny;
  const int y = act ? m.y : 0, x0 = act ? m.x0 : 0;
  const T* f0 = psi;
  const T* f1 = psi + n;
  const T* f2 = psi + 2 * n;
  const bool yin = act && y > 0 && y < ny - 1;
  const int row = y * sy + x0;
  // windows of psi0 and psi1 (the components differentiated along z)
  V p0m = vzero<T, VEC>(), p1m = vzero<T, VEC>(), p0c = vzero<T, VEC>(), p1c = vzero<T, VEC>();
  if (act && m.z0 - 1 >= 0) {
    p0m = vload<T, VEC>(f0, (m.z0 - 1) * sz + row);
    p1m = vload<T, VEC>(f1, (m.z0 - 1) * sz + row);
  }
  if (act) {
    p0c = vload<T, VEC>(f0, m.z0 * sz + row);
    p1c = vload<T, VEC>(f1, m.z0 * sz + row);
  }
  const T U[3] = {U0, U1, U2};
#pragma unroll 2
  for (int z = m.z0; z < m.z1; ++z) {
    const int i = z * sz + row;
    const bool zin = yin && z > 0 && z < nz - 1;
    const V p0p = (act && z + 1 < nz) ? vload<T, VEC>(f0, i + sz) : vzero<T, VEC>();
    const V p1p = (act && z + 1 < nz) ? vload<T, VEC>(f1, i + sz) : vzero<T, VEC>();
    const V q2c = act ? vload<T, VEC>(f2, i) : vzero<T, VEC>();
    const V q2p = zin ? vload<T, VEC>(f2, i + sy) : vzero<T, VEC>();
    const V q2m = zin ? vload<T, VEC>(f2, i - sy) : vzero<T, VEC>();
    const V q0p = zin ? vload<T, VEC>(f0, i + sy) : vzero<T, VEC>();
    const V q0m = zin ? vload<T, VEC>(f0, i - sy) : vzero<T, VEC>();
    T l1, r1, l2, r2;
    x_nbrs<T, VEC>(p1c, f1 + z * sz + y * sy, x0, nx, m.tx, act, l1, r1);
    x_nbrs<T, VEC>(q2c, f2 + z * sz + y * sy, x0, nx, m.tx, act, l2, r2);
    V o[3];
    for (int k = 0; k < VEC; ++k) {
      const int x = x0 + k;
      T c0 = T(0), c1 = T(0), c2 = T(0);
      if (zin && x > 0 && x < nx - 1) {
        const T f1r = k < VEC - 1 ? p1c.v[k + 1] : r1, f1l = k > 0 ? p1c.v[k - 1] : l1;
        const T f2r = k < VEC - 1 ? q2c.v[k + 1] : r2, f2l = k > 0 ? q2c.v[k - 1] : l2;
        c0 = p * (q2p.v[k] - q2m.v[k] - p1p.v[k] + p1m.v[k]);
        c1 = p * (p0p.v[k] - p0m.v[k] - f2r + f2l);
        c2 = p * (f1r - f1l - q0p.v[k] + q0m.v[k]);
      }
      o[0].v[k] = c0 + U[0];
      o[1].v[k] = c1 + U[1];
      o[2].v[k] = c2 + U[2];
      // passive_transport_flow_simulators.py:151: sum over components of |u_i|, numpy's order
      const T sabs = (fabs(o[0].v[k]) + fabs(o[1].v[k])) + fabs(o[2].v[k]);
      vmax = sabs > vmax ? sabs : vmax;
    }
    if (act)
      for (int c = 0; c < 3; ++c) vstore_stream<T, VEC>(u + c * n, i, o[c]);
    p0m = p0c;
    p1m = p1c;
    p0c = p0p;
    p1c = p1p;
  }
  if (umax) {  // one partial maximum per workgroup (inactive lanes hold 0)
    __shared__ T red[kMarchThreads];
    red[threadIdx.x] = act ? vmax : T(0);
    __syncthreads();
    for (int k = kMarchThreads / 2; k > 0; k >>= 1) {
      if (int(threadIdx.x) < k) red[threadIdx.x] = red[threadIdx.x + k] > red[threadIdx.x] ? red[threadIdx.x + k] : red[threadIdx.x];
      __syncthreads();
    }
    if (threadIdx.x == 0) umax[item] = red[0];  // one slot per work item
  }
}

template <typename T, int VEC>
__global__ void __launch_bounds__(kMarchThreads) k_s2(Grid g, const T* __restrict__ psi,
                                                      T* __restrict__ u, T p, T U0, T U1, T U2,
                                                      int tiles_x, int tiles_y, int remap, int chunk,
                                                      T* umax) {
  s2_body<T, VEC>(g, psi, u, p, U0, U1, U2, tiles_x, tiles_y, xcd_item(remap), chunk, umax);
}

// vector width for a row length and penalty width: whole 16-byte vectors per row, and the x
// clamp of the penalty zone inside one vector
template <typename T>
inline int march_vec(int64_t nx, int width) {
  const int v = 16 / int(sizeof(T));
  return (nx % v == 0 && width <= v && nx >= 2 * v) ? v : 1;
}

template <typename T>
void launch_sc(const MarchLaunch& l, Grid g, const T* w2, T* w, T* F, T pD, const Pen& pen, int wf,
               hipStream_t s, const int* fflags = nullptr, S1Map s1 = S1Map{0, 0, 1, kS1Chunk}) {
  if (l.vec == 1)
    k_sc<T, 1><<<l.blocks, kMarchThreads, 0, s>>>(g, w2, w, F, pD, pen, wf, l.tiles_x, l.tiles_y, l.remap,
                                                  l.chunk, fflags, s1);
  else
    k_sc<T, 16 / sizeof(T)><<<l.blocks, kMarchThreads, 0, s>>>(g, w2, w, F, pD, pen, wf, l.tiles_x,
                                                                l.tiles_y, l.remap, l.chunk, fflags, s1);
}

template <typename T>
void launch_s2(const MarchLaunch& l, Grid g, const T* psi, T* u, T p, T U0, T U1, T U2,
               hipStream_t s, T* umax = nullptr) {
  if (l.vec == 1)
    k_s2<T, 1><<<l.blocks, kMarchThreads, 0, s>>>(g, psi, u, p, U0, U1, U2, l.tiles_x, l.tiles_y, l.remap,
                                                  l.chunk, umax);
  else
    k_s2<T, 16 / sizeof(T)><<<l.blocks, kMarchThreads, 0, s>>>(g, psi, u, p, U0, U1, U2, l.tiles_x,
                                                                l.tiles_y, l.remap, l.chunk, umax);
}

// max over cells of |u0| + |u1| + |u2| (sum in component order, as numpy's axis-0 sum)
template <typename T>
__global__ void __launch_bounds__(256) k_max_abs_sum(const T* u, int64_t n, T* partial) {
  __shared__ T red[256];
  T m = T(0);
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const T s = (fabs(u[i]) + fabs(u[n + i])) + fabs(u[2 * n + i]);
    m = s > m ? s : m;
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] = red[threadIdx.x + k] > red[threadIdx.x] ? red[threadIdx.x + k] : red[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

template <typename T>
__global__ void __launch_bounds__(256) k_max_final(const T* partial, int nb, T* out) {
  __shared__ T red[256];
  T m = T(0);
  for (int i = threadIdx.x; i < nb; i += 256) m = partial[i] > m ? partial[i] : m;
  red[threadIdx.x] = m;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] = red[threadIdx.x + k] > red[threadIdx.x] ? red[threadIdx.x + k] : red[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}


}  // namespace
}  // namespace sopht
