// Eulerian <-> Lagrangian transfer with the 4^3 delta kernel and the virtual-boundary
// penalty force (EulerianLagrangianGridCommunicator3D.py, VirtualBoundaryForcing.py), gfx950.
//
// Interpolation is a per-node gather. Spreading is deterministic and reproduces the
// reference's serial order (EulerianLagrangianGridCommunicator3D.py:365-380: nodes
// n = 0..N-1 each doing F[block(n)] += F_n w_n) without float atomics:
//   1. keys:   per node, the linear index of its block's base cell (idx - 1 per axis);
//   2. sort:   a stable sort of (key, node) pairs (bitonic on packed 64-bit keys);
//   3. gather: one thread per (sorted node, slot). The thread whose node comes first in
//      sorted order among all contributors of its target cell owns the cell; it collects
//      the contributors (16 binary searches: for each (dz, dy) the 4 candidate bases with
//      dx = 3..0 are consecutive keys) and adds their products to F[c] in increasing node
//      order, i.e. acc = F[c]; acc += F_n w_n for n ascending; F[c] = acc.
// Blocks that leave the grid set an integer flag (the reference leaves this unchecked,
// EulerianLagrangianGridCommunicator3D.py:123-124) and are skipped.
#include <algorithm>
#include <cmath>
#include <new>

#include "common.h"

struct sopht_spread3d {
  int64_t capacity;   // max nodes
  int64_t padded;     // power of two >= capacity
  uint64_t* keys;     // (key << 32 | node), sorted, padded
  uint64_t* prev;     // unsorted keys of the previous call, padded
  int* flag;          // [0] out-of-domain, [1] keys changed since the plan was built,
                      // [2] fill counter of the plan's contributor list
  // Gather plan, rebuilt only when the keys change: for support slot t = 64 p + s of sorted
  // node p that owns its cell, pcount[t] contributors starting at plist[poff[t]], each
  // (node << 6 | slot) in increasing node order; pcount[t] = 0 for the other slots.
  int* pcount;
  int* poff;
  uint32_t* plist;
  int64_t plan_nz, plan_ny, plan_nx;  // grid the plan was built for (host copy)
};

namespace sopht {
namespace {

constexpr int W4 = 4;  // 2 * interp_kernel_width

// numpy floor_divide for doubles (npy_divmod): exact floor of a / b including the
// cases where a / b rounds onto an integer.
__device__ __forceinline__ double np_floor_divide(double a, double b) {
  const double mod = fmod(a, b);
  double div = (a - mod) / b;
  if (mod != 0.0 && ((b < 0.0) != (mod < 0.0))) div -= 1.0;
  double fl;
  if (div != 0.0) {
    fl = floor(div);
    if (div - fl > 0.5) fl += 1.0;
  } else {
    fl = copysign(0.0, a / b);
  }
  return fl;
}

template <typename T>
__device__ __forceinline__ T cosine_phi(T s) {
  return T(1) + cos(T(0.5 * M_PI) * s);
}

template <typename T>
__device__ __forceinline__ T peskin_phi(T r) {
  // Peskin (2002) eq. 6.27 in |d|/dx (EulerianLagrangianGridCommunicator3D.py:434-516)
  T v = T(0);
  if (r < T(1)) v = T(3) - T(2) * r + sqrt(fabs(T(1) + T(4) * r - T(4) * r * r));
  else if (r < T(2)) v = T(5) - T(2) * r - sqrt(fabs(T(-7) + T(12) * r - T(4) * r * r));
  return v;
}

// support + weights for one node; writes idx, support (already divided by dx, or |d|/dx for
// Peskin, as the reference leaves it) and weights; returns block base (idx - 1).
template <typename T>
__device__ void node_support_weights(int kernel_type, const double* X, int64_t n, int64_t i,
                                     double dx, double shift, int64_t* idx_out, T* support,
                                     T* weights, int64_t base[3], T wloc[64]) {
  int64_t idx[3];
  double d[3][W4];
  for (int a = 0; a < 3; ++a) {
    const double xa = X[a * n + i];
    idx[a] = int64_t(np_floor_divide(xa - shift, dx));
    if (idx_out) idx_out[a * n + i] = idx[a];
    base[a] = idx[a] - 1;
    for (int o = 0; o < W4; ++o) d[a][o] = double(idx[a] + (o - 1)) * dx + shift - xa;
  }
  // reference: support stored in real_t, then divided by dx (cosine) / |.|/dx (peskin)
  const T tdx = T(dx);
  T s[3][W4];
  for (int a = 0; a < 3; ++a)
    for (int o = 0; o < W4; ++o) {
      const T dr = T(d[a][o]);
      s[a][o] = kernel_type == 0 ? dr / tdx : fabs(dr) / tdx;
    }
  T ph[3][W4];
  for (int a = 0; a < 3; ++a)
    for (int o = 0; o < W4; ++o) ph[a][o] = kernel_type == 0 ? cosine_phi<T>(s[a][o]) : peskin_phi<T>(s[a][o]);
  const T pre = kernel_type == 0 ? T((T(0.25) / tdx) * (T(0.25) / tdx) * (T(0.25) / tdx))
                                 : T((T(0.125) / tdx) * (T(0.125) / tdx) * (T(0.125) / tdx));
  for (int kz = 0; kz < W4; ++kz)
    for (int ky = 0; ky < W4; ++ky)
      for (int kx = 0; kx < W4; ++kx) {
        const int slot = (kz * W4 + ky) * W4 + kx;
        const T w = pre * ph[0][kx] * ph[1][ky] * ph[2][kz];
        wloc[slot] = w;
        if (weights) weights[int64_t(slot) * n + i] = w;
        if (support) {
          support[(int64_t(0 * 64) + slot) * n + i] = s[0][kx];
          support[(int64_t(1 * 64) + slot) * n + i] = s[1][ky];
          support[(int64_t(2 * 64) + slot) * n + i] = s[2][kz];
        }
      }
}

template <typename T>
__global__ void k_support(const double* X, int64_t n, double dx, double shift, int64_t* idx,
                          T* support) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  for (int a = 0; a < 3; ++a) {
    const double xa = X[a * n + i];
    const int64_t ia = int64_t(np_floor_divide(xa - shift, dx));
    idx[a * n + i] = ia;
    for (int kz = 0; kz < W4; ++kz)
      for (int ky = 0; ky < W4; ++ky)
        for (int kx = 0; kx < W4; ++kx) {
          const int o = a == 0 ? kx : (a == 1 ? ky : kz);
          const int slot = (kz * W4 + ky) * W4 + kx;
          support[(int64_t(a * 64) + slot) * n + i] = T(double(ia + (o - 1)) * dx + shift - xa);
        }
  }
}

template <typename T>
__global__ void k_weights(int kernel_type, T* weights, T* support, int64_t n, T dx) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const T pre = kernel_type == 0 ? T((T(0.25) / dx) * (T(0.25) / dx) * (T(0.25) / dx))
                                 : T((T(0.125) / dx) * (T(0.125) / dx) * (T(0.125) / dx));
  for (int slot = 0; slot < 64; ++slot) {
    T w = pre;
    for (int a = 0; a < 3; ++a) {
      T& sv = support[(int64_t(a * 64) + slot) * n + i];
      const T s = kernel_type == 0 ? sv / dx : fabs(sv) / dx;
      sv = s;
      w = w * (kernel_type == 0 ? cosine_phi<T>(s) : peskin_phi<T>(s));
    }
    weights[int64_t(slot) * n + i] = w;
  }
}

template <typename T>
struct Eul3 {
  const T* p[3];
  int64_t sz, sy, sx;
};

template <typename T>
__device__ __forceinline__ bool block_inside(const int64_t base[3], int64_t nz, int64_t ny,
                                             int64_t nx) {
  return base[0] >= 0 && base[1] >= 0 && base[2] >= 0 && base[0] + W4 <= nx &&
         base[1] + W4 <= ny && base[2] + W4 <= nz;
}

// Gather: U_n = dx^3 * sum over the block (kz, ky, kx order) of u * w.
template <typename T>
__device__ void gather_node(const Eul3<T>& u, int ncomp, const int64_t base[3], const T* w,
                            int64_t wstride, T dx3, T out[3]) {
  for (int c = 0; c < ncomp; ++c) {
    T acc = T(0);
    const T* f = u.p[c];
    for (int kz = 0; kz < W4; ++kz)
      for (int ky = 0; ky < W4; ++ky)
        for (int kx = 0; kx < W4; ++kx) {
          const int slot = (kz * W4 + ky) * W4 + kx;
          acc += f[(base[2] + kz) * u.sz + (base[1] + ky) * u.sy + (base[0] + kx) * u.sx] *
                 w[slot * wstride];
        }
    out[c] = acc * dx3;
  }
}

template <typename T>
__global__ void k_interp(int ncomp, T* lag, Eul3<T> u, int64_t nz, int64_t ny, int64_t nx,
                         const T* weights, const int64_t* idx, int64_t n, T dx3, int* flag) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const int64_t base[3] = {idx[i] - 1, idx[n + i] - 1, idx[2 * n + i] - 1};
  T out[3] = {T(0), T(0), T(0)};
  if (block_inside<T>(base, nz, ny, nx)) {
    gather_node<T>(u, ncomp, base, weights + i, n, dx3, out);
  } else if (flag) {
    atomicOr(flag, 1);
  }
  for (int c = 0; c < ncomp; ++c) lag[c * n + i] = out[c];
}

// Fused virtual-boundary force on the Lagrangian grid: one wavefront per node, lane =
// support slot (kz, ky, kx). Each lane computes its weight, reads u at its cell, and the
// 64 products are summed with a butterfly reduction; lane 0 forms U, dV and F.
template <typename T>
__global__ void __launch_bounds__(256) k_vbf_lag_force(int kernel_type, Eul3<T> u, int64_t nz,
                                                       int64_t ny, int64_t nx, const double* X,
                                                       const double* V, int64_t n, double dx,
                                                       double shift, double k, double c,
                                                       int64_t* idx, T* support, T* weights,
                                                       T* lag_u, T* dV, const T* dX, T* F) {
  const int lane = threadIdx.x & 63;
  const int64_t i = int64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= n) return;  // whole wavefronts leave together; no barriers follow
  const int off[3] = {lane & 3, (lane >> 2) & 3, lane >> 4};  // kx, ky, kz
  const T tdx = T(dx);
  int64_t base[3];
  T s[3];
  for (int a = 0; a < 3; ++a) {
    const double xa = X[a * n + i];
    const int64_t ia = int64_t(np_floor_divide(xa - shift, dx));
    base[a] = ia - 1;
    if (lane == 0) idx[a * n + i] = ia;
    // reference: support stored in real_t, then divided by dx (cosine) or |.|/dx (peskin)
    const T dr = T(double(ia + off[a] - 1) * dx + shift - xa);
    s[a] = kernel_type == 0 ? dr / tdx : fabs(dr) / tdx;
    support[(int64_t(a) * 64 + lane) * n + i] = s[a];
  }
  const T pre = kernel_type == 0 ? T((T(0.25) / tdx) * (T(0.25) / tdx) * (T(0.25) / tdx))
                                 : T((T(0.125) / tdx) * (T(0.125) / tdx) * (T(0.125) / tdx));
  T w = pre;
  for (int a = 0; a < 3; ++a) w = w * (kernel_type == 0 ? cosine_phi<T>(s[a]) : peskin_phi<T>(s[a]));
  weights[int64_t(lane) * n + i] = w;
  T p[3] = {T(0), T(0), T(0)};
  if (block_inside<T>(base, nz, ny, nx)) {
    const int64_t cell = (base[2] + off[2]) * u.sz + (base[1] + off[1]) * u.sy + (base[0] + off[0]) * u.sx;
    for (int a = 0; a < 3; ++a) p[a] = u.p[a][cell] * w;
  }
  for (int m = 32; m >= 1; m >>= 1)
    for (int a = 0; a < 3; ++a) p[a] += __shfl_xor(p[a], m, 64);
  if (lane == 0) {
    const T dx3 = tdx * tdx * tdx;
    for (int a = 0; a < 3; ++a) {
      const T U = p[a] * dx3;
      lag_u[a * n + i] = U;
      const T dv = T(U - V[a * n + i]);
      dV[a * n + i] = dv;
      // F = k dX + c dV evaluated in FP64 (the reference's rescaled coefficients are float64)
      F[a * n + i] = T(k * double(dX[a * n + i]) + c * double(dv));
    }
  }
}

template <typename T>
__global__ void k_vbf_time_step(T* dX, const T* dV, int64_t m, T dt) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < m) dX[i] = dX[i] + dt * dV[i];
}

// -------------------------------------------------------------------- spreading
__global__ void k_spread_keys(const int64_t* idx, int64_t n, int64_t padded, int64_t nz,
                              int64_t ny, int64_t nx, uint64_t* prev, int* flag, int force) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (force && i == 0) atomicOr(flag + 1, 1);
  if (i >= padded) return;
  uint64_t key = ~uint64_t(0);
  if (i < n) {
    const int64_t base[3] = {idx[i] - 1, idx[n + i] - 1, idx[2 * n + i] - 1};
    if (block_inside<float>(base, nz, ny, nx)) {
      const uint64_t cell = uint64_t((base[2] * ny + base[1]) * nx + base[0]);
      key = (cell << 32) | uint64_t(i);
    } else {
      atomicOr(flag, 1);
    }
  }
  // the sorted order and the gather plan are reused while every node keeps its support block
  if (prev[i] != key) {
    atomicOr(flag + 1, 1);
    prev[i] = key;
  }
}

// Stable sort of the packed keys (bitonic on unique 64-bit keys) in one workgroup, only
// when some node changed its support block since the plan was built; clears the plan's fill
// counter for the rebuild.
__global__ void __launch_bounds__(1024) k_bitonic_lds(const uint64_t* prev, uint64_t* keys,
                                                      int padded, int* flag) {
  extern __shared__ uint64_t sk[];
  if (atomicAdd(flag + 1, 0) == 0) return;  // uniform for the workgroup
  if (threadIdx.x == 0) atomicExch(flag + 2, 0);
  for (int i = threadIdx.x; i < padded; i += blockDim.x) sk[i] = prev[i];
  __syncthreads();
  for (int k = 2; k <= padded; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < padded; i += blockDim.x) {
        const int p = i ^ j;
        if (p > i) {
          const bool up = (i & k) == 0;
          const uint64_t a = sk[i], b = sk[p];
          if ((a > b) == up) {
            sk[i] = b;
            sk[p] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < padded; i += blockDim.x) keys[i] = sk[i];
}

__global__ void k_copy_keys(const uint64_t* src, uint64_t* dst, int64_t m, int* flag) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < m) dst[i] = src[i];
  if (i == 0) atomicExch(flag + 2, 0);
}

// One compare-exchange step of a global-memory bitonic sort (for large node counts).
__global__ void k_bitonic_step(uint64_t* keys, int64_t padded, int64_t j, int64_t k) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= padded) return;
  const int64_t p = i ^ j;
  if (p > i) {
    const bool up = (i & k) == 0;
    const uint64_t a = keys[i], b = keys[p];
    if ((a > b) == up) {
      keys[i] = b;
      keys[p] = a;
    }
  }
}

// first position in keys[0, n) with (key >> 32) >= cell
__device__ __forceinline__ int64_t lower_bound_cell(const uint64_t* keys, int64_t n,
                                                    uint64_t cell) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((keys[mid] >> 32) < cell) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

template <typename T>
struct EulW3 {
  T* p[3];
  int64_t sz, sy, sx;
};

// Plan build (only when the keys changed): support slot t = 64 p + s of sorted node p owns
// its cell c when p is c's first contributor in key order. Contributors of c are the nodes
// whose base cell lies in c - [0, 3]^3: for each (dz, dy) they form one contiguous range of
// the sorted keys (base x from cx - 3 to cx). The owner lists them in increasing node order.
template <bool LDS_KEYS>
__global__ void __launch_bounds__(512) k_spread_plan(int64_t nz, int64_t ny, int64_t nx,
                                                     const uint64_t* gkeys, int64_t n, int* flag,
                                                     int* pcount, int* poff, uint32_t* plist) {
  if (atomicAdd(flag + 1, 0) == 0) return;  // plan still valid (uniform for the launch)
  extern __shared__ uint64_t skeys[];
  const uint64_t* keys = gkeys;
  if constexpr (LDS_KEYS) {
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) skeys[i] = gkeys[i];
    __syncthreads();
    keys = skeys;
  }
  const int64_t gid = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  const int64_t p = gid >> 6;
  const int slot = int(gid & 63);
  if (p >= n) return;
  const uint64_t kp = keys[p];
  if (kp == ~uint64_t(0)) {  // out-of-domain node (sorted to the end)
    pcount[gid] = 0;
    return;
  }
  const int64_t cell0 = int64_t(kp >> 32);
  const int64_t bx = cell0 % nx, by = (cell0 / nx) % ny, bz = cell0 / (nx * ny);
  const int kz = slot >> 4, ky = (slot >> 2) & 3, kx = slot & 3;
  const int64_t cz = bz + kz, cy = by + ky, cx = bx + kx;
  int64_t lo[16], hi[16];
  bool owner = true;
  bool found_first = false;
  int m = 0;
#pragma unroll
  for (int g = 0; g < 16; ++g) {
    const int dz = 3 - (g >> 2), dy = 3 - (g & 3);  // increasing key order
    const int64_t qz = cz - dz, qy = cy - dy;
    lo[g] = hi[g] = 0;
    if (qz < 0 || qy < 0 || qz + W4 > nz || qy + W4 > ny) continue;
    int64_t qx0 = cx - 3, qx1 = cx;  // inclusive base x range
    if (qx0 < 0) qx0 = 0;
    if (qx1 + W4 > nx) qx1 = nx - W4;
    if (qx0 > qx1) continue;
    const uint64_t c0 = uint64_t((qz * ny + qy) * nx + qx0);
    const uint64_t c1 = uint64_t((qz * ny + qy) * nx + qx1);
    const int64_t a = lower_bound_cell(keys, n, c0);
    const int64_t b = lower_bound_cell(keys, n, c1 + 1);
    lo[g] = a;
    hi[g] = b;
    m += int(b - a);
    if (!found_first && a < b) {
      found_first = true;
      if (a != p) {
        owner = false;
        break;
      }
    }
  }
  if (!owner) {
    pcount[gid] = 0;
    return;
  }
  // the list's place in plist depends on scheduling; its contents do not
  const int base = atomicAdd(flag + 2, m);
  pcount[gid] = m;
  poff[gid] = base;
  int64_t last = -1;
  for (int k = 0; k < m; ++k) {
    int64_t best = INT64_MAX;
    int64_t best_pos = -1;
#pragma unroll
    for (int g = 0; g < 16; ++g)
      for (int64_t q = lo[g]; q < hi[g]; ++q) {
        const int64_t node = int64_t(keys[q] & 0xffffffffu);
        if (node > last && node < best) {
          best = node;
          best_pos = q;
        }
      }
    const int64_t cell = int64_t(keys[best_pos] >> 32);
    const int64_t qx = cell % nx, qy = (cell / nx) % ny, qz = cell / (nx * ny);
    const int sl = int(((cz - qz) * W4 + (cy - qy)) * W4 + (cx - qx));
    plist[base + k] = (uint32_t(best) << 6) | uint32_t(sl);
    last = best;
  }
}

// Plan replay (every call): each owned cell adds its contributors F_lag[node] w[slot][node]
// in increasing node order, every product and sum rounded separately, exactly as the
// reference's serial `F[block] += F_n * w_n`. Clears the keys-changed flag for the next call.
template <typename T>
__global__ void __launch_bounds__(256) k_spread_replay(int ncomp, EulW3<T> F, int64_t nz,
                                                       int64_t ny, int64_t nx, const T* lag,
                                                       const T* weights, const uint64_t* keys,
                                                       int64_t n, const int* pcount,
                                                       const int* poff, const uint32_t* plist,
                                                       int* flag) {
#pragma clang fp contract(off)
  const int64_t gid = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (gid == 0) atomicExch(flag + 1, 0);
  if (gid >= 64 * n) return;
  const int m = pcount[gid];
  if (m == 0) return;
  const int64_t cell0 = int64_t(keys[gid >> 6] >> 32);
  const int slot = int(gid & 63);
  const int64_t cx = cell0 % nx + (slot & 3), cy = (cell0 / nx) % ny + ((slot >> 2) & 3),
                cz = cell0 / (nx * ny) + (slot >> 4);
  const int64_t cidx = cz * F.sz + cy * F.sy + cx * F.sx;
  T acc[3];
  for (int c = 0; c < ncomp; ++c) acc[c] = F.p[c][cidx];
  const int o = poff[gid];
  for (int k = 0; k < m; ++k) {
    const uint32_t e = plist[o + k];
    const int64_t node = int64_t(e >> 6);
    const T w = weights[int64_t(e & 63u) * n + node];
    for (int c = 0; c < ncomp; ++c) acc[c] = acc[c] + lag[c * n + node] * w;
  }
  for (int c = 0; c < ncomp; ++c) F.p[c][cidx] = acc[c];
}

inline int blocks_for(int64_t n, int t) { return int((n + t - 1) / t); }

}  // namespace
}  // namespace sopht

using namespace sopht;

#define IB_DISPATCH(dtype, CALL)                                       \
  do {                                                                 \
    if ((dtype) == SOPHT_F32) {                                        \
      using T = float;                                                 \
      CALL;                                                            \
    } else if ((dtype) == SOPHT_F64) {                                 \
      using T = double;                                                \
      CALL;                                                            \
    } else {                                                           \
      set_last_error("dtype must be SOPHT_F32 or SOPHT_F64", __FILE__, __LINE__); \
      return SOPHT_ERR_INVALID;                                        \
    }                                                                  \
  } while (0)

extern "C" {

int sopht_lag_support_3d(int dtype, const double* X, int64_t* idx, void* support, int64_t n,
                         double dx, double shift, void* stream) {
  SOPHT_REQUIRE(n >= 0 && dx > 0, "invalid node count or grid spacing");
  if (n == 0) return SOPHT_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  IB_DISPATCH(dtype, (k_support<T><<<blocks_for(n, 128), 128, 0, s>>>(X, n, dx, shift, idx,
                                                                     static_cast<T*>(support))));
  SOPHT_CHECK_LAUNCH();
  return SOPHT_OK;
}

int sopht_lag_interp_weights_3d(int dtype, int kernel_type, void* weights, void* support,
                                int64_t n, double dx, void* stream) {
  SOPHT_REQUIRE(kernel_type == 0 || kernel_type == 1, "kernel_type must be 0 (cosine) or 1 (peskin)");
  if (n == 0) return SOPHT_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  IB_DISPATCH(dtype, (k_weights<T><<<blocks_for(n, 128), 128, 0, s>>>(
                         kernel_type, static_cast<T*>(weights), static_cast<T*>(support), n, T(dx))));
  SOPHT_CHECK_LAUNCH();
  return SOPHT_OK;
}

int sopht_eul_to_lag_interp_3d(int dtype, int ncomp, void* lag, const void* const eul[3],
                               const int64_t* es, int64_t nz, int64_t ny, int64_t nx,
                               const void* weights, const int64_t* idx, int64_t n, double dx,
                               void* stream) {
  SOPHT_REQUIRE(ncomp == 1 || ncomp == 3, "Invalid number of components for interpolation, must be either 1 or 3");
  if (n == 0) return SOPHT_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  IB_DISPATCH(dtype, ([&] {
                Eul3<T> u;
                for (int c = 0; c < 3; ++c) u.p[c] = c < ncomp ? static_cast<const T*>(eul[c]) : nullptr;
                u.sz = es[0]; u.sy = es[1]; u.sx = es[2];
                const T tdx = T(dx);
                k_interp<T><<<blocks_for(n, 128), 128, 0, s>>>(ncomp, static_cast<T*>(lag), u, nz, ny, nx,
                                                              static_cast<const T*>(weights), idx, n,
                                                              tdx * tdx * tdx, nullptr);
              }()));
  SOPHT_CHECK_LAUNCH();
  return SOPHT_OK;
}

int sopht_spread3d_create(int64_t max_nodes, sopht_spread3d** out) {
  SOPHT_REQUIRE(out != nullptr && max_nodes >= 0, "invalid spread workspace arguments");
  SOPHT_REQUIRE(max_nodes < (int64_t(1) << 25), "too many Lagrangian nodes (at most 2^25)");
  sopht_spread3d* w = new (std::nothrow) sopht_spread3d();
  SOPHT_REQUIRE(w != nullptr, "out of host memory");
  w->capacity = max_nodes;
  int64_t p = 1;
  while (p < max_nodes) p <<= 1;
  w->padded = p;
  const int init_flags[3] = {0, 1, 0};  // no out-of-domain node yet; first call builds
  const int64_t slots = 64 * (max_nodes > 0 ? max_nodes : 1);
  if (hipMalloc(&w->keys, sizeof(uint64_t) * p) != hipSuccess ||
      hipMalloc(&w->prev, sizeof(uint64_t) * p) != hipSuccess ||
      hipMalloc(&w->flag, 3 * sizeof(int)) != hipSuccess ||
      hipMalloc(&w->pcount, sizeof(int) * slots) != hipSuccess ||
      hipMalloc(&w->poff, sizeof(int) * slots) != hipSuccess ||
      hipMalloc(&w->plist, sizeof(uint32_t) * slots) != hipSuccess ||
      hipMemset(w->prev, 0xff, sizeof(uint64_t) * p) != hipSuccess ||
      hipMemcpy(w->flag, init_flags, sizeof(init_flags), hipMemcpyHostToDevice) != hipSuccess) {
    sopht_spread3d_destroy(w);
    set_last_error("hipMalloc of the spread workspace failed", __FILE__, __LINE__);
    return SOPHT_ERR_HIP;
  }
  *out = w;
  return SOPHT_OK;
}

int sopht_spread3d_destroy(sopht_spread3d* w) {
  if (!w) return SOPHT_OK;
  if (w->keys) (void)hipFree(w->keys);
  if (w->prev) (void)hipFree(w->prev);
  if (w->flag) (void)hipFree(w->flag);
  if (w->pcount) (void)hipFree(w->pcount);
  if (w->poff) (void)hipFree(w->poff);
  if (w->plist) (void)hipFree(w->plist);
  delete w;
  return SOPHT_OK;
}

int sopht_lag_to_eul_spread_3d(int dtype, int ncomp, sopht_spread3d* ws, void* const eul[3],
                               const int64_t* es, int64_t nz, int64_t ny, int64_t nx,
                               const void* lag, const void* weights, const int64_t* idx,
                               int64_t n, void* stream) {
  SOPHT_REQUIRE(ws != nullptr, "spread workspace is null");
  SOPHT_REQUIRE(ncomp == 1 || ncomp == 3, "Invalid number of components for interpolation, must be either 1 or 3");
  SOPHT_REQUIRE(n <= ws->capacity, "more Lagrangian nodes than the spread workspace holds");
  SOPHT_REQUIRE(nz * ny * nx < (int64_t(1) << 32), "grid too large for 32-bit cell keys");
  if (n == 0) return SOPHT_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  int64_t padded = 1;
  while (padded < n) padded <<= 1;
  // keys of this call into ws->prev (and the "changed" flag) over the workspace's padded
  // length, so a later call with fewer nodes still sees its stale tail as changed
  // a plan built for another grid is rebuilt even if the packed keys happen to agree
  const int force = (ws->plan_nz != nz || ws->plan_ny != ny || ws->plan_nx != nx) ? 1 : 0;
  ws->plan_nz = nz;
  ws->plan_ny = ny;
  ws->plan_nx = nx;
  k_spread_keys<<<blocks_for(ws->padded, 256), 256, 0, s>>>(idx, n, ws->padded, nz, ny, nx,
                                                            ws->prev, ws->flag, force);
  SOPHT_CHECK_LAUNCH();
  if (padded <= 8192) {
    k_bitonic_lds<<<1, 1024, padded * sizeof(uint64_t), s>>>(ws->prev, ws->keys, int(padded), ws->flag);
    SOPHT_CHECK_LAUNCH();
  } else {
    k_copy_keys<<<blocks_for(padded, 256), 256, 0, s>>>(ws->prev, ws->keys, padded, ws->flag);
    SOPHT_CHECK_LAUNCH();
    for (int64_t k = 2; k <= padded; k <<= 1)
      for (int64_t j = k >> 1; j > 0; j >>= 1) {
        k_bitonic_step<<<blocks_for(padded, 256), 256, 0, s>>>(ws->keys, padded, j, k);
        SOPHT_CHECK_LAUNCH();
      }
  }
  // sorted keys in LDS for the plan build's binary searches when they fit (128 KiB)
  if (n <= 16384)
    k_spread_plan<true><<<blocks_for(n * 64, 512), 512, n * sizeof(uint64_t), s>>>(
        nz, ny, nx, ws->keys, n, ws->flag, ws->pcount, ws->poff, ws->plist);
  else
    k_spread_plan<false><<<blocks_for(n * 64, 512), 512, 0, s>>>(nz, ny, nx, ws->keys, n, ws->flag,
                                                                 ws->pcount, ws->poff, ws->plist);
  SOPHT_CHECK_LAUNCH();
  IB_DISPATCH(dtype, ([&] {
                EulW3<T> F;
                for (int c = 0; c < 3; ++c) F.p[c] = c < ncomp ? static_cast<T*>(eul[c]) : nullptr;
                F.sz = es[0]; F.sy = es[1]; F.sx = es[2];
                k_spread_replay<T><<<blocks_for(n * 64, 256), 256, 0, s>>>(
                    ncomp, F, nz, ny, nx, static_cast<const T*>(lag), static_cast<const T*>(weights),
                    ws->keys, n, ws->pcount, ws->poff, ws->plist, ws->flag);
              }()));
  SOPHT_CHECK_LAUNCH();
  return SOPHT_OK;
}

int sopht_spread3d_check(sopht_spread3d* ws, void* stream) {
  SOPHT_REQUIRE(ws != nullptr, "spread workspace is null");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int h = 0;
  SOPHT_CHECK_HIP(hipMemcpyAsync(&h, ws->flag, sizeof(int), hipMemcpyDeviceToHost, s));
  SOPHT_CHECK_HIP(hipStreamSynchronize(s));
  if (h) {
    SOPHT_CHECK_HIP(hipMemsetAsync(ws->flag, 0, sizeof(int), s));
    set_last_error("a Lagrangian node's 4^3 support leaves the Eulerian grid", __FILE__, __LINE__);
    return SOPHT_ERR_OUT_OF_DOMAIN;
  }
  return SOPHT_OK;
}

int sopht_vbf_lag_force_3d(int dtype, int kernel_type, const void* const vel[3],
                           const int64_t* vs, int64_t nz, int64_t ny, int64_t nx,
                           const double* X, const double* V, int64_t n, double dx, double shift,
                           double k, double c, int64_t* idx, void* support, void* weights,
                           void* lag_u, void* dV, const void* dX, void* F, void* stream) {
  SOPHT_REQUIRE(kernel_type == 0 || kernel_type == 1, "kernel_type must be 0 (cosine) or 1 (peskin)");
  if (n == 0) return SOPHT_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  IB_DISPATCH(dtype, ([&] {
                Eul3<T> u;
                for (int a = 0; a < 3; ++a) u.p[a] = static_cast<const T*>(vel[a]);
                u.sz = vs[0]; u.sy = vs[1]; u.sx = vs[2];
                k_vbf_lag_force<T><<<blocks_for(n, 4), 256, 0, s>>>(
                    kernel_type, u, nz, ny, nx, X, V, n, dx, shift, k, c, idx,
                    static_cast<T*>(support), static_cast<T*>(weights), static_cast<T*>(lag_u),
                    static_cast<T*>(dV), static_cast<const T*>(dX), static_cast<T*>(F));
              }()));
  SOPHT_CHECK_LAUNCH();
  return SOPHT_OK;
}

int sopht_vbf_time_step_3d(int dtype, void* dX, const void* dV, int64_t n, double dt,
                           void* stream) {
  if (n == 0) return SOPHT_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  IB_DISPATCH(dtype, (k_vbf_time_step<T><<<blocks_for(3 * n, 256), 256, 0, s>>>(
                         static_cast<T*>(dX), static_cast<const T*>(dV), 3 * n, T(dt))));
  SOPHT_CHECK_LAUNCH();
  return SOPHT_OK;
}

}  // extern "C"
