// Device code of the Eulerian-Lagrangian kernels (support, weights, interpolation, the fused
// virtual-boundary force, and the deterministic spreading), shared by the single-body entry
// points (ib_3d.hip) and the batched step (batch_3d.hip). Bodies take one node (or one
// support slot) of one body; the kernels map workgroups onto them.
#pragma once

#include <cmath>
#include <cstdint>
#include <type_traits>

#include "common.h"

struct sopht_spread3d {
  int64_t capacity;   // max nodes
  int64_t padded;     // power of two >= capacity
  uint64_t* keys;     // (key << 32 | node), sorted, padded
  uint64_t* prev;     // unsorted keys of the previous call, padded
  int* flag;          // [0] out-of-domain, [1] keys changed since the plan was built,
                      // [2] fill counter of the plan's contributor list, [3] owner count,
                      // [4] largest contributor count of an owned cell (8 ints)
  // Gather plan, rebuilt only when the keys change: for support slot t = 64 p + s of sorted
  // node p that owns its cell, pcount[t] contributors starting at plist[poff[t]], each
  // (node << 6 | slot) in increasing node order; pcount[t] = 0 for the other slots.
  int* pcount;
  int* poff;
  uint32_t* plist;
  uint32_t* owners;   // owning slots t of the plan (flag[3] of them), for the list merge
  // Lower-bound table of the sorted keys over the bounding box of the node base cells (built
  // with the sort when the box fits, else bbox[0] = 0 and the plan build searches the keys):
  // bbox = {valid, z0, y0, x0, Z, Y, X}; table[l] = number of sorted keys whose cell comes
  // before box cell l (lexicographic (z, y, x) in the box), table[Z Y X] = valid keys.
  int* table;
  int* bbox;
  int64_t plan_nz, plan_ny, plan_nx;  // grid the plan was built for (host copy)
  int64_t last_n;     // node count of the last call (host copy)
  int reuse_next;     // the caller declared the next call's support indices unchanged
};

namespace sopht {
namespace {

constexpr int W4 = 4;  // 2 * interp_kernel_width

// Plain load of a flag written by an earlier launch (no read-modify-write: an atomic "read"
// from every thread of a launch serialises on the flag's address).
__device__ __forceinline__ int flag_load(const int* f) { return __atomic_load_n(f, __ATOMIC_RELAXED); }

// Wavefront-scope ordering of LDS accesses: one wavefront's LDS operations execute in order;
// the fences keep the compiler from moving accesses across this point.
__device__ __forceinline__ void ib_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Value of v in lane j (j uniform over the wavefront).
__device__ __forceinline__ float lane_value(float v, int j) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), j));
}
__device__ __forceinline__ double lane_value(double v, int j) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(b)), j));
  const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int(uint32_t(b >> 32)), j));
  return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}

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
struct VbfArgs {
  int kernel_type;
  Eul3<T> u;
  int64_t nz, ny, nx;
  const double* X;
  const double* V;
  int64_t n;
  double dx, shift, k, c;
  int64_t* idx;
  T* support;
  T* weights;
  T* lag_u;
  T* dV;
  const T* dX;
  T* F;
};

// node i of one body, lanes 0..63 of one wavefront
template <typename T>
__device__ __forceinline__ void vbf_lag_force_node(const VbfArgs<T>& A, int64_t i, int lane) {
  const int kernel_type = A.kernel_type;
  const Eul3<T>& u = A.u;
  const int64_t nz = A.nz, ny = A.ny, nx = A.nx, n = A.n;
  const double* X = A.X;
  const double* V = A.V;
  const double dx = A.dx, shift = A.shift, k = A.k, c = A.c;
  int64_t* idx = A.idx;
  T* support = A.support;
  T* weights = A.weights;
  T* lag_u = A.lag_u;
  T* dV = A.dV;
  const T* dX = A.dX;
  T* F = A.F;
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
__global__ void __launch_bounds__(256) k_vbf_lag_force(VbfArgs<T> A) {
  const int64_t i = int64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= A.n) return;  // whole wavefronts leave together; no barriers follow
  vbf_lag_force_node<T>(A, i, threadIdx.x & 63);
}

template <typename T>
__global__ void k_vbf_time_step(T* dX, const T* dV, int64_t m, T dt) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < m) dX[i] = dX[i] + dt * dV[i];
}

// -------------------------------------------------------------------- spreading
__device__ __forceinline__ void spread_key(const int64_t* idx, int64_t n, int64_t nz, int64_t ny,
                                           int64_t nx, uint64_t* prev, int* flag, int64_t i,
                                           int guard) {
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
    if (guard) atomicOr(flag + 5, 1);  // declared unchanged, but it changed
    prev[i] = key;
  }
}

// guard: the caller declared the keys unchanged and the plan kernels are not launched; a key
// that changed anyway sets flag[5], which stops the replay and is reported by the check
__global__ void k_spread_keys(const int64_t* idx, int64_t n, int64_t padded, int64_t nz,
                              int64_t ny, int64_t nx, uint64_t* prev, int* flag, int force,
                              int guard) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (force && i == 0) atomicOr(flag + 1, 1);
  if (i < padded) spread_key(idx, n, nz, ny, nx, prev, flag, i, guard);
}

// Sort of the packed keys (unique 64-bit keys: the order is stable by construction) in one
// workgroup, only when some node changed its support block since the plan was built; clears
// the plan's fill counter for the rebuild.
//
// Bitonic sort of the padded keys (a power of two, at most 8192) by one workgroup of 1024
// threads: thread t holds the E = max(1, padded / 1024) keys t E .. t E + E - 1 in registers.
// Compare-exchange steps of stride j < E run inside the thread, strides below 64 E between the
// lanes of a wavefront (shuffles), and only the longer strides through LDS with workgroup
// barriers: 10 such steps for 8192 keys (and for 2048) instead of the 91 (66) of a sort that
// takes every step through LDS. The sorted keys end in `keys` and in the LDS copy sk (read
// by the box table).
template <int B, int E, typename F>
__device__ __forceinline__ void static_for_sort(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for_sort<B + 1, E>(f);
  }
}

constexpr int kSortThreads = 1024;

template <int J, int E>
__device__ __forceinline__ void sort_in_thread(uint64_t (&v)[E], int base, int k) {
  static_for_sort<0, E>([&](auto e_) {
    constexpr int e = decltype(e_)::value;
    if constexpr ((e & J) == 0 && e + J < E) {
      const bool up = ((base + e) & k) == 0;
      const uint64_t a = v[e], b = v[e + J];
      if ((a > b) == up) {
        v[e] = b;
        v[e + J] = a;
      }
    }
  });
}

// x of lane ^ M: an immediate swizzle inside each 32-lane half for M < 32 (bit mask mode:
// and 0x1f, or 0, xor M), a permute across the halves for M = 32
template <int M>
__device__ __forceinline__ int xor_lane(int x) {
  if constexpr (M < 32) return __builtin_amdgcn_ds_swizzle(x, (M << 10) | 0x1f);
  else return __shfl_xor(x, M, 64);
}

template <int M>
__device__ __forceinline__ uint64_t xor_lane_u64(uint64_t x) {
  const int lo = int(unsigned(x)), hi = int(unsigned(x >> 32));
  return (uint64_t(unsigned(xor_lane<M>(hi))) << 32) | unsigned(xor_lane<M>(lo));
}

template <int E>
__device__ __forceinline__ void bitonic_regs(const uint64_t* prev, uint64_t* keys, int padded,
                                             uint64_t* sk) {
  const int t = int(threadIdx.x);
  const bool act = t * E < padded;
  const int base = t * E;
  uint64_t v[E];
  static_for_sort<0, E>([&](auto e_) {
    constexpr int e = decltype(e_)::value;
    v[e] = act ? prev[base + e] : ~uint64_t(0);
  });
  for (int k = 2; k <= padded; k <<= 1) {
    // strides j = k / 2, k / 4, ... in decreasing order: through LDS while j >= 64 E ...
    for (int j = k >> 1; j >= 64 * E; j >>= 1) {
      if (act)
        static_for_sort<0, E>([&](auto e_) { sk[base + decltype(e_)::value] = v[decltype(e_)::value]; });
      __syncthreads();
      if (act) {
        const bool lower = (base & j) == 0;
        static_for_sort<0, E>([&](auto e_) {
          constexpr int e = decltype(e_)::value;
          const bool up = ((base + e) & k) == 0;
          const uint64_t a = v[e], b = sk[(base + e) ^ j];
          v[e] = (lower == up) ? (a < b ? a : b) : (a < b ? b : a);
        });
      }
      __syncthreads();
    }
    // ... between the lanes of a wavefront for j = M E, M = 32, 16, ..., 1 ...
    static_for_sort<0, 6>([&](auto s_) {
      constexpr int M = 32 >> decltype(s_)::value;
      if (M * E < k) {  // j = M E <= k / 2
        const bool lower = (t & M) == 0;
        static_for_sort<0, E>([&](auto e_) {
          constexpr int e = decltype(e_)::value;
          const bool up = ((base + e) & k) == 0;
          const uint64_t a = v[e], b = xor_lane_u64<M>(v[e]);
          v[e] = (lower == up) ? (a < b ? a : b) : (a < b ? b : a);
        });
      }
    });
    // ... and inside the thread for j < E
    if constexpr (E > 4) if (4 < k) sort_in_thread<4, E>(v, base, k);
    if constexpr (E > 2) if (2 < k) sort_in_thread<2, E>(v, base, k);
    if constexpr (E > 1) sort_in_thread<1, E>(v, base, k);
  }
  if (act) {
    static_for_sort<0, E>([&](auto e_) {
      constexpr int e = decltype(e_)::value;
      sk[base + e] = v[e];
      keys[base + e] = v[e];
    });
  }
  __syncthreads();
}

// blockDim.x must be kSortThreads and padded <= 8 kSortThreads (the launches ensure both)
__device__ __forceinline__ void bitonic_block(const uint64_t* prev, uint64_t* keys, int padded,
                                              int* flag, uint64_t* sk) {
  if (flag_load(flag + 1) == 0) return;  // uniform for the workgroup
  if (threadIdx.x == 0) {
    atomicExch(flag + 2, 0);
    atomicExch(flag + 3, 0);
  }
  if (padded >= 8 * kSortThreads) bitonic_regs<8>(prev, keys, padded, sk);
  else if (padded >= 4 * kSortThreads) bitonic_regs<4>(prev, keys, padded, sk);
  else if (padded >= 2 * kSortThreads) bitonic_regs<2>(prev, keys, padded, sk);
  else bitonic_regs<1>(prev, keys, padded, sk);
}

// Box table of the sorted keys (see sopht_spread3d::table), built by the sorting workgroup
// from its LDS copy of the keys: the valid keys' bounding box (z from the first and last key,
// y and x by LDS min/max), per-cell counts in LDS, and their exclusive scan. A box of more
// than `cap` cells leaves bbox[0] = 0. Keys in key order are in lexicographic (z, y, x) order
// inside the box, so a lower bound of any cell is one table entry after clamping.
constexpr int kTableMax = 32768;

__device__ void cell_table_block(const uint64_t* sk, int padded, int64_t ny, int64_t nx, int cap,
                                 int* cnt, int* scan, int* sh, int* table, int* bb) {
  const int tid = int(threadIdx.x), nt = int(blockDim.x);
  if (tid == 0) {
    sh[0] = 0;
    sh[1] = INT32_MAX; sh[2] = -1;  // y min, max
    sh[3] = INT32_MAX; sh[4] = -1;  // x min, max
  }
  __syncthreads();
  int ymn = INT32_MAX, ymx = -1, xmn = INT32_MAX, xmx = -1;
  for (int i = tid; i < padded; i += nt) {
    const uint64_t k = sk[i];
    if (k == ~uint64_t(0)) continue;
    if (i + 1 == padded || sk[i + 1] == ~uint64_t(0)) sh[0] = i + 1;  // the last valid key
    const int64_t cell = int64_t(k >> 32);
    const int y = int((cell / nx) % ny), x = int(cell % nx);
    ymn = y < ymn ? y : ymn;
    ymx = y > ymx ? y : ymx;
    xmn = x < xmn ? x : xmn;
    xmx = x > xmx ? x : xmx;
  }
  for (int sft = 1; sft < 64; sft <<= 1) {  // wavefront reduction, then one LDS atomic each
    const int a0 = __shfl_xor(ymn, sft, 64), a1 = __shfl_xor(ymx, sft, 64);
    const int a2 = __shfl_xor(xmn, sft, 64), a3 = __shfl_xor(xmx, sft, 64);
    ymn = a0 < ymn ? a0 : ymn;
    ymx = a1 > ymx ? a1 : ymx;
    xmn = a2 < xmn ? a2 : xmn;
    xmx = a3 > xmx ? a3 : xmx;
  }
  if ((tid & 63) == 0) {
    atomicMin(&sh[1], ymn);
    atomicMax(&sh[2], ymx);
    atomicMin(&sh[3], xmn);
    atomicMax(&sh[4], xmx);
  }
  __syncthreads();
  const int nv = sh[0];
  const int64_t plane = nx * ny;
  const int z0 = nv > 0 ? int(int64_t(sk[0] >> 32) / plane) : 0;
  const int Z = nv > 0 ? int(int64_t(sk[nv - 1] >> 32) / plane) - z0 + 1 : 0;
  const int Y = sh[2] - sh[1] + 1, X = sh[4] - sh[3] + 1;
  const int64_t vol = int64_t(Z) * Y * X;
  if (nv == 0 || vol > cap) {  // uniform
    if (tid == 0) bb[0] = 0;
    return;
  }
  const int y0 = sh[1], x0 = sh[3];
  for (int l = tid; l < vol; l += nt) cnt[l] = 0;
  __syncthreads();
  for (int i = tid; i < nv; i += nt) {
    const int64_t cell = int64_t(sk[i] >> 32);
    const int z = int(cell / plane), y = int((cell / nx) % ny), x = int(cell % nx);
    atomicAdd(&cnt[((z - z0) * Y + (y - y0)) * X + (x - x0)], 1);
  }
  __syncthreads();
  const int v32 = int(vol);
  const int chunk = (v32 + nt - 1) / nt;
  const int lb = tid * chunk < v32 ? tid * chunk : v32;
  const int le = lb + chunk < v32 ? lb + chunk : v32;
  int sum = 0;
  for (int l = lb; l < le; ++l) sum += cnt[l];
  // exclusive scan of the chunk sums: inside each wavefront by shuffles, then over the
  // wavefront totals (one barrier)
  const int lane = tid & 63, wave = tid >> 6;
  int inc = sum;
  for (int sft = 1; sft < 64; sft <<= 1) {
    const int v = __shfl_up(inc, sft, 64);
    if (lane >= sft) inc += v;
  }
  if (lane == 63) scan[wave] = inc;
  __syncthreads();
  int run = inc - sum;
  for (int w2 = 0; w2 < wave; ++w2) run += scan[w2];
  for (int l = lb; l < le; ++l) {
    table[l] = run;
    run += cnt[l];
  }
  if (tid == 0) {
    table[v32] = nv;
    bb[0] = 1; bb[1] = z0; bb[2] = y0; bb[3] = x0; bb[4] = Z; bb[5] = Y; bb[6] = X;
  }
}

// Dynamic LDS of k_bitonic_lds: the keys, then cap counts, the 1024 scan slots and 8 ints.
inline int table_cap_for(int64_t padded) {
  const int64_t budget = 160 * 1024 - 512 - padded * 8 - 1024 * 4 - 8 * 4;
  const int64_t cap = budget / 4;
  return int(cap < kTableMax ? (cap > 0 ? cap : 0) : kTableMax);
}
inline size_t bitonic_lds_bytes(int64_t padded) {
  return size_t(padded) * 8 + size_t(table_cap_for(padded)) * 4 + 1024 * 4 + 8 * 4;
}

__global__ void __launch_bounds__(1024) k_bitonic_lds(const uint64_t* prev, uint64_t* keys,
                                                      int padded, int* flag, int64_t ny,
                                                      int64_t nx, int* table, int* bb, int cap) {
  extern __shared__ uint64_t sk[];
  if (flag_load(flag + 1) == 0) return;  // keys unchanged: sort and table stay valid
  if (threadIdx.x == 0) atomicExch(flag + 4, 0);
  bitonic_block(prev, keys, padded, flag, sk);
  int* cnt = reinterpret_cast<int*>(sk + padded);
  cell_table_block(sk, padded, ny, nx, cap, cnt, cnt + cap, cnt + cap + 1024, table, bb);
}

// Lower bound of cell (z, y, x) in the sorted keys from the box table: the number of valid
// keys whose cell comes before it.
__device__ __forceinline__ int64_t table_lb(const int* T, const int* bb, int64_t z, int64_t y,
                                            int64_t x) {
  const int64_t Z = bb[4], Y = bb[5], X = bb[6];
  const int64_t zz = z - bb[1];
  if (zz < 0) return 0;
  if (zz >= Z) return T[Z * Y * X];
  const int64_t yy = y - bb[2];
  if (yy < 0) return T[zz * Y * X];
  if (yy >= Y) return T[(zz + 1) * Y * X];
  int64_t xx = x - bb[3];
  xx = xx < 0 ? 0 : (xx > X ? X : xx);  // xx = X is the next row's start
  return T[(zz * Y + yy) * X + xx];
}

// Large-node path of the key sort (only when the keys changed, as bitonic_block): copy, then
// global compare-exchange steps.
__global__ void k_copy_keys(const uint64_t* src, uint64_t* dst, int64_t m, int* flag, int* bb) {
  if (flag_load(flag + 1) == 0) return;  // plan still valid: keys stay sorted
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i < m) dst[i] = src[i];
  if (i == 0) {
    atomicExch(flag + 2, 0);
    atomicExch(flag + 3, 0);
    atomicExch(flag + 4, 0);
    bb[0] = 0;  // no box table on this path: the plan build searches the keys
  }
}

// One compare-exchange step of a global-memory bitonic sort (for large node counts).
__global__ void k_bitonic_step(uint64_t* keys, int64_t padded, int64_t j, int64_t k, const int* flag) {
  if (flag_load(flag + 1) == 0) return;
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
__device__ __forceinline__ void spread_plan_slot(int64_t nz, int64_t ny, int64_t nx,
                                                 const uint64_t* keys, int64_t n, int* flag,
                                                 int* pcount, int* poff, uint32_t* plist,
                                                 int64_t gid) {
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

// Plan build for moving bodies, in two phases (only when the keys changed).
// Phase 1, one thread per support slot t = 64 p + s: slot t owns its cell c when sorted node p
// is the first of c's contributors in key order, i.e. when the first non-empty range among
// c's candidate base-cell ranges (searched in increasing key order) starts at p. An owner
// also counts its contributors m (the lengths of all 16 ranges); a workgroup then reserves
// its owners' places in the owner list and their m-entry windows of the contributor list
// with one atomic each (same-address atomics from every owner serialise). pcount[t] = m for
// owners, 0 otherwise. Ranges: a box-table lookup (two per range) when the table is valid,
// else binary searches of the sorted keys.
__device__ __forceinline__ int spread_owner_count(int64_t nz, int64_t ny, int64_t nx,
                                                  const uint64_t* keys, int64_t n, int64_t gid,
                                                  const int* table, const int* bb, bool use_table) {
  const int64_t p = gid >> 6;
  const int slot = int(gid & 63);
  const uint64_t kp = keys[p];
  if (kp == ~uint64_t(0)) return 0;  // out-of-domain node (sorted to the end)
  const int64_t cell0 = int64_t(kp >> 32);
  const int64_t bx = cell0 % nx, by = (cell0 / nx) % ny, bz = cell0 / (nx * ny);
  const int64_t cz = bz + (slot >> 4), cy = by + ((slot >> 2) & 3), cx = bx + (slot & 3);
  int64_t qx0 = cx - 3, qx1 = cx;  // inclusive base x range, the same for every row
  if (qx0 < 0) qx0 = 0;
  if (qx1 + W4 > nx) qx1 = nx - W4;
  if (use_table) {  // uniform: all 32 lookups issued before any is used
    int a[16], b[16];
#pragma unroll
    for (int g = 0; g < 16; ++g) {
      const int64_t qz = cz - (3 - (g >> 2)), qy = cy - (3 - (g & 3));  // increasing key order
      const bool row = qz >= 0 && qy >= 0 && qz + W4 <= nz && qy + W4 <= ny && qx0 <= qx1;
      a[g] = row ? int(table_lb(table, bb, qz, qy, qx0)) : 0;
      b[g] = row ? int(table_lb(table, bb, qz, qy, qx1 + 1)) : 0;
    }
    int m = 0;
    bool first = true;
#pragma unroll
    for (int g = 0; g < 16; ++g) {
      if (b[g] > a[g]) {
        if (first && a[g] != p) return 0;  // an earlier node is the first contributor
        first = false;
        m += b[g] - a[g];
      }
    }
    return m;  // p itself lies in the first non-empty range, so m >= 1
  }
  bool first = true;
  int m = 0;
  for (int g = 0; g < 16; ++g) {
    const int dz = 3 - (g >> 2), dy = 3 - (g & 3);  // increasing key order
    const int64_t qz = cz - dz, qy = cy - dy;
    if (qz < 0 || qy < 0 || qz + W4 > nz || qy + W4 > ny) continue;
    if (qx0 > qx1) continue;
    const int64_t a = lower_bound_cell(keys, n, uint64_t((qz * ny + qy) * nx + qx0));
    const int64_t b = lower_bound_cell(keys, n, uint64_t((qz * ny + qy) * nx + qx1) + 1);
    if (b > a) {
      if (first) {
        if (a != p) return 0;  // an earlier node is the first contributor
        first = false;
      }
      m += int(b - a);
    }
  }
  return m;  // p itself lies in the first non-empty range, so m >= 1
}

template <bool LDS_KEYS>
__global__ void __launch_bounds__(512) k_spread_owner(int64_t nz, int64_t ny, int64_t nx,
                                                      const uint64_t* gkeys, int64_t n, int* flag,
                                                      int* pcount, int* poff, uint32_t* owners,
                                                      const int* table, const int* bb) {
  if (flag_load(flag + 1) == 0) return;  // plan still valid (uniform for the launch)
  extern __shared__ uint64_t skeys[];
  __shared__ int wo[8], wm[8], wx[8], base_o, base_m;
  const bool use_table = bb[0] != 0;     // uniform
  const uint64_t* keys = gkeys;
  if constexpr (LDS_KEYS) {
    if (!use_table) {  // the binary searches then run on LDS
      for (int64_t i = threadIdx.x; i < n; i += blockDim.x) skeys[i] = gkeys[i];
      __syncthreads();
      keys = skeys;
    }
  }
  const int tid = int(threadIdx.x), lane = tid & 63, w = tid >> 6;
  const int64_t gid = blockIdx.x * int64_t(blockDim.x) + tid;
  const bool live = (gid >> 6) < n;
  const int m = live ? spread_owner_count(nz, ny, nx, keys, n, gid, table, bb, use_table) : 0;
  int so = m > 0 ? 1 : 0, sm = m, mx = m;  // inclusive scans (and the maximum) over the wavefront
  for (int sft = 1; sft < 64; sft <<= 1) {
    const int a = __shfl_up(so, sft, 64), b = __shfl_up(sm, sft, 64);
    const int c = __shfl_xor(mx, sft, 64);
    if (lane >= sft) {
      so += a;
      sm += b;
    }
    mx = c > mx ? c : mx;
  }
  if (lane == 63) {
    wo[w] = so;
    wm[w] = sm;
    wx[w] = mx;
  }
  __syncthreads();
  if (tid == 0) {
    int to = 0, tm = 0, tx = 0;
    for (int i = 0; i < int(blockDim.x >> 6); ++i) {
      const int a = wo[i], b = wm[i];
      wo[i] = to;
      wm[i] = tm;
      to += a;
      tm += b;
      tx = wx[i] > tx ? wx[i] : tx;
    }
    base_o = to > 0 ? atomicAdd(flag + 3, to) : 0;
    base_m = tm > 0 ? atomicAdd(flag + 2, tm) : 0;
    if (tx > 0) atomicMax(flag + 4, tx);
  }
  __syncthreads();
  if (live) {
    pcount[gid] = m;
    if (m > 0) {
      owners[base_o + wo[w] + so - 1] = uint32_t(gid);
      poff[gid] = base_m + wm[w] + sm - m;
    }
  }
}

// Phase 2, one wavefront per owner: the contributors of cell c are the nodes whose base cell
// is c - (dz, dy, dx) for (dz, dy, dx) in [0, 3]^3; lane (dz, dy, dx) holds that base cell's
// run of the sorted keys (in increasing node order). A prefix sum over the lanes places each
// run in the owner's window of m entries (reserved by phase 1); the runs are staged in the wavefront's
// LDS window when they fit, and the wavefront merges the 64 runs by node: each output is the
// minimum head over the lanes (node numbers are unique), and the contribution's support
// slot in its node's block is the lane number itself.
constexpr int kMergeWaves = 4, kMergeCap = 1024;  // LDS entries per wavefront

__device__ __forceinline__ void spread_merge_owner(int64_t nz, int64_t ny, int64_t nx,
                                                   const uint64_t* keys, int64_t n, int* flag,
                                                   int* pcount, int* poff, uint32_t* plist,
                                                   int64_t gid, int lane, uint32_t* stage,
                                                   const int* table, const int* bb, bool use_table) {
  const int64_t p = gid >> 6;
  const int slot = int(gid & 63);
  const int64_t cell0 = int64_t(keys[p] >> 32);
  const int64_t cx = cell0 % nx + (slot & 3), cy = (cell0 / nx) % ny + ((slot >> 2) & 3),
                cz = cell0 / (nx * ny) + (slot >> 4);
  const int64_t qz = cz - (lane >> 4), qy = cy - ((lane >> 2) & 3), qx = cx - (lane & 3);
  int64_t pos = 0, hi = 0;
  if (qz >= 0 && qy >= 0 && qx >= 0 && qz + W4 <= nz && qy + W4 <= ny && qx + W4 <= nx) {
    if (use_table) {
      pos = table_lb(table, bb, qz, qy, qx);
      hi = table_lb(table, bb, qz, qy, qx + 1);
    } else {
      const uint64_t qc = uint64_t((qz * ny + qy) * nx + qx);
      pos = lower_bound_cell(keys, n, qc);
      hi = lower_bound_cell(keys, n, qc + 1);
    }
  }
  const int len = int(hi - pos);
  int incl = len;  // inclusive prefix sum over the lanes
  for (int sft = 1; sft < 64; sft <<= 1) {
    const int v = __shfl_up(incl, sft, 64);
    if (lane >= sft) incl += v;
  }
  // the lanes' runs telescope to the owner pass's ranges, so the total equals pcount[gid];
  // the window reserved there is never exceeded
  const int mc = __shfl(incl, 63, 64);
  const int reserved = pcount[gid];
  const int m = mc < reserved ? mc : reserved;
  const int base = poff[gid];
  const int mine = incl - len;  // this lane's window in the stage
  if (mc <= kMergeCap) {        // uniform: rank the staged entries
    // entries (node << 6 | slot) with unique nodes: their order is the node order, and an
    // entry's place in the list is the number of staged entries below it
    ib_wave_sync();  // the previous owner's reads of the stage are done
    for (int i = 0; i < len; ++i)
      stage[mine + i] = (uint32_t(keys[pos + i] & 0xffffffffu) << 6) | uint32_t(lane);
    ib_wave_sync();
    for (int t0 = 0; t0 < m; t0 += 64) {
      if (t0 + lane < m) {
        const uint32_t e = stage[t0 + lane];
        int rank = 0;
        for (int u = 0; u < m; ++u) rank += stage[u] < e ? 1 : 0;  // broadcast reads
        plist[base + rank] = e;
      }
    }
    return;
  }
  // more contributors than the stage holds: merge the lanes' runs from the keys, the minimum
  // head over the wavefront at a time
  int i = 0;  // position within the lane's run
  for (int k = 0; k < m; ++k) {  // m is uniform over the wavefront
    uint32_t head = 0xffffffffu;
    if (i < len) head = uint32_t(keys[pos + i] & 0xffffffffu);
    uint32_t mn = head;
    for (int sft = 1; sft < 64; sft <<= 1) {
      const uint32_t o = __shfl_xor(mn, sft, 64);
      mn = o < mn ? o : mn;
    }
    if (i < len && head == mn) {
      plist[base + k] = (mn << 6) | uint32_t(lane);
      ++i;
    }
  }
}

// Keys are read from global memory (L2): with the box table the runs need two lookups per lane;
// without it the binary searches run there too.
__global__ void __launch_bounds__(64 * kMergeWaves) k_spread_merge(
    int64_t nz, int64_t ny, int64_t nx, const uint64_t* keys, int64_t n, int* flag, int* pcount,
    int* poff, const uint32_t* owners, uint32_t* plist, const int* table, const int* bb) {
  __shared__ uint32_t stage[kMergeWaves][kMergeCap];
  if (flag_load(flag + 1) == 0) return;  // plan still valid (uniform for the launch)
  const bool use_table = bb[0] != 0;     // uniform
  const int lane = int(threadIdx.x & 63);
  const int w = int(threadIdx.x >> 6);
  const int64_t waves = int64_t(gridDim.x) * kMergeWaves;
  const int64_t count = flag_load(flag + 3);
  for (int64_t o = int64_t(blockIdx.x) * kMergeWaves + w; o < count; o += waves)
    spread_merge_owner(nz, ny, nx, keys, n, flag, pcount, poff, plist, int64_t(owners[o]), lane,
                       stage[w], table, bb, use_table);
}

// contributors per cell whose list entries a replay thread loads all at once
constexpr int kReplayGather = 32;

// Plan replay for the batched step (every call; one thread per support slot): each owned cell
// adds its contributors F_lag[node] w[slot][node] in increasing node order, every product and
// sum rounded separately, exactly as the reference's serial `F[block] += F_n * w_n`. GATHER
// (the single-simulator replay, whose cells have at most kReplayGather contributors): a cell's
// list entries are loaded at once and its products formed eight at a time (19.8 -> 16.6 us at
// the 3000-node sphere). The batched replay keeps the plain loop: its dense small-grid members
// have cells with hundreds of contributors, and the gathered form made that launch slower
// (4.13 -> 4.18 ms per batched step).
template <typename T, bool GATHER>
__device__ __forceinline__ void spread_replay_slot(int ncomp, const EulW3<T>& F, int64_t nz,
                                                   int64_t ny, int64_t nx, const T* lag,
                                                   const T* weights, const uint64_t* keys,
                                                   int64_t n, const int* pcount, const int* poff,
                                                   const uint32_t* plist, int64_t gid) {
#pragma clang fp contract(off)
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
  if (GATHER && m <= kReplayGather) {
    // every list entry of the cell is loaded first and then the products in groups of eight,
    // so the loads of many contributions are in flight together; the sums still run one after
    // another in list order
    uint32_t es[kReplayGather];
    static_for_sort<0, kReplayGather>([&](auto k_) {
      constexpr int k = decltype(k_)::value;
      es[k] = k < m ? plist[o + k] : 0u;
    });
    static_for_sort<0, kReplayGather / 8>([&](auto g_) {
      constexpr int g0 = 8 * decltype(g_)::value;
      if (g0 < m) {
        T pr[8][3];
        static_for_sort<0, 8>([&](auto k_) {
          constexpr int k = decltype(k_)::value;
          const int64_t node = int64_t(es[g0 + k] >> 6);
          const bool live = g0 + k < m;
          const T w = live ? weights[int64_t(es[g0 + k] & 63u) * n + node] : T(0);
          static_for_sort<0, 3>([&](auto c_) {
            constexpr int c = decltype(c_)::value;
            pr[k][c] = (live && c < ncomp) ? lag[c * n + node] * w : T(0);
          });
        });
        static_for_sort<0, 8>([&](auto k_) {
          constexpr int k = decltype(k_)::value;
          if (g0 + k < m)
            static_for_sort<0, 3>([&](auto c_) {
              constexpr int c = decltype(c_)::value;
              if (c < ncomp) acc[c] = acc[c] + pr[k][c];
            });
        });
      }
    });
  } else {
#pragma unroll 4
    for (int k = 0; k < m; ++k) {
      const uint32_t e = plist[o + k];
      const int64_t node = int64_t(e >> 6);
      const T w = weights[int64_t(e & 63u) * n + node];
      for (int c = 0; c < ncomp; ++c) acc[c] = acc[c] + lag[c * n + node] * w;
    }
  }
  for (int c = 0; c < ncomp; ++c) F.p[c][cidx] = acc[c];
}

// Plan replay over the owner list of the plan build, in one of two forms chosen on the device
// from the plan's largest contributor count (uniform for the launch):
// - at most kThreadReplayMax contributors per cell: one thread per owned cell, the per-slot sum
//   of spread_replay_slot without the idle slots;
// - more: one wavefront per owned cell. The lanes load up to 64 contributions at once and form
//   their products F_lag[node] w[slot][node]; the products are then added to the cell's value
//   one after another in increasing node order (every lane keeps the same running sum).
// Either way every product and sum is rounded separately, exactly as the reference's serial
// `F[block] += F_n * w_n`. Clears the keys-changed flag for the next call.
constexpr int kThreadReplayMax = 32;

template <typename T>
__global__ void __launch_bounds__(256) k_spread_replay(
    int ncomp, EulW3<T> F, int64_t nz, int64_t ny, int64_t nx, const T* lag, const T* weights,
    const uint64_t* keys, int64_t n, const int* pcount, const int* poff, const uint32_t* plist,
    const uint32_t* owners, int* flag) {
#pragma clang fp contract(off)
  if (flag_load(flag + 5) != 0) return;  // a plan declared valid whose keys changed (uniform)
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicExch(flag + 1, 0);
  const int64_t count = flag_load(flag + 3);
  if (flag_load(flag + 4) <= kThreadReplayMax) {
    for (int64_t o = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; o < count;
         o += int64_t(gridDim.x) * blockDim.x)
      spread_replay_slot<T, true>(ncomp, F, nz, ny, nx, lag, weights, keys, n, pcount, poff,
                                  plist, int64_t(owners[o]));
    return;
  }
  const int lane = int(threadIdx.x & 63);
  const int64_t waves = int64_t(gridDim.x) * (blockDim.x >> 6);
  for (int64_t o = int64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6); o < count; o += waves) {
    const int64_t gid = int64_t(owners[o]);
    const int m = pcount[gid];
    const int base = poff[gid];
    const int64_t cell0 = int64_t(keys[gid >> 6] >> 32);
    const int slot = int(gid & 63);
    const int64_t cx = cell0 % nx + (slot & 3), cy = (cell0 / nx) % ny + ((slot >> 2) & 3),
                  cz = cell0 / (nx * ny) + (slot >> 4);
    const int64_t cidx = cz * F.sz + cy * F.sy + cx * F.sx;
    T acc[3] = {T(0), T(0), T(0)};
    for (int c = 0; c < ncomp; ++c) acc[c] = F.p[c][cidx];
    for (int k0 = 0; k0 < m; k0 += 64) {
      T p[3] = {T(0), T(0), T(0)};
      if (k0 + lane < m) {
        const uint32_t e = plist[base + k0 + lane];
        const int64_t node = int64_t(e >> 6);
        const T w = weights[int64_t(e & 63u) * n + node];
        for (int c = 0; c < ncomp; ++c) p[c] = lag[c * n + node] * w;
      }
      const int cnt = m - k0 < 64 ? m - k0 : 64;
      for (int j = 0; j < cnt; ++j)
        for (int c = 0; c < ncomp; ++c) acc[c] = acc[c] + lane_value(p[c], j);
    }
    if (lane == 0)
      for (int c = 0; c < ncomp; ++c) F.p[c][cidx] = acc[c];
  }
}

inline int blocks_for(int64_t n, int t) { return int((n + t - 1) / t); }

}  // namespace
}  // namespace sopht
