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
#include <mutex>
#include <new>

#include "common.h"
#include "ib_kernels.h"



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
  const int init_flags[8] = {0, 1, 0, 0, 0, 0, 0, 0};  // no out-of-domain node yet; first call builds
  const int64_t slots = 64 * (max_nodes > 0 ? max_nodes : 1);
  if (hipMalloc(&w->keys, sizeof(uint64_t) * p) != hipSuccess ||
      hipMalloc(&w->prev, sizeof(uint64_t) * p) != hipSuccess ||
      hipMalloc(&w->flag, 8 * sizeof(int)) != hipSuccess ||
      hipMalloc(&w->pcount, sizeof(int) * slots) != hipSuccess ||
      hipMalloc(&w->poff, sizeof(int) * slots) != hipSuccess ||
      hipMalloc(&w->plist, sizeof(uint32_t) * slots) != hipSuccess ||
      hipMalloc(&w->owners, sizeof(uint32_t) * slots) != hipSuccess ||
      hipMalloc(&w->table, sizeof(int) * (kTableMax + 1)) != hipSuccess ||
      hipMalloc(&w->bbox, sizeof(int) * 8) != hipSuccess ||
      hipMemset(w->bbox, 0, sizeof(int) * 8) != hipSuccess ||
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
  if (w->owners) (void)hipFree(w->owners);
  if (w->table) (void)hipFree(w->table);
  if (w->bbox) (void)hipFree(w->bbox);
  delete w;
  return SOPHT_OK;
}

}  // extern "C"

namespace {

// Spreading through a workspace (the kernels of both spreading entry points). force: rebuild
// the plan whatever the keys (a workspace whose previous contents are not this caller's).
int spread_impl(int dtype, int ncomp, sopht_spread3d* ws, void* const eul[3], const int64_t* es,
                int64_t nz, int64_t ny, int64_t nx, const void* lag, const void* weights,
                const int64_t* idx, int64_t n, hipStream_t s, int force) {
  int64_t padded = 1;
  while (padded < n) padded <<= 1;
  // the caller's declaration (sopht_spread3d_reuse_plan) holds for this call only, and only
  // for the same node count and grid as the call that left the plan
  const int reuse = (ws->reuse_next && !force && ws->last_n == n) ? 1 : 0;
  ws->reuse_next = 0;
  ws->last_n = n;
  // keys of this call into ws->prev (and the "changed" flag) over the workspace's padded
  // length, so a later call with fewer nodes still sees its stale tail as changed
  k_spread_keys<<<blocks_for(ws->padded, 256), 256, 0, s>>>(idx, n, ws->padded, nz, ny, nx,
                                                            ws->prev, ws->flag, force, reuse);
  SOPHT_CHECK_LAUNCH();
  if (reuse) {
    // sort, owners and merge would find the keys unchanged and return: not launched
  } else if (padded <= 8192) {
    k_bitonic_lds<<<1, 1024, bitonic_lds_bytes(padded), s>>>(ws->prev, ws->keys, int(padded),
                                                             ws->flag, ny, nx, ws->table, ws->bbox,
                                                             table_cap_for(padded));
    SOPHT_CHECK_LAUNCH();
  } else {
    k_copy_keys<<<blocks_for(padded, 256), 256, 0, s>>>(ws->prev, ws->keys, padded, ws->flag, ws->bbox);
    SOPHT_CHECK_LAUNCH();
    for (int64_t k = 2; k <= padded; k <<= 1)
      for (int64_t j = k >> 1; j > 0; j >>= 1) {
        k_bitonic_step<<<blocks_for(padded, 256), 256, 0, s>>>(ws->keys, padded, j, k, ws->flag);
        SOPHT_CHECK_LAUNCH();
      }
  }
  // plan rebuild (a no-op unless the keys changed): owners, then the merged contributor
  // lists; sorted keys in LDS for the owners' binary searches when they fit (128 KiB)
  if (reuse) {
  } else if (n <= 16384)
    k_spread_owner<true><<<blocks_for(n * 64, 512), 512, n * sizeof(uint64_t), s>>>(
        nz, ny, nx, ws->keys, n, ws->flag, ws->pcount, ws->poff, ws->owners, ws->table, ws->bbox);
  else
    k_spread_owner<false><<<blocks_for(n * 64, 512), 512, 0, s>>>(nz, ny, nx, ws->keys, n, ws->flag,
                                                                  ws->pcount, ws->poff, ws->owners,
                                                                  ws->table, ws->bbox);
  SOPHT_CHECK_LAUNCH();
  // about one wavefront per node (an owned cell has ~16 contributors, so a node owns ~4 cells
  // on average): the owners' latency chains then overlap instead of queueing in a few waves
  if (!reuse) {
    const int mblocks = std::min(8192, blocks_for(n * 64, 64 * kMergeWaves / 4));
    k_spread_merge<<<mblocks, 64 * kMergeWaves, 0, s>>>(nz, ny, nx, ws->keys, n, ws->flag, ws->pcount,
                                                       ws->poff, ws->owners, ws->plist, ws->table,
                                                       ws->bbox);
    SOPHT_CHECK_LAUNCH();
  }
  IB_DISPATCH(dtype, ([&] {
                EulW3<T> F;
                for (int c = 0; c < 3; ++c) F.p[c] = c < ncomp ? static_cast<T*>(eul[c]) : nullptr;
                F.sz = es[0]; F.sy = es[1]; F.sx = es[2];
                const T* L = static_cast<const T*>(lag);
                const T* W = static_cast<const T*>(weights);
                // about four wavefronts per node for the wavefront form (a node owns ~4 cells)
                k_spread_replay<T><<<std::min(8192, blocks_for(n * 4, 4)), 256, 0, s>>>(
                    ncomp, F, nz, ny, nx, L, W, ws->keys, n, ws->pcount, ws->poff, ws->plist,
                    ws->owners, ws->flag);
              }()));
  SOPHT_CHECK_LAUNCH();
  return SOPHT_OK;
}

// ---------------------------------------------------------------- spreading without allocation
// The library-owned scratch of sopht_lag_to_eul_spread_3d_noalloc: a workspace for
// kNoAllocNodes nodes carved out of one static device array. Each GPU that runs the entry
// point holds its own copy for the life of the process; its address on a device comes from
// hipGetSymbolAddress (the array's name is not a device address in host code).
constexpr int64_t kNoAllocNodes = 8192;
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }
constexpr size_t kKeysBytes = align_up(sizeof(uint64_t) * kNoAllocNodes);
constexpr size_t kFlagBytes = align_up(8 * sizeof(int));
constexpr size_t kSlotBytes = align_up(sizeof(int) * 64 * kNoAllocNodes);
constexpr size_t kTableBytes = align_up(sizeof(int) * (kTableMax + 1));
constexpr size_t kBboxBytes = align_up(sizeof(int) * 8);
constexpr size_t kScratchBytes = 2 * kKeysBytes + kFlagBytes + 4 * kSlotBytes + kTableBytes + kBboxBytes;

}  // namespace

// 8.5 MB per GPU (the workspace of 8192 nodes), held until the process ends
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// The scratch's device address on the current device, looked up once per device.
int scratch_address(void** out) {
  int dev = 0;
  SOPHT_CHECK_HIP(hipGetDevice(&dev));
  SOPHT_REQUIRE(dev >= 0 && dev < kMaxDevices, "device ordinal beyond the library's table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    SOPHT_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SOPHT_OK;
}

// Flags start as a fresh workspace's ({no out-of-domain node, keys changed}, empty box table);
// a kernel, not a memset, so the call can sit in a captured graph.
__global__ void k_noalloc_reset(int* flag, int* bbox) {
  const int i = threadIdx.x;
  if (i < 8) {
    flag[i] = i == 1 ? 1 : 0;
    bbox[i] = 0;
  }
}

__global__ void k_noalloc_report(const int* flag, int* out_of_domain) {
  if (threadIdx.x == 0) *out_of_domain = flag[0] != 0 ? 1 : 0;
}

}  // namespace

extern "C" {

int sopht_lag_to_eul_spread_3d(int dtype, int ncomp, sopht_spread3d* ws, void* const eul[3],
                               const int64_t* es, int64_t nz, int64_t ny, int64_t nx,
                               const void* lag, const void* weights, const int64_t* idx,
                               int64_t n, void* stream) {
  SOPHT_REQUIRE(ws != nullptr, "spread workspace is null");
  SOPHT_REQUIRE(ncomp == 1 || ncomp == 3, "Invalid number of components for interpolation, must be either 1 or 3");
  SOPHT_REQUIRE(n <= ws->capacity, "more Lagrangian nodes than the spread workspace holds");
  SOPHT_REQUIRE(nz * ny * nx < (int64_t(1) << 32), "grid too large for 32-bit cell keys");
  if (n == 0) return SOPHT_OK;
  // a plan built for another grid is rebuilt even if the packed keys happen to agree
  const int force = (ws->plan_nz != nz || ws->plan_ny != ny || ws->plan_nx != nx) ? 1 : 0;
  ws->plan_nz = nz;
  ws->plan_ny = ny;
  ws->plan_nx = nx;
  return spread_impl(dtype, ncomp, ws, eul, es, nz, ny, nx, lag, weights, idx, n,
                     static_cast<hipStream_t>(stream), force);
}

int sopht_lag_to_eul_spread_3d_noalloc(int dtype, int ncomp, void* const eul[3], const int64_t* es,
                                       int64_t nz, int64_t ny, int64_t nx, const void* lag,
                                       const void* weights, const int64_t* idx, int64_t n,
                                       int* out_of_domain, void* stream) {
  SOPHT_REQUIRE(ncomp == 1 || ncomp == 3, "Invalid number of components for interpolation, must be either 1 or 3");
  SOPHT_REQUIRE(n >= 0, "negative number of Lagrangian nodes");
  SOPHT_REQUIRE(n <= kNoAllocNodes,
                "more Lagrangian nodes than the allocation-free spreading holds (at most 8192); "
                "use sopht_lag_to_eul_spread_3d with a workspace");
  SOPHT_REQUIRE(nz * ny * nx < (int64_t(1) << 32), "grid too large for 32-bit cell keys");
  if (n == 0) return SOPHT_OK;
  void* base = nullptr;
  int rc = scratch_address(&base);
  if (rc) return rc;
  unsigned char* p = static_cast<unsigned char*>(base);
  sopht_spread3d ws{};
  ws.capacity = kNoAllocNodes;
  ws.padded = kNoAllocNodes;  // a power of two
  ws.keys = reinterpret_cast<uint64_t*>(p);
  p += kKeysBytes;
  ws.prev = reinterpret_cast<uint64_t*>(p);
  p += kKeysBytes;
  ws.flag = reinterpret_cast<int*>(p);
  p += kFlagBytes;
  ws.pcount = reinterpret_cast<int*>(p);
  p += kSlotBytes;
  ws.poff = reinterpret_cast<int*>(p);
  p += kSlotBytes;
  ws.plist = reinterpret_cast<uint32_t*>(p);
  p += kSlotBytes;
  ws.owners = reinterpret_cast<uint32_t*>(p);
  p += kSlotBytes;
  ws.table = reinterpret_cast<int*>(p);
  p += kTableBytes;
  ws.bbox = reinterpret_cast<int*>(p);
  hipStream_t s = static_cast<hipStream_t>(stream);
  // the scratch holds whatever the last call on this GPU left: reset the flags and rebuild
  // the plan from this call's keys (every key slot up to the capacity is rewritten)
  k_noalloc_reset<<<1, 64, 0, s>>>(ws.flag, ws.bbox);
  SOPHT_CHECK_LAUNCH();
  rc = spread_impl(dtype, ncomp, &ws, eul, es, nz, ny, nx, lag, weights, idx, n, s, 1);
  if (rc) return rc;
  if (out_of_domain) {
    k_noalloc_report<<<1, 64, 0, s>>>(ws.flag, out_of_domain);
    SOPHT_CHECK_LAUNCH();
  }
  return SOPHT_OK;
}

int sopht_spread3d_check(sopht_spread3d* ws, void* stream) {
  SOPHT_REQUIRE(ws != nullptr, "spread workspace is null");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int h[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  SOPHT_CHECK_HIP(hipMemcpyAsync(h, ws->flag, sizeof(h), hipMemcpyDeviceToHost, s));
  SOPHT_CHECK_HIP(hipStreamSynchronize(s));
  if (h[5]) {  // a call declared its plan reusable although its keys changed: nothing spread
    const int zero = 0;
    SOPHT_CHECK_HIP(hipMemcpyAsync(ws->flag + 5, &zero, sizeof(int), hipMemcpyHostToDevice, s));
    SOPHT_CHECK_HIP(hipStreamSynchronize(s));
    set_last_error("sopht_spread3d_reuse_plan was declared for a call whose support indices changed; "
                   "that call spread nothing", __FILE__, __LINE__);
    return SOPHT_ERR_INVALID;
  }
  if (h[0]) {
    SOPHT_CHECK_HIP(hipMemsetAsync(ws->flag, 0, sizeof(int), s));
    set_last_error("a Lagrangian node's 4^3 support leaves the Eulerian grid", __FILE__, __LINE__);
    return SOPHT_ERR_OUT_OF_DOMAIN;
  }
  return SOPHT_OK;
}

int sopht_spread3d_reuse_plan(sopht_spread3d* ws) {
  SOPHT_REQUIRE(ws != nullptr, "spread workspace is null");
  ws->reuse_next = 1;
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
                VbfArgs<T> A{kernel_type, u, nz, ny, nx, X, V, n, dx, shift, k, c, idx,
                             static_cast<T*>(support), static_cast<T*>(weights),
                             static_cast<T*>(lag_u), static_cast<T*>(dV),
                             static_cast<const T*>(dX), static_cast<T*>(F)};
                k_vbf_lag_force<T><<<blocks_for(n, 4), 256, 0, s>>>(A);
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
