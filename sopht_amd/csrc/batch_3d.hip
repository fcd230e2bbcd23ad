// Batched coupled IB flow step: many independent simulators of any sizes advanced together,
// one launch per stage for all of them (the Poisson passes once per distinct grid size). It is
// the reference driver's loop body (flow_past_sphere_case.py:191-197) per member:
//
//   dt = compute_stable_timestep()       CFL reduction and dt on the device (k_cfl_*)
//   interactor.time_step(dt)             dX += dt dV                     (k_vbf_time_step_batch)
//   interactor()                         VB force and deterministic spreading into F
//   flow.time_step(dt, U)                S1, SC, Poisson, S2
//
// Everything a step needs lives on the device (dt included), so a step can be captured into
// a hipGraph once and replayed. Work items of all members are laid end to end; a workgroup
// finds its member by binary search in a table of item offsets. Each member runs exactly the
// arithmetic of the single-simulator path, so the results agree bit for bit with stepping
// the members one by one.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <new>
#include <tuple>
#include <vector>

#include "common.h"
#include "ib_kernels.h"
#include "sweeps_3d.h"

namespace sopht {
namespace {

// member descriptor on the device
template <typename T>
struct Member {
  Grid g;
  T* w;
  T* u;
  T* psi;
  T* F;
  double dx, nu, rho, cfl;
  double U[3];
  int with_forcing;
  Pen pen;
  int s1_tx, s1_ty, m_tx, m_ty;
  double dt;    // this step's dt (written by k_cfl_dt_batch)
  double time;  // simulator time, accumulated in real_t as the reference does
  // virtual-boundary body (n = 0: none)
  VbfArgs<T> vb;
  EulW3<T> Fw;
  uint64_t* keys;
  uint64_t* prev;
  int* flag;
  int* pcount;
  int* poff;
  uint32_t* plist;
  int padded;
};

// workgroup -> member: the largest m with off[m] <= item (off has count + 1 entries)
__device__ __forceinline__ int find_member(const int64_t* off, int count, int64_t item) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= item) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// ---- CFL: per-member maximum of |u0| + |u1| + |u2|, then dt as
// passive_transport_flow_simulators.py:139-155 forms it in real_t
constexpr int kCflBlock = 256, kCflPerThread = 16;

template <typename T>
__global__ void __launch_bounds__(kCflBlock) k_cfl_partial_batch(const Member<T>* M, const int64_t* off,
                                                                 int count, T* partial) {
  __shared__ T red[kCflBlock];
  const int m = find_member(off, count, blockIdx.x);
  const Member<T>& S = M[m];
  const int64_t n = S.g.nz * S.g.ny * S.g.nx;
  const int64_t b = blockIdx.x - off[m];
  const int64_t nb = off[m + 1] - off[m];
  T mx = T(0);
  for (int64_t i = b * kCflBlock + threadIdx.x; i < n; i += nb * kCflBlock) {
    const T v = (fabs(S.u[i]) + fabs(S.u[n + i])) + fabs(S.u[2 * n + i]);
    mx = v > mx ? v : mx;
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int k = kCflBlock / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] = red[threadIdx.x + k] > red[threadIdx.x] ? red[threadIdx.x + k] : red[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

template <typename T>
__global__ void __launch_bounds__(kCflBlock) k_cfl_dt_batch(Member<T>* M, const int64_t* off,
                                                            const T* partial) {
#pragma clang fp contract(off)
  __shared__ T red[kCflBlock];
  const int m = blockIdx.x;
  T mx = T(0);
  for (int64_t i = off[m] + threadIdx.x; i < off[m + 1]; i += kCflBlock) mx = partial[i] > mx ? partial[i] : mx;
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int k = kCflBlock / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] = red[threadIdx.x + k] > red[threadIdx.x] ? red[threadIdx.x + k] : red[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    Member<T>& S = M[m];
    const T vmax = red[0];
    const T dx = T(S.dx);
    const T tol = T(10) * T(sizeof(T) == 4 ? 1.1920928955078125e-07 : 2.220446049250313e-16);
    const T a = (T(S.cfl) * dx) / (vmax + tol);
    const T b = ((T(0.9) * (dx * dx)) / T(6)) / T(S.nu) + tol;
    const T dt = b < a ? b : a;
    S.dt = double(dt);
    S.time = double(T(T(S.time) + dt));
  }
}

// ---- virtual boundary
template <typename T>
__global__ void __launch_bounds__(256) k_vbf_time_step_batch(const Member<T>* M) {
  const Member<T>& S = M[blockIdx.x];
  const int64_t m3 = 3 * S.vb.n;
  const T dt = T(S.dt);
  T* dX = const_cast<T*>(S.vb.dX);
  for (int64_t i = threadIdx.x; i < m3; i += 256) dX[i] = dX[i] + dt * S.vb.dV[i];
}

// one wavefront per node; off = node offsets
template <typename T>
__global__ void __launch_bounds__(256) k_vbf_lag_force_batch(const Member<T>* M, const int64_t* off,
                                                             int count, int64_t total) {
  const int64_t node = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (node >= total) return;  // whole wavefronts leave together; no barriers follow
  const int m = find_member(off, count, node);
  vbf_lag_force_node<T>(M[m].vb, node - off[m], threadIdx.x & 63);
}

// one workgroup per member: keys, then the in-LDS bitonic sort (only when they changed)
template <typename T>
__global__ void __launch_bounds__(1024) k_spread_sort_batch(const Member<T>* M) {
  extern __shared__ uint64_t sk[];
  const Member<T>& S = M[blockIdx.x];
  if (S.vb.n == 0) return;
  for (int64_t i = threadIdx.x; i < S.padded; i += blockDim.x)
    spread_key(S.vb.idx, S.vb.n, S.g.nz, S.g.ny, S.g.nx, S.prev, S.flag, i, 0);
  __syncthreads();
  bitonic_block(S.prev, S.keys, S.padded, S.flag, sk);
}

// plan build and replay: 8 nodes (512 support slots) per workgroup, each member's nodes
// starting on a workgroup boundary (off = workgroup offsets)
template <typename T>
__global__ void __launch_bounds__(512) k_spread_plan_batch(const Member<T>* M, const int64_t* off,
                                                           int count) {
  extern __shared__ uint64_t sk[];
  const int m = find_member(off, count, blockIdx.x);
  const Member<T>& S = M[m];
  if (flag_load(S.flag + 1) == 0) return;  // plan still valid (uniform for the workgroup)
  const int64_t n = S.vb.n;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) sk[i] = S.keys[i];
  __syncthreads();
  const int64_t gid = (int64_t(blockIdx.x) - off[m]) * 512 + threadIdx.x;
  if ((gid >> 6) < n)
    spread_plan_slot(S.g.nz, S.g.ny, S.g.nx, sk, n, S.flag, S.pcount, S.poff, S.plist, gid);
}

template <typename T>
__global__ void __launch_bounds__(512) k_spread_replay_batch(const Member<T>* M, const int64_t* off,
                                                             int count) {
  const int m = find_member(off, count, blockIdx.x);
  const Member<T>& S = M[m];
  const int64_t gid = (int64_t(blockIdx.x) - off[m]) * 512 + threadIdx.x;
  if (gid == 0) atomicExch(S.flag + 1, 0);
  if (gid < 64 * S.vb.n)
    spread_replay_slot<T, false>(3, S.Fw, S.g.nz, S.g.ny, S.g.nx, S.vb.F, S.vb.weights, S.keys, S.vb.n,
                          S.pcount, S.poff, S.plist, gid);
}

// F = 0 on every member (once, before the first step; afterwards SC leaves F zeroed)
template <typename T>
__global__ void __launch_bounds__(256) k_zero_forcing_batch(const Member<T>* M, const int64_t* off,
                                                            int count) {
  const int m = find_member(off, count, blockIdx.x);
  const Member<T>& S = M[m];
  if (!S.F) return;
  const int64_t n3 = 3 * S.g.nz * S.g.ny * S.g.nx;
  const int64_t nb = off[m + 1] - off[m];
  for (int64_t i = (int64_t(blockIdx.x) - off[m]) * 256 + threadIdx.x; i < n3; i += nb * 256)
    S.F[i] = T(0);
}

// ---- flow sweeps; prefactors as the single-simulator step forms them from dt
template <typename T>
__global__ void __launch_bounds__(S1Tile<T>::NT) __attribute__((amdgpu_waves_per_eu(6)))
k_s1_batch(const Member<T>* M, const int64_t* off, int count, int remap, int* fflags) {
  const int64_t item = xcd_item(remap);
  const int m = find_member(off, count, item);
  const Member<T>& S = M[m];
  const T pF = T(S.dt / (2.0 * S.dx * S.rho));
  const T pB = T(S.dt / (2.0 * S.dx));
  // per S1 tile: did it see a bit of F set (member m's flags start at its first S1 item)
  s1_body<T>(S.g, S.w, S.u, S.F, S.psi, pF, pB, S.with_forcing, kS1Chunk, S.s1_tx, S.s1_ty,
             item - off[m], nullptr, fflags + off[m]);
}

template <typename T>
__global__ void __launch_bounds__(S1VTile<T>::NT) __attribute__((amdgpu_waves_per_eu(2)))
k_s1v_batch(const Member<T>* M, const int64_t* off, int count, int remap, int* fflags) {
  const int64_t item = xcd_item(remap);
  const int m = find_member(off, count, item);
  const Member<T>& S = M[m];
  const T pF = T(S.dt / (2.0 * S.dx * S.rho));
  const T pB = T(S.dt / (2.0 * S.dx));
  s1v_body<T>(S.g, S.w, S.u, S.F, S.psi, pF, pB, S.with_forcing, kS1Chunk, S.s1_tx, S.s1_ty,
              item - off[m], nullptr, fflags + off[m]);
}

template <typename T, int VEC>
__global__ void __launch_bounds__(kMarchThreads) k_sc_batch(const Member<T>* M, const int64_t* off,
                                                            int count, int remap, const int* fflags,
                                                            const int64_t* off_s1, int s1_rows) {
  const int64_t item = xcd_item(remap);
  const int m = find_member(off, count, item);
  const Member<T>& S = M[m];
  const T pD = T(S.nu * S.dt / S.dx / S.dx);
  // F is zeroed only in the S1 tiles that saw a bit of it set (elsewhere it is zero already)
  sc_body<T, VEC>(S.g, S.psi, S.w, S.F, pD, S.pen, S.with_forcing, S.m_tx, S.m_ty, item - off[m],
                  kMarchChunk, fflags + off_s1[m], S1Map{S.s1_tx, S.s1_ty, s1_rows, kS1Chunk});
}

template <typename T, int VEC>
__global__ void __launch_bounds__(kMarchThreads) k_s2_batch(const Member<T>* M, const int64_t* off,
                                                            int count, int remap, T* umax) {
  const int64_t item = xcd_item(remap);
  const int m = find_member(off, count, item);
  const Member<T>& S = M[m];
  // per work item maximum of |u0| + |u1| + |u2| at slot item: member m's slots are
  // off[m] .. off[m+1], which the next step's dt reduces
  s2_body<T, VEC>(S.g, S.psi, S.u, T(0.5 / S.dx), T(S.U[0]), T(S.U[1]), T(S.U[2]), S.m_tx, S.m_ty,
                  item - off[m], kMarchChunk, umax + off[m]);
}

inline unsigned nblocks(int64_t n) { return unsigned(n); }

}  // namespace
}  // namespace sopht

using namespace sopht;

// one Poisson plan per distinct grid of the batch, with device tables of its members' fields
struct BatchClass {
  sopht_poisson3d* plan = nullptr;
  int64_t shape[3] = {0, 0, 0};  // (nz, ny, nx)
  int ncomp = 0;
  void** sol_tab = nullptr;  // device: psi components
  void** rhs_tab = nullptr;  // device: omega components
  int64_t strides[3] = {0, 0, 0};
};

struct sopht_ns3d_batch {
  int dtype = SOPHT_F32;
  int count = 0;
  int vec = 1;
  void* members = nullptr;  // device Member<T>[count]
  size_t member_bytes = 0;
  int64_t* off = nullptr;   // device: 6 offset tables of count + 1 entries
  int64_t tot[6] = {0, 0, 0, 0, 0, 0};  // S1 items, march items, CFL blocks, nodes, plan groups, zero blocks
  void* partial = nullptr;  // CFL partial maxima
  void* s2max = nullptr;    // per S2 work item maxima of the last step's velocity
  int plans_built = 0;      // the spreading plans of the (static) bodies exist
  int* s1flags = nullptr;   // per S1 work item: did it see a bit of F set
  int s1v = 0;              // S1 with 16-byte accesses (k_s1v_batch)
  int s1_rows = 0;          // rows per S1 tile
  std::vector<BatchClass> classes;
  std::vector<sopht_spread3d*> ws;  // spread workspaces of the members with bodies
  int max_padded = 0;
  int forcing_zeroed = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // graphs are captured and replayed on a stream of the batch's own (the caller's stream may
  // be the null stream, which cannot be captured), ordered against the caller's with events
  hipStream_t cs = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
};

namespace {

enum { kOffS1 = 0, kOffMarch, kOffCfl, kOffNode, kOffPlan, kOffZero, kNumOff };

template <typename T>
int batch_build(sopht_ns3d_batch* B, const sopht_batch_member* mem, hipStream_t s) {
  const int count = B->count;
  std::vector<Member<T>> host(count);
  std::vector<int64_t> offs(kNumOff * (count + 1), 0);
  auto OFF = [&](int t, int m) -> int64_t& { return offs[t * (count + 1) + m]; };
  // one vector width for the march sweeps of every member
  int vec = 16 / int(sizeof(T));
  for (int m = 0; m < count; ++m) vec = std::min(vec, march_vec<T>(mem[m].nx, mem[m].width));
  B->vec = vec;
  // the S1 sweep with 16-byte accesses when every member's rows allow it (the same bits as
  // the scalar sweep; component byte offsets must stay below 2^32)
  B->s1v = vec > 1;
  for (int m = 0; m < count; ++m)
    if (mem[m].nz * mem[m].ny * mem[m].nx * int64_t(sizeof(T)) >= (int64_t(1) << 32)) B->s1v = 0;
  if (const char* scalar = std::getenv("SOPHT_S1_SCALAR"))
    if (std::atoi(scalar) == 1) B->s1v = 0;
  const int s1_rows = B->s1v ? S1VTile<T>::TY : S1Tile<T>::TY;
  B->s1_rows = s1_rows;
  std::map<std::tuple<int64_t, int64_t, int64_t, double>, std::vector<int>> by_grid;
  for (int m = 0; m < count; ++m) {
    const sopht_batch_member& q = mem[m];
    SOPHT_REQUIRE(q.nz >= 3 && q.ny >= 3 && q.nx >= 3, "member grid must be at least 3 cells per axis");
    SOPHT_REQUIRE(3 * q.nz * q.ny * q.nx < (int64_t(1) << 31), "member grid too large");
    SOPHT_REQUIRE(q.width >= 0 && q.width <= 8 && 2 * q.width <= std::min(q.nz, std::min(q.ny, q.nx)),
                  "invalid penalty width");
    SOPHT_REQUIRE(q.vorticity && q.velocity && q.psi && (!q.with_forcing || q.forcing), "null member field");
    SOPHT_REQUIRE(q.n_lag == 0 || (q.with_forcing && q.X && q.V && q.idx && q.support && q.weights &&
                                   q.lag_u && q.dV && q.dX && q.Fl),
                  "a member with a body needs forcing and every virtual-boundary buffer");
    SOPHT_REQUIRE(q.n_lag <= 8192, "at most 8192 Lagrangian nodes per member in a batch");
    Member<T>& S = host[m];
    S.g = Grid{q.nz, q.ny, q.nx};
    S.w = static_cast<T*>(q.vorticity);
    S.u = static_cast<T*>(q.velocity);
    S.psi = static_cast<T*>(q.psi);
    S.F = q.with_forcing ? static_cast<T*>(q.forcing) : nullptr;
    S.dx = q.x_range / double(q.nx);
    if (sizeof(T) == 4) S.dx = double(float(S.dx));
    S.nu = q.nu;
    S.rho = q.rho;
    S.cfl = q.cfl;
    for (int a = 0; a < 3; ++a) S.U[a] = q.with_free_stream ? q.free_stream[a] : 0.0;
    S.with_forcing = q.with_forcing ? 1 : 0;
    S.pen.w = q.width;
    for (int a = 0; a < 3; ++a)
      for (int i = 0; i < q.width; ++i) {
        const double pre = (M_PI / 2.0) / (double(q.width) * S.dx);
        S.pen.f[a][i] = std::sin(pre * (double(i) * S.dx));
        S.pen.f[a][q.width + i] = std::sin(pre * (double(q.width - 1 - i) * S.dx));
      }
    S.s1_tx = int(ceil_div(q.nx, 64));
    S.s1_ty = int(ceil_div(q.ny, s1_rows));
    const MarchLaunch ml = march_launch(q.nz, q.ny, q.nx, vec);
    S.m_tx = ml.tiles_x;
    S.m_ty = ml.tiles_y;
    S.dt = 0.0;
    S.time = double(T(q.time));
    const int64_t ncell = q.nz * q.ny * q.nx;
    OFF(kOffS1, m + 1) = OFF(kOffS1, m) + int64_t(S.s1_tx) * S.s1_ty * ceil_div(q.nz, kS1Chunk);
    OFF(kOffMarch, m + 1) = OFF(kOffMarch, m) + int64_t(ml.blocks);
    OFF(kOffCfl, m + 1) = OFF(kOffCfl, m) + std::max<int64_t>(1, ceil_div(ncell, kCflBlock * kCflPerThread));
    OFF(kOffNode, m + 1) = OFF(kOffNode, m) + q.n_lag;
    OFF(kOffPlan, m + 1) = OFF(kOffPlan, m) + ceil_div(q.n_lag, 8);
    OFF(kOffZero, m + 1) = OFF(kOffZero, m) + std::max<int64_t>(1, ceil_div(3 * ncell, 256 * 16));
    // virtual boundary
    S.vb = VbfArgs<T>{};
    S.vb.n = q.n_lag;
    S.padded = 0;
    if (q.n_lag > 0) {
      S.vb.kernel_type = q.kernel_type;
      for (int a = 0; a < 3; ++a) S.vb.u.p[a] = S.u + a * ncell;
      S.vb.u.sz = q.ny * q.nx;
      S.vb.u.sy = q.nx;
      S.vb.u.sx = 1;
      S.vb.nz = q.nz;
      S.vb.ny = q.ny;
      S.vb.nx = q.nx;
      S.vb.X = q.X;
      S.vb.V = q.V;
      S.vb.dx = q.dx;
      S.vb.shift = q.shift;
      S.vb.k = q.k;
      S.vb.c = q.c;
      S.vb.idx = q.idx;
      S.vb.support = static_cast<T*>(q.support);
      S.vb.weights = static_cast<T*>(q.weights);
      S.vb.lag_u = static_cast<T*>(q.lag_u);
      S.vb.dV = static_cast<T*>(q.dV);
      S.vb.dX = static_cast<T*>(q.dX);
      S.vb.F = static_cast<T*>(q.Fl);
      for (int a = 0; a < 3; ++a) S.Fw.p[a] = S.F + a * ncell;
      S.Fw.sz = q.ny * q.nx;
      S.Fw.sy = q.nx;
      S.Fw.sx = 1;
      sopht_spread3d* w = nullptr;
      int rc = sopht_spread3d_create(q.n_lag, &w);
      if (rc) return rc;
      B->ws.push_back(w);
      S.keys = w->keys;
      S.prev = w->prev;
      S.flag = w->flag;
      S.pcount = w->pcount;
      S.poff = w->poff;
      S.plist = w->plist;
      S.padded = int(w->padded);
      B->max_padded = std::max(B->max_padded, S.padded);
    }
    by_grid[std::make_tuple(q.nz, q.ny, q.nx, q.x_range)].push_back(m);
  }
  for (int t = 0; t < kNumOff; ++t) B->tot[t] = OFF(t, count);
  SOPHT_CHECK_HIP(hipMalloc(&B->members, sizeof(Member<T>) * count));
  B->member_bytes = sizeof(Member<T>) * count;
  SOPHT_CHECK_HIP(hipMalloc(&B->off, sizeof(int64_t) * offs.size()));
  SOPHT_CHECK_HIP(hipMalloc(&B->partial, sizeof(T) * std::max<int64_t>(1, B->tot[kOffCfl])));
  SOPHT_CHECK_HIP(hipMalloc(&B->s2max, sizeof(T) * std::max<int64_t>(1, B->tot[kOffMarch])));
  SOPHT_CHECK_HIP(hipMalloc(&B->s1flags, sizeof(int) * std::max<int64_t>(1, B->tot[kOffS1])));
  SOPHT_CHECK_HIP(hipMemcpyAsync(B->members, host.data(), B->member_bytes, hipMemcpyHostToDevice, s));
  SOPHT_CHECK_HIP(hipMemcpyAsync(B->off, offs.data(), sizeof(int64_t) * offs.size(), hipMemcpyHostToDevice, s));
  // Poisson classes
  for (auto& kv : by_grid) {
    const auto& ms = kv.second;
    const sopht_batch_member& q = mem[ms[0]];
    BatchClass C;
    C.ncomp = 3 * int(ms.size());
    C.shape[0] = q.nz;
    C.shape[1] = q.ny;
    C.shape[2] = q.nx;
    int rc = poisson_create(B->dtype, q.nz, q.ny, q.nx, q.x_range, C.ncomp, s, &C.plan);
    if (rc) return rc;
    std::vector<void*> sol(C.ncomp), rhs(C.ncomp);
    const int64_t n = q.nz * q.ny * q.nx;
    for (size_t i = 0; i < ms.size(); ++i)
      for (int c = 0; c < 3; ++c) {
        sol[3 * i + c] = static_cast<T*>(mem[ms[i]].psi) + c * n;
        rhs[3 * i + c] = static_cast<T*>(mem[ms[i]].vorticity) + c * n;
      }
    C.strides[0] = q.ny * q.nx;
    C.strides[1] = q.nx;
    C.strides[2] = 1;
    B->classes.push_back(C);
    BatchClass& D = B->classes.back();
    SOPHT_CHECK_HIP(hipMalloc(&D.sol_tab, sizeof(void*) * C.ncomp));
    SOPHT_CHECK_HIP(hipMalloc(&D.rhs_tab, sizeof(void*) * C.ncomp));
    SOPHT_CHECK_HIP(hipMemcpyAsync(D.sol_tab, sol.data(), sizeof(void*) * C.ncomp, hipMemcpyHostToDevice, s));
    SOPHT_CHECK_HIP(hipMemcpyAsync(D.rhs_tab, rhs.data(), sizeof(void*) * C.ncomp, hipMemcpyHostToDevice, s));
  }
  SOPHT_CHECK_HIP(hipStreamSynchronize(s));
  return SOPHT_OK;
}

// One batched step. fused_cfl: dt from the maxima that the previous step's S2 left per work
// item (the velocity is still the one S2 wrote); otherwise from a full read of every member's
// velocity (the first step of a call: the fields may have changed in between).
template <typename T>
int batch_step(sopht_ns3d_batch* B, hipStream_t s, bool fused_cfl) {
  const Member<T>* M = static_cast<const Member<T>*>(B->members);
  Member<T>* Mw = static_cast<Member<T>*>(B->members);
  const int count = B->count;
  const int64_t* off = B->off;
  auto O = [&](int t) { return off + t * (count + 1); };
  T* partial = static_cast<T*>(B->partial);
  T* s2max = static_cast<T*>(B->s2max);
  // dt
  if (fused_cfl) {
    k_cfl_dt_batch<T><<<count, kCflBlock, 0, s>>>(Mw, O(kOffMarch), s2max);
    SOPHT_CHECK_LAUNCH();
  } else {
    k_cfl_partial_batch<T><<<nblocks(B->tot[kOffCfl]), kCflBlock, 0, s>>>(M, O(kOffCfl), count, partial);
    SOPHT_CHECK_LAUNCH();
    k_cfl_dt_batch<T><<<count, kCflBlock, 0, s>>>(Mw, O(kOffCfl), partial);
    SOPHT_CHECK_LAUNCH();
  }
  // interactors
  if (B->tot[kOffNode] > 0) {
    k_vbf_time_step_batch<T><<<count, 256, 0, s>>>(M);
    SOPHT_CHECK_LAUNCH();
  }
  if (!B->forcing_zeroed) {
    k_zero_forcing_batch<T><<<nblocks(B->tot[kOffZero]), 256, 0, s>>>(M, O(kOffZero), count);
    SOPHT_CHECK_LAUNCH();
    B->forcing_zeroed = 1;
  }
  if (B->tot[kOffNode] > 0) {
    k_vbf_lag_force_batch<T><<<nblocks(ceil_div(B->tot[kOffNode], 4)), 256, 0, s>>>(
        M, O(kOffNode), count, B->tot[kOffNode]);
    SOPHT_CHECK_LAUNCH();
    // the bodies keep the positions they had when the batch was built, so the plans built by
    // the first step stay valid: later steps only replay them
    if (!B->plans_built) {
      k_spread_sort_batch<T><<<count, 1024, size_t(std::max(B->max_padded, 1)) * sizeof(uint64_t), s>>>(M);
      SOPHT_CHECK_LAUNCH();
      k_spread_plan_batch<T><<<nblocks(B->tot[kOffPlan]), 512, size_t(std::max(B->max_padded, 1)) * sizeof(uint64_t), s>>>(
          M, O(kOffPlan), count);
      SOPHT_CHECK_LAUNCH();
      B->plans_built = 1;
    }
    k_spread_replay_batch<T><<<nblocks(B->tot[kOffPlan]), 512, 0, s>>>(M, O(kOffPlan), count);
    SOPHT_CHECK_LAUNCH();
  }
  // flow
  const int r1 = B->tot[kOffS1] % 8 == 0 ? 1 : 0;
  if (B->s1v)
    k_s1v_batch<T><<<nblocks(B->tot[kOffS1]), S1VTile<T>::NT, 0, s>>>(M, O(kOffS1), count, r1, B->s1flags);
  else
    k_s1_batch<T><<<nblocks(B->tot[kOffS1]), S1Tile<T>::NT, 0, s>>>(M, O(kOffS1), count, r1, B->s1flags);
  SOPHT_CHECK_LAUNCH();
  const int rm = B->tot[kOffMarch] % 8 == 0 ? 1 : 0;
  if (B->vec == 1)
    k_sc_batch<T, 1><<<nblocks(B->tot[kOffMarch]), kMarchThreads, 0, s>>>(M, O(kOffMarch), count, rm,
                                                                         B->s1flags, O(kOffS1), B->s1_rows);
  else
    k_sc_batch<T, 16 / sizeof(T)><<<nblocks(B->tot[kOffMarch]), kMarchThreads, 0, s>>>(
        M, O(kOffMarch), count, rm, B->s1flags, O(kOffS1), B->s1_rows);
  SOPHT_CHECK_LAUNCH();
  for (BatchClass& C : B->classes) {
    int rc = poisson_table_solve(C.plan, C.ncomp, C.sol_tab, C.strides, C.rhs_tab, C.strides, s);
    if (rc) return rc;
  }
  if (B->vec == 1)
    k_s2_batch<T, 1><<<nblocks(B->tot[kOffMarch]), kMarchThreads, 0, s>>>(M, O(kOffMarch), count, rm, s2max);
  else
    k_s2_batch<T, 16 / sizeof(T)><<<nblocks(B->tot[kOffMarch]), kMarchThreads, 0, s>>>(M, O(kOffMarch), count, rm, s2max);
  SOPHT_CHECK_LAUNCH();
  return SOPHT_OK;
}

int batch_step_any(sopht_ns3d_batch* B, hipStream_t s, bool fused_cfl) {
  return B->dtype == SOPHT_F32 ? batch_step<float>(B, s, fused_cfl) : batch_step<double>(B, s, fused_cfl);
}

}  // namespace

extern "C" {

int sopht_ns3d_batch_create(int dtype, int count, const sopht_batch_member* members, void* stream,
                            sopht_ns3d_batch** out) {
  SOPHT_REQUIRE(out != nullptr, "batch pointer is null");
  *out = nullptr;
  SOPHT_REQUIRE(dtype == SOPHT_F32 || dtype == SOPHT_F64, "dtype must be SOPHT_F32 or SOPHT_F64");
  SOPHT_REQUIRE(count >= 1 && members != nullptr, "a batch needs at least one member");
  sopht_ns3d_batch* B = new (std::nothrow) sopht_ns3d_batch();
  SOPHT_REQUIRE(B != nullptr, "out of host memory");
  B->dtype = dtype;
  B->count = count;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int rc = dtype == SOPHT_F32 ? batch_build<float>(B, members, s) : batch_build<double>(B, members, s);
  if (rc) {
    sopht_ns3d_batch_destroy(B);
    return rc;
  }
  *out = B;
  return SOPHT_OK;
}

int sopht_ns3d_batch_step(sopht_ns3d_batch* B, int steps, int use_graph, void* stream) {
  SOPHT_REQUIRE(B != nullptr, "batch is null");
  SOPHT_REQUIRE(steps >= 0, "steps must be non-negative");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int k = 0;
  // the first step of a call runs eagerly with dt from a full read of the velocities (the
  // caller may have changed them since the last call; the very first step also zeroes F);
  // the following steps take dt from the maxima their predecessor's S2 kept
  if (steps > 0) {
    int rc = batch_step_any(B, s, false);
    if (rc) return rc;
    ++k;
  }
  if (!use_graph) {
    for (; k < steps; ++k) {
      int rc = batch_step_any(B, s, true);
      if (rc) return rc;
    }
    return SOPHT_OK;
  }
  if (k >= steps) return SOPHT_OK;
  if (!B->cs) {
    SOPHT_CHECK_HIP(hipStreamCreateWithFlags(&B->cs, hipStreamNonBlocking));
    SOPHT_CHECK_HIP(hipEventCreateWithFlags(&B->ev_in, hipEventDisableTiming));
    SOPHT_CHECK_HIP(hipEventCreateWithFlags(&B->ev_out, hipEventDisableTiming));
  }
  SOPHT_CHECK_HIP(hipEventRecord(B->ev_in, s));
  SOPHT_CHECK_HIP(hipStreamWaitEvent(B->cs, B->ev_in, 0));
  if (!B->exec) {
    SOPHT_CHECK_HIP(hipStreamBeginCapture(B->cs, hipStreamCaptureModeThreadLocal));
    int rc = batch_step_any(B, B->cs, true);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(B->cs, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    SOPHT_CHECK_HIP(e);
    B->graph = g;
    SOPHT_CHECK_HIP(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
  }
  for (; k < steps; ++k) SOPHT_CHECK_HIP(hipGraphLaunch(B->exec, B->cs));
  SOPHT_CHECK_HIP(hipEventRecord(B->ev_out, B->cs));
  SOPHT_CHECK_HIP(hipStreamWaitEvent(s, B->ev_out, 0));
  return SOPHT_OK;
}

int sopht_ns3d_batch_state(sopht_ns3d_batch* B, double* dt, double* time, void* stream) {
  SOPHT_REQUIRE(B != nullptr, "batch is null");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t bytes = B->dtype == SOPHT_F32 ? sizeof(Member<float>) : sizeof(Member<double>);
  std::vector<unsigned char> host(bytes * B->count);
  SOPHT_CHECK_HIP(hipMemcpyAsync(host.data(), B->members, host.size(), hipMemcpyDeviceToHost, s));
  SOPHT_CHECK_HIP(hipStreamSynchronize(s));
  for (int m = 0; m < B->count; ++m) {
    const unsigned char* p = host.data() + bytes * m;
    if (B->dtype == SOPHT_F32) {
      const Member<float>* S = reinterpret_cast<const Member<float>*>(p);
      if (dt) dt[m] = S->dt;
      if (time) time[m] = S->time;
    } else {
      const Member<double>* S = reinterpret_cast<const Member<double>*>(p);
      if (dt) dt[m] = S->dt;
      if (time) time[m] = S->time;
    }
  }
  return SOPHT_OK;
}

int sopht_ns3d_batch_class_schedule(const sopht_ns3d_batch* B, int i, int64_t* shape, int* ncomp,
                                    int* passes, int* zpass_jobs) {
  SOPHT_REQUIRE(B != nullptr, "batch is null");
  const int n = int(B->classes.size());
  if (i < 0 || i >= n) return n;
  const BatchClass& cl = B->classes[size_t(i)];
  if (shape)
    for (int a = 0; a < 3; ++a) shape[a] = cl.shape[a];
  if (ncomp) *ncomp = cl.ncomp;
  const int rc = sopht_poisson3d_schedule(cl.plan, passes, zpass_jobs);
  return rc ? rc : n;
}

int sopht_ns3d_batch_destroy(sopht_ns3d_batch* B) {
  if (!B) return SOPHT_OK;
  if (B->cs) (void)hipStreamSynchronize(B->cs);
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  if (B->ev_in) (void)hipEventDestroy(B->ev_in);
  if (B->ev_out) (void)hipEventDestroy(B->ev_out);
  if (B->cs) (void)hipStreamDestroy(B->cs);
  for (BatchClass& C : B->classes) {
    sopht_poisson3d_destroy(C.plan);
    if (C.sol_tab) (void)hipFree(C.sol_tab);
    if (C.rhs_tab) (void)hipFree(C.rhs_tab);
  }
  for (sopht_spread3d* w : B->ws) sopht_spread3d_destroy(w);
  if (B->members) (void)hipFree(B->members);
  if (B->off) (void)hipFree(B->off);
  if (B->partial) (void)hipFree(B->partial);
  if (B->s2max) (void)hipFree(B->s2max);
  if (B->s1flags) (void)hipFree(B->s1flags);
  delete B;
  return SOPHT_OK;
}

}  // extern "C"
