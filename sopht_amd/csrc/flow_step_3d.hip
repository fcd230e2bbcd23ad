// The 3D unbounded Navier-Stokes step with forcing (navier_stokes_flow_simulators.py:449-498)
// as fused sweeps on gfx950. Fields are contiguous (3, nz, ny, nx), x fastest.
//
//   SA  : omega1 = omega + pF curl_h(F)  (interior; border omega unchanged), written back to
//         omega in place (pointwise in omega), and B = u x omega1 into the buffer (all cells).
//   SB  : omega2 = omega1 + pB curl_h(B) (interior) into the stream-function field, which the
//         Poisson solve overwrites later in the step.
//   SC  : omega3 = omega2 + nu dt/dx^2 lap_h(omega2) (interior; flux border 0), penalised
//         (value at the clamped index times the per-axis sine factors) into omega; F = 0.
//   Poisson (poisson_3d.hip) omega -> psi.
//   S2  : u = (0.5/dx) curl_h(psi) (interior), 0 on the border, + U_inf everywhere.
// Each sweep reads its neighbours through L2 with x-fastest lanes (coalesced rows).
#include <algorithm>
#include <cmath>
#include <new>

#include "common.h"

struct sopht_ns3d {
  int dtype;
  int64_t nz, ny, nx;
  double x_range, dx, nu, rho;
  int width, with_forcing, with_free_stream;
  sopht_poisson3d* poisson;
  double sine[6 * 8];  // x, y, z: width front factors then width back factors
  void* partial;       // per-block partial maxima for the CFL reduction
  int64_t partial_cap;
  sopht::StageTimer timer;
};

// Stage order of one step, for sopht_ns3d_stage_times.
static const char* const kStageNames[] = {
    "vorticity_forcing_cross",  // SA
    "vorticity_advect",         // SB
    "vorticity_diffuse_penalise",  // SC
    "poisson_x_r2c",            // K1
    "poisson_y_fwd",            // K2
    "poisson_z_conv",           // K3
    "poisson_y_inv",            // K4
    "poisson_x_c2r",            // K5
    "velocity_curl",            // S2
};
static constexpr int kNumStages = 9;

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

constexpr int BX = 64, BY = 4;

inline dim3 cell_grid(int64_t nz, int64_t ny, int64_t nx) {
  return dim3(unsigned((nx + BX - 1) / BX), unsigned((ny + BY - 1) / BY), unsigned(nz));
}

// curl_h of a vector field f (component planes f0, f1, f2) at an interior cell, unscaled.
template <typename T>
__device__ __forceinline__ void curl_at(const T* f0, const T* f1, const T* f2, const Grid& g,
                                        int64_t z, int64_t y, int64_t x, T c[3]) {
  const int64_t sy = g.nx, sz = g.nx * g.ny;
  const int64_t i = g.at(z, y, x);
  c[0] = f2[i + sy] - f2[i - sy] - f1[i + sz] + f1[i - sz];
  c[1] = f0[i + sz] - f0[i - sz] - f2[i + 1] + f2[i - 1];
  c[2] = f1[i + 1] - f1[i - 1] - f0[i + sy] + f0[i - sy];
}

template <typename T>
__global__ void __launch_bounds__(BX * BY) k_sa(Grid g, T* w, const T* u, const T* F, T* B, T pF,
                                                int with_forcing) {
  const int64_t x = blockIdx.x * int64_t(BX) + threadIdx.x;
  const int64_t y = blockIdx.y * int64_t(BY) + threadIdx.y;
  const int64_t z = blockIdx.z;
  if (x >= g.nx || y >= g.ny) return;
  const int64_t n = g.nz * g.ny * g.nx;
  const int64_t i = g.at(z, y, x);
  T w0 = w[i], w1 = w[n + i], w2 = w[2 * n + i];
  if (with_forcing && g.interior(z, y, x)) {
    T c[3];
    curl_at<T>(F, F + n, F + 2 * n, g, z, y, x, c);
    w0 = w0 + pF * c[0];
    w1 = w1 + pF * c[1];
    w2 = w2 + pF * c[2];
    w[i] = w0;
    w[n + i] = w1;
    w[2 * n + i] = w2;
  }
  const T u0 = u[i], u1 = u[n + i], u2 = u[2 * n + i];
  // elementwise_ops_3d.py:427-447: B_i = u_j w_k - w_j u_k
  B[i] = u1 * w2 - w1 * u2;
  B[n + i] = u2 * w0 - w2 * u0;
  B[2 * n + i] = u0 * w1 - w0 * u1;
}

template <typename T>
__global__ void __launch_bounds__(BX * BY) k_sb(Grid g, const T* w, const T* B, T* w2out, T pB) {
  const int64_t x = blockIdx.x * int64_t(BX) + threadIdx.x;
  const int64_t y = blockIdx.y * int64_t(BY) + threadIdx.y;
  const int64_t z = blockIdx.z;
  if (x >= g.nx || y >= g.ny) return;
  const int64_t n = g.nz * g.ny * g.nx;
  const int64_t i = g.at(z, y, x);
  T a0 = w[i], a1 = w[n + i], a2 = w[2 * n + i];
  if (g.interior(z, y, x)) {
    T c[3];
    curl_at<T>(B, B + n, B + 2 * n, g, z, y, x, c);
    a0 = a0 + pB * c[0];
    a1 = a1 + pB * c[1];
    a2 = a2 + pB * c[2];
  }
  w2out[i] = a0;
  w2out[n + i] = a1;
  w2out[2 * n + i] = a2;
}

struct Pen {
  int w;
  // per-axis factors: index 0..w-1 front, then w back entries for n-w..n-1
  double f[3][16];
};

template <typename T>
__device__ __forceinline__ T diffused(const T* a, const Grid& g, int64_t z, int64_t y, int64_t x,
                                      T p) {
  const int64_t i = g.at(z, y, x);
  if (!g.interior(z, y, x)) return a[i];
  const int64_t sy = g.nx, sz = g.nx * g.ny;
  const T flux = p * (a[i + sz] + a[i - sz] + a[i + sy] + a[i - sy] + a[i + 1] + a[i - 1] -
                      T(6) * a[i]);
  return a[i] + flux;
}

__device__ __forceinline__ int64_t clampi(int64_t v, int64_t lo, int64_t hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

template <typename T>
__global__ void __launch_bounds__(BX * BY) k_sc(Grid g, const T* w2, T* w, T* F, T pD, Pen pen,
                                                int zero_forcing) {
  const int64_t x = blockIdx.x * int64_t(BX) + threadIdx.x;
  const int64_t y = blockIdx.y * int64_t(BY) + threadIdx.y;
  const int64_t z = blockIdx.z;
  if (x >= g.nx || y >= g.ny) return;
  const int64_t n = g.nz * g.ny * g.nx;
  const int64_t i = g.at(z, y, x);
  const int W = pen.w;
  int64_t cz = z, cy = y, cx = x;
  T sx = T(1), sy = T(1), sz = T(1);
  if (W > 0) {
    cx = clampi(x, W - 1, g.nx - W);
    cy = clampi(y, W - 1, g.ny - W);
    cz = clampi(z, W - 1, g.nz - W);
    if (x < W) sx = T(pen.f[0][x]);
    else if (x >= g.nx - W) sx = T(pen.f[0][W + x - (g.nx - W)]);
    if (y < W) sy = T(pen.f[1][y]);
    else if (y >= g.ny - W) sy = T(pen.f[1][W + y - (g.ny - W)]);
    if (z < W) sz = T(pen.f[2][z]);
    else if (z >= g.nz - W) sz = T(pen.f[2][W + z - (g.nz - W)]);
  }
  const bool pen_cell = (sx != T(1)) || (sy != T(1)) || (sz != T(1)) || cx != x || cy != y || cz != z;
  for (int c = 0; c < 3; ++c) {
    const T* a = w2 + c * n;
    T v = diffused<T>(a, g, cz, cy, cx, pD);
    if (pen_cell) v = ((v * sx) * sy) * sz;  // penalise_field_boundary_3d.py:384-403, x then y then z
    w[c * n + i] = v;
  }
  if (zero_forcing) {
    F[i] = T(0);
    F[n + i] = T(0);
    F[2 * n + i] = T(0);
  }
}

template <typename T>
__global__ void __launch_bounds__(BX * BY) k_s2(Grid g, const T* psi, T* u, T p, T U0, T U1, T U2) {
  const int64_t x = blockIdx.x * int64_t(BX) + threadIdx.x;
  const int64_t y = blockIdx.y * int64_t(BY) + threadIdx.y;
  const int64_t z = blockIdx.z;
  if (x >= g.nx || y >= g.ny) return;
  const int64_t n = g.nz * g.ny * g.nx;
  const int64_t i = g.at(z, y, x);
  T c[3] = {T(0), T(0), T(0)};
  if (g.interior(z, y, x)) {
    curl_at<T>(psi, psi + n, psi + 2 * n, g, z, y, x, c);
    c[0] = p * c[0];
    c[1] = p * c[1];
    c[2] = p * c[2];
  }
  u[i] = c[0] + U0;
  u[n + i] = c[1] + U1;
  u[2 * n + i] = c[2] + U2;
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

Pen make_pen(const sopht_ns3d* P) {
  Pen pen;
  pen.w = P->width;
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < 2 * P->width; ++i) pen.f[a][i] = P->sine[a * 2 * P->width + i];
  return pen;
}

template <typename T>
int step_impl(sopht_ns3d* P, void* vort, void* vel, void* psi, void* forcing, void* buffer,
              double dt, const double* U, hipStream_t s) {
  Grid g{P->nz, P->ny, P->nx};
  const dim3 grid = cell_grid(g.nz, g.ny, g.nx);
  const dim3 block(BX, BY);
  T* w = static_cast<T*>(vort);
  T* u = static_cast<T*>(vel);
  T* ps = static_cast<T*>(psi);
  T* F = static_cast<T*>(forcing);
  T* B = static_cast<T*>(buffer);
  const T dx = T(P->dx);
  const int wf = P->with_forcing && F != nullptr;
  // prefactors as the reference forms them: real_t(dt / (2 * dx * rho)), real_t(dt / (2 dx)),
  // real_t(nu * dt / dx / dx), real_t(0.5 / dx)  (navier_stokes_flow_simulators.py:463,469,482,495)
  const T pF = T(dt / (2.0 * double(dx) * P->rho));
  const T pB = T(dt / (2.0 * double(dx)));
  const T pD = T(P->nu * dt / double(dx) / double(dx));
  const T pC = T(0.5 / double(dx));
  StageTimer& tm = P->timer;
  tm.begin(s);
  k_sa<T><<<grid, block, 0, s>>>(g, w, u, F, B, pF, wf);
  SOPHT_CHECK_LAUNCH();
  tm.mark(0, s);
  k_sb<T><<<grid, block, 0, s>>>(g, w, B, ps, pB);
  SOPHT_CHECK_LAUNCH();
  tm.mark(1, s);
  k_sc<T><<<grid, block, 0, s>>>(g, ps, w, F, pD, make_pen(P), wf);
  SOPHT_CHECK_LAUNCH();
  tm.mark(2, s);
  const int64_t n = g.nz * g.ny * g.nx;
  void* const sol[3] = {ps, ps + n, ps + 2 * n};
  const void* const rhs[3] = {w, w + n, w + 2 * n};
  const int64_t st[3] = {g.ny * g.nx, g.nx, 1};
  int rc = poisson_vector_solve_timed(P->poisson, sol, st, rhs, st, s, &tm, 3);
  if (rc) return rc;
  const T U0 = P->with_free_stream ? T(U[0]) : T(0);
  const T U1 = P->with_free_stream ? T(U[1]) : T(0);
  const T U2 = P->with_free_stream ? T(U[2]) : T(0);
  k_s2<T><<<grid, block, 0, s>>>(g, ps, u, pC, U0, U1, U2);
  SOPHT_CHECK_LAUNCH();
  tm.mark(8, s);
  tm.end();
  return SOPHT_OK;
}

template <typename T>
int recover_impl(sopht_ns3d* P, void* vort, void* vel, void* psi, const double* U, hipStream_t s) {
  // Penalise in place with the race-free standalone kernel, then Poisson, then S2.
  Grid g{P->nz, P->ny, P->nx};
  const int64_t st[3] = {g.ny * g.nx, g.nx, 1};
  const int64_t n = g.nz * g.ny * g.nx;
  T* w = static_cast<T*>(vort);
  T* ps = static_cast<T*>(psi);
  for (int c = 0; c < 3; ++c) {
    int rc = sopht_penalise_field_boundary_3d(P->dtype, w + c * n, st, g.nz, g.ny, g.nx, P->width,
                                              P->sine, s);
    if (rc) return rc;
  }
  void* const sol[3] = {ps, ps + n, ps + 2 * n};
  const void* const rhs[3] = {w, w + n, w + 2 * n};
  int rc = sopht_poisson3d_vector_solve(P->poisson, sol, st, rhs, st, s);
  if (rc) return rc;
  const T U0 = P->with_free_stream ? T(U[0]) : T(0);
  const T U1 = P->with_free_stream ? T(U[1]) : T(0);
  const T U2 = P->with_free_stream ? T(U[2]) : T(0);
  k_s2<T><<<cell_grid(g.nz, g.ny, g.nx), dim3(BX, BY), 0, s>>>(g, ps, static_cast<T*>(vel),
                                                                T(0.5 / P->dx), U0, U1, U2);
  SOPHT_CHECK_LAUNCH();
  return SOPHT_OK;
}

}  // namespace
}  // namespace sopht

using namespace sopht;

extern "C" {

int sopht_ns3d_create(int dtype, int64_t nz, int64_t ny, int64_t nx, double x_range, double nu,
                      double rho, int width, int with_forcing, int with_free_stream,
                      void* stream, sopht_ns3d** out) {
  SOPHT_REQUIRE(out != nullptr, "plan pointer is null");
  *out = nullptr;
  SOPHT_REQUIRE(dtype == SOPHT_F32 || dtype == SOPHT_F64, "dtype must be SOPHT_F32 or SOPHT_F64");
  SOPHT_REQUIRE(nz >= 3 && ny >= 3 && nx >= 3, "grid must be at least 3 cells per axis");
  SOPHT_REQUIRE(width >= 0 && width <= 8, "penalty_zone_width must be in [0, 8]");
  SOPHT_REQUIRE(2 * width <= nz && 2 * width <= ny && 2 * width <= nx, "grid too small for penalty width");
  sopht_ns3d* P = new (std::nothrow) sopht_ns3d();
  SOPHT_REQUIRE(P != nullptr, "out of host memory");
  P->dtype = dtype;
  P->nz = nz; P->ny = ny; P->nx = nx;
  P->x_range = x_range;
  P->dx = x_range / double(nx);
  if (dtype == SOPHT_F32) P->dx = double(float(P->dx));
  P->nu = nu;
  P->rho = rho;
  P->width = width;
  P->with_forcing = with_forcing;
  P->with_free_stream = with_free_stream;
  // sine factors sin(pi/(2 w dx) * distance from the first / last cell centre); cell
  // centres are (i + 1/2) dx, so the distance of slab cell i is i dx
  // (penalise_field_boundary_3d.py:257-264,282-284,294-296)
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < width; ++i) {
      const double pre = (M_PI / 2.0) / (double(width) * P->dx);
      P->sine[a * 2 * width + i] = std::sin(pre * (double(i) * P->dx));
      P->sine[a * 2 * width + width + i] = std::sin(pre * (double(width - 1 - i) * P->dx));
    }
  int rc = sopht_poisson3d_create(dtype, nz, ny, nx, x_range, stream, &P->poisson);
  if (rc) {
    delete P;
    return rc;
  }
  P->partial_cap = 4096;
  if (hipMalloc(&P->partial, P->partial_cap * 8) != hipSuccess) {
    sopht_ns3d_destroy(P);
    set_last_error("hipMalloc of the reduction scratch failed", __FILE__, __LINE__);
    return SOPHT_ERR_HIP;
  }
  *out = P;
  return SOPHT_OK;
}

int sopht_ns3d_enable_timing(sopht_ns3d* P, int enable) {
  SOPHT_REQUIRE(P != nullptr, "plan is null");
  if (enable) {
    SOPHT_CHECK_HIP(P->timer.enable(kNumStages) == SOPHT_OK ? hipSuccess : hipErrorUnknown);
  } else {
    P->timer.disable();
  }
  return SOPHT_OK;
}

int sopht_ns3d_stage_times(sopht_ns3d* P, double* ms, int n, int64_t* steps, int reset) {
  SOPHT_REQUIRE(P != nullptr, "plan is null");
  P->timer.harvest(0);
  P->timer.harvest(1);
  for (int i = 0; i < n && i < kNumStages; ++i) ms[i] = P->timer.acc[i];
  if (steps) *steps = P->timer.steps;
  if (reset) {
    for (double& a : P->timer.acc) a = 0.0;
    P->timer.steps = 0;
  }
  return kNumStages;
}

const char* sopht_ns3d_stage_name(int i) {
  return (i >= 0 && i < kNumStages) ? kStageNames[i] : "";
}

int sopht_ns3d_destroy(sopht_ns3d* P) {
  if (!P) return SOPHT_OK;
  P->timer.disable();
  sopht_poisson3d_destroy(P->poisson);
  if (P->partial) (void)hipFree(P->partial);
  delete P;
  return SOPHT_OK;
}

int sopht_ns3d_step(sopht_ns3d* P, void* vort, void* vel, void* psi, void* forcing, void* buffer,
                    double dt, const double* U, void* stream) {
  SOPHT_REQUIRE(P != nullptr, "plan is null");
  SOPHT_REQUIRE(vort && vel && psi && buffer, "null field pointer");
  SOPHT_REQUIRE(!P->with_forcing || forcing, "forcing field required when with_forcing");
  SOPHT_REQUIRE(!P->with_free_stream || U, "free stream velocity required");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return P->dtype == SOPHT_F32 ? step_impl<float>(P, vort, vel, psi, forcing, buffer, dt, U, s)
                               : step_impl<double>(P, vort, vel, psi, forcing, buffer, dt, U, s);
}

int sopht_ns3d_recover_velocity(sopht_ns3d* P, void* vort, void* vel, void* psi, const double* U,
                                void* stream) {
  SOPHT_REQUIRE(P != nullptr, "plan is null");
  SOPHT_REQUIRE(!P->with_free_stream || U, "free stream velocity required");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return P->dtype == SOPHT_F32 ? recover_impl<float>(P, vort, vel, psi, U, s)
                               : recover_impl<double>(P, vort, vel, psi, U, s);
}

int sopht_ns3d_max_abs_velocity_sum(sopht_ns3d* P, const void* vel, void* out, void* stream) {
  SOPHT_REQUIRE(P != nullptr, "plan is null");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t n = P->nz * P->ny * P->nx;
  const int nb = int(std::min<int64_t>(P->partial_cap, (n + 255) / 256));
  if (P->dtype == SOPHT_F32) {
    k_max_abs_sum<float><<<nb, 256, 0, s>>>(static_cast<const float*>(vel), n,
                                           static_cast<float*>(P->partial));
    SOPHT_CHECK_LAUNCH();
    k_max_final<float><<<1, 256, 0, s>>>(static_cast<const float*>(P->partial), nb, static_cast<float*>(out));
  } else {
    k_max_abs_sum<double><<<nb, 256, 0, s>>>(static_cast<const double*>(vel), n,
                                            static_cast<double*>(P->partial));
    SOPHT_CHECK_LAUNCH();
    k_max_final<double><<<1, 256, 0, s>>>(static_cast<const double*>(P->partial), nb, static_cast<double*>(out));
  }
  SOPHT_CHECK_LAUNCH();
  return SOPHT_OK;
}

}  // extern "C"
