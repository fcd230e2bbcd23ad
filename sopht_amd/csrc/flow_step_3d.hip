// The 3D unbounded Navier-Stokes step with forcing (navier_stokes_flow_simulators.py:449-498)
// as fused sweeps on gfx950. Fields are contiguous (3, nz, ny, nx), x fastest.
//
//   S1  : omega1 = omega + pF curl_h(F) (interior; border omega unchanged), B = u x omega1,
//         omega2 = omega1 + pB curl_h(B) (interior) into the stream-function field, which the
//         Poisson solve overwrites later in the step; omega1 and B live only in LDS and
//         registers.
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
    "vorticity_forcing_advect",    // S1
    "vorticity_diffuse_penalise",  // SC
    "poisson_x_r2c",            // K1
    "poisson_y_fwd",            // K2
    "poisson_z_conv",           // K3
    "poisson_y_inv",            // K4
    "poisson_x_c2r",            // K5
    "velocity_curl",            // S2
};
static constexpr int kNumStages = 8;

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

// z planes per S1 work item (two extra planes of B and three of F are recomputed per item)
constexpr int kS1Chunk = 16;

// S1: omega2 = omega1 + pB curl_h(u x omega1), omega1 = omega +
// pF curl_h(F), with neither omega1 nor B stored in HBM. A workgroup owns a TX x TY tile of
// (x, y) columns over a chunk of ZC planes and marches up z. Iteration i
//   (1) stages F plane i+2 (tile plus a 2-cell ring) in a 4-plane LDS ring,
//   (2) produces omega1 and B on plane i over the tile plus its 1-cell edge ring (B goes to a
//       3-plane LDS ring; the tile's own omega1, B(i-1), B(i-2) stay in registers),
//   (3) after the one barrier, consumes plane i-1: its curl_h(B) takes the in-plane
//       neighbours from LDS and B(i), B(i-2) from registers.
// Ring depths make one barrier per plane enough: the slot written at iteration i was last
// read before the previous iteration's barrier.
template <typename T>
struct S1Tile {
  static constexpr int TX = 64, TY = sizeof(T) == 4 ? 8 : 4, NT = TX * TY;
  static constexpr int FX = TX + 4, FY = TY + 4, FP = FX * FY;  // F staging box (2-cell ring)
  static constexpr int GX = TX + 2, GY = TY + 2, GP = GX * GY;  // B box (1-cell ring)
  static constexpr int NEDGE = 2 * (TX + TY);                    // edge-ring items (no corners)
};

template <typename T>
__global__ void __launch_bounds__(S1Tile<T>::NT) k_s1(Grid g, const T* w, const T* u, const T* F,
                                                      T* w2out, T pF, T pB, int with_forcing,
                                                      int zchunk, int tiles_x, int tiles_y,
                                                      int xcd_remap) {
  using S = S1Tile<T>;
  constexpr int TX = S::TX, TY = S::TY, NT = S::NT, FX = S::FX, FP = S::FP, GX = S::GX, GP = S::GP;
  __shared__ T fs[4][3][FP];
  __shared__ T bs[3][3][GP];
  // Work item of this workgroup. Workgroups are dealt to the 8 XCDs in turn; with the remap,
  // XCD k takes a contiguous run of items (neighbouring y tiles share halo rows in its L2).
  int64_t item = blockIdx.x;
  if (xcd_remap) item = (item % 8) * (gridDim.x / 8) + item / 8;
  const int64_t tx_i = item % tiles_x;
  const int64_t ty_i = (item / tiles_x) % tiles_y;
  const int64_t zc_i = item / (int64_t(tiles_x) * tiles_y);
  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  const int64_t x0 = tx_i * TX, y0 = ty_i * TY;
  const int64_t z0 = zc_i * zchunk;
  const int64_t z1 = z0 + zchunk < g.nz ? z0 + zchunk : g.nz;
  const int64_t n = g.nz * g.ny * g.nx, sz = g.ny * g.nx, sy = g.nx;
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

  auto stage_f = [&](int64_t q) {  // F plane q -> ring slot q & 3 (zeros outside the grid)
    T(*dst)[FP] = fs[q & 3];
    const bool qin = q >= 0 && q < g.nz;
    for (int k = tid; k < FP; k += NT) {
      const int64_t x = x0 + (k % FX) - 2, y = y0 + (k / FX) - 2;
      const bool in = qin && x >= 0 && x < g.nx && y >= 0 && y < g.ny;
      const int64_t i = in ? q * sz + y * sy + x : 0;
      dst[0][k] = in ? F[i] : T(0);
      dst[1][k] = in ? F[n + i] : T(0);
      dst[2][k] = in ? F[2 * n + i] : T(0);
    }
  };
  // omega1 and B = u x omega1 at (x0 + xr, y0 + yr, q)
  auto produce = [&](int xr, int yr, int64_t q, T b[3], T a[3]) {
    const int64_t x = x0 + xr, y = y0 + yr;
    const int64_t i = q * sz + y * sy + x;
    a[0] = w[i];
    a[1] = w[n + i];
    a[2] = w[2 * n + i];
    if (with_forcing && g.interior(q, y, x)) {
      const int c = (yr + 2) * FX + xr + 2;
      const T(*Fc)[FP] = fs[q & 3];
      const T(*Fm)[FP] = fs[(q - 1) & 3];
      const T(*Fp)[FP] = fs[(q + 1) & 3];
      const T c0 = Fc[2][c + FX] - Fc[2][c - FX] - Fp[1][c] + Fm[1][c];
      const T c1 = Fp[0][c] - Fm[0][c] - Fc[2][c + 1] + Fc[2][c - 1];
      const T c2 = Fc[1][c + 1] - Fc[1][c - 1] - Fc[0][c + FX] + Fc[0][c - FX];
      a[0] = a[0] + pF * c0;
      a[1] = a[1] + pF * c1;
      a[2] = a[2] + pF * c2;
    }
    const T u0 = u[i], u1 = u[n + i], u2 = u[2 * n + i];
    // elementwise_ops_3d.py:427-447: B_i = u_j w_k - w_j u_k
    b[0] = u1 * a[2] - a[1] * u2;
    b[1] = u2 * a[0] - a[2] * u0;
    b[2] = u0 * a[1] - a[0] * u1;
  };

  if (with_forcing) {
    stage_f(z0 - 2);
    stage_f(z0 - 1);
    stage_f(z0);
  }
  __syncthreads();
  T bm2[3] = {T(0), T(0), T(0)}, bm1[3] = {T(0), T(0), T(0)}, a1prev[3] = {T(0), T(0), T(0)};
  const int gt = (ty + 1) * GX + tx + 1;
  for (int64_t i = z0 - 1; i <= z1; ++i) {
    if (with_forcing) stage_f(i + 2);
    T b0[3] = {T(0), T(0), T(0)}, a1[3] = {T(0), T(0), T(0)};
    if (i >= 0 && i < g.nz) {
      T(*Bq)[GP] = bs[i % 3];
      if (tv) {
        produce(tx, ty, i, b0, a1);
        Bq[0][gt] = b0[0];
        Bq[1][gt] = b0[1];
        Bq[2][gt] = b0[2];
      }
      if (ev) {
        T eb[3], ea[3];
        produce(ex, ey, i, eb, ea);
        const int ge = (ey + 1) * GX + ex + 1;
        Bq[0][ge] = eb[0];
        Bq[1][ge] = eb[1];
        Bq[2][ge] = eb[2];
      }
    }
    __syncthreads();
    const int64_t r = i - 1;  // consumed plane: B(r-1) = bm2, B(r+1) = b0, omega1(r) = a1prev
    if (r >= z0 && tv) {
      const int64_t x = x0 + tx, y = y0 + ty;
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
      const int64_t j = r * sz + y * sy + x;
      w2out[j] = o0;
      w2out[n + j] = o1;
      w2out[2 * n + j] = o2;
    }
    for (int c = 0; c < 3; ++c) {
      bm2[c] = bm1[c];
      bm1[c] = b0[c];
      a1prev[c] = a1[c];
    }
  }
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
  (void)buffer;  // u x omega1 never leaves the S1 sweep
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
  {
    using ST = S1Tile<T>;
    const int zchunk = kS1Chunk;
    const int tiles_x = int(ceil_div(g.nx, ST::TX)), tiles_y = int(ceil_div(g.ny, ST::TY));
    const int64_t items = int64_t(tiles_x) * tiles_y * ceil_div(g.nz, zchunk);
    k_s1<T><<<unsigned(items), ST::NT, 0, s>>>(g, w, u, F, ps, pF, pB, wf, zchunk, tiles_x, tiles_y,
                                              items % 8 == 0 ? 1 : 0);
    SOPHT_CHECK_LAUNCH();
  }
  tm.mark(0, s);
  k_sc<T><<<grid, block, 0, s>>>(g, ps, w, F, pD, make_pen(P), wf);
  SOPHT_CHECK_LAUNCH();
  tm.mark(1, s);
  const int64_t n = g.nz * g.ny * g.nx;
  void* const sol[3] = {ps, ps + n, ps + 2 * n};
  const void* const rhs[3] = {w, w + n, w + 2 * n};
  const int64_t st[3] = {g.ny * g.nx, g.nx, 1};
  int rc = poisson_vector_solve_timed(P->poisson, sol, st, rhs, st, s, &tm, 2);
  if (rc) return rc;
  const T U0 = P->with_free_stream ? T(U[0]) : T(0);
  const T U1 = P->with_free_stream ? T(U[1]) : T(0);
  const T U2 = P->with_free_stream ? T(U[2]) : T(0);
  k_s2<T><<<grid, block, 0, s>>>(g, ps, u, pC, U0, U1, U2);
  SOPHT_CHECK_LAUNCH();
  tm.mark(7, s);
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
