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
// The sweeps march up z in (x, y) tiles with x-fastest lanes (coalesced rows).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <new>

#include "common.h"
#include "ns3d_plan.h"
#include "sweeps_3d.h"

// Stage order of one step, for sopht_ns3d_stage_times: the two vorticity sweeps, the
// Poisson solver's launches (five or three passes, or the fast-diagonalisation stages), the
// velocity sweep.
static const char* const kStagesFivePass[] = {
    "vorticity_forcing_advect", "vorticity_diffuse_penalise", "poisson_x_r2c", "poisson_y_fwd",
    "poisson_z_conv", "poisson_y_inv", "poisson_x_c2r", "velocity_curl"};
static const char* const kStagesThreePass[] = {
    "vorticity_forcing_advect", "vorticity_diffuse_penalise", "poisson_xy_fwd", "poisson_z_conv",
    "poisson_yx_inv", "velocity_curl"};
static const char* const kStagesFastdiag[] = {
    "vorticity_forcing_advect", "vorticity_diffuse_penalise", "fastdiag_z_fwd", "fastdiag_y_fwd",
    "fastdiag_x_fwd_scale", "fastdiag_z_bwd", "fastdiag_yx_bwd", "velocity_curl"};

static int stage_count(const sopht_ns3d* P) {
  return P->fastdiag ? 8 : 3 + sopht::poisson_stage_count(P->poisson);
}
static const char* const* stage_names(const sopht_ns3d* P) {
  if (P->fastdiag) return kStagesFastdiag;
  return sopht::poisson_stage_count(P->poisson) == 3 ? kStagesThreePass : kStagesFivePass;
}

namespace sopht {
namespace {

// The velocity sweep's launch: the SC launch's tiles, with 32-plane items for FP32 grids that
// still give 512 of them (256^3: 0.092 ms against 0.098 at 16 planes; FP64 and the 192^3 grids
// were no faster with longer items).
template <typename T>
MarchLaunch s2_launch(const MarchLaunch& ml, int64_t nz) {
  MarchLaunch l = ml;
  const int64_t items = int64_t(ml.tiles_x) * ml.tiles_y * ceil_div(nz, int64_t(32));
  if (sizeof(T) == 4 && ml.chunk == kMarchChunk && items >= 512) {
    l.chunk = 32;
    l.blocks = unsigned(items);
    l.remap = items % 8 == 0 ? 1 : 0;
  }
  return l;
}

// One step from (w_in, u_in, F) to (w_out, u_out); w_out may be w_in and u_out u_in (the
// in-place step of the reference). zero_forcing clears F in the SC sweep.
template <typename T>
int step_impl(sopht_ns3d* P, const void* vort_in, void* vort_out, const void* vel_in,
              void* vel_out, void* psi, void* forcing, int zero_forcing, double dt,
              const double* U, hipStream_t s, void* vort1_out = nullptr, void* buffer = nullptr) {
  Grid g{P->nz, P->ny, P->nx};
  const T* w_in = static_cast<const T*>(vort_in);
  T* w = static_cast<T*>(vort_out);
  const T* u_in = static_cast<const T*>(vel_in);
  T* u = static_cast<T*>(vel_out);
  T* ps = static_cast<T*>(psi);
  T* F = static_cast<T*>(forcing);
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
  // S1 with 16-byte accesses when the rows allow them, else the scalar sweep
  const bool s1v = P->s1_vec != 0;
  const int s1_rows = s1v ? P->s1_rows : S1Tile<T>::TY;
  const int s1_cols = s1v ? P->s1_cols : S1Tile<T>::TX;
  const int s1_chunk = s1v ? P->s1_chunk : kS1Chunk;
  const S1Map s1m{int(ceil_div(g.nx, s1_cols)), int(ceil_div(g.ny, s1_rows)), s1_rows, s1_chunk, s1_cols};
  const int64_t s1_items = int64_t(s1m.tiles_x) * s1m.tiles_y * ceil_div(g.nz, s1_chunk);
  const bool flags_ok = s1_items <= P->s1flags_cap;
  int* const fflags = flags_ok ? P->s1flags : nullptr;
  {
    const int remap = s1_items % 8 == 0 ? 1 : 0;
    if constexpr (sizeof(T) == 4) {
      if (s1v && s1_cols == 128) {
        using W = S1VTile<T, 128, 8>;
        k_s1v<T, W::TX, W::TY><<<unsigned(s1_items), W::NT, 0, s>>>(g, w_in, u_in, F, ps, pF, pB, wf, s1_chunk,
                                                                    s1m.tiles_x, s1m.tiles_y, remap,
                                                                    static_cast<T*>(vort1_out), fflags);
      }
    }
    if (s1v && s1_cols == 64)
      k_s1v<T><<<unsigned(s1_items), S1VTile<T>::NT, 0, s>>>(g, w_in, u_in, F, ps, pF, pB, wf, s1_chunk,
                                                            s1m.tiles_x, s1m.tiles_y, remap,
                                                            static_cast<T*>(vort1_out), fflags);
    else if (!s1v)
      k_s1<T><<<unsigned(s1_items), S1Tile<T>::NT, 0, s>>>(g, w_in, u_in, F, ps, pF, pB, wf, kS1Chunk,
                                                          s1m.tiles_x, s1m.tiles_y, remap,
                                                          static_cast<T*>(vort1_out), fflags);
    SOPHT_CHECK_LAUNCH();
  }
  tm.mark(0, s);
  const MarchLaunch ml = march_launch(g.nz, g.ny, g.nx, march_vec<T>(g.nx, P->width), true);
  if (P->filter_order >= 0) {
    // diffusion alone, the filter (buffer components 0 and 1, as the reference), then the
    // penalisation (navier_stokes_flow_simulators.py:466-474)
    Pen none = make_pen(P);
    none.w = 0;
    launch_sc<T>(ml, g, ps, w, F, pD, none, wf && zero_forcing, s, fflags, s1m);
    SOPHT_CHECK_LAUNCH();
    const int64_t n = g.nz * g.ny * g.nx;
    const int64_t st[3] = {g.ny * g.nx, g.nx, 1};
    T* bf = static_cast<T*>(buffer);
    if (fused_filter_supported(P->filter_order, g.nx)) {
      // one launch per axis for the three components (the buffer holds the multiplicative
      // filter's intermediate passes)
      int rc = fused_vorticity_filter(P->dtype, P->filter_order, P->filter_type, w, bf, g.nz, g.ny,
                                      g.nx, s);
      if (rc) return rc;
    } else {
      for (int c = 0; c < 3; ++c) {
        int rc = laplacian_filter_field(P->dtype, P->filter_order, P->filter_type, w + c * n, st,
                                        g.nz, g.ny, g.nx, bf, st, bf + n, st, 1, s);
        if (rc) return rc;
      }
    }
    for (int c = 0; c < 3; ++c) {
      int rc = sopht_penalise_field_boundary_3d(P->dtype, w + c * n, st, g.nz, g.ny, g.nx, P->width,
                                                P->sine, s);
      if (rc) return rc;
    }
  } else {
    launch_sc<T>(ml, g, ps, w, F, pD, make_pen(P), wf && zero_forcing, s, fflags, s1m);
    SOPHT_CHECK_LAUNCH();
  }
  tm.mark(1, s);
  const int64_t n = g.nz * g.ny * g.nx;
  void* const sol[3] = {ps, ps + n, ps + 2 * n};
  const void* const rhs[3] = {w, w + n, w + 2 * n};
  const int64_t st[3] = {g.ny * g.nx, g.nx, 1};
  int rc = P->fastdiag ? fastdiag_vector_solve_timed(P->fastdiag, sol, st, rhs, st, s, &tm, 2)
                       : poisson_vector_solve_timed(P->poisson, sol, st, rhs, st, s, &tm, 2);
  if (rc) return rc;
  const T U0 = P->with_free_stream ? T(U[0]) : T(0);
  const T U1 = P->with_free_stream ? T(U[1]) : T(0);
  const T U2 = P->with_free_stream ? T(U[2]) : T(0);
  const MarchLaunch l2 = s2_launch<T>(ml, g.nz);
  const bool fused_max = int64_t(l2.blocks) <= P->s2max_cap;
  launch_s2<T>(l2, g, ps, u, pC, U0, U1, U2, s, fused_max ? static_cast<T*>(P->s2max) : nullptr);
  SOPHT_CHECK_LAUNCH();
  P->s2_partials = fused_max ? int64_t(l2.blocks) : 0;
  P->s2_velocity = u;
  tm.mark(stage_count(P) - 1, s);
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
  int rc = P->fastdiag ? fastdiag_vector_solve_timed(P->fastdiag, sol, st, rhs, st, s, nullptr, 0)
                       : sopht_poisson3d_vector_solve(P->poisson, sol, st, rhs, st, s);
  if (rc) return rc;
  const T U0 = P->with_free_stream ? T(U[0]) : T(0);
  const T U1 = P->with_free_stream ? T(U[1]) : T(0);
  const T U2 = P->with_free_stream ? T(U[2]) : T(0);
  const MarchLaunch ml = s2_launch<T>(march_launch(g.nz, g.ny, g.nx, march_vec<T>(g.nx, P->width), true), g.nz);
  const bool fused_max = int64_t(ml.blocks) <= P->s2max_cap;
  launch_s2<T>(ml, g, ps, static_cast<T*>(vel), T(0.5 / P->dx), U0, U1, U2, s,
               fused_max ? static_cast<T*>(P->s2max) : nullptr);
  SOPHT_CHECK_LAUNCH();
  P->s2_partials = fused_max ? int64_t(ml.blocks) : 0;
  P->s2_velocity = vel;
  return SOPHT_OK;
}

}  // namespace
}  // namespace sopht

using namespace sopht;

extern "C" {

int sopht_ns3d_create(int dtype, int64_t nz, int64_t ny, int64_t nx, double x_range, double nu,
                      double rho, int width, int with_forcing, int with_free_stream,
                      void* stream, sopht_ns3d** out) {
  return sopht_ns3d_create_with_poisson_solver(dtype, nz, ny, nx, x_range, nu, rho, width,
                                               with_forcing, with_free_stream, 0, stream, out);
}

int sopht_ns3d_create_with_poisson_solver(int dtype, int64_t nz, int64_t ny, int64_t nx,
                                          double x_range, double nu, double rho, int width,
                                          int with_forcing, int with_free_stream,
                                          int poisson_solver_type, void* stream, sopht_ns3d** out) {
  SOPHT_REQUIRE(out != nullptr, "plan pointer is null");
  *out = nullptr;
  SOPHT_REQUIRE(poisson_solver_type == 0 || poisson_solver_type == 1,
                "poisson_solver_type must be 0 (greens function) or 1 (fast diagonalisation)");
  SOPHT_REQUIRE(dtype == SOPHT_F32 || dtype == SOPHT_F64, "dtype must be SOPHT_F32 or SOPHT_F64");
  SOPHT_REQUIRE(nz >= 3 && ny >= 3 && nx >= 3, "grid must be at least 3 cells per axis");
  SOPHT_REQUIRE(width >= 0 && width <= 8, "penalty_zone_width must be in [0, 8]");
  SOPHT_REQUIRE(2 * width <= nz && 2 * width <= ny && 2 * width <= nx, "grid too small for penalty width");
  SOPHT_REQUIRE(3 * nz * ny * nx < (int64_t(1) << 31), "grid too large (3 nz ny nx must be below 2^31)");
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
  // (penalise_field_boundary_3d.py:69,85-102)
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < width; ++i) {
      const double pre = (M_PI / 2.0) / (double(width) * P->dx);
      P->sine[a * 2 * width + i] = std::sin(pre * (double(i) * P->dx));
      P->sine[a * 2 * width + width + i] = std::sin(pre * (double(width - 1 - i) * P->dx));
    }
  int rc = poisson_solver_type == 1
               ? sopht_fastdiag3d_create(dtype, nz, ny, nx, P->dx, stream, &P->fastdiag)
               : sopht_poisson3d_create(dtype, nz, ny, nx, x_range, stream, &P->poisson);
  if (rc) {
    delete P;
    return rc;
  }
  P->partial_cap = 4096;
  P->s2max_cap = int64_t(march_launch(nz, ny, nx, 1, true).blocks);  // VEC 1 has the most
  P->s2_partials = 0;
  P->s2_velocity = nullptr;
  P->filter_order = -1;
  P->filter_type = 0;
  {
    // The vector sweep needs whole 16-byte row vectors and component byte offsets below 2^32.
    // Its work items hold the most planes (64, 32 or 16) that still make 384 items, one round
    // of workgroups at one and a half to three per CU. Measured S1 times: 256^3 FP32 (128 x 8
    // tiles) 0.21 ms at 32 planes against 0.24 at 16 (two chunks per XCD resident at once
    // doubled the L2 misses) and 0.29 at 64; 256^3 FP64 0.38 ms at 64 against 0.43 at 32;
    // 192^3 FP64 0.19 ms at 32 against 0.22 at 16; 192^3 FP32 0.11 ms at 16 against 0.16 at
    // 32. Grids with fewer than 256 items of 16 planes keep the scalar sweep (twice the
    // workgroups). The environment can force either: SOPHT_S1_SCALAR=1 (the two sweeps give
    // the same bits) and SOPHT_S1_CHUNK=16/32/64. FP32 rows that are a multiple of 128 cells take tiles of 128
    // cells by 8 rows, others 64 by 16 (FP64: 64 by 8); SOPHT_S1_COLS=64/128 forces the width.
    const int64_t vec = 16 / (dtype == SOPHT_F32 ? 4 : 8);
    const int64_t bytes = nz * ny * nx * (dtype == SOPHT_F32 ? 4 : 8);
    P->s1_cols = dtype == SOPHT_F32 && nx % 128 == 0 ? 128 : 64;
    if (const char* cols = std::getenv("SOPHT_S1_COLS")) {
      const int v = std::atoi(cols);
      if (dtype == SOPHT_F32 && (v == 64 || v == 128)) P->s1_cols = v;
    }
    P->s1_rows = P->s1_cols == 128 ? S1VTile<float, 128, 8>::TY
                                   : (dtype == SOPHT_F32 ? S1VTile<float>::TY : S1VTile<double>::TY);
    const int64_t tiles = ceil_div(nx, P->s1_cols) * ceil_div(ny, P->s1_rows);
    const char* scalar = std::getenv("SOPHT_S1_SCALAR");
    const char* chunk = std::getenv("SOPHT_S1_CHUNK");
    P->s1_chunk = 16;
    for (const int c : {64, 32})
      if (tiles * ceil_div(nz, int64_t(c)) >= 384) {
        P->s1_chunk = c;
        break;
      }
    bool fill = tiles * ceil_div(nz, 16) >= 256;
    if (chunk) {
      const int v = std::atoi(chunk);
      if (v == 16 || v == 32 || v == 64) {
        P->s1_chunk = v;
        fill = true;
      }
    }
    P->s1_vec = nx % vec == 0 && bytes < (int64_t(1) << 32) && fill && !(scalar && std::atoi(scalar) == 1);
  }
  {
    const int64_t i32 = ceil_div(nx, S1Tile<float>::TX) * ceil_div(ny, S1Tile<float>::TY) * ceil_div(nz, kS1Chunk);
    const int64_t i64 = ceil_div(nx, S1Tile<double>::TX) * ceil_div(ny, S1Tile<double>::TY) * ceil_div(nz, kS1Chunk);
    P->s1flags_cap = dtype == SOPHT_F32 ? i32 : i64;
  }
  if (hipMalloc(&P->partial, P->partial_cap * 8) != hipSuccess ||
      hipMalloc(&P->s2max, P->s2max_cap * 8) != hipSuccess ||
      hipMalloc(&P->s1flags, P->s1flags_cap * sizeof(int)) != hipSuccess) {
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
    SOPHT_CHECK_HIP(P->timer.enable(stage_count(P)) == SOPHT_OK ? hipSuccess : hipErrorUnknown);
  } else {
    P->timer.disable();
  }
  return SOPHT_OK;
}

int sopht_ns3d_stage_times(sopht_ns3d* P, double* ms, int n, int64_t* steps, int reset) {
  SOPHT_REQUIRE(P != nullptr, "plan is null");
  P->timer.harvest(0);
  P->timer.harvest(1);
  const int ns = stage_count(P);
  for (int i = 0; i < n && i < ns; ++i) ms[i] = P->timer.acc[i];
  if (steps) *steps = P->timer.steps;
  if (reset) {
    for (double& a : P->timer.acc) a = 0.0;
    P->timer.steps = 0;
  }
  return ns;
}

int sopht_ns3d_poisson_schedule(const sopht_ns3d* P, int* passes, int* zpass_jobs) {
  SOPHT_REQUIRE(P != nullptr, "plan is null");
  if (!P->poisson) {  // the fast-diagonalisation solver
    if (passes) *passes = 0;
    if (zpass_jobs) *zpass_jobs = 0;
    return SOPHT_OK;
  }
  return sopht_poisson3d_schedule(P->poisson, passes, zpass_jobs);
}

int sopht_ns3d_s1_schedule(const sopht_ns3d* P, int* vector, int* zchunk, int* cols) {
  SOPHT_REQUIRE(P != nullptr, "plan is null");
  if (vector) *vector = P->s1_vec;
  if (zchunk) *zchunk = P->s1_vec ? P->s1_chunk : sopht::kS1Chunk;
  if (cols) *cols = P->s1_vec ? P->s1_cols : sopht::S1Tile<float>::TX;
  return SOPHT_OK;
}

const char* sopht_ns3d_stage_name(const sopht_ns3d* P, int i) {
  return (P && i >= 0 && i < stage_count(P)) ? stage_names(P)[i] : "";
}

int sopht_ns3d_destroy(sopht_ns3d* P) {
  if (!P) return SOPHT_OK;
  P->timer.disable();
  sopht_poisson3d_destroy(P->poisson);
  sopht_fastdiag3d_destroy(P->fastdiag);
  if (P->partial) (void)hipFree(P->partial);
  if (P->s2max) (void)hipFree(P->s2max);
  if (P->s1flags) (void)hipFree(P->s1flags);
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
  // buffer: u x omega1 never leaves the S1 sweep; the vorticity filter uses it
  return P->dtype == SOPHT_F32
             ? step_impl<float>(P, vort, vort, vel, vel, psi, forcing, 1, dt, U, s, nullptr, buffer)
             : step_impl<double>(P, vort, vort, vel, vel, psi, forcing, 1, dt, U, s, nullptr, buffer);
}

int sopht_ns3d_trajectory_step(sopht_ns3d* P, const void* vort_in, const void* vel_in,
                               const void* forcing, void* vort_out, void* vel_out, void* psi,
                               void* vort1_out, double dt, const double* U, void* stream) {
  SOPHT_REQUIRE(P != nullptr, "plan is null");
  SOPHT_REQUIRE(vort_in && vel_in && vort_out && vel_out && psi, "null field pointer");
  SOPHT_REQUIRE(vort_in != vort_out && vel_in != vel_out, "trajectory step is out of place");
  SOPHT_REQUIRE(P->filter_order < 0, "trajectories (and the adjoint) do not cover the vorticity filter");
  SOPHT_REQUIRE(!P->with_forcing || forcing, "forcing field required when with_forcing");
  SOPHT_REQUIRE(!P->with_free_stream || U, "free stream velocity required");
  hipStream_t s = static_cast<hipStream_t>(stream);
  void* F = const_cast<void*>(forcing);  // read only: zero_forcing is 0
  SOPHT_REQUIRE(vort1_out != vort_in && vort1_out != vort_out, "omega1 needs its own buffer");
  return P->dtype == SOPHT_F32
             ? step_impl<float>(P, vort_in, vort_out, vel_in, vel_out, psi, F, 0, dt, U, s, vort1_out)
             : step_impl<double>(P, vort_in, vort_out, vel_in, vel_out, psi, F, 0, dt, U, s, vort1_out);
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

int sopht_ns3d_set_filter(sopht_ns3d* P, int order, int type) {
  SOPHT_REQUIRE(P != nullptr, "plan is null");
  SOPHT_REQUIRE(type == 0 || type == 1, "Invalid filter type");
  P->filter_order = order < 0 ? -1 : order;
  P->filter_type = type;
  return SOPHT_OK;
}

int sopht_ns3d_last_velocity_max(sopht_ns3d* P, const void* vel, void* out, void* stream) {
  SOPHT_REQUIRE(P != nullptr, "plan is null");
  SOPHT_REQUIRE(P->s2_partials > 0 && vel == P->s2_velocity,
                "no fused maxima for this velocity field (the last sweep wrote another one)");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = int(P->s2_partials);
  if (P->dtype == SOPHT_F32)
    k_max_final<float><<<1, 256, 0, s>>>(static_cast<const float*>(P->s2max), nb, static_cast<float*>(out));
  else
    k_max_final<double><<<1, 256, 0, s>>>(static_cast<const double*>(P->s2max), nb, static_cast<double*>(out));
  SOPHT_CHECK_LAUNCH();
  return SOPHT_OK;
}

}  // extern "C"
