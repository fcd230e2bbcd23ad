// Passive transport step of a scalar or vector field on gfx950: one fused sweep per call
// computes, per component phi,
//   phi_a   = phi + E(phi, u)                  conservative ENO3 advection with -dt/dx on
//                                              [2:-2]^3, E = 0 elsewhere
//                                              (advection_timestep_3d.py:12-99,
//                                              advection_flux_3d.py:12-233)
//   phi_new = phi_a + pD lap_h(phi_a)          on [1:-1]^3, phi_a on the border
//                                              (diffusion_timestep_3d.py:12-80)
// i.e. the reference's advection step followed by its diffusion step, with the advected field
// never stored: it exists per z plane in LDS (tile + 1-cell ring) and, for the thread's own
// column, in registers. Arithmetic follows the reference's operation order (the six ENO
// sub-stencils accumulate into a zeroed flux in the order x front/back, y, z; the diffusion
// sum runs z+1, z-1, y+1, y-1, x+1, x-1, then -6 phi).
//
// A workgroup owns a TX x TY tile of (x, y) columns and marches up a chunk of z planes.
// Iteration i loads plane i of phi and u_x, u_y over the tile plus a 3-cell halo into LDS
// (the in-plane ENO stencils reach 2 cells beyond the ring cells the diffusion needs),
// slides each column's z-window of phi and u_z (planes i-2 .. i+2, registers), forms phi_a on
// plane i for the tile and its ring, then diffuses plane i-1. The result goes to a separate
// output (neighbouring workgroups still read the input's halo), which the entry point copies
// back into the field.
#include <algorithm>

#include "common.h"

namespace sopht {
namespace {

template <typename T>
struct TransportTile {
  static constexpr int TX = 64, TY = sizeof(T) == 4 ? 8 : 4, NT = TX * TY;
  static constexpr int H = 3;                        // halo of the input planes
  static constexpr int LX = TX + 2 * H, LY = TY + 2 * H, LP = LX * LY;
  static constexpr int RX = TX + 2, RY = TY + 2, RP = RX * RY;  // phi_a plane: tile + ring
  static constexpr int NRING = 2 * (TX + TY);       // ring cells (no corners)
};
constexpr int kTransportChunk = 16;

template <typename T>
__device__ __forceinline__ T eno_front(const T* F, const T* U) {  // F[k], U[k] for k = -2..2 at [k + 2]
  const T c13 = T(1.0 / 3.0), c56 = T(5.0 / 6.0), c16 = T(1.0 / 6.0);
  return (U[2] > -U[3]) ? (c13 * F[3] * U[3] + c56 * F[2] * U[2] - c16 * F[1] * U[1])
                        : (c13 * F[2] * U[2] + c56 * F[3] * U[3] - c16 * F[4] * U[4]);
}
template <typename T>
__device__ __forceinline__ T eno_back(const T* F, const T* U) {
  const T c13 = T(1.0 / 3.0), c56 = T(5.0 / 6.0), c16 = T(1.0 / 6.0);
  return (U[2] > -U[1]) ? (c13 * F[2] * U[2] + c56 * F[1] * U[1] - c16 * F[0] * U[0])
                        : (c13 * F[1] * U[1] + c56 * F[2] * U[2] - c16 * F[3] * U[3]);
}

template <typename T, int NC>
struct Comps {
  const T* in[NC];
  T* out[NC];
};

template <typename T, int NC>
__global__ void __launch_bounds__(TransportTile<T>::NT)
k_transport(Comps<T, NC> f, const T* __restrict__ ux, const T* __restrict__ uy,
            const T* __restrict__ uz, int nz, int ny, int nx, T inv_dx, T pD, int zchunk,
            int tiles_x, int tiles_y, int remap) {
  using S = TransportTile<T>;
  constexpr int TX = S::TX, TY = S::TY, H = S::H, LX = S::LX, LP = S::LP, RX = S::RX;
  __shared__ T sphi[NC][LP];
  __shared__ T sux[LP], suy[LP];
  __shared__ T sa[2][NC][S::RP];
  int64_t item = blockIdx.x;
  if (remap) item = (item % 8) * (gridDim.x / 8) + item / 8;
  const int x0 = int(item % tiles_x) * TX;
  const int y0 = int((item / tiles_x) % tiles_y) * TY;
  const int z0 = int(item / (int64_t(tiles_x) * tiles_y)) * zchunk;
  const int z1 = z0 + zchunk < nz ? z0 + zchunk : nz;
  const int sz = ny * nx, sy = nx;
  const int tid = threadIdx.x;
  // own cell, and (threads below NRING) one ring cell: x = -1 / TX columns, y = -1 / TY rows
  const int ox = tid % TX, oy = tid / TX;
  int ex = 0, ey = 0;
  if (tid < 2 * TY) {
    ex = (tid & 1) ? TX : -1;
    ey = tid >> 1;
  } else if (tid < S::NRING) {
    ex = (tid - 2 * TY) >> 1;
    ey = ((tid - 2 * TY) & 1) ? TY : -1;
  }
  const bool has_ring = tid < S::NRING;
  const int cx[2] = {x0 + ox, x0 + ex}, cy[2] = {y0 + oy, y0 + ey};
  const bool cin[2] = {cx[0] < nx && cy[0] < ny,
                       has_ring && cx[1] >= 0 && cx[1] < nx && cy[1] >= 0 && cy[1] < ny};
  // z-windows (planes i-2 .. i+2) of phi and u_z for the own [0] and ring [1] columns
  T wf[2][NC][5], wu[2][5];
  auto col_load = [&](int k, const T* p, int q) -> T {
    return (cin[k] && q >= 0 && q < nz) ? p[q * sz + cy[k] * sy + cx[k]] : T(0);
  };
  for (int k = 0; k < 2; ++k)
    for (int d = 0; d < 4; ++d) {  // planes z0-3 .. z0 into slots 1..4 (slot 0 is shifted out)
      const int q = z0 - 3 + d;
      wu[k][d + 1] = col_load(k, uz, q);
      for (int c = 0; c < NC; ++c) wf[k][c][d + 1] = col_load(k, f.in[c], q);
    }
  // the own column's phi_a on planes r-1, r, r+1 (r = i - 1)
  T am[NC], a0[NC], ap[NC];
  for (int c = 0; c < NC; ++c) am[c] = a0[c] = ap[c] = T(0);
  auto in_eno = [&](int z, int y, int x) {
    return z >= 2 && z < nz - 2 && y >= 2 && y < ny - 2 && x >= 2 && x < nx - 2;
  };
  for (int i = z0 - 1; i <= z1; ++i) {
    // (1) plane i of phi, u_x, u_y over the tile and its halo (zero outside the grid)
    for (int p = tid; p < LP; p += S::NT) {
      const int x = x0 - H + p % LX, y = y0 - H + p / LX;
      const bool v = i >= 0 && i < nz && x >= 0 && x < nx && y >= 0 && y < ny;
      const int j = i * sz + y * sy + x;
      sux[p] = v ? ux[j] : T(0);
      suy[p] = v ? uy[j] : T(0);
      for (int c = 0; c < NC; ++c) sphi[c][p] = v ? f.in[c][j] : T(0);
    }
    for (int k = 0; k < 2; ++k) {  // slide the z-windows to planes i-2 .. i+2
      for (int d = 0; d < 4; ++d) wu[k][d] = wu[k][d + 1];
      wu[k][4] = col_load(k, uz, i + 2);
      for (int c = 0; c < NC; ++c) {
        for (int d = 0; d < 4; ++d) wf[k][c][d] = wf[k][c][d + 1];
        wf[k][c][4] = col_load(k, f.in[c], i + 2);
      }
    }
    __syncthreads();
    // (2) phi_a on plane i for the own cell (k = 0) and the ring cell (k = 1); each call site
    // passes a constant k, so the register windows are indexed statically
    T* const sap = &sa[(i + 2) & 1][0][0];
    auto advect = [&](int k, int xr, int yr, T* av) {
      const int l = (yr + H) * LX + xr + H;
      for (int c = 0; c < NC; ++c) {
        T v = wf[k][c][2];
        if (cin[k] && i >= 0 && i < nz && in_eno(i, cy[k], cx[k])) {
          T F[5], U[5], acc = T(0);
          for (int d = 0; d < 5; ++d) {  // x
            F[d] = sphi[c][l + d - 2];
            U[d] = sux[l + d - 2];
          }
          acc = acc + inv_dx * eno_front(F, U);
          acc = acc - inv_dx * eno_back(F, U);
          for (int d = 0; d < 5; ++d) {  // y
            F[d] = sphi[c][l + (d - 2) * LX];
            U[d] = suy[l + (d - 2) * LX];
          }
          acc = acc + inv_dx * eno_front(F, U);
          acc = acc - inv_dx * eno_back(F, U);
          acc = acc + inv_dx * eno_front(wf[k][c], wu[k]);  // z
          acc = acc - inv_dx * eno_back(wf[k][c], wu[k]);
          v = v + acc;
        }
        av[c] = v;
        sap[c * S::RP + (yr + 1) * RX + xr + 1] = v;
      }
    };
    T av[NC];
    advect(0, ox, oy, av);
    for (int c = 0; c < NC; ++c) {
      am[c] = a0[c];
      a0[c] = ap[c];
      ap[c] = av[c];
    }
    if (has_ring) {
      T rv[NC];
      advect(1, ex, ey, rv);
    }
    __syncthreads();
    // (3) diffusion of plane r = i - 1 (phi_a(r) = a0, r -+ 1 in am / ap, in-plane from LDS)
    const int r = i - 1;
    if (r >= z0 && cin[0]) {
      const T* const sar = &sa[(r + 2) & 1][0][0];
      const int g = (oy + 1) * RX + ox + 1;
      const bool inner = r > 0 && r < nz - 1 && cy[0] > 0 && cy[0] < ny - 1 && cx[0] > 0 && cx[0] < nx - 1;
      const int j = r * sz + cy[0] * sy + cx[0];
      for (int c = 0; c < NC; ++c) {
        T v = a0[c];
        if (inner) {
          const T* A = sar + c * S::RP;
          v = v + pD * (ap[c] + am[c] + A[g + RX] + A[g - RX] + A[g + 1] + A[g - 1] - T(6) * a0[c]);
        }
        f.out[c][j] = v;
      }
    }
  }
}

template <typename T, int NC>
int transport_launch(void* const field[3], void* const scratch[3], const void* const vel[3],
                     int64_t nz, int64_t ny, int64_t nx, double dt_by_dx, double nu_dt_by_dx2,
                     hipStream_t s) {
  using S = TransportTile<T>;
  Comps<T, NC> f;
  for (int c = 0; c < NC; ++c) {
    f.in[c] = static_cast<const T*>(field[c]);
    f.out[c] = static_cast<T*>(scratch[c]);
  }
  const int tiles_x = ceil_div(nx, S::TX), tiles_y = ceil_div(ny, S::TY);
  int chunk = kTransportChunk;
  while (chunk > 2 && int64_t(tiles_x) * tiles_y * ceil_div(nz, chunk) < 1024) chunk /= 2;
  const int64_t items = int64_t(tiles_x) * tiles_y * ceil_div(nz, chunk);
  // the reference forms inv_dx = -real_t(dt / dx) and nu dt / dx^2 in real_t
  const T inv_dx = -T(dt_by_dx), pD = T(nu_dt_by_dx2);
  k_transport<T, NC><<<unsigned(items), S::NT, 0, s>>>(
      f, static_cast<const T*>(vel[0]), static_cast<const T*>(vel[1]), static_cast<const T*>(vel[2]),
      int(nz), int(ny), int(nx), inv_dx, pD, chunk, tiles_x, tiles_y, items % 8 == 0 ? 1 : 0);
  SOPHT_CHECK_LAUNCH();
  const size_t bytes = size_t(nz * ny * nx) * sizeof(T);
  for (int c = 0; c < NC; ++c)
    SOPHT_CHECK_HIP(hipMemcpyAsync(field[c], scratch[c], bytes, hipMemcpyDeviceToDevice, s));
  return SOPHT_OK;
}

// max over cells of |u0| + |u1| + |u2| (numpy's axis-0 sum order); optionally stores the sum
template <typename T>
__global__ void __launch_bounds__(256) k_abs_sum_max(const T* __restrict__ u0, const T* __restrict__ u1,
                                                     const T* __restrict__ u2, int64_t n,
                                                     T* __restrict__ mag, T* __restrict__ partial) {
  __shared__ T red[256];
  T m = T(0);
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const T v = (fabs(u0[i]) + fabs(u1[i])) + fabs(u2[i]);
    if (mag) mag[i] = v;
    m = v > m ? v : m;
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (int(threadIdx.x) < k) red[threadIdx.x] = red[threadIdx.x + k] > red[threadIdx.x] ? red[threadIdx.x + k] : red[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

template <typename T>
__global__ void __launch_bounds__(256) k_partial_max(const T* __restrict__ partial, int nb, T* __restrict__ out) {
  __shared__ T red[256];
  T m = T(0);
  for (int i = threadIdx.x; i < nb; i += 256) m = partial[i] > m ? partial[i] : m;
  red[threadIdx.x] = m;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (int(threadIdx.x) < k) red[threadIdx.x] = red[threadIdx.x + k] > red[threadIdx.x] ? red[threadIdx.x + k] : red[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

template <typename T>
int abs_sum_max(const void* const vel[3], int64_t n, void* mag, void* partial, int64_t cap, void* out,
                hipStream_t s) {
  const int nb = int(std::min<int64_t>(cap, (n + 255) / 256));
  k_abs_sum_max<T><<<nb, 256, 0, s>>>(static_cast<const T*>(vel[0]), static_cast<const T*>(vel[1]),
                                      static_cast<const T*>(vel[2]), n, static_cast<T*>(mag),
                                      static_cast<T*>(partial));
  SOPHT_CHECK_LAUNCH();
  k_partial_max<T><<<1, 256, 0, s>>>(static_cast<const T*>(partial), nb, static_cast<T*>(out));
  SOPHT_CHECK_LAUNCH();
  return SOPHT_OK;
}

}  // namespace
}  // namespace sopht

using namespace sopht;

extern "C" {

int sopht_passive_transport_step_3d(int dtype, int ncomp, void* const field[3], void* const scratch[3],
                                    const void* const vel[3], int64_t nz, int64_t ny, int64_t nx,
                                    double dt_by_dx, double nu_dt_by_dx2, void* stream) {
  SOPHT_REQUIRE(dtype == SOPHT_F32 || dtype == SOPHT_F64, "dtype must be SOPHT_F32 or SOPHT_F64");
  SOPHT_REQUIRE(ncomp == 1 || ncomp == 3, "ncomp must be 1 (scalar) or 3 (vector)");
  SOPHT_REQUIRE(nz >= 1 && ny >= 1 && nx >= 1, "grid sizes must be positive");
  SOPHT_REQUIRE(nz * ny * nx < (int64_t(1) << 31), "grid too large (nz ny nx must be below 2^31)");
  SOPHT_REQUIRE(field && scratch && vel && vel[0] && vel[1] && vel[2], "null pointer");
  for (int c = 0; c < ncomp; ++c) {
    SOPHT_REQUIRE(field[c] && scratch[c], "null component pointer");
    SOPHT_REQUIRE(field[c] != scratch[c], "the scratch field must not alias the field");
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == SOPHT_F32)
    return ncomp == 1 ? transport_launch<float, 1>(field, scratch, vel, nz, ny, nx, dt_by_dx, nu_dt_by_dx2, s)
                      : transport_launch<float, 3>(field, scratch, vel, nz, ny, nx, dt_by_dx, nu_dt_by_dx2, s);
  return ncomp == 1 ? transport_launch<double, 1>(field, scratch, vel, nz, ny, nx, dt_by_dx, nu_dt_by_dx2, s)
                    : transport_launch<double, 3>(field, scratch, vel, nz, ny, nx, dt_by_dx, nu_dt_by_dx2, s);
}

int sopht_velocity_abs_sum_max_3d(int dtype, const void* const vel[3], int64_t n, void* magnitude,
                                  void* partial, int64_t partial_cap, void* out, void* stream) {
  SOPHT_REQUIRE(dtype == SOPHT_F32 || dtype == SOPHT_F64, "dtype must be SOPHT_F32 or SOPHT_F64");
  SOPHT_REQUIRE(vel && vel[0] && vel[1] && vel[2] && partial && out, "null pointer");
  SOPHT_REQUIRE(n > 0 && partial_cap > 0, "sizes must be positive");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return dtype == SOPHT_F32 ? abs_sum_max<float>(vel, n, magnitude, partial, partial_cap, out, s)
                            : abs_sum_max<double>(vel, n, magnitude, partial, partial_cap, out, s);
}

}  // extern "C"
