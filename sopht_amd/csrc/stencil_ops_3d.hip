// Standalone drop-in grid stencils (the gen_*_pyst_kernel_3d family) for gfx950.
//
// These keep the reference's per-op semantics exactly (which cells each op writes, border
// resets, strided views) and run one thread per written cell with x-fastest lanes, so
// every neighbour access along x is coalesced and y/z neighbours come from L2. The flow
// step itself does not use them: it runs the fused sweeps in flow_step_3d.hip.
#include <cstdio>
#include <cstring>
#include <mutex>

#include "common.h"

namespace sopht {

static thread_local char g_last_error[512] = "";

void set_last_error(const char* msg, const char* file, int line) {
  snprintf(g_last_error, sizeof(g_last_error), "%s (%s:%d)", msg, file, line);
}

namespace {

// Visit every cell of the box [z0, z0+bz) x [y0, y0+by) x [x0, x0+bx).
template <typename Fn>
__global__ void __launch_bounds__(256) box_kernel(int64_t z0, int64_t y0, int64_t x0, int64_t bz,
                                                  int64_t by, int64_t bx, Fn fn) {
  const int64_t total = bz * by * bx;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < total;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t x = i % bx;
    const int64_t t = i / bx;
    const int64_t y = t % by;
    const int64_t z = t / by;
    fn(z0 + z, y0 + y, x0 + x);
  }
}

template <typename Fn>
int launch_box(int64_t z0, int64_t y0, int64_t x0, int64_t bz, int64_t by, int64_t bx, Fn fn,
               hipStream_t s) {
  const int64_t total = bz * by * bx;
  if (total <= 0) return SOPHT_OK;
  int blocks = ceil_div(total, 256);
  if (blocks > 65536) blocks = 65536;
  box_kernel<<<blocks, 256, 0, s>>>(z0, y0, x0, bz, by, bx, fn);
  SOPHT_CHECK_LAUNCH();
  return SOPHT_OK;
}

// The six width-w boundary slabs of an (nz, ny, nx) box, written in the reference's order
// (z front/back, y front/back, x front/back; elementwise_ops_3d.py:228-233). Overlaps are
// written more than once with the same value.
template <typename Fn>
int launch_boundary_slabs(int64_t nz, int64_t ny, int64_t nx, int64_t w, Fn fn, hipStream_t s) {
  int rc;
  const int64_t wz = w < nz ? w : nz, wy = w < ny ? w : ny, wx = w < nx ? w : nx;
  if ((rc = launch_box(0, 0, 0, wz, ny, nx, fn, s))) return rc;
  if ((rc = launch_box(nz - wz, 0, 0, wz, ny, nx, fn, s))) return rc;
  if ((rc = launch_box(0, 0, 0, nz, wy, nx, fn, s))) return rc;
  if ((rc = launch_box(0, ny - wy, 0, nz, wy, nx, fn, s))) return rc;
  if ((rc = launch_box(0, 0, 0, nz, ny, wx, fn, s))) return rc;
  return launch_box(0, 0, nx - wx, nz, ny, wx, fn, s);
}

// The surface cells of the box [z0, z0 + bz) x [y0, y0 + by) x [x0, x0 + bx) (bz, by, bx >= 2),
// one thread each: the two z faces, then the two y faces between them, then the two x faces.
template <typename Fn>
__global__ void __launch_bounds__(256) shell_kernel(int64_t z0, int64_t y0, int64_t x0, int64_t bz,
                                                    int64_t by, int64_t bx, Fn fn) {
  const int64_t nzf = by * bx, nyf = (bz - 2) * bx, nxf = (bz - 2) * (by - 2);
  const int64_t total = 2 * (nzf + nyf + nxf);
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < total;
       i += int64_t(gridDim.x) * blockDim.x) {
    int64_t z, y, x, r = i;
    if (r < 2 * nzf) {
      z = z0 + (r < nzf ? 0 : bz - 1);
      r %= nzf;
      y = y0 + r / bx;
      x = x0 + r % bx;
    } else if ((r -= 2 * nzf) < 2 * nyf) {
      y = y0 + (r < nyf ? 0 : by - 1);
      r %= nyf;
      z = z0 + 1 + r / bx;
      x = x0 + r % bx;
    } else {
      r -= 2 * nyf;
      x = x0 + (r < nxf ? 0 : bx - 1);
      r %= nxf;
      z = z0 + 1 + r / (by - 2);
      y = y0 + 1 + r % (by - 2);
    }
    fn(z, y, x);
  }
}

template <typename Fn>
int launch_shell(int64_t z0, int64_t y0, int64_t x0, int64_t bz, int64_t by, int64_t bx, Fn fn,
                 hipStream_t s) {
  if (bz < 2 || by < 2 || bx < 2) return launch_box(z0, y0, x0, bz, by, bx, fn, s);
  const int64_t total = 2 * (by * bx + (bz - 2) * bx + (bz - 2) * (by - 2));
  int blocks = ceil_div(total, 256);
  if (blocks > 65536) blocks = 65536;
  shell_kernel<<<blocks, 256, 0, s>>>(z0, y0, x0, bz, by, bx, fn);
  SOPHT_CHECK_LAUNCH();
  return SOPHT_OK;
}

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

// ------------------------------------------------------------------ curl-type operators
template <typename T>
int curl_impl(void* const out[3], const int64_t* os, const void* const in[3], const int64_t* is,
              int64_t nz, int64_t ny, int64_t nx, T p, int reset, bool accumulate,
              hipStream_t s) {
  View3<T> o0 = make_view<T>(out[0], os), o1 = make_view<T>(out[1], os),
           o2 = make_view<T>(out[2], os);
  View3<T> fx = make_view<T>(in[0], is), fy = make_view<T>(in[1], is), fz = make_view<T>(in[2], is);
  int rc = SOPHT_OK;
  if (nz > 2 && ny > 2 && nx > 2) {
    rc = launch_box(
        1, 1, 1, nz - 2, ny - 2, nx - 2,
        [=] __device__(int64_t z, int64_t y, int64_t x) {
          // curl_x = df_z/dy - df_y/dz, curl_y = df_x/dz - df_z/dx, curl_z = df_y/dx - df_x/dy
          const T cx = p * (fz(z, y + 1, x) - fz(z, y - 1, x) - fy(z + 1, y, x) + fy(z - 1, y, x));
          const T cy = p * (fx(z + 1, y, x) - fx(z - 1, y, x) - fz(z, y, x + 1) + fz(z, y, x - 1));
          const T cz = p * (fy(z, y, x + 1) - fy(z, y, x - 1) - fx(z, y + 1, x) + fx(z, y - 1, x));
          if (accumulate) {
            o0(z, y, x) = o0(z, y, x) + cx;
            o1(z, y, x) = o1(z, y, x) + cy;
            o2(z, y, x) = o2(z, y, x) + cz;
          } else {
            o0(z, y, x) = cx;
            o1(z, y, x) = cy;
            o2(z, y, x) = cz;
          }
        },
        s);
    if (rc) return rc;
  }
  if (reset) {
    rc = launch_boundary_slabs(
        nz, ny, nx, 1,
        [=] __device__(int64_t z, int64_t y, int64_t x) {
          o0(z, y, x) = T(0);
          o1(z, y, x) = T(0);
          o2(z, y, x) = T(0);
        },
        s);
  }
  return rc;
}

template <typename T>
int update_penalised_impl(void* const w[3], const int64_t* ws, const void* const up[3],
                          const int64_t* ups, const void* const u[3], const int64_t* us,
                          int64_t nz, int64_t ny, int64_t nx, T p, hipStream_t s) {
  View3<T> w0 = make_view<T>(w[0], ws), w1 = make_view<T>(w[1], ws), w2 = make_view<T>(w[2], ws);
  View3<T> px = make_view<T>(up[0], ups), py = make_view<T>(up[1], ups), pz = make_view<T>(up[2], ups);
  View3<T> vx = make_view<T>(u[0], us), vy = make_view<T>(u[1], us), vz = make_view<T>(u[2], us);
  if (nz <= 2 || ny <= 2 || nx <= 2) return SOPHT_OK;
  return launch_box(
      1, 1, 1, nz - 2, ny - 2, nx - 2,
      [=] __device__(int64_t z, int64_t y, int64_t x) {
        // update_vorticity_from_velocity_forcing_3d.py:175-184 (and the y, z analogues)
        w0(z, y, x) = w0(z, y, x) + p * (pz(z, y + 1, x) - vz(z, y + 1, x) - pz(z, y - 1, x) +
                                         vz(z, y - 1, x) - py(z + 1, y, x) + vy(z + 1, y, x) +
                                         py(z - 1, y, x) - vy(z - 1, y, x));
        w1(z, y, x) = w1(z, y, x) + p * (px(z + 1, y, x) - vx(z + 1, y, x) - px(z - 1, y, x) +
                                         vx(z - 1, y, x) - pz(z, y, x + 1) + vz(z, y, x + 1) +
                                         pz(z, y, x - 1) - vz(z, y, x - 1));
        w2(z, y, x) = w2(z, y, x) + p * (py(z, y, x + 1) - vy(z, y, x + 1) - py(z, y, x - 1) +
                                         vy(z, y, x - 1) - px(z, y + 1, x) + vx(z, y + 1, x) +
                                         px(z, y - 1, x) - vx(z, y - 1, x));
      },
      s);
}

template <typename T>
int cross_impl(void* const r[3], const int64_t* rs, const void* const a[3], const int64_t* as,
               const void* const b[3], const int64_t* bs, int64_t nz, int64_t ny, int64_t nx,
               hipStream_t s) {
  View3<T> r0 = make_view<T>(r[0], rs), r1 = make_view<T>(r[1], rs), r2 = make_view<T>(r[2], rs);
  View3<T> a0 = make_view<T>(a[0], as), a1 = make_view<T>(a[1], as), a2 = make_view<T>(a[2], as);
  View3<T> b0 = make_view<T>(b[0], bs), b1 = make_view<T>(b[1], bs), b2 = make_view<T>(b[2], bs);
  return launch_box(
      0, 0, 0, nz, ny, nx,
      [=] __device__(int64_t z, int64_t y, int64_t x) {
        const T ax = a0(z, y, x), ay = a1(z, y, x), az = a2(z, y, x);
        const T bx = b0(z, y, x), by = b1(z, y, x), bz = b2(z, y, x);
        r0(z, y, x) = ay * bz - by * az;
        r1(z, y, x) = az * bx - bz * ax;
        r2(z, y, x) = ax * by - bx * ay;
      },
      s);
}

template <typename T>
int diffusion_flux_impl(void* flux, const int64_t* fs, const void* field, const int64_t* is,
                        int64_t nz, int64_t ny, int64_t nx, T p, int reset, hipStream_t s) {
  View3<T> o = make_view<T>(flux, fs), f = make_view<T>(field, is);
  int rc = SOPHT_OK;
  if (nz > 2 && ny > 2 && nx > 2) {
    rc = launch_box(
        1, 1, 1, nz - 2, ny - 2, nx - 2,
        [=] __device__(int64_t z, int64_t y, int64_t x) {
          o(z, y, x) = p * (f(z + 1, y, x) + f(z - 1, y, x) + f(z, y + 1, x) + f(z, y - 1, x) +
                            f(z, y, x + 1) + f(z, y, x - 1) - T(6) * f(z, y, x));
        },
        s);
    if (rc) return rc;
  }
  if (reset)
    rc = launch_boundary_slabs(
        nz, ny, nx, 1, [=] __device__(int64_t z, int64_t y, int64_t x) { o(z, y, x) = T(0); }, s);
  return rc;
}

template <typename T>
int divergence_impl(void* div, const int64_t* ds, const void* const in[3], const int64_t* is,
                    int64_t nz, int64_t ny, int64_t nx, T inv_dx, int reset, hipStream_t s) {
  View3<T> o = make_view<T>(div, ds);
  View3<T> fx = make_view<T>(in[0], is), fy = make_view<T>(in[1], is), fz = make_view<T>(in[2], is);
  int rc = SOPHT_OK;
  if (nz > 2 && ny > 2 && nx > 2) {
    rc = launch_box(
        1, 1, 1, nz - 2, ny - 2, nx - 2,
        [=] __device__(int64_t z, int64_t y, int64_t x) {
          o(z, y, x) = T(0.5) * inv_dx *
                       (fx(z, y, x + 1) - fx(z, y, x - 1) + fy(z, y + 1, x) - fy(z, y - 1, x) +
                        fz(z + 1, y, x) - fz(z - 1, y, x));
        },
        s);
    if (rc) return rc;
  }
  if (reset)
    rc = launch_boundary_slabs(
        nz, ny, nx, 1, [=] __device__(int64_t z, int64_t y, int64_t x) { o(z, y, x) = T(0); }, s);
  return rc;
}

// Penalisation, race-free in place: every border cell t takes the ORIGINAL value of its
// clamped source c(t) (clamp of each index to [w-1, n-w]) times the three slab factors
// (penalise_field_boundary_3d.py:191-208 collapsed to one pass). Targets are grouped by
// source: one thread per source cell on the surface of the core box [w-1, n-w]^3 reads
// its value once and writes its whole group (itself last), so no thread reads a cell
// another thread writes.
template <typename T>
struct SineTables {
  T f[3][8];  // front factors per axis (x, y, z), index i < w
  T b[3][8];  // back factors per axis, index n-w+i
};

template <typename T>
int penalise_impl(void* field, const int64_t* fs, int64_t nz, int64_t ny, int64_t nx, int w,
                  const double* tables, hipStream_t s) {
  if (w == 0) return SOPHT_OK;
  SOPHT_REQUIRE(w > 0 && w <= 8, "penalty width must be in [0, 8]");
  SOPHT_REQUIRE(2 * w <= nz && 2 * w <= ny && 2 * w <= nx, "grid too small for penalty width");
  SineTables<T> st;
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < w; ++i) {
      st.f[a][i] = T(tables[a * 2 * w + i]);
      st.b[a][i] = T(tables[a * 2 * w + w + i]);
    }
  View3<T> f = make_view<T>(field, fs);
  const int64_t cz0 = w - 1, cz1 = nz - w, cy0 = w - 1, cy1 = ny - w, cx0 = w - 1, cx1 = nx - w;
  const int64_t bz = cz1 - cz0 + 1, by = cy1 - cy0 + 1, bx = cx1 - cx0 + 1;
  // only the surface of the core box [w-1, n-w]^3 has work: one thread per surface cell
  return launch_shell(
      cz0, cy0, cx0, bz, by, bx,
      [=] __device__(int64_t z, int64_t y, int64_t x) {
        const bool onz = (z == cz0) || (z == cz1), ony = (y == cy0) || (y == cy1),
                   onx = (x == cx0) || (x == cx1);
        if (!(onz || ony || onx)) return;
        const T v = f(z, y, x);
        // target ranges and factors per axis: a source on the low face fans out to [0, w),
        // on the high face to [n-w, n), else only itself (factor 1).
        auto range = [&](int64_t c, int64_t c0, int64_t c1, int64_t n, int64_t& lo,
                         int64_t& hi) {
          if (c == c0) { lo = 0; hi = w; }
          else if (c == c1) { lo = n - w; hi = n; }
          else { lo = c; hi = c + 1; }
          if (c0 == c1) { lo = 0; hi = n; }  // degenerate core (n == 2w - 1 is impossible here)
        };
        auto factor = [&](int a, int64_t i, int64_t n) -> T {
          if (i < w) return st.f[a][i];
          if (i >= n - w) return st.b[a][i - (n - w)];
          return T(1);
        };
        int64_t zl, zh, yl, yh, xl, xh;
        range(z, cz0, cz1, nz, zl, zh);
        range(y, cy0, cy1, ny, yl, yh);
        range(x, cx0, cx1, nx, xl, xh);
        for (int64_t tz = zl; tz < zh; ++tz) {
          const T sz = factor(2, tz, nz);
          for (int64_t ty = yl; ty < yh; ++ty) {
            const T sy = factor(1, ty, ny);
            for (int64_t tx = xl; tx < xh; ++tx) {
              if (tz == z && ty == y && tx == x) continue;
              f(tz, ty, tx) = ((v * factor(0, tx, nx)) * sy) * sz;
            }
          }
        }
        f(z, y, x) = ((v * factor(0, x, nx)) * factor(1, y, ny)) * factor(2, z, nz);
      },
      s);
}

// Strides of one (z, y, x) field, with the neighbour step along each axis (0 = x, 1 = y, 2 = z).
struct Strides3 {
  int64_t z, y, x;
  __device__ __forceinline__ int64_t at(int64_t zz, int64_t yy, int64_t xx) const {
    return zz * z + yy * y + xx * x;
  }
  __device__ __forceinline__ int64_t step(int a) const { return a == 0 ? x : (a == 1 ? y : z); }
};

inline Strides3 strides_of(const int64_t* s) { return Strides3{s[0], s[1], s[2]}; }

// ENO3 launches: 64 x 4 threads per workgroup, one z plane per grid slice, so each thread
// finds its cell without division and forms every neighbour address from one base offset.
constexpr int kEnoTx = 64, kEnoTy = 4;

inline bool eno3_grid_fits(int64_t bz, int64_t by) {
  return bz > 0 && bz <= 65535 && ceil_div(by, kEnoTy) <= 65535;
}

// ENO3 conservative advection flux (advection_flux_3d.py:12-233): the six sub-stencils
// accumulate in the reference order on [2:-2]^3; one thread computes all six for its cell.
template <typename T>
__device__ __forceinline__ T eno3_cell(const T* f, int64_t fo, Strides3 fs, const T* const u[3],
                                       int64_t uo, Strides3 us, T acc, T inv_dx) {
  const T c13 = T(1.0 / 3.0), c56 = T(5.0 / 6.0), c16 = T(1.0 / 6.0);
  for (int a = 0; a < 3; ++a) {
    const T* ua = u[a];
    const int64_t sf = fs.step(a), su = us.step(a);
    auto F = [&](int k) -> T { return f[fo + k * sf]; };
    auto U = [&](int k) -> T { return ua[uo + k * su]; };
    T front = (U(0) > -U(1)) ? (c13 * F(1) * U(1) + c56 * F(0) * U(0) - c16 * F(-1) * U(-1))
                             : (c13 * F(0) * U(0) + c56 * F(1) * U(1) - c16 * F(2) * U(2));
    acc = acc + inv_dx * front;
    T back = (U(0) > -U(-1)) ? (c13 * F(0) * U(0) + c56 * F(-1) * U(-1) - c16 * F(-2) * U(-2))
                             : (c13 * F(-1) * U(-1) + c56 * F(0) * U(0) - c16 * F(1) * U(1));
    acc = acc - inv_dx * back;
  }
  return acc;
}

// The same flux marching z: a thread owns one (y, x) column of kEnoZc planes and keeps f
// and u_z at z-2..z+2 in registers, so the z terms cost two loads per cell instead of nine.
// The x and y terms are read as in eno3_cell, and every expression is the same, so the
// results are the same bits.
constexpr int kEnoZc = 16;

template <typename T>
__global__ void __launch_bounds__(kEnoTx* kEnoTy)
    k_eno3_zmarch(T* flux, Strides3 os, const T* f, Strides3 fs, const T* u0, const T* u1,
                  const T* u2, Strides3 us, int64_t nz, int64_t ny, int64_t nx, T inv_dx) {
  const int64_t x = 2 + int64_t(blockIdx.x) * kEnoTx + threadIdx.x;
  const int64_t y = 2 + int64_t(blockIdx.y) * kEnoTy + threadIdx.y;
  const int64_t z0 = 2 + int64_t(blockIdx.z) * kEnoZc;
  if (x >= nx - 2 || y >= ny - 2 || z0 >= nz - 2) return;
  const int64_t z1 = z0 + kEnoZc < nz - 2 ? z0 + kEnoZc : nz - 2;
  const T c13 = T(1.0 / 3.0), c56 = T(5.0 / 6.0), c16 = T(1.0 / 6.0);
  T fw[5], uw[5];
  {
    const int64_t fo = fs.at(z0, y, x), uo = us.at(z0, y, x);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      fw[k] = f[fo + (k - 2) * fs.z];
      uw[k] = u2[uo + (k - 2) * us.z];
    }
  }
  for (int64_t z = z0; z < z1; ++z) {
    const int64_t fo = fs.at(z, y, x), uo = us.at(z, y, x), oo = os.at(z, y, x);
    T acc = flux[oo];
    for (int a = 0; a < 2; ++a) {
      const T* ua = a == 0 ? u0 : u1;
      const int64_t sf = a == 0 ? fs.x : fs.y, su = a == 0 ? us.x : us.y;
      auto F = [&](int k) -> T { return f[fo + k * sf]; };
      auto U = [&](int k) -> T { return ua[uo + k * su]; };
      T front = (U(0) > -U(1)) ? (c13 * F(1) * U(1) + c56 * F(0) * U(0) - c16 * F(-1) * U(-1))
                               : (c13 * F(0) * U(0) + c56 * F(1) * U(1) - c16 * F(2) * U(2));
      acc = acc + inv_dx * front;
      T back = (U(0) > -U(-1)) ? (c13 * F(0) * U(0) + c56 * F(-1) * U(-1) - c16 * F(-2) * U(-2))
                               : (c13 * F(-1) * U(-1) + c56 * F(0) * U(0) - c16 * F(1) * U(1));
      acc = acc - inv_dx * back;
    }
    {
      auto F = [&](int k) -> T { return fw[k + 2]; };
      auto U = [&](int k) -> T { return uw[k + 2]; };
      T front = (U(0) > -U(1)) ? (c13 * F(1) * U(1) + c56 * F(0) * U(0) - c16 * F(-1) * U(-1))
                               : (c13 * F(0) * U(0) + c56 * F(1) * U(1) - c16 * F(2) * U(2));
      acc = acc + inv_dx * front;
      T back = (U(0) > -U(-1)) ? (c13 * F(0) * U(0) + c56 * F(-1) * U(-1) - c16 * F(-2) * U(-2))
                               : (c13 * F(-1) * U(-1) + c56 * F(0) * U(0) - c16 * F(1) * U(1));
      acc = acc - inv_dx * back;
    }
    flux[oo] = acc;
    if (z + 1 < z1) {
      // plane z + 3 exists: z + 1 < z1 <= nz - 2
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        fw[k] = fw[k + 1];
        uw[k] = uw[k + 1];
      }
      fw[4] = f[fo + 3 * fs.z];
      uw[4] = u2[uo + 3 * us.z];
    }
  }
}

template <typename T>
int eno3_impl(void* flux, const int64_t* fs, const void* field, const int64_t* is,
              const void* const vel[3], const int64_t* vs, int64_t nz, int64_t ny, int64_t nx,
              T inv_dx, hipStream_t s) {
  if (nz <= 4 || ny <= 4 || nx <= 4) return SOPHT_OK;
  const Strides3 os3 = strides_of(fs), is3 = strides_of(is), vs3 = strides_of(vs);
  T* o = static_cast<T*>(flux);
  const T* f = static_cast<const T*>(field);
  const T* u0 = static_cast<const T*>(vel[0]);
  const T* u1 = static_cast<const T*>(vel[1]);
  const T* u2 = static_cast<const T*>(vel[2]);
  if (eno3_grid_fits(ceil_div(nz - 4, kEnoZc), ny - 4)) {
    dim3 grid(unsigned(ceil_div(nx - 4, kEnoTx)), unsigned(ceil_div(ny - 4, kEnoTy)),
              unsigned(ceil_div(nz - 4, kEnoZc)));
    k_eno3_zmarch<T><<<grid, dim3(kEnoTx, kEnoTy), 0, s>>>(o, os3, f, is3, u0, u1, u2, vs3, nz, ny,
                                                         nx, inv_dx);
    SOPHT_CHECK_LAUNCH();
    return SOPHT_OK;
  }
  return launch_box(
      2, 2, 2, nz - 4, ny - 4, nx - 4,
      [=] __device__(int64_t z, int64_t y, int64_t x) {
        const T* const u[3] = {u0, u1, u2};
        const int64_t oo = os3.at(z, y, x);
        o[oo] = eno3_cell<T>(f, is3.at(z, y, x), is3, u, vs3.at(z, y, x), vs3, o[oo], inv_dx);
      },
      s);
}

// Weight of g[c + o] in the ENO3 flux at c along one axis (front face minus back face), for
// the face switches up_front = u[c] > -u[c+1] and up_back = u[c] > -u[c-1] of eno3_cell.
template <typename T>
__device__ __forceinline__ T eno3_weight(bool up_front, bool up_back, int o) {
  const T c13 = T(1.0 / 3.0), c56 = T(5.0 / 6.0), c16 = T(1.0 / 6.0);
  T w = T(0);
  if (up_front) {
    w = o == 1 ? c13 : (o == 0 ? c56 : (o == -1 ? -c16 : T(0)));
  } else {
    w = o == 0 ? c13 : (o == 1 ? c56 : (o == 2 ? -c16 : T(0)));
  }
  if (up_back) {
    w -= o == 0 ? c13 : (o == -1 ? c56 : (o == -2 ? -c16 : T(0)));
  } else {
    w -= o == -1 ? c13 : (o == 0 ? c56 : (o == 1 ? -c16 : T(0)));
  }
  return w;
}

// Transpose of the ENO3 flux with the switches frozen at the velocity, for one cell m:
// gbar_axis(m) = inv_dx * sum_o W(m - o, o) a(m - o) over m - o in [2:-2]^3, then
// field_adj += sum_axis gbar_axis * u_axis and vel_adj[axis] += gbar_axis * field. One thread
// gathers for its own cell (no atomics, the same bits on every run). Either output may be
// absent (null pointer).
template <typename T>
__device__ __forceinline__ void eno3_adjoint_cell(int64_t z, int64_t y, int64_t x, int64_t nz,
                                                  int64_t ny, int64_t nx, T* fa, Strides3 fas,
                                                  T* const va[3], Strides3 vas, const T* a,
                                                  Strides3 as, const T* f, Strides3 is,
                                                  const T* const u[3], Strides3 us, T inv_dx) {
  const int64_t m[3] = {x, y, z}, n[3] = {nx, ny, nz};
  const bool in_j[3] = {x >= 2 && x < nx - 2, y >= 2 && y < ny - 2, z >= 2 && z < nz - 2};
  const int64_t ao = as.at(z, y, x), uo = us.at(z, y, x);
  T gbar[3] = {T(0), T(0), T(0)};
  for (int ax = 0; ax < 3; ++ax) {
    // the flux cells that read this one along ax share its other two coordinates
    if (!in_j[(ax + 1) % 3] || !in_j[(ax + 2) % 3]) continue;
    const T* ua = u[ax];
    const int64_t su = us.step(ax), sa = as.step(ax);
    auto U = [&](int64_t k) -> T { return ua[uo + k * su]; };
    T g = T(0);
    for (int o = -2; o <= 2; ++o) {
      const int64_t c = m[ax] - o;
      if (c < 2 || c >= n[ax] - 2) continue;
      const T uc = U(-o);
      g += eno3_weight<T>(uc > -U(1 - o), uc > -U(-1 - o), o) * a[ao - o * sa];
    }
    gbar[ax] = inv_dx * g;
  }
  if (fa != nullptr) {
    fa[fas.at(z, y, x)] += gbar[0] * u[0][uo] + gbar[1] * u[1][uo] + gbar[2] * u[2][uo];
  }
  if (va[0] != nullptr) {
    const T fv = f[is.at(z, y, x)];
    const int64_t vo = vas.at(z, y, x);
    va[0][vo] += gbar[0] * fv;
    va[1][vo] += gbar[1] * fv;
    va[2][vo] += gbar[2] * fv;
  }
}

template <typename T>
__global__ void __launch_bounds__(kEnoTx* kEnoTy)
    k_eno3_adjoint(T* fa, Strides3 fas, T* va0, T* va1, T* va2, Strides3 vas, const T* a,
                   Strides3 as, const T* f, Strides3 is, const T* u0, const T* u1, const T* u2,
                   Strides3 us, int64_t nz, int64_t ny, int64_t nx, T inv_dx) {
  const int64_t x = int64_t(blockIdx.x) * kEnoTx + threadIdx.x;
  const int64_t y = int64_t(blockIdx.y) * kEnoTy + threadIdx.y;
  const int64_t z = blockIdx.z;
  if (x >= nx || y >= ny) return;
  T* const va[3] = {va0, va1, va2};
  const T* const u[3] = {u0, u1, u2};
  eno3_adjoint_cell<T>(z, y, x, nz, ny, nx, fa, fas, va, vas, a, as, f, is, u, us, inv_dx);
}

template <typename T>
int eno3_adjoint_impl(void* field_adj, const int64_t* fas, void* const vel_adj[3],
                      const int64_t* vas, const void* flux_adj, const int64_t* as,
                      const void* field, const int64_t* is, const void* const vel[3],
                      const int64_t* vs, int64_t nz, int64_t ny, int64_t nx, T inv_dx,
                      hipStream_t s) {
  if (nz <= 4 || ny <= 4 || nx <= 4) return SOPHT_OK;
  const bool has_f = field_adj != nullptr, has_v = vel_adj != nullptr;
  if (!has_f && !has_v) return SOPHT_OK;
  const int64_t zero3[3] = {0, 0, 0};
  const Strides3 fas3 = strides_of(has_f ? fas : zero3), vas3 = strides_of(has_v ? vas : zero3);
  const Strides3 as3 = strides_of(as), is3 = strides_of(is), us3 = strides_of(vs);
  T* fa = static_cast<T*>(field_adj);
  T* va0 = has_v ? static_cast<T*>(vel_adj[0]) : nullptr;
  T* va1 = has_v ? static_cast<T*>(vel_adj[1]) : nullptr;
  T* va2 = has_v ? static_cast<T*>(vel_adj[2]) : nullptr;
  const T* a = static_cast<const T*>(flux_adj);
  const T* f = static_cast<const T*>(field);
  const T* u0 = static_cast<const T*>(vel[0]);
  const T* u1 = static_cast<const T*>(vel[1]);
  const T* u2 = static_cast<const T*>(vel[2]);
  if (eno3_grid_fits(nz, ny)) {
    dim3 grid(unsigned(ceil_div(nx, kEnoTx)), unsigned(ceil_div(ny, kEnoTy)), unsigned(nz));
    k_eno3_adjoint<T><<<grid, dim3(kEnoTx, kEnoTy), 0, s>>>(fa, fas3, va0, va1, va2, vas3, a, as3, f,
                                                          is3, u0, u1, u2, us3, nz, ny, nx, inv_dx);
    SOPHT_CHECK_LAUNCH();
    return SOPHT_OK;
  }
  return launch_box(
      0, 0, 0, nz, ny, nx,
      [=] __device__(int64_t z, int64_t y, int64_t x) {
        T* const va[3] = {va0, va1, va2};
        const T* const u[3] = {u0, u1, u2};
        eno3_adjoint_cell<T>(z, y, x, nz, ny, nx, fa, fas3, va, vas3, a, as3, f, is3, u, us3, inv_dx);
      },
      s);
}

template <typename T>
int stretching_impl(void* const flux[3], const int64_t* fs, const void* const w[3],
                    const int64_t* ws, const void* const u[3], const int64_t* us, int64_t nz,
                    int64_t ny, int64_t nx, T p, hipStream_t s) {
  View3<T> o0 = make_view<T>(flux[0], fs), o1 = make_view<T>(flux[1], fs), o2 = make_view<T>(flux[2], fs);
  View3<T> wx = make_view<T>(w[0], ws), wy = make_view<T>(w[1], ws), wz = make_view<T>(w[2], ws);
  View3<T> u0 = make_view<T>(u[0], us), u1 = make_view<T>(u[1], us), u2 = make_view<T>(u[2], us);
  int rc = SOPHT_OK;
  if (nz > 2 && ny > 2 && nx > 2) {
    rc = launch_box(
        1, 1, 1, nz - 2, ny - 2, nx - 2,
        [=] __device__(int64_t z, int64_t y, int64_t x) {
          const T a = wx(z, y, x), b = wy(z, y, x), c = wz(z, y, x);
          auto g = [&](const View3<T>& v) -> T {
            return p * (a * (v(z, y, x + 1) - v(z, y, x - 1)) + b * (v(z, y + 1, x) - v(z, y - 1, x)) +
                        c * (v(z + 1, y, x) - v(z - 1, y, x)));
          };
          o0(z, y, x) = g(u0);
          o1(z, y, x) = g(u1);
          o2(z, y, x) = g(u2);
        },
        s);
    if (rc) return rc;
  }
  return launch_boundary_slabs(
      nz, ny, nx, 1,
      [=] __device__(int64_t z, int64_t y, int64_t x) {
        o0(z, y, x) = T(0);
        o1(z, y, x) = T(0);
        o2(z, y, x) = T(0);
      },
      s);
}


// gen_laplacian_filter_kernel_3d on one scalar field (laplacian_filter_3d.py:13-194), with the
// reference's sequence of kernels: the flux buffer's boundary of width bw is zeroed; each 1D
// pass writes flux = 0.25 (-f[+1] - f[-1] + 2 f) on [1:-1]^3 (ghost width 1 on every axis,
// :60-80); the field buffer becomes a full copy of the flux after every pass.
// type 0 multiplicative (:98-122): order x (x, y, z) passes, then field -= flux.
// type 1 convolution (:124-166): per axis x, y, z: order passes from a copy of the field, then
// field -= flux.
template <typename T>
int laplacian_filter_impl(int order, int type, void* field, const int64_t* fs, int64_t nz,
                          int64_t ny, int64_t nx, void* flux, const int64_t* xs, void* fbuf,
                          const int64_t* bs, int bw, hipStream_t s) {
  View3<T> f = make_view<T>(field, fs), F = make_view<T>(flux, xs), B = make_view<T>(fbuf, bs);
  int rc = launch_boundary_slabs(
      nz, ny, nx, bw, [=] __device__(int64_t z, int64_t y, int64_t x) { F(z, y, x) = T(0); }, s);
  if (rc) return rc;
  auto copy = [&](View3<T> dst, View3<T> src) {
    return launch_box(0, 0, 0, nz, ny, nx,
                      [=] __device__(int64_t z, int64_t y, int64_t x) { dst(z, y, x) = src(z, y, x); }, s);
  };
  auto pass = [&](int axis) {  // F = filter_axis(B) on the interior
    if (nz <= 2 || ny <= 2 || nx <= 2) return int(SOPHT_OK);
    const int64_t dz = axis == 2, dy = axis == 1, dx = axis == 0;  // axis 0 = x
    return launch_box(1, 1, 1, nz - 2, ny - 2, nx - 2,
                      [=] __device__(int64_t z, int64_t y, int64_t x) {
                        F(z, y, x) = T(0.25) * ((-B(z + dz, y + dy, x + dx) - B(z - dz, y - dy, x - dx)) +
                                                T(2) * B(z, y, x));
                      },
                      s);
  };
  auto subtract = [&]() {  // elementwise_saxpby(field, 1.0 * field + (-1.0) * flux)
    return launch_box(0, 0, 0, nz, ny, nx,
                      [=] __device__(int64_t z, int64_t y, int64_t x) {
                        f(z, y, x) = T(1) * f(z, y, x) + T(-1) * F(z, y, x);
                      },
                      s);
  };
  if (type == 0) {
    if ((rc = copy(B, f))) return rc;
    for (int k = 0; k < order; ++k)
      for (int axis = 0; axis < 3; ++axis) {
        if ((rc = pass(axis))) return rc;
        if ((rc = copy(B, F))) return rc;
      }
    return subtract();
  }
  for (int axis = 0; axis < 3; ++axis) {
    if ((rc = copy(B, f))) return rc;
    for (int k = 0; k < order; ++k) {
      if ((rc = pass(axis))) return rc;
      if ((rc = copy(B, F))) return rc;
    }
    if ((rc = subtract())) return rc;
  }
  return SOPHT_OK;
}
}  // namespace

int laplacian_filter_field(int dtype, int order, int type, void* field, const int64_t* fs,
                           int64_t nz, int64_t ny, int64_t nx, void* flux, const int64_t* xs,
                           void* fbuf, const int64_t* bs, int bw, hipStream_t s) {
  if (dtype == SOPHT_F32)
    return laplacian_filter_impl<float>(order, type, field, fs, nz, ny, nx, flux, xs, fbuf, bs, bw, s);
  return laplacian_filter_impl<double>(order, type, field, fs, nz, ny, nx, flux, xs, fbuf, bs, bw, s);
}
}  // namespace sopht

using namespace sopht;

#define DISPATCH(dtype, CALL)                                          \
  do {                                                                 \
    if ((dtype) == SOPHT_F32) {                                        \
      using T = float;                                                 \
      return CALL;                                                     \
    } else if ((dtype) == SOPHT_F64) {                                 \
      using T = double;                                                \
      return CALL;                                                     \
    }                                                                  \
    set_last_error("dtype must be SOPHT_F32 or SOPHT_F64", __FILE__, __LINE__); \
    return SOPHT_ERR_INVALID;                                          \
  } while (0)

#define REQUIRE_GRID(nz, ny, nx) \
  SOPHT_REQUIRE((nz) > 0 && (ny) > 0 && (nx) > 0, "grid sizes must be positive")

extern "C" {

const char* sopht_last_error(void) { return g_last_error; }
int sopht_version(void) { return 1; }

int sopht_curl_3d(int dtype, void* const curl[3], const int64_t* cs, const void* const field[3],
                  const int64_t* fs, int64_t nz, int64_t ny, int64_t nx, double prefactor,
                  int reset, void* stream) {
  REQUIRE_GRID(nz, ny, nx);
  DISPATCH(dtype, curl_impl<T>(curl, cs, field, fs, nz, ny, nx, T(prefactor), reset, false,
                               as_stream(stream)));
}

int sopht_update_vorticity_from_velocity_forcing_3d(int dtype, void* const w[3],
                                                    const int64_t* ws, const void* const f[3],
                                                    const int64_t* fs, int64_t nz, int64_t ny,
                                                    int64_t nx, double prefactor, void* stream) {
  REQUIRE_GRID(nz, ny, nx);
  DISPATCH(dtype, curl_impl<T>(w, ws, f, fs, nz, ny, nx, T(prefactor), 0, true, as_stream(stream)));
}

int sopht_update_vorticity_from_penalised_velocity_3d(int dtype, void* const w[3],
                                                      const int64_t* ws,
                                                      const void* const up[3],
                                                      const int64_t* ups, const void* const u[3],
                                                      const int64_t* us, int64_t nz, int64_t ny,
                                                      int64_t nx, double prefactor,
                                                      void* stream) {
  REQUIRE_GRID(nz, ny, nx);
  DISPATCH(dtype, update_penalised_impl<T>(w, ws, up, ups, u, us, nz, ny, nx, T(prefactor),
                                           as_stream(stream)));
}

int sopht_elementwise_cross_product_3d(int dtype, void* const r[3], const int64_t* rs,
                                       const void* const a[3], const int64_t* as,
                                       const void* const b[3], const int64_t* bs, int64_t nz,
                                       int64_t ny, int64_t nx, void* stream) {
  REQUIRE_GRID(nz, ny, nx);
  DISPATCH(dtype, cross_impl<T>(r, rs, a, as, b, bs, nz, ny, nx, as_stream(stream)));
}

int sopht_diffusion_flux_3d(int dtype, void* flux, const int64_t* fs, const void* field,
                            const int64_t* is, int64_t nz, int64_t ny, int64_t nx,
                            double prefactor, int reset, void* stream) {
  REQUIRE_GRID(nz, ny, nx);
  DISPATCH(dtype, diffusion_flux_impl<T>(flux, fs, field, is, nz, ny, nx, T(prefactor), reset,
                                         as_stream(stream)));
}

int sopht_elementwise_sum_3d(int dtype, void* sum, const int64_t* ss, const void* f1,
                             const int64_t* s1, const void* f2, const int64_t* s2, int64_t nz,
                             int64_t ny, int64_t nx, void* stream);

int sopht_diffusion_timestep_euler_forward_3d(int dtype, void* field, const int64_t* fs,
                                              void* flux, const int64_t* xs, int64_t nz,
                                              int64_t ny, int64_t nx, double p, void* stream) {
  int rc = sopht_diffusion_flux_3d(dtype, flux, xs, field, fs, nz, ny, nx, p, 1, stream);
  if (rc) return rc;
  return sopht_elementwise_sum_3d(dtype, field, fs, field, fs, flux, xs, nz, ny, nx, stream);
}

int sopht_penalise_field_boundary_3d(int dtype, void* field, const int64_t* fs, int64_t nz,
                                     int64_t ny, int64_t nx, int width, const double* tables,
                                     void* stream) {
  REQUIRE_GRID(nz, ny, nx);
  SOPHT_REQUIRE(width >= 0, "Invalid width for boundary zone, must be a non-negative integer");
  DISPATCH(dtype, penalise_impl<T>(field, fs, nz, ny, nx, width, tables, as_stream(stream)));
}

int sopht_divergence_3d(int dtype, void* div, const int64_t* ds, const void* const field[3],
                        const int64_t* fs, int64_t nz, int64_t ny, int64_t nx, double inv_dx,
                        int reset, void* stream) {
  REQUIRE_GRID(nz, ny, nx);
  DISPATCH(dtype, divergence_impl<T>(div, ds, field, fs, nz, ny, nx, T(inv_dx), reset,
                                     as_stream(stream)));
}

int sopht_set_fixed_val_3d(int dtype, void* field, const int64_t* st, int64_t nz, int64_t ny,
                           int64_t nx, double value, void* stream) {
  REQUIRE_GRID(nz, ny, nx);
  DISPATCH(dtype, ([&] {
             View3<T> f = make_view<T>(field, st);
             const T v = T(value);
             return launch_box(0, 0, 0, nz, ny, nx,
                               [=] __device__(int64_t z, int64_t y, int64_t x) { f(z, y, x) = v; },
                               as_stream(stream));
           }()));
}

int sopht_add_fixed_val_3d(int dtype, void* sum, const int64_t* ss, const void* field,
                           const int64_t* fs, int64_t nz, int64_t ny, int64_t nx, double value,
                           void* stream) {
  REQUIRE_GRID(nz, ny, nx);
  DISPATCH(dtype, ([&] {
             View3<T> o = make_view<T>(sum, ss), f = make_view<T>(field, fs);
             const T v = T(value);
             return launch_box(
                 0, 0, 0, nz, ny, nx,
                 [=] __device__(int64_t z, int64_t y, int64_t x) { o(z, y, x) = f(z, y, x) + v; },
                 as_stream(stream));
           }()));
}

int sopht_elementwise_saxpby_3d(int dtype, void* sum, const int64_t* ss, const void* f1,
                                const int64_t* s1, const void* f2, const int64_t* s2, int64_t nz,
                                int64_t ny, int64_t nx, double a, double b, void* stream) {
  REQUIRE_GRID(nz, ny, nx);
  DISPATCH(dtype, ([&] {
             View3<T> o = make_view<T>(sum, ss), x1 = make_view<T>(f1, s1), x2 = make_view<T>(f2, s2);
             const T ta = T(a), tb = T(b);
             return launch_box(
                 0, 0, 0, nz, ny, nx,
                 [=] __device__(int64_t z, int64_t y, int64_t x) {
                   o(z, y, x) = ta * x1(z, y, x) + tb * x2(z, y, x);
                 },
                 as_stream(stream));
           }()));
}

int sopht_elementwise_sum_3d(int dtype, void* sum, const int64_t* ss, const void* f1,
                             const int64_t* s1, const void* f2, const int64_t* s2, int64_t nz,
                             int64_t ny, int64_t nx, void* stream) {
  REQUIRE_GRID(nz, ny, nx);
  DISPATCH(dtype, ([&] {
             View3<T> o = make_view<T>(sum, ss), x1 = make_view<T>(f1, s1), x2 = make_view<T>(f2, s2);
             return launch_box(
                 0, 0, 0, nz, ny, nx,
                 [=] __device__(int64_t z, int64_t y, int64_t x) {
                   o(z, y, x) = x1(z, y, x) + x2(z, y, x);
                 },
                 as_stream(stream));
           }()));
}

int sopht_elementwise_copy_3d(int dtype, void* field, const int64_t* st, const void* rhs,
                              const int64_t* rs, int64_t nz, int64_t ny, int64_t nx,
                              void* stream) {
  REQUIRE_GRID(nz, ny, nx);
  DISPATCH(dtype, ([&] {
             View3<T> o = make_view<T>(field, st), r = make_view<T>(rhs, rs);
             return launch_box(
                 0, 0, 0, nz, ny, nx,
                 [=] __device__(int64_t z, int64_t y, int64_t x) { o(z, y, x) = r(z, y, x); },
                 as_stream(stream));
           }()));
}

int sopht_set_fixed_val_at_boundaries_3d(int dtype, void* field, const int64_t* st, int64_t nz,
                                         int64_t ny, int64_t nx, int width, double value,
                                         void* stream) {
  REQUIRE_GRID(nz, ny, nx);
  SOPHT_REQUIRE(width > 0, "Invalid width for boundary zone, must be a positive integer");
  DISPATCH(dtype, ([&] {
             View3<T> f = make_view<T>(field, st);
             const T v = T(value);
             return launch_boundary_slabs(
                 nz, ny, nx, width,
                 [=] __device__(int64_t z, int64_t y, int64_t x) { f(z, y, x) = v; },
                 as_stream(stream));
           }()));
}

// Complex fields are interleaved (re, im); strides are in complex elements.
int sopht_elementwise_complex_product_3d(int dtype, void* prod, const int64_t* ps,
                                         const void* f1, const int64_t* s1, const void* f2,
                                         const int64_t* s2, int64_t nz, int64_t ny, int64_t nx,
                                         void* stream) {
  REQUIRE_GRID(nz, ny, nx);
  DISPATCH(dtype, ([&] {
             View3<Cplx<T>> o = make_view<Cplx<T>>(prod, ps), a = make_view<Cplx<T>>(f1, s1),
                            b = make_view<Cplx<T>>(f2, s2);
             return launch_box(
                 0, 0, 0, nz, ny, nx,
                 [=] __device__(int64_t z, int64_t y, int64_t x) {
                   const Cplx<T> u = a(z, y, x), v = b(z, y, x);
                   o(z, y, x) = Cplx<T>{u.re * v.re - u.im * v.im, u.re * v.im + u.im * v.re};
                 },
                 as_stream(stream));
           }()));
}

int sopht_advection_flux_conservative_eno3_3d(int dtype, void* flux, const int64_t* fs,
                                              const void* field, const int64_t* is,
                                              const void* const vel[3], const int64_t* vs,
                                              int64_t nz, int64_t ny, int64_t nx, double inv_dx,
                                              void* stream) {
  REQUIRE_GRID(nz, ny, nx);
  DISPATCH(dtype, eno3_impl<T>(flux, fs, field, is, vel, vs, nz, ny, nx, T(inv_dx),
                               as_stream(stream)));
}

int sopht_advection_flux_conservative_eno3_adjoint_3d(
    int dtype, void* field_adj, const int64_t* fas, void* const velocity_adj[3],
    const int64_t* vas, const void* flux_adj, const int64_t* as, const void* field,
    const int64_t* is, const void* const vel[3], const int64_t* vs, int64_t nz, int64_t ny,
    int64_t nx, double inv_dx, void* stream) {
  REQUIRE_GRID(nz, ny, nx);
  SOPHT_REQUIRE(field_adj == nullptr || fas != nullptr, "field_adj needs its strides");
  SOPHT_REQUIRE(velocity_adj == nullptr || (vas != nullptr && velocity_adj[0] != nullptr &&
                                            velocity_adj[1] != nullptr && velocity_adj[2] != nullptr),
                "velocity_adj needs three components and their strides");
  DISPATCH(dtype, eno3_adjoint_impl<T>(field_adj, fas, velocity_adj, vas, flux_adj, as, field, is,
                                       vel, vs, nz, ny, nx, T(inv_dx), as_stream(stream)));
}

int sopht_vorticity_stretching_flux_3d(int dtype, void* const flux[3], const int64_t* fs,
                                       const void* const w[3], const int64_t* ws,
                                       const void* const u[3], const int64_t* us, int64_t nz,
                                       int64_t ny, int64_t nx, double p, void* stream) {
  REQUIRE_GRID(nz, ny, nx);
  DISPATCH(dtype, stretching_impl<T>(flux, fs, w, ws, u, us, nz, ny, nx, T(p), as_stream(stream)));
}

int sopht_laplacian_filter_3d(int dtype, int filter_order, int filter_type, void* field,
                              const int64_t* field_strides, int64_t nz, int64_t ny, int64_t nx,
                              void* filter_flux_buffer, const int64_t* flux_strides,
                              void* field_buffer, const int64_t* buffer_strides,
                              int flux_boundary_width, void* stream) {
  REQUIRE_GRID(nz, ny, nx);
  SOPHT_REQUIRE(filter_order >= 0, "Invalid filter order, must be a non-negative integer");
  SOPHT_REQUIRE(flux_boundary_width > 0,
                "Invalid value for filter flux buffer boundary zone, must be a positive integer");
  SOPHT_REQUIRE(filter_type == 0 || filter_type == 1, "Invalid filter type");
  SOPHT_REQUIRE(dtype == SOPHT_F32 || dtype == SOPHT_F64, "dtype must be SOPHT_F32 or SOPHT_F64");
  return laplacian_filter_field(dtype, filter_order, filter_type, field, field_strides, nz, ny, nx,
                                filter_flux_buffer, flux_strides, field_buffer, buffer_strides,
                                flux_boundary_width, as_stream(stream));
}

}  // extern "C"
