// Reverse-mode adjoint of the coupled 3D flow step (no counterpart in the reference, which
// has no differentiation): the transposes of the step of flow_step_3d.hip
// (navier_stokes_flow_simulators.py:449-498) and of the virtual-boundary forcing
// (VirtualBoundaryForcing.py:149-283 with the cosine kernel of
// EulerianLagrangianGridCommunicator3D.py:387-412), with the time step frozen.
//
// Forward step (M: interior mask of the radius-1 stencils, C: unnormalised central curl,
// L: 7-point Laplacian, P: clamp-and-sine penalisation, G: free-space Green's convolution)
//   w1 = w + pF M C F ; w2 = w1 + pB M C (u x w1) ; w3 = w2 + pD M L w2 ; w' = P w3 ;
//   u' = pC M C G w' + U_inf.
// Adjoint (C^T = C and L^T = L on every cell with zero outside the grid; G^T = G):
//   psib = pC C(M ub') ; t = P^T(wb' + G psib) ; g2 = t + pD L(M t) ; Bb = pB C(M g2) ;
//   ub = w1 x Bb ; g1 = g2 + Bb x u ; wb = g1 ; Fb = pF C(M g1) (at the body's support only).
// w1 comes from the recorded trajectory (sopht_ns3d_trajectory_step writes it).
#include <algorithm>
#include <cmath>

#include "common.h"
#include "ib_kernels.h"
#include "ns3d_plan.h"
#include "sweeps_3d.h"

namespace sopht {
namespace {

// ------------------------------------------------------------------ masked row access

// Row (z, y) of f at x0 .. x0 + VEC - 1 with the cells outside the interior set to zero.
template <typename T, int VEC>
__device__ __forceinline__ VecT<T, VEC> mrow(const T* __restrict__ f, int z, int y, int x0,
                                             int nz, int ny, int nx, int sz, int sy, bool act) {
  VecT<T, VEC> v = vzero<T, VEC>();
  if (act && z >= 1 && z <= nz - 2 && y >= 1 && y <= ny - 2) {
    v = vload<T, VEC>(f, z * sz + y * sy + x0);
    for (int k = 0; k < VEC; ++k) {
      const int x = x0 + k;
      if (x < 1 || x > nx - 2) v.v[k] = T(0);
    }
  }
  return v;
}

// x neighbours (masked) of a masked row vector held across the MXT lanes of a row. Lanes
// hold adjacent columns (no clamping here); tile edges load. Every lane calls this.
template <typename T, int VEC>
__device__ __forceinline__ void mx_nbrs(const VecT<T, VEC>& mc, const T* __restrict__ f, int z,
                                        int y, int x0, int nz, int ny, int nx, int sz, int sy,
                                        int tx, bool act, T& left, T& right) {
  constexpr int MXT = MarchGeom<VEC>::MXT;
  left = __shfl_up(mc.v[VEC - 1], 1, MXT);
  right = __shfl_down(mc.v[0], 1, MXT);
  const bool row_in = act && z >= 1 && z <= nz - 2 && y >= 1 && y <= ny - 2;
  if (tx == 0) {
    const int x = x0 - 1;
    left = (row_in && x >= 1 && x <= nx - 2) ? f[z * sz + y * sy + x] : T(0);
  }
  if (tx == MXT - 1) {
    const int x = x0 + VEC;
    right = (row_in && x >= 1 && x <= nx - 2) ? f[z * sz + y * sy + x] : T(0);
  }
}

// ------------------------------------------------------------------ curl transpose

// out = p curl_h(M a) on every cell (neighbours outside the grid are zero).
template <typename T, int VEC>
__global__ void __launch_bounds__(kMarchThreads) k_curl_adj(Grid g, const T* __restrict__ a,
                                                            T* __restrict__ out, T p, int tiles_x,
                                                            int tiles_y, int remap, int chunk) {
  using V = VecT<T, VEC>;
  const MarchTile m = march_tile<VEC>(g, tiles_x, tiles_y, xcd_item(remap), chunk);
  const int nz = int(g.nz), ny = int(g.ny), nx = int(g.nx);
  const int n = nz * ny * nx, sz = ny * nx, sy = nx;
  const bool act = m.x0 < nx && m.y < ny;
  const int y = act ? m.y : 0, x0 = act ? m.x0 : 0;
  const T* f0 = a;
  const T* f1 = a + n;
  const T* f2 = a + 2 * n;
  V p0m = mrow<T, VEC>(f0, m.z0 - 1, y, x0, nz, ny, nx, sz, sy, act);
  V p1m = mrow<T, VEC>(f1, m.z0 - 1, y, x0, nz, ny, nx, sz, sy, act);
  V p0c = mrow<T, VEC>(f0, m.z0, y, x0, nz, ny, nx, sz, sy, act);
  V p1c = mrow<T, VEC>(f1, m.z0, y, x0, nz, ny, nx, sz, sy, act);
  for (int z = m.z0; z < m.z1; ++z) {
    const V p0p = mrow<T, VEC>(f0, z + 1, y, x0, nz, ny, nx, sz, sy, act);
    const V p1p = mrow<T, VEC>(f1, z + 1, y, x0, nz, ny, nx, sz, sy, act);
    const V q2c = mrow<T, VEC>(f2, z, y, x0, nz, ny, nx, sz, sy, act);
    const V q2p = mrow<T, VEC>(f2, z, y + 1, x0, nz, ny, nx, sz, sy, act);
    const V q2m = mrow<T, VEC>(f2, z, y - 1, x0, nz, ny, nx, sz, sy, act);
    const V q0p = mrow<T, VEC>(f0, z, y + 1, x0, nz, ny, nx, sz, sy, act);
    const V q0m = mrow<T, VEC>(f0, z, y - 1, x0, nz, ny, nx, sz, sy, act);
    T l1, r1, l2, r2;
    mx_nbrs<T, VEC>(p1c, f1, z, y, x0, nz, ny, nx, sz, sy, m.tx, act, l1, r1);
    mx_nbrs<T, VEC>(q2c, f2, z, y, x0, nz, ny, nx, sz, sy, m.tx, act, l2, r2);
    V o[3];
    for (int k = 0; k < VEC; ++k) {
      const T f1r = k < VEC - 1 ? p1c.v[k + 1] : r1, f1l = k > 0 ? p1c.v[k - 1] : l1;
      const T f2r = k < VEC - 1 ? q2c.v[k + 1] : r2, f2l = k > 0 ? q2c.v[k - 1] : l2;
      o[0].v[k] = p * ((q2p.v[k] - q2m.v[k]) - (p1p.v[k] - p1m.v[k]));
      o[1].v[k] = p * ((p0p.v[k] - p0m.v[k]) - (f2r - f2l));
      o[2].v[k] = p * ((f1r - f1l) - (q0p.v[k] - q0m.v[k]));
    }
    if (act)
      for (int c = 0; c < 3; ++c) vstore<T, VEC>(out + c * n, z * sz + y * sy + x0, o[c]);
    p0m = p0c;
    p1m = p1c;
    p0c = p0p;
    p1c = p1p;
  }
}

// ------------------------------------------------------------------ diffusion transpose

// out_c = t_c + pD L(M t_c) on every cell, per component.
template <typename T, int VEC>
__global__ void __launch_bounds__(kMarchThreads) k_diff_adj(Grid g, const T* __restrict__ t,
                                                            T* __restrict__ out, T pD, int tiles_x,
                                                            int tiles_y, int remap, int chunk) {
  using V = VecT<T, VEC>;
  const MarchTile m = march_tile<VEC>(g, tiles_x, tiles_y, xcd_item(remap), chunk);
  const int nz = int(g.nz), ny = int(g.ny), nx = int(g.nx);
  const int n = nz * ny * nx, sz = ny * nx, sy = nx;
  const bool act = m.x0 < nx && m.y < ny;
  const int y = act ? m.y : 0, x0 = act ? m.x0 : 0;
  V am[3], ac[3];
  for (int c = 0; c < 3; ++c) {
    am[c] = mrow<T, VEC>(t + c * n, m.z0 - 1, y, x0, nz, ny, nx, sz, sy, act);
    ac[c] = mrow<T, VEC>(t + c * n, m.z0, y, x0, nz, ny, nx, sz, sy, act);
  }
  for (int z = m.z0; z < m.z1; ++z) {
    V o[3];
    for (int c = 0; c < 3; ++c) {
      const T* f = t + c * n;
      const V ap = mrow<T, VEC>(f, z + 1, y, x0, nz, ny, nx, sz, sy, act);
      const V yp = mrow<T, VEC>(f, z, y + 1, x0, nz, ny, nx, sz, sy, act);
      const V ym = mrow<T, VEC>(f, z, y - 1, x0, nz, ny, nx, sz, sy, act);
      const V own = act ? vload<T, VEC>(f, z * sz + y * sy + x0) : vzero<T, VEC>();
      T l, r;
      mx_nbrs<T, VEC>(ac[c], f, z, y, x0, nz, ny, nx, sz, sy, m.tx, act, l, r);
      for (int k = 0; k < VEC; ++k) {
        const T xr = k < VEC - 1 ? ac[c].v[k + 1] : r, xl = k > 0 ? ac[c].v[k - 1] : l;
        o[c].v[k] = own.v[k] + pD * (((am[c].v[k] + ap.v[k]) + (ym.v[k] + yp.v[k])) + (xl + xr) -
                                     T(6) * ac[c].v[k]);
      }
      am[c] = ac[c];
      ac[c] = ap;
    }
    if (act)
      for (int c = 0; c < 3; ++c) vstore<T, VEC>(out + c * n, z * sz + y * sy + x0, o[c]);
  }
}

// ------------------------------------------------------------------ penalisation transpose

// Cells p with clamp(p) = q along one axis and their sine factors: the whole front slab for
// q = w-1, the whole back slab for q = n-w, q itself inside, nothing for q in a slab.
struct PenRange {
  int lo, hi;
  int front, back;  // lo is in the front / back slab (factor table offset)
};

__device__ __forceinline__ PenRange pen_range(int q, int n, int w) {
  PenRange r{q, q, 0, 0};
  if (w == 0) return r;
  if (q < w - 1 || q > n - w) {
    r.lo = 0;
    r.hi = -1;
  } else if (q == w - 1) {
    r.lo = 0;
    r.hi = w - 1;
    r.front = 1;
  } else if (q == n - w) {
    r.lo = n - w;
    r.hi = n - 1;
    r.back = 1;
  }
  return r;
}

template <typename T>
__device__ __forceinline__ T pen_factor(const Pen& pen, int axis, int p, int n) {
  const int w = pen.w;
  if (w == 0) return T(1);
  if (p < w) return T(pen.f[axis][p]);
  if (p >= n - w) return T(pen.f[axis][w + p - (n - w)]);
  return T(1);
}

// ------------------------------------------------------------------ fused sweeps

// P^T along one axis, in place, one thread per line of that axis (c and the two other
// coordinates): the clamp planes w-1 and n-w gather their slabs with the sine factors, the
// slab cells become zero. P = P_z P_y P_x is separable, so P^T is the three in any order.
template <typename T>
__global__ void __launch_bounds__(256) k_pen_adj_axis(T* __restrict__ f, int axis, int nz, int ny,
                                                      int nx, Pen pen) {
  const int n = nz * ny * nx;
  const int len = axis == 0 ? nx : (axis == 1 ? ny : nz);
  const int lines = 3 * (n / len);
  const int i = int(blockIdx.x) * 256 + int(threadIdx.x);
  if (i >= lines) return;
  // line base and stride: lines enumerate (c, other coordinates) with x fastest where x is
  // not the axis, so neighbouring threads touch neighbouring cells
  int base, stride;
  const int c = i / (n / len), r = i % (n / len);
  if (axis == 0) {          // rows (c, z, y)
    base = c * n + r * nx;
    stride = 1;
  } else if (axis == 1) {   // (c, z, x)
    base = c * n + (r / nx) * ny * nx + r % nx;
    stride = nx;
  } else {                  // (c, y, x)
    base = c * n + r;
    stride = ny * nx;
  }
  const int w = pen.w;
  T front = T(0), back = T(0);
  for (int p = 0; p < w; ++p) {
    front += T(pen.f[axis][p]) * f[base + p * stride];
    back += T(pen.f[axis][w + p]) * f[base + (len - w + p) * stride];
  }
  for (int p = 0; p < w - 1; ++p) {
    f[base + p * stride] = T(0);
    f[base + (len - 1 - p) * stride] = T(0);
  }
  f[base + (w - 1) * stride] = front;
  f[base + (len - w) * stride] = back;
}

// Bb = pB C(M g2) on every cell, then ub = w1 x Bb and g1 = g2 + Bb x u, with w1 the
// forward's omega1 = omega + pF M C F.
template <typename T, int VEC>
__global__ void __launch_bounds__(kMarchThreads) k_curl_cross_adj(
    Grid g, const T* __restrict__ g2, const T* __restrict__ w1, const T* __restrict__ u,
    T* __restrict__ ub, T* __restrict__ g1, T pB, int tiles_x, int tiles_y, int remap, int chunk) {
  using V = VecT<T, VEC>;
  const MarchTile m = march_tile<VEC>(g, tiles_x, tiles_y, xcd_item(remap), chunk);
  const int nz = int(g.nz), ny = int(g.ny), nx = int(g.nx);
  const int n = nz * ny * nx, sz = ny * nx, sy = nx;
  const bool act = m.x0 < nx && m.y < ny;
  const int y = act ? m.y : 0, x0 = act ? m.x0 : 0;
  const T* f0 = g2;
  const T* f1 = g2 + n;
  const T* f2 = g2 + 2 * n;
  V p0m = mrow<T, VEC>(f0, m.z0 - 1, y, x0, nz, ny, nx, sz, sy, act);
  V p1m = mrow<T, VEC>(f1, m.z0 - 1, y, x0, nz, ny, nx, sz, sy, act);
  V p0c = mrow<T, VEC>(f0, m.z0, y, x0, nz, ny, nx, sz, sy, act);
  V p1c = mrow<T, VEC>(f1, m.z0, y, x0, nz, ny, nx, sz, sy, act);
  for (int z = m.z0; z < m.z1; ++z) {
    const V p0p = mrow<T, VEC>(f0, z + 1, y, x0, nz, ny, nx, sz, sy, act);
    const V p1p = mrow<T, VEC>(f1, z + 1, y, x0, nz, ny, nx, sz, sy, act);
    const V q2c = mrow<T, VEC>(f2, z, y, x0, nz, ny, nx, sz, sy, act);
    const V q2p = mrow<T, VEC>(f2, z, y + 1, x0, nz, ny, nx, sz, sy, act);
    const V q2m = mrow<T, VEC>(f2, z, y - 1, x0, nz, ny, nx, sz, sy, act);
    const V q0p = mrow<T, VEC>(f0, z, y + 1, x0, nz, ny, nx, sz, sy, act);
    const V q0m = mrow<T, VEC>(f0, z, y - 1, x0, nz, ny, nx, sz, sy, act);
    T l1, r1, l2, r2;
    mx_nbrs<T, VEC>(p1c, f1, z, y, x0, nz, ny, nx, sz, sy, m.tx, act, l1, r1);
    mx_nbrs<T, VEC>(q2c, f2, z, y, x0, nz, ny, nx, sz, sy, m.tx, act, l2, r2);
    const int i = z * sz + y * sy + x0;
    if (act) {
      V A[3], G2[3], U[3], o[3], q[3];
      for (int c = 0; c < 3; ++c) {
        A[c] = vload<T, VEC>(w1 + c * n, i);
        G2[c] = vload<T, VEC>(g2 + c * n, i);
        U[c] = vload<T, VEC>(u + c * n, i);
      }
      for (int k = 0; k < VEC; ++k) {
        const T f1r = k < VEC - 1 ? p1c.v[k + 1] : r1, f1l = k > 0 ? p1c.v[k - 1] : l1;
        const T f2r = k < VEC - 1 ? q2c.v[k + 1] : r2, f2l = k > 0 ? q2c.v[k - 1] : l2;
        const T b0 = pB * ((q2p.v[k] - q2m.v[k]) - (p1p.v[k] - p1m.v[k]));
        const T b1 = pB * ((p0p.v[k] - p0m.v[k]) - (f2r - f2l));
        const T b2 = pB * ((f1r - f1l) - (q0p.v[k] - q0m.v[k]));
        const T a0 = A[0].v[k], a1 = A[1].v[k], a2 = A[2].v[k];
        const T u0 = U[0].v[k], u1 = U[1].v[k], u2 = U[2].v[k];
        o[0].v[k] = a1 * b2 - a2 * b1;
        o[1].v[k] = a2 * b0 - a0 * b2;
        o[2].v[k] = a0 * b1 - a1 * b0;
        q[0].v[k] = G2[0].v[k] + (b1 * u2 - b2 * u1);
        q[1].v[k] = G2[1].v[k] + (b2 * u0 - b0 * u2);
        q[2].v[k] = G2[2].v[k] + (b0 * u1 - b1 * u0);
      }
      for (int c = 0; c < 3; ++c) {
        vstore<T, VEC>(ub + c * n, i, o[c]);
        vstore<T, VEC>(g1 + c * n, i, q[c]);
      }
    }
    p0m = p0c;
    p1m = p1c;
    p0c = p0p;
    p1c = p1p;
  }
}

template <typename T>
int launch_curl_adj(const MarchLaunch& l, Grid g, const T* in, T* out, T p, hipStream_t s) {
  if (l.vec == 1)
    k_curl_adj<T, 1><<<l.blocks, kMarchThreads, 0, s>>>(g, in, out, p, l.tiles_x, l.tiles_y, l.remap, l.chunk);
  else
    k_curl_adj<T, 16 / sizeof(T)><<<l.blocks, kMarchThreads, 0, s>>>(g, in, out, p, l.tiles_x,
                                                                    l.tiles_y, l.remap, l.chunk);
  SOPHT_CHECK_LAUNCH();
  return SOPHT_OK;
}

// vector width of the adjoint sweeps: whole 16-byte vectors per row (no penalty clamp here)
template <typename T>
inline int adj_vec(int64_t nx) {
  const int v = 16 / int(sizeof(T));
  return (nx % v == 0 && nx >= 2 * v) ? v : 1;
}

template <typename T>
int adjoint_step_impl(sopht_ns3d* P, const void* vort1, const void* vel, double dt,
                      void* vort_adj, void* vel_adj, void* work, hipStream_t s) {
  Grid g{P->nz, P->ny, P->nx};
  const int64_t n = g.nz * g.ny * g.nx;
  const T dx = T(P->dx);
  // the forward's prefactors (flow_step_3d.hip step_impl)
  const T pB = T(dt / (2.0 * double(dx)));
  const T pD = T(P->nu * dt / double(dx) / double(dx));
  const T pC = T(0.5 / double(dx));
  T* wb = static_cast<T*>(vort_adj);
  T* ub = static_cast<T*>(vel_adj);
  T* W0 = static_cast<T*>(work);
  const MarchLaunch ml = march_launch(g.nz, g.ny, g.nx, adj_vec<T>(g.nx), true);
  int rc;
  // psib = pC C(M ub)
  if ((rc = launch_curl_adj<T>(ml, g, ub, W0, pC, s))) return rc;
  // wb += G psib (the Poisson's last pass accumulates)
  {
    void* const sol[3] = {wb, wb + n, wb + 2 * n};
    const void* const rhs[3] = {W0, W0 + n, W0 + 2 * n};
    const int64_t st[3] = {g.ny * g.nx, g.nx, 1};
    if ((rc = poisson_vector_solve_accumulate(P->poisson, sol, st, rhs, st, s))) return rc;
  }
  // t = P^T wb, in place, axis by axis
  const Pen pen = make_pen(P);
  if (pen.w > 0) {
    const int dims[3] = {int(g.nx), int(g.ny), int(g.nz)};
    for (int axis = 0; axis < 3; ++axis) {
      const int lines = int(3 * n / dims[axis]);
      k_pen_adj_axis<T><<<unsigned(ceil_div(lines, 256)), 256, 0, s>>>(wb, axis, int(g.nz), int(g.ny),
                                                                       int(g.nx), pen);
      SOPHT_CHECK_LAUNCH();
    }
  }
  // g2 = t + pD L(M t)
  const T* w1 = static_cast<const T*>(vort1);
  const T* un = static_cast<const T*>(vel);
  constexpr int VV = 16 / sizeof(T);
  if (ml.vec == 1)
    k_diff_adj<T, 1><<<ml.blocks, kMarchThreads, 0, s>>>(g, wb, W0, pD, ml.tiles_x, ml.tiles_y, ml.remap, ml.chunk);
  else
    k_diff_adj<T, VV><<<ml.blocks, kMarchThreads, 0, s>>>(g, wb, W0, pD, ml.tiles_x, ml.tiles_y, ml.remap, ml.chunk);
  SOPHT_CHECK_LAUNCH();
  // Bb = pB C(M g2) ; ub = w1 x Bb ; wb = g2 + Bb x u
  if (ml.vec == 1)
    k_curl_cross_adj<T, 1><<<ml.blocks, kMarchThreads, 0, s>>>(g, W0, w1, un, ub, wb, pB, ml.tiles_x,
                                                               ml.tiles_y, ml.remap, ml.chunk);
  else
    k_curl_cross_adj<T, VV><<<ml.blocks, kMarchThreads, 0, s>>>(g, W0, w1, un, ub, wb, pB, ml.tiles_x,
                                                                ml.tiles_y, ml.remap, ml.chunk);
  SOPHT_CHECK_LAUNCH();
  return SOPHT_OK;
}

// ------------------------------------------------------------------ virtual-boundary adjoint

template <typename T>
struct VbfAdjArgs {
  const T* wb;      // omega_bar after the flow-step adjoint (g1)
  const T* u;       // u_n
  int64_t nz, ny, nx;
  const double* X;  // node positions of the step
  const T* Fl;      // the step's Lagrangian forcing
  int64_t n;
  double dx, shift, k, c;
  T pF, dt;
  const T* fbar;    // adjoint of the body force (3)
  T* dXb;           // in/out
  T* dVb;           // in/out
  T* Xb;            // out
  T* Vb;            // out
  T* weights;       // out (for spreading)
  int64_t* idx;     // out
  T* Us;            // out: dx^3 * U_bar (for spreading)
};

// pF C(M wb) at cell (z, y, x): the forcing adjoint at one support cell
template <typename T>
__device__ __forceinline__ void forcing_adj_at(const VbfAdjArgs<T>& A, int64_t z, int64_t y,
                                               int64_t x, T out[3]) {
  const int64_t nz = A.nz, ny = A.ny, nx = A.nx, n = nz * ny * nx;
  auto m = [&](int c, int64_t zz, int64_t yy, int64_t xx) -> T {
    return (zz >= 1 && zz <= nz - 2 && yy >= 1 && yy <= ny - 2 && xx >= 1 && xx <= nx - 2)
               ? A.wb[c * n + (zz * ny + yy) * nx + xx]
               : T(0);
  };
  out[0] = A.pF * ((m(2, z, y + 1, x) - m(2, z, y - 1, x)) - (m(1, z + 1, y, x) - m(1, z - 1, y, x)));
  out[1] = A.pF * ((m(0, z + 1, y, x) - m(0, z - 1, y, x)) - (m(2, z, y, x + 1) - m(2, z, y, x - 1)));
  out[2] = A.pF * ((m(1, z, y, x + 1) - m(1, z, y, x - 1)) - (m(0, z, y + 1, x) - m(0, z, y - 1, x)));
}

// One wavefront per node, lane = support slot (kz, ky, kx). Cosine kernel only.
template <typename T>
__global__ void __launch_bounds__(256) k_vbf_adjoint(VbfAdjArgs<T> A) {
  const int64_t i = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (i >= A.n) return;  // whole wavefronts leave together; no barriers follow
  const int lane = threadIdx.x & 63;
  const int64_t n = A.n;
  const int off[3] = {lane & 3, (lane >> 2) & 3, lane >> 4};  // kx, ky, kz
  const double dx = A.dx, shift = A.shift;
  const T tdx = T(dx);
  int64_t base[3];
  T s[3], ph[3], dph[3];
  for (int a = 0; a < 3; ++a) {
    const double xa = A.X[a * n + i];
    const int64_t ia = int64_t(np_floor_divide(xa - shift, dx));
    base[a] = ia - 1;
    if (lane == 0) A.idx[a * n + i] = ia;
    const T dr = T(double(ia + off[a] - 1) * dx + shift - xa);
    s[a] = dr / tdx;
    ph[a] = cosine_phi<T>(s[a]);
    // d phi(s)/dX = -phi'(s) / dx = (pi / (2 dx)) sin(pi s / 2)
    dph[a] = (T(0.5 * M_PI) / tdx) * sin(T(0.5 * M_PI) * s[a]);
  }
  const T pre = T((T(0.25) / tdx) * (T(0.25) / tdx) * (T(0.25) / tdx));
  const T w = ((pre * ph[0]) * ph[1]) * ph[2];
  A.weights[int64_t(lane) * n + i] = w;
  const T dw[3] = {pre * dph[0] * ph[1] * ph[2], pre * ph[0] * dph[1] * ph[2],
                   pre * ph[0] * ph[1] * dph[2]};
  // r[0..2] = Fb w ; r[3 + 3c + a] = Fb_c dw_a ; r[12 + 3c + a] = u_c dw_a
  T r[21];
  for (int j = 0; j < 21; ++j) r[j] = T(0);
  if (block_inside<T>(base, A.nz, A.ny, A.nx)) {
    const int64_t z = base[2] + off[2], y = base[1] + off[1], x = base[0] + off[0];
    T Fb[3];
    forcing_adj_at<T>(A, z, y, x, Fb);
    const int64_t nn = A.nz * A.ny * A.nx;
    const int64_t cell = (z * A.ny + y) * A.nx + x;
    for (int c = 0; c < 3; ++c) {
      const T uc = A.u[c * nn + cell];
      r[c] = Fb[c] * w;
      for (int a = 0; a < 3; ++a) {
        r[3 + 3 * c + a] = Fb[c] * dw[a];
        r[12 + 3 * c + a] = uc * dw[a];
      }
    }
  }
  for (int mm = 32; mm >= 1; mm >>= 1)
    for (int j = 0; j < 21; ++j) r[j] += __shfl_xor(r[j], mm, 64);
  if (lane == 0) {
    const T dx3 = tdx * tdx * tdx;
    T Flb[3], Ub[3];
    for (int c = 0; c < 3; ++c) {
      Flb[c] = r[c] - A.fbar[c];
      const T dxb = A.dXb[c * n + i] + T(A.k) * Flb[c];
      Ub[c] = A.dVb[c * n + i] + T(A.c) * Flb[c];
      A.Vb[c * n + i] = -Ub[c];
      A.Us[c * n + i] = Ub[c] * dx3;
      A.dXb[c * n + i] = dxb;
      A.dVb[c * n + i] = A.dt * dxb;
    }
    for (int a = 0; a < 3; ++a) {
      T xb = T(0);
      for (int c = 0; c < 3; ++c)
        xb += A.Fl[c * n + i] * r[3 + 3 * c + a] + dx3 * Ub[c] * r[12 + 3 * c + a];
      A.Xb[a * n + i] = xb;
    }
  }
}

}  // namespace
}  // namespace sopht

using namespace sopht;

extern "C" {

int sopht_ns3d_adjoint_step(sopht_ns3d* P, const void* vort1, const void* vel, double dt,
                            void* vort_adj, void* vel_adj, void* work, void* stream) {
  SOPHT_REQUIRE(P != nullptr, "plan is null");
  SOPHT_REQUIRE(vort1 && vel && vort_adj && vel_adj && work, "null field pointer");
  SOPHT_REQUIRE(P->filter_order < 0, "the adjoint does not cover the vorticity filter");
  SOPHT_REQUIRE(P->poisson != nullptr, "the adjoint covers the greens-function Poisson solver only");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return P->dtype == SOPHT_F32
             ? adjoint_step_impl<float>(P, vort1, vel, dt, vort_adj, vel_adj, work, s)
             : adjoint_step_impl<double>(P, vort1, vel, dt, vort_adj, vel_adj, work, s);
}

int sopht_vbf_adjoint_3d(int dtype, const void* vort_adj, double forcing_prefactor,
                         const void* velocity, int64_t nz, int64_t ny, int64_t nx,
                         const double* lag_positions, const void* lag_forcing,
                         int64_t num_lag_nodes, double dx, double eul_grid_coord_shift,
                         double stiffness, double damping, double dt, const void* body_force_adj,
                         void* position_mismatch_adj, void* velocity_mismatch_adj,
                         void* lag_position_adj, void* lag_velocity_adj, void* weights,
                         int64_t* nearest_index, void* lag_scratch, sopht_spread3d* ws,
                         void* velocity_adj, void* stream) {
  SOPHT_REQUIRE(dtype == SOPHT_F32 || dtype == SOPHT_F64, "dtype must be SOPHT_F32 or SOPHT_F64");
  SOPHT_REQUIRE(vort_adj && velocity && lag_positions && lag_forcing && body_force_adj &&
                    position_mismatch_adj && velocity_mismatch_adj && lag_position_adj &&
                    lag_velocity_adj && weights && nearest_index && lag_scratch && ws &&
                    velocity_adj,
                "null pointer");
  SOPHT_REQUIRE(num_lag_nodes >= 0, "negative node count");
  SOPHT_REQUIRE(3 * nz * ny * nx < (int64_t(1) << 31), "grid too large (3 nz ny nx must be below 2^31)");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (num_lag_nodes == 0) return SOPHT_OK;
  const unsigned blocks = unsigned((num_lag_nodes + 3) / 4);
  if (dtype == SOPHT_F32) {
    using T = float;
    VbfAdjArgs<T> A{static_cast<const T*>(vort_adj), static_cast<const T*>(velocity), nz, ny, nx,
                    lag_positions, static_cast<const T*>(lag_forcing), num_lag_nodes, dx,
                    eul_grid_coord_shift, stiffness, damping, T(forcing_prefactor), T(dt),
                    static_cast<const T*>(body_force_adj), static_cast<T*>(position_mismatch_adj),
                    static_cast<T*>(velocity_mismatch_adj), static_cast<T*>(lag_position_adj),
                    static_cast<T*>(lag_velocity_adj), static_cast<T*>(weights), nearest_index,
                    static_cast<T*>(lag_scratch)};
    k_vbf_adjoint<T><<<blocks, 256, 0, s>>>(A);
  } else {
    using T = double;
    VbfAdjArgs<T> A{static_cast<const T*>(vort_adj), static_cast<const T*>(velocity), nz, ny, nx,
                    lag_positions, static_cast<const T*>(lag_forcing), num_lag_nodes, dx,
                    eul_grid_coord_shift, stiffness, damping, T(forcing_prefactor), T(dt),
                    static_cast<const T*>(body_force_adj), static_cast<T*>(position_mismatch_adj),
                    static_cast<T*>(velocity_mismatch_adj), static_cast<T*>(lag_position_adj),
                    static_cast<T*>(lag_velocity_adj), static_cast<T*>(weights), nearest_index,
                    static_cast<T*>(lag_scratch)};
    k_vbf_adjoint<T><<<blocks, 256, 0, s>>>(A);
  }
  SOPHT_CHECK_LAUNCH();
  // u_bar += S(dx^3 U_bar): spreading is the transpose of interpolation up to dx^3
  const int64_t n = nz * ny * nx;
  const size_t es = dtype == SOPHT_F32 ? 4 : 8;
  char* ub = static_cast<char*>(velocity_adj);
  void* const eul[3] = {ub, ub + n * es, ub + 2 * n * es};
  const int64_t st[3] = {ny * nx, nx, 1};
  return sopht_lag_to_eul_spread_3d(dtype, 3, ws, eul, st, nz, ny, nx, lag_scratch, weights,
                                    nearest_index, num_lag_nodes, stream);
}

}  // extern "C"
