// The Laplacian vorticity filter of the flow step (laplacian_filter_3d.py:13-194, applied at
// navier_stokes_flow_simulators.py:473) as one launch per axis for all components, instead of
// the reference's sequence of flux, copy and saxpby kernels per pass and per component.
//
// The reference's 1D filter pass along axis a writes, on the interior [1:-1]^3 of the grid,
//   P_a(g)[i] = 0.25 * ((-g[i+1] - g[i-1]) + 2 g[i])          (laplacian_filter_3d.py:60-80)
// and 0 on the border (the flux buffer's border is zeroed and copied into the next pass's
// input). So along a line that lies on the border of another axis every pass gives 0, and
// along an interior line the first pass reads the raw border values while later passes read
// the zeros the previous pass left there.
//   convolution (:124-166): for a = x, y, z in turn: f <- 1 * f + (-1) * P_a^k(f)
//   multiplicative (:98-122): g = f; k times: g <- P_x g, g <- P_y g, g <- P_z g; then
//                              f <- 1 * f + (-1) * g
// Kernels:
//   k_filter_rows   x: one wavefront per row, the row across the lanes (x = v * 64 + lane),
//                   every pass in registers with lane shuffles;
//   k_filter_march  y or z: one thread per line, marching along it with the k passes as a
//                   register pipeline (pass j produces its value j cells behind the read
//                   position from the three latest values of pass j - 1).
// The fused step uses these for orders 1..kMaxFusedOrder; the standalone factory keeps the
// reference's kernel sequence, whose buffers are part of its observable result.
#include "common.h"

namespace sopht {
namespace {

constexpr int kMaxFusedOrder = 8;
constexpr int kMaxRowVec = 16;  // rows up to 1024 cells

enum FilterMode {
  kConv = 0,       // f <- f - P^k f, in place (lines on another axis' border unchanged)
  kMultPass = 1,   // out <- P in (one pass), out may be in
  kMultFinal = 2,  // f <- f - P in (one pass)
};

template <typename T>
__device__ __forceinline__ T lap_pass(T left, T mid, T right) {
  return T(0.25) * ((-right - left) + T(2) * mid);
}

template <typename T>
struct FilterComps {
  const T* in[3];
  T* out[3];
  T* f[3];  // the filtered field (kMultFinal: f <- f - P in)
};

// x passes. Row r = blockIdx.x * 4 + wave of component blockIdx.y; lane l holds the VEC
// consecutive cells x = VEC l + i, so a pass exchanges only the two cells at the lane's edges
// with its neighbours. Rows of a multiple of 4 cells load and store 16 bytes per access.
template <typename T, int VEC, int MODE>
__global__ void __launch_bounds__(256) k_filter_rows(FilterComps<T> c, int nz, int ny, int nx, int order) {
  const int lane = threadIdx.x & 63;
  const int64_t row = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (row >= int64_t(nz) * ny) return;  // uniform per wavefront
  const int comp = blockIdx.y;
  const int z = int(row / ny), y = int(row % ny);
  const bool line_in = z > 0 && z < nz - 1 && y > 0 && y < ny - 1;
  const int64_t base = row * nx;
  if (MODE == kConv && !line_in) return;  // the flux is 0 along the whole row
  const T* in = c.in[comp] + base;
  const int x0 = VEC * lane;
  // 16-byte accesses need 4 | nx (the row base is then 16-byte aligned) and 4 | VEC
  const bool vec4 = sizeof(T) == 4 && VEC % 4 == 0 && nx % 4 == 0;
  T g[VEC];
  if (vec4) {
    typedef T quad __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int i = 0; i < VEC; i += 4) {
      if (x0 + i < nx) {
        const quad q = *reinterpret_cast<const quad*>(in + x0 + i);
        g[i] = q.x; g[i + 1] = q.y; g[i + 2] = q.z; g[i + 3] = q.w;
      } else {
        g[i] = g[i + 1] = g[i + 2] = g[i + 3] = T(0);
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < VEC; ++i) g[i] = x0 + i < nx ? in[x0 + i] : T(0);
  }
  T g0[VEC];  // the row's input (kConv updates it in place: no second read)
#pragma unroll
  for (int i = 0; i < VEC; ++i) g0[i] = g[i];
  const int passes = MODE == kConv ? order : 1;
  for (int p = 0; p < passes; ++p) {
    // x0 - 1 from the lane below (its last cell), x0 + VEC from the lane above (its first)
    const T left = __shfl_up(g[VEC - 1], 1, 64);
    const T right = __shfl_down(g[0], 1, 64);
    T h[VEC];
#pragma unroll
    for (int i = 0; i < VEC; ++i) {
      const int x = x0 + i;
      const T l = i > 0 ? g[i > 0 ? i - 1 : 0] : left;
      const T r = i < VEC - 1 ? g[i < VEC - 1 ? i + 1 : 0] : right;
      h[i] = (line_in && x > 0 && x < nx - 1) ? lap_pass(l, g[i], r) : T(0);
    }
#pragma unroll
    for (int i = 0; i < VEC; ++i) g[i] = h[i];
  }
  if (MODE == kMultPass) {
    T* o = c.out[comp] + base;
#pragma unroll
    for (int i = 0; i < VEC; ++i)
      if (x0 + i < nx) o[x0 + i] = g[i];
  } else {
    T* f = (MODE == kConv ? c.out[comp] : c.f[comp]) + base;
    if (vec4) {
      typedef T quad __attribute__((ext_vector_type(4)));
#pragma unroll
      for (int i = 0; i < VEC; i += 4) {
        if (x0 + i < nx) {
          quad q;
          if (MODE == kConv) {
            q.x = g0[i]; q.y = g0[i + 1]; q.z = g0[i + 2]; q.w = g0[i + 3];
          } else {
            q = *reinterpret_cast<const quad*>(f + x0 + i);
          }
          q.x = T(1) * q.x + T(-1) * g[i];
          q.y = T(1) * q.y + T(-1) * g[i + 1];
          q.z = T(1) * q.z + T(-1) * g[i + 2];
          q.w = T(1) * q.w + T(-1) * g[i + 3];
          *reinterpret_cast<quad*>(f + x0 + i) = q;
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < VEC; ++i)
        if (x0 + i < nx) f[x0 + i] = T(1) * (MODE == kConv ? g0[i] : f[x0 + i]) + T(-1) * g[i];
    }
  }
}

// y (AXIS 1) or z (AXIS 2) passes: one thread per line (the other two coordinates), lanes
// along x for coalesced accesses; K passes pipelined along the line. The line is read in
// batches of kMarchBatch values, all in flight together; the convolution keeps the last K + 1
// inputs in registers for its update f - P^K f (the value at q = p - K is still the input
// there: this thread writes the line only behind its reads).
constexpr int kMarchBatch = 8;

template <typename T, int AXIS, int K, int MODE>
__global__ void __launch_bounds__(256) k_filter_march(FilterComps<T> c, int nz, int ny, int nx) {
  constexpr int D = kMarchBatch;
  const int comp = blockIdx.z;
  const int x = blockIdx.x * 256 + threadIdx.x;
  const int other = blockIdx.y;  // z for the y march, y for the z march
  if (x >= nx) return;
  const int n = AXIS == 1 ? ny : nz;
  const int64_t stride = AXIS == 1 ? int64_t(nx) : int64_t(ny) * nx;
  const int64_t start = AXIS == 1 ? int64_t(other) * ny * nx + x : int64_t(other) * nx + x;
  const int no = AXIS == 1 ? nz : ny;  // extent of the line's other coordinate
  const bool line_in = other > 0 && other < no - 1 && x > 0 && x < nx - 1;
  if (MODE == kConv && !line_in) return;
  const T* in = c.in[comp] + start;
  T* out = (MODE == kMultPass ? c.out[comp] : (MODE == kConv ? c.out[comp] : c.f[comp])) + start;
  // w[j][0..2]: the three latest values of pass j (w[0]: the input)
  T w[K + 1][3];
#pragma unroll
  for (int j = 0; j <= K; ++j) w[j][0] = w[j][1] = w[j][2] = T(0);
  T hist[K + 1];  // inputs at p - K .. p (kConv)
#pragma unroll
  for (int j = 0; j <= K; ++j) hist[j] = T(0);
  for (int p0 = 0; p0 < n + K; p0 += D) {
    T buf[D], fq[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int p = p0 + d;
      buf[d] = p < n ? in[p * stride] : T(0);
    }
    if (MODE == kMultFinal) {
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const int q = p0 + d - K;
        fq[d] = (q >= 0 && q < n) ? out[q * stride] : T(0);
      }
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int p = p0 + d;
      w[0][0] = w[0][1];
      w[0][1] = w[0][2];
      w[0][2] = buf[d];
#pragma unroll
      for (int j = 0; j < K; ++j) hist[j] = hist[j + 1];
      hist[K] = buf[d];
#pragma unroll
      for (int j = 1; j <= K; ++j) {
        const int q = p - j;  // pass j's position this iteration; its inputs q - 1, q, q + 1
        T v = T(0);
        if (line_in && q > 0 && q < n - 1) v = lap_pass(w[j - 1][0], w[j - 1][1], w[j - 1][2]);
        w[j][0] = w[j][1];
        w[j][1] = w[j][2];
        w[j][2] = v;
      }
      const int q = p - K;
      if (q >= 0 && q < n) {
        const T flux = w[K][2];
        if (MODE == kMultPass) {
          out[q * stride] = flux;
        } else if (MODE == kConv) {
          out[q * stride] = T(1) * hist[0] + T(-1) * flux;
        } else {
          out[q * stride] = T(1) * fq[d] + T(-1) * flux;
        }
      }
    }
  }
}

template <typename T, int MODE>
int launch_rows(const FilterComps<T>& c, int nz, int ny, int nx, int order, hipStream_t s) {
  const unsigned blocks = unsigned((int64_t(nz) * ny + 3) / 4);
  const dim3 grid(blocks, 3);
  const int vec = (nx + 63) / 64;
#define ROWS(V)                                                                      \
  if (vec <= V) {                                                                    \
    k_filter_rows<T, V, MODE><<<grid, 256, 0, s>>>(c, nz, ny, nx, order);            \
    SOPHT_CHECK_LAUNCH();                                                            \
    return SOPHT_OK;                                                                 \
  }
  ROWS(1) ROWS(2) ROWS(4) ROWS(8) ROWS(16)
#undef ROWS
  set_last_error("row too long for the fused filter", __FILE__, __LINE__);
  return SOPHT_ERR_UNSUPPORTED;
}

template <typename T, int AXIS, int MODE>
int launch_march(const FilterComps<T>& c, int nz, int ny, int nx, int order, hipStream_t s) {
  const dim3 grid(unsigned((nx + 255) / 256), unsigned(AXIS == 1 ? nz : ny), 3);
#define MARCH(KK)                                                                      \
  if (order == KK) {                                                                   \
    k_filter_march<T, AXIS, KK, MODE><<<grid, 256, 0, s>>>(c, nz, ny, nx);            \
    SOPHT_CHECK_LAUNCH();                                                              \
    return SOPHT_OK;                                                                   \
  }
  if (MODE != kConv) {
    MARCH(1)
  } else {
    MARCH(1) MARCH(2) MARCH(3) MARCH(4) MARCH(5) MARCH(6) MARCH(7) MARCH(8)
  }
#undef MARCH
  set_last_error("filter order outside the fused filter's range", __FILE__, __LINE__);
  return SOPHT_ERR_UNSUPPORTED;
}

template <typename T>
int fused_filter_impl(int order, int type, T* w, T* buf, int64_t nz, int64_t ny, int64_t nx,
                      hipStream_t s) {
  const int64_t n = nz * ny * nx;
  const int Z = int(nz), Y = int(ny), X = int(nx);
  FilterComps<T> c;
  for (int k = 0; k < 3; ++k) c.f[k] = w + k * n;
  int rc;
  if (type == 1) {  // convolution: x, y, z in place on the field
    for (int k = 0; k < 3; ++k) {
      c.in[k] = w + k * n;
      c.out[k] = w + k * n;
    }
    if ((rc = launch_rows<T, kConv>(c, Z, Y, X, order, s))) return rc;
    if ((rc = launch_march<T, 1, kConv>(c, Z, Y, X, order, s))) return rc;
    return launch_march<T, 2, kConv>(c, Z, Y, X, order, s);
  }
  // multiplicative: the passes go through the buffer g; the last one updates the field
  for (int p = 0; p < order; ++p) {
    for (int k = 0; k < 3; ++k) {
      c.in[k] = (p == 0 ? w : buf) + k * n;
      c.out[k] = buf + k * n;
    }
    if ((rc = launch_rows<T, kMultPass>(c, Z, Y, X, 1, s))) return rc;
    for (int k = 0; k < 3; ++k) c.in[k] = buf + k * n;
    if ((rc = launch_march<T, 1, kMultPass>(c, Z, Y, X, 1, s))) return rc;
    if (p + 1 < order) {
      if ((rc = launch_march<T, 2, kMultPass>(c, Z, Y, X, 1, s))) return rc;
    } else {
      if ((rc = launch_march<T, 2, kMultFinal>(c, Z, Y, X, 1, s))) return rc;
    }
  }
  return SOPHT_OK;
}

}  // namespace

bool fused_filter_supported(int order, int64_t nx) {
  return order >= 1 && order <= kMaxFusedOrder && nx <= 64 * kMaxRowVec && nx >= 1;
}

int fused_vorticity_filter(int dtype, int order, int type, void* w, void* buf, int64_t nz, int64_t ny,
                           int64_t nx, hipStream_t s) {
  return dtype == SOPHT_F32
             ? fused_filter_impl<float>(order, type, static_cast<float*>(w), static_cast<float*>(buf), nz, ny, nx, s)
             : fused_filter_impl<double>(order, type, static_cast<double*>(w), static_cast<double*>(buf), nz, ny, nx, s);
}

}  // namespace sopht
