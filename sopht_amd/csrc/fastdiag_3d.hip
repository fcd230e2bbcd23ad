// Fast-diagonalisation Poisson solver in 3D with homogeneous Neumann conditions on every
// face (FastDiagPoissonSolver3D.py:12-202).
//
// The reference builds the second-difference matrix of each axis, sets its corner entries
// to 1/dx^2 (cell-centred Neumann, :80-96), takes a dense eigendecomposition with the
// eigenvalues sorted in decreasing order (:98-145) and solves -lap(psi) = rhs as
//   psi = Vz Vy Vx [ (Vz^-1 Vy^-1 Vx^-1 rhs) / (lz + ly + lx) ],
// one dense N x N contraction per axis and direction, with the constant (null) mode set to
// zero (:147-201).
//
// That matrix is diagonalised exactly by the cell-centred cosines (DCT-II):
//   v_k(i) = c_k cos(pi k (i + 1/2) / N),   l_k = (4 / dx^2) sin^2(pi k / (2 N)),
// with c_0 = sqrt(1/N), c_k = sqrt(2/N), so the eigenbasis is built here in closed form in
// double precision, orthonormal (inverse = transpose), in the reference's decreasing order
// (column p holds mode k = N-1-p; the null mode is the last column). The operator
// V diag(1/l) V^-1 does not depend on the signs or scaling of the eigenvectors, so the
// result equals the reference's to rounding.
//
// Device work per solve: six GEMMs on the MFMA units through rocBLAS (the three components
// of a vector field batched into each) and one elementwise spectral scaling. For an N^3
// grid each GEMM does 2 N^4 flops per component on N^3 inputs: compute-bound (64 flop/B in
// FP32 at N = 256), so the roofline is the dense FP32 / FP64 MFMA peak.
#include <rocblas/rocblas.h>

#include <cmath>
#include <new>
#include <vector>

#include "common.h"

struct sopht_fastdiag3d {
  int dtype;
  int64_t n[3];               // z, y, x
  double dx;
  std::vector<double> lam[3];  // per axis, decreasing
  std::vector<double> vec[3];  // per axis, row-major N x N: vec[i * N + p] = v_p(i)
  void* fwd[3] = {};           // device, row-major M[p][i] = v_p(i)   (V^-1 = V^T)
  void* bwd[3] = {};           // device, row-major M[i][p] = v_p(i)   (V)
  void* lam_dev[3] = {};       // device eigenvalues in the working precision
  void* work = nullptr;        // 3 N^3 spectral buffer
  rocblas_handle handle = nullptr;
};

namespace sopht {
namespace {

// f[c][kz][ky][kx] *= 1 / ((lz + ly) + lx); the mode with all three indices last (the
// constant eigenvector, eigenvalue 0) is set to 0 (FastDiagPoissonSolver3D.py:140-145).
// Grid: y over (component, kz) planes, x over bands of kScaleRows rows of a plane. A thread
// owns V consecutive kx (one 16-byte access when V * sizeof(T) == 16) in one row at a time;
// the 256 threads cover 256 / (nx / V) rows at once (or loop along a row when nx / V > 256).
constexpr int kScaleRows = 16;

template <typename T, int V>
struct alignas(V * sizeof(T)) Pack {
  T v[V];
};

template <typename T, int V>
__global__ void __launch_bounds__(256) k_spectral_scale(T* __restrict__ f, const T* __restrict__ lz,
                                                        const T* __restrict__ ly,
                                                        const T* __restrict__ lx, int nz, int ny,
                                                        int nx) {
  const int vx = nx / V;                    // packs per row
  const int rows = vx >= 256 ? 1 : 256 / vx;
  const int tr = int(threadIdx.x) / (vx >= 256 ? 256 : vx);
  const int tc = int(threadIdx.x) - tr * (vx >= 256 ? 256 : vx);
  if (tr >= rows) return;
  const int kz = int(blockIdx.y) % nz;
  const T a = lz[kz];
  const int y0 = int(blockIdx.x) * kScaleRows;
  const int y1 = min(ny, y0 + kScaleRows);
  Pack<T, V>* plane = reinterpret_cast<Pack<T, V>*>(f + int64_t(blockIdx.y) * ny * nx);
  for (int ky = y0 + tr; ky < y1; ky += rows) {
    const T ab = a + ly[ky];
    const bool last_row = kz == nz - 1 && ky == ny - 1;
    for (int c = tc; c < vx; c += 256) {
      Pack<T, V> p = plane[int64_t(ky) * vx + c];
#pragma unroll
      for (int j = 0; j < V; ++j) {
        const int kx = c * V + j;
        const T inv = (last_row && kx == nx - 1) ? T(0) : T(1) / (ab + lx[kx]);
        p.v[j] = p.v[j] * inv;
      }
      plane[int64_t(ky) * vx + c] = p;
    }
  }
}

// C = op(A) B or A B in rocBLAS's column-major convention, strided over a batch.
template <typename T>
rocblas_status gemm(rocblas_handle h, rocblas_operation ta, rocblas_int m, rocblas_int n,
                    rocblas_int k, const T* A, rocblas_int lda, rocblas_stride sa, const T* B,
                    rocblas_int ldb, rocblas_stride sb, T* C, rocblas_int ldc, rocblas_stride sc,
                    rocblas_int batch);
template <>
rocblas_status gemm<float>(rocblas_handle h, rocblas_operation ta, rocblas_int m, rocblas_int n,
                           rocblas_int k, const float* A, rocblas_int lda, rocblas_stride sa,
                           const float* B, rocblas_int ldb, rocblas_stride sb, float* C,
                           rocblas_int ldc, rocblas_stride sc, rocblas_int batch) {
  const float one = 1.f, zero = 0.f;
  return rocblas_sgemm_strided_batched(h, ta, rocblas_operation_none, m, n, k, &one, A, lda, sa,
                                       B, ldb, sb, &zero, C, ldc, sc, batch);
}
template <>
rocblas_status gemm<double>(rocblas_handle h, rocblas_operation ta, rocblas_int m, rocblas_int n,
                            rocblas_int k, const double* A, rocblas_int lda, rocblas_stride sa,
                            const double* B, rocblas_int ldb, rocblas_stride sb, double* C,
                            rocblas_int ldc, rocblas_stride sc, rocblas_int batch) {
  const double one = 1.0, zero = 0.0;
  return rocblas_dgemm_strided_batched(h, ta, rocblas_operation_none, m, n, k, &one, A, lda, sa,
                                       B, ldb, sb, &zero, C, ldc, sc, batch);
}

#define SOPHT_CHECK_BLAS(expr)                                                          \
  do {                                                                                  \
    rocblas_status _st = (expr);                                                        \
    if (_st != rocblas_status_success) {                                                \
      ::sopht::set_last_error(rocblas_status_to_string(_st), __FILE__, __LINE__);       \
      return SOPHT_ERR_HIP;                                                             \
    }                                                                                   \
  } while (0)

// out[.., k, ..] = sum_i M[k][i] in[.., i, ..] along one axis of ncomp contiguous
// (nz, ny, nx) fields spaced nz*ny*nx apart; M is a row-major N x N device matrix.
// Row-major data read column-major: x is the leading dimension.
template <typename T>
int apply_axis(sopht_fastdiag3d* P, int axis, const T* M, const T* in, T* out, int ncomp) {
  const int64_t nz = P->n[0], ny = P->n[1], nx = P->n[2];
  const int64_t vol = nz * ny * nx;
  rocblas_handle h = P->handle;
  if (axis == 2) {
    // columns of length nx: Out = M * X, with M read column-major as M^T -> op = transpose
    SOPHT_CHECK_BLAS(gemm<T>(h, rocblas_operation_transpose, rocblas_int(nx),
                             rocblas_int(ncomp * nz * ny), rocblas_int(nx), M, rocblas_int(nx), 0,
                             in, rocblas_int(nx), 0, out, rocblas_int(nx), 0, 1));
  } else if (axis == 1) {
    // per (component, z) slab, column-major (nx x ny): Out = Y * M^T, M^T read as M
    SOPHT_CHECK_BLAS(gemm<T>(h, rocblas_operation_none, rocblas_int(nx), rocblas_int(ny),
                             rocblas_int(ny), in, rocblas_int(nx), ny * nx, M, rocblas_int(ny), 0,
                             out, rocblas_int(nx), ny * nx, rocblas_int(ncomp * nz)));
  } else {
    // per component, column-major (ny*nx x nz): Out = Z * M^T
    SOPHT_CHECK_BLAS(gemm<T>(h, rocblas_operation_none, rocblas_int(ny * nx), rocblas_int(nz),
                             rocblas_int(nz), in, rocblas_int(ny * nx), vol, M, rocblas_int(nz), 0,
                             out, rocblas_int(ny * nx), vol, rocblas_int(ncomp)));
  }
  return SOPHT_OK;
}

// Forward x, y, z (:155-170), spectral scaling (:172-177), backward x, y, z (:179-189).
// rhs is read only by the first GEMM, so sol may alias it. Stage marks first..first+4.
template <typename T>
int solve_impl(sopht_fastdiag3d* P, T* sol, const T* rhs, int ncomp, hipStream_t s,
               StageTimer* tm, int first) {
  SOPHT_CHECK_BLAS(rocblas_set_stream(P->handle, s));
  T* w = static_cast<T*>(P->work);
  const T* const* F = reinterpret_cast<const T* const*>(P->fwd);
  const T* const* B = reinterpret_cast<const T* const*>(P->bwd);
  int rc;
  if ((rc = apply_axis<T>(P, 2, F[2], rhs, w, ncomp))) return rc;
  if (tm) tm->mark(first, s);
  if ((rc = apply_axis<T>(P, 1, F[1], w, sol, ncomp))) return rc;
  if (tm) tm->mark(first + 1, s);
  if ((rc = apply_axis<T>(P, 0, F[0], sol, w, ncomp))) return rc;
  {
    const dim3 grid(ceil_div(P->n[1], kScaleRows), unsigned(ncomp * P->n[0]));
    const T* lz = static_cast<const T*>(P->lam_dev[0]);
    const T* ly = static_cast<const T*>(P->lam_dev[1]);
    const T* lx = static_cast<const T*>(P->lam_dev[2]);
    constexpr int V = 16 / sizeof(T);
    if (P->n[2] % V == 0)  // rows of whole 16-byte packs: every plane starts 16-byte aligned
      k_spectral_scale<T, V><<<grid, 256, 0, s>>>(w, lz, ly, lx, int(P->n[0]), int(P->n[1]), int(P->n[2]));
    else
      k_spectral_scale<T, 1><<<grid, 256, 0, s>>>(w, lz, ly, lx, int(P->n[0]), int(P->n[1]), int(P->n[2]));
  }
  SOPHT_CHECK_LAUNCH();
  if (tm) tm->mark(first + 2, s);
  if ((rc = apply_axis<T>(P, 2, B[2], w, sol, ncomp))) return rc;
  if (tm) tm->mark(first + 3, s);
  if ((rc = apply_axis<T>(P, 1, B[1], sol, w, ncomp))) return rc;
  if ((rc = apply_axis<T>(P, 0, B[0], w, sol, ncomp))) return rc;
  if (tm) tm->mark(first + 4, s);
  return SOPHT_OK;
}

bool canonical(const int64_t* st, const sopht_fastdiag3d* P) {
  return st[2] == 1 && st[1] == P->n[2] && st[0] == P->n[1] * P->n[2];
}

template <typename T>
int upload(void** dst, const std::vector<double>& src) {
  std::vector<T> h(src.begin(), src.end());
  SOPHT_CHECK_HIP(hipMalloc(dst, h.size() * sizeof(T)));
  SOPHT_CHECK_HIP(hipMemcpy(*dst, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return SOPHT_OK;
}

template <typename T>
int create_device(sopht_fastdiag3d* P) {
  for (int a = 0; a < 3; ++a) {
    const int64_t N = P->n[a];
    std::vector<double> f(N * N), b(N * N);
    for (int64_t i = 0; i < N; ++i)
      for (int64_t p = 0; p < N; ++p) {
        const double v = P->vec[a][i * N + p];
        f[p * N + i] = v;
        b[i * N + p] = v;
      }
    int rc;
    if ((rc = upload<T>(&P->fwd[a], f))) return rc;
    if ((rc = upload<T>(&P->bwd[a], b))) return rc;
    if ((rc = upload<T>(&P->lam_dev[a], P->lam[a]))) return rc;
  }
  SOPHT_CHECK_HIP(hipMalloc(&P->work, 3 * P->n[0] * P->n[1] * P->n[2] * sizeof(T)));
  return SOPHT_OK;
}

}  // namespace

int fastdiag_vector_solve_timed(sopht_fastdiag3d* P, void* const sol[3], const int64_t* ss,
                                const void* const rhs[3], const int64_t* rs, hipStream_t s,
                                StageTimer* tm, int first) {
  SOPHT_REQUIRE(P != nullptr, "plan is null");
  SOPHT_REQUIRE(sol && rhs && ss && rs, "null argument");
  for (int c = 0; c < 3; ++c) SOPHT_REQUIRE(sol[c] && rhs[c], "null field pointer");
  SOPHT_REQUIRE(canonical(ss, P) && canonical(rs, P),
                "fast-diagonalisation solve needs C-contiguous (nz, ny, nx) components");
  const size_t es = P->dtype == SOPHT_F32 ? 4 : 8;
  const int64_t vb = P->n[0] * P->n[1] * P->n[2] * int64_t(es);
  auto spaced = [vb](const char* p0, const char* p1, const char* p2) {
    return p1 - p0 == vb && p2 - p1 == vb;
  };
  const bool batched = spaced(static_cast<const char*>(sol[0]), static_cast<const char*>(sol[1]),
                              static_cast<const char*>(sol[2])) &&
                       spaced(static_cast<const char*>(rhs[0]), static_cast<const char*>(rhs[1]),
                              static_cast<const char*>(rhs[2]));
  if (batched) {
    return P->dtype == SOPHT_F32
               ? solve_impl<float>(P, static_cast<float*>(sol[0]), static_cast<const float*>(rhs[0]), 3, s, tm, first)
               : solve_impl<double>(P, static_cast<double*>(sol[0]), static_cast<const double*>(rhs[0]), 3, s, tm, first);
  }
  for (int c = 0; c < 3; ++c) {
    const int rc = P->dtype == SOPHT_F32
        ? solve_impl<float>(P, static_cast<float*>(sol[c]), static_cast<const float*>(rhs[c]), 1, s,
                            c == 2 ? tm : nullptr, first)
        : solve_impl<double>(P, static_cast<double*>(sol[c]), static_cast<const double*>(rhs[c]), 1, s,
                             c == 2 ? tm : nullptr, first);
    if (rc) return rc;
  }
  return SOPHT_OK;
}

}  // namespace sopht

using namespace sopht;

extern "C" {

int sopht_fastdiag3d_create(int dtype, int64_t nz, int64_t ny, int64_t nx, double dx, void* stream,
                            sopht_fastdiag3d** out) {
  SOPHT_REQUIRE(out != nullptr, "plan pointer is null");
  *out = nullptr;
  SOPHT_REQUIRE(dtype == SOPHT_F32 || dtype == SOPHT_F64, "dtype must be SOPHT_F32 or SOPHT_F64");
  SOPHT_REQUIRE(nz >= 2 && ny >= 2 && nx >= 2, "grid must be at least 2 cells per axis");
  SOPHT_REQUIRE(nz <= 2048 && ny <= 2048 && nx <= 2048, "grid sizes above 2048 per axis are not supported");
  SOPHT_REQUIRE(3 * nz * ny * nx < (int64_t(1) << 31), "grid too large (3 nz ny nx must be below 2^31)");
  SOPHT_REQUIRE(dx > 0.0 && std::isfinite(dx), "dx must be positive");
  (void)stream;  // set-up copies are synchronous
  sopht_fastdiag3d* P = new (std::nothrow) sopht_fastdiag3d();
  SOPHT_REQUIRE(P != nullptr, "out of host memory");
  P->dtype = dtype;
  P->n[0] = nz; P->n[1] = ny; P->n[2] = nx;
  P->dx = dx;
  const double inv_dx2 = 1.0 / dx / dx;
  for (int a = 0; a < 3; ++a) {
    const int64_t N = P->n[a];
    P->lam[a].resize(N);
    P->vec[a].resize(N * N);
    for (int64_t p = 0; p < N; ++p) {
      const int64_t k = N - 1 - p;  // decreasing eigenvalues (:105-107)
      const double half = M_PI * double(k) / (2.0 * double(N));
      const double sk = std::sin(half);
      P->lam[a][p] = 4.0 * inv_dx2 * sk * sk;
      const double ck = k == 0 ? std::sqrt(1.0 / double(N)) : std::sqrt(2.0 / double(N));
      for (int64_t i = 0; i < N; ++i)
        P->vec[a][i * N + p] = ck * std::cos(M_PI * double(k) * (double(i) + 0.5) / double(N));
    }
  }
  int rc = dtype == SOPHT_F32 ? create_device<float>(P) : create_device<double>(P);
  if (rc == SOPHT_OK && rocblas_create_handle(&P->handle) != rocblas_status_success) {
    set_last_error("rocblas_create_handle failed", __FILE__, __LINE__);
    rc = SOPHT_ERR_HIP;
  }
  if (rc) {
    sopht_fastdiag3d_destroy(P);
    return rc;
  }
  *out = P;
  return SOPHT_OK;
}

int sopht_fastdiag3d_solve(sopht_fastdiag3d* P, void* solution, const int64_t* solution_strides,
                           const void* rhs, const int64_t* rhs_strides, void* stream) {
  SOPHT_REQUIRE(P != nullptr, "plan is null");
  SOPHT_REQUIRE(solution && rhs && solution_strides && rhs_strides, "null argument");
  SOPHT_REQUIRE(canonical(solution_strides, P) && canonical(rhs_strides, P),
                "fast-diagonalisation solve needs C-contiguous (nz, ny, nx) fields");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return P->dtype == SOPHT_F32
             ? solve_impl<float>(P, static_cast<float*>(solution), static_cast<const float*>(rhs), 1, s, nullptr, 0)
             : solve_impl<double>(P, static_cast<double*>(solution), static_cast<const double*>(rhs), 1, s, nullptr, 0);
}

int sopht_fastdiag3d_vector_solve(sopht_fastdiag3d* P, void* const solution[3],
                                  const int64_t* solution_strides, const void* const rhs[3],
                                  const int64_t* rhs_strides, void* stream) {
  return fastdiag_vector_solve_timed(P, solution, solution_strides, rhs, rhs_strides,
                                     static_cast<hipStream_t>(stream), nullptr, 0);
}

int sopht_fastdiag3d_eigenbasis(const sopht_fastdiag3d* P, int axis, double* eig_vals,
                                double* eig_vecs) {
  SOPHT_REQUIRE(P != nullptr, "plan is null");
  SOPHT_REQUIRE(axis >= 0 && axis <= 2, "axis must be 0 (z), 1 (y) or 2 (x)");
  const int64_t N = P->n[axis];
  if (eig_vals)
    for (int64_t p = 0; p < N; ++p) eig_vals[p] = P->lam[axis][p];
  if (eig_vecs)
    for (int64_t i = 0; i < N * N; ++i) eig_vecs[i] = P->vec[axis][i];
  return SOPHT_OK;
}

int sopht_fastdiag3d_destroy(sopht_fastdiag3d* P) {
  if (!P) return SOPHT_OK;
  if (P->handle) (void)rocblas_destroy_handle(P->handle);
  for (int a = 0; a < 3; ++a) {
    if (P->fwd[a]) (void)hipFree(P->fwd[a]);
    if (P->bwd[a]) (void)hipFree(P->bwd[a]);
    if (P->lam_dev[a]) (void)hipFree(P->lam_dev[a]);
  }
  if (P->work) (void)hipFree(P->work);
  delete P;
  return SOPHT_OK;
}

}  // extern "C"
