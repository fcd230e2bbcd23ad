// Shared definitions for the sopht_amd HIP library (gfx950 / CDNA4).
//
// Field layout: C-contiguous (3, Nz, Ny, Nx) vector fields or (Nz, Ny, Nx) scalar fields,
// x fastest. Standalone drop-in kernels take element strides (z, y, x) per field so
// strided views behave as in the reference; the fused step kernels assume contiguous
// fields with the component stride Nz*Ny*Nx.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sopht_amd.h"

namespace sopht {

// Return-code helpers ------------------------------------------------------------------
#define SOPHT_CHECK_HIP(expr)                                  \
  do {                                                         \
    hipError_t _e = (expr);                                    \
    if (_e != hipSuccess) {                                    \
      ::sopht::set_last_error(hipGetErrorString(_e), __FILE__, __LINE__); \
      return SOPHT_ERR_HIP;                                    \
    }                                                          \
  } while (0)

#define SOPHT_CHECK_LAUNCH() SOPHT_CHECK_HIP(hipGetLastError())

#define SOPHT_REQUIRE(cond, msg)                               \
  do {                                                         \
    if (!(cond)) {                                             \
      ::sopht::set_last_error((msg), __FILE__, __LINE__);      \
      return SOPHT_ERR_INVALID;                                \
    }                                                          \
  } while (0)

void set_last_error(const char* msg, const char* file, int line);

// Strided 3D view --------------------------------------------------------------------------
template <typename T>
struct View3 {
  T* p;
  int64_t sz, sy, sx;
  __device__ __forceinline__ T& operator()(int64_t z, int64_t y, int64_t x) const {
    return p[z * sz + y * sy + x * sx];
  }
};

template <typename T>
inline View3<T> make_view(void* p, const int64_t* s) {
  View3<T> v;
  v.p = static_cast<T*>(p);
  v.sz = s[0];
  v.sy = s[1];
  v.sx = s[2];
  return v;
}
template <typename T>
inline View3<T> make_view(const void* p, const int64_t* s) {
  return make_view<T>(const_cast<void*>(p), s);
}

// Complex arithmetic in registers -----------------------------------------------------------
template <typename T>
struct alignas(2 * sizeof(T)) Cplx {  // aligned as a pair: one 64/128-bit access, not two
  T re, im;
};

template <typename T>
__device__ __forceinline__ Cplx<T> cadd(Cplx<T> a, Cplx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ Cplx<T> csub(Cplx<T> a, Cplx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__device__ __forceinline__ Cplx<T> cmul(Cplx<T> a, Cplx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ Cplx<T> cmulc(Cplx<T> a, Cplx<T> b) {
  return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
template <typename T>
__device__ __forceinline__ Cplx<T> cscale(Cplx<T> a, T s) { return {a.re * s, a.im * s}; }
template <typename T>
__device__ __forceinline__ Cplx<T> cconj(Cplx<T> a) { return {a.re, -a.im}; }
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV>
__device__ __forceinline__ Cplx<T> mul_mi(Cplx<T> a) {
  return INV ? Cplx<T>{-a.im, a.re} : Cplx<T>{a.im, -a.re};
}

inline int ceil_div(int64_t a, int64_t b) { return static_cast<int>((a + b - 1) / b); }

// Per-stage timing with HIP events recorded on the launch stream between consecutive
// kernels of a step. Two event sets alternate between steps, so harvesting the set of
// step k-2 before reusing it never waits on work that is still queued behind step k-1.
struct StageTimer {
  static constexpr int kMax = 16;
  hipEvent_t ev[2][kMax + 1] = {};
  bool pending[2] = {false, false};
  bool enabled = false;
  int nstages = 0;
  int set = 0;
  double acc[kMax] = {};
  int64_t steps = 0;

  int enable(int n) {
    if (enabled) return SOPHT_OK;
    nstages = n;
    for (int k = 0; k < 2; ++k)
      for (int i = 0; i <= n; ++i)
        if (hipEventCreate(&ev[k][i]) != hipSuccess) return SOPHT_ERR_HIP;
    enabled = true;
    return SOPHT_OK;
  }
  void disable() {
    if (!enabled) return;
    harvest(0);
    harvest(1);
    for (int k = 0; k < 2; ++k)
      for (int i = 0; i <= nstages; ++i) (void)hipEventDestroy(ev[k][i]);
    enabled = false;
  }
  void harvest(int k) {
    if (!pending[k]) return;
    (void)hipEventSynchronize(ev[k][nstages]);
    for (int i = 0; i < nstages; ++i) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, ev[k][i], ev[k][i + 1]);
      acc[i] += ms;
    }
    ++steps;
    pending[k] = false;
  }
  void begin(hipStream_t s) {
    if (!enabled) return;
    set ^= 1;
    harvest(set);
    (void)hipEventRecord(ev[set][0], s);
  }
  // stage i (0-based) has just been launched
  void mark(int i, hipStream_t s) {
    if (enabled) (void)hipEventRecord(ev[set][i + 1], s);
  }
  void end() {
    if (enabled) pending[set] = true;
  }
};

}  // namespace sopht

struct sopht_poisson3d;
namespace sopht {
// Vector solve with stage marks first_stage .. first_stage+4 after its five kernels.
int poisson_vector_solve_timed(sopht_poisson3d* P, void* const sol[3], const int64_t* ss,
                               const void* const rhs[3], const int64_t* rs, hipStream_t s,
                               StageTimer* timer, int first_stage);
// The Laplacian filter of one scalar field (gen_laplacian_filter_kernel_3d), type 0
// multiplicative, 1 convolution, with the reference's buffers and kernel sequence.
int laplacian_filter_field(int dtype, int order, int type, void* field, const int64_t* fs,
                           int64_t nz, int64_t ny, int64_t nx, void* flux, const int64_t* xs,
                           void* fbuf, const int64_t* bs, int bw, hipStream_t s);
// The step's Laplacian vorticity filter on the three contiguous components of w: one launch
// per axis (buf: three scratch components for the multiplicative filter's passes). The
// standalone factory's kernel sequence covers the orders and rows this does not.
bool fused_filter_supported(int order, int64_t nx);
int fused_vorticity_filter(int dtype, int order, int type, void* w, void* buf, int64_t nz, int64_t ny,
                           int64_t nx, hipStream_t s);
// Fast-diagonalisation (Neumann) vector solve with stage marks first .. first+4.
int fastdiag_vector_solve_timed(sopht_fastdiag3d* P, void* const sol[3], const int64_t* ss,
                                const void* const rhs[3], const int64_t* rs, hipStream_t s,
                                StageTimer* tm, int first);
// Launches (timed stages) of one Poisson solve: 3 (x and y passes fused) or 5.
int poisson_stage_count(const sopht_poisson3d* P);
// As the vector solve, adding the solution into sol (sol += G rhs).
int poisson_vector_solve_accumulate(sopht_poisson3d* P, void* const sol[3], const int64_t* ss,
                                    const void* const rhs[3], const int64_t* rs, hipStream_t s);
// A plan whose slab holds `comps` components (one Poisson plan per grid size of a batch).
int poisson_create(int dtype, int64_t nz, int64_t ny, int64_t nx, double x_range, int comps,
                   void* stream, sopht_poisson3d** out);
// Solve ncomp components whose fields come from device tables of pointers (sol_tab[c],
// rhs_tab[c]); all share the element strides ss and rs.
int poisson_table_solve(sopht_poisson3d* P, int ncomp, void* const* sol_tab, const int64_t* ss,
                        const void* const* rhs_tab, const int64_t* rs, hipStream_t s);
}  // namespace sopht
