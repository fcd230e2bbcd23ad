// Plan of the fused 3D flow step (sopht_ns3d), shared by the step (flow_step_3d.hip) and its
// adjoint (adjoint_3d.hip).
#pragma once

#include "common.h"
#include "sweeps_3d.h"

struct sopht_ns3d {
  int dtype;
  int64_t nz, ny, nx;
  double x_range, dx, nu, rho;
  int width, with_forcing, with_free_stream;
  sopht_poisson3d* poisson;     // free-space solver (the default); null when fastdiag is set
  sopht_fastdiag3d* fastdiag;   // Neumann fast-diagonalisation solver, or null
  double sine[6 * 8];  // x, y, z: width front factors then width back factors
  void* partial;       // per-block partial maxima for the CFL reduction
  int64_t partial_cap;
  void* s2max;         // per-workgroup maxima of sum_i |u_i| written by the last S2 sweep
  int64_t s2max_cap;
  int64_t s2_partials; // how many (0: none)
  const void* s2_velocity;  // the velocity field that S2 wrote
  int* s1flags;        // per S1 tile: did it see a bit of F set (SC zeroes F only there)
  int64_t s1flags_cap;
  int filter_order;    // vorticity filter: order (< 0: off) and type (0 multiplicative, 1 convolution)
  int filter_type;
  int s1_vec;          // S1 with 16-byte accesses (rows a multiple of 16 bytes), else the scalar sweep
  int s1_chunk;        // z planes per work item of the vector S1 sweep (16, 32 or 64)
  int s1_cols;         // its tile: s1_cols cells (64, or 128 for FP32) by s1_rows rows
  int s1_rows;
  sopht::StageTimer timer;
};

namespace sopht {
namespace {

Pen make_pen(const sopht_ns3d* P) {
  Pen pen;
  pen.w = P->width;
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < 2 * P->width; ++i) pen.f[a][i] = P->sine[a * 2 * P->width + i];
  return pen;
}

}  // namespace
}  // namespace sopht
