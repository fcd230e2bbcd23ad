/*
 * sopht_amd.h — C ABI of the MI355X (gfx950) backend for SophT's 3D unbounded vorticity
 * flow step with virtual-boundary immersed-boundary coupling.
 *
 * Conventions
 *   - dtype: SOPHT_F32 or SOPHT_F64 (the reference's real_t = np.float32 / np.float64).
 *   - Grid sizes are (nz, ny, nx); fields are x-fastest. A "strides" argument is three
 *     int64 element strides (z, y, x) of a scalar field or of every component of a
 *     vector field; vector fields are passed as three component pointers.
 *   - stream: a hipStream_t passed as void* (0 = the null stream). Every call is
 *     stream-ordered and asynchronous unless stated; no call synchronises the device
 *     unless stated.
 *   - Return value: 0 on success, a negative SOPHT_ERR_* code otherwise;
 *     sopht_last_error() describes the last failure of the calling thread.
 *   - The library holds no global mutable state besides the per-thread error string;
 *     plans own their device scratch and are not shared across streams concurrently.
 *
 * Each entry point names the reference (SophT) interface it replaces as file:line,
 * paths relative to the SophT repository root.
 */
#ifndef SOPHT_AMD_H
#define SOPHT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SOPHT_F32 = 0, SOPHT_F64 = 1 };
enum {
  SOPHT_OK = 0,
  SOPHT_ERR_INVALID = -1,     /* bad argument (maps to ValueError)            */
  SOPHT_ERR_UNSUPPORTED = -2, /* size or option not supported (ValueError)    */
  SOPHT_ERR_HIP = -3,         /* HIP runtime error (RuntimeError)             */
  SOPHT_ERR_OUT_OF_DOMAIN = -4 /* Lagrangian support leaves the Eulerian grid */
};

const char* sopht_last_error(void);
int sopht_version(void);

/* ---------------------------------------------------------------- stencils (grid ops) */

/* gen_curl_pyst_kernel_3d — sopht/numeric/eulerian_grid_ops/stencil_ops_3d/curl_3d.py:13-132.
 * curl[1:-1]^3 = prefactor * curl_h(field); border set to 0 when reset_ghost_zone. */
int sopht_curl_3d(int dtype, void* const curl[3], const int64_t* curl_strides,
                  const void* const field[3], const int64_t* field_strides, int64_t nz,
                  int64_t ny, int64_t nx, double prefactor, int reset_ghost_zone, void* stream);

/* gen_update_vorticity_from_velocity_forcing_pyst_kernel_3d —
 * .../update_vorticity_from_velocity_forcing_3d.py:12-132. w[1:-1]^3 += p * curl_h(f). */
int sopht_update_vorticity_from_velocity_forcing_3d(int dtype, void* const vorticity[3],
                                                    const int64_t* vorticity_strides,
                                                    const void* const forcing[3],
                                                    const int64_t* forcing_strides, int64_t nz,
                                                    int64_t ny, int64_t nx, double prefactor,
                                                    void* stream);

/* gen_update_vorticity_from_penalised_velocity_pyst_kernel_3d —
 * .../update_vorticity_from_velocity_forcing_3d.py:135-291. w[1:-1]^3 += p*curl_h(u_pen - u). */
int sopht_update_vorticity_from_penalised_velocity_3d(
    int dtype, void* const vorticity[3], const int64_t* vorticity_strides,
    const void* const penalised_velocity[3], const int64_t* penalised_strides,
    const void* const velocity[3], const int64_t* velocity_strides, int64_t nz, int64_t ny,
    int64_t nx, double prefactor, void* stream);

/* gen_elementwise_cross_product_pyst_kernel_3d — .../elementwise_ops_3d.py:390-449. */
int sopht_elementwise_cross_product_3d(int dtype, void* const result[3],
                                       const int64_t* result_strides,
                                       const void* const field_1[3], const int64_t* f1_strides,
                                       const void* const field_2[3], const int64_t* f2_strides,
                                       int64_t nz, int64_t ny, int64_t nx, void* stream);

/* gen_diffusion_flux_pyst_kernel_3d (scalar) — .../diffusion_flux_3d.py:14-75.
 * flux[1:-1]^3 = p*(sum of 6 neighbours - 6 c); border 0 when reset_ghost_zone. */
int sopht_diffusion_flux_3d(int dtype, void* flux, const int64_t* flux_strides,
                            const void* field, const int64_t* field_strides, int64_t nz,
                            int64_t ny, int64_t nx, double prefactor, int reset_ghost_zone,
                            void* stream);

/* gen_diffusion_timestep_euler_forward_pyst_kernel_3d (scalar) —
 * .../diffusion_timestep_3d.py:12-80. flux (border reset) then field += flux. */
int sopht_diffusion_timestep_euler_forward_3d(int dtype, void* field,
                                              const int64_t* field_strides, void* flux,
                                              const int64_t* flux_strides, int64_t nz,
                                              int64_t ny, int64_t nx, double nu_dt_by_dx2,
                                              void* stream);

/* gen_penalise_field_boundary_pyst_kernel_3d (scalar) —
 * .../penalise_field_boundary_3d.py:13-240. In place. sine_tables holds, for x then y then
 * z, `width` front factors followed by `width` back factors (6*width doubles, host
 * memory): sin(pi/(2 w dx) * distance of the slab cell from the first/last cell centre). */
int sopht_penalise_field_boundary_3d(int dtype, void* field, const int64_t* field_strides,
                                     int64_t nz, int64_t ny, int64_t nx, int width,
                                     const double* sine_tables, void* stream);

/* gen_divergence_pyst_kernel_3d — .../divergence_3d.py:13-96. */
int sopht_divergence_3d(int dtype, void* divergence, const int64_t* divergence_strides,
                        const void* const field[3], const int64_t* field_strides, int64_t nz,
                        int64_t ny, int64_t nx, double inv_dx, int reset_ghost_zone,
                        void* stream);

/* Elementwise family — .../elementwise_ops_3d.py: set_fixed_val :60-119, add_fixed_val
 * :271-334, elementwise_sum :13-57, elementwise_copy :122-143, saxpby :337-387,
 * set_fixed_val_at_boundaries :200-268, complex_product :146-197 (interleaved complex). */
int sopht_set_fixed_val_3d(int dtype, void* field, const int64_t* strides, int64_t nz,
                           int64_t ny, int64_t nx, double value, void* stream);
int sopht_add_fixed_val_3d(int dtype, void* sum_field, const int64_t* sum_strides,
                           const void* field, const int64_t* field_strides, int64_t nz,
                           int64_t ny, int64_t nx, double value, void* stream);
int sopht_elementwise_saxpby_3d(int dtype, void* sum_field, const int64_t* sum_strides,
                                const void* field_1, const int64_t* f1_strides,
                                const void* field_2, const int64_t* f2_strides, int64_t nz,
                                int64_t ny, int64_t nx, double field_1_prefac,
                                double field_2_prefac, void* stream);
int sopht_elementwise_sum_3d(int dtype, void* sum_field, const int64_t* sum_strides,
                             const void* field_1, const int64_t* f1_strides,
                             const void* field_2, const int64_t* f2_strides, int64_t nz,
                             int64_t ny, int64_t nx, void* stream);
int sopht_elementwise_copy_3d(int dtype, void* field, const int64_t* strides,
                              const void* rhs_field, const int64_t* rhs_strides, int64_t nz,
                              int64_t ny, int64_t nx, void* stream);
int sopht_set_fixed_val_at_boundaries_3d(int dtype, void* field, const int64_t* strides,
                                         int64_t nz, int64_t ny, int64_t nx, int width,
                                         double value, void* stream);
int sopht_elementwise_complex_product_3d(int dtype, void* product, const int64_t* product_strides,
                                         const void* field_1, const int64_t* f1_strides,
                                         const void* field_2, const int64_t* f2_strides,
                                         int64_t nz, int64_t ny, int64_t nx, void* stream);

/* PassiveTransportFlowSimulator._advection_and_diffusion_time_step —
 * sopht/simulator/flow/passive_transport_flow_simulators.py:100-113 (3D): the conservative
 * ENO3 advection step (advection_timestep_3d.py:12-99, inv_dx = -dt_by_dx) followed by the
 * Euler diffusion step (diffusion_timestep_3d.py:12-80, prefactor nu_dt_by_dx2) of ncomp
 * (1 or 3) contiguous (nz, ny, nx) component fields, in place, in one fused sweep; scratch
 * holds ncomp fields of the same shape (contents undefined on return). velocity: three
 * contiguous components. */
int sopht_passive_transport_step_3d(int dtype, int ncomp, void* const field[3], void* const scratch[3],
                                    const void* const velocity[3], int64_t nz, int64_t ny,
                                    int64_t nx, double dt_by_dx, double nu_dt_by_dx2, void* stream);

/* compute_advection_diffusion_stable_timestep's reduction —
 * sopht/simulator/flow/passive_transport_flow_simulators.py:139-155: magnitude (optional, n
 * elements) = |u0| + |u1| + |u2| per cell, *out = its maximum (real_t, on the device), over n
 * contiguous cells; partial holds partial_cap real_t words of scratch. */
int sopht_velocity_abs_sum_max_3d(int dtype, const void* const velocity[3], int64_t n, void* magnitude,
                                  void* partial, int64_t partial_cap, void* out, void* stream);

/* gen_advection_flux_conservative_eno3_pyst_kernel_3d — .../advection_flux_3d.py:12-233.
 * Accumulates into flux on [2:-2]^3 (six sub-stencils in the reference order). */
int sopht_advection_flux_conservative_eno3_3d(int dtype, void* flux, const int64_t* flux_strides,
                                              const void* field, const int64_t* field_strides,
                                              const void* const velocity[3],
                                              const int64_t* velocity_strides, int64_t nz,
                                              int64_t ny, int64_t nx, double inv_dx,
                                              void* stream);

/* Adjoint of gen_advection_flux_conservative_eno3_pyst_kernel_3d
 * (.../advection_flux_3d.py:12-233; the reference has no adjoint): the transpose of the flux
 * with every face's upwind stencil choice frozen at `velocity`. For a flux adjoint read on
 * [2:-2]^3, accumulates field_adj += J_field^T flux_adj and velocity_adj[c] += J_u_c^T flux_adj
 * over the whole grid. Either output may be NULL. Deterministic (gather, no atomics). */
int sopht_advection_flux_conservative_eno3_adjoint_3d(
    int dtype, void* field_adj, const int64_t* field_adj_strides, void* const velocity_adj[3],
    const int64_t* velocity_adj_strides, const void* flux_adj, const int64_t* flux_adj_strides,
    const void* field, const int64_t* field_strides, const void* const velocity[3],
    const int64_t* velocity_strides, int64_t nz, int64_t ny, int64_t nx, double inv_dx,
    void* stream);

/* gen_vorticity_stretching_flux_pyst_kernel_3d — .../vorticity_stretching_flux_3d.py:13-107. */
int sopht_vorticity_stretching_flux_3d(int dtype, void* const flux[3], const int64_t* flux_strides,
                                       const void* const vorticity[3],
                                       const int64_t* vorticity_strides,
                                       const void* const velocity[3],
                                       const int64_t* velocity_strides, int64_t nz, int64_t ny,
                                       int64_t nx, double prefactor, void* stream);

/* ---------------------------------------------------------------- free-space Poisson */

typedef struct sopht_poisson3d sopht_poisson3d;

/* UnboundedPoissonSolverPYFFTW3D.__init__ — .../poisson_solver_3d/UnboundedPoissonSolverPYFFTW3D.py:12-54.
 * Builds the Green's function octant (FP64 on the device) and the pass scratch. Lengths
 * 2*n per axis must factor into 2, 3 and 5 with 8 <= 2n <= 1024. */
int sopht_poisson3d_create(int dtype, int64_t nz, int64_t ny, int64_t nx, double x_range,
                           void* stream, sopht_poisson3d** plan);
/* solve — :111-149 (scalar). rhs is read only. */
int sopht_poisson3d_solve(sopht_poisson3d* plan, void* solution, const int64_t* solution_strides,
                          const void* rhs, const int64_t* rhs_strides, void* stream);
/* vector_field_solve — :151-172. Solves the three components in one pass sequence. */
int sopht_poisson3d_vector_solve(sopht_poisson3d* plan, void* const solution[3],
                                 const int64_t* solution_strides, const void* const rhs[3],
                                 const int64_t* rhs_strides, void* stream);
int sopht_poisson3d_destroy(sopht_poisson3d* plan);
/* Bytes of device scratch the plan holds (for capacity planning). */
int64_t sopht_poisson3d_scratch_bytes(const sopht_poisson3d* plan);
/* The plan's FFT schedule (no counterpart in the reference, which leaves it to FFTW's
 * planner): *passes is 3 (x transform fused with the inner y sums; FP32, nx = ny) or 5, and
 * *zpass_jobs the jobs per workgroup of the three-pass z pass, timed over {1, 2, 3, 6} when
 * the plan was created (SOPHT_POISSON_ZJOBS in the environment fixes it; 0 with five passes).
 * Results do not depend on it. Either pointer may be null. */
int sopht_poisson3d_schedule(const sopht_poisson3d* plan, int* passes, int* zpass_jobs);

/* ---------------------------------------------------------------- Neumann Poisson (fast diagonalisation) */

typedef struct sopht_fastdiag3d sopht_fastdiag3d;

/* FastDiagPoissonSolver3D.__init__ — .../poisson_solver_3d/FastDiagPoissonSolver3D.py:15-145,
 * bc_type "homogenous_neumann_along_xyz". The per-axis eigenbasis of the Neumann
 * second-difference matrix is the cell-centred cosine basis, built in closed form (double),
 * orthonormal, eigenvalues in decreasing order. 2 <= n <= 2048 per axis. */
int sopht_fastdiag3d_create(int dtype, int64_t nz, int64_t ny, int64_t nx, double dx, void* stream,
                            sopht_fastdiag3d** plan);
/* solve — :147-189: -lap(solution) = rhs with the constant mode removed. Both fields
 * C-contiguous (nz, ny, nx); rhs is read only; solution may alias rhs. */
int sopht_fastdiag3d_solve(sopht_fastdiag3d* plan, void* solution, const int64_t* solution_strides,
                           const void* rhs, const int64_t* rhs_strides, void* stream);
/* vector_field_solve — :191-202. Components C-contiguous; components spaced nz*ny*nx apart
 * (a (3, nz, ny, nx) tensor) are solved in one batched sequence. */
int sopht_fastdiag3d_vector_solve(sopht_fastdiag3d* plan, void* const solution[3],
                                  const int64_t* solution_strides, const void* const rhs[3],
                                  const int64_t* rhs_strides, void* stream);
/* Host copy of one axis' eigenbasis (axis 0 z, 1 y, 2 x): eig_vals (n), eig_vecs row-major
 * (n, n) with column p the eigenvector of eig_vals[p] (the reference's eig_vecs_*). Either
 * output may be null. */
int sopht_fastdiag3d_eigenbasis(const sopht_fastdiag3d* plan, int axis, double* eig_vals,
                                double* eig_vecs);
int sopht_fastdiag3d_destroy(sopht_fastdiag3d* plan);

/* ---------------------------------------------------------------- Eulerian <-> Lagrangian */

/* local_eulerian_grid_support_of_lagrangian_grid_kernel_3d —
 * sopht/numeric/immersed_boundary_ops/EulerianLagrangianGridCommunicator3D.py:68-135.
 * lag_positions: (3, n) float64, device. nearest_index: (3, n) int64. support:
 * (3, 4, 4, 4, n) real. */
int sopht_lag_support_3d(int dtype, const double* lag_positions, int64_t* nearest_index,
                         void* support, int64_t num_lag_nodes, double dx,
                         double eul_grid_coord_shift, void* stream);

/* cosine / peskin interpolation weights — :387-518. kernel_type 0 = cosine, 1 = peskin.
 * support is divided by dx in place (cosine, :404) or replaced by |d|/dx (peskin, :432). */
int sopht_lag_interp_weights_3d(int dtype, int kernel_type, void* weights, void* support,
                                int64_t num_lag_nodes, double dx, void* stream);

/* eulerian_to_lagrangian_grid_interpolation_kernel_3d — :180-299. n_components 1 or 3. */
int sopht_eul_to_lag_interp_3d(int dtype, int n_components, void* lag_field,
                               const void* const eul_field[3], const int64_t* eul_strides,
                               int64_t nz, int64_t ny, int64_t nx, const void* weights,
                               const int64_t* nearest_index, int64_t num_lag_nodes, double dx,
                               void* stream);

typedef struct sopht_spread3d sopht_spread3d;
/* Workspace for deterministic spreading of up to max_lag_nodes nodes. */
int sopht_spread3d_create(int64_t max_lag_nodes, sopht_spread3d** ws);
int sopht_spread3d_destroy(sopht_spread3d* ws);
/* lagrangian_to_eulerian_grid_interpolation_kernel_3d — :302-384. F[block(n)] += F_n w_n,
 * summed per cell in increasing node order (the reference's serial order) without float
 * atomics. Returns SOPHT_ERR_OUT_OF_DOMAIN (after the launch, on the next call that
 * synchronises) if a support block leaves the grid; such contributions are dropped. */
int sopht_lag_to_eul_spread_3d(int dtype, int n_components, sopht_spread3d* ws,
                               void* const eul_field[3], const int64_t* eul_strides, int64_t nz,
                               int64_t ny, int64_t nx, const void* lag_field,
                               const void* weights, const int64_t* nearest_index,
                               int64_t num_lag_nodes, void* stream);
/* Reads back (synchronising `stream`) and clears the out-of-domain flag of ws; returns
 * SOPHT_ERR_INVALID (and clears it) if a call declared by sopht_spread3d_reuse_plan found
 * changed support indices. */
int sopht_spread3d_check(sopht_spread3d* ws, void* stream);
/* No counterpart in the reference (its spreading has no plan): declares that the next
 * sopht_lag_to_eul_spread_3d through ws has the node count, grid and support indices of the
 * previous one (a body at rest), so that call launches only the key check and the replay,
 * not the sort and plan kernels. A key that changed anyway is detected on the device: that
 * call spreads nothing and the next sopht_spread3d_check returns SOPHT_ERR_INVALID. */
int sopht_spread3d_reuse_plan(sopht_spread3d* ws);

/* No counterpart in the reference: sopht_lag_to_eul_spread_3d for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, result and order of summation, with no workspace: the kernels
 * use a library-owned static device scratch of 8.5 MB per GPU (held until the process ends)
 * sized for at most 8192 Lagrangian nodes; larger n is rejected with SOPHT_ERR_INVALID. The
 * gather plan is rebuilt on every call (the scratch keeps only the last call's contents).
 * Calls that overlap on one GPU (several streams, host threads) race on the scratch: order
 * them on one stream. out_of_domain (optional device int) receives 1 when a node's 4^3
 * support left the grid (those contributions are dropped, as in the workspace form). The
 * call allocates nothing. */
int sopht_lag_to_eul_spread_3d_noalloc(int dtype, int n_components, void* const eul[3],
                                       const int64_t* eul_strides, int64_t nz, int64_t ny,
                                       int64_t nx, const void* lag_field, const void* weights,
                                       const int64_t* nearest_index, int64_t n_lag,
                                       int* out_of_domain, void* stream);

/* VirtualBoundaryForcing.compute_interaction_force_on_lag_grid —
 * sopht/numeric/immersed_boundary_ops/VirtualBoundaryForcing.py:187-230 fused:
 * support, weights, interpolation of the velocity, velocity mismatch (flow - body) and
 * penalty force k*dX + c*dV in one pass over the nodes. Lagrangian arrays are (3, n)
 * real except positions/body velocity (float64). */
int sopht_vbf_lag_force_3d(int dtype, int kernel_type, const void* const velocity[3],
                           const int64_t* velocity_strides, int64_t nz, int64_t ny, int64_t nx,
                           const double* lag_positions, const double* lag_body_velocity,
                           int64_t num_lag_nodes, double dx, double eul_grid_coord_shift,
                           double stiffness, double damping, int64_t* nearest_index,
                           void* support, void* weights, void* lag_flow_velocity,
                           void* velocity_mismatch, const void* position_mismatch,
                           void* lag_forcing, void* stream);

/* VirtualBoundaryForcing.time_step — :149-164,276-283: dX += dt * dV. */
int sopht_vbf_time_step_3d(int dtype, void* position_mismatch, const void* velocity_mismatch,
                           int64_t num_lag_nodes, double dt, void* stream);

/* gen_laplacian_filter_kernel_3d (scalar field) —
 * sopht/numeric/eulerian_grid_ops/stencil_ops_3d/laplacian_filter_3d.py:13-194.
 * filter_type 0 = "multiplicative", 1 = "convolution". Buffers as the reference's
 * (filter_flux_buffer, field_buffer); their contents afterwards are those the reference leaves.
 * A vector field is its three components in turn (:176-189). */
int sopht_laplacian_filter_3d(int dtype, int filter_order, int filter_type, void* field,
                              const int64_t* field_strides, int64_t nz, int64_t ny, int64_t nx,
                              void* filter_flux_buffer, const int64_t* flux_strides,
                              void* field_buffer, const int64_t* buffer_strides,
                              int flux_boundary_width, void* stream);

/* ---------------------------------------------------------------- the fused flow step */

typedef struct sopht_ns3d sopht_ns3d;

/* UnboundedNavierStokesFlowSimulator3D — sopht/simulator/flow/navier_stokes_flow_simulators.py:226-447
 * (greens-function Poisson solver, filter off). Owns the Poisson plan and step scratch. */
int sopht_ns3d_create(int dtype, int64_t nz, int64_t ny, int64_t nx, double x_range,
                      double kinematic_viscosity, double flow_density, int penalty_zone_width,
                      int with_forcing, int with_free_stream_flow, void* stream,
                      sopht_ns3d** plan);
/* As sopht_ns3d_create with the simulator's poisson_solver_type (:296-358): 0
 * "greens_function_convolution", 1 "fast_diagonalisation" (FastDiagPoissonSolver3D with the
 * plan's dx, homogeneous Neumann; any grid of 3 or more cells per axis). Only the chosen
 * solver's tables are built. The adjoint step needs type 0. */
int sopht_ns3d_create_with_poisson_solver(int dtype, int64_t nz, int64_t ny, int64_t nx,
                                          double x_range, double kinematic_viscosity,
                                          double flow_density, int penalty_zone_width,
                                          int with_forcing, int with_free_stream_flow,
                                          int poisson_solver_type, void* stream, sopht_ns3d** plan);
int sopht_ns3d_destroy(sopht_ns3d* plan);

/* Vorticity filtering of the step (filter_vorticity=True with filter_setting_dict,
 * navier_stokes_flow_simulators.py:274-292,396-408, applied at :473 after diffusion and before
 * penalisation). order < 0 turns it off (the default); type 0 multiplicative, 1 convolution.
 * The step then uses the components 0 and 1 of its buffer argument as the filter buffers. */
int sopht_ns3d_set_filter(sopht_ns3d* plan, int filter_order, int filter_type);

/* _navier_stokes_with_forcing_time_step / _navier_stokes_time_step — :449-498.
 * Fields are contiguous (3, nz, ny, nx). forcing may be NULL when with_forcing is 0 and is
 * zeroed after use. buffer is the simulator's buffer_vector_field (scratch; its first
 * component holds the last diffusion flux as in the reference only in spirit: contents are
 * unspecified after the call). free_stream: 3 doubles (host). */
int sopht_ns3d_step(sopht_ns3d* plan, void* vorticity, void* velocity, void* stream_function,
                    void* forcing, void* buffer, double dt, const double* free_stream,
                    void* stream);

/* The velocity recovery half of the step (:475-485): penalise vorticity (in place),
 * Poisson solve into stream_function, curl into velocity with border reset, + free stream. */
int sopht_ns3d_recover_velocity(sopht_ns3d* plan, void* vorticity, void* velocity,
                                void* stream_function, const double* free_stream, void* stream);

/* compute_advection_diffusion_stable_timestep's reduction —
 * sopht/simulator/flow/passive_transport_flow_simulators.py:139-155:
 * max over cells of |u_x|+|u_y|+|u_z| into *out (device scalar of dtype). */
int sopht_ns3d_max_abs_velocity_sum(sopht_ns3d* plan, const void* velocity, void* out,
                                    void* stream);

/* The same maximum for the velocity field that the last sopht_ns3d_step /
 * sopht_ns3d_recover_velocity of this plan wrote, from the per-workgroup maxima its velocity
 * sweep kept (no pass over the field). SOPHT_ERR_INVALID if the plan's last velocity sweep
 * wrote another field. The caller must know that the field is unchanged since. */
int sopht_ns3d_last_velocity_max(sopht_ns3d* plan, const void* velocity, void* out, void* stream);

/* Per-kernel timing of sopht_ns3d_step with HIP events on the launch stream (no
 * counterpart in the reference, which has no profiling hooks). enable_timing creates the
 * events; stage_times harvests (synchronising on the last recorded step), writes the
 * accumulated milliseconds of up to n stages into ms, the number of timed steps into
 * *steps, optionally resets, and returns the stage count. stage_name names stage i of the
 * plan (its Poisson solver sets how many stages a step has). */
int sopht_ns3d_enable_timing(sopht_ns3d* plan, int enable);
int sopht_ns3d_stage_times(sopht_ns3d* plan, double* ms, int n, int64_t* steps, int reset);
/* sopht_poisson3d_schedule of the step's Poisson plan (0 and 0 with the fast-diagonalisation
 * solver; no counterpart in the reference). */
int sopht_ns3d_poisson_schedule(const sopht_ns3d* plan, int* passes, int* zpass_jobs);
/* How the step's first sweep (vorticity forcing, advection and stretching) runs: *vector = 1
 * for 16-byte row accesses, 0 for the scalar sweep; *zchunk = z planes per work item; *cols =
 * cells per tile row (64, or 128 for FP32 rows a multiple of 128 cells). Every choice gives
 * the same bits (no counterpart in the reference). Null outputs are skipped. */
int sopht_ns3d_s1_schedule(const sopht_ns3d* plan, int* vector, int* zchunk, int* cols);
const char* sopht_ns3d_stage_name(const sopht_ns3d* plan, int stage);

/* ---- batched step (no counterpart in the reference: many independent simulators, the
 * reference driver's loop body flow_past_sphere_case.py:191-197 per member, advanced together
 * with one launch per stage; each member's results are those of stepping it alone) */
typedef struct sopht_ns3d_batch sopht_ns3d_batch;
typedef struct sopht_batch_member {
  int64_t nz, ny, nx;
  double x_range, nu, rho, cfl;  /* UnboundedNavierStokesFlowSimulator3D arguments */
  int width, with_forcing, with_free_stream, kernel_type;
  double free_stream[3];         /* U_inf (x, y, z), used when with_free_stream */
  double time;                   /* simulator time before the first step */
  void* vorticity;               /* (3, nz, ny, nx) contiguous real_t device fields */
  void* velocity;
  void* psi;
  void* forcing;
  /* virtual-boundary body (VirtualBoundaryForcing.py:20-283); n_lag = 0 for none */
  int64_t n_lag;                 /* at most 8192 */
  double dx, shift, k, c;        /* grid spacing, coordinate shift, rescaled stiffness, damping */
  const double* X;               /* (3, n_lag) float64 positions, device */
  const double* V;               /* (3, n_lag) float64 body velocities, device */
  int64_t* idx;                  /* VirtualBoundaryForcing buffers, device, real_t unless noted */
  void* support;
  void* weights;
  void* lag_u;
  void* dV;
  void* dX;
  void* Fl;
} sopht_batch_member;

/* Copies the member descriptions (fields stay owned by the caller and must outlive the batch). */
int sopht_ns3d_batch_create(int dtype, int count, const sopht_batch_member* members, void* stream,
                            sopht_ns3d_batch** batch);
/* `steps` steps of every member: dt from each member's own CFL condition (on the device),
 * its virtual-boundary time step and forcing, then the flow step. With use_graph the steps
 * after the first are replays of one captured hipGraph. */
int sopht_ns3d_batch_step(sopht_ns3d_batch* batch, int steps, int use_graph, void* stream);
/* Synchronises and reads each member's last dt and its time (either pointer may be NULL). */
int sopht_ns3d_batch_state(sopht_ns3d_batch* batch, double* dt, double* time, void* stream);
int sopht_ns3d_batch_destroy(sopht_ns3d_batch* batch);
/* The batch's size classes (one Poisson plan each; no counterpart in the reference): returns
 * their number; for 0 <= i < that number also writes class i's grid (nz, ny, nx), its
 * component count and its plan's schedule (sopht_poisson3d_schedule). Pointers may be null. */
int sopht_ns3d_batch_class_schedule(const sopht_ns3d_batch* batch, int i, int64_t* shape, int* ncomp,
                                    int* passes, int* zpass_jobs);

/* ---- reverse-mode adjoint of the coupled step (no counterpart in the reference, which has
 * no differentiation; the step it differentiates is navier_stokes_flow_simulators.py:449-498
 * with the forcing of VirtualBoundaryForcing.py:149-283 and the driver loop of
 * flow_past_sphere_case.py:191-197; dt is frozen) */

/* The flow step of sopht_ns3d_step out of place, for recording trajectories: reads
 * (vorticity_in, velocity_in, forcing), writes (vorticity_out, velocity_out); forcing is left
 * unchanged (the in-place step zeroes it). stream_function receives psi as in the step.
 * vorticity1_out (may be NULL) receives omega1 = vorticity + dt/(2 dx rho) curl_h(forcing)
 * (interior), the state the adjoint step needs. */
int sopht_ns3d_trajectory_step(sopht_ns3d* plan, const void* vorticity_in, const void* velocity_in,
                               const void* forcing, void* vorticity_out, void* velocity_out,
                               void* stream_function, void* vorticity1_out, double dt,
                               const double* free_stream, void* stream);
/* Transpose of one flow step with step dt, given the step's omega1 (vorticity1: the vorticity
 * after the forcing update, from sopht_ns3d_trajectory_step; the vorticity itself when
 * forcing is off) and its input velocity.
 * vorticity_adj: in, the adjoint of the vorticity after the step; out, the adjoint of the
 * vorticity before it. velocity_adj: in, the adjoint of the velocity after the step; out, the
 * adjoint of the velocity before it through the advection term (the virtual-boundary
 * interpolation part is added by sopht_vbf_adjoint_3d). The adjoint of the forcing is not
 * stored: sopht_vbf_adjoint_3d evaluates it at the body's support from vorticity_adj.
 * work: (3, nz, ny, nx) real_t scratch. Fields contiguous. */
int sopht_ns3d_adjoint_step(sopht_ns3d* plan, const void* vorticity1, const void* velocity,
                            double dt, void* vorticity_adj, void* velocity_adj, void* work,
                            void* stream);
/* Transpose of the virtual-boundary time step and forcing of one step (cosine kernel):
 * dX += dt dV_prev; U = dx^3 I(u; w(X)); dV = U - V; Fl = k dX + c dV; F = S(Fl; w(X));
 * body force f = -sum Fl. Inputs: vorticity_adj (the flow-step adjoint's output; the forcing
 * adjoint is forcing_prefactor * curl_h of it at the support cells), velocity (u of the step),
 * lag_positions X (float64), lag_forcing Fl of the step, body_force_adj (3 real_t, device).
 * position_mismatch_adj / velocity_mismatch_adj: the adjoint state of dX and dV_prev, updated
 * in place. Outputs: lag_position_adj (dJ/dX), lag_velocity_adj (dJ/dV), and
 * velocity_adj += dx^3 S(dJ/dU) through ws (deterministic). weights (64 n), nearest_index
 * (3 n) and lag_scratch (3 n) are scratch. */
int sopht_vbf_adjoint_3d(int dtype, const void* vorticity_adj, double forcing_prefactor,
                         const void* velocity, int64_t nz, int64_t ny, int64_t nx,
                         const double* lag_positions, const void* lag_forcing,
                         int64_t num_lag_nodes, double dx, double eul_grid_coord_shift,
                         double stiffness, double damping, double dt, const void* body_force_adj,
                         void* position_mismatch_adj, void* velocity_mismatch_adj,
                         void* lag_position_adj, void* lag_velocity_adj, void* weights,
                         int64_t* nearest_index, void* lag_scratch, sopht_spread3d* ws,
                         void* velocity_adj, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SOPHT_AMD_H */
