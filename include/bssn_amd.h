/*
 * bssn_amd.h -- C ABI of the MI355X (gfx950) BSSN evolution library.
 *
 * Plain C: device pointers, sizes and a HIP stream passed as void*.  Every
 * entry point is asynchronous on the given stream (NULL = the default
 * stream), retains nothing after it returns, never allocates unless stated,
 * and returns 0 on success or a negative status; bssn_last_error_string()
 * describes the last failure of the calling thread.
 *
 * State layout ("packed state"): 21 FP64 component arrays of nx*ny*nz
 * points each, back to back (component c at state + c*nx*ny*nz), every
 * array C-order with z fastest -- the reference layout of each field
 * (reference docs/bssn.rst:11-12, bssn/variables.py:8-16) with the symmetric
 * tensors stored once:
 *
 *    0..5   conformal_metric  xx xy xz yy yz zz
 *    6      conformal_factor  W
 *    7..12  traceless_K       xx xy xz yy yz zz
 *    13     trace_K
 *    14..16 conformal_connection x y z
 *    17     lapse
 *    18..20 shift x y z
 *
 * Arrays are the whole periodic domain with no ghost points: periodic axes
 * wrap (the reference's jnp.roll), Sommerfeld axes use one-sided closures,
 * exactly like the reference unigrid path.  Guard cells / owned ranges of
 * the reference's tiled path are described by bssn_params.
 */
#ifndef BSSN_AMD_H
#define BSSN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BSSN_NCOMP 21
#define BSSN_OK 0
#define BSSN_ERR_INVALID (-1)   /* invalid configuration (Python: ValueError) */
#define BSSN_ERR_HIP (-2)       /* HIP runtime failure   (Python: RuntimeError) */
#define BSSN_ERR_UNSUPPORTED (-3)

/* Mirrors reference BSSNParameters (bssn/variables.py:19-46) field by field.
 * owned_shape[d] < 0 means "None" (the whole array extent), reduced_axes[0]
 * < 0 means "None" (legacy singleton inference). */
typedef struct bssn_params {
    double eta, kappa, nu, g, dx, dt;
    double x_min, y_min, z_min;
    double mad_q;
    int32_t zero_shift;      /* 1: the shift RHS is zero                */
    int32_t gauge;           /* 0 harmonic, 1 one-plus-log              */
    int32_t bc[6];           /* xl xr yl yr zl zr: 0 periodic, 1 Sommerfeld */
    int32_t guard_cells;
    int32_t owned_shape[3];
    int32_t owned_faces[6];
    int32_t reduced_axes[3];
} bssn_params;

/* Array extents of one evolution. */
typedef struct bssn_grid {
    int32_t nx, ny, nz;
} bssn_grid;

/* Arithmetic of the RHS: BSSN_PREC_F64 everything in FP64; BSSN_PREC_F32RHS
 * derivatives and RHS algebra in FP32 while the RK4 state, accumulators,
 * projections and reductions stay FP64. */
#define BSSN_PREC_F64 0
#define BSSN_PREC_F32RHS 1

const char *bssn_last_error_string(void);

/* Library version string ("bssn_amd <semver> gfx950"). */
const char *bssn_version(void);

/* Caller-owned scratch, in doubles: what = 0 for bssn_rhs / bssn_rk4_stage
 * (non-zero only when kappa != 0: the derived grid functions of the
 * constraint-damping term), what = 1 for bssn_evolve (3 state-sized RK4
 * buffers plus the former). */
int64_t bssn_work_size(const bssn_params *p, const bssn_grid *grid, int what);

/* Replaces reference evolution/time_evolve.py:41-57 compute_bssn_rhs.
 * rhs[21*n] <- RHS(state[21*n]); state and rhs must not alias.  work holds
 * bssn_work_size(p, grid, 0) doubles (may be NULL when that is 0). */
int bssn_rhs(const bssn_params *p, const bssn_grid *grid, int precision,
             const double *state, double *rhs, double *work, void *stream);

/* Replaces reference time_evolve.py:60-91 compute_bssn_rhs_with_matter:
 * matter[10*n] = rho, S_x, S_y, S_z, S_xx, S_xy, S_xz, S_yy, S_yz, S_zz
 * (reference layouts rho[n], S_i[3,n], S_ij[3,3,n] with S_ij stored once).
 * Other arguments as bssn_rhs. */
int bssn_rhs_with_matter(const bssn_params *p, const bssn_grid *grid, int precision,
                         const double *state, const double *matter, double *rhs,
                         double *work, void *stream);

/* Replaces reference time_evolve.py:134-141 enforce_algebraic_constraints:
 * out <- P(in) pointwise (det = 1 rescale, then trace-free A).  in == out
 * is allowed. */
int bssn_enforce_constraints(const bssn_grid *grid, const double *in,
                             double *out, void *stream);

/* One fused RK4 stage (reference time_evolve.py:144-217, one of the four
 * RHS evaluations and the combination that follows it):
 *   k = RHS(u_s);
 *   stage 1: ustar  = u0 + dt/6 k;   u_next = P(u0 + dt/2 k)
 *   stage 2: ustar += dt/3 k;        u_next = P(u0 + dt/2 k)
 *   stage 3: ustar += dt/3 k;        u_next = P(u0 + dt k)
 *   stage 4:                         u_next = P(ustar + dt/6 k)
 *                                    (P(P(...)) when reproject != 0)
 * u_next must alias neither u_s nor u0 (except stage 4, where u_next ==
 * u0 is allowed); u0 is unused in stage 4; ustar is unused in stage 4's
 * write set.  work as for bssn_rhs. */
int bssn_rk4_stage(const bssn_params *p, const bssn_grid *grid, int precision,
                   int stage, const double *u_s, const double *u0,
                   double *ustar, double *u_next, int reproject, double *work,
                   void *stream);

/* Replaces reference time_evolve.py:144-217 rk4_step applied nsteps times:
 * state <- rk4_step^nsteps(state), in place.  work must hold
 * bssn_work_size(p, grid, 1) doubles (caller-owned; no allocation).  Identical to calling rk4_step
 * nsteps times (the projection at the start of each step included). */
int bssn_evolve(const bssn_params *p, const bssn_grid *grid, int precision,
                int nsteps, double *state, double *work, void *stream);

/* Replaces reference bssn/constraints.py:387-412 compute_all_constraints:
 * out[9*n] = H, M_x, M_y, M_z, det-1, tr A, G^x, G^y, G^z (each n points). */
int bssn_constraints(const bssn_params *p, const bssn_grid *grid,
                     const double *state, double *out, void *stream);

/* Replaces reference bssn/constraints.py:415-440: as bssn_constraints with
 * H - 16 pi rho and M_i - 8 pi S_i (matter layout as bssn_rhs_with_matter,
 * NULL = vacuum). */
int bssn_constraints_with_matter(const bssn_params *p, const bssn_grid *grid,
                                 const double *state, const double *matter,
                                 double *out, void *stream);

/* Replaces reference diagnostics/constraints.py:9-36 compute_constraint_norms
 * over the constraint fields bssn_constraints writes.
 * norms[10] = L2 and Linf of (H, M, det, trA, Gamma) in the order
 * hamiltonian_l2, momentum_l2, det_gamma_l2, trace_A_l2, gamma_l2,
 * hamiltonian_linf, momentum_linf, det_gamma_linf, trace_A_linf, gamma_linf
 * (device pointer).  partial must hold bssn_norm_partial_size(grid) doubles. */
int64_t bssn_norm_partial_size(const bssn_grid *grid);
int bssn_constraint_norms(const bssn_grid *grid, const double *constraints,
                          double *partial, double *norms, void *stream);

/* No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  bssn_constraint_norms with the reduction's partial sums in a
 * static device array of the library (1024 x 10 doubles = 80 KB on each
 * device it runs on, held until the process ends) instead of a caller
 * buffer; calls no allocating routine.  Grids of at most 2^31 - 1 points per
 * component (BSSN_ERR_INVALID beyond).  Every call on one device shares that
 * device's array: calls that overlap on one device (concurrent streams)
 * race. */
int bssn_constraint_norms_noalloc(const bssn_grid *grid, const double *constraints,
                                  double *norms, void *stream);

/* Fixed mesh refinement (vertex-centred, 2:1).  A child patch with parent
 * active bounds lo..hi (inclusive, per axis) is allocated with
 * 2*(hi-lo)+1 active vertices plus `ghost` points on each side; the parent
 * allocation has `parent_ghost` points before its active vertex 0.
 *
 * Replaces reference fmr/transfer.py:237-244 fill_tiled_fine_ghosts (and the
 * frozen serial tests/support/fmr_reference.py:126-143 fill_fine_ghosts):
 * every child point outside the active box <- separable degree-5 Lagrange
 * interpolation of the parent on six nodes per axis around the parent
 * position parent_ghost + lo + (i - ghost)/2.  periodic[3] marks parent
 * axes that wrap (a ghost-free periodic root); on other axes the six nodes
 * must lie inside the parent allocation.  All 21 components. */
int bssn_fmr_prolong(const bssn_grid *parent_grid, int32_t parent_ghost,
                     const int32_t *periodic, const double *parent,
                     const bssn_grid *child_grid, const int32_t *lo, int32_t ghost,
                     double *child, void *stream);

/* Replaces reference fmr/transfer.py:247-252 restrict_tiled_to_coarse
 * (frozen fmr_reference.py:146-160 restrict_to_coarse): parent vertices
 * lo..hi <- the coincident active child vertices (every second one). */
int bssn_fmr_restrict(const bssn_grid *child_grid, const int32_t *lo, int32_t ghost,
                      const double *child, const bssn_grid *parent_grid,
                      int32_t parent_ghost, double *parent, void *stream);

/* RHS on the box lo..hi-1 of the grid only (every other point of rhs is
 * left untouched), with the same stencils, closures and Sommerfeld faces as
 * bssn_rhs on the whole grid.  Serves callers that need the RHS on part of
 * a support, like the axisymmetric Cartoon step, which keeps only the
 * centre y plane of its nine-plane support (reference
 * cartoon/axisymmetry/evolution.py:18-28 evaluates all nine). */
int bssn_rhs_region(const bssn_params *p, const bssn_grid *grid, int precision,
                    const int32_t *lo, const int32_t *hi, const double *state,
                    double *rhs, double *work, void *stream);

/* The RK4 combination of one stage for an RHS k computed elsewhere
 * (reference time_evolve.py:144-217; weights as bssn_rk4_stage):
 *   stage 1: ustar = u0 + dt/6 k,  u_next = P(u0 + dt/2 k)
 *   stage 2: ustar += dt/3 k,      u_next = P(u0 + dt/2 k)
 *   stage 3: ustar += dt/3 k,      u_next = P(u0 + dt k)
 *   stage 4:                       u_next = P(ustar + dt/6 k)
 * Pointwise; u_next may alias u0 in stage 4 only. */
int bssn_rk4_combine(const bssn_grid *grid, double dt, int stage, const double *k,
                     const double *u0, double *ustar, double *u_next, int reproject,
                     void *stream);

/* Axisymmetric Cartoon (rotation about z).  Compact state: extents
 * (nx, 1, nz), x index 0..3 parity ghosts, 4..nx-1 positive radial
 * half-cell centres rho = (i - 3.5) dx; support: (nx, 9, nz), y = (j - 4) dx.
 *
 * Replaces reference cartoon/axisymmetry/reconstruction.py:131-145
 * fill_axisymmetric_ghosts: the four ghosts from the positive half by the
 * parity of a rotation by pi (vectors (-1, -1, +1)), in place. */
int bssn_cartoon_axi_fill_ghosts(const bssn_grid *compact, double *state, void *stream);

/* Replaces reconstruction.py:255-292 reconstruct_axisymmetric_support:
 * support <- every component interpolated (six-point Lagrange on the
 * parity-reflected radial profile with a three-sample 1/r outer buffer,
 * z_min the z coordinate of index 0) at rho = sqrt(x^2 + y^2) and rotated
 * by phi about z. */
int bssn_cartoon_axi_reconstruct(const bssn_grid *compact, double dx, double z_min,
                                 const double *state, double *support, void *stream);

/* Replaces reconstruction.py:295-304 project_axisymmetric_rhs: compact rhs
 * <- the centre plane (j = 4) of a support RHS at positive rho, ghosts by
 * parity.  Only the centre plane's points x >= 4 of support_rhs are read. */
int bssn_cartoon_axi_project(const bssn_grid *compact, const double *support_rhs,
                             double *rhs, void *stream);

/* Tile layout of a tiled field (reference parallel/layout.py:27-78
 * TileLayout): counts tiles per axis, each owning tile_shape samples
 * (+1 on the last tile of an axis with an upper endpoint) padded by
 * guard_cells on both sides.  On one GPU the tiles are blocks of one buffer,
 * tile-major: tile (ix, iy, iz) at ((ix*cy + iy)*cz + iz) * ncomp * prod(alloc)
 * with alloc = tile_shape + 2 guard_cells, components inside a tile back to
 * back, each C-order with z fastest. */
typedef struct bssn_tile_layout {
    int32_t counts[3];
    int32_t tile_shape[3];
    int32_t guard_cells;
    int32_t upper_endpoints[3];
} bssn_tile_layout;

/* Replaces reference parallel/communication.py:36-147 (make_tiled_halo_op,
 * refresh_tiled_field): every guard slab of every tile refreshed in an
 * x -> y -> z sweep, from the neighbouring tile's owned slab (periodic
 * wrap across the tile grid) or, on a face with boundary code 1 (bc[6] =
 * xl xr yl yr zl zr), by the exterior rule: 0 linear extrapolation from
 * the face sample and its inward neighbour, 1 constant, 2 left as it was.
 * ncomp components per tile, in place. */
int bssn_tiled_ghost_fill(const bssn_tile_layout *layout, const int32_t *bc, int exterior,
                          int ncomp, double *field, void *stream);

/* Batched evolution (no reference counterpart: the reference evolves one
 * grid per call; this advances many independent evolutions of uneven sizes
 * with shared launches per RK4 stage: one over the lean tiles of every
 * instance, four over the face tiles of the instances with Sommerfeld
 * faces, one workgroup per tile).  Each instance has its own parameters (periodic or Sommerfeld
 * faces, the Sommerfeld ones on the tile grid of the chosen tile shape,
 * whole grids, mad_q = 1, kappa = 0; others return BSSN_ERR_UNSUPPORTED)
 * and packed state.
 *
 * bssn_batch_prepare uploads the instance descriptors into work (doubles:
 * bssn_batch_work_size) and fills *plan; it waits for the upload.
 * tile_variant selects the tile shape (0: 4x8x8, 1: 8x8x4, 2: 8x4x8),
 * grid_blocks the most workgroups per launch (a list of more tiles is
 * covered by successive launches).
 * bssn_batch_step then applies rk4_step (reference time_evolve.py:144-217)
 * once to every instance, in place: one projection launch and four stages
 * of five tile lists each, nothing on the host in between (capturable in a
 * HIP graph).  plan->total_tiles / face_tiles count the lean / face tiles,
 * face_list_tiles the face tiles with closures along x, y, z only and the
 * edge and corner tiles.
 * The states and work must stay where they were when the plan was made. */
typedef struct bssn_batch_plan {
    int32_t ninst, total_tiles, precision, tile_variant, grid_blocks, face_tiles;
    int64_t total_cells, desc_offset, prefix_offset, proj_offset;
    int32_t face_list_tiles[4];
} bssn_batch_plan;

int64_t bssn_batch_work_size(const bssn_grid *grids, int n);
int bssn_batch_prepare(const bssn_params *params, const bssn_grid *grids, double *const *states,
                       int n, int precision, int tile_variant, int grid_blocks, double *work,
                       bssn_batch_plan *plan, void *stream);
int bssn_batch_step(const bssn_batch_plan *plan, double *work, void *stream);

/* Forward-mode differentiation (no reference counterpart: the reference
 * differentiates with jax autodiff).  Exact tangent-linear maps of the
 * operators above along ndir = 1 or 2 directions at once, kappa = 0 and
 * vacuum only (BSSN_ERR_UNSUPPORTED otherwise).  A tangent buffer holds
 * ndir state-layout arrays back to back (ndir * 21 * n doubles; direction l
 * at offset l * 21 * n):
 *
 * bssn_rhs_tangent: rhs <- RHS(state), drhs <- dRHS(state)[dstate]
 * (compute_bssn_rhs, time_evolve.py:41-57). */
int bssn_rhs_tangent(const bssn_params *p, const bssn_grid *grid, int ndir, const double *state,
                     const double *dstate, double *rhs, double *drhs, void *stream);

/* Doubles of scratch bssn_evolve_tangent needs: 3 (1 + ndir) * 21 * n. */
int64_t bssn_tangent_work_size(const bssn_grid *grid, int ndir);

/* bssn_evolve with the tangents carried along: state <- rk4_step^nsteps
 * (state), dstate <- its derivatives along dstate, both in place. */
int bssn_evolve_tangent(const bssn_params *p, const bssn_grid *grid, int nsteps, int ndir, double *state,
                        double *dstate, double *work, void *stream);

/* The Hamiltonian constraint (bssn/constraints.py:271-303) and its
 * tangents: H[n], dH[ndir * n]. */
int bssn_hamiltonian_tangent(const bssn_params *p, const bssn_grid *grid, int ndir, const double *state,
                             const double *dstate, double *H, double *dH, void *stream);

/* Reverse-mode differentiation (no reference counterpart: the reference
 * differentiates with jax autodiff): the transposed linearisations an
 * adjoint (reverse) sweep of rk4_step (time_evolve.py:144-217) is built
 * from.  kappa = 0, vacuum, whole grids (no guard cells; status
 * BSSN_ERR_UNSUPPORTED otherwise).
 *
 * bssn_rhs_vjp: state_bar (= or += when accumulate) <- J_F(state)^T rhs_bar
 * + J_H(state)^T h_bar, with F the RHS of compute_bssn_rhs (time_evolve.py:
 * 41-57) and H the Hamiltonian constraint (bssn/constraints.py:271-303);
 * rhs_bar (21 * n) or h_bar (n) may be null.  work holds
 * bssn_vjp_work_size(grid) doubles (the features of every point, their
 * adjoints and per-axis tables of the transposed stencil rows).  The upwind
 * side of each point is held
 * fixed, as jax differentiates jnp.where. */
int64_t bssn_vjp_work_size(const bssn_grid *grid);
int bssn_rhs_vjp(const bssn_params *p, const bssn_grid *grid, const double *state, const double *rhs_bar,
                 const double *h_bar, double *state_bar, int accumulate, double *work, void *stream);

/* state_bar <- J_P(state)^T state_bar in place, P the pointwise projection
 * of enforce_algebraic_constraints (time_evolve.py:94-141) at the
 * pre-projection state. */
int bssn_project_vjp(const bssn_grid *grid, const double *state, double *state_bar, void *stream);

/* Launch parts of one RK4 stage, for the measurement helpers below:
 * TILES    both tile kernels (every point the tile grid owns);
 * SHELL    the Sommerfeld closure-shell kernel (tiles not aligned to faces);
 * LEAN     the tile kernel without closure rows only (interior tiles when the
 *          face tiles take closures, otherwise every tile);
 * CLOSURE  the closure-capable tile kernel only (face tiles). */
#define BSSN_PART_TILES 1
#define BSSN_PART_SHELL 2
#define BSSN_PART_LEAN 4
#define BSSN_PART_CLOSURE 8

/* Measurement helper (no reference counterpart): launches one part
 * (BSSN_PART_*) of RK4 stage `stage` `reps` times between two hipEvents
 * recorded on `stream`, waits for them and stores the mean duration per
 * launch in *ms.  Buffers as bssn_rk4_stage. */
int bssn_time_stage_part(const bssn_params *p, const bssn_grid *grid, int precision, int stage, int part,
                         const double *u_s, const double *u0, double *ustar, double *u_next, int reps,
                         float *ms, double *work, void *stream);

/* Number of grid points one part (BSSN_PART_*) of a stage updates
 * (negative status on invalid input). */
int64_t bssn_stage_points(const bssn_params *p, const bssn_grid *grid, int part);

#ifdef __cplusplus
}
#endif
#endif /* BSSN_AMD_H */
