/*
 * ibis_gpu.h -- C ABI of the MI355X cell-residual library (libibis_gpu.so).
 *
 * The reference's boundary is the C++ class template FiniteVolume<T>
 * (src/finite_volume/finite_volume/finite_volume.h:28-142) constructed from a
 * GridBlock<T> (src/grid/grid/grid.h:26) and a json config.  This header is
 * the same operator surface as plain C: handles, pointers and sizes, no C++
 * types and no exceptions across it.  Every entry point returns an
 * ibis_status; ibis_last_error() gives the message of the last failure on the
 * calling thread.
 *
 * Conventions
 *   - FP64 arrays are structure-of-arrays with the row stride given below.
 *   - prim  (device, caller-owned): 7 x n_total, rows rho, p, T, e, vx, vy, vz
 *     (GasStates {rho,p,T,e} + velocity, src/gas/gas/gas_state.h:40-46).
 *     Ghost rows [n_cells, n_total) are written by the boundary actions,
 *     exactly as the reference mutates its FlowStates.
 *   - cons / dudt (device, caller-owned): (dim+2) x n_total, rows mass,
 *     mom_x, mom_y, [mom_z], energy (conserved_quantities.cpp:16-23); the
 *     residual writes rows [0, n_cells).
 *   - Cells and faces are in the numbering of the ibis_grid the mesh was
 *     made from.  ibis_grid_build with reorder=1 renumbers cells and faces
 *     for memory locality and reports the permutation.
 *   - Calls taking a hipStream_t are asynchronous on that stream unless
 *     documented as synchronising.  Handles are single-stream and not
 *     thread-safe, like the reference's FiniteVolume.
 */
#ifndef IBIS_GPU_H
#define IBIS_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* ibis_stream_t; /* == hipStream_t */

typedef enum {
    IBIS_OK = 0,
    IBIS_ERR_INVALID = 1, /* bad argument or configuration */
    IBIS_ERR_HIP = 2,     /* a HIP runtime call failed */
    IBIS_ERR_GRID = 3     /* grid construction failed */
} ibis_status;

#define IBIS_MAX_TAGS 16
#define IBIS_MAX_ACT 4
#define IBIS_NPAR 8
#define IBIS_MAX_MESH_TAGS 4096 /* a mesh may hold more tags than a config (batches) */

/* --------------------------------------------------------------------------
 * Host grid (replaces GridBlock construction, src/grid/grid/grid.h:89-167,
 * and the SU2 ingest of src/grid/grid/grid_io.cpp:218-348).
 * ------------------------------------------------------------------------ */
typedef struct {
    int32_t dim;
    int32_t ntags;            /* boundary tags, in the reference's sorted order */
    int64_t nv;
    const double* xyz;        /* nv x 3 (AoS) */
    int64_t nc;
    const int32_t* ctype;     /* SU2 element code per cell */
    const int64_t* cvtx;      /* nc x 8, padded with -1 */
    const int32_t* tag_ghost; /* per tag: 1 -> boundary has ghost cells */
    int64_t nb;
    const int32_t* btag;      /* boundary element -> tag */
    const int32_t* btype;     /* SU2 element code of the boundary element */
    const int64_t* bvtx;      /* nb x 4, padded with -1 */
} ibis_grid_io;

typedef struct ibis_grid ibis_grid;

/* Build connectivity, geometry and LSQ factors exactly as the reference
 * does (faces in first-appearance order, ghosts by sorted tag).  With
 * reorder != 0 the result is renumbered (valid cells in Morton order of
 * their centroids, ghosts after their valid neighbour, faces by owner). */
ibis_status ibis_grid_build(const ibis_grid_io* io, int reorder, ibis_grid** out);
void ibis_grid_destroy(ibis_grid* g);
/* out[8] = dim, n_cells, n_ghost, n_faces, n_cell_faces, n_lsq, ntags, n_bfaces */
ibis_status ibis_grid_sizes(const ibis_grid* g, int64_t* out);
/* Permutations: new id of reference cell c (n_total) and of reference face f. */
ibis_status ibis_grid_permutation(const ibis_grid* g, int64_t* cell_new_of_old,
                                  int64_t* face_new_of_old);
/* Geometry/connectivity in the grid's numbering (any pointer may be NULL):
 * face_geom 13 x F (n, t1, t2, centre, area), face_lr 2 x F, cell_off N+1,
 * cell_faces, outsigns, neighbours (n_cell_faces), centroid 3 x n_total,
 * volume n_total, lsq n_lsq x N, tag_off ntags+1, bfaces, ghosts. */
ibis_status ibis_grid_export(const ibis_grid* g, double* face_geom, int64_t* face_lr,
                             int64_t* cell_off, int64_t* cell_faces, int32_t* outsigns,
                             int64_t* neighbours, double* centroid, double* volume, double* lsq,
                             int64_t* tag_off, int64_t* bfaces, int64_t* ghosts);

/* --------------------------------------------------------------------------
 * Device mesh: a GridBlock's arrays uploaded to HBM in the kernels' layout.
 * ------------------------------------------------------------------------ */
typedef struct {
    int32_t dim, ntags;
    int64_t n_cells, n_ghost, n_faces;
    const int64_t* face_left;       /* F (GR/interface.h:335-336) */
    const int64_t* face_right;      /* F */
    const double* face_geom;        /* 13 x F: n(3) t1(3) t2(3) centre(3) area */
    const int64_t* cell_face_off;   /* n_cells + 1 (GR/cell.h:112-114) */
    const int64_t* cell_faces;      /* cell -> faces, in the cell's face order */
    const int32_t* cell_outsigns;
    const int64_t* cell_neighbours; /* GR/cell.h:327 */
    const double* centroid;         /* 3 x n_total */
    const double* volume;           /* n_total */
    const double* lsq;              /* (dim==2 ? 3 : 6) x n_cells (GR/gradient.h:56-61) */
    const int64_t* tag_off;         /* ntags + 1 */
    const int64_t* bfaces;          /* boundary faces of each tag */
} ibis_mesh_desc;

typedef struct ibis_mesh ibis_mesh;

/* Upload a mesh described by host arrays (any GridBlock). */
ibis_status ibis_mesh_create(const ibis_mesh_desc* desc, ibis_mesh** out);
/* Convenience: upload an ibis_grid. */
ibis_status ibis_mesh_from_grid(const ibis_grid* g, ibis_mesh** out);
void ibis_mesh_destroy(ibis_mesh* m);

/* --------------------------------------------------------------------------
 * Residual operator (replaces FiniteVolume<T>,
 * src/finite_volume/finite_volume/finite_volume.cpp:13-54).
 * ------------------------------------------------------------------------ */
enum { IBIS_FLUX_HANEL = 0, IBIS_FLUX_AUSMDV = 1, IBIS_FLUX_LDFSS2 = 2, IBIS_FLUX_RUSANOV = 3 };
enum { IBIS_THERMO_RHO_P = 0, IBIS_THERMO_RHO_T = 1, IBIS_THERMO_RHO_U = 2, IBIS_THERMO_P_T = 3 };
enum { IBIS_LIM_UNLIMITED = 0, IBIS_LIM_BARTH_JESPERSEN = 1, IBIS_LIM_VENKATAKRISHNAN = 2 };
/* boundary actions (src/finite_volume/finite_volume/boundaries/boundary.cpp:449-492) */
enum {
    IBIS_BC_NONE = 0,
    IBIS_BC_FLOW_STATE_COPY = 1, /* par: T, p, rho, e, vx, vy, vz */
    IBIS_BC_INTERNAL_COPY = 2,
    IBIS_BC_REFLECT_NORMAL = 3,
    IBIS_BC_VEL_REFLECT = 4,
    IBIS_BC_FIX_TEMPERATURE = 5, /* par: Twall */
    IBIS_BC_SUBSONIC_INFLOW = 6, /* par: T, p, rho, e, vx, vy, vz */
    IBIS_BC_SUBSONIC_OUTFLOW = 7,/* par: p */
    IBIS_BC_CONSTANT_FLUX = 8    /* post-convective flux action; par as flow_state_copy */
};
enum { IBIS_PREC_FP64 = 0, IBIS_PREC_FP32_FLUX = 1, IBIS_PREC_FP32_RECON = 2 };

typedef struct {
    int32_t flux, order, thermo, limiter, viscous, ntags;
    double ldfss_delta, lim_eps, venkat_K, signal_factor;
    double R, Cv, Cp, gamma, mu0, T0, Ts, Pr;
    int32_t pre_kind[IBIS_MAX_TAGS * IBIS_MAX_ACT];
    double pre_par[IBIS_MAX_TAGS * IBIS_MAX_ACT * IBIS_NPAR];
    int32_t visc_kind[IBIS_MAX_TAGS * IBIS_MAX_ACT];
    double visc_par[IBIS_MAX_TAGS * IBIS_MAX_ACT * IBIS_NPAR];
    int32_t post_kind[IBIS_MAX_TAGS * IBIS_MAX_ACT];
    double post_par[IBIS_MAX_TAGS * IBIS_MAX_ACT * IBIS_NPAR];
} ibis_fv_desc;

typedef struct ibis_residual ibis_residual;

ibis_status ibis_residual_create(const ibis_mesh* m, const ibis_fv_desc* desc,
                                 ibis_residual** out);
void ibis_residual_destroy(ibis_residual* r);

/* Arithmetic of the residual.  IBIS_PREC_FP64 (default) is the reference's
 * double arithmetic.  IBIS_PREC_FP32_FLUX stores the reconstructed face
 * values as float and evaluates the thermo closure, frame rotation and flux
 * calculator in float; gradients, limiters, the face-to-cell sum, the
 * division by the volume, estimate_dt and every solver reduction stay FP64.
 * IBIS_PREC_FP32_RECON also computes the LSQ gradients, the limiter and the
 * reconstructed values in float (differences of the FP64 inputs are formed
 * in FP64 and rounded once); accumulation and reductions stay FP64.
 * Both are supported on the inviscid second-order residual without
 * flux-override boundaries (ibis_compute_dudt fails otherwise).  No
 * counterpart in the reference (its Ibis::real is double). */
ibis_status ibis_residual_set_precision(ibis_residual* r, int mode);

/* dim, n_cells, n_total, n_faces, n_ghost of the residual's mesh. */
ibis_status ibis_residual_sizes(const ibis_residual* r, int64_t* sizes);

/* FiniteVolume<T>::compute_dudt (finite_volume.cpp:57-96, static grid).
 * Writes ghost rows of prim, rows [0, n_cells) of dudt.  Asynchronous. */
ibis_status ibis_compute_dudt(ibis_residual* r, double* prim, double* dudt,
                              int allow_reconstruction, ibis_stream_t stream);

/* Tangent of the residual: FiniteVolume<Ibis::dual>::compute_dudt as
 * SteadyStateLinearisation::matrix_vector_product runs it
 * (src/solvers/solvers/steady_state.cpp:52-128), with the reference's Dual
 * arithmetic (src/util/util/dual.h:10-346).  prim / prim_dot are the value
 * and tangent of the state (7 x n_total; ghost rows of both are written),
 * dudt / dudt_dot receive rows [0, n_cells) of the residual and of its
 * directional derivative.  The tangent scratch (twice the FP64 scratch) is
 * allocated on the first call.  Asynchronous after that first call. */
ibis_status ibis_compute_dudt_jvp(ibis_residual* r, double* prim, double* prim_dot, double* dudt,
                                  double* dudt_dot, int allow_reconstruction,
                                  ibis_stream_t stream);

/* Discrete adjoint of the residual (no counterpart in the reference, which
 * differentiates only forward, through Dual numbers).  With prim the state
 * (ghost rows are written as by ibis_compute_dudt), computes
 *   prim_bar = (d dU/dt / d prim)^T dudt_bar + (d J / d prim)^T
 * on the valid cells (ghost rows of prim_bar are zero), where
 *   J = sum over the boundary faces of the tags in force_tags (bit t = tag t)
 *       of force_dir . (momentum flux * area, oriented out of the valid
 *       cell): the pressure + viscous force on those boundaries.
 * dudt_bar ((dim+2) x n_total, rows [0, n_cells)) may be NULL; force_tags may
 * be 0.  limiter_mode: IBIS_LIMITER_ACTIVE differentiates through the
 * limiter (the branch the Dual residual takes), IBIS_LIMITER_FROZEN holds it
 * constant (match it with the option "freeze_limiter" of the tangent
 * residual).  First- or second-order residual, AUSMDV / Hanel / Rusanov (the
 * reference's own tangent through LDFSS2 is NaN on subsonic faces), FP64,
 * every boundary action (a constant-flux face passes no cotangent to the
 * state).  Reconstructed states are recomputed, not stored.  Asynchronous. */
enum { IBIS_LIMITER_FROZEN = 0, IBIS_LIMITER_ACTIVE = 1 };
ibis_status ibis_compute_dudt_vjp(ibis_residual* r, double* prim, const double* dudt_bar,
                                  uint64_t force_tags, const double* force_dir, double* prim_bar,
                                  int limiter_mode, ibis_stream_t stream);

/* conserved_to_primatives backwards: cons_bar = (d prim / d cons)^T prim_bar,
 * all n_total rows. */
ibis_status ibis_cons_to_prim_vjp(ibis_residual* r, const double* cons, const double* prim_bar,
                                  double* cons_bar, ibis_stream_t stream);

/* The force functional J above (x, y, z components, force_dir = e_k) for
 * the state prim: runs the residual and sums its boundary face fluxes.
 * Synchronises. */
ibis_status ibis_wall_force(ibis_residual* r, double* prim, uint64_t force_tags, double* force,
                            ibis_stream_t stream);

/* conserved_to_primatives on Dual values (primative_conserved_conversion.cpp:
 * 5-30), all n_total rows: value and tangent of the state from those of the
 * conserved quantities. */
ibis_status ibis_cons_to_prim_jvp(ibis_residual* r, const double* cons, const double* cons_dot,
                                  double* prim, double* prim_dot, ibis_stream_t stream);

/* FiniteVolume<T>::estimate_dt (finite_volume.cpp:131-184).  Synchronises
 * the stream and writes the global minimum to *dt_host. */
ibis_status ibis_estimate_dt(ibis_residual* r, const double* prim, double* dt_host,
                             ibis_stream_t stream);

/* FiniteVolume<T>::count_bad_cells (finite_volume.cpp:258-270).  Synchronises. */
ibis_status ibis_count_bad_cells(ibis_residual* r, const double* prim, uint64_t* n_host,
                                 ibis_stream_t stream);

/* conserved_to_primatives / primatives_to_conserved
 * (primative_conserved_conversion.cpp:5-51), all n_total rows. */
ibis_status ibis_cons_to_prim(ibis_residual* r, const double* cons, double* prim,
                              ibis_stream_t stream);
ibis_status ibis_prim_to_cons(ibis_residual* r, const double* prim, double* cons,
                              ibis_stream_t stream);

/* u1 = u0 + k * dt over all (dim+2) x n_total entries
 * (apply_time_derivative, conserved_quantities.cpp:80-94); u1 may alias u0. */
ibis_status ibis_apply_time_derivative(ibis_residual* r, const double* u0, double* u1,
                                       const double* k, double dt, ibis_stream_t stream);

/* One explicit Runge-Kutta step (RungeKutta::take_step + estimate_dt,
 * src/solvers/solvers/runge_kutta.cpp:108-203).  a is nstages x nstages
 * (a[i*nstages+j] = stage-i weight of k_j), b has nstages entries.  cons and
 * prim are the device solution; *dt_prev is read and updated.  Stage
 * buffers are library-owned.  Synchronises once (the dt reduction). */
ibis_status ibis_rk_step(ibis_residual* r, int nstages, const double* a, const double* b,
                         double cfl, double* cons, double* prim, double* dt_prev, double t,
                         double max_time, double plot_remaining, double* dt_taken,
                         ibis_stream_t stream);

/* ibis_rk_step without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight time-stepping or serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and the same results, bit for bit, as
 * ibis_rk_step; the stage buffers ((nstages + 1) x (dim+2) x n_total +
 * 7 x n_total doubles) live in a fixed library array of
 * IBIS_RK_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends, and
 * larger meshes are refused with IBIS_ERR_INVALID.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (two streams, two
 * threads) race and must not be issued. */
#define IBIS_RK_NOALLOC_SCRATCH_BYTES (64u << 20)
ibis_status ibis_rk_step_noalloc(ibis_residual* r, int nstages, const double* a,
                                 const double* b, double cfl, double* cons, double* prim,
                                 double* dt_prev, double t, double max_time,
                                 double plot_remaining, double* dt_taken, ibis_stream_t stream);

/* FiniteVolume<T>::compute_convective_gradient and
 * compute_viscous_gradient (finite_volume.h:104-112, finite_volume.cpp:
 * 273-288), the post-processing gradients read by IO/accessor.cpp:115-250.
 * Both run the pre-reconstruction boundary actions on prim (ghost rows are
 * written; the viscous one also the pre-viscous-gradient actions), then
 * overwrite slots of the residual's gradient storage with the LSQ
 * gradients of the state:
 *   convective: the thermo interpolator's pair (rho_p: p, rho; rho_u: rho,
 *     u; rho_T: rho, T; p_T: p, T), vx, vy and, in 3D only, vz
 *     (convective_flux.cpp:181-209);
 *   viscous: T, vx, vy, vz (viscous_flux.cpp:122-130).
 * Other slots keep their values.  Asynchronous. */
ibis_status ibis_compute_convective_gradient(ibis_residual* r, double* prim,
                                             ibis_stream_t stream);
ibis_status ibis_compute_viscous_gradient(ibis_residual* r, double* prim, ibis_stream_t stream);

/* FiniteVolume<T>::cell_gradients() (finite_volume.h:102): device pointer
 * of the gradient storage, 7 x 3 x n_cells FP64, slot-major (slots p, rho,
 * T, u, vx, vy, vz; components x, y, z of each slot over the cells). */
ibis_status ibis_cell_gradients(const ibis_residual* r, const double** grads);

/* Device pointers of the library's intermediates: gradients (as
 * ibis_cell_gradients), limiters 7 x n_cells, face flux (dim+2) x n_faces.
 * What they hold after a call:
 *   - ibis_compute_convective_gradient / ibis_compute_viscous_gradient:
 *     the gradient slots those write (above);
 *   - ibis_compute_dudt on the reference-shaped path (viscous,
 *     constant-flux boundaries, order 1 or allow_reconstruction = 0):
 *     gradients, limiters and face flux as the reference leaves cell_grad_,
 *     its limiters and flux_ (viscous gradients overwrite vx, vy, vz);
 *   - ibis_compute_dudt on the fused second-order inviscid path: gradients
 *     and limiters only with the option "keep_gradients" = 1 (FP64 and
 *     FP32-flux modes); otherwise, and for the face flux always, the
 *     buffers keep what an earlier call left.
 * The face flux is multiplied by the face area when the configuration is
 * inviscid. */
ibis_status ibis_residual_scratch(ibis_residual* r, double** grads, double** limiters,
                                  double** flux, int* flux_has_area);

/* Implementation choices of the 2nd-order path; every choice gives the same
 * bits.  A second-order residual without flux-override boundaries is created
 * with face_state_layout 2, pack_cells 0, precomputed_offsets 1, face_kernel
 * 1, stage_cells 1 and tile 256 (the autotuner's choice on the 3D and 2D
 * workloads); others start at 0 / layout 0 / tile 256.  Keys:
 *   "face_state_layout"   0: face states indexed by face (left/right arrays)
 *                         1: face states stored per (cell, face slot)
 *                         2: per (face slot, variable) planes over cells
 *   "precomputed_offsets" 0: x_face - x_cell computed from face centres
 *                         1: precomputed once per (cell, face slot)
 *   "pack_cells"          1: cell values + centroid packed into 64-byte
 *                            records before the cell kernel
 *                         0: gathered from the state rows (layout 2 only)
 *   "face_kernel"         0: face flux + integral over per-tile face lists
 *                         1: the same, with each list entry's face data
 *                            (value indices, geometry) stored per entry
 *                         64, 128, 256: slot-major kernel with that many
 *                            cells per block (layout 2 only)
 *   "stage_offsets"       1: the cell kernel keeps the face offsets of its
 *                            limiter pass in LDS for the reconstruction
 *                            pass (limited schemes, at most 10 faces/cell)
 *                         0: both passes read them from memory
 *   "stage_cells"         1: each block of the cell kernel copies its tile's
 *                            cells' and their outside neighbours' values
 *                            and centroids into LDS once, and the stencil
 *                            loop reads them there (layout 2, pack_cells 0;
 *                            with stage_offsets 1 the offsets then reuse
 *                            that LDS after the stencil loop)
 *   "cell_prefetch"       1: the cell kernel loads all neighbour indices
 *                            first and each neighbour's values while the
 *                            previous one is summed (layout 2, pack_cells 0,
 *                            inviscid; no effect when stage_cells stages
 *                            the stencil in LDS)
 *   "freeze_limiter"      1: the tangent (JVP) treats the limiter values
 *                            as constants; 0 (default) differentiates them
 *   "keep_gradients"      1: the fused second-order path also writes the
 *                            gradient and limiter storage (the reference's
 *                            cell_grad_ and limiters after compute_dudt);
 *                            0 (default) skips those stores
 *   "tile"                128, 256 (default) or 512 cells per tile, which
 *                            is also the block size of the tile kernels;
 *                            the mesh builds a size's tile tables on first
 *                            use (kept until ibis_mesh_destroy)
 *   "fuse_viscous"        1 (default): a second-order viscous residual runs
 *                            as four kernels -- boundary actions of both
 *                            passes, cell kernel (viscous gradients,
 *                            convective gradients, limiter, face values),
 *                            viscous terms per face, face kernel
 *                            (convective flux minus the viscous terms,
 *                            surface integral) -- with layout 2, offsets,
 *                            face_kernel 1, stage_cells, pack_cells 0; 0:
 *                            the reference-shaped general path
 *   "fold_boundary"       1 (the default on 2D meshes of at most 2^20
 *                            cells): on the fused paths
 *                            in FP64 with the cell stencil in LDS, the
 *                            boundary actions (both passes when viscous),
 *                            the ghosts' face values and the ghost rows of
 *                            prim are formed by the cell kernel's block
 *                            whose tile has the ghost in its stencil halo,
 *                            before it stages the rest of the stencil,
 *                            instead of by a kernel of their own; 0 (the
 *                            default otherwise: the action code's registers
 *                            cost the cell kernel occupancy, a cost per
 *                            cell that outweighs the boundary kernel saved
 *                            on larger and on 3D meshes): the boundary
 *                            kernel
 * Setting precomputed_offsets, face_kernel 1 or a new tile size allocates
 * and synchronises once. */
ibis_status ibis_residual_set_option(ibis_residual* r, const char* key, int64_t value);

/* Per-kernel timing with HIP events recorded around every launch of
 * ibis_compute_dudt on its stream.  Kernel slots:
 *   0 boundary actions before reconstruction (k_bc; on the fused path
 *     k_bc_ghost_fs, or k_bc, k_pack and k_ghost_fs; nothing when the
 *     option fold_boundary puts them in the cell kernel)
 *   0 also k_bc_visc_fs (both boundary passes) on the fused viscous path
 *   1 cell stencil: k_cell_grad_var, or k_cell_fs on the fused path
 *     (gradients, limiter and face values; viscous gradients too on the
 *     fused viscous path)
 *   2 face flux: k_face_flux, or k_face_*_fs on the fused path (flux and
 *     surface integral)
 *   3 k_post_flux   4 k_bc (pre-viscous)   5 k_cell_grad_var (viscous)
 *   6 k_visc_flux (k_visc_terms on the fused viscous path)   7 k_integral
 * ibis_residual_kernel_times synchronises, adds the elapsed times of all
 * completed launches into total_ms[slot] / counts[slot] (IBIS_NKERNELS
 * entries each) and, with reset != 0, clears the accumulators. */
#define IBIS_NKERNELS 8
ibis_status ibis_residual_profile(ibis_residual* r, int enable);
ibis_status ibis_residual_kernel_times(ibis_residual* r, double* total_ms, int64_t* counts,
                                       int reset);

/* ------------------------------------------------------------------------
 * Steady-state linearisation, GMRES / FGMRES and JFNK
 * (src/solvers/solvers/steady_state.cpp:52-165, src/solvers/solvers/jfnk.cpp,
 * src/linear_algebra/linear_algebra/gmres.cpp).  Vectors have n_cells *
 * (dim+2) FP64 entries in the reference's order: index cell*(dim+2) + q
 * (steady_state.cpp:60-68).  All vectors are device memory.
 * ------------------------------------------------------------------------ */

/* SteadyStateLinearisation::matrix_vector_product: result = vec / dt_star -
 * dual(R(U + eps*vec)) with the tangent residual, cons the (dim+2) x n_total
 * conserved state.  allow_reconstruction = 0 is the preconditioner's
 * system (steady_state.cpp:47-50).  Asynchronous. */
ibis_status ibis_steady_matvec(ibis_residual* r, const double* cons, const double* vec,
                               double* result, double dt_star, int allow_reconstruction,
                               ibis_stream_t stream);

/* SteadyStateLinearisation::eval_rhs: residual of prim into dudt and into
 * rhs in vector order.  Asynchronous. */
ibis_status ibis_steady_rhs(ibis_residual* r, double* prim, double* dudt, double* rhs,
                            int allow_reconstruction, ibis_stream_t stream);

/* A linear operator for the Krylov solvers: w = A v on the given stream. */
typedef ibis_status (*ibis_matvec_fn)(void* ctx, const double* v, double* w,
                                      ibis_stream_t stream);

/* Two ready-made operators (pass the function as A or P and a pointer to
 * the struct as its context):
 *   ibis_steady_op_matvec: the steady-state linearisation above;
 *   ibis_dense_matvec: w = a v for a dense row-major n x n device matrix
 *   (the LinearSystem of the reference's GMRES tests, gmres.cpp:320-500). */
typedef struct {
    ibis_residual* r;
    const double* cons;
    double dt_star;
    int32_t allow_reconstruction;
} ibis_steady_op;
ibis_status ibis_steady_op_matvec(void* ctx, const double* v, double* w, ibis_stream_t stream);

typedef struct {
    int64_t n;
    const double* a;
} ibis_dense_op;
ibis_status ibis_dense_matvec(void* ctx, const double* v, double* w, ibis_stream_t stream);

/* LinearSolveResult (src/linear_algebra/linear_algebra/linear_system.h) */
typedef struct {
    int32_t success;
    int32_t n_iters;
    double tol;
    double residual; /* |g_{m}| / beta */
} ibis_linear_result;

typedef struct ibis_gmres ibis_gmres;

/* Gmres(system, max_iters, tol) (gmres.cpp:87-116): Krylov basis of
 * min(num_vars, max_iters) + 1 device vectors. */
ibis_status ibis_gmres_create(int64_t num_vars, int32_t max_iters, double tol,
                              ibis_gmres** out);
void ibis_gmres_destroy(ibis_gmres* g);

/* Gmres::solve (gmres.cpp:119-182): x += the GMRES correction for A x = rhs
 * from the initial guess x.  Modified Gram-Schmidt on the device (one host
 * synchronisation per Arnoldi step for the Givens rotations, kept on the
 * host as in the reference).  Deterministic reductions. */
ibis_status ibis_gmres_solve(ibis_gmres* g, ibis_matvec_fn A, void* ctx, const double* rhs,
                             double* x, ibis_linear_result* result, ibis_stream_t stream);

/* FGmres::solve (gmres.cpp:234-303): right-preconditioned flexible GMRES;
 * each preconditioner application solves P z = v with the inner solver
 * `inner` from z = 0. */
ibis_status ibis_fgmres_solve(ibis_gmres* g, ibis_matvec_fn A, void* ctx_a, const double* rhs,
                              double* x, ibis_gmres* inner, ibis_matvec_fn P, void* ctx_p,
                              ibis_linear_result* result, ibis_stream_t stream);

/* CFL schedules (src/solvers/solvers/cfl.cpp:9-83) */
enum { IBIS_CFL_CONSTANT = 0, IBIS_CFL_LINEAR = 1, IBIS_CFL_RESIDUAL = 2 };
#define IBIS_MAX_CFL_POINTS 16

/* Jfnk configuration (jfnk.cpp:7-22 and the linear_solver block read by
 * make_linear_solver, gmres.cpp:305-318) */
typedef struct {
    int32_t linear_solver; /* 0 gmres, 1 fgmres */
    int32_t max_iters;
    double tol;
    int32_t max_precond_iters;
    double precond_tol;
    int32_t cfl_kind;
    int32_t n_cfl_points;
    double cfl_value;                          /* constant */
    double cfl_times[IBIS_MAX_CFL_POINTS];     /* linear_interpolate */
    double cfl_values[IBIS_MAX_CFL_POINTS];
    double growth_threshold, power, start_cfl, max_cfl; /* residual_based */
} ibis_jfnk_desc;

typedef struct ibis_jfnk ibis_jfnk;

/* ConservedQuantitiesNorm: mass, momentum x, y, z, energy, global */
#define IBIS_NNORMS 6

ibis_status ibis_jfnk_create(ibis_residual* r, const ibis_jfnk_desc* desc, ibis_jfnk** out);
void ibis_jfnk_destroy(ibis_jfnk* j);

/* Jfnk::initialise (jfnk.cpp:24-29): residual of the state, its L2 norms
 * (ConservedQuantities::L2_norms, conserved_quantities.cpp:41-69, over the
 * valid cells) into norms[IBIS_NNORMS]; they become the initial norms of the
 * residual-based CFL.  Synchronises. */
ibis_status ibis_jfnk_initialise(ibis_jfnk* j, double* cons, double* prim, double* norms,
                                 ibis_stream_t stream);

/* Jfnk::step (jfnk.cpp:39-93): dt* = cfl(step) * estimate_dt, linear solve
 * of (I/dt* - dR/dU) dU = R, U += dU on the valid cells, conversion to
 * primitives, new residual and its norms.  Synchronises. */
ibis_status ibis_jfnk_step(ibis_jfnk* j, double* cons, double* prim, int64_t step,
                           ibis_linear_result* result, double* norms, double* dt_star,
                           ibis_stream_t stream);

/* CflSchedule::eval for a Jfnk's schedule (host only; for tests). */
ibis_status ibis_cfl_eval(const ibis_jfnk_desc* desc, double t, double* previous_cfl,
                          double* previous_residual, double* cfl);

/* ------------------------------------------------------------------------
 * Batches: many independent instances (RungeKutta::take_step of each,
 * src/solvers/solvers/runge_kutta.cpp:108-203) advanced together.  No
 * counterpart in the reference, which steps one block per process.
 * ------------------------------------------------------------------------ */
typedef struct ibis_batch ibis_batch;

/* m: the instances' grids concatenated (valid cells of instance 0, 1, ...,
 * then their ghosts in the same order; tags of instance 0, then 1, ...).
 * descs[i]: instance i's configuration (its ntags tags); all instances must
 * share flux, reconstruction, limiter, gas and transport settings.
 * cell_off / ghost_off (n_inst + 1): each instance's range of valid cells
 * and of ghost cells (ghost g is cell n_cells + g). */
ibis_status ibis_batch_create(const ibis_mesh* m, int n_inst, const ibis_fv_desc* descs,
                              const int64_t* cell_off, const int64_t* ghost_off,
                              ibis_batch** out);
void ibis_batch_destroy(ibis_batch* b);
/* The residual over all instances (owned by the batch). */
ibis_residual* ibis_batch_residual(ibis_batch* b);

/* Bind the device state (cons, prim over the concatenated mesh) and the
 * tableau (as ibis_rk_step); every instance starts at t = 0 with
 * dt_prev = dt_init.  Plotting clips are not applied. */
ibis_status ibis_batch_rk_init(ibis_batch* b, int nstages, const double* a, const double* bw,
                               double cfl, double max_time, double dt_init, double* cons,
                               double* prim);
/* nsteps explicit steps of every instance on the batch's own stream, each
 * instance with its own dt; with use_graph the first call captures one step
 * into a HIP graph that later steps replay.  Synchronises the device before
 * and the batch's stream after.  The caller may write the bound state
 * between calls: each call takes its first time step from the state as it
 * finds it (with the stage updates fused, later steps take theirs from the
 * local time steps the previous step's last stage wrote, the same bits). */
ibis_status ibis_batch_rk_steps(ibis_batch* b, int nsteps, int use_graph);
/* Per-instance time and last dt (n_inst each; either may be NULL). */
ibis_status ibis_batch_times(ibis_batch* b, double* t, double* dt_prev);
/* Batch options (any value gives the same bits):
 *   "fuse_stages" (default 1): each stage update and primitive conversion
 *   runs in the epilogue of the residual's face kernel, and the time step
 *   before the first residual, where the residual's path allows it (the
 *   inviscid second-order FP64 path with the per-entry face kernel);
 *   0 runs them as separate passes. */
ibis_status ibis_batch_set_option(ibis_batch* b, const char* key, int64_t value);

const char* ibis_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
