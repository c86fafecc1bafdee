/*
 * ibis_oracle.h -- CPU restatement of ibis's cell residual (dU/dt).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the MI355X path is checked
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker or as the timed CPU baseline --
 * never as part of the product path.
 *
 * Every function restates the reference loop by loop, in the same arithmetic
 * order (see the citations in ibis_oracle.cpp).  Build flags keep IEEE
 * semantics: -O2 -ffp-contract=off, no fast-math.
 *
 * Array conventions (all structure-of-arrays, FP64, original reference
 * numbering):
 *   prim  : 7 x N_total  rows rho, p, T, e, vx, vy, vz   (GasStates 0..3 +
 *           Vector3s, src/gas/gas/gas_state.h:40-46)
 *   cons  : ncons x N_total rows mass, mom_x, mom_y, [mom_z], energy
 *           (src/finite_volume/finite_volume/conserved_quantities.cpp:16-23)
 */
#ifndef IBIS_ORACLE_H
#define IBIS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_TAGS 16
#define OR_MAX_ACT 4
#define OR_NPAR 8

/* SU2-like grid description (GridIO, src/grid/grid/grid_io.h:55) */
typedef struct {
    int32_t dim;
    int32_t ntags;           /* boundary tags, already in sorted (json) order */
    int64_t nv;
    const double* xyz;       /* nv*3 */
    int64_t nc;
    const int32_t* ctype;    /* SU2 element code: 3,5,9,10,12,13,14 */
    const int64_t* cvtx;     /* nc*8, padded with -1 */
    const int32_t* tag_ghost;/* per tag: 1 -> ghost cells */
    int64_t nb;
    const int32_t* btag;     /* boundary element -> tag index */
    const int32_t* btype;    /* SU2 element code of the boundary element */
    const int64_t* bvtx;     /* nb*4, padded with -1 */
} or_grid_io;

/* Boundary action kinds (src/finite_volume/finite_volume/boundaries/boundary.cpp:449-479) */
enum {
    OR_BC_NONE = 0,
    OR_BC_FLOW_STATE_COPY = 1,  /* par: T, p, rho, e, vx, vy, vz */
    OR_BC_INTERNAL_COPY = 2,
    OR_BC_REFLECT_NORMAL = 3,
    OR_BC_VEL_REFLECT = 4,
    OR_BC_FIX_TEMPERATURE = 5,  /* par: Twall */
    OR_BC_SUBSONIC_INFLOW = 6,  /* par: T, p, rho, e, vx, vy, vz */
    OR_BC_SUBSONIC_OUTFLOW = 7, /* par: p */
    OR_BC_CONSTANT_FLUX = 8     /* post-convective; par: T, p, rho, e, vx, vy, vz */
};

enum { OR_FLUX_HANEL = 0, OR_FLUX_AUSMDV = 1, OR_FLUX_LDFSS2 = 2, OR_FLUX_RUSANOV = 3 };
enum { OR_THERMO_RHO_P = 0, OR_THERMO_RHO_T = 1, OR_THERMO_RHO_U = 2, OR_THERMO_P_T = 3 };
enum { OR_LIM_UNLIMITED = 0, OR_LIM_BARTH_JESPERSEN = 1, OR_LIM_VENKATAKRISHNAN = 2 };

typedef struct {
    int32_t flux, order, thermo, limiter, viscous, ntags;
    double ldfss_delta, lim_eps, venkat_K, signal_factor;
    double R, Cv, Cp, gamma, mu0, T0, Ts, Pr;
    int32_t pre_kind[OR_MAX_TAGS * OR_MAX_ACT];
    double pre_par[OR_MAX_TAGS * OR_MAX_ACT * OR_NPAR];
    int32_t visc_kind[OR_MAX_TAGS * OR_MAX_ACT];
    double visc_par[OR_MAX_TAGS * OR_MAX_ACT * OR_NPAR];
    int32_t post_kind[OR_MAX_TAGS * OR_MAX_ACT];
    double post_par[OR_MAX_TAGS * OR_MAX_ACT * OR_NPAR];
} or_fv_config;

/* grid */
void* or_grid_build(const or_grid_io* io);
void or_grid_free(void* g);
/* out: dim, N, G, F, nnz, n_lsq, ntags, n_bfaces */
void or_grid_sizes(const void* g, int64_t* out);
void or_grid_export(const void* g, double* face_geom /* 13 x F: n t1 t2 c area */,
                    int64_t* face_lr /* 2 x F */, int64_t* cell_off /* N+1 */,
                    int64_t* cell_faces, int32_t* outsigns, int64_t* neighbours,
                    double* centroid /* 3 x Ntot */, double* volume /* Ntot */,
                    double* lsq /* n_lsq x N */, int64_t* tag_off /* ntags+1 */,
                    int64_t* bfaces, int64_t* ghosts);
/* LSQ gradient of an arbitrary N_total field (GR/gradient.h:66-116) */
void or_grid_gradient(const void* g, const double* values, double* grad /* 3 x N */);

/* finite volume */
void* or_fv_create(void* g, const or_fv_config* cfg);
void or_fv_free(void* fv);
int or_compute_dudt(void* fv, double* prim, double* dudt, int allow_reconstruction);
int or_compute_dudt_jvp(void* fv, double* prim, double* prim_dot, double* dudt,
                        double* dudt_dot, int allow_reconstruction);
double or_estimate_dt(void* fv, const double* prim);
int64_t or_count_bad_cells(void* fv, const double* prim);
void or_cons_to_prim(void* fv, const double* cons, double* prim);
void or_cons_to_prim_jvp(void* fv, const double* cons, const double* cons_dot,
                         double* prim, double* prim_dot);
void or_prim_to_cons(void* fv, const double* prim, double* cons);
/* grads: 7 vars (p, rho, T, u, vx, vy, vz) x 3 x N; lim: 7 x N; flux: ncons x F;
   left/right: 7 x F.  Any pointer may be NULL. */
/* FiniteVolume::compute_convective_gradient / compute_viscous_gradient
 * (FV/finite_volume.cpp:273-288): the boundary actions, then the gradients
 * into the gradient slots (read back with or_fv_intermediates). */
void or_compute_convective_gradient(void* fv, double* prim);
void or_compute_viscous_gradient(void* fv, double* prim);
void or_fv_intermediates(void* fv, double* grads, double* lim, double* flux,
                         double* left, double* right);
/* face gradients of the last viscous residual: 4 vars (T, vx, vy, vz) x 3 x F */
void or_fv_face_gradients(void* fv, double* out);
/* one flux calculator on one face; L, R: rho, p, T, e, u_n, u_t1, u_t2 in
   the face frame; out: mass, n-, t1-, t2-momentum, energy */
void or_flux_face(const or_fv_config* cfg, int dim, const double* L, const double* R,
                  double* out);
/* Dual<double> operations (see the op codes in ibis_oracle.cpp) */
void or_dual_eval(int op, const double* a, const double* b, double* out);
/* One explicit RK step (SO/runge_kutta.cpp:131-203).  cons/prim are the
   solution; dt_prev in/out.  Returns the dt taken. */
double or_rk_step(void* fv, int nstages, const double* a /* nstages x nstages, row i col j */,
                  const double* b, double cfl, double* cons, double* prim,
                  double* dt_prev, double t, double max_time, double plot_remaining);
void or_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
