// ibis_internal.h -- shared host-side definitions of libibis_gpu.
#ifndef IBIS_INTERNAL_H
#define IBIS_INTERNAL_H

#include <cstdint>
#include <vector>

#include "../../include/ibis_gpu.h"

struct ibis_grid {
    int dim = 2;
    int64_t N = 0, G = 0, F = 0;
    int ntags = 0, n_lsq = 3;
    std::vector<double> geom;  // 13 x F
    std::vector<int64_t> left, right;
    std::vector<int64_t> cf_off, cf, nbr;
    std::vector<int32_t> outsign;
    std::vector<double> centroid, volume, lsq;
    std::vector<int64_t> tag_off, bfaces, ghosts;
    std::vector<int64_t> cell_new_of_old, face_new_of_old;
};

void ibis_grid_fill_desc(const ibis_grid* g, ibis_mesh_desc* d);

// Library-internal entry points shared between the compilation units.
//
// A residual whose boundary tables hold `ntags` tags (kinds3: pre, viscous,
// post tables of ntags * IBIS_MAX_ACT entries each; pars3 likewise with
// IBIS_NPAR values per entry), for meshes with more tags than an
// ibis_fv_desc holds (a batch of instances); the rest of the desc applies.
ibis_status ibis_residual_create_tables(const ibis_mesh* m, const ibis_fv_desc* desc, int ntags,
                                        const int32_t* kinds3, const double* pars3,
                                        ibis_residual** out);
int ibis_mesh_ntags(const ibis_mesh* m);
// local time step of every valid cell (the terms estimate_dt minimises)
ibis_status ibis_residual_cell_dt(ibis_residual* r, const double* prim, double* dt_cells,
                                  ibis_stream_t stream);

// Runge-Kutta stage epilogue of the residual (batched stepping): with a
// valid cell's dU/dt in registers, the face kernel writes it to `dudt`
// (unless null), evaluates the stage value
//   u + k_0 (c_0 dt) + ... + k_{nk-1} (c_{nk-1} dt) [+ dU/dt (c_reg dt)]
// of each conserved row left to right (dt = dt[inst_of_cell[cell]]),
// writes it to cons_out (unless null) and its primitive rows to prim_out.
// With dt_cells_out set (ibis_residual_stage_dt_fusable), it also writes
// each valid cell's local time step of the new primitive state, the same
// bits ibis_residual_cell_dt gives for that state.  Ghost rows are not
// touched.
constexpr int IBIS_EPI_MAXK = 8;
struct ibis_stage_epilogue {
    const double* u = nullptr;
    const double* k[IBIS_EPI_MAXK] = {};
    double c[IBIS_EPI_MAXK] = {};
    int nk = 0;
    int use_reg = 0;  // the residual being computed enters as the last term
    double c_reg = 0.0;
    const int32_t* inst_of_cell = nullptr;
    const double* dt = nullptr;
    double* cons_out = nullptr;
    double* prim_out = nullptr;
    double R = 0.0, Cv = 0.0;
    double* dt_cells_out = nullptr;
    // 1: the ghost rows too, with the residual's own dU/dt term as +0 (its
    // ghost rows are zero), as a separate stage pass over them gives
    int ghost_rows = 0;
    // set by ibis_compute_dudt_stage: tile t takes rows N + t*ghost_per_tile ...
    int ghost_per_tile = 0;
    // filled in by ibis_compute_dudt_stage from the residual's description
    double dt_gamma = 0.0, dt_R = 0.0, dt_sf = 0.0;
};
// 1 when ibis_compute_dudt_stage fuses the epilogue into this residual's
// kernels (the inviscid second-order path with the per-entry face kernel
// in FP64); else the caller runs the stage update separately.
int ibis_residual_stage_fusable(const ibis_residual* r);
// 1 when, in addition, the epilogue can write the local time steps
// (dt_cells_out): an inviscid residual whose tile face lists fit in LDS with
// each entry's normal and area beside its flux.
int ibis_residual_stage_dt_fusable(const ibis_residual* r);
ibis_status ibis_compute_dudt_stage(ibis_residual* r, double* prim, double* dudt,
                                    const ibis_stage_epilogue* epi, ibis_stream_t stream);

// 1 when the value starts of the per-entry face kernel (face_kernel 1), at
// most maxf*5*n_cells + 5*n_ghost, fit in a 32-bit index; residuals of
// larger meshes keep the tile-list face kernel.  Exported for host tests.
extern "C" int ibis_entry_indices_fit(int64_t maxf, int64_t n_cells, int64_t n_ghost);
// Counter bumped by every ibis_residual_set_option / set_precision call, so
// that a captured HIP graph of the residual can tell it is stale.
uint64_t ibis_residual_generation(const ibis_residual* r);

namespace ibis {
// Record a failure message for ibis_last_error() and return the status.
ibis_status fail(ibis_status s, const char* msg);
}  // namespace ibis

#endif
