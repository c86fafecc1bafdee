// batch.hip -- many independent residual instances stepped together.
//
// The instances' grids are concatenated into one mesh (valid cells of
// instance 0, 1, ... then their ghost cells in the same order; faces and
// boundary tags concatenated; no face joins two instances), so one launch
// of each residual kernel covers every instance.  The explicit Runge-Kutta
// step of RungeKutta::take_step (src/solvers/solvers/runge_kutta.cpp:108-203)
// then needs only per-instance time steps: a segmented minimum of the local
// time steps, the dt rules of runge_kutta.cpp:108-119 per instance on the
// device, and stage updates that read their instance's dt.  A whole step
// (s residuals, the dt logic and the stage algebra) is captured once into
// a HIP graph and replayed, with no host round trip between steps.
//
// Every instance's arithmetic is the same as stepping it alone: the
// residual of a cell depends only on its own stencil, the minimum is
// order-independent, and the stage updates evaluate the same expressions
// element by element in the same order (u + k_j * (a_ij * dt)).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "ibis_internal.h"
#include "state.hpp"

namespace {

constexpr int BB = 256;
constexpr int MAXS = 8;

#define HIP_TRY(expr)                                                             \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) {                                                   \
            std::string m_ = std::string(#expr) + ": " + hipGetErrorString(e_);   \
            return ibis::fail(IBIS_ERR_HIP, m_.c_str());                          \
        }                                                                         \
    } while (0)

#define IBIS_TRY(expr)                      \
    do {                                    \
        ibis_status s_ = (expr);            \
        if (s_ != IBIS_OK) return s_;       \
    } while (0)

inline unsigned nblk(int64_t n) { return (unsigned)((n + BB - 1) / BB); }

__device__ __forceinline__ double kmin(double a, double b) { return (b < a) ? b : a; }

constexpr double DMAX = 1.79769313486231570815e+308;

// stable_dt[i] = min over the valid cells of instance i, in two passes:
// every chunk (at most CHUNK cells of one instance) reduces to partial[k],
// then each instance reduces its chunks chunk_off[i] .. chunk_off[i+1].
// 1024 cells: four loads per thread of a 256-thread block, so that even one
// small instance spreads its minimum over many blocks
constexpr int CHUNK = 1024;

__global__ void __launch_bounds__(BB)
    k_chunk_min(const int64_t* __restrict__ chunk_lo, const double* __restrict__ dt_cells,
                double* __restrict__ partial) {
    __shared__ double sm[BB / 64];
    const int k = blockIdx.x;
    const int64_t lo = chunk_lo[k], hi = chunk_lo[k + 1];
    double v = DMAX;
    for (int64_t c = lo + threadIdx.x; c < hi; c += BB) v = kmin(dt_cells[c], v);
    for (int o = 32; o > 0; o >>= 1) v = kmin(__shfl_down(v, o, 64), v);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < BB / 64; w++) v = kmin(sm[w], v);
        partial[k] = v;
    }
}

__global__ void __launch_bounds__(64)
    k_seg_min(const int32_t* __restrict__ chunk_off, const double* __restrict__ partial,
              double* __restrict__ stable) {
    const int inst = blockIdx.x;
    double v = DMAX;
    for (int k = chunk_off[inst] + threadIdx.x; k < chunk_off[inst + 1]; k += 64)
        v = kmin(partial[k], v);
    for (int o = 32; o > 0; o >>= 1) v = kmin(__shfl_down(v, o, 64), v);
    if (threadIdx.x == 0) stable[inst] = v;
}

// RungeKutta::estimate_dt (runge_kutta.cpp:108-119) per instance
__global__ void k_dt_rules(int n, const double* __restrict__ stable, double cfl,
                           const double* __restrict__ dt_prev, const double* __restrict__ t,
                           double max_time, double* __restrict__ dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = kmin(cfl * stable[i], 1.5 * dt_prev[i]);
    dt[i] = kmin(d, max_time - t[i]);
}

__global__ void k_advance(int n, const double* __restrict__ dt, double* __restrict__ dt_prev,
                          double* __restrict__ t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dt_prev[i] = dt[i];
    t[i] += dt[i];
}

struct StageArgs {
    const double* k[MAXS];
    double c[MAXS];  // tableau weight of k[j]
    int nk;
};

// Stage value u + k_0 * (c_0 * dt) + k_1 * (c_1 * dt) + ... of every
// conserved row of cell i, evaluated left to right as successive
// apply_time_derivative passes do (conserved_quantities.cpp:80-94,
// runge_kutta.cpp:151-191), then converted to the primitive rows by the
// residual's cons_to_prim expressions without storing the intermediate
// conserved state; cons_out (may alias u) receives it when not null.
// Cells [i0, i1) of the Nt rows.
template <int DIM>
__global__ void __launch_bounds__(BB)
    k_stage_prim(int Nt, int i0, int i1, const int32_t* __restrict__ inst_of_cell,
                 const double* __restrict__ dt, const double* u, StageArgs sa, double* cons_out,
                 double R, double Cv, double* __restrict__ prim) {
    const int i = i0 + blockIdx.x * BB + threadIdx.x;
    if (i >= i1) return;
    constexpr int NC = DIM + 2;
    const double d = dt[inst_of_cell[i]];
    double q[NC], out[ibis::NPRIM];
#pragma unroll
    for (int r = 0; r < NC; r++) {
        const size_t e = (size_t)r * Nt + i;
        double v = u[e];
        for (int j = 0; j < sa.nk; j++) v = v + sa.k[j][e] * (sa.c[j] * d);
        q[r] = v;
        if (cons_out) cons_out[e] = v;
    }
    ibis::cons_to_prim_cell<double, DIM>(q, R, Cv, out);
#pragma unroll
    for (int r = 0; r < ibis::NPRIM; r++) prim[(size_t)r * Nt + i] = out[r];
}

}  // namespace

struct ibis_batch {
    ibis_residual* r = nullptr;
    int n_inst = 0;
    int dim = 0, N = 0, Nt = 0, nc = 0;
    int64_t* cell_off = nullptr;    // device, n_inst + 1
    int32_t* inst_of_cell = nullptr;  // device, Nt
    int nchunks = 0;
    int64_t* chunk_lo = nullptr;    // device, nchunks + 1: cell range of each chunk
    int32_t* chunk_off = nullptr;   // device, n_inst + 1: chunk range of each instance
    double* partial = nullptr;      // device, nchunks
    double R = 0.0, Cv = 0.0;
    double* dt_cells = nullptr;
    double* stable = nullptr;
    double* dt = nullptr;
    double* dt_prev = nullptr;
    double* t = nullptr;
    // Runge-Kutta
    int s = 0;
    double a[MAXS * MAXS] = {0};
    double bw[MAXS] = {0};
    double cfl = 0.0, max_time = 0.0;
    double* cons = nullptr;
    double* prim = nullptr;
    std::vector<double*> k;
    double* flow_tmp = nullptr;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t graph_generation = 0;  // the residual's option generation at capture
    int fuse_stages = 1;  // stage updates in the residual's epilogue where its path allows
};

void ibis_batch_destroy(ibis_batch* b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    for (void* p : {(void*)b->cell_off, (void*)b->inst_of_cell, (void*)b->chunk_lo,
                    (void*)b->chunk_off, (void*)b->partial, (void*)b->dt_cells,
                    (void*)b->stable, (void*)b->dt, (void*)b->dt_prev, (void*)b->t,
                    (void*)b->flow_tmp})
        if (p) (void)hipFree(p);
    for (double* p : b->k) (void)hipFree(p);
    ibis_residual_destroy(b->r);
    delete b;
}

ibis_residual* ibis_batch_residual(ibis_batch* b) { return b ? b->r : nullptr; }

ibis_status ibis_batch_create(const ibis_mesh* m, int n_inst, const ibis_fv_desc* descs,
                              const int64_t* cell_off, const int64_t* ghost_off,
                              ibis_batch** out) {
    if (!m || n_inst <= 0 || !descs || !cell_off || !ghost_off || !out)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_batch_create: bad argument");
    // every instance shares the numerics; only the boundary actions differ
    const ibis_fv_desc& d0 = descs[0];
    int ntags = 0;
    for (int i = 0; i < n_inst; i++) {
        const ibis_fv_desc& d = descs[i];
        if (d.flux != d0.flux || d.order != d0.order || d.thermo != d0.thermo ||
            d.limiter != d0.limiter || d.viscous != d0.viscous || d.lim_eps != d0.lim_eps ||
            d.venkat_K != d0.venkat_K || d.ldfss_delta != d0.ldfss_delta ||
            d.signal_factor != d0.signal_factor || d.R != d0.R || d.Cv != d0.Cv ||
            d.Cp != d0.Cp || d.gamma != d0.gamma || d.mu0 != d0.mu0 || d.T0 != d0.T0 ||
            d.Ts != d0.Ts || d.Pr != d0.Pr)
            return ibis::fail(IBIS_ERR_INVALID,
                              "ibis_batch_create: instances must share flux, reconstruction, "
                              "limiter, gas and transport settings");
        if (d.ntags < 0 || d.ntags > IBIS_MAX_TAGS)
            return ibis::fail(IBIS_ERR_INVALID, "ibis_batch_create: bad tag count");
        ntags += d.ntags;
    }
    if (ntags != ibis_mesh_ntags(m))
        return ibis::fail(IBIS_ERR_INVALID,
                          "ibis_batch_create: the mesh's tags are not the instances' tags");
    const int ntab = ntags * IBIS_MAX_ACT;
    std::vector<int32_t> kinds(3 * (size_t)ntab, 0);
    std::vector<double> pars(3 * (size_t)ntab * IBIS_NPAR, 0.0);
    int t0 = 0;
    for (int i = 0; i < n_inst; i++) {
        const ibis_fv_desc& d = descs[i];
        const int nq = d.ntags * IBIS_MAX_ACT;
        const int q0 = t0 * IBIS_MAX_ACT;
        for (int q = 0; q < nq; q++) {
            kinds[q0 + q] = d.pre_kind[q];
            kinds[ntab + q0 + q] = d.visc_kind[q];
            kinds[2 * (size_t)ntab + q0 + q] = d.post_kind[q];
            for (int p = 0; p < IBIS_NPAR; p++) {
                pars[(size_t)(q0 + q) * IBIS_NPAR + p] = d.pre_par[q * IBIS_NPAR + p];
                pars[(size_t)(ntab + q0 + q) * IBIS_NPAR + p] = d.visc_par[q * IBIS_NPAR + p];
                pars[(size_t)(2 * ntab + q0 + q) * IBIS_NPAR + p] = d.post_par[q * IBIS_NPAR + p];
            }
        }
        t0 += d.ntags;
    }
    ibis_batch* b = new ibis_batch();
    ibis_status st = ibis_residual_create_tables(m, &d0, ntags, kinds.data(), pars.data(), &b->r);
    if (st != IBIS_OK) {
        delete b;
        return st;
    }
    int64_t sz[5];
    (void)ibis_residual_sizes(b->r, sz);
    b->dim = (int)sz[0];
    b->N = (int)sz[1];
    b->Nt = (int)sz[2];
    b->nc = b->dim + 2;
    b->n_inst = n_inst;
    const int64_t G = sz[4];
    if (cell_off[0] != 0 || cell_off[n_inst] != b->N || ghost_off[0] != 0 || ghost_off[n_inst] != G) {
        ibis_batch_destroy(b);
        return ibis::fail(IBIS_ERR_INVALID, "ibis_batch_create: instance ranges do not cover the mesh");
    }
    b->R = d0.R;
    b->Cv = d0.Cv;
    std::vector<int32_t> inst(b->Nt, -1);
    std::vector<int64_t> clo;
    std::vector<int32_t> coff(n_inst + 1, 0);
    for (int i = 0; i < n_inst; i++) {
        if (cell_off[i + 1] < cell_off[i] || ghost_off[i + 1] < ghost_off[i]) {
            ibis_batch_destroy(b);
            return ibis::fail(IBIS_ERR_INVALID, "ibis_batch_create: instance ranges not increasing");
        }
        for (int64_t c = cell_off[i]; c < cell_off[i + 1]; c++) inst[c] = i;
        for (int64_t g = ghost_off[i]; g < ghost_off[i + 1]; g++) inst[b->N + g] = i;
        for (int64_t c = cell_off[i]; c < cell_off[i + 1]; c += CHUNK) clo.push_back(c);
        coff[i + 1] = (int32_t)clo.size();
    }
    b->nchunks = (int)clo.size();
    clo.push_back(b->N);
    hipError_t e = hipMalloc((void**)&b->cell_off, sizeof(int64_t) * (n_inst + 1));
    if (e == hipSuccess) e = hipMemcpy(b->cell_off, cell_off, sizeof(int64_t) * (n_inst + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void**)&b->inst_of_cell, sizeof(int32_t) * b->Nt);
    if (e == hipSuccess) e = hipMemcpy(b->inst_of_cell, inst.data(), sizeof(int32_t) * b->Nt, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void**)&b->chunk_lo, sizeof(int64_t) * clo.size());
    if (e == hipSuccess) e = hipMemcpy(b->chunk_lo, clo.data(), sizeof(int64_t) * clo.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void**)&b->chunk_off, sizeof(int32_t) * (n_inst + 1));
    if (e == hipSuccess) e = hipMemcpy(b->chunk_off, coff.data(), sizeof(int32_t) * (n_inst + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void**)&b->partial, sizeof(double) * std::max(b->nchunks, 1));
    if (e == hipSuccess) e = hipMalloc((void**)&b->dt_cells, sizeof(double) * b->N);
    for (double** p : {&b->stable, &b->dt, &b->dt_prev, &b->t})
        if (e == hipSuccess) e = hipMalloc((void**)p, sizeof(double) * n_inst);
    if (e == hipSuccess) e = hipStreamCreate(&b->stream);
    if (e != hipSuccess) {
        ibis_batch_destroy(b);
        return ibis::fail(IBIS_ERR_HIP, hipGetErrorString(e));
    }
    *out = b;
    return IBIS_OK;
}

ibis_status ibis_batch_rk_init(ibis_batch* b, int nstages, const double* a, const double* bw,
                               double cfl, double max_time, double dt_init, double* cons,
                               double* prim) {
    if (!b || !a || !bw || !cons || !prim || nstages < 1 || nstages > MAXS)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_batch_rk_init: bad argument");
    b->s = nstages;
    std::memcpy(b->a, a, sizeof(double) * nstages * nstages);
    std::memcpy(b->bw, bw, sizeof(double) * nstages);
    b->cfl = cfl;
    b->max_time = max_time;
    b->cons = cons;
    b->prim = prim;
    const size_t ncv = (size_t)b->nc * b->Nt;
    while ((int)b->k.size() < nstages) {
        double* p = nullptr;
        HIP_TRY(hipMalloc(&p, ncv * sizeof(double)));
        HIP_TRY(hipMemset(p, 0, ncv * sizeof(double)));
        b->k.push_back(p);
    }
    if (!b->flow_tmp) {
        HIP_TRY(hipMalloc(&b->flow_tmp, (size_t)7 * b->Nt * sizeof(double)));
        HIP_TRY(hipMemset(b->flow_tmp, 0, (size_t)7 * b->Nt * sizeof(double)));
    }
    std::vector<double> v(b->n_inst, dt_init), z(b->n_inst, 0.0);
    HIP_TRY(hipMemcpy(b->dt_prev, v.data(), sizeof(double) * b->n_inst, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->t, z.data(), sizeof(double) * b->n_inst, hipMemcpyHostToDevice));
    // a new binding invalidates a captured step
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    return IBIS_OK;
}

// stage values u + sum_j k_j (c_j dt) of cells [i0, i1) -> primitive rows
static ibis_status stage_prim(ibis_batch* b, const StageArgs& sa, double* cons_out, double* prim,
                              int i0, int i1) {
    if (i1 <= i0) return IBIS_OK;
    const unsigned g = nblk(i1 - i0);
    if (b->dim == 3)
        k_stage_prim<3><<<g, BB, 0, b->stream>>>(b->Nt, i0, i1, b->inst_of_cell, b->dt, b->cons,
                                                  sa, cons_out, b->R, b->Cv, prim);
    else
        k_stage_prim<2><<<g, BB, 0, b->stream>>>(b->Nt, i0, i1, b->inst_of_cell, b->dt, b->cons,
                                                  sa, cons_out, b->R, b->Cv, prim);
    HIP_TRY(hipGetLastError());
    return IBIS_OK;
}

static ibis_status stage_prim(ibis_batch* b, const StageArgs& sa, double* cons_out, double* prim) {
    return stage_prim(b, sa, cons_out, prim, 0, b->Nt);
}

// the time step of every instance from the state in b->prim
static void batch_dt(ibis_batch* b, hipStream_t st) {
    if (b->nchunks) k_chunk_min<<<b->nchunks, BB, 0, st>>>(b->chunk_lo, b->dt_cells, b->partial);
    k_seg_min<<<b->n_inst, 64, 0, st>>>(b->chunk_off, b->partial, b->stable);
    k_dt_rules<<<nblk(b->n_inst), BB, 0, st>>>(b->n_inst, b->stable, b->cfl, b->dt_prev, b->t,
                                              b->max_time, b->dt);
}

// the stored stages of stage i's value (runge_kutta.cpp:151-191): k_0
// always, k_j (0 < j < i) unless a_ij < 1e-14
static StageArgs stage_terms(const ibis_batch* b, int i) {
    const int s = b->s;
    StageArgs sa{};
    sa.k[0] = b->k[0];
    sa.c[0] = b->a[i * s + 0];
    sa.nk = 1;
    for (int j = 1; j < i; j++) {
        if (b->a[i * s + j] < 1e-14) continue;  // runge_kutta.cpp:166
        sa.k[sa.nk] = b->k[j];
        sa.c[sa.nk] = b->a[i * s + j];
        sa.nk++;
    }
    return sa;
}

// 1 when a step runs with the stage updates in the face kernel and its
// last stage also writes the local time steps of the new state
static bool step_carries_dt(ibis_batch* b) {
    return b->fuse_stages && ibis_residual_stage_fusable(b->r) && b->s <= IBIS_EPI_MAXK &&
           ibis_residual_stage_dt_fusable(b->r);
}

// One step with the stage updates in the residual's face kernel: the
// time step first (it reads only the valid cells' state, which the
// residual does not change), then each residual's epilogue evaluates the
// next stage's value (or the step's final one) from the stored stages and
// its own dU/dt in registers, with the same expressions in the same order
// as the separate stage pass, and the ghost rows (spread over the tiles) as
// the separate pass leaves them.  The last stage's dU/dt is not stored.
// When step_carries_dt, the last stage's epilogue also writes the local
// time steps of the step's result (the same bits as ibis_residual_cell_dt
// of that state), so the next step starts from them and no step reads the
// state and face geometry again for its time step; ibis_batch_rk_steps
// computes them for the first step of each call.
static ibis_status batch_step_fused(ibis_batch* b) {
    hipStream_t st = b->stream;
    ibis_stream_t sh = reinterpret_cast<ibis_stream_t>(st);
    const int s = b->s;
    const bool carry_dt = step_carries_dt(b);
    if (!carry_dt) IBIS_TRY(ibis_residual_cell_dt(b->r, b->prim, b->dt_cells, sh));
    batch_dt(b, st);
    for (int i = 0; i < s; i++) {
        ibis_stage_epilogue e;
        e.u = b->cons;
        e.inst_of_cell = b->inst_of_cell;
        e.dt = b->dt;
        e.R = b->R;
        e.Cv = b->Cv;
        double* dudt = b->k[i];
        if (i + 1 < s) {
            // stage i+1: the stored k_0 .. k_{i-1}, then this residual's k_i
            const StageArgs all = stage_terms(b, i + 1);
            const bool reg = all.nk > 0 && all.k[all.nk - 1] == b->k[i];
            for (int j = 0; j < all.nk - (reg ? 1 : 0); j++) {
                e.k[e.nk] = all.k[j];
                e.c[e.nk] = all.c[j];
                e.nk++;
            }
            e.use_reg = reg ? 1 : 0;
            e.c_reg = reg ? all.c[all.nk - 1] : 0.0;
            e.prim_out = b->flow_tmp;
        } else {
            // the step's value: k_0 .. k_{s-2} stored, k_{s-1} in registers
            for (int j = 0; j + 1 < s; j++) {
                e.k[e.nk] = b->k[j];
                e.c[e.nk] = b->bw[j];
                e.nk++;
            }
            e.use_reg = 1;
            e.c_reg = b->bw[s - 1];
            e.cons_out = b->cons;
            e.prim_out = b->prim;
            if (carry_dt) e.dt_cells_out = b->dt_cells;
            dudt = nullptr;
        }
        // the ghost rows (their dU/dt rows stay zero) in the same epilogue,
        // as the separate pass evaluates them
        e.ghost_rows = 1;
        double* src = i == 0 ? b->prim : b->flow_tmp;
        IBIS_TRY(ibis_compute_dudt_stage(b->r, src, dudt, &e, sh));
    }
    k_advance<<<nblk(b->n_inst), BB, 0, st>>>(b->n_inst, b->dt, b->dt_prev, b->t);
    HIP_TRY(hipGetLastError());
    return IBIS_OK;
}

// one step of every instance on b->stream
static ibis_status batch_step(ibis_batch* b) {
    if (b->fuse_stages && ibis_residual_stage_fusable(b->r) && b->s <= IBIS_EPI_MAXK)
        return batch_step_fused(b);
    hipStream_t st = b->stream;
    ibis_stream_t sh = reinterpret_cast<ibis_stream_t>(st);
    const int s = b->s;
    IBIS_TRY(ibis_compute_dudt(b->r, b->prim, b->k[0], 1, sh));
    IBIS_TRY(ibis_residual_cell_dt(b->r, b->prim, b->dt_cells, sh));
    batch_dt(b, st);
    for (int i = 1; i < s; i++) {
        const StageArgs sa = stage_terms(b, i);
        IBIS_TRY(stage_prim(b, sa, nullptr, b->flow_tmp));
        IBIS_TRY(ibis_compute_dudt(b->r, b->flow_tmp, b->k[i], 1, sh));
    }
    StageArgs sa{};
    for (int i = 0; i < s; i++) {
        sa.k[i] = b->k[i];
        sa.c[i] = b->bw[i];
    }
    sa.nk = s;
    IBIS_TRY(stage_prim(b, sa, b->cons, b->prim));
    k_advance<<<nblk(b->n_inst), BB, 0, st>>>(b->n_inst, b->dt, b->dt_prev, b->t);
    HIP_TRY(hipGetLastError());
    return IBIS_OK;
}

ibis_status ibis_batch_rk_steps(ibis_batch* b, int nsteps, int use_graph) {
    if (!b || !b->cons || nsteps < 0)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_batch_rk_steps: call ibis_batch_rk_init first");
    // the state may have been written on any stream of the device
    HIP_TRY(hipDeviceSynchronize());
    // the local time steps of the state as the caller left it, for a first
    // step that takes them from the previous one's last stage
    if (nsteps > 0 && step_carries_dt(b))
        IBIS_TRY(ibis_residual_cell_dt(b->r, b->prim, b->dt_cells,
                                       reinterpret_cast<ibis_stream_t>(b->stream)));
    if (!use_graph) {
        for (int n = 0; n < nsteps; n++) IBIS_TRY(batch_step(b));
    } else {
        // an option or precision change of the residual since the capture
        // makes the captured kernels stale: capture the step again
        if (b->exec && b->graph_generation != ibis_residual_generation(b->r)) {
            (void)hipGraphExecDestroy(b->exec);
            (void)hipGraphDestroy(b->graph);
            b->exec = nullptr;
            b->graph = nullptr;
        }
        if (!b->exec) {
            b->graph_generation = ibis_residual_generation(b->r);
            HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            ibis_status s = batch_step(b);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (s != IBIS_OK) {
                if (g) (void)hipGraphDestroy(g);
                return s;
            }
            HIP_TRY(e);
            b->graph = g;
            HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        }
        for (int n = 0; n < nsteps; n++) HIP_TRY(hipGraphLaunch(b->exec, b->stream));
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    return IBIS_OK;
}

ibis_status ibis_batch_set_option(ibis_batch* b, const char* key, int64_t value) {
    if (!b || !key) return ibis::fail(IBIS_ERR_INVALID, "ibis_batch_set_option: null argument");
    if (std::string(key) != "fuse_stages")
        return ibis::fail(IBIS_ERR_INVALID, ("unknown batch option " + std::string(key)).c_str());
    b->fuse_stages = value ? 1 : 0;
    // a captured step holds the other form
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    return IBIS_OK;
}

ibis_status ibis_batch_times(ibis_batch* b, double* t, double* dt_prev) {
    if (!b) return ibis::fail(IBIS_ERR_INVALID, "ibis_batch_times: null batch");
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (t) HIP_TRY(hipMemcpy(t, b->t, sizeof(double) * b->n_inst, hipMemcpyDeviceToHost));
    if (dt_prev)
        HIP_TRY(hipMemcpy(dt_prev, b->dt_prev, sizeof(double) * b->n_inst, hipMemcpyDeviceToHost));
    return IBIS_OK;
}
