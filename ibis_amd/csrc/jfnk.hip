// jfnk.hip -- FP64 Krylov solvers and the JFNK pseudo-transient step on the
// MI355X, around the residual and its tangent (residual.hip).
//
// Restates src/linear_algebra/linear_algebra/gmres.cpp:11-303 (GMRES and
// flexible GMRES with modified Gram-Schmidt Arnoldi, Givens rotations of
// the Hessenberg matrix on the host) and src/solvers/solvers/jfnk.cpp:24-93
// with the CFL schedules of src/solvers/solvers/cfl.cpp:9-83.
//
// Device design: the Krylov basis is one contiguous (m+1) x n FP64 block.
// Every inner product of the Arnoldi step is a deterministic two-stage
// reduction (fixed grid, fixed tree) that leaves its result in device
// memory; the following axpy reads the coefficient from there, so a whole
// Arnoldi step runs without a host round trip, and the host reads the new
// Hessenberg column once, for the rotations.  All arithmetic is FP64.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "ibis_internal.h"

namespace {

constexpr int RB = 256;      // threads per reduction block
constexpr int RNB = 1024;    // blocks of the first reduction stage (fixed)

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

inline unsigned nblk(int64_t n) { return (unsigned)((n + RB - 1) / RB); }

__device__ __forceinline__ double block_sum(double v, double* sm) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sm[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < RB / 64; k++) v += sm[k];
    }
    return v;
}

// stage 1: partial[b] = sum over the grid-stride slice of block b of a*b
__global__ void __launch_bounds__(RB)
    k_dot_partial(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                  double* __restrict__ partial) {
    __shared__ double sm[RB / 64];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)RNB * RB)
        s += a[i] * b[i];
    s = block_sum(s, sm);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// stage 2: out[0] = sum of the partials (sqrt of it when take_sqrt)
__global__ void __launch_bounds__(RB)
    k_dot_final(const double* __restrict__ partial, int take_sqrt, double* __restrict__ out) {
    __shared__ double sm[RB / 64];
    double s = 0.0;
    for (int k = threadIdx.x; k < RNB; k += RB) s += partial[k];
    s = block_sum(s, sm);
    if (threadIdx.x == 0) out[0] = take_sqrt ? sqrt(s) : s;
}

// add_scaled_vector(w, v, -h) with h in device memory (gmres.cpp:147)
__global__ void __launch_bounds__(RB)
    k_axpy_neg(int64_t n, double* __restrict__ w, const double* __restrict__ v,
               const double* __restrict__ h) {
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (i < n) w[i] += v[i] * (-h[0]);
}

// scale(w, v, 1 / h) with h in device memory (gmres.cpp:150)
__global__ void __launch_bounds__(RB)
    k_scale_inv(int64_t n, const double* __restrict__ w, double* __restrict__ v,
                const double* __restrict__ h) {
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (i < n) v[i] = w[i] * (1.0 / h[0]);
}

// scale(r0, v, 1 / beta) with a host factor, r0 = rhs - w (gmres.cpp:17-19)
__global__ void __launch_bounds__(RB)
    k_sub(int64_t n, const double* __restrict__ rhs, const double* __restrict__ w,
          double* __restrict__ r0) {
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (i < n) r0[i] = rhs[i] - w[i];
}

__global__ void __launch_bounds__(RB)
    k_scale(int64_t n, const double* __restrict__ a, double* __restrict__ out, double f) {
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (i < n) out[i] = a[i] * f;
}

__global__ void __launch_bounds__(RB) k_fill(int64_t n, double* __restrict__ a, double v) {
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (i < n) a[i] = v;
}

// w = V[:, 0:k] y (gemv, in column order), then x += w * 1.0
// (gmres.cpp:177-179)
__global__ void __launch_bounds__(RB)
    k_combine_add(int64_t n, int k, const double* __restrict__ V, const double* __restrict__ y,
                  double* __restrict__ w, double* __restrict__ x) {
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < k; c++) s += V[(size_t)c * n + i] * y[c];
    w[i] = s;
    x[i] += w[i] * 1.0;
}

// Jfnk::apply_update_ (jfnk.cpp:67-74): cons(cell, q) += dU[cell*nc + q]
__global__ void __launch_bounds__(RB)
    k_apply_update(int N, int Nt, int nc, const double* __restrict__ dU, double* __restrict__ cons) {
    const int i = blockIdx.x * RB + threadIdx.x;
    if (i >= N) return;
    for (int q = 0; q < nc; q++) cons[(size_t)q * Nt + i] += dU[(size_t)i * nc + q];
}

// ConservedQuantities::L2_norms partial sums (conserved_quantities.cpp:41-69):
// per block, the six sums (mass, mx, my, mz, energy, global)
__global__ void __launch_bounds__(RB)
    k_norm_partial(int N, int Nt, int dim, const double* __restrict__ dudt,
                   double* __restrict__ partial) {
    __shared__ double sm[RB / 64];
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = blockIdx.x * RB + threadIdx.x; i < N; i += RNB * RB) {
        const double m = dudt[i];
        const double mx = dudt[(size_t)Nt + i];
        const double my = dudt[2 * (size_t)Nt + i];
        const double mz = dim == 3 ? dudt[3 * (size_t)Nt + i] : 0.0;
        const double e = dudt[(size_t)(dim + 1) * Nt + i];
        s[0] += m * m;
        s[1] += mx * mx;
        s[2] += my * my;
        s[3] += mz * mz;
        s[4] += e * e;
        s[5] += m * m + mx * mx + my * my + mz * mz + e * e;
    }
    for (int q = 0; q < 6; q++) {
        const double t = block_sum(s[q], sm);
        if (threadIdx.x == 0) partial[(size_t)q * RNB + blockIdx.x] = t;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(RB)
    k_norm_final(const double* __restrict__ partial, double* __restrict__ out) {
    __shared__ double sm[RB / 64];
    for (int q = 0; q < 6; q++) {
        double s = 0.0;
        for (int k = threadIdx.x; k < RNB; k += RB) s += partial[(size_t)q * RNB + k];
        s = block_sum(s, sm);
        if (threadIdx.x == 0) out[q] = sqrt(s);
        __syncthreads();
    }
}

}  // namespace

// ===========================================================================
// GMRES / FGMRES
// ===========================================================================
struct ibis_gmres {
    int64_t n = 0;
    int m = 0;
    double tol = 0.0;
    double* V = nullptr;        // (m+1) x n Krylov basis
    double* Z = nullptr;        // m x n preconditioned vectors (FGMRES, on demand)
    double* r0 = nullptr;
    double* w = nullptr;
    double* partial = nullptr;  // RNB
    double* hdev = nullptr;     // m + 2 Hessenberg column entries (device)
    double* ydev = nullptr;     // m
    double* hhost = nullptr;    // pinned copy of hdev
    std::vector<double> H;      // (m+1) x m, row-major
    std::vector<double> cs, sn, g;
};

static ibis_status dot_dev(ibis_gmres* s, const double* a, const double* b, double* out,
                           bool take_sqrt, hipStream_t st) {
    k_dot_partial<<<RNB, RB, 0, st>>>(s->n, a, b, s->partial);
    k_dot_final<<<1, RB, 0, st>>>(s->partial, take_sqrt ? 1 : 0, out);
    HIP_TRY(hipGetLastError());
    return IBIS_OK;
}

static ibis_status read_dev(ibis_gmres* s, const double* dev, int k, hipStream_t st) {
    HIP_TRY(hipMemcpyAsync(s->hhost, dev, sizeof(double) * k, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return IBIS_OK;
}

void ibis_gmres_destroy(ibis_gmres* g) {
    if (!g) return;
    for (void* p : {(void*)g->V, (void*)g->Z, (void*)g->r0, (void*)g->w, (void*)g->partial,
                    (void*)g->hdev, (void*)g->ydev})
        if (p) (void)hipFree(p);
    if (g->hhost) (void)hipHostFree(g->hhost);
    delete g;
}

ibis_status ibis_gmres_create(int64_t num_vars, int32_t max_iters, double tol, ibis_gmres** out) {
    if (!out || num_vars <= 0 || max_iters <= 0)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_gmres_create: invalid argument");
    ibis_gmres* g = new ibis_gmres();
    g->n = num_vars;
    g->m = (int)std::min<int64_t>(num_vars, max_iters);  // gmres.cpp:90
    g->tol = tol;
    const size_t n = (size_t)num_vars, m = (size_t)g->m;
    hipError_t e = hipMalloc((void**)&g->V, sizeof(double) * n * (m + 1));
    if (e == hipSuccess) e = hipMalloc((void**)&g->r0, sizeof(double) * n);
    if (e == hipSuccess) e = hipMalloc((void**)&g->w, sizeof(double) * n);
    if (e == hipSuccess) e = hipMalloc((void**)&g->partial, sizeof(double) * 6 * RNB);
    if (e == hipSuccess) e = hipMalloc((void**)&g->hdev, sizeof(double) * (m + 2));
    if (e == hipSuccess) e = hipMalloc((void**)&g->ydev, sizeof(double) * (m + 1));
    if (e == hipSuccess) e = hipHostMalloc((void**)&g->hhost, sizeof(double) * (m + 2), hipHostMallocDefault);
    if (e != hipSuccess) {
        ibis_gmres_destroy(g);
        return ibis::fail(IBIS_ERR_HIP, hipGetErrorString(e));
    }
    g->H.assign((m + 1) * m, 0.0);
    g->cs.assign(m, 0.0);
    g->sn.assign(m, 0.0);
    g->g.assign(m + 1, 0.0);
    *out = g;
    return IBIS_OK;
}

// One (F)GMRES solve.  With inner != nullptr the Krylov vectors are
// preconditioned by an inner GMRES solve of P z = v from z = 0.
static ibis_status gmres_run(ibis_gmres* s, ibis_matvec_fn A, void* ctx_a, const double* rhs,
                             double* x, ibis_gmres* inner, ibis_matvec_fn P, void* ctx_p,
                             ibis_linear_result* res, hipStream_t st) {
    const int64_t n = s->n;
    const int m = s->m;
    ibis_stream_t sh = reinterpret_cast<ibis_stream_t>(st);
    if (inner && !s->Z) HIP_TRY(hipMalloc((void**)&s->Z, sizeof(double) * (size_t)n * m));
    std::fill(s->H.begin(), s->H.end(), 0.0);
    std::fill(s->g.begin(), s->g.end(), 0.0);
    auto Hm = [&](int i, int j) -> double& { return s->H[(size_t)i * m + j]; };
    // r0 = rhs - A x0, beta = |r0| (gmres.cpp:11-20, 130-135)
    IBIS_TRY(A(ctx_a, x, s->w, sh));
    k_sub<<<nblk(n), RB, 0, st>>>(n, rhs, s->w, s->r0);
    IBIS_TRY(dot_dev(s, s->r0, s->r0, s->hdev, true, st));
    IBIS_TRY(read_dev(s, s->hdev, 1, st));
    const double beta = s->hhost[0];
    ibis_linear_result r{0, 0, s->tol, beta};
    if (beta < s->tol) {
        r.success = 1;
        *res = r;
        return IBIS_OK;
    }
    s->g[0] = beta;
    k_scale<<<nblk(n), RB, 0, st>>>(n, s->r0, s->V, 1.0 / beta);
    for (int j = 0; j < m; j++) {
        double* vj = s->V + (size_t)j * n;
        if (inner) {  // gmres.cpp:260-265
            double* zj = s->Z + (size_t)j * n;
            k_fill<<<nblk(n), RB, 0, st>>>(n, zj, 0.0);
            ibis_linear_result ri;
            IBIS_TRY(gmres_run(inner, P, ctx_p, vj, zj, nullptr, nullptr, nullptr, &ri, st));
            IBIS_TRY(A(ctx_a, zj, s->w, sh));
        } else {
            IBIS_TRY(A(ctx_a, vj, s->w, sh));
        }
        // modified Gram-Schmidt, coefficients kept on the device
        for (int i = 0; i <= j; i++) {
            const double* vi = s->V + (size_t)i * n;
            IBIS_TRY(dot_dev(s, s->w, vi, s->hdev + i, false, st));
            k_axpy_neg<<<nblk(n), RB, 0, st>>>(n, s->w, vi, s->hdev + i);
        }
        IBIS_TRY(dot_dev(s, s->w, s->w, s->hdev + j + 1, true, st));
        k_scale_inv<<<nblk(n), RB, 0, st>>>(n, s->w, s->V + (size_t)(j + 1) * n, s->hdev + j + 1);
        IBIS_TRY(read_dev(s, s->hdev, j + 2, st));
        for (int i = 0; i <= j + 1; i++) Hm(i, j) = s->hhost[i];
        // previous rotations on the new column, then this step's rotation
        // (apply_rotations_to_hessenberg_, gmres.cpp:22-74)
        for (int i = 0; i < j; i++) {
            const double a = Hm(i, j), b = Hm(i + 1, j);
            Hm(i, j) = s->cs[i] * a + s->sn[i] * b;
            Hm(i + 1, j) = -s->sn[i] * a + s->cs[i] * b;
        }
        const double denom = std::sqrt(Hm(j, j) * Hm(j, j) + Hm(j + 1, j) * Hm(j + 1, j));
        const double si = Hm(j + 1, j) / denom;
        const double ci = Hm(j, j) / denom;
        s->cs[j] = ci;
        s->sn[j] = si;
        Hm(j, j) = ci * Hm(j, j) + si * Hm(j + 1, j);
        Hm(j + 1, j) = 0.0;
        const double gj = s->g[j];
        s->g[j] = ci * gj;
        s->g[j + 1] = -si * gj;
        const double residual = std::fabs(s->g[j + 1]);
        r.residual = residual / beta;
        r.n_iters = j + 1;
        if (residual < s->tol * beta) {
            r.success = 1;
            break;
        }
    }
    // upper-triangular solve H y = g, then x += V y (or Z y)
    const int k = r.n_iters;
    std::vector<double> y(k, 0.0);
    y[k - 1] = s->g[k - 1] / Hm(k - 1, k - 1);
    for (int i = k - 2; i >= 0; i--) {
        double sum = s->g[i];
        for (int c = i + 1; c < k; c++) sum -= Hm(i, c) * y[c];
        y[i] = sum / Hm(i, i);
    }
    HIP_TRY(hipMemcpyAsync(s->ydev, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, st));
    k_combine_add<<<nblk(n), RB, 0, st>>>(n, k, inner ? s->Z : s->V, s->ydev, s->w, x);
    HIP_TRY(hipGetLastError());
    // y is host memory: finish the copy before it goes out of scope
    HIP_TRY(hipStreamSynchronize(st));
    *res = r;
    return IBIS_OK;
}

ibis_status ibis_gmres_solve(ibis_gmres* g, ibis_matvec_fn A, void* ctx, const double* rhs,
                             double* x, ibis_linear_result* result, ibis_stream_t stream) {
    if (!g || !A || !rhs || !x || !result)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_gmres_solve: null argument");
    return gmres_run(g, A, ctx, rhs, x, nullptr, nullptr, nullptr, result,
                     reinterpret_cast<hipStream_t>(stream));
}

ibis_status ibis_fgmres_solve(ibis_gmres* g, ibis_matvec_fn A, void* ctx_a, const double* rhs,
                              double* x, ibis_gmres* inner, ibis_matvec_fn P, void* ctx_p,
                              ibis_linear_result* result, ibis_stream_t stream) {
    if (!g || !A || !rhs || !x || !result || !inner || !P)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_fgmres_solve: null argument");
    if (inner->n != g->n)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_fgmres_solve: inner solver size differs");
    return gmres_run(g, A, ctx_a, rhs, x, inner, P, ctx_p, result,
                     reinterpret_cast<hipStream_t>(stream));
}

// ===========================================================================
// CFL schedules and JFNK
// ===========================================================================
static ibis_status cfl_eval(const ibis_jfnk_desc& d, double t, double* prev_cfl,
                            double* prev_res, double* out) {
    switch (d.cfl_kind) {
        case IBIS_CFL_CONSTANT:  // cfl.cpp:9-14
            *out = d.cfl_value;
            return IBIS_OK;
        case IBIS_CFL_LINEAR: {  // cfl.cpp:16-41
            const int n = d.n_cfl_points;
            if (t <= d.cfl_times[0]) {
                *out = d.cfl_values[0];
                return IBIS_OK;
            }
            if (t >= d.cfl_times[n - 1]) {
                *out = d.cfl_values[n - 1];
                return IBIS_OK;
            }
            for (int i = 0; i + 1 < n; i++) {
                if (t >= d.cfl_times[i] && t < d.cfl_times[i + 1]) {
                    const double frac = (t - d.cfl_times[i]) / (d.cfl_times[i + 1] - d.cfl_times[i]);
                    *out = d.cfl_values[i] + frac * (d.cfl_values[i + 1] - d.cfl_values[i]);
                    return IBIS_OK;
                }
            }
            return ibis::fail(IBIS_ERR_INVALID, "linear CFL schedule: times not increasing");
        }
        case IBIS_CFL_RESIDUAL: {  // cfl.cpp:43-68
            if (t > d.growth_threshold) {
                *prev_res = t;
                *out = *prev_cfl;
                return IBIS_OK;
            }
            const double ratio = *prev_res / t;
            double c = *prev_cfl * std::pow(ratio, d.power);
            c = (d.max_cfl < c) ? d.max_cfl : c;      // Ibis::min
            c = (d.start_cfl < c) ? c : d.start_cfl;  // Ibis::max
            *prev_res = t;
            *prev_cfl = c;
            *out = c;
            return IBIS_OK;
        }
        default:
            return ibis::fail(IBIS_ERR_INVALID, "Unknown CFL schedule");
    }
}

ibis_status ibis_cfl_eval(const ibis_jfnk_desc* desc, double t, double* previous_cfl,
                          double* previous_residual, double* cfl) {
    if (!desc || !previous_cfl || !previous_residual || !cfl)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_cfl_eval: null argument");
    return cfl_eval(*desc, t, previous_cfl, previous_residual, cfl);
}

typedef ibis_steady_op SteadyCtx;

static ibis_status steady_mv(void* ctx, const double* v, double* w, ibis_stream_t st) {
    return ibis_steady_op_matvec(ctx, v, w, st);
}

ibis_status ibis_steady_op_matvec(void* ctx, const double* v, double* w, ibis_stream_t stream) {
    const ibis_steady_op* c = static_cast<const ibis_steady_op*>(ctx);
    if (!c) return ibis::fail(IBIS_ERR_INVALID, "ibis_steady_op_matvec: null operator");
    return ibis_steady_matvec(c->r, c->cons, v, w, c->dt_star, c->allow_reconstruction, stream);
}

namespace {
__global__ void __launch_bounds__(RB)
    k_dense_mv(int64_t n, const double* __restrict__ a, const double* __restrict__ v,
               double* __restrict__ w) {
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int64_t k = 0; k < n; k++) s += a[(size_t)i * n + k] * v[k];
    w[i] = s;
}
}  // namespace

ibis_status ibis_dense_matvec(void* ctx, const double* v, double* w, ibis_stream_t stream) {
    const ibis_dense_op* c = static_cast<const ibis_dense_op*>(ctx);
    if (!c || !c->a || !v || !w) return ibis::fail(IBIS_ERR_INVALID, "ibis_dense_matvec: null argument");
    k_dense_mv<<<nblk(c->n), RB, 0, reinterpret_cast<hipStream_t>(stream)>>>(c->n, c->a, v, w);
    HIP_TRY(hipGetLastError());
    return IBIS_OK;
}

struct ibis_jfnk {
    ibis_residual* r = nullptr;
    ibis_jfnk_desc d{};
    int dim = 0, N = 0, Nt = 0, nc = 0;
    int64_t nv = 0;
    ibis_gmres* outer = nullptr;
    ibis_gmres* inner = nullptr;
    double* rhs = nullptr;
    double* dU = nullptr;
    double* dudt = nullptr;
    double* partial = nullptr;
    double* norms_dev = nullptr;
    double* norms_host = nullptr;  // pinned
    double norms[IBIS_NNORMS] = {0};
    double initial_norms[IBIS_NNORMS] = {0};
    double prev_cfl = 0.0, prev_res = 1.0;  // ResidualBasedCfl state (cfl.cpp:43-50)
};

void ibis_jfnk_destroy(ibis_jfnk* j) {
    if (!j) return;
    ibis_gmres_destroy(j->outer);
    ibis_gmres_destroy(j->inner);
    for (void* p : {(void*)j->rhs, (void*)j->dU, (void*)j->dudt, (void*)j->partial,
                    (void*)j->norms_dev})
        if (p) (void)hipFree(p);
    if (j->norms_host) (void)hipHostFree(j->norms_host);
    delete j;
}

ibis_status ibis_jfnk_create(ibis_residual* r, const ibis_jfnk_desc* desc, ibis_jfnk** out) {
    if (!r || !desc || !out) return ibis::fail(IBIS_ERR_INVALID, "ibis_jfnk_create: null argument");
    const ibis_jfnk_desc& d = *desc;
    if (d.linear_solver != 0 && d.linear_solver != 1)
        return ibis::fail(IBIS_ERR_INVALID, "Unknown linear solver");
    if (d.cfl_kind < 0 || d.cfl_kind > 2) return ibis::fail(IBIS_ERR_INVALID, "Unknown CFL schedule");
    if (d.cfl_kind == IBIS_CFL_LINEAR && (d.n_cfl_points < 1 || d.n_cfl_points > IBIS_MAX_CFL_POINTS))
        return ibis::fail(IBIS_ERR_INVALID, "linear CFL schedule: bad number of points");
    int64_t sizes[5];
    IBIS_TRY(ibis_residual_sizes(r, sizes));
    ibis_jfnk* j = new ibis_jfnk();
    j->r = r;
    j->d = d;
    j->dim = (int)sizes[0];
    j->N = (int)sizes[1];
    j->Nt = (int)sizes[2];
    j->nc = j->dim + 2;
    j->nv = (int64_t)j->N * j->nc;
    j->prev_cfl = d.start_cfl;
    ibis_status s = ibis_gmres_create(j->nv, d.max_iters, d.tol, &j->outer);
    if (s == IBIS_OK && d.linear_solver == 1)
        s = ibis_gmres_create(j->nv, d.max_precond_iters, d.precond_tol, &j->inner);
    if (s != IBIS_OK) {
        ibis_jfnk_destroy(j);
        return s;
    }
    hipError_t e = hipMalloc((void**)&j->rhs, sizeof(double) * j->nv);
    if (e == hipSuccess) e = hipMalloc((void**)&j->dU, sizeof(double) * j->nv);
    if (e == hipSuccess) e = hipMalloc((void**)&j->dudt, sizeof(double) * j->nc * (size_t)j->Nt);
    if (e == hipSuccess) e = hipMemset(j->dudt, 0, sizeof(double) * j->nc * (size_t)j->Nt);
    if (e == hipSuccess) e = hipMalloc((void**)&j->partial, sizeof(double) * 6 * RNB);
    if (e == hipSuccess) e = hipMalloc((void**)&j->norms_dev, sizeof(double) * IBIS_NNORMS);
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&j->norms_host, sizeof(double) * IBIS_NNORMS, hipHostMallocDefault);
    if (e != hipSuccess) {
        ibis_jfnk_destroy(j);
        return ibis::fail(IBIS_ERR_HIP, hipGetErrorString(e));
    }
    *out = j;
    return IBIS_OK;
}

// eval_rhs + L2_norms (jfnk.cpp:24-29, 60-63)
static ibis_status jfnk_rhs_norms(ibis_jfnk* j, double* prim, hipStream_t st) {
    IBIS_TRY(ibis_steady_rhs(j->r, prim, j->dudt, j->rhs, 1, reinterpret_cast<ibis_stream_t>(st)));
    k_norm_partial<<<RNB, RB, 0, st>>>(j->N, j->Nt, j->dim, j->dudt, j->partial);
    k_norm_final<<<1, RB, 0, st>>>(j->partial, j->norms_dev);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(j->norms_host, j->norms_dev, sizeof(double) * IBIS_NNORMS,
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    // kernel order mass, mx, my, mz, energy, global
    for (int q = 0; q < IBIS_NNORMS; q++) j->norms[q] = j->norms_host[q];
    return IBIS_OK;
}

ibis_status ibis_jfnk_initialise(ibis_jfnk* j, double* cons, double* prim, double* norms,
                                 ibis_stream_t stream) {
    if (!j || !cons || !prim) return ibis::fail(IBIS_ERR_INVALID, "ibis_jfnk_initialise: null argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    IBIS_TRY(jfnk_rhs_norms(j, prim, st));
    std::memcpy(j->initial_norms, j->norms, sizeof(j->norms));
    if (norms) std::memcpy(norms, j->norms, sizeof(j->norms));
    return IBIS_OK;
}

ibis_status ibis_jfnk_step(ibis_jfnk* j, double* cons, double* prim, int64_t step,
                           ibis_linear_result* result, double* norms, double* dt_star_out,
                           ibis_stream_t stream) {
    if (!j || !cons || !prim) return ibis::fail(IBIS_ERR_INVALID, "ibis_jfnk_step: null argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    k_fill<<<nblk(j->nv), RB, 0, st>>>(j->nv, j->dU, 0.0);
    double stable_dt = 0.0;
    IBIS_TRY(ibis_estimate_dt(j->r, prim, &stable_dt, stream));
    // Jfnk::calculate_cfl (jfnk.h:39-47)
    double cfl = 0.0;
    if (j->d.cfl_kind == IBIS_CFL_RESIDUAL) {
        const double rel = j->norms[5] / j->initial_norms[5];
        IBIS_TRY(cfl_eval(j->d, rel, &j->prev_cfl, &j->prev_res, &cfl));
    } else {
        IBIS_TRY(cfl_eval(j->d, (double)step, &j->prev_cfl, &j->prev_res, &cfl));
    }
    const double dt_star = cfl * stable_dt;
    SteadyCtx sys{j->r, cons, dt_star, 1};
    SteadyCtx pre{j->r, cons, dt_star, 0};
    ibis_linear_result lr;
    if (j->d.linear_solver == 1)
        IBIS_TRY(gmres_run(j->outer, steady_mv, &sys, j->rhs, j->dU, j->inner, steady_mv, &pre,
                           &lr, st));
    else
        IBIS_TRY(gmres_run(j->outer, steady_mv, &sys, j->rhs, j->dU, nullptr, nullptr, nullptr,
                           &lr, st));
    k_apply_update<<<nblk(j->N), RB, 0, st>>>(j->N, j->Nt, j->nc, j->dU, cons);
    HIP_TRY(hipGetLastError());
    IBIS_TRY(ibis_cons_to_prim(j->r, cons, prim, stream));
    IBIS_TRY(jfnk_rhs_norms(j, prim, st));
    if (result) *result = lr;
    if (norms) std::memcpy(norms, j->norms, sizeof(j->norms));
    if (dt_star_out) *dt_star_out = dt_star;
    return IBIS_OK;
}
