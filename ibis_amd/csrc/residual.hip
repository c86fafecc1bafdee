// residual.hip -- MI355X (gfx950) kernels and C ABI for ibis's cell residual.
//
// One residual evaluation (FiniteVolume<T>::compute_dudt,
// src/finite_volume/finite_volume/finite_volume.cpp:57-96) is:
//
//   k_bc          boundary faces -> ghost rows of prim (pre-reconstruction
//                 actions; boundaries/boundary.cpp:15-399)
//   k_cell_grad   per cell: LSQ gradients of the reconstruction variables
//                 and their limiters in one pass over the neighbour stencil
//                 (gradient.h:66-116, limiter.cpp:24-62)
//   k_face_flux   per face: MUSCL reconstruction of both sides, thermo
//                 closure, rotation into the face frame, flux calculator,
//                 rotation back (convective_flux.cpp:85-420)
//   [k_post_flux] constant-flux boundary override (boundary.cpp:405-443)
//   [k_bc, k_visc_grad, k_visc_flux]  viscous path (viscous_flux.cpp:16-184)
//   k_integral    per cell: deterministic gather of its faces' fluxes in the
//                 cell's own face order (finite_volume.cpp:221-255)
//
// Layout in HBM: all state, geometry and scratch are structure-of-arrays
// FP64 with int32 indices.  The cell->face and cell->neighbour lists are
// stored column-major ("ELL": entry j of cell i at j*N + i) so that a
// wavefront reading entry j of 64 consecutive cells issues one coalesced
// load.  Cells and faces are numbered in a locality order by the grid
// builder, so neighbour and face gathers hit the same cache lines.
//
// There are no atomics in the residual: the face->cell accumulation is a
// gather, so results are deterministic and independent of launch shape.
// All code is compiled without floating-point contraction so that it
// evaluates the reference's expression trees operation by operation.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "flux.hpp"
#include "state.hpp"
#include "flux_adj.hpp"
#include "ibis_internal.h"

namespace {

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            std::string m_ = std::string(#expr) + ": " + hipGetErrorString(e_);         \
            return ibis::fail(IBIS_ERR_HIP, m_.c_str());                                \
        }                                                                               \
    } while (0)

#define IBIS_TRY(expr)                \
    do {                              \
        ibis_status s_ = (expr);      \
        if (s_ != IBIS_OK) return s_; \
    } while (0)

// primitive rows (ibis::P_*, state.hpp) and gradient/limiter slots
using ibis::P_RHO;
using ibis::P_P;
using ibis::P_T;
using ibis::P_E;
using ibis::P_VX;
using ibis::P_VY;
using ibis::P_VZ;
using ibis::NPRIM;
enum { G_P = 0, G_RHO = 1, G_T = 2, G_U = 3, G_VX = 4, G_VY = 5, G_VZ = 6, NGRAD = 7 };

constexpr int BLOCK = 256;
constexpr int MAXF = 6;

struct DevMesh {
    int dim, nl, maxf;
    int N, G, Nt, F, nb;
    const int* L;
    const int* R;
    const double* geom;  // 13 x F
    const double* cen;   // 3 x Nt
    const double* vol;   // Nt
    const double* lsq;   // nl x N
    const int* ell_face; // maxf x N
    const int* ell_nbr;  // maxf x N
    const signed char* ell_sign;  // maxf x N
    const unsigned char* nfaces;  // N
    const int* b_face;   // nb
    const int* b_ghost;
    const int* b_valid;
    const int* b_tag;
    const int* b_of_ghost;  // G: boundary-face index of each ghost cell
    // tiles of `tile` consecutive cells and the faces touching them
    int ntiles, maxtf, ntf;           // ntf: entries of all tile lists
    const int* tile_face_off;         // ntiles + 1
    const int* tile_faces;            // face ids, ascending within a tile
    const unsigned short* cell_lds;   // maxf x N: position of a cell's face in its tile list
    const int* ent_faces;             // the tile lists in entry order (face_kernel 1)
    const unsigned short* ent_lds;    // maxf x N: position in the entry order
    // stencil halo of each tile (neighbours outside it) and, per (cell,
    // slot), the neighbour's index in the tile's LDS stencil
    int maxhalo;
    const int* halo_off;              // ntiles + 1
    const int* halo_ids;
    const unsigned short* nbr_lds;    // maxf x N
    // face-state records: slot of face f in the face list of its left/right cell
    const unsigned char* slotL;
    const unsigned char* slotR;
    int tile;      // cells per tile (and threads per block of the tile kernels)
    int lsq_from_cen;  // lsq equals its expression of the centroids bit for bit, so a
                       // kernel with the stencil's centroids in LDS forms it there
    int n_from_t;  // face normals equal cross(t1, t2) bit for bit (grid-built meshes)
    int planar2d;  // 2D, and every face has t1z = +0 and t2 = (+0, +0, +1) bit for bit
    int flat_z;    // 2D, and every cell and face centroid has the same finite z bits: every
                   // z difference of centroids (the face offsets' z, e_z) is +0, and the
                   // 2D kernels take it as 0 without loading it
};

constexpr int TILE = 256;     // cells per tile of a new mesh (the option "tile" picks others)
constexpr int FOLD_MAX_CELLS = 1 << 20;  // 2D meshes up to this size start with fold_boundary on
constexpr int MAXTILE = 512;  // largest tile: the tile kernels' block size is the tile
constexpr size_t MAX_LDS_BYTES = 160 * 1024;  // LDS of one gfx950 workgroup
constexpr int NXCD = 8;  // MI355X: 8 XCDs, workgroups dispatched round-robin across them

// Logical block for XCD locality: hardware block b runs on XCD b % 8, so
// handing XCD x the logical blocks [x*per, (x+1)*per) keeps neighbouring
// cells and faces in the same XCD's L2.  Launch grids are padded to 8*per
// blocks; a returned value >= nlogical means "no work".
__device__ __forceinline__ int xcd_block(int nlogical) {
    const int per = (nlogical + NXCD - 1) / NXCD;
    const int b = blockIdx.x;
    return (b % NXCD) * per + b / NXCD;
}

__device__ __forceinline__ double gget(const double* p, int stride, int row, int i) {
    return p[(size_t)row * stride + i];
}

// ---------------------------------------------------------------------------
// Boundary ghost-cell actions (src/finite_volume/finite_volume/boundaries/boundary.cpp)
// ---------------------------------------------------------------------------
struct GasParams {
    double R, Cv, Cp, gamma;
};

// S is double for the residual and ibis::Dual for its tangent (JVP); the
// geometry and the action parameters are plain doubles in both.
// The actions of tag t on boundary face f, in the order of the tag's list,
// through an accessor X: X.v(row) reads the valid cell's prim row (the
// actions only read it), X.g(row) and X.set(row, x) read and write the
// ghost's (entries no action writes keep what the ghost held).
template <class S, class Acc>
__device__ __forceinline__ void bc_actions_on(const DevMesh& m, const int* __restrict__ kinds,
                                              const double* __restrict__ pars,
                                              const GasParams& gp, int f, int t, Acc& X) {
    const int F = m.F;
    // one copy of the action switch (unrolled over the list it costs about
    // a hundred registers)
#pragma unroll 1
    for (int a = 0; a < IBIS_MAX_ACT; a++) {
        const int kind = kinds[t * IBIS_MAX_ACT + a];
        if (kind == IBIS_BC_NONE) continue;
        const double* par = pars + (t * IBIS_MAX_ACT + a) * IBIS_NPAR;
        switch (kind) {
            case IBIS_BC_FLOW_STATE_COPY:  // boundary.cpp:15-45
                X.set(P_T, S(par[0]));
                X.set(P_P, S(par[1]));
                X.set(P_RHO, S(par[2]));
                X.set(P_E, S(par[3]));
                X.set(P_VX, S(par[4]));
                X.set(P_VY, S(par[5]));
                X.set(P_VZ, S(par[6]));
                break;
            case IBIS_BC_INTERNAL_COPY:  // boundary.cpp:100-136
                X.set(P_T, X.v(P_T));
                X.set(P_P, X.v(P_P));
                X.set(P_RHO, X.v(P_RHO));
                X.set(P_E, X.v(P_E));
                X.set(P_VX, X.v(P_VX));
                X.set(P_VY, X.v(P_VY));
                X.set(P_VZ, X.v(P_VZ));
                break;
            case IBIS_BC_REFLECT_NORMAL: {  // boundary.cpp:142-206
                X.set(P_T, X.v(P_T));
                X.set(P_P, X.v(P_P));
                X.set(P_RHO, X.v(P_RHO));
                X.set(P_E, X.v(P_E));
                const S x = X.v(P_VX), y = X.v(P_VY), z = X.v(P_VZ);
                double nx = m.geom[0 * F + f], ny = m.geom[1 * F + f], nz = m.geom[2 * F + f];
                double t1x = m.geom[3 * F + f], t1y = m.geom[4 * F + f], t1z = m.geom[5 * F + f];
                double t2x = m.geom[6 * F + f], t2y = m.geom[7 * F + f], t2z = m.geom[8 * F + f];
                const S xs = -(x * nx + y * ny + z * nz);
                const S ys = x * t1x + y * t1y + z * t1z;
                const S zs = x * t2x + y * t2y + z * t2z;
                X.set(P_VX, xs * nx + ys * t1x + zs * t2x);
                X.set(P_VY, xs * ny + ys * t1y + zs * t2y);
                X.set(P_VZ, xs * nz + ys * t1z + zs * t2z);
                break;
            }
            case IBIS_BC_VEL_REFLECT:  // boundary.cpp:209-241
                X.set(P_VX, -X.v(P_VX));
                X.set(P_VY, -X.v(P_VY));
                X.set(P_VZ, -X.v(P_VZ));
                break;
            case IBIS_BC_FIX_TEMPERATURE:  // boundary.cpp:244-277
                X.set(P_T, 2.0 * S(par[0]) - X.v(P_T));
                break;
            case IBIS_BC_SUBSONIC_INFLOW: {  // boundary.cpp:286-342
                double nx = m.geom[0 * F + f], ny = m.geom[1 * F + f], nz = m.geom[2 * F + f];
                const S in_vx = S(par[4]), in_vy = S(par[5]), in_vz = S(par[6]);
                const S vel_correct =
                    nx * (in_vx - X.v(P_VX)) + ny * (in_vy - X.v(P_VY)) + nz * (in_vz - X.v(P_VZ));
                const S a = sqrt(S(gp.gamma * gp.R) * X.v(P_T));
                const S rho_ref = X.v(P_RHO);
                const S p_in = S(par[1]);
                const S p_face = 0.5 * (p_in + X.v(P_P) - rho_ref * a * vel_correct);
                const S rho_face = S(par[2]) + (p_face - p_in) / a / a;
                const S vx_face = in_vx - nx * (p_in - p_face) / (rho_ref * a);
                const S vy_face = in_vy - ny * (p_in - p_face) / (rho_ref * a);
                const S vz_face = in_vz - nz * (p_in - p_face) / (rho_ref * a);
                X.set(P_P, 2.0 * p_face - X.v(P_P));
                X.set(P_RHO, 2.0 * rho_face - X.v(P_RHO));
                X.set(P_T, X.g(P_P) / (X.g(P_RHO) * gp.R));
                X.set(P_E, gp.Cv * X.g(P_T));
                X.set(P_VX, 2.0 * vx_face - X.v(P_VX));
                X.set(P_VY, 2.0 * vy_face - X.v(P_VY));
                X.set(P_VZ, 2.0 * vz_face - X.v(P_VZ));
                break;
            }
            case IBIS_BC_SUBSONIC_OUTFLOW: {  // boundary.cpp:348-399
                double nx = m.geom[0 * F + f], ny = m.geom[1 * F + f], nz = m.geom[2 * F + f];
                const S p_face = S(par[0]);
                const S rho_ref = X.v(P_RHO);
                const S p_ref = X.v(P_P);
                const S a = sqrt(S(gp.gamma * gp.R) * X.v(P_T));
                const S rho_face = X.v(P_RHO) + (p_face - p_ref) / a / a;
                const S vx_face = X.v(P_VX) + nx * (p_ref - p_face) / (rho_ref * a);
                const S vy_face = X.v(P_VY) + ny * (p_ref - p_face) / (rho_ref * a);
                const S vz_face = X.v(P_VZ) + nz * (p_ref - p_face) / (rho_ref * a);
                X.set(P_P, 2.0 * p_face - X.v(P_P));
                X.set(P_RHO, 2.0 * rho_face - X.v(P_RHO));
                X.set(P_T, X.g(P_P) / (X.g(P_RHO) * gp.R));
                X.set(P_E, gp.Cv * X.g(P_T));
                X.set(P_VX, 2.0 * vx_face - X.v(P_VX));
                X.set(P_VY, 2.0 * vy_face - X.v(P_VY));
                X.set(P_VZ, 2.0 * vz_face - X.v(P_VZ));
                break;
            }
            default:
                break;
        }
    }
}

// the accessor of prim's rows: each action reads and writes only what it uses
template <class S>
struct PrimRows {
    S* prim;
    int Nt, va, gh;
    __device__ __forceinline__ S v(int row) const { return prim[(size_t)row * Nt + va]; }
    __device__ __forceinline__ S g(int row) const { return prim[(size_t)row * Nt + gh]; }
    __device__ __forceinline__ void set(int row, S x) { prim[(size_t)row * Nt + gh] = x; }
};

// the actions of boundary face b, in the order of its tag's list, on prim
template <class S>
__device__ __forceinline__ void bc_actions(const DevMesh& m, const int* __restrict__ kinds,
                                           const double* __restrict__ pars, const GasParams& gp,
                                           S* __restrict__ prim, int b) {
    const int f = m.b_face[b], t = m.b_tag[b];
    PrimRows<S> X{prim, m.Nt, m.b_valid[b], m.b_ghost[b]};
    bc_actions_on<S>(m, kinds, pars, gp, f, t, X);
}

template <class S>
__global__ void __launch_bounds__(BLOCK) k_bc(DevMesh m, const int* __restrict__ kinds,
                                              const double* __restrict__ pars, GasParams gp,
                                              S* __restrict__ prim) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.nb) return;
    bc_actions<S>(m, kinds, pars, gp, prim, b);
}

// ---------------------------------------------------------------------------
// Cell kernel: LSQ gradients (GR/gradient.h:66-116) of the reconstruction
// variables, then the limiter of each (FV/limiter.cpp:24-62, or the
// Venkatakrishnan limiter).
//   row[v]  : prim row of variable v
//   slot[v] : gradient/limiter slot of variable v
//   ngrad   : variables [0, ngrad) get their gradient computed here; the
//             rest use the gradient already stored in their slot (2D vz)
//   LIM     : 0 none, 1 Barth-Jespersen, 2 Venkatakrishnan
// ---------------------------------------------------------------------------
struct VarSet {
    int n, ngrad;
    int row[5];
    int slot[5];
};

// Cell kernel, one wavefront per (variable, 64 cells): wave v of a block
// handles variable v of cells [64*b, 64*b + 64).  The weight formulae are
// evaluated by each wave (they depend on geometry only); per-thread state is
// one variable, which keeps register use low enough for 8 waves per SIMD.
template <class S, int DIM, int LIM>
__global__ void __launch_bounds__(320)
    k_cell_grad_var(DevMesh m, VarSet vs, const S* __restrict__ prim, S* __restrict__ grad,
                    S* __restrict__ lim, double eps, double venkat_K, int nlogical) {
    const int lb = xcd_block(nlogical);
    if (lb >= nlogical) return;
    const int v = threadIdx.x >> 6;
    const int i = lb * 64 + (threadIdx.x & 63);
    const int N = m.N, Nt = m.Nt, F = m.F;
    if (i >= N) return;
    const int row = vs.row[v], slot = vs.slot[v];
    const double* cx = m.cen;
    const double* cy = m.cen + Nt;
    const double* cz = m.cen + 2 * (size_t)Nt;
    const S* u = prim + (size_t)row * Nt;
    const double xi = cx[i], yi = cy[i], zi = cz[i];
    const int nf = m.nfaces[i];
    const double r11 = m.lsq[i], r12 = m.lsq[N + i], r22 = m.lsq[2 * (size_t)N + i];
    double r13 = 0.0, r23 = 0.0, r33 = 0.0;
    if (DIM == 3) {
        r13 = m.lsq[3 * (size_t)N + i];
        r23 = m.lsq[4 * (size_t)N + i];
        r33 = m.lsq[5 * (size_t)N + i];
    }
    const double beta = (r12 * r23 - r13 * r23) / (r11 * r22);  // GR/gradient.h:89 (sic)
    const double r11s = r11 * r11;
    const double i22 = 1.0 / (r22 * r22);
    const double q12 = r12 / r11;
    const double i33 = 1.0 / (r33 * r33);
    const double q23 = r23 / r22;
    const S ui = u[i];
    S gx = S(0.0), gy = S(0.0), gz = S(0.0), umin = ui, umax = ui;
    for (int j = 0; j < nf; j++) {
        const int n = m.ell_nbr[(size_t)j * N + i];
        const double dx = cx[n] - xi;
        const double dy = cy[n] - yi;
        const double dz = cz[n] - zi;
        const double alpha_1 = dx / r11s;
        const double alpha_2 = i22 * (dy - q12 * dx);
        double alpha_3 = 0.0;
        if (DIM == 3) alpha_3 = i33 * (dz - q23 * dy + beta * dx);
        const double w_1 = alpha_1 - q12 * alpha_2 + beta * alpha_3;
        const double w_2 = alpha_2 - q23 * alpha_3;
        const double w_3 = alpha_3;
        const S un = u[n];
        const S diff = un - ui;
        gx += w_1 * diff;
        gy += w_2 * diff;
        gz += w_3 * diff;
        umin = ibis::kmin(umin, un);
        umax = ibis::kmax(umax, un);
    }
    S* g = grad + (size_t)slot * 3 * N;
    if (v < vs.ngrad) {
        g[i] = gx;
        g[N + i] = gy;
        g[2 * (size_t)N + i] = gz;
    } else {
        gx = g[i];
        gy = g[N + i];
        gz = g[2 * (size_t)N + i];
    }
    if (LIM == 0) return;
    S phi = S(1.0);
    S bn = S(1.0), bd = S(1.0);  // Venkatakrishnan: the minimum as a fraction
    double eps2 = 0.0;
    if (LIM == 2) {
        const double V = fabs(m.vol[i]);  // cell size h^dim = |V|
        const double K3 = venkat_K * venkat_K * venkat_K;
        eps2 = (DIM == 3) ? K3 * V : K3 * (V * sqrt(V));
    }
    for (int j = 0; j < nf; j++) {
        const int f = m.ell_face[(size_t)j * N + i];
        const double dx = m.geom[9 * (size_t)F + f] - xi;
        const double dy = m.geom[10 * (size_t)F + f] - yi;
        const double dz = m.geom[11 * (size_t)F + f] - zi;
        S d2 = gx * dx + gy * dy + gz * dz;
        if (LIM == 1) {  // Barth-Jespersen, FV/limiter.cpp:24-62
            const int sg = (d2 > S(0.0)) - (d2 < S(0.0));
            d2 = (double)sg * (fabs(d2) + eps);
            if (sg > 0)
                phi = ibis::kmin(phi, (umax - ui) / d2);
            else if (sg < 0)
                phi = ibis::kmin(phi, (umin - ui) / d2);
        } else {  // Venkatakrishnan
            S d1;
            if (d2 > S(0.0))
                d1 = umax - ui;
            else if (d2 < S(0.0))
                d1 = umin - ui;
            else
                continue;
            const S d1s = d1 * d1;
            const S d2s = d2 * d2;
            const S num = d1s + eps2 + 2.0 * (d1 * d2);
            const S den = d1s + 2.0 * d2s + d1 * d2 + eps2;
            if (ibis::venkat_take(num, den, bn, bd)) {
                bn = num;
                bd = den;
            }
        }
    }
    if (LIM == 2) phi = bn / bd;
    lim[(size_t)slot * N + i] = phi;
}

// ---------------------------------------------------------------------------
// Face kernel: reconstruction + flux (FV/convective_flux.cpp:85-420)
// ---------------------------------------------------------------------------
struct FaceParams {
    int thermo;   // 0 rho_p, 1 rho_T, 2 rho_u, 3 p_T
    int limited;  // limiter enabled
    int slotA, rowA, slotB, rowB;  // the two thermo reconstruction variables
    double R, Cv, gamma, delta;
};

// value + limiter * (grad . d), FV/convective_flux.cpp:243-255
template <class S>
__device__ __forceinline__ S lin_interp(S value, const S* grad, int N, int slot, int c,
                                        bool valid, S L, double dx, double dy, double dz) {
    S gx = S(0.0), gy = S(0.0), gz = S(0.0);
    if (valid) {
        const S* g = grad + (size_t)slot * 3 * N;
        gx = g[c];
        gy = g[N + c];
        gz = g[2 * (size_t)N + c];
    }
    return value + L * (gx * dx + gy * dy + gz * dz);
}

template <class S, bool RECON>
__device__ __forceinline__ void reconstruct(const DevMesh& m, const FaceParams& fp,
                                            const S* __restrict__ prim, const S* __restrict__ grad,
                                            const S* __restrict__ lim, int f, int c, S* st) {
    const int N = m.N, Nt = m.Nt, F = m.F;
    if (!RECON) {  // copy_reconstruct, FV/convective_flux.cpp:212-240
#pragma unroll
        for (int v = 0; v < NPRIM; v++) st[v] = prim[(size_t)v * Nt + c];
        return;
    }
    const bool valid = c < N;
    const bool limit = fp.limited && valid;
    const double dx = m.geom[9 * (size_t)F + f] - m.cen[c];
    const double dy = m.geom[10 * (size_t)F + f] - m.cen[Nt + c];
    const double dz = m.geom[11 * (size_t)F + f] - m.cen[2 * (size_t)Nt + c];
    const S vA = prim[(size_t)fp.rowA * Nt + c];
    const S vB = prim[(size_t)fp.rowB * Nt + c];
    const S LA = limit ? lim[(size_t)fp.slotA * N + c] : S(1.0);
    const S LB = limit ? lim[(size_t)fp.slotB * N + c] : S(1.0);
    const S a = lin_interp<S>(vA, grad, N, fp.slotA, c, valid, LA, dx, dy, dz);
    const S b = lin_interp<S>(vB, grad, N, fp.slotB, c, valid, LB, dx, dy, dz);
    S rho = S(0.0), p = S(0.0), T = S(0.0), e = S(0.0);
    switch (fp.thermo) {
        case IBIS_THERMO_RHO_P: p = a; rho = b; break;
        case IBIS_THERMO_RHO_T: rho = a; T = b; break;
        case IBIS_THERMO_RHO_U: rho = a; e = b; break;
        default: p = a; T = b; break;
    }
    ibis::thermo_update<S>(fp.thermo, fp.R, fp.Cv, rho, p, T, e);
    st[ibis::S_RHO] = rho;
    st[ibis::S_P] = p;
    st[ibis::S_T] = T;
    st[ibis::S_E] = e;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int slot = G_VX + k;
        const S L = limit ? lim[(size_t)slot * N + c] : S(1.0);
        st[ibis::S_U + k] = lin_interp<S>(prim[(size_t)(P_VX + k) * Nt + c], grad, N, slot, c,
                                          valid, L, dx, dy, dz);
    }
}

// Flux of face f in the global frame, rows in conserved order (mass,
// momentum..., energy): reconstruction of both sides, rotation to the face
// frame, flux calculator, rotation back (FV/convective_flux.cpp:85-178).
template <class S, int DIM, int FLUX, bool RECON>
__device__ __forceinline__ void face_flux_eval(const DevMesh& m, const FaceParams& fp,
                                               const S* __restrict__ prim,
                                               const S* __restrict__ grad,
                                               const S* __restrict__ lim, int f, S* res) {
    const int F = m.F;
    const int l = m.L[f], r = m.R[f];
    S Ls[NPRIM], Rs[NPRIM];
    reconstruct<S, RECON>(m, fp, prim, grad, lim, f, l, Ls);
    reconstruct<S, RECON>(m, fp, prim, grad, lim, f, r, Rs);
    const double nx = m.geom[f], ny = m.geom[(size_t)F + f], nz = m.geom[2 * (size_t)F + f];
    const double t1x = m.geom[3 * (size_t)F + f], t1y = m.geom[4 * (size_t)F + f],
                 t1z = m.geom[5 * (size_t)F + f];
    const double t2x = m.geom[6 * (size_t)F + f], t2y = m.geom[7 * (size_t)F + f],
                 t2z = m.geom[8 * (size_t)F + f];
    // transform_to_local_frame, UT/vector3.cpp:142-153
#pragma unroll
    for (int side = 0; side < 2; side++) {
        S* s = side ? Rs : Ls;
        const S a = s[4], b = s[5], c = s[6];
        s[4] = a * nx + b * ny + c * nz;
        s[5] = a * t1x + b * t1y + c * t1z;
        s[6] = a * t2x + b * t2y + c * t2z;
    }
    S out[5];
    constexpr bool T3 = DIM == 3;
    if (FLUX == IBIS_FLUX_AUSMDV)
        ibis::flux_ausmdv<S, T3>(Ls, Rs, fp.R, fp.Cv, fp.gamma, out);
    else if (FLUX == IBIS_FLUX_HANEL)
        ibis::flux_hanel<S, T3>(Ls, Rs, fp.R, fp.Cv, fp.gamma, out);
    else if (FLUX == IBIS_FLUX_LDFSS2)
        ibis::flux_ldfss2<S, T3>(Ls, Rs, fp.R, fp.Cv, fp.gamma, fp.delta, out);
    else
        ibis::flux_rusanov<S, T3>(Ls, Rs, fp.R, fp.Cv, fp.gamma, out);
    // flux::transform_to_global, FV/convective_flux.cpp:148-177
    const S px = out[1], py = out[2];
    S pz = S(0.0);
    if (T3) pz = out[3];
    res[0] = out[0];
    res[1] = px * nx + py * t1x + pz * t2x;
    res[2] = px * ny + py * t1y + pz * t2y;
    if (T3) {
        res[3] = px * nz + py * t1z + pz * t2z;
        res[4] = out[4];
    } else {
        res[3] = out[4];
    }
}

template <class S, int DIM, int FLUX, bool RECON, bool PREMUL>
__global__ void __launch_bounds__(BLOCK)
    k_face_flux(DevMesh m, FaceParams fp, const S* __restrict__ prim, const S* __restrict__ grad,
                const S* __restrict__ lim, S* __restrict__ flux) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    const int F = m.F;
    if (f >= F) return;
    constexpr int NC = DIM + 2;
    S res[NC];
    face_flux_eval<S, DIM, FLUX, RECON>(m, fp, prim, grad, lim, f, res);
    const double A = PREMUL ? m.geom[12 * (size_t)F + f] : 1.0;
#pragma unroll
    for (int q = 0; q < NC; q++) flux[(size_t)q * F + f] = PREMUL ? res[q] * A : res[q];
}

// Fused face flux + surface integral for one tile of m.tile cells.
// Phase 1: every face touching the tile is evaluated once per tile and its
// flux*area kept in LDS (faces on the tile border are evaluated by both
// neighbouring tiles).  Phase 2: each cell sums its faces from LDS in its
// own face order, so the result is the gather of
// FV/finite_volume.cpp:221-255 with no flux array in HBM and no atomics.
template <int DIM, int FLUX, bool RECON>
__global__ void __launch_bounds__(MAXTILE)
    k_face_tile(DevMesh m, FaceParams fp, const double* __restrict__ prim,
                const double* __restrict__ grad, const double* __restrict__ lim,
                double* __restrict__ dudt) {
    extern __shared__ double sflux[];  // NC x maxtf
    const int t = xcd_block(m.ntiles);
    if (t >= m.ntiles) return;
    constexpr int NC = DIM + 2;
    const int F = m.F, N = m.N, Nt = m.Nt, S = m.maxtf;
    const int f0 = m.tile_face_off[t];
    const int nfl = m.tile_face_off[t + 1] - f0;
    for (int k = threadIdx.x; k < nfl; k += m.tile) {
        const int f = m.tile_faces[f0 + k];
        double res[NC];
        face_flux_eval<double, DIM, FLUX, RECON>(m, fp, prim, grad, lim, f, res);
        const double A = m.geom[12 * (size_t)F + f];
#pragma unroll
        for (int q = 0; q < NC; q++) sflux[q * S + k] = res[q] * A;
    }
    __syncthreads();
    const int c = t * m.tile + threadIdx.x;
    if (c >= N) return;
    double d[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) d[q] = 0.0;
    const int nf = m.nfaces[c];
    for (int j = 0; j < nf; j++) {
        const int k = m.cell_lds[(size_t)j * N + c];
        const int s = m.ell_sign[(size_t)j * N + c];
#pragma unroll
        for (int q = 0; q < NC; q++) {
            const double fa = sflux[q * S + k];
            d[q] += (s > 0) ? -fa : fa;  // == flux * (-area * outsign), exactly
        }
    }
    const double V = m.vol[c];
#pragma unroll
    for (int q = 0; q < NC; q++) dudt[(size_t)q * Nt + c] = d[q] / V;
}

// ---------------------------------------------------------------------------
// Face-state path for the 2nd-order residual without viscous or
// flux-override actions.
//
//   pack[c]  (8 doubles, every cell incl. ghosts): the 5 reconstruction
//            variables (thermo pair A, B, then vx, vy, vz) and the centroid,
//            so that a neighbour is read with four 16-byte loads.
//   FL[f], FR[f] (5 doubles each): the reconstructed values on the left and
//            right side of face f (value + limiter * grad . (x_f - x_c),
//            FV/convective_flux.cpp:243-255).
//
// The cell kernel computes gradients and limiters and writes the face
// values it owns; the face kernel reads FL/FR in face order (coalesced),
// applies the thermo closure, and never touches gradients or centroids.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(BLOCK)
    k_pack_geom(int Nt, const double* __restrict__ cen, double* __restrict__ pack) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nt) return;
    pack[(size_t)i * 8 + 5] = cen[i];
    pack[(size_t)i * 8 + 6] = cen[Nt + i];
    pack[(size_t)i * 8 + 7] = cen[2 * (size_t)Nt + i];
}

__global__ void __launch_bounds__(BLOCK)
    k_pack(int Nt, VarSet vs, const double* __restrict__ prim, double* __restrict__ pack) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nt) return;
#pragma unroll
    for (int v = 0; v < 5; v++) pack[(size_t)i * 8 + v] = prim[(size_t)vs.row[v] * Nt + i];
}

// Ghost side of boundary faces: grad = 0 and limiter = 1 for a ghost cell,
// so the reconstructed value is value + 1.0 * (0.0*dx + 0.0*dy + 0.0*dz).
// LAYOUT 0: face-indexed FL[f] / FR[f].  LAYOUT 1: cell-major records
// (record of valid cell c, slot j at (c*S + j); ghost g at (N*S + g - N),
// S = maxf) in FL, read through the per-face slot of each side.
__device__ __forceinline__ size_t rec_index(const DevMesh& m, int c, int slot) {
    return c < m.N ? (size_t)c * m.maxf + slot : (size_t)m.N * m.maxf + (c - m.N);
}
// LAYOUT 2: slot-major structure of arrays, value v of slot j of valid cell c
// at (j*5 + v)*N + c (one store instruction of a wavefront writes whole cache
// lines); ghost g at maxf*5*N + v*G + (g - N).
__device__ __forceinline__ size_t soa_index(const DevMesh& m, int c, int slot, int v) {
    return c < m.N ? ((size_t)slot * 5 + v) * m.N + c
                   : (size_t)m.maxf * 5 * m.N + (size_t)v * m.G + (c - m.N);
}

template <int LAYOUT, class W>
__device__ __forceinline__ void ghost_face_values(const DevMesh& m, const VarSet& vs,
                                                  const double* __restrict__ prim,
                                                  W* __restrict__ FL, W* __restrict__ FR, int b) {
    const int f = m.b_face[b], g = m.b_ghost[b];
    const int F = m.F, Nt = m.Nt;
    const double dx = m.geom[9 * (size_t)F + f] - m.cen[g];
    const double dy = m.geom[10 * (size_t)F + f] - m.cen[Nt + g];
    const double dz = m.geom[11 * (size_t)F + f] - m.cen[2 * (size_t)Nt + g];
    W* rec = LAYOUT == 0 ? (m.L[f] == g ? FL : FR) + (size_t)f * 5 : FL + rec_index(m, g, 0) * 5;
#pragma unroll
    for (int v = 0; v < 5; v++) {
        const double gx = 0.0, gy = 0.0, gz = 0.0;
        const double val = prim[(size_t)vs.row[v] * Nt + g] + 1.0 * (gx * dx + gy * dy + gz * dz);
        if (LAYOUT == 2)
            FL[soa_index(m, g, 0, v)] = W(val);
        else
            rec[v] = W(val);
    }
}

template <int LAYOUT, class W = double>
__global__ void __launch_bounds__(BLOCK)
    k_ghost_fs(DevMesh m, VarSet vs, const double* __restrict__ prim, W* __restrict__ FL,
               W* __restrict__ FR) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.nb) return;
    ghost_face_values<LAYOUT, W>(m, vs, prim, FL, FR, b);
}

// k_bc then k_ghost_fs for each boundary face in one thread: the ghost's
// face values are computed from the state its actions have just written
// (every ghost belongs to exactly one boundary face)
template <int LAYOUT, class W = double>
__global__ void __launch_bounds__(BLOCK)
    k_bc_ghost_fs(DevMesh m, const int* __restrict__ kinds, const double* __restrict__ pars,
                  GasParams gp, VarSet vs, double* __restrict__ prim, W* __restrict__ FL,
                  W* __restrict__ FR) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.nb) return;
    bc_actions<double>(m, kinds, pars, gp, prim, b);
    ghost_face_values<LAYOUT, W>(m, vs, prim, FL, FR, b);
}

// Boundary faces of the fused viscous path, one thread per face: the
// pre-reconstruction actions, the ghost's face values (layout 2), its
// reconstruction variables as they are now (gconv, 5 x G: the convective
// gradients read these), then the viscous actions, whose result stays in
// prim's ghost rows as after the reference's second ghost fill.  The
// actions of a face read only its valid cell and write only its ghost, so
// running them back to back per face is running each pass over all faces.
// prim_c (the adjoint's record of the ghosts between the two passes): when
// not null, the ghost's whole row is also stored there, 7 x G.
template <class W = double>
__global__ void __launch_bounds__(BLOCK)
    k_bc_visc_fs(DevMesh m, const int* __restrict__ kinds, const double* __restrict__ pars,
                 int ntab, int has_pre, int has_visc, GasParams gp, VarSet vs,
                 double* __restrict__ prim, W* __restrict__ FL, double* __restrict__ gconv,
                 double* __restrict__ prim_c) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.nb) return;
    if (has_pre) bc_actions<double>(m, kinds, pars, gp, prim, b);
    ghost_face_values<2, W>(m, vs, prim, FL, FL, b);
    const int g = m.b_ghost[b];
#pragma unroll
    for (int v = 0; v < 5; v++)
        gconv[(size_t)v * m.G + (g - m.N)] = prim[(size_t)vs.row[v] * m.Nt + g];
    if (prim_c) {
#pragma unroll
        for (int k = 0; k < NPRIM; k++)
            prim_c[(size_t)k * m.G + (g - m.N)] = prim[(size_t)k * m.Nt + g];
    }
    if (has_visc) bc_actions<double>(m, kinds + ntab, pars + (size_t)ntab * IBIS_NPAR, gp, prim, b);
}

// Offsets x_face - x_cell per (cell, face slot), column-major like the other
// cell lists: offs[(j*3 + k)*N + i].  Same expression as the kernels use.
__global__ void __launch_bounds__(BLOCK) k_offsets(DevMesh m, double* __restrict__ offs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = m.N, F = m.F, Nt = m.Nt;
    if (i >= N) return;
    const double xi = m.cen[i], yi = m.cen[Nt + i], zi = m.cen[2 * (size_t)Nt + i];
    for (int j = 0; j < m.nfaces[i]; j++) {
        const int f = m.ell_face[(size_t)j * N + i];
        offs[(size_t)(j * 3 + 0) * N + i] = m.geom[9 * (size_t)F + f] - xi;
        offs[(size_t)(j * 3 + 1) * N + i] = m.geom[10 * (size_t)F + f] - yi;
        offs[(size_t)(j * 3 + 2) * N + i] = m.geom[11 * (size_t)F + f] - zi;
    }
}

// W is the storage type of the face values (float in the reduced-precision
// mode, layout 2 only); gradients and limiters are FP64 in both.
// values (A, B, vx, vy, vz) and centroid of cell c: from the packed
// 64-byte record, or (PACKED = false) straight from the prim rows and the
// centroid rows
template <bool PACKED>
__device__ __forceinline__ void load_cell(const DevMesh& m, const VarSet& vs,
                                          const double* __restrict__ pack,
                                          const double* __restrict__ prim, int c, double* u,
                                          double& x, double& y, double& z, bool loadz = true) {
    if (PACKED) {
        const double2* P = reinterpret_cast<const double2*>(pack);
        const double2 a0 = P[(size_t)c * 4 + 0], a1 = P[(size_t)c * 4 + 1];
        const double2 a2 = P[(size_t)c * 4 + 2], a3 = P[(size_t)c * 4 + 3];
        u[0] = a0.x;
        u[1] = a0.y;
        u[2] = a1.x;
        u[3] = a1.y;
        u[4] = a2.x;
        x = a2.y;
        y = a3.x;
        z = a3.y;
    } else {
        const int Nt = m.Nt;
#pragma unroll
        for (int v = 0; v < 5; v++) u[v] = prim[(size_t)vs.row[v] * Nt + c];
        x = m.cen[c];
        y = m.cen[Nt + c];
        z = loadz ? m.cen[2 * (size_t)Nt + c] : 0.0;
    }
}

// Stage block lb's cell stencil in LDS: its m.tile cells' and its halo's
// values and centroids as 8 planes of SS = m.tile + maxhalo doubles (own
// cells first, then the halo in the tile's halo order).  Each thread
// gathers its own cell and its first halo cell together, then writes both;
// further halo cells one per pass.  Ends with the block barrier.
// The boundary pass folded into the cell kernel's stencil staging (the
// option fold_boundary): a ghost's only neighbour is the valid cell of its
// boundary face, so exactly one tile has it in its stencil halo, and the
// block of that tile runs the ghost's boundary actions while it stages it.
struct BcFold {
    const int* kinds;     // the pre-reconstruction then the viscous action tables
    const double* pars;
    int ntab, has_pre, has_visc;
    GasParams gp;
    double* prim;         // the state: the ghost rows are written through this
    double* FL;           // the face values (layout 2, FP64): the ghost entries
    double* prim_cg;      // when set (the adjoint's forward pass): each ghost's row after
                          // the pre-reconstruction actions, 7 x G
};

// the accessor of the folded pass: the ghost's row in registers, loaded
// before the actions; the valid cell's read from the state
struct GhostRegs {
    const double* prim;
    int Nt, va;
    double gr[NPRIM];
    __device__ __forceinline__ double v(int row) const { return prim[(size_t)row * Nt + va]; }
    __device__ __forceinline__ double g(int row) const { return gr[row]; }
    __device__ __forceinline__ void set(int row, double x) { gr[row] = x; }
};

// prim row `row` of a register copy (row is a kernel argument, the same in
// every lane: a chain of selects keeps the copy in registers)
__device__ __forceinline__ double row_entry(const double (&r)[NPRIM], int row) {
    double x = r[0];
#pragma unroll
    for (int k = 1; k < NPRIM; k++) x = row == k ? r[k] : x;
    return x;
}

// Ghost c of the stencil halo: its boundary face's pre-reconstruction
// actions, its face values (as ghost_face_values), its viscous actions and
// its prim row written back, as k_bc_ghost_fs / k_bc_visc_fs do.  u: the
// reconstruction variables after the pre-reconstruction actions; tv: T,
// vx, vy, vz after the viscous ones.
__device__ __forceinline__ void fold_ghost(const DevMesh& m, const VarSet& vs, const BcFold& bf,
                                           int c, double (&u)[5], double (&tv)[4]) {
    const int N = m.N, Nt = m.Nt, F = m.F;
    const int b = m.b_of_ghost[c - N];
    const int f = m.b_face[b], t = m.b_tag[b];
    GhostRegs X{bf.prim, Nt, m.b_valid[b], {}};
#pragma unroll
    for (int k = 0; k < NPRIM; k++) X.gr[k] = bf.prim[(size_t)k * Nt + c];
    if (bf.has_pre) bc_actions_on<double>(m, bf.kinds, bf.pars, bf.gp, f, t, X);
    const double dx = m.geom[9 * (size_t)F + f] - m.cen[c];
    const double dy = m.geom[10 * (size_t)F + f] - m.cen[Nt + c];
    const double dz = m.geom[11 * (size_t)F + f] - m.cen[2 * (size_t)Nt + c];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        u[v] = row_entry(X.gr, vs.row[v]);
        const double gx = 0.0, gy = 0.0, gz = 0.0;
        bf.FL[soa_index(m, c, 0, v)] = u[v] + 1.0 * (gx * dx + gy * dy + gz * dz);
    }
    if (bf.prim_cg) {
#pragma unroll
        for (int k = 0; k < NPRIM; k++) bf.prim_cg[(size_t)k * m.G + (c - N)] = X.gr[k];
    }
    if (bf.has_visc)
        bc_actions_on<double>(m, bf.kinds + bf.ntab, bf.pars + (size_t)bf.ntab * IBIS_NPAR, bf.gp,
                              f, t, X);
    tv[0] = X.gr[P_T];
    tv[1] = X.gr[P_VX];
    tv[2] = X.gr[P_VY];
    tv[3] = X.gr[P_VZ];
#pragma unroll
    for (int k = 0; k < NPRIM; k++) bf.prim[(size_t)k * Nt + c] = X.gr[k];
}

// FOLD: the ghosts of the tile's stencil halo (a suffix of the ascending
// halo list) through fold_ghost into their stencil entries, before the
// cell kernel loads anything else, so that the actions' registers are not
// added to those the kernel keeps live; VISC: also T, vx, vy, vz.  The
// staging that follows leaves these entries alone; its barrier publishes
// them.
template <bool VISC>
__device__ __forceinline__ void fold_halo_ghosts(const DevMesh& m, const VarSet& vs,
                                                 const BcFold& bf, int lb, double* soff) {
    const int N = m.N, Nt = m.Nt, SS = m.tile + m.maxhalo;
    const int h0 = m.halo_off[lb], nh = m.halo_off[lb + 1] - h0;
    for (int k = threadIdx.x; k < nh; k += m.tile) {
        const int c = m.halo_ids[h0 + k];
        if (c < N) continue;
        double u[5], tv[4];
        fold_ghost(m, vs, bf, c, u, tv);
        const int e = m.tile + k;
#pragma unroll
        for (int v = 0; v < 5; v++) soff[v * SS + e] = u[v];
        soff[5 * SS + e] = m.cen[c];
        soff[6 * SS + e] = m.cen[Nt + c];
        soff[7 * SS + e] = m.cen[2 * (size_t)Nt + c];
        if (VISC) {
#pragma unroll
            for (int q = 0; q < 4; q++) soff[(8 + q) * SS + e] = tv[q];
        }
    }
}

template <bool PACKED, bool FOLD>
__device__ __forceinline__ void stage_stencil(const DevMesh& m, const VarSet& vs,
                                              const double* __restrict__ pack,
                                              const double* __restrict__ prim, int lb, int i,
                                              double* soff) {
    // one stencil entry per thread and pass: the block is the tile.  On a
    // flat 2D mesh the centroids' z plane is not staged (every z difference
    // is +0 there and the kernels take it so).
    const int N = m.N, SS = m.tile + m.maxhalo;
    const int h0 = m.halo_off[lb], nh = m.halo_off[lb + 1] - h0;
    const bool loadz = !(m.dim == 2 && m.flat_z);
    auto put = [&](int k, const double* u, double x, double y, double z) {
#pragma unroll
        for (int v = 0; v < 5; v++) soff[v * SS + k] = u[v];
        soff[5 * SS + k] = x;
        soff[6 * SS + k] = y;
        if (loadz) soff[7 * SS + k] = z;
    };
    // FOLD: the halo's ghosts are staged already (fold_halo_ghosts)
    const int kh = threadIdx.x;
    const int ch = kh < nh ? m.halo_ids[h0 + kh] : 0;
    const bool hv = kh < nh && !(FOLD && ch >= N);
    double u0[5], x0, y0, z0, u1[5], x1, y1, z1;
    if (i < N) load_cell<PACKED>(m, vs, pack, prim, i, u0, x0, y0, z0, loadz);
    if (hv) load_cell<PACKED>(m, vs, pack, prim, ch, u1, x1, y1, z1, loadz);
    if (i < N) put(threadIdx.x, u0, x0, y0, z0);
    if (hv) put(m.tile + kh, u1, x1, y1, z1);
    for (int k = kh + m.tile; k < nh; k += m.tile) {
        const int c = m.halo_ids[h0 + k];
        if (FOLD && c >= N) continue;
        double u[5], x, y, z;
        load_cell<PACKED>(m, vs, pack, prim, c, u, x, y, z, loadz);
        put(m.tile + k, u, x, y, z);
    }
    __syncthreads();
}

// The stencil of the fused viscous path: 12 planes of SS doubles -- the
// five reconstruction variables (a ghost's from gconv, its state before the
// viscous boundary actions, which the convective gradients read), the
// centroid, then T, vx, vy, vz as the viscous gradients read them (a
// ghost's after the viscous actions, in prim).  Ends with the block barrier.
template <bool FOLD>
__device__ __forceinline__ void stage_stencil_visc(const DevMesh& m, const VarSet& vs,
                                                   const double* __restrict__ prim,
                                                   const double* __restrict__ gconv, int lb, int i,
                                                   double* soff) {
    const int N = m.N, Nt = m.Nt, G = m.G, SS = m.tile + m.maxhalo;
    const int h0 = m.halo_off[lb], nh = m.halo_off[lb + 1] - h0;
    const bool loadz = !(m.dim == 2 && m.flat_z);  // as stage_stencil
    auto stage = [&](int c, int k) {
        if (FOLD && c >= N) return;  // staged by fold_halo_ghosts
        double u[5], tv[4];
#pragma unroll
        for (int v = 0; v < 5; v++)
            u[v] = c < N ? prim[(size_t)vs.row[v] * Nt + c] : gconv[(size_t)v * G + (c - N)];
        tv[0] = prim[(size_t)P_T * Nt + c];
        tv[1] = prim[(size_t)P_VX * Nt + c];
        tv[2] = prim[(size_t)P_VY * Nt + c];
        tv[3] = prim[(size_t)P_VZ * Nt + c];
#pragma unroll
        for (int v = 0; v < 5; v++) soff[v * SS + k] = u[v];
        soff[5 * SS + k] = m.cen[c];
        soff[6 * SS + k] = m.cen[Nt + c];
        if (loadz) soff[7 * SS + k] = m.cen[2 * (size_t)Nt + c];
#pragma unroll
        for (int q = 0; q < 4; q++) soff[(8 + q) * SS + k] = tv[q];
    };
    if (i < N) stage(i, threadIdx.x);
    for (int k = threadIdx.x; k < nh; k += m.tile) stage(m.halo_ids[h0 + k], m.tile + k);
    __syncthreads();
}

// In 2D every LSQ gradient's z component is +0 exactly (its weight is 0 and
// the sum starts at +0; the gradient buffers start zeroed), so it is taken as
// 0 there instead of being loaded (the fused cell kernel does not store it).
template <class S, int DIM>
__device__ __forceinline__ S grad_z(const S* __restrict__ g, int N, int c) {
    return DIM == 3 ? g[2 * (size_t)N + c] : S(0.0);
}

__device__ __forceinline__ bool viscous_slot(int slot) {
    return slot == G_T || slot == G_VX || slot == G_VY || slot == G_VZ;
}

// VISC (with SN, layout 2): the fused viscous path's cell kernel -- the
// stencil of stage_stencil_visc, and before the convective work the viscous
// gradients of T, vx, vy, vz (the LSQ gradient of k_cell_grad_var, same
// expressions and order) into their slots of vgrad.  When vgrad is grad,
// the convective slots do not overwrite them afterwards (the viscous
// gradients are computed after the convective ones in the reference and
// overwrite them); the adjoint passes its own vgrad and keeps both.
template <int DIM, int LIM, int LAYOUT, bool OFFS, class W = double, bool PACKED = true,
          bool PF = false, bool SN = false, bool VISC = false, bool FOLD = false>
__global__ void __launch_bounds__(MAXTILE)
    k_cell_fs(DevMesh m, VarSet vs, const double* __restrict__ pack, double* __restrict__ grad,
              double* __restrict__ lim, double eps, double venkat_K, int nlogical,
              int write_slots, W* __restrict__ FL, W* __restrict__ FR,
              const double* __restrict__ offs, const double* __restrict__ prim, int stage,
              const double* __restrict__ gconv, double* __restrict__ vgrad,
              BcFold bf = BcFold{}) {
    // FOLD (FP64, SN): the boundary pass runs first, on the tile's halo ghosts
    // (fold_halo_ghosts, bf), then the staging of the rest of the stencil.
    // stage != 0: the limiter pass keeps each face offset in LDS
    // (maxf x 3 doubles per thread) for the reconstruction pass, so the
    // offsets are read from memory once.
    // SN: the block (one tile of m.tile cells) first copies its cells' and its
    // stencil halo's values and centroids into LDS (8 planes of
    // m.tile + maxhalo doubles) with one load per cell, and the neighbour loop
    // reads them from there.  With stage != 0 as well, the offsets take the
    // same LDS once the whole block has finished the neighbour loop.
    extern __shared__ double soff[];
    const int lb = xcd_block(nlogical);
    if (lb >= nlogical) return;
    static_assert(!FOLD || (SN && std::is_same<W, double>::value && !PACKED),
                  "the folded boundary pass: LDS stencil, FP64 face values from the state");
    if (FOLD) fold_halo_ghosts<VISC>(m, vs, bf, lb, soff);
    const int TS = m.tile;  // cells per block: one tile
    const int i = lb * TS + threadIdx.x;
    const int N = m.N, F = m.F;
    const int SS = TS + m.maxhalo;  // plane length of the LDS stencil
    // per-cell inputs that do not depend on the staged stencil, loaded
    // before the staging loop and its barrier so that their latency
    // overlaps them (a thread past the last cell reads cell 0's and
    // discards them)
    const int il = i < N ? i : 0;
    const int nf = m.nfaces[il];
    // the LSQ factors: formed from the staged stencil's centroids below
    // when the mesh's equal that expression (SN), else read
    const bool lsq_lds = SN && m.lsq_from_cen;
    double r11 = 1.0, r12 = 0.0, r22 = 1.0, r13 = 0.0, r23 = 0.0, r33 = 0.0;
    if (!lsq_lds) {
        r11 = m.lsq[il];
        r12 = m.lsq[N + il];
        r22 = m.lsq[2 * (size_t)N + il];
        if (DIM == 3) {
            r13 = m.lsq[3 * (size_t)N + il];
            r23 = m.lsq[4 * (size_t)N + il];
            r33 = m.lsq[5 * (size_t)N + il];
        }
    }
    int kst[MAXF];  // the neighbours' positions in the LDS stencil
    if (SN) {
#pragma unroll
        for (int j = 0; j < MAXF; j++) kst[j] = j < m.maxf ? m.nbr_lds[(size_t)j * N + il] : 0;
    }
    if (SN && VISC)
        stage_stencil_visc<FOLD>(m, vs, prim, gconv, lb, i, soff);
    else if (SN)
        stage_stencil<PACKED, FOLD>(m, vs, pack, prim, lb, i, soff);
    // with the offsets staged over the stencil, a thread past the last cell
    // stays to the second barrier (it works on LDS slots that are not read
    // back and stores nothing)
    const bool reuse = SN && LIM != 0 && stage;
    if (i >= N && !reuse) return;
    constexpr int NG = DIM == 3 ? 5 : 4;  // variables whose gradient is computed here
    double ui[5];
    double xi, yi, zi;
    if (SN) {
#pragma unroll
        for (int v = 0; v < 5; v++) ui[v] = soff[v * SS + threadIdx.x];
        xi = soff[5 * SS + threadIdx.x];
        yi = soff[6 * SS + threadIdx.x];
        zi = soff[7 * SS + threadIdx.x];
    } else {
        load_cell<PACKED>(m, vs, pack, prim, i, ui, xi, yi, zi);
    }
    if (SN && lsq_lds) {  // GR/gradient.h:119-162, the grid builder's expressions and order
        double sxx = 0.0, sxy = 0.0, sxz = 0.0, syy = 0.0, syz = 0.0, szz = 0.0;
#pragma unroll
        for (int j = 0; j < MAXF; j++) {
            if (j < nf) {
                const int k = kst[j];
                const double dx = soff[5 * SS + k] - xi;
                const double dy = soff[6 * SS + k] - yi;
                const double dz = soff[7 * SS + k] - zi;
                sxx += dx * dx;
                sxy += dx * dy;
                sxz += dx * dz;
                syy += dy * dy;
                syz += dy * dz;
                szz += dz * dz;
            }
        }
        r11 = sqrt(sxx);
        r12 = 1.0 / r11 * sxy;
        r22 = sqrt(syy - r12 * r12);
        if (DIM == 3) {
            r13 = 1.0 / r11 * sxz;
            r23 = 1.0 / r22 * (syz - r12 / r11 * sxz);
            r33 = sqrt(szz - (r13 * r13 + r23 * r23));
        }
    }
    const double beta = (r12 * r23 - r13 * r23) / (r11 * r22);  // GR/gradient.h:89 (sic)
    const double r11s = r11 * r11;
    const double i22 = 1.0 / (r22 * r22);
    const double q12 = r12 / r11;
    const double i33 = 1.0 / (r33 * r33);
    const double q23 = r23 / r22;
    double gx[5], gy[5], gz[5], umin[5], umax[5];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        gx[v] = gy[v] = gz[v] = 0.0;
        umin[v] = umax[v] = ui[v];
    }
    // one neighbour's contribution to the LSQ gradient and the min / max
    auto accumulate = [&](const double* un, double xn, double yn, double zn) {
        const double dx = xn - xi;
        const double dy = yn - yi;
        const double dz = zn - zi;
        const double alpha_1 = dx / r11s;
        const double alpha_2 = i22 * (dy - q12 * dx);
        double alpha_3 = 0.0;
        if (DIM == 3) alpha_3 = i33 * (dz - q23 * dy + beta * dx);
        const double w_1 = alpha_1 - q12 * alpha_2 + beta * alpha_3;
        const double w_2 = alpha_2 - q23 * alpha_3;
        const double w_3 = alpha_3;
#pragma unroll
        for (int v = 0; v < 5; v++) {
            if (v < NG) {
                const double diff = un[v] - ui[v];
                gx[v] += w_1 * diff;
                gy[v] += w_2 * diff;
                gz[v] += w_3 * diff;
            }
            umin[v] = ibis::kmin(umin[v], un[v]);
            umax[v] = ibis::kmax(umax[v], un[v]);
        }
    };
    // VISC: the 2D convective vz gradient is its slot's value from before
    // this call's viscous gradients
    double vz_old[3] = {0.0, 0.0, 0.0};
    if (VISC && DIM == 2 && i < N) {
        const double* g = grad + (size_t)G_VZ * 3 * N;
        vz_old[0] = g[i];
        vz_old[1] = g[N + i];
        vz_old[2] = grad_z<double, DIM>(g, N, i);
    }
    if (VISC && SN && i < N) {
        double tx[4] = {0.0, 0.0, 0.0, 0.0}, ty[4] = {0.0, 0.0, 0.0, 0.0},
               tz[4] = {0.0, 0.0, 0.0, 0.0}, ti[4];
#pragma unroll
        for (int q = 0; q < 4; q++) ti[q] = soff[(8 + q) * SS + threadIdx.x];
#pragma unroll
        for (int j = 0; j < MAXF; j++) {
            if (j < nf) {
                const int k = kst[j];
                const double dx = soff[5 * SS + k] - xi;
                const double dy = soff[6 * SS + k] - yi;
                const double dz = soff[7 * SS + k] - zi;
                const double alpha_1 = dx / r11s;
                const double alpha_2 = i22 * (dy - q12 * dx);
                double alpha_3 = 0.0;
                if (DIM == 3) alpha_3 = i33 * (dz - q23 * dy + beta * dx);
                const double w_1 = alpha_1 - q12 * alpha_2 + beta * alpha_3;
                const double w_2 = alpha_2 - q23 * alpha_3;
                const double w_3 = alpha_3;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const double diff = soff[(8 + q) * SS + k] - ti[q];
                    tx[q] += w_1 * diff;
                    ty[q] += w_2 * diff;
                    if (DIM == 3) tz[q] += w_3 * diff;
                }
            }
        }
        // in 2D the z components are not stored: they are +0 (grad_z)
        const int vslot[4] = {G_T, G_VX, G_VY, G_VZ};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            double* g = vgrad + (size_t)vslot[q] * 3 * N;
            g[i] = tx[q];
            g[N + i] = ty[q];
            if (DIM == 3) g[2 * (size_t)N + i] = tz[q];
        }
    }
    if (SN) {
#pragma unroll
        for (int j = 0; j < MAXF; j++) {
            if (j < nf) {
                const int k = kst[j];
                double un[5];
#pragma unroll
                for (int v = 0; v < 5; v++) un[v] = soff[v * SS + k];
                accumulate(un, soff[5 * SS + k], soff[6 * SS + k], soff[7 * SS + k]);
            }
        }
    } else if (PF) {
        // every neighbour index first, then each neighbour's values loaded
        // while the previous neighbour is accumulated (same order of sums)
        int nbr[MAXF];
#pragma unroll
        for (int j = 0; j < MAXF; j++) nbr[j] = j < nf ? m.ell_nbr[(size_t)j * N + i] : i;
        double un[5], xn, yn, zn;
        load_cell<PACKED>(m, vs, pack, prim, nbr[0], un, xn, yn, zn);
#pragma unroll
        for (int j = 0; j < MAXF; j++) {
            if (j < nf) {
                double cu[5];
#pragma unroll
                for (int v = 0; v < 5; v++) cu[v] = un[v];
                const double cx = xn, cy = yn, cz = zn;
                if (j + 1 < MAXF && j + 1 < nf)
                    load_cell<PACKED>(m, vs, pack, prim, nbr[j + 1 < MAXF ? j + 1 : j], un, xn, yn,
                                      zn);
                accumulate(cu, cx, cy, cz);
            }
        }
    } else {
        for (int j = 0; j < nf; j++) {
            const int n = m.ell_nbr[(size_t)j * N + i];
            double un[5], xn, yn, zn;
            load_cell<PACKED>(m, vs, pack, prim, n, un, xn, yn, zn);
            accumulate(un, xn, yn, zn);
        }
    }
    if (reuse) {
        __syncthreads();  // the stencil is read: its LDS now takes the offsets
        if (i >= N) return;
    }
#pragma unroll
    for (int v = 0; v < 5; v++) {
        double* g = grad + (size_t)vs.slot[v] * 3 * N;
        if (v < NG) {
            if (write_slots && !(VISC && vgrad == grad && viscous_slot(vs.slot[v]))) {
                g[i] = gx[v];
                g[N + i] = gy[v];
                if (DIM == 3) g[2 * (size_t)N + i] = gz[v];  // 2D: +0, not stored (grad_z)
            }
        } else if (VISC) {  // 2D vz, VISC: the slot's value read above
            gx[v] = vz_old[0];
            gy[v] = vz_old[1];
            gz[v] = vz_old[2];
        } else {  // 2D vz: the gradient stored in its slot
            gx[v] = g[i];
            gy[v] = g[N + i];
            gz[v] = g[2 * (size_t)N + i];
        }
    }
    // x_face - x_cell of the cell's face in slot j (z: +0 on a flat 2D mesh)
    const bool zflat = DIM == 2 && m.flat_z;
    auto face_offset = [&](int j, double& dx, double& dy, double& dz) {
        if (OFFS) {
            dx = offs[(size_t)(j * 3 + 0) * N + i];
            dy = offs[(size_t)(j * 3 + 1) * N + i];
            dz = zflat ? 0.0 : offs[(size_t)(j * 3 + 2) * N + i];
        } else {
            const int f = m.ell_face[(size_t)j * N + i];
            dx = m.geom[9 * (size_t)F + f] - xi;
            dy = m.geom[10 * (size_t)F + f] - yi;
            dz = zflat ? 0.0 : m.geom[11 * (size_t)F + f] - zi;
        }
    };
    double phi[5] = {1.0, 1.0, 1.0, 1.0, 1.0};
    if (LIM != 0) {
        double eps2 = 0.0;
        // Venkatakrishnan: each variable's minimum ratio as the fraction
        // bn / bd, divided out after the face loop
        double bn[5] = {1.0, 1.0, 1.0, 1.0, 1.0}, bd[5] = {1.0, 1.0, 1.0, 1.0, 1.0};
        if (LIM == 2) {
            const double V = fabs(m.vol[i]);  // cell size h^dim = |V|
            const double K3 = venkat_K * venkat_K * venkat_K;
            eps2 = (DIM == 3) ? K3 * V : K3 * (V * sqrt(V));
        }
        // the next face's offset is loaded while this face is limited
        double ndx, ndy, ndz;
        face_offset(0, ndx, ndy, ndz);
#pragma unroll
        for (int j = 0; j < MAXF; j++) {
            if (j >= nf) break;
            const double dx = ndx, dy = ndy, dz = ndz;
            if (j + 1 < MAXF) face_offset(j + 1 < nf ? j + 1 : j, ndx, ndy, ndz);
            if (stage) {
                soff[(j * 3 + 0) * TS + threadIdx.x] = dx;
                soff[(j * 3 + 1) * TS + threadIdx.x] = dy;
                soff[(j * 3 + 2) * TS + threadIdx.x] = dz;
            }
#pragma unroll
            for (int v = 0; v < 5; v++) {
                double d2 = gx[v] * dx + gy[v] * dy + gz[v] * dz;
                // written without branches: the quotient of a face that the
                // reference skips (d2 zero or NaN) is formed and discarded,
                // so every lane runs the same instructions
                if (LIM == 1) {  // Barth-Jespersen, FV/limiter.cpp:24-62
                    const int sg = (d2 > 0) - (d2 < 0);
                    d2 = sg * (fabs(d2) + eps);
                    const double q = (sg > 0 ? umax[v] - ui[v] : umin[v] - ui[v]) / d2;
                    const double cand = ibis::kmin(phi[v], q);
                    phi[v] = sg != 0 ? cand : phi[v];
                } else {  // Venkatakrishnan
                    const bool pos = d2 > 0.0, neg = d2 < 0.0;
                    const double d1 = pos ? umax[v] - ui[v] : umin[v] - ui[v];
                    const double d1s = d1 * d1;
                    const double d2s = d2 * d2;
                    const double num = d1s + eps2 + 2.0 * (d1 * d2);
                    const double den = d1s + 2.0 * d2s + d1 * d2 + eps2;
                    const bool take = (pos || neg) && ibis::venkat_take(num, den, bn[v], bd[v]);
                    bn[v] = take ? num : bn[v];
                    bd[v] = take ? den : bd[v];
                }
            }
        }
        if (LIM == 2) {
#pragma unroll
            for (int v = 0; v < 5; v++) phi[v] = bn[v] / bd[v];
        }
        if (write_slots) {
#pragma unroll
            for (int v = 0; v < 5; v++) lim[(size_t)vs.slot[v] * N + i] = phi[v];
        }
    }
    // the reconstruction's offsets: from LDS when the limiter staged them,
    // else all loaded before the face loop (a slot past the cell's last face
    // reads its last face's and does not use them)
    constexpr bool FROM_LDS = LIM != 0;
    double ox[MAXF], oy[MAXF], oz[MAXF];
    if (!(FROM_LDS && stage)) {
#pragma unroll
        for (int j = 0; j < MAXF; j++) face_offset(j < nf ? j : nf - 1, ox[j], oy[j], oz[j]);
    }
#pragma unroll
    for (int j = 0; j < MAXF; j++) {
        if (j >= nf) break;
        const int f = m.ell_face[(size_t)j * N + i];
        double dx = ox[j], dy = oy[j], dz = oz[j];
        if (FROM_LDS && stage) {
            dx = soff[(j * 3 + 0) * TS + threadIdx.x];
            dy = soff[(j * 3 + 1) * TS + threadIdx.x];
            dz = soff[(j * 3 + 2) * TS + threadIdx.x];
        }
        W* rec = LAYOUT == 0 ? (m.ell_sign[(size_t)j * N + i] > 0 ? FL : FR) + (size_t)f * 5
                             : FL + ((size_t)i * m.maxf + j) * 5;
#pragma unroll
        for (int v = 0; v < 5; v++) {
            const double val = ui[v] + phi[v] * (gx[v] * dx + gy[v] * dy + gz[v] * dz);
            if (LAYOUT == 2)
                __builtin_nontemporal_store(W(val), FL + ((size_t)j * 5 + v) * N + i);
            else
                rec[v] = W(val);
        }
    }
}

// Cell kernel of the FP32 reconstruction mode (layout 2, float face values):
// the steps of k_cell_fs -- LSQ gradient, limiter, MUSCL value on each face
// -- in float arithmetic.  Differences of the FP64 inputs (centroid offsets,
// value differences) are formed in FP64 and rounded once; the extremes are
// kept as differences from the cell value, so the limiter's d1 is one
// rounding away from the FP64 one.  stage != 0 keeps the face offsets of
// the limiter pass in LDS as floats for the reconstruction pass.
template <int DIM, int LIM, bool OFFS, bool SN = false>
__global__ void __launch_bounds__(MAXTILE)
    k_cell_fs32(DevMesh m, VarSet vs, const double* __restrict__ prim,
                const double* __restrict__ grad, double eps, double venkat_K, int nlogical,
                float* __restrict__ FL, const double* __restrict__ offs, int stage,
                const float* __restrict__ offs32) {
    // SN: the cell stencil is staged in LDS as doubles, as in k_cell_fs.
    // stage != 0 keeps the limiter pass's face offsets in LDS as floats for
    // the reconstruction pass; with SN they take the stencil's LDS once the
    // whole block has finished the neighbour loop
    extern __shared__ double sh32[];
    float* soff32 = reinterpret_cast<float*>(sh32);
    const int lb = xcd_block(nlogical);
    if (lb >= nlogical) return;
    const int TS = m.tile;  // cells per block: one tile
    const int i = lb * TS + threadIdx.x;
    const int N = m.N, F = m.F;
    const int SS = TS + m.maxhalo;
    // per-cell inputs that do not depend on the stencil, loaded before the
    // staging barrier (a thread past the last cell reads cell 0's)
    const int il = i < N ? i : 0;
    const int nf = m.nfaces[il];
    // the LSQ factors in FP64: formed from the LDS stencil as k_cell_fs
    // does when the mesh's equal that expression, else read
    const bool lsq_lds = SN && m.lsq_from_cen;
    double l11 = 1.0, l12 = 0.0, l22 = 1.0, l13 = 0.0, l23 = 0.0, l33 = 0.0;
    if (!lsq_lds) {
        l11 = m.lsq[il];
        l12 = m.lsq[N + il];
        l22 = m.lsq[2 * (size_t)N + il];
        if (DIM == 3) {
            l13 = m.lsq[3 * (size_t)N + il];
            l23 = m.lsq[4 * (size_t)N + il];
            l33 = m.lsq[5 * (size_t)N + il];
        }
    }
    int kst[MAXF];  // the neighbours' positions in the LDS stencil
    if (SN) {
#pragma unroll
        for (int j = 0; j < MAXF; j++) kst[j] = j < m.maxf ? m.nbr_lds[(size_t)j * N + il] : 0;
        stage_stencil<false, false>(m, vs, nullptr, prim, lb, i, sh32);
    }
    // with the offsets staged over the stencil, a thread past the last cell
    // stays to the second barrier (it works on LDS slots that are not read
    // back and stores nothing)
    const bool reuse = SN && LIM != 0 && stage;
    if (i >= N && !reuse) return;
    constexpr int NG = DIM == 3 ? 5 : 4;
    double ui[5];
    double xi, yi, zi;
    if (SN) {
#pragma unroll
        for (int v = 0; v < 5; v++) ui[v] = sh32[v * SS + threadIdx.x];
        xi = sh32[5 * SS + threadIdx.x];
        yi = sh32[6 * SS + threadIdx.x];
        zi = sh32[7 * SS + threadIdx.x];
    } else {
        load_cell<false>(m, vs, nullptr, prim, i, ui, xi, yi, zi);
    }
    if (SN && lsq_lds) {  // GR/gradient.h:119-162, the grid builder's expressions and order
        double sxx = 0.0, sxy = 0.0, sxz = 0.0, syy = 0.0, syz = 0.0, szz = 0.0;
#pragma unroll
        for (int j = 0; j < MAXF; j++) {
            if (j < nf) {
                const int k = kst[j];
                const double dx = sh32[5 * SS + k] - xi;
                const double dy = sh32[6 * SS + k] - yi;
                const double dz = sh32[7 * SS + k] - zi;
                sxx += dx * dx;
                sxy += dx * dy;
                sxz += dx * dz;
                syy += dy * dy;
                syz += dy * dz;
                szz += dz * dz;
            }
        }
        l11 = sqrt(sxx);
        l12 = 1.0 / l11 * sxy;
        l22 = sqrt(syy - l12 * l12);
        if (DIM == 3) {
            l13 = 1.0 / l11 * sxz;
            l23 = 1.0 / l22 * (syz - l12 / l11 * sxz);
            l33 = sqrt(szz - (l13 * l13 + l23 * l23));
        }
    }
    const float r11 = (float)l11, r12 = (float)l12, r22 = (float)l22;
    const float r13 = (float)l13, r23 = (float)l23, r33 = (float)l33;
    const float beta = (r12 * r23 - r13 * r23) / (r11 * r22);  // GR/gradient.h:89 (sic)
    const float r11s = r11 * r11;
    const float i22 = 1.0f / (r22 * r22);
    const float q12 = r12 / r11;
    const float i33 = 1.0f / (r33 * r33);
    const float q23 = r23 / r22;
    float gx[5], gy[5], gz[5], dmin[5], dmax[5];
#pragma unroll
    for (int v = 0; v < 5; v++) gx[v] = gy[v] = gz[v] = dmin[v] = dmax[v] = 0.0f;
    // one neighbour's contribution to the gradient and the extremes
    auto accumulate = [&](const double* un, double xn, double yn, double zn) {
        const float dx = (float)(xn - xi), dy = (float)(yn - yi), dz = (float)(zn - zi);
        const float alpha_1 = dx / r11s;
        const float alpha_2 = i22 * (dy - q12 * dx);
        float alpha_3 = 0.0f;
        if (DIM == 3) alpha_3 = i33 * (dz - q23 * dy + beta * dx);
        const float w_1 = alpha_1 - q12 * alpha_2 + beta * alpha_3;
        const float w_2 = alpha_2 - q23 * alpha_3;
        const float w_3 = alpha_3;
#pragma unroll
        for (int v = 0; v < 5; v++) {
            const float diff = (float)(un[v] - ui[v]);
            if (v < NG) {
                gx[v] += w_1 * diff;
                gy[v] += w_2 * diff;
                gz[v] += w_3 * diff;
            }
            dmin[v] = diff < dmin[v] ? diff : dmin[v];
            dmax[v] = dmax[v] < diff ? diff : dmax[v];
        }
    };
    if (SN) {
#pragma unroll
        for (int j = 0; j < MAXF; j++) {
            if (j < nf) {
                const int k = kst[j];
                double un[5];
#pragma unroll
                for (int v = 0; v < 5; v++) un[v] = sh32[v * SS + k];
                accumulate(un, sh32[5 * SS + k], sh32[6 * SS + k], sh32[7 * SS + k]);
            }
        }
    } else {
        for (int j = 0; j < nf; j++) {
            const int n = m.ell_nbr[(size_t)j * N + i];
            double un[5], xn, yn, zn;
            load_cell<false>(m, vs, nullptr, prim, n, un, xn, yn, zn);
            accumulate(un, xn, yn, zn);
        }
    }
    if (reuse) {
        __syncthreads();  // the stencil is read: its LDS now takes the offsets
        if (i >= N) return;
    }
#pragma unroll
    for (int v = NG; v < 5; v++) {  // 2D vz: the gradient stored in its slot
        const double* g = grad + (size_t)vs.slot[v] * 3 * N;
        gx[v] = (float)g[i];
        gy[v] = (float)g[N + i];
        gz[v] = (float)g[2 * (size_t)N + i];
    }
    // offs32: the offsets already rounded to float (the same values), when
    // the mode made them.  z: +0 on a flat 2D mesh (whose stencil staging
    // leaves the centroids' z out)
    const bool zflat = DIM == 2 && m.flat_z;
    auto offset = [&](int j, float& dx, float& dy, float& dz) {
        if (OFFS && offs32) {
            dx = offs32[(size_t)(j * 3 + 0) * N + i];
            dy = offs32[(size_t)(j * 3 + 1) * N + i];
            dz = zflat ? 0.0f : offs32[(size_t)(j * 3 + 2) * N + i];
        } else if (OFFS) {
            dx = (float)offs[(size_t)(j * 3 + 0) * N + i];
            dy = (float)offs[(size_t)(j * 3 + 1) * N + i];
            dz = zflat ? 0.0f : (float)offs[(size_t)(j * 3 + 2) * N + i];
        } else {
            const int f = m.ell_face[(size_t)j * N + i];
            dx = (float)(m.geom[9 * (size_t)F + f] - xi);
            dy = (float)(m.geom[10 * (size_t)F + f] - yi);
            dz = zflat ? 0.0f : (float)(m.geom[11 * (size_t)F + f] - zi);
        }
    };
    const bool staged = LIM != 0 && stage;
    float phi[5] = {1.0f, 1.0f, 1.0f, 1.0f, 1.0f};
    if (LIM != 0) {
        float eps2 = 0.0f;
        // Venkatakrishnan: the minimum ratio as a fraction, as in k_cell_fs
        float bn[5] = {1.0f, 1.0f, 1.0f, 1.0f, 1.0f}, bd[5] = {1.0f, 1.0f, 1.0f, 1.0f, 1.0f};
        if (LIM == 2) {
            const double V = fabs(m.vol[i]);  // cell size h^dim = |V|
            const double K3 = venkat_K * venkat_K * venkat_K;
            eps2 = (float)((DIM == 3) ? K3 * V : K3 * (V * sqrt(V)));
        }
        const float epsf = (float)eps;
        // the next face's offset is loaded while this face is limited
        float ndx, ndy, ndz;
        offset(0, ndx, ndy, ndz);
#pragma unroll
        for (int j = 0; j < MAXF; j++) {
            if (j >= nf) break;
            const float dx = ndx, dy = ndy, dz = ndz;
            if (j + 1 < MAXF) offset(j + 1 < nf ? j + 1 : j, ndx, ndy, ndz);
            if (staged) {
                soff32[(j * 3 + 0) * TS + threadIdx.x] = dx;
                soff32[(j * 3 + 1) * TS + threadIdx.x] = dy;
                soff32[(j * 3 + 2) * TS + threadIdx.x] = dz;
            }
#pragma unroll
            for (int v = 0; v < 5; v++) {
                float d2 = gx[v] * dx + gy[v] * dy + gz[v] * dz;
                if (LIM == 1) {  // Barth-Jespersen, FV/limiter.cpp:24-62
                    const int sg = (d2 > 0.0f) - (d2 < 0.0f);
                    d2 = sg * (fabsf(d2) + epsf);
                    const float q = (sg > 0 ? dmax[v] : dmin[v]) / d2;
                    if (sg != 0) phi[v] = q < phi[v] ? q : phi[v];
                } else {  // Venkatakrishnan
                    if (d2 == 0.0f) continue;
                    const float d1 = d2 > 0.0f ? dmax[v] : dmin[v];
                    const float d1s = d1 * d1;
                    const float d2s = d2 * d2;
                    const float num = d1s + eps2 + 2.0f * (d1 * d2);
                    const float den = d1s + 2.0f * d2s + d1 * d2 + eps2;
                    const bool take = num * bd[v] < bn[v] * den;
                    bn[v] = take ? num : bn[v];
                    bd[v] = take ? den : bd[v];
                }
            }
        }
        if (LIM == 2) {
#pragma unroll
            for (int v = 0; v < 5; v++) phi[v] = bn[v] / bd[v];
        }
    }
    // the reconstruction's offsets: from LDS when staged, else all loaded
    // before the face loop (a slot past the cell's last face reads its last
    // face's and does not use them)
    float ox[MAXF], oy[MAXF], oz[MAXF];
    if (!staged) {
#pragma unroll
        for (int j = 0; j < MAXF; j++) offset(j < nf ? j : nf - 1, ox[j], oy[j], oz[j]);
    }
#pragma unroll
    for (int j = 0; j < MAXF; j++) {
        if (j >= nf) break;
        float dx = ox[j], dy = oy[j], dz = oz[j];
        if (staged) {
            dx = soff32[(j * 3 + 0) * TS + threadIdx.x];
            dy = soff32[(j * 3 + 1) * TS + threadIdx.x];
            dz = soff32[(j * 3 + 2) * TS + threadIdx.x];
        }
#pragma unroll
        for (int v = 0; v < 5; v++) {
            const float val = (float)ui[v] + phi[v] * (gx[v] * dx + gy[v] * dy + gz[v] * dz);
            __builtin_nontemporal_store(val, FL + ((size_t)j * 5 + v) * N + i);
        }
    }
}

// Flux * area of face f from the two sides' reconstructed values (A, B,
// vx, vy, vz): thermo closure, rotation to the face frame, flux calculator,
// rotation back (FV/convective_flux.cpp:85-178).  R is the arithmetic of
// these steps: double, or float in the reduced-precision mode; the product
// with the area is FP64.
// t1, t2 and area of face f (geom rows 3-8 and 12)
__device__ __forceinline__ void load_face_geom(const DevMesh& m, int f, double* g) {
    const int F = m.F;
#pragma unroll
    for (int k = 0; k < 6; k++) g[k] = m.geom[(3 + k) * (size_t)F + f];
    g[6] = m.geom[12 * (size_t)F + f];
}

template <int DIM, int FLUX, class R>
__device__ __forceinline__ void face_flux_values_g(const DevMesh& m, const FaceParams& fp, int f,
                                                   const double* g, const R* lv, const R* rv,
                                                   double* res_a, const double* gn = nullptr);

template <int DIM, int FLUX, class R>
__device__ __forceinline__ void face_flux_values(const DevMesh& m, const FaceParams& fp, int f,
                                                 const R* lv, const R* rv, double* res_a) {
    double g[7];
    load_face_geom(m, f, g);
    face_flux_values_g<DIM, FLUX, R>(m, fp, f, g, lv, rv, res_a);
}

// The face flux in the global frame before the product with the area
// (the flux rows of FV/convective_flux.cpp:85-178).  gn: the face normal
// when the caller has loaded it (else read by face id); unused when the
// normal is cross(t1, t2)
template <int DIM, int FLUX, class R>
__device__ __forceinline__ void face_flux_raw_g(const DevMesh& m, const FaceParams& fp, int f,
                                                const double* g, const R* lv, const R* rv,
                                                double* res_d, const double* gn) {
    constexpr int NC = DIM + 2;
    constexpr bool T3 = DIM == 3;
    const int F = m.F;
    R Ls[NPRIM], Rs[NPRIM];
#pragma unroll
    for (int side = 0; side < 2; side++) {
        const R* rec = side ? rv : lv;
        R* st = side ? Rs : Ls;
        const R a = rec[0], b = rec[1];
        R rho = R(0.0), p = R(0.0), T = R(0.0), e = R(0.0);
        switch (fp.thermo) {
            case IBIS_THERMO_RHO_P: p = a; rho = b; break;
            case IBIS_THERMO_RHO_T: rho = a; T = b; break;
            case IBIS_THERMO_RHO_U: rho = a; e = b; break;
            default: p = a; T = b; break;
        }
        ibis::thermo_update<R>(fp.thermo, R(fp.R), R(fp.Cv), rho, p, T, e);
        st[ibis::S_RHO] = rho;
        st[ibis::S_P] = p;
        st[ibis::S_T] = T;
        st[ibis::S_E] = e;
        st[4] = rec[2];
        st[5] = rec[3];
        st[6] = rec[4];
    }
    const R t1x = R(g[0]), t1y = R(g[1]), t1z = R(g[2]);
    const R t2x = R(g[3]), t2y = R(g[4]), t2z = R(g[5]);
    R nx, ny, nz;
    if (m.n_from_t) {  // cross(tan1, tan2), GR/interface.h:249
        nx = t1y * t2z - t1z * t2y;
        ny = t1z * t2x - t1x * t2z;
        nz = t1x * t2y - t1y * t2x;
    } else if (gn) {
        nx = R(gn[0]);
        ny = R(gn[1]);
        nz = R(gn[2]);
    } else {
        nx = R(m.geom[f]);
        ny = R(m.geom[(size_t)F + f]);
        nz = R(m.geom[2 * (size_t)F + f]);
    }
#pragma unroll
    for (int side = 0; side < 2; side++) {
        R* s = side ? Rs : Ls;
        const R a = s[4], b = s[5], c = s[6];
        s[4] = a * nx + b * ny + c * nz;
        s[5] = a * t1x + b * t1y + c * t1z;
        s[6] = a * t2x + b * t2y + c * t2z;
    }
    R out[5];
    if (FLUX == IBIS_FLUX_AUSMDV)
        ibis::flux_ausmdv<R, T3>(Ls, Rs, R(fp.R), R(fp.Cv), R(fp.gamma), out);
    else if (FLUX == IBIS_FLUX_HANEL)
        ibis::flux_hanel<R, T3>(Ls, Rs, R(fp.R), R(fp.Cv), R(fp.gamma), out);
    else if (FLUX == IBIS_FLUX_LDFSS2)
        ibis::flux_ldfss2<R, T3>(Ls, Rs, R(fp.R), R(fp.Cv), R(fp.gamma), R(fp.delta), out);
    else
        ibis::flux_rusanov<R, T3>(Ls, Rs, R(fp.R), R(fp.Cv), R(fp.gamma), out);
    const R px = out[1], py = out[2];
    R pz = R(0.0);
    if (T3) pz = out[3];
    R res[NC];
    res[0] = out[0];
    res[1] = px * nx + py * t1x + pz * t2x;
    res[2] = px * ny + py * t1y + pz * t2y;
    if (T3) {
        res[3] = px * nz + py * t1z + pz * t2z;
        res[4] = out[4];
    } else {
        res[3] = out[4];
    }
#pragma unroll
    for (int q = 0; q < NC; q++) res_d[q] = (double)res[q];
}

// flux * area of face f
template <int DIM, int FLUX, class R>
__device__ __forceinline__ void face_flux_values_g(const DevMesh& m, const FaceParams& fp, int f,
                                                   const double* g, const R* lv, const R* rv,
                                                   double* res_a, const double* gn) {
    constexpr int NC = DIM + 2;
    double res[NC];
    face_flux_raw_g<DIM, FLUX, R>(m, fp, f, g, lv, rv, res, gn);
    const double A = g[6];
#pragma unroll
    for (int q = 0; q < NC; q++) res_a[q] = res[q] * A;
}

// Fused face flux + surface integral from the face states, one tile of
// m.tile cells per block, driven by the tile's face list.
template <int DIM, int FLUX, int LAYOUT, class R = double>
__global__ void __launch_bounds__(MAXTILE)
    k_face_tile_fs(DevMesh m, FaceParams fp, const R* __restrict__ FL,
                   const R* __restrict__ FR, double* __restrict__ dudt) {
    extern __shared__ double sflux[];  // NC x maxtf
    const int t = xcd_block(m.ntiles);
    if (t >= m.ntiles) return;
    constexpr int NC = DIM + 2;
    const int N = m.N, Nt = m.Nt, S = m.maxtf;
    const int f0 = m.tile_face_off[t];
    const int nfl = m.tile_face_off[t + 1] - f0;
    // the inputs of this thread's next face are loaded before the current
    // face is evaluated, so their latency overlaps the flux arithmetic
    auto load = [&](int k, int& f, R (&v)[2][5], double (&g)[7]) {
        f = m.tile_faces[f0 + k];
#pragma unroll
        for (int side = 0; side < 2; side++) {
            if (LAYOUT == 0) {
                const R* rec = (side ? FR : FL) + (size_t)f * 5;
#pragma unroll
                for (int q = 0; q < 5; q++) v[side][q] = rec[q];
            } else if (LAYOUT == 1) {
                const R* rec =
                    FL + (side ? rec_index(m, m.R[f], m.slotR[f]) : rec_index(m, m.L[f], m.slotL[f])) * 5;
#pragma unroll
                for (int q = 0; q < 5; q++) v[side][q] = rec[q];
            } else {
                const int c = side ? m.R[f] : m.L[f];
                const int sl = side ? m.slotR[f] : m.slotL[f];
#pragma unroll
                for (int q = 0; q < 5; q++) v[side][q] = FL[soa_index(m, c, sl, q)];
            }
        }
        load_face_geom(m, f, g);
    };
    int k = threadIdx.x;
    int fn = 0;
    R vn[2][5];
    double gn[7];
    if (k < nfl) load(k, fn, vn, gn);
    while (k < nfl) {
        const int f = fn;
        R v[2][5];
        double g[7];
#pragma unroll
        for (int side = 0; side < 2; side++)
#pragma unroll
            for (int q = 0; q < 5; q++) v[side][q] = vn[side][q];
#pragma unroll
        for (int q = 0; q < 7; q++) g[q] = gn[q];
        const int kn = k + m.tile;
        if (kn < nfl) load(kn, fn, vn, gn);
        double res[NC];
        face_flux_values_g<DIM, FLUX, R>(m, fp, f, g, v[0], v[1], res);
#pragma unroll
        for (int q = 0; q < NC; q++) sflux[q * S + k] = res[q];
        k = kn;
    }
    __syncthreads();
    const int c = t * m.tile + threadIdx.x;
    if (c >= N) return;
    double d[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) d[q] = 0.0;
    const int nf = m.nfaces[c];
    for (int j = 0; j < nf; j++) {
        const int k = m.cell_lds[(size_t)j * N + c];
        const int s = m.ell_sign[(size_t)j * N + c];
#pragma unroll
        for (int q = 0; q < NC; q++) {
            const double fa = sflux[q * S + k];
            d[q] += (s > 0) ? -fa : fa;  // == flux * (-area * outsign), exactly
        }
    }
    const double V = m.vol[c];
#pragma unroll
    for (int q = 0; q < NC; q++) dudt[(size_t)q * Nt + c] = d[q] / V;
}

// Per-entry copies of the tile lists' face data (face_kernel 1): for list
// entry e (face f = tile_faces[e]) the start of each side's face values in
// the layout-2 array as one int2, the lower-numbered cell's first and the
// other's second, one's-complemented when the lower cell is the face's right
// cell (values at start + q * stride, stride N for a valid cell's record and
// G for a ghost's, told apart by start >= maxf*5*N), and the geometry as
// 16-byte pairs over the entries:
// planes (t1x, t1y), (t1z, t2x), (t2y, t2z) [, (nx, ny) when the normal is
// not cross(t1, t2)], then the area [and nz] as plain planes.  A face's
// inputs are then 1 + 4 (+1 +1) contiguous loads at e plus the gathers of
// the two sides' values.  On a planar 2D mesh (with_n 2) only (t1x, t1y)
// and the area are stored: the other tangent components are 0, 0, 0, 1.
__global__ void __launch_bounds__(BLOCK)
    k_entries(DevMesh m, int with_n, int2* __restrict__ eidx, double* __restrict__ egeom) {
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    const int NE = m.ntf, F = m.F;
    if (e >= NE) return;
    const int f = m.ent_faces[e];
    const int bl = (int)soa_index(m, m.L[f], m.slotL[f], 0);
    const int br = (int)soa_index(m, m.R[f], m.slotR[f], 0);
    eidx[e] = m.L[f] < m.R[f] ? make_int2(bl, br) : make_int2(br, ~bl);
    const double* gf = m.geom;
    double2* g2 = reinterpret_cast<double2*>(egeom);
    g2[e] = make_double2(gf[3 * (size_t)F + f], gf[4 * (size_t)F + f]);
    if (with_n == 2) {
        egeom[(size_t)2 * NE + e] = gf[12 * (size_t)F + f];
        return;
    }
    g2[(size_t)NE + e] = make_double2(gf[5 * (size_t)F + f], gf[6 * (size_t)F + f]);
    g2[2 * (size_t)NE + e] = make_double2(gf[7 * (size_t)F + f], gf[8 * (size_t)F + f]);
    const int np = with_n ? 4 : 3;  // pair planes
    if (with_n) g2[3 * (size_t)NE + e] = make_double2(gf[f], gf[(size_t)F + f]);
    egeom[(size_t)2 * np * NE + e] = gf[12 * (size_t)F + f];
    if (with_n) egeom[(size_t)(2 * np + 1) * NE + e] = gf[2 * (size_t)F + f];
}

// a[k] rounded to float (the FP32 reconstruction mode's face offsets)
__global__ void __launch_bounds__(BLOCK) k_to_float(size_t n, const double* __restrict__ a,
                                                   float* __restrict__ b) {
    const size_t k = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (k < n) b[k] = (float)a[k];
}

// Float copies of the entry tangents for the reduced-precision modes, whose
// face kernel rounds t1 and t2 to float before use: planes of float4
// (t1x, t1y, t1z, t2x) and float2 (t2y, t2z) over the entries, then the area
// as doubles (32 bytes per entry instead of 56).
__global__ void __launch_bounds__(BLOCK) k_entries32(DevMesh m, char* __restrict__ g) {
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    const int NE = m.ntf, F = m.F;
    if (e >= NE) return;
    const int f = m.ent_faces[e];
    const double* gf = m.geom;
    float4* a = reinterpret_cast<float4*>(g);
    float2* b = reinterpret_cast<float2*>(g + (size_t)16 * NE);
    double* c = reinterpret_cast<double*>(g + (size_t)24 * NE);
    a[e] = make_float4((float)gf[3 * (size_t)F + f], (float)gf[4 * (size_t)F + f],
                       (float)gf[5 * (size_t)F + f], (float)gf[6 * (size_t)F + f]);
    b[e] = make_float2((float)gf[7 * (size_t)F + f], (float)gf[8 * (size_t)F + f]);
    c[e] = gf[12 * (size_t)F + f];
}

struct ViscParams {
    double R, Cv, Cp, mu0, T0, Ts, Pr;
};

// GAS/transport_properties.h:27-42; the constants have the scalar type
template <class S>
__device__ __forceinline__ S sutherland(S temp, const ViscParams& vp) {
    const S mu0 = S(vp.mu0), T0 = S(vp.T0), Ts = S(vp.Ts);
    return mu0 * pow(temp / T0, S(3.0 / 2.0)) * (T0 + Ts) / (temp + Ts);
}

// GAS/transport_properties.h:64-75
template <class S>
__device__ __forceinline__ S conductivity(S temp, const ViscParams& vp) {
    const S mu = sutherland<S>(temp, vp);
    return vp.Cp * mu / S(vp.Pr);
}

template <class S, int DIM>
__device__ __forceinline__ void hasselbacher(const S* __restrict__ grad, int N, int slot, S vl,
                                             S vr, int l, int r, const double* ehat,
                                             const double* n, double len_e, double edn, S* out) {
    const S* g = grad + (size_t)slot * 3 * N;
    const S ax = 0.5 * (g[l] + g[r]);
    const S ay = 0.5 * (g[N + l] + g[N + r]);
    const S az = 0.5 * (grad_z<S, DIM>(g, N, l) + grad_z<S, DIM>(g, N, r));
    const S avg_dot_ehat = ax * ehat[0] + ay * ehat[1] + az * ehat[2];
    const S correction = avg_dot_ehat - (vr - vl) / len_e;
    out[0] = ax - correction * n[0] / edn;
    out[1] = ay - correction * n[1] / edn;
    out[2] = az - correction * n[2] / edn;
}

// The viscous flux of face f (viscous_flux.cpp:133-184) between cells l
// and r, whose gradients are in the slots G_T, G_VX, G_VY, G_VZ of grad:
// Hasselbacher face gradients between two valid cells, the valid cell's own
// gradients at a boundary face (viscous_flux.cpp:16-80); the face state is
// the average of the two cells' p, T, v (GAS/flow_state.h:89-91).  vt: the
// terms the flux rows lose, momentum then energy (dim+1 values); rho_f and
// T_f: the face density and temperature that estimate_dt reads.
template <class S, int DIM>
__device__ __forceinline__ void visc_face(const DevMesh& m, const ViscParams& vp,
                                          const S* __restrict__ prim, const S* __restrict__ grad,
                                          int f, int l, int r, S* vt, double* rho_f,
                                          double* T_f) {
    const int F = m.F, N = m.N, Nt = m.Nt;
    const bool lv = l < N, rv = r < N;
    S gT[3], gvx[3], gvy[3], gvz[3];
    const double n[3] = {m.geom[f], m.geom[(size_t)F + f], m.geom[2 * (size_t)F + f]};
    if (!lv || !rv) {  // copy_gradients_to_face, viscous_flux.cpp:16-34
        const int c = lv ? l : r;
        for (int k = 0; k < 2; k++) {
            gT[k] = grad[((size_t)G_T * 3 + k) * N + c];
            gvx[k] = grad[((size_t)G_VX * 3 + k) * N + c];
            gvy[k] = grad[((size_t)G_VY * 3 + k) * N + c];
            gvz[k] = grad[((size_t)G_VZ * 3 + k) * N + c];
        }
        gT[2] = grad_z<S, DIM>(grad + (size_t)G_T * 3 * N, N, c);
        gvx[2] = grad_z<S, DIM>(grad + (size_t)G_VX * 3 * N, N, c);
        gvy[2] = grad_z<S, DIM>(grad + (size_t)G_VY * 3 * N, N, c);
        gvz[2] = grad_z<S, DIM>(grad + (size_t)G_VZ * 3 * N, N, c);
    } else {  // hasselbacher_average, viscous_flux.cpp:52-80
        const double ex = m.cen[r] - m.cen[l];
        const double ey = m.cen[Nt + r] - m.cen[Nt + l];
        const double ez =
            DIM == 2 && m.flat_z ? 0.0 : m.cen[2 * (size_t)Nt + r] - m.cen[2 * (size_t)Nt + l];
        const double len_e = sqrt(ex * ex + ey * ey + ez * ez);
        const double ehat[3] = {ex / len_e, ey / len_e, ez / len_e};
        const double edn = ehat[0] * n[0] + ehat[1] * n[1] + ehat[2] * n[2];
        hasselbacher<S, DIM>(grad, N, G_VX, prim[(size_t)P_VX * Nt + l], prim[(size_t)P_VX * Nt + r], l,
                        r, ehat, n, len_e, edn, gvx);
        hasselbacher<S, DIM>(grad, N, G_VY, prim[(size_t)P_VY * Nt + l], prim[(size_t)P_VY * Nt + r], l,
                        r, ehat, n, len_e, edn, gvy);
        hasselbacher<S, DIM>(grad, N, G_VZ, prim[(size_t)P_VZ * Nt + l], prim[(size_t)P_VZ * Nt + r], l,
                        r, ehat, n, len_e, edn, gvz);
        hasselbacher<S, DIM>(grad, N, G_T, prim[(size_t)P_T * Nt + l], prim[(size_t)P_T * Nt + r], l, r,
                        ehat, n, len_e, edn, gT);
    }
    // average_flow_states_pT + update_thermo_from_pT (GAS/flow_state.h:89-91)
    const S p = 0.5 * (prim[(size_t)P_P * Nt + l] + prim[(size_t)P_P * Nt + r]);
    const S temp = 0.5 * (prim[(size_t)P_T * Nt + l] + prim[(size_t)P_T * Nt + r]);
    const S vx = 0.5 * (prim[(size_t)P_VX * Nt + l] + prim[(size_t)P_VX * Nt + r]);
    const S vy = 0.5 * (prim[(size_t)P_VY * Nt + l] + prim[(size_t)P_VY * Nt + r]);
    const S vz = 0.5 * (prim[(size_t)P_VZ * Nt + l] + prim[(size_t)P_VZ * Nt + r]);
    *rho_f = ibis::real_part(p / (vp.R * temp));
    *T_f = ibis::real_part(temp);
    const S mu = sutherland<S>(temp, vp);
    const S k = conductivity<S>(temp, vp);
    const S lambda = -2.0 / 3.0 * mu;
    const S bulk = lambda * (gvx[0] + gvy[1] + gvz[2]);
    const S tau_xx = 2.0 * mu * gvx[0] + bulk;
    const S tau_yy = 2.0 * mu * gvy[1] + bulk;
    const S tau_zz = 2.0 * mu * gvz[2] + bulk;
    const S tau_xy = mu * (gvx[1] + gvy[0]);
    const S tau_xz = mu * (gvx[2] + gvz[0]);
    const S tau_yz = mu * (gvy[2] + gvz[1]);
    const S theta_x = vx * tau_xx + vy * tau_xy + vz * tau_xz + k * gT[0];
    const S theta_y = vx * tau_xy + vy * tau_yy + vz * tau_yz + k * gT[1];
    const S theta_z = vx * tau_xz + vy * tau_yz + vz * tau_zz + k * gT[2];
    vt[0] = n[0] * tau_xx + n[1] * tau_xy + n[2] * tau_xz;
    vt[1] = n[0] * tau_xy + n[1] * tau_yy + n[2] * tau_yz;
    if (DIM == 3) {
        vt[2] = n[0] * tau_xz + n[1] * tau_yz + n[2] * tau_zz;
        vt[3] = n[0] * theta_x + n[1] * theta_y + n[2] * theta_z;
    } else {
        vt[2] = n[0] * theta_x + n[1] * theta_y + n[2] * theta_z;
    }
}

// estimate_dt's per-cell term (FV/finite_volume.cpp:143-176): the local
// time step of one cell from its velocity and temperature, with
// face(j, nx, ny, nz, area) giving the normal and area of its face slot j
// (slot order) and face_id(j) the face index for the viscous term.
template <class Face, class FaceId>
__device__ __forceinline__ double cell_dt_of(double vx, double vy, double vz, double temp,
                                             double volume, int nf, const double* __restrict__ face_rho,
                                             const double* __restrict__ face_T, int viscous,
                                             double sf, const ViscParams& vp, double gamma,
                                             Face face, FaceId face_id) {
    const double a = sqrt(gamma * vp.R * temp);
    double src = 0.0, srv = 0.0;
    for (int j = 0; j < nf; j++) {
        double nx, ny, nz, area;
        face(j, nx, ny, nz, area);
        const double dot = vx * nx + vy * ny + vz * nz;
        const double sig_vel = fabs(dot) + a;
        src += sig_vel * area;
        if (viscous) {
            const int f = face_id(j);
            const double tf = face_T[f];
            const double mu = sutherland(tf, vp);
            const double k = vp.Cp * sutherland(tf, vp) / vp.Pr;
            const double rho = face_rho[f];
            const double Pr = mu * vp.Cp / k;
            const double tmp = (gamma / rho) * (mu / Pr) * area * area;
            srv += tmp / volume;
        }
    }
    return volume / (src + sf * srv);
}

// k_face_tile_fs over the per-entry face data (layout 2 values); NRM: the
// normals are stored (they are not cross(t1, t2)).  EPI: the Runge-Kutta
// stage epilogue (ibis_stage_epilogue) follows each cell's dU/dt; dudt may
// then be null.
// The viscous inputs of the fused viscous face kernel: the state after the
// viscous boundary actions, the cells' viscous gradients (slots G_T, G_VX,
// G_VY, G_VZ) and the face density and temperature it writes for estimate_dt.
struct ViscFace {
    const double* vterm;  // (dim+1) x F: k_visc_terms of each face
};

// The viscous terms of every face (visc_face), in face order, and the face
// density and temperature for estimate_dt.
template <int DIM>
__global__ void __launch_bounds__(BLOCK)
    k_visc_terms(DevMesh m, ViscParams vp, const double* __restrict__ prim,
                 const double* __restrict__ grad, double* __restrict__ vterm,
                 double* __restrict__ face_rho, double* __restrict__ face_T) {
    const int f = blockIdx.x * BLOCK + threadIdx.x;
    const int F = m.F;
    if (f >= F) return;
    double vt[DIM + 1], rho_f, T_f;
    visc_face<double, DIM>(m, vp, prim, grad, f, m.L[f], m.R[f], vt, &rho_f, &T_f);
    face_rho[f] = rho_f;
    face_T[f] = T_f;
#pragma unroll
    for (int q = 0; q <= DIM; q++) vterm[(size_t)q * F + f] = vt[q];
}

// VISC: each entry's viscous terms (k_visc_terms of its face) are
// subtracted from its convective flux before the product with the area, as
// the reference's viscous pass subtracts them from the stored face flux.
template <int DIM, int FLUX, bool NRM, class R = double, bool P2 = false, bool G32 = false,
          bool EPI = false, bool VISC = false>
__global__ void __launch_bounds__(MAXTILE)
    k_face_entry_fs(DevMesh m, FaceParams fp, const int2* __restrict__ eidx,
                    const double* __restrict__ egeom, const R* __restrict__ FL,
                    double* __restrict__ dudt, ibis_stage_epilogue epi, ViscFace vf) {
    extern __shared__ double sflux[];  // NC x maxtf [+ 4 x maxtf: normals, areas]
    const int t = xcd_block(m.ntiles);
    if (t >= m.ntiles) return;
    constexpr int NC = DIM + 2;
    const int N = m.N, Nt = m.Nt, S = m.maxtf, NE = m.ntf;
    const int f0 = m.tile_face_off[t];
    const int nfl = m.tile_face_off[t + 1] - f0;
    const int gbase = m.maxf * 5 * N;
    constexpr int NG = NRM ? 10 : 7;
    const double2* eg2 = reinterpret_cast<const double2*>(egeom);
    constexpr int NP = NRM ? 4 : 3;  // pair planes of the entry geometry
    auto load = [&](int k, R (&v)[2][5], double (&g)[NG]) {
        const int e = f0 + k;
        const int2 bb = eidx[e];
        const bool swap = bb.y < 0;  // the first start is the right cell's
        R u[2][5];
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const int b = side ? (swap ? ~bb.y : bb.y) : bb.x;
            const size_t stride = b >= gbase ? (size_t)m.G : (size_t)N;
#pragma unroll
            for (int q = 0; q < 5; q++) u[side][q] = FL[(size_t)b + q * stride];
        }
#pragma unroll
        for (int q = 0; q < 5; q++) {
            v[0][q] = swap ? u[1][q] : u[0][q];
            v[1][q] = swap ? u[0][q] : u[1][q];
        }
        // g: t1 (0-2), t2 (3-5), area (6) [, n (7-9)]
        if (G32) {  // the float copies of k_entries32 (egeom is their base)
            const char* gb = reinterpret_cast<const char*>(egeom);
            const float4 a = reinterpret_cast<const float4*>(gb)[e];
            const float2 b = reinterpret_cast<const float2*>(gb + (size_t)16 * NE)[e];
            g[0] = a.x;
            g[1] = a.y;
            g[2] = a.z;
            g[3] = a.w;
            g[4] = b.x;
            g[5] = b.y;
            g[6] = reinterpret_cast<const double*>(gb + (size_t)24 * NE)[e];
            return;
        }
        if (P2) {  // planar 2D: t1z = 0, t2 = (0, 0, 1), the same bits as stored
            const double2 p0 = eg2[e];
            g[0] = p0.x;
            g[1] = p0.y;
            g[2] = 0.0;
            g[3] = 0.0;
            g[4] = 0.0;
            g[5] = 1.0;
            g[6] = egeom[(size_t)2 * NE + e];
            return;
        }
        const double2 p0 = eg2[e], p1 = eg2[(size_t)NE + e], p2 = eg2[2 * (size_t)NE + e];
        g[0] = p0.x;
        g[1] = p0.y;
        g[2] = p1.x;
        g[3] = p1.y;
        g[4] = p2.x;
        g[5] = p2.y;
        g[6] = egeom[(size_t)2 * NP * NE + e];
        if (NRM) {
            const double2 p3 = eg2[3 * (size_t)NE + e];
            g[7] = p3.x;
            g[8] = p3.y;
            g[NG - 1] = egeom[(size_t)(2 * NP + 1) * NE + e];
        }
    };
    int k = threadIdx.x;
    R vn[2][5];
    double gn[NG];
    if (k < nfl) load(k, vn, gn);
    while (k < nfl) {
        R v[2][5];
        double g[NG];
#pragma unroll
        for (int side = 0; side < 2; side++)
#pragma unroll
            for (int q = 0; q < 5; q++) v[side][q] = vn[side][q];
#pragma unroll
        for (int q = 0; q < NG; q++) g[q] = gn[q];
        const int kn = k + m.tile;
        if (kn < nfl) load(kn, vn, gn);
        double res[NC];
        if (VISC) {
            face_flux_raw_g<DIM, FLUX, R>(m, fp, 0, g, v[0], v[1], res, NRM ? g + 7 : nullptr);
            const size_t f = m.ent_faces[f0 + k];
            const double A = g[6];
#pragma unroll
            for (int q = 0; q < NC; q++)
                res[q] = (q == 0 ? res[q] : res[q] - vf.vterm[(size_t)(q - 1) * m.F + f]) * A;
        } else {
            face_flux_values_g<DIM, FLUX, R>(m, fp, 0, g, v[0], v[1], res, NRM ? g + 7 : nullptr);
        }
#pragma unroll
        for (int q = 0; q < NC; q++) sflux[q * S + k] = res[q];
        if (EPI && epi.dt_cells_out) {
            // the normal as the flux takes it (cross(t1, t2), GR/interface.h:249,
            // equal to the stored one for these meshes) and the area, for the
            // cells' local time steps below
            double* sgeo = sflux + NC * S;
            sgeo[k] = NRM ? g[NG - 3] : g[1] * g[5] - g[2] * g[4];
            sgeo[S + k] = NRM ? g[NG - 2] : g[2] * g[3] - g[0] * g[5];
            sgeo[2 * S + k] = NRM ? g[NG - 1] : g[0] * g[4] - g[1] * g[3];
            sgeo[3 * S + k] = g[6];
        }
        k = kn;
    }
    __syncthreads();
    if (EPI && epi.ghost_per_tile > 0) {
        // this tile's share of the ghost rows: the stage value and its
        // primitives, as the separate stage pass evaluates them
        const int g1 = min(Nt, N + (t + 1) * epi.ghost_per_tile);
        for (int g = N + t * epi.ghost_per_tile + threadIdx.x; g < g1; g += blockDim.x) {
            const double dtg = epi.dt[epi.inst_of_cell[g]];
            double qv[NC], out[NPRIM];
#pragma unroll
            for (int q = 0; q < NC; q++) {
                const size_t e = (size_t)q * Nt + g;
                double v = epi.u[e];
                for (int j = 0; j < epi.nk; j++) v = v + epi.k[j][e] * (epi.c[j] * dtg);
                if (epi.use_reg) v = v + 0.0 * (epi.c_reg * dtg);
                qv[q] = v;
                if (epi.cons_out) epi.cons_out[e] = v;
            }
            ibis::cons_to_prim_cell<double, DIM>(qv, epi.R, epi.Cv, out);
#pragma unroll
            for (int q = 0; q < NPRIM; q++) epi.prim_out[(size_t)q * Nt + g] = out[q];
        }
    }
    const int c = t * m.tile + threadIdx.x;
    if (c >= N) return;
    double d[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) d[q] = 0.0;
    const int nf = m.nfaces[c];
    for (int j = 0; j < nf; j++) {
        const int k = m.ent_lds[(size_t)j * N + c];
        const int s = m.ell_sign[(size_t)j * N + c];
#pragma unroll
        for (int q = 0; q < NC; q++) {
            const double fa = sflux[q * S + k];
            d[q] += (s > 0) ? -fa : fa;  // == flux * (-area * outsign), exactly
        }
    }
    const double V = m.vol[c];
    if (!EPI) {
#pragma unroll
        for (int q = 0; q < NC; q++) dudt[(size_t)q * Nt + c] = d[q] / V;
        return;
    }
    // stage epilogue: the same expressions, in the same order, as a
    // separate stage-update pass over the stored dU/dt
    double kr[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) {
        kr[q] = d[q] / V;
        if (dudt) dudt[(size_t)q * Nt + c] = kr[q];
    }
    const double dtc = epi.dt[epi.inst_of_cell[c]];
    double qv[NC], out[NPRIM];
#pragma unroll
    for (int q = 0; q < NC; q++) {
        const size_t e = (size_t)q * Nt + c;
        double v = epi.u[e];
        for (int j = 0; j < epi.nk; j++) v = v + epi.k[j][e] * (epi.c[j] * dtc);
        if (epi.use_reg) v = v + kr[q] * (epi.c_reg * dtc);
        qv[q] = v;
        if (epi.cons_out) epi.cons_out[e] = v;
    }
    ibis::cons_to_prim_cell<double, DIM>(qv, epi.R, epi.Cv, out);
#pragma unroll
    for (int q = 0; q < NPRIM; q++) epi.prim_out[(size_t)q * Nt + c] = out[q];
    if (epi.dt_cells_out) {
        // the local time step of the new state (inviscid): cell_dt's terms
        // with the faces' normals and areas from LDS, in the cell's slot order
        const double* sgeo = sflux + NC * S;
        auto face = [&](int j, double& nx, double& ny, double& nz, double& area) {
            const int k = m.ent_lds[(size_t)j * N + c];
            nx = sgeo[k];
            ny = sgeo[S + k];
            nz = sgeo[2 * S + k];
            area = sgeo[3 * S + k];
        };
        auto face_id = [](int) { return 0; };
        ViscParams vp{};
        vp.R = epi.dt_R;
        epi.dt_cells_out[c] = cell_dt_of(out[P_VX], out[P_VY], out[P_VZ], out[P_T], V, nf, nullptr,
                                         nullptr, 0, epi.dt_sf, vp, epi.dt_gamma, face, face_id);
    }
}

// Fused face flux + surface integral, slot-major variant for face values in
// (slot, variable) planes (layout 2).  Block t owns cells [t*TS, t*TS + TS).
// The thread of cell c evaluates, for each slot j, the face f = ell_face[j][c]
// when c is its left cell, or when c is its right cell and the left cell
// lies outside the block: its own side's values are one coalesced load per
// plane across the wavefront, only the other side's are gathered.  The
// flux * area goes to LDS at (j, c); each cell then sums its faces from LDS
// in its own face order (a face both of whose cells are in the block is
// read from its left cell's slot).  No face lists are needed, so TS is
// free: 64, 128 or 256.
template <int DIM, int FLUX, int TS, class R = double>
__global__ void __launch_bounds__(TS)
    k_face_slot_fs(DevMesh m, FaceParams fp, int nlogical, const R* __restrict__ FL,
                   double* __restrict__ dudt) {
    extern __shared__ double sflux[];  // NC x maxf x TS
    const int t = xcd_block(nlogical);
    if (t >= nlogical) return;
    constexpr int NC = DIM + 2;
    const int N = m.N, Nt = m.Nt, MF = m.maxf;
    const int c0 = t * TS;
    const int cend = min(c0 + TS, N);  // valid cells of the block (ghost ids are >= N)
    const int c = c0 + threadIdx.x;
    const int nf = c < N ? m.nfaces[c] : 0;
    for (int j = 0; j < nf; j++) {
        const int f = m.ell_face[(size_t)j * N + c];
        const int l = m.L[f];
        const bool left = l == c;
        if (!left && l >= c0 && l < cend) continue;  // evaluated by its left cell
        const int o = left ? m.R[f] : l;
        const int so = left ? m.slotR[f] : m.slotL[f];
        R own[5], oth[5];
#pragma unroll
        for (int q = 0; q < 5; q++) {
            own[q] = FL[((size_t)j * 5 + q) * N + c];
            oth[q] = FL[soa_index(m, o, so, q)];
        }
        double res[NC];
        if (left)
            face_flux_values<DIM, FLUX, R>(m, fp, f, own, oth, res);
        else
            face_flux_values<DIM, FLUX, R>(m, fp, f, oth, own, res);
#pragma unroll
        for (int q = 0; q < NC; q++) sflux[((size_t)q * MF + j) * TS + threadIdx.x] = res[q];
    }
    __syncthreads();
    if (c >= N) return;
    double d[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) d[q] = 0.0;
    for (int j = 0; j < nf; j++) {
        const int f = m.ell_face[(size_t)j * N + c];
        const int l = m.L[f];
        const int s = m.ell_sign[(size_t)j * N + c];
        const bool mine = l == c || l < c0 || l >= cend;
        const int jj = mine ? j : m.slotL[f];
        const int tt = mine ? threadIdx.x : l - c0;
#pragma unroll
        for (int q = 0; q < NC; q++) {
            const double fa = sflux[((size_t)q * MF + jj) * TS + tt];
            d[q] += (s > 0) ? -fa : fa;  // == flux * (-area * outsign), exactly
        }
    }
    const double V = m.vol[c];
#pragma unroll
    for (int q = 0; q < NC; q++) dudt[(size_t)q * Nt + c] = d[q] / V;
}

// ConstantFlux::apply (boundary.cpp:405-443), static grid.  The override
// depends on the action parameters only, so its tangent is zero.
template <class S>
__global__ void __launch_bounds__(BLOCK)
    k_post_flux(DevMesh m, const int* __restrict__ kinds, const double* __restrict__ pars,
                GasParams gp, int premul, S* __restrict__ flux) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.nb) return;
    const int f = m.b_face[b], t = m.b_tag[b], F = m.F;
    const bool three_d = m.dim == 3;
    for (int a = 0; a < IBIS_MAX_ACT; a++) {
        if (kinds[t * IBIS_MAX_ACT + a] != IBIS_BC_CONSTANT_FLUX) continue;
        const double* par = pars + (t * IBIS_MAX_ACT + a) * IBIS_NPAR;
        const double nx = m.geom[f], ny = m.geom[(size_t)F + f], nz = m.geom[2 * (size_t)F + f];
        const double u = gp.Cv * par[0];
        const double fvx = par[4], fvy = par[5], fvz = par[6];
        const double rvx = fvx - 0.0, rvy = fvy - 0.0, rvz = fvz - 0.0;
        const double mass_flux = par[2] * (rvx * nx + rvy * ny + rvz * nz);
        const double p = par[1];
        double v[5];
        v[0] = mass_flux;
        v[1] = p * nx + fvx * mass_flux;
        v[2] = p * ny + fvy * mass_flux;
        v[3] = p * nz + fvz * mass_flux;
        v[4] = mass_flux * (u + 0.5 * (fvx * fvx + fvy * fvy + fvz * fvz)) +
               p * (fvx * nx + fvy * nx + fvz * nz);  // v_y * n_x as in the reference
        const double A = premul ? m.geom[12 * (size_t)F + f] : 1.0;
        const int ne = three_d ? 5 : 4;
        for (int q = 0; q < ne; q++) {
            const double val = (!three_d && q == 3) ? v[4] : v[q];
            flux[(size_t)q * F + f] = S(premul ? val * A : val);
        }
    }
}

// ---------------------------------------------------------------------------
// Viscous flux (FV/viscous_flux.cpp:36-184)
// ---------------------------------------------------------------------------
template <class S, int DIM>
__global__ void __launch_bounds__(BLOCK)
    k_visc_flux(DevMesh m, ViscParams vp, const S* __restrict__ prim, const S* __restrict__ grad,
                S* __restrict__ flux, double* __restrict__ face_rho, double* __restrict__ face_T) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    const int F = m.F;
    if (f >= F) return;
    S vt[DIM + 1];
    double rho_f, T_f;
    visc_face<S, DIM>(m, vp, prim, grad, f, m.L[f], m.R[f], vt, &rho_f, &T_f);
    if (face_rho) {
        face_rho[f] = rho_f;
        face_T[f] = T_f;
    }
#pragma unroll
    for (int q = 0; q <= DIM; q++) flux[(size_t)(q + 1) * F + f] -= vt[q];
}

// ---------------------------------------------------------------------------
// Surface integral (FV/finite_volume.cpp:221-255): per cell, in the cell's
// face order.  With PREMUL the flux already holds flux*area and
// flux*(-area*outsign) == (outsign > 0 ? -(flux*area) : flux*area) exactly.
// ---------------------------------------------------------------------------
template <class S, int DIM, bool PREMUL>
__global__ void __launch_bounds__(BLOCK)
    k_integral(DevMesh m, const S* __restrict__ flux, S* __restrict__ dudt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = m.N, Nt = m.Nt, F = m.F;
    if (i >= N) return;
    constexpr int NC = DIM + 2;
    S d[NC];
#pragma unroll
    for (int q = 0; q < NC; q++) d[q] = S(0.0);
    const int nf = m.nfaces[i];
    for (int j = 0; j < nf; j++) {
        const int f = m.ell_face[(size_t)j * N + i];
        const int s = m.ell_sign[(size_t)j * N + i];
        if (PREMUL) {
#pragma unroll
            for (int q = 0; q < NC; q++) {
                const S fa = flux[(size_t)q * F + f];
                d[q] += (s > 0) ? -fa : fa;
            }
        } else {
            const double area = -m.geom[12 * (size_t)F + f] * s;
#pragma unroll
            for (int q = 0; q < NC; q++) d[q] += flux[(size_t)q * F + f] * area;
        }
    }
    const double V = m.vol[i];
#pragma unroll
    for (int q = 0; q < NC; q++) dudt[(size_t)q * Nt + i] = d[q] / V;
}

// ---------------------------------------------------------------------------
// estimate_dt (FV/finite_volume.cpp:131-184): block minimum, then a single
// block reduces the partial minima.  Minimum is order independent.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_min(double v) {
    __shared__ double sm[BLOCK / 64];
    for (int o = 32; o > 0; o >>= 1) v = ibis::kmin(__shfl_down(v, o, 64), v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sm[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < BLOCK / 64; k++) v = ibis::kmin(sm[k], v);
    }
    return v;
}

// local time step of valid cell i (FV/finite_volume.cpp:143-176)
__device__ __forceinline__ double cell_dt(const DevMesh& m, const double* __restrict__ prim,
                                          const double* __restrict__ face_rho,
                                          const double* __restrict__ face_T, int viscous,
                                          double sf, const ViscParams& vp, double gamma, int i) {
    const int N = m.N, Nt = m.Nt;
    const size_t F = m.F;
    auto face_id = [&](int j) { return m.ell_face[(size_t)j * N + i]; };
    auto face = [&](int j, double& nx, double& ny, double& nz, double& area) {
        const int f = face_id(j);
        area = m.geom[12 * F + f];
        nx = m.geom[f];
        ny = m.geom[F + f];
        nz = m.geom[2 * F + f];
    };
    return cell_dt_of(prim[(size_t)P_VX * Nt + i], prim[(size_t)P_VY * Nt + i],
                      prim[(size_t)P_VZ * Nt + i], prim[(size_t)P_T * Nt + i], m.vol[i],
                      m.nfaces[i], face_rho, face_T, viscous, sf, vp, gamma, face, face_id);
}

__global__ void __launch_bounds__(BLOCK)
    k_dt_partial(DevMesh m, const double* __restrict__ prim, const double* __restrict__ face_rho,
                 const double* __restrict__ face_T, int viscous, double sf, ViscParams vp,
                 double gamma, double* __restrict__ partial) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double dt = 1.79769313486231570815e+308;
    if (i < m.N) dt = cell_dt(m, prim, face_rho, face_T, viscous, sf, vp, gamma, i);
    dt = block_min(dt);
    if (threadIdx.x == 0) partial[blockIdx.x] = dt;
}

__global__ void __launch_bounds__(BLOCK)
    k_dt_cells(DevMesh m, const double* __restrict__ prim, const double* __restrict__ face_rho,
               const double* __restrict__ face_T, int viscous, double sf, ViscParams vp,
               double gamma, double* __restrict__ dt_cells) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m.N) dt_cells[i] = cell_dt(m, prim, face_rho, face_T, viscous, sf, vp, gamma, i);
}

__global__ void __launch_bounds__(BLOCK) k_min_final(const double* __restrict__ partial, int n,
                                                     double* __restrict__ out) {
    double v = 1.79769313486231570815e+308;
    for (int k = threadIdx.x; k < n; k += BLOCK) v = ibis::kmin(partial[k], v);
    v = block_min(v);
    if (threadIdx.x == 0) out[0] = v;
}

// count_bad_cells (FV/finite_volume.cpp:258-270): integer sum, order independent
__global__ void __launch_bounds__(BLOCK) k_zero_u64(unsigned long long* p) { p[0] = 0ull; }

__global__ void __launch_bounds__(BLOCK)
    k_count_bad(int N, int Nt, const double* __restrict__ prim, unsigned long long* count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int bad = 0;
    if (i < N) {
        const double temp = prim[(size_t)P_T * Nt + i], rho = prim[(size_t)P_RHO * Nt + i];
        bad = (temp < 0.0 || rho < 0.0 || isnan(rho) || isinf(rho)) ? 1 : 0;
    }
    unsigned long long ballot = __ballot(bad);
    if ((threadIdx.x & 63) == 0 && ballot) atomicAdd(count, (unsigned long long)__popcll(ballot));
}

// conversions, FV/primative_conserved_conversion.cpp:5-51
template <class S, int DIM>
__global__ void __launch_bounds__(BLOCK)
    k_cons_to_prim(int Nt, double R, double Cv, const S* __restrict__ cq, S* __restrict__ fs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nt) return;
    S q[DIM + 2], out[NPRIM];
#pragma unroll
    for (int k = 0; k < DIM + 2; k++) q[k] = cq[(size_t)k * Nt + i];
    ibis::cons_to_prim_cell<S, DIM>(q, R, Cv, out);
#pragma unroll
    for (int k = 0; k < NPRIM; k++) fs[(size_t)k * Nt + i] = out[k];
}

template <int DIM>
__global__ void __launch_bounds__(BLOCK)
    k_prim_to_cons(int Nt, const double* __restrict__ fs, double* __restrict__ cq) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nt) return;
    const double vx = fs[(size_t)P_VX * Nt + i], vy = fs[(size_t)P_VY * Nt + i],
                 vz = fs[(size_t)P_VZ * Nt + i];
    const double rho = fs[(size_t)P_RHO * Nt + i];
    cq[i] = rho;
    cq[(size_t)Nt + i] = rho * vx;
    cq[2 * (size_t)Nt + i] = rho * vy;
    if (DIM == 3) cq[3 * (size_t)Nt + i] = rho * vz;
    const double ke = 0.5 * (vx * vx + vy * vy + vz * vz);
    cq[(size_t)(DIM + 1) * Nt + i] = rho * (ke + fs[(size_t)P_E * Nt + i]);
}

// SteadyStateLinearisation::matrix_vector_product set_dual
// (steady_state.cpp:58-67): value = cons, tangent = vec[cell*ncons + q] for
// valid cells; ghost rows hold zero (they are filled by the boundary
// actions after the conversion).
__global__ void __launch_bounds__(BLOCK)
    k_seed_cons(int N, int Nt, int nc, const double* __restrict__ cons,
                const double* __restrict__ vec, ibis::Dual* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nt) return;
    for (int q = 0; q < nc; q++) {
        const size_t k = (size_t)q * Nt + i;
        out[k] = i < N ? ibis::Dual(cons[k], vec[(size_t)i * nc + q]) : ibis::Dual(0.0, 0.0);
    }
}

// set_vector (steady_state.cpp:102-111): 1/dt* * v - dual(R)
__global__ void __launch_bounds__(BLOCK)
    k_jv_result(int N, int Nt, int nc, double dt_star, const double* __restrict__ vec,
                const ibis::Dual* __restrict__ dudt, double* __restrict__ result) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int q = 0; q < nc; q++) {
        const size_t v = (size_t)i * nc + q;
        result[v] = 1 / dt_star * vec[v] - dudt[(size_t)q * Nt + i].d;
    }
}

// eval_rhs (steady_state.cpp:141-148): rhs[cell*ncons + q] = R(cell, q)
__global__ void __launch_bounds__(BLOCK)
    k_rhs_vector(int N, int Nt, int nc, const double* __restrict__ dudt,
                 double* __restrict__ rhs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int q = 0; q < nc; q++) rhs[(size_t)i * nc + q] = dudt[(size_t)q * Nt + i];
}

// (value, tangent) pairs <-> separate value and tangent arrays
__global__ void __launch_bounds__(BLOCK)
    k_to_dual(long n, const double* __restrict__ x, const double* __restrict__ xd,
              ibis::Dual* __restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = ibis::Dual(x[i], xd[i]);
}

__global__ void __launch_bounds__(BLOCK)
    k_from_dual(int rows, int n, int stride, const ibis::Dual* __restrict__ in,
                double* __restrict__ x, double* __restrict__ xd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int q = 0; q < rows; q++) {
        const ibis::Dual v = in[(size_t)q * stride + i];
        x[(size_t)q * stride + i] = v.r;
        xd[(size_t)q * stride + i] = v.d;
    }
}

// u1 = u0 + k*dt  (apply_time_derivative, FV/conserved_quantities.cpp:80-94)
__global__ void __launch_bounds__(BLOCK)
    k_axpy(long n, const double* __restrict__ u0, double* u1, const double* __restrict__ k,
           double dt) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u1[i] = u0[i] + k[i] * dt;
}

#include "adjoint.hpp"

// zero the tangent of the limiters (frozen-limiter tangent residual)
__global__ void __launch_bounds__(BLOCK) k_freeze_tangent(long n, ibis::Dual* __restrict__ a) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i].d = 0.0;
}

// per boundary face of the selected tags: momentum flux * area, oriented
// out of the valid cell (the face force on the boundary)
__global__ void __launch_bounds__(BLOCK)
    k_force_terms(DevMesh m, const double* __restrict__ flux, unsigned long long tags,
                  double* __restrict__ out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.nb) return;
    const int f = m.b_face[b], F = m.F;
    const bool on = (tags >> m.b_tag[b]) & 1ull;
    const double A = m.geom[12 * (size_t)F + f];
    const double s = m.L[f] < m.N ? A : -A;
    for (int k = 0; k < 3; k++)
        out[(size_t)k * m.nb + b] = (on && k < m.dim) ? flux[(size_t)(1 + k) * F + f] * s : 0.0;
}

inline unsigned nblocks(long n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }
inline unsigned xcd_grid(long nlogical) {
    const long per = (nlogical + NXCD - 1) / NXCD;
    return (unsigned)(per * NXCD);
}

}  // namespace

// ===========================================================================
// host side
// ===========================================================================
int ibis_entry_indices_fit(int64_t maxf, int64_t n_cells, int64_t n_ghost) {
    return maxf * 5 * n_cells + 5 * n_ghost < (int64_t)INT32_MAX ? 1 : 0;
}

struct ibis_mesh {
    DevMesh d;
    int ntags = 0;
    std::vector<int> b_in_tag_order;  // device index of the k-th boundary face in tag order
    std::vector<void*> allocs;
    // tile tables of every tile size built so far (d holds those of TILE)
    std::map<int, DevMesh> tiles;
    std::mutex tiles_mu;
};

int ibis_mesh_ntags(const ibis_mesh* m) { return m ? m->ntags : 0; }

template <class T>
static ibis_status upload(ibis_mesh* m, const std::vector<T>& h, const T** out) {
    void* p = nullptr;
    size_t bytes = std::max<size_t>(h.size() * sizeof(T), 8);
    HIP_TRY(hipMalloc(&p, bytes));
    m->allocs.push_back(p);
    if (!h.empty()) HIP_TRY(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    *out = static_cast<const T*>(p);
    return IBIS_OK;
}

// Tile tables of one tile size ts: tiles of ts consecutive cells, the faces
// touching each tile (ascending, and in the per-entry face kernel's order),
// and each tile's stencil halo.
struct TileHost {
    int ts = 0, ntiles = 0, maxtf = 0, maxhalo = 0;
    std::vector<int> toff, tfaces, hoff, hids, efaces;
    std::vector<unsigned short> clds, nlds, elds;
};

// From the cells' face, neighbour and face-count lists (maxf x N, column
// per slot) and each face's cells and its slot in their lists.
static ibis_status build_tiles(int ts, int64_t N, int maxf, const std::vector<int>& ef,
                               const std::vector<int>& en, const std::vector<unsigned char>& nf,
                               const std::vector<int>& L, const std::vector<int>& R,
                               const std::vector<unsigned char>& slotL,
                               const std::vector<unsigned char>& slotR, TileHost* out) {
    TileHost& T = *out;
    T.ts = ts;
    const int ntiles = (int)((N + ts - 1) / ts);
    T.ntiles = ntiles;
    // faces touching each tile, and for every (cell, face slot) the face's
    // position in its tile's list
    T.toff.assign(ntiles + 1, 0);
    T.clds.assign((size_t)maxf * N, 0);
    {
        std::vector<int> fl;
        for (int t = 0; t < ntiles; t++) {
            fl.clear();
            const int64_t c0 = (int64_t)t * ts, c1 = std::min<int64_t>(N, c0 + ts);
            for (int64_t c = c0; c < c1; c++)
                for (int j = 0; j < nf[c]; j++) fl.push_back(ef[(size_t)j * N + c]);
            std::sort(fl.begin(), fl.end());
            fl.erase(std::unique(fl.begin(), fl.end()), fl.end());
            if (fl.size() > 65535)
                return ibis::fail(IBIS_ERR_INVALID, "tile face list too large for 16-bit positions");
            T.maxtf = std::max<int>(T.maxtf, (int)fl.size());
            for (int64_t c = c0; c < c1; c++)
                for (int j = 0; j < nf[c]; j++) {
                    auto it = std::lower_bound(fl.begin(), fl.end(), ef[(size_t)j * N + c]);
                    T.clds[(size_t)j * N + c] = (unsigned short)(it - fl.begin());
                }
            T.tfaces.insert(T.tfaces.end(), fl.begin(), fl.end());
            T.toff[t + 1] = (int)T.tfaces.size();
        }
    }
    // stencil halo of each tile: the neighbours (valid or ghost) of its
    // cells that lie outside it, ascending; for every (cell, face slot) the
    // neighbour's position in the tile's LDS stencil (own cells first, then
    // the halo)
    T.hoff.assign(ntiles + 1, 0);
    T.nlds.assign((size_t)maxf * N, 0);
    {
        std::vector<int> h;
        for (int t = 0; t < ntiles; t++) {
            h.clear();
            const int64_t c0 = (int64_t)t * ts, c1 = std::min<int64_t>(N, c0 + ts);
            for (int64_t c = c0; c < c1; c++)
                for (int j = 0; j < nf[c]; j++) {
                    const int64_t n = en[(size_t)j * N + c];
                    if (n < c0 || n >= c1) h.push_back((int)n);
                }
            std::sort(h.begin(), h.end());
            h.erase(std::unique(h.begin(), h.end()), h.end());
            T.maxhalo = std::max<int>(T.maxhalo, (int)h.size());
            if (ts + (int64_t)h.size() > 65535)
                return ibis::fail(IBIS_ERR_INVALID, "ibis_mesh_create: tile stencil too large");
            for (int64_t c = c0; c < c1; c++)
                for (int j = 0; j < nf[c]; j++) {
                    const int64_t n = en[(size_t)j * N + c];
                    const int64_t k = (n >= c0 && n < c1)
                                          ? n - c0
                                          : ts + (std::lower_bound(h.begin(), h.end(), (int)n) - h.begin());
                    T.nlds[(size_t)j * N + c] = (unsigned short)k;
                }
            T.hids.insert(T.hids.end(), h.begin(), h.end());
            T.hoff[t + 1] = (int)T.hids.size();
        }
    }
    // entry order of the tile lists for the per-entry face kernel: within a
    // tile, faces by (slot of the lower-numbered cell, slot of the other,
    // lower cell), so that neighbouring lanes read neighbouring cells of one
    // face-value plane on both sides whichever side is the face's left;
    // elds is each (cell, slot)'s position in that order
    T.efaces.assign(T.tfaces.size(), 0);
    T.elds.assign((size_t)maxf * N, 0);
    {
        std::vector<int> idx, inv;
        for (int t = 0; t < ntiles; t++) {
            const int a = T.toff[t], n = T.toff[t + 1] - a;
            idx.resize(n);
            inv.resize(n);
            std::iota(idx.begin(), idx.end(), 0);
            auto key = [&](int k) {
                const int f = T.tfaces[a + k];
                const bool lo_is_left = L[f] < R[f];
                return std::make_tuple(lo_is_left ? slotL[f] : slotR[f],
                                       lo_is_left ? slotR[f] : slotL[f], std::min(L[f], R[f]), f);
            };
            std::sort(idx.begin(), idx.end(), [&](int x, int y) { return key(x) < key(y); });
            for (int k = 0; k < n; k++) {
                T.efaces[a + k] = T.tfaces[a + idx[k]];
                inv[idx[k]] = k;
            }
            const int64_t c0 = (int64_t)t * ts, c1 = std::min<int64_t>(N, c0 + ts);
            for (int64_t c = c0; c < c1; c++)
                for (int j = 0; j < nf[c]; j++)
                    T.elds[(size_t)j * N + c] = (unsigned short)inv[T.clds[(size_t)j * N + c]];
        }
    }
    return IBIS_OK;
}

static ibis_status upload_tiles(ibis_mesh* m, const TileHost& T, DevMesh* D) {
    IBIS_TRY(upload(m, T.toff, &D->tile_face_off));
    IBIS_TRY(upload(m, T.tfaces, &D->tile_faces));
    IBIS_TRY(upload(m, T.clds, &D->cell_lds));
    IBIS_TRY(upload(m, T.efaces, &D->ent_faces));
    IBIS_TRY(upload(m, T.elds, &D->ent_lds));
    IBIS_TRY(upload(m, T.hoff, &D->halo_off));
    IBIS_TRY(upload(m, T.hids, &D->halo_ids));
    IBIS_TRY(upload(m, T.nlds, &D->nbr_lds));
    D->tile = T.ts;
    D->ntiles = T.ntiles;
    D->maxtf = T.maxtf;
    D->ntf = T.toff[T.ntiles];
    D->maxhalo = T.maxhalo;
    return IBIS_OK;
}

void ibis_mesh_destroy(ibis_mesh* m) {
    if (!m) return;
    for (void* p : m->allocs) (void)hipFree(p);
    delete m;
}

ibis_status ibis_mesh_create(const ibis_mesh_desc* d, ibis_mesh** out) {
    if (!d || !out) return ibis::fail(IBIS_ERR_INVALID, "ibis_mesh_create: null argument");
    if (d->dim != 2 && d->dim != 3) return ibis::fail(IBIS_ERR_INVALID, "dim must be 2 or 3");
    const int64_t N = d->n_cells, G = d->n_ghost, F = d->n_faces, Nt = N + G;
    if (N <= 0 || F <= 0 || Nt >= (int64_t)INT32_MAX || F >= (int64_t)INT32_MAX)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_mesh_create: mesh size out of range");
    if (d->ntags < 0 || d->ntags > IBIS_MAX_MESH_TAGS)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_mesh_create: too many boundary tags");
    // validate connectivity on the host: every index the kernels follow
    int maxf = 0;
    for (int64_t c = 0; c < N; c++) {
        int64_t n = d->cell_face_off[c + 1] - d->cell_face_off[c];
        if (n <= 0 || n > MAXF)
            return ibis::fail(IBIS_ERR_INVALID, "ibis_mesh_create: cells must have 1..6 faces");
        maxf = std::max<int>(maxf, (int)n);
    }
    const int64_t nnz = d->cell_face_off[N];
    for (int64_t j = 0; j < nnz; j++) {
        if (d->cell_faces[j] < 0 || d->cell_faces[j] >= F || d->cell_neighbours[j] < 0 ||
            d->cell_neighbours[j] >= Nt)
            return ibis::fail(IBIS_ERR_INVALID, "ibis_mesh_create: cell list index out of range");
    }
    for (int64_t f = 0; f < F; f++)
        if (d->face_left[f] < 0 || d->face_left[f] >= Nt || d->face_right[f] < 0 ||
            d->face_right[f] >= Nt)
            return ibis::fail(IBIS_ERR_INVALID, "ibis_mesh_create: face cell index out of range");
    // outsign +1 <=> the cell is the face's left cell (GR/grid.h:204-212)
    for (int64_t c = 0; c < N; c++)
        for (int64_t j = d->cell_face_off[c]; j < d->cell_face_off[c + 1]; j++) {
            const int64_t f = d->cell_faces[j];
            const bool left = d->face_left[f] == c;
            if (!(left ? d->cell_outsigns[j] == 1 : (d->cell_outsigns[j] == -1 && d->face_right[f] == c)))
                return ibis::fail(IBIS_ERR_INVALID,
                                  "ibis_mesh_create: outsigns disagree with face left/right cells");
        }
    const int64_t nb = d->tag_off[d->ntags];
    std::vector<int> b_face(nb), b_ghost(nb), b_valid(nb), b_tag(nb), b_of_ghost(G > 0 ? G : 1, -1);
    for (int t = 0; t < d->ntags; t++)
        for (int64_t k = d->tag_off[t]; k < d->tag_off[t + 1]; k++) {
            int64_t f = d->bfaces[k];
            if (f < 0 || f >= F) return ibis::fail(IBIS_ERR_INVALID, "boundary face out of range");
            int64_t l = d->face_left[f], r = d->face_right[f];
            b_face[k] = (int)f;
            b_valid[k] = (int)(l < N ? l : r);
            b_ghost[k] = (int)(l < N ? r : l);
            b_tag[k] = t;
            if (b_ghost[k] < N)
                return ibis::fail(IBIS_ERR_INVALID, "boundary face without a ghost cell");
            if (b_of_ghost[b_ghost[k] - N] != -1)
                return ibis::fail(IBIS_ERR_INVALID, "a ghost cell on two boundary faces");
            b_of_ghost[b_ghost[k] - N] = (int)k;
        }
    // every ghost a face or a cell list reaches must lie on a tagged
    // boundary face: the boundary and adjoint kernels index their tables
    // with the ghost's boundary face
    for (int64_t f = 0; f < F; f++)
        for (int64_t c : {d->face_left[f], d->face_right[f]})
            if (c >= N && b_of_ghost[c - N] == -1)
                return ibis::fail(IBIS_ERR_INVALID,
                                  "ibis_mesh_create: a face reaches a ghost cell that is on no "
                                  "tagged boundary face");
    for (int64_t j = 0; j < nnz; j++)
        if (d->cell_neighbours[j] >= N && b_of_ghost[d->cell_neighbours[j] - N] == -1)
            return ibis::fail(IBIS_ERR_INVALID,
                              "ibis_mesh_create: a cell's neighbour is a ghost cell that is on no "
                              "tagged boundary face");
    // device order of the boundary faces: by ghost id, so that neighbouring
    // threads of the boundary kernels write neighbouring ghost rows and read
    // nearby valid cells (the locality renumbering places each ghost after
    // its valid neighbour); b_in_tag_order keeps the tag order for sums
    std::vector<int> ord(nb), b_in_tag_order(nb);
    std::iota(ord.begin(), ord.end(), 0);
    std::sort(ord.begin(), ord.end(), [&](int a, int b) { return b_ghost[a] < b_ghost[b]; });
    {
        std::vector<int> pf(nb), pg(nb), pv(nb), pt(nb);
        for (int64_t i = 0; i < nb; i++) {
            const int k = ord[i];
            pf[i] = b_face[k];
            pg[i] = b_ghost[k];
            pv[i] = b_valid[k];
            pt[i] = b_tag[k];
            b_of_ghost[pg[i] - N] = (int)i;
            b_in_tag_order[k] = (int)i;
        }
        b_face.swap(pf);
        b_ghost.swap(pg);
        b_valid.swap(pv);
        b_tag.swap(pt);
    }
    ibis_mesh* m = new ibis_mesh();
    m->ntags = d->ntags;
    m->b_in_tag_order = b_in_tag_order;
    DevMesh& D = m->d;
    D.dim = d->dim;
    D.nl = d->dim == 2 ? 3 : 6;
    D.maxf = maxf;
    D.N = (int)N;
    D.G = (int)G;
    D.Nt = (int)Nt;
    D.F = (int)F;
    D.nb = (int)nb;
    std::vector<int> L(F), R(F);
    for (int64_t f = 0; f < F; f++) {
        L[f] = (int)d->face_left[f];
        R[f] = (int)d->face_right[f];
    }
    std::vector<int> ef((size_t)maxf * N, 0), en((size_t)maxf * N, 0);
    std::vector<signed char> es((size_t)maxf * N, 0);
    std::vector<unsigned char> nf(N);
    for (int64_t c = 0; c < N; c++) {
        int64_t o = d->cell_face_off[c];
        nf[c] = (unsigned char)(d->cell_face_off[c + 1] - o);
        for (int j = 0; j < nf[c]; j++) {
            ef[(size_t)j * N + c] = (int)d->cell_faces[o + j];
            en[(size_t)j * N + c] = (int)d->cell_neighbours[o + j];
            es[(size_t)j * N + c] = (signed char)d->cell_outsigns[o + j];
        }
    }
    std::vector<unsigned char> slotL(F, 0), slotR(F, 0);
    for (int64_t c = 0; c < N; c++)
        for (int64_t j = d->cell_face_off[c]; j < d->cell_face_off[c + 1]; j++) {
            const int64_t f = d->cell_faces[j];
            if (d->face_left[f] == c)
                slotL[f] = (unsigned char)(j - d->cell_face_off[c]);
            else
                slotR[f] = (unsigned char)(j - d->cell_face_off[c]);
        }
    TileHost tiles;
    {
        const ibis_status ts = build_tiles(TILE, N, maxf, ef, en, nf, L, R, slotL, slotR, &tiles);
        if (ts != IBIS_OK) {
            delete m;
            return ts;
        }
    }
    // the LSQ factors as the grid builder forms them from the centroids
    // (WLSGradient::compute_weights, GR/gradient.h:119-162): when a caller's
    // mesh holds exactly these, the staged cell kernel recomputes them from
    // its LDS stencil instead of reading 8 * nl bytes per cell
    int lsq_from_cen = 1;
    for (int64_t i = 0; i < N && lsq_from_cen; i++) {
        const double* c = d->centroid;
        double sxx = 0.0, sxy = 0.0, sxz = 0.0, syy = 0.0, syz = 0.0, szz = 0.0;
        for (int64_t j = d->cell_face_off[i]; j < d->cell_face_off[i + 1]; j++) {
            const int64_t n = d->cell_neighbours[j];
            const double dx = c[n] - c[i], dy = c[Nt + n] - c[Nt + i],
                         dz = c[2 * Nt + n] - c[2 * Nt + i];
            sxx += dx * dx;
            sxy += dx * dy;
            sxz += dx * dz;
            syy += dy * dy;
            syz += dy * dz;
            szz += dz * dz;
        }
        const double r11 = std::sqrt(sxx);
        const double r12 = 1.0 / r11 * sxy;
        const double r22 = std::sqrt(syy - r12 * r12);
        const double r13 = 1.0 / r11 * sxz;
        const double r23 = 1.0 / r22 * (syz - r12 / r11 * sxz);
        const double r33 = std::sqrt(szz - (r13 * r13 + r23 * r23));
        const double want[6] = {r11, r12, r22, r13, r23, r33};
        for (int a = 0; a < D.nl; a++)
            if (std::memcmp(&want[a], d->lsq + (size_t)a * N + i, 8)) lsq_from_cen = 0;
    }
    int n_from_t = 1;
    for (int64_t f = 0; f < F && n_from_t; f++) {
        const double* gg = d->face_geom;
        const double t1x = gg[3 * F + f], t1y = gg[4 * F + f], t1z = gg[5 * F + f];
        const double t2x = gg[6 * F + f], t2y = gg[7 * F + f], t2z = gg[8 * F + f];
        const double nx = t1y * t2z - t1z * t2y;
        const double ny = t1z * t2x - t1x * t2z;
        const double nz = t1x * t2y - t1y * t2x;
        if (std::memcmp(&nx, gg + f, 8) || std::memcmp(&ny, gg + F + f, 8) ||
            std::memcmp(&nz, gg + 2 * F + f, 8))
            n_from_t = 0;
    }
    int planar2d = d->dim == 2 ? 1 : 0;
    for (int64_t f = 0; f < F && planar2d; f++) {
        const double zero = 0.0, one = 1.0;
        const double* gg = d->face_geom;
        if (std::memcmp(gg + 5 * F + f, &zero, 8) || std::memcmp(gg + 6 * F + f, &zero, 8) ||
            std::memcmp(gg + 7 * F + f, &zero, 8) || std::memcmp(gg + 8 * F + f, &one, 8))
            planar2d = 0;
    }
    int flat_z = d->dim == 2 && std::isfinite(d->centroid[2 * Nt]) ? 1 : 0;
    for (int64_t c = 0; c < Nt && flat_z; c++)
        if (std::memcmp(d->centroid + 2 * Nt + c, d->centroid + 2 * Nt, 8)) flat_z = 0;
    for (int64_t f = 0; f < F && flat_z; f++)
        if (std::memcmp(d->face_geom + 11 * F + f, d->centroid + 2 * Nt, 8)) flat_z = 0;
    std::vector<double> geom(d->face_geom, d->face_geom + 13 * F);
    std::vector<double> cen(d->centroid, d->centroid + 3 * Nt);
    std::vector<double> vol(d->volume, d->volume + Nt);
    std::vector<double> lsq(d->lsq, d->lsq + (size_t)D.nl * N);
    ibis_status s;
#define UP(vec, field)                                      \
    if ((s = upload(m, vec, &D.field)) != IBIS_OK) {        \
        ibis_mesh_destroy(m);                               \
        return s;                                           \
    }
    UP(L, L);
    UP(R, R);
    UP(geom, geom);
    UP(cen, cen);
    UP(vol, vol);
    UP(lsq, lsq);
    UP(ef, ell_face);
    UP(en, ell_nbr);
    UP(es, ell_sign);
    UP(nf, nfaces);
    UP(b_face, b_face);
    UP(b_ghost, b_ghost);
    UP(b_valid, b_valid);
    UP(b_tag, b_tag);
    UP(b_of_ghost, b_of_ghost);
    UP(slotL, slotL);
    UP(slotR, slotR);
#undef UP
    if ((s = upload_tiles(m, tiles, &D)) != IBIS_OK) {
        ibis_mesh_destroy(m);
        return s;
    }
    m->tiles[TILE] = D;
    D.n_from_t = n_from_t;
    D.planar2d = planar2d;
    D.flat_z = flat_z;
    D.lsq_from_cen = lsq_from_cen;
    *out = m;
    return IBIS_OK;
}

// The tile fields of a tile size: built on first use from the device copies
// of the cell lists (the mesh keeps every size it has built until it is
// destroyed), then copied into *D.
static ibis_status mesh_tiles(const ibis_mesh* mc, int ts, DevMesh* D) {
    ibis_mesh* m = const_cast<ibis_mesh*>(mc);
    std::lock_guard<std::mutex> lock(m->tiles_mu);
    auto it = m->tiles.find(ts);
    if (it == m->tiles.end()) {
        const DevMesh& M = m->d;
        const size_t N = M.N, F = M.F, E = (size_t)M.maxf * N;
        std::vector<int> ef(E), en(E), L(F), R(F);
        std::vector<unsigned char> nf(N), slotL(F), slotR(F);
        HIP_TRY(hipMemcpy(ef.data(), M.ell_face, E * sizeof(int), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(en.data(), M.ell_nbr, E * sizeof(int), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(nf.data(), M.nfaces, N, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(L.data(), M.L, F * sizeof(int), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(R.data(), M.R, F * sizeof(int), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(slotL.data(), M.slotL, F, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(slotR.data(), M.slotR, F, hipMemcpyDeviceToHost));
        TileHost T;
        IBIS_TRY(build_tiles(ts, (int64_t)N, M.maxf, ef, en, nf, L, R, slotL, slotR, &T));
        DevMesh t = M;
        IBIS_TRY(upload_tiles(m, T, &t));
        it = m->tiles.emplace(ts, t).first;
    }
    const DevMesh& t = it->second;
    D->tile = t.tile;
    D->ntiles = t.ntiles;
    D->maxtf = t.maxtf;
    D->ntf = t.ntf;
    D->tile_face_off = t.tile_face_off;
    D->tile_faces = t.tile_faces;
    D->cell_lds = t.cell_lds;
    D->ent_faces = t.ent_faces;
    D->ent_lds = t.ent_lds;
    D->maxhalo = t.maxhalo;
    D->halo_off = t.halo_off;
    D->halo_ids = t.halo_ids;
    D->nbr_lds = t.nbr_lds;
    return IBIS_OK;
}

ibis_status ibis_mesh_from_grid(const ibis_grid* g, ibis_mesh** out) {
    if (!g) return ibis::fail(IBIS_ERR_INVALID, "ibis_mesh_from_grid: null grid");
    ibis_mesh_desc d;
    ibis_grid_fill_desc(g, &d);
    return ibis_mesh_create(&d, out);
}

struct ibis_residual {
    const ibis_mesh* m = nullptr;
    DevMesh d;  // the mesh's device view with the tile tables of option "tile"
    ibis_fv_desc desc;
    bool premul = false;
    double* grad = nullptr;
    double* lim = nullptr;
    double* flux = nullptr;
    double* pack = nullptr;  // per-cell AoS values + centroid (face-state path)
    double* FL = nullptr;    // left / right face states (face-state path)
    double* FR = nullptr;
    double* offs = nullptr;  // precomputed face offsets (option precomputed_offsets)
    int opt_layout = 0;      // option face_state_layout
    int prec = IBIS_PREC_FP64;  // ibis_residual_set_precision
    int ntab = 0;               // tag x action entries of each boundary table
    bool has_pre = false, has_visc = false, has_post = false;
    bool has_subsonic = false;  // a subsonic inflow/outflow action in the pre or viscous tables
    uint64_t generation = 0;    // bumped by every option or precision change
    int opt_offsets = 0;     // option precomputed_offsets
    int opt_pack = 1;        // option pack_cells
    int opt_stage = 1;       // option stage_offsets
    int opt_cell_prefetch = 0;  // option cell_prefetch (layout 2, pack_cells 0)
    int opt_stage_cells = 0;    // option stage_cells (layout 2, pack_cells 0)
    int opt_keep_grads = 0;     // option keep_gradients: the fused path writes the
                                // gradient and limiter slots too
    int opt_fuse_visc = 1;      // option fuse_viscous: the four-kernel viscous path
    int opt_fold = 0;           // option fold_boundary: the fused paths' boundary pass
                                // runs in the cell kernel's stencil staging
    int force_slots = 0;        // the adjoint's forward pass: the fused cell kernel
                                // writes the gradient and limiter slots
    double* gconv = nullptr;    // viscous: the ghosts' reconstruction variables (5 x G)
    double* vterm = nullptr;    // viscous: (dim+1) x F viscous terms of the fused path
    int opt_face_kernel = 0; // option face_kernel: 0 tile face lists, 1 the same with
                             // per-entry face data, 64/128/256 slot-major
    int2* ent_idx = nullptr;     // per tile-list entry face data (face_kernel 1)
    void* ent_geom32 = nullptr;  // float entry tangents of the reduced-precision modes
    float* offs32 = nullptr;     // float face offsets of the FP32 reconstruction mode
    double* ent_geom = nullptr;
    double* face_rho = nullptr;
    double* face_T = nullptr;
    int* kinds = nullptr;    // pre, visc, post
    double* pars = nullptr;  // pre, visc, post
    double* partial = nullptr;
    double* dt_dev = nullptr;
    unsigned long long* count_dev = nullptr;
    double* host_buf = nullptr;  // pinned: [0] dt, [1] count
    // per-kernel event timing
    bool profiling = false;
    struct Ev {
        int slot;
        hipEvent_t a, b;
    };
    std::vector<Ev> pending;
    std::vector<hipEvent_t> pool;
    double acc_ms[IBIS_NKERNELS] = {0};
    int64_t acc_n[IBIS_NKERNELS] = {0};
    // RK stage storage
    std::vector<double*> k;
    double* k_tmp = nullptr;
    double* flow_tmp = nullptr;
    // adjoint (VJP) scratch, allocated on first use
    std::vector<void*> adj;
    double *adj_prim_c = nullptr, *adj_vgrad = nullptr, *adj_face = nullptr, *adj_cell = nullptr;
    double *adj_bcb = nullptr, *adj_zero = nullptr, *adj_dudt = nullptr, *adj_force = nullptr;
    double* adj_lsqw = nullptr;  // LSQ weight table of the reverse kernels
    int freeze_lim = 0;  // option freeze_limiter (tangent residual only)
    // tangent (JVP) state and scratch, allocated on first use
    ibis::Dual* dual_prim = nullptr;
    ibis::Dual* dual_dudt = nullptr;
    ibis::Dual* dual_grad = nullptr;
    ibis::Dual* dual_lim = nullptr;
    ibis::Dual* dual_flux = nullptr;
};

uint64_t ibis_residual_generation(const ibis_residual* r) { return r ? r->generation : 0; }

void ibis_residual_destroy(ibis_residual* r) {
    if (!r) return;
    for (void* p : {(void*)r->grad, (void*)r->lim, (void*)r->flux, (void*)r->pack, (void*)r->FL, (void*)r->FR, (void*)r->offs, (void*)r->ent_idx, (void*)r->ent_geom, r->ent_geom32, (void*)r->offs32, (void*)r->gconv, (void*)r->vterm, (void*)r->face_rho,
                    (void*)r->face_T, (void*)r->kinds, (void*)r->pars, (void*)r->partial,
                    (void*)r->dt_dev, (void*)r->count_dev, (void*)r->k_tmp, (void*)r->flow_tmp,
                    (void*)r->dual_prim, (void*)r->dual_dudt, (void*)r->dual_grad,
                    (void*)r->dual_lim, (void*)r->dual_flux})
        if (p) (void)hipFree(p);
    for (double* p : r->k) (void)hipFree(p);
    for (void* p : r->adj) (void)hipFree(p);
    for (auto& e : r->pending) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    for (hipEvent_t e : r->pool) (void)hipEventDestroy(e);
    if (r->host_buf) (void)hipHostFree(r->host_buf);
    delete r;
}

static bool kinds_valid(const int32_t* pre, const int32_t* visc, const int32_t* post, int n) {
    for (int q = 0; q < n; q++) {
        if (pre[q] < 0 || pre[q] > 7 || visc[q] < 0 || visc[q] > 7 ||
            (post[q] != 0 && post[q] != IBIS_BC_CONSTANT_FLUX))
            return false;
    }
    return true;
}

static ibis_status check_desc(const ibis_fv_desc& d) {
    if (d.flux < 0 || d.flux > 3) return ibis::fail(IBIS_ERR_INVALID, "Unknown flux calculator");
    if (d.order != 1 && d.order != 2) return ibis::fail(IBIS_ERR_INVALID, "Invalid reconstruction order");
    if (d.order > 1 && (d.thermo < 0 || d.thermo > 3))
        return ibis::fail(IBIS_ERR_INVALID, "Unknown thermo interpolator");
    if (d.order > 1 && (d.limiter < 0 || d.limiter > 2))
        return ibis::fail(IBIS_ERR_INVALID, "Unknown limiter");
    if (!kinds_valid(d.pre_kind, d.visc_kind, d.post_kind, IBIS_MAX_TAGS * IBIS_MAX_ACT))
        return ibis::fail(IBIS_ERR_INVALID, "Unknown boundary action");
    return IBIS_OK;
}

ibis_status ibis_residual_create(const ibis_mesh* m, const ibis_fv_desc* desc,
                                 ibis_residual** out) {
    if (!m || !desc || !out) return ibis::fail(IBIS_ERR_INVALID, "ibis_residual_create: null argument");
    ibis_status st = check_desc(*desc);
    if (st != IBIS_OK) return st;
    if (m->ntags > IBIS_MAX_TAGS)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_residual_create: the mesh has more boundary tags "
                                            "than a configuration holds");
    const int ntab = IBIS_MAX_TAGS * IBIS_MAX_ACT;
    std::vector<int32_t> kinds(3 * ntab);
    std::vector<double> pars(3 * (size_t)ntab * IBIS_NPAR);
    std::memcpy(&kinds[0], desc->pre_kind, ntab * sizeof(int32_t));
    std::memcpy(&kinds[ntab], desc->visc_kind, ntab * sizeof(int32_t));
    std::memcpy(&kinds[2 * ntab], desc->post_kind, ntab * sizeof(int32_t));
    std::memcpy(&pars[0], desc->pre_par, ntab * IBIS_NPAR * sizeof(double));
    std::memcpy(&pars[(size_t)ntab * IBIS_NPAR], desc->visc_par, ntab * IBIS_NPAR * sizeof(double));
    std::memcpy(&pars[2 * (size_t)ntab * IBIS_NPAR], desc->post_par, ntab * IBIS_NPAR * sizeof(double));
    return ibis_residual_create_tables(m, desc, IBIS_MAX_TAGS, kinds.data(), pars.data(), out);
}

ibis_status ibis_residual_create_tables(const ibis_mesh* m, const ibis_fv_desc* desc, int ntags,
                                        const int32_t* kinds3, const double* pars3,
                                        ibis_residual** out) {
    if (!m || !desc || !kinds3 || !pars3 || !out || ntags < m->ntags || ntags <= 0)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_residual_create_tables: bad argument");
    const ibis_fv_desc& d = *desc;
    const int ntab = ntags * IBIS_MAX_ACT;
    if (!kinds_valid(kinds3, kinds3 + ntab, kinds3 + 2 * ntab, ntab))
        return ibis::fail(IBIS_ERR_INVALID, "Unknown boundary action");
    const DevMesh& D = m->d;
    ibis_residual* r = new ibis_residual();
    r->m = m;
    r->d = m->d;
    r->desc = d;
    // the unfused path keeps the flux unmultiplied (the viscous kernel adds to it)
    r->premul = false;
    const size_t N = D.N, F = D.F;
    const int nc = D.dim + 2;
    auto alloc0 = [&](void** p, size_t bytes) -> hipError_t {
        hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 8));
        if (e != hipSuccess) return e;
        return hipMemset(*p, 0, std::max<size_t>(bytes, 8));
    };
#define AL(ptr, bytes)                                                         \
    do {                                                                       \
        hipError_t e_ = alloc0((void**)&(ptr), (bytes));                       \
        if (e_ != hipSuccess) {                                                \
            ibis_residual_destroy(r);                                          \
            return ibis::fail(IBIS_ERR_HIP, hipGetErrorString(e_));            \
        }                                                                      \
    } while (0)
    AL(r->grad, sizeof(double) * NGRAD * 3 * N);
    AL(r->lim, sizeof(double) * NGRAD * N);
    AL(r->flux, sizeof(double) * nc * F);
    AL(r->pack, sizeof(double) * 8 * (size_t)D.Nt);
    AL(r->FL, sizeof(double) * 5 * std::max<size_t>(F, (size_t)D.N * D.maxf + D.G));
    AL(r->FR, sizeof(double) * 5 * F);
    if (d.viscous) {
        AL(r->face_rho, sizeof(double) * F);
        AL(r->face_T, sizeof(double) * F);
        AL(r->gconv, sizeof(double) * 5 * std::max<size_t>(D.G, 1));
        AL(r->vterm, sizeof(double) * (D.dim + 1) * F);
    }
    r->ntab = ntab;
    AL(r->kinds, sizeof(int) * 3 * ntab);
    AL(r->pars, sizeof(double) * 3 * ntab * IBIS_NPAR);
    AL(r->partial, sizeof(double) * (nblocks(D.N) + 1));
    AL(r->dt_dev, sizeof(double));
    AL(r->count_dev, sizeof(unsigned long long));
#undef AL
    // which action lists are non-empty (tags the mesh does not have hold none)
    for (int q = 0; q < m->ntags * IBIS_MAX_ACT; q++) {
        r->has_pre = r->has_pre || kinds3[q] != IBIS_BC_NONE;
        r->has_visc = r->has_visc || kinds3[ntab + q] != IBIS_BC_NONE;
        r->has_post = r->has_post || kinds3[2 * ntab + q] == IBIS_BC_CONSTANT_FLUX;
    }
    for (int q = 0; q < 2 * ntab; q++)
        r->has_subsonic = r->has_subsonic || kinds3[q] == IBIS_BC_SUBSONIC_INFLOW ||
                          kinds3[q] == IBIS_BC_SUBSONIC_OUTFLOW;
    r->has_pre = r->has_pre && D.nb > 0;
    r->has_visc = r->has_visc && D.nb > 0;
    r->has_post = r->has_post && D.nb > 0;
    // the boundary pass in the cell kernel's staging saves the boundary
    // kernel (a fixed latency) and costs the cell kernel occupancy (a cost per
    // cell): it pays on 2D meshes up to about a million cells; in 3D the
    // action code's registers cost more than it saves
    r->opt_fold = (D.dim == 2 && D.N <= FOLD_MAX_CELLS) ? 1 : 0;
    if (hipMemcpy(r->kinds, kinds3, 3 * (size_t)ntab * sizeof(int), hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(r->pars, pars3, 3 * (size_t)ntab * IBIS_NPAR * sizeof(double),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipHostMalloc((void**)&r->host_buf, 2 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        ibis_residual_destroy(r);
        return ibis::fail(IBIS_ERR_HIP, "ibis_residual_create: upload failed");
    }
    k_pack_geom<<<nblocks(D.Nt), BLOCK>>>(D.Nt, D.cen, r->pack);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        ibis_residual_destroy(r);
        return ibis::fail(IBIS_ERR_HIP, "ibis_residual_create: pack initialisation failed");
    }
    // the fused second-order inviscid path starts from the options that the
    // autotuner picks on the 3D and 2D workloads (every choice gives the
    // same bits; ibis_residual_set_option changes them)
    if (d.order == 2 && !r->has_post) {
        // face_kernel 1 indexes the face values with 32-bit entries; a mesh
        // too large for them keeps the tile-list kernel (face_kernel 0)
        const int fk = ibis_entry_indices_fit(D.maxf, D.N, D.G) ? 1 : 0;
        for (const auto& kv : {std::make_pair("face_state_layout", 2), std::make_pair("pack_cells", 0),
                               std::make_pair("precomputed_offsets", 1),
                               std::make_pair("face_kernel", fk), std::make_pair("stage_cells", 1)}) {
            const ibis_status st = ibis_residual_set_option(r, kv.first, kv.second);
            if (st != IBIS_OK) {
                ibis_residual_destroy(r);
                return st;
            }
        }
    }
    *out = r;
    return IBIS_OK;
}

namespace {

template <class S, int DIM>
void launch_cell_grad(const DevMesh& D, const VarSet& vs, int lim, const S* prim, S* grad,
                      S* limv, double eps, double K, hipStream_t st) {
    const int nlog = (D.N + 63) / 64;
    const unsigned g = xcd_grid(nlog);
    const unsigned b = 64 * vs.n;
    if (lim == 0)
        k_cell_grad_var<S, DIM, 0><<<g, b, 0, st>>>(D, vs, prim, grad, limv, eps, K, nlog);
    else if (lim == 1)
        k_cell_grad_var<S, DIM, 1><<<g, b, 0, st>>>(D, vs, prim, grad, limv, eps, K, nlog);
    else
        k_cell_grad_var<S, DIM, 2><<<g, b, 0, st>>>(D, vs, prim, grad, limv, eps, K, nlog);
}

template <int DIM, int FLUX>
void launch_face_tile_f(const DevMesh& D, const FaceParams& fp, bool recon, const double* prim,
                        const double* grad, const double* lim, double* dudt, hipStream_t st) {
    const unsigned g = xcd_grid(D.ntiles);
    const size_t shm = (size_t)(DIM + 2) * D.maxtf * sizeof(double);
    if (recon)
        k_face_tile<DIM, FLUX, true><<<g, D.tile, shm, st>>>(D, fp, prim, grad, lim, dudt);
    else
        k_face_tile<DIM, FLUX, false><<<g, D.tile, shm, st>>>(D, fp, prim, grad, lim, dudt);
}

template <int DIM>
void launch_face_tile(const DevMesh& D, int flux_kind, const FaceParams& fp, bool recon,
                      const double* prim, const double* grad, const double* lim, double* dudt,
                      hipStream_t st) {
    switch (flux_kind) {
        case IBIS_FLUX_HANEL:
            launch_face_tile_f<DIM, IBIS_FLUX_HANEL>(D, fp, recon, prim, grad, lim, dudt, st);
            break;
        case IBIS_FLUX_AUSMDV:
            launch_face_tile_f<DIM, IBIS_FLUX_AUSMDV>(D, fp, recon, prim, grad, lim, dudt, st);
            break;
        case IBIS_FLUX_LDFSS2:
            launch_face_tile_f<DIM, IBIS_FLUX_LDFSS2>(D, fp, recon, prim, grad, lim, dudt, st);
            break;
        default:
            launch_face_tile_f<DIM, IBIS_FLUX_RUSANOV>(D, fp, recon, prim, grad, lim, dudt, st);
            break;
    }
}

template <class S, int DIM, int FLUX>
void launch_face_flux_f(const DevMesh& D, const FaceParams& fp, bool recon, bool premul,
                        const S* prim, const S* grad, const S* lim, S* flux, hipStream_t st) {
    const unsigned nb = nblocks(D.F);
    if (recon) {
        if (premul)
            k_face_flux<S, DIM, FLUX, true, true><<<nb, BLOCK, 0, st>>>(D, fp, prim, grad, lim, flux);
        else
            k_face_flux<S, DIM, FLUX, true, false><<<nb, BLOCK, 0, st>>>(D, fp, prim, grad, lim, flux);
    } else {
        if (premul)
            k_face_flux<S, DIM, FLUX, false, true><<<nb, BLOCK, 0, st>>>(D, fp, prim, grad, lim, flux);
        else
            k_face_flux<S, DIM, FLUX, false, false><<<nb, BLOCK, 0, st>>>(D, fp, prim, grad, lim, flux);
    }
}

template <class S, int DIM>
void launch_face_flux(const DevMesh& D, int flux_kind, const FaceParams& fp, bool recon,
                      bool premul, const S* prim, const S* grad, const S* lim, S* flux,
                      hipStream_t st) {
    switch (flux_kind) {
        case IBIS_FLUX_HANEL:
            launch_face_flux_f<S, DIM, IBIS_FLUX_HANEL>(D, fp, recon, premul, prim, grad, lim, flux, st);
            break;
        case IBIS_FLUX_AUSMDV:
            launch_face_flux_f<S, DIM, IBIS_FLUX_AUSMDV>(D, fp, recon, premul, prim, grad, lim, flux, st);
            break;
        case IBIS_FLUX_LDFSS2:
            launch_face_flux_f<S, DIM, IBIS_FLUX_LDFSS2>(D, fp, recon, premul, prim, grad, lim, flux, st);
            break;
        default:
            launch_face_flux_f<S, DIM, IBIS_FLUX_RUSANOV>(D, fp, recon, premul, prim, grad, lim, flux, st);
            break;
    }
}

// LDS bytes of the cell kernel's offset staging: maxf x 3 doubles per
// thread, only with a limiter (two passes over the faces) and only while
// the block's share stays within 64 KiB (maxf <= 10); 0 = not staged
size_t cell_stage_bytes(const DevMesh& D, int lim, int opt) {
    const size_t b = (size_t)D.maxf * 3 * D.tile * sizeof(double);
    return (opt && lim != 0 && b <= 65536) ? b : 0;
}

// the cell stencil goes to LDS: the option stage_cells, when its 8 planes
// fit one block's 64 KiB
bool stencil_in_lds(const ibis_residual* r) {
    const DevMesh& D = r->d;
    return r->opt_stage_cells && (size_t)8 * (D.tile + D.maxhalo) * sizeof(double) <= 65536;
}

// the inviscid fused path's boundary pass runs in the cell kernel's staging
// (option fold_boundary): FP64, layout 2 gathered from the state, stencil in LDS
bool fold_on(const ibis_residual* r) {
    return r->opt_fold && r->prec == IBIS_PREC_FP64 && r->opt_layout == 2 && !r->opt_pack &&
           stencil_in_lds(r);
}

// bf: the folded boundary pass (only when fold_on), or null
template <int DIM, int LIM, int LAYOUT, bool OFFS>
void launch_cell_fs_t(const DevMesh& D, const VarSet& vs, const ibis_fv_desc& d, ibis_residual* r,
                      int nlog, const double* prim, hipStream_t st, const BcFold* bf) {
    const size_t shm = cell_stage_bytes(D, LIM, r->opt_stage);
    const int stage = shm ? 1 : 0;
    const int ws = (r->opt_keep_grads || r->force_slots) ? 1 : 0;
    // the LDS cell stencil (option stage_cells), in all precisions; the
    // FP64 and FP32-flux kernels reuse it for the offsets (stage_offsets)
    const size_t shs0 = (size_t)8 * (D.tile + D.maxhalo) * sizeof(double);
    const bool sn = stencil_in_lds(r);
    const size_t shs = std::max(shs0, shm);
    if (r->prec == IBIS_PREC_FP32_RECON) {
        float* fl = reinterpret_cast<float*>(r->FL);
        if (sn)
            k_cell_fs32<DIM, LIM, OFFS, true><<<xcd_grid(nlog), D.tile, std::max(shs0, shm / 2), st>>>(
                D, vs, prim, r->grad, d.lim_eps, d.venkat_K, nlog, fl, r->offs, stage, r->offs32);
        else
            k_cell_fs32<DIM, LIM, OFFS><<<xcd_grid(nlog), D.tile, shm / 2, st>>>(
                D, vs, prim, r->grad, d.lim_eps, d.venkat_K, nlog, fl, r->offs, stage, r->offs32);
        return;
    }
    if (r->prec == IBIS_PREC_FP32_FLUX) {
        float* fl = reinterpret_cast<float*>(r->FL);
        if (sn)
            k_cell_fs<DIM, LIM, 2, OFFS, float, false, false, true><<<xcd_grid(nlog), D.tile, shs, st>>>(
                D, vs, r->pack, r->grad, r->lim, d.lim_eps, d.venkat_K, nlog, ws, fl, fl, r->offs,
                prim, stage, nullptr, nullptr);
        else
            k_cell_fs<DIM, LIM, 2, OFFS, float, false><<<xcd_grid(nlog), D.tile, shm, st>>>(
                D, vs, r->pack, r->grad, r->lim, d.lim_eps, d.venkat_K, nlog, ws, fl, fl, r->offs,
                prim, stage, nullptr, nullptr);
        return;
    }
    if (!r->opt_pack && LAYOUT == 2) {
        if (sn) {
            if (bf)
                k_cell_fs<DIM, LIM, 2, OFFS, double, false, false, true, false, true>
                    <<<xcd_grid(nlog), D.tile, shs, st>>>(
                        D, vs, r->pack, r->grad, r->lim, d.lim_eps, d.venkat_K, nlog, ws, r->FL,
                        r->FR, r->offs, prim, stage, nullptr, nullptr, *bf);
            else
                k_cell_fs<DIM, LIM, 2, OFFS, double, false, false, true>
                    <<<xcd_grid(nlog), D.tile, shs, st>>>(
                        D, vs, r->pack, r->grad, r->lim, d.lim_eps, d.venkat_K, nlog, ws, r->FL,
                        r->FR, r->offs, prim, stage, nullptr, nullptr);
            return;
        }
        if (r->opt_cell_prefetch)
            k_cell_fs<DIM, LIM, 2, OFFS, double, false, true><<<xcd_grid(nlog), D.tile, shm, st>>>(
                D, vs, r->pack, r->grad, r->lim, d.lim_eps, d.venkat_K, nlog, ws, r->FL, r->FR,
                r->offs, prim, stage, nullptr, nullptr);
        else
            k_cell_fs<DIM, LIM, 2, OFFS, double, false><<<xcd_grid(nlog), D.tile, shm, st>>>(
                D, vs, r->pack, r->grad, r->lim, d.lim_eps, d.venkat_K, nlog, ws, r->FL, r->FR,
                r->offs, prim, stage, nullptr, nullptr);
        return;
    }
    k_cell_fs<DIM, LIM, LAYOUT, OFFS><<<xcd_grid(nlog), D.tile, shm, st>>>(
        D, vs, r->pack, r->grad, r->lim, d.lim_eps, d.venkat_K, nlog, ws, r->FL, r->FR, r->offs,
        prim, stage, nullptr, nullptr);
}

template <int DIM, int LIM>
void launch_cell_fs_l(const DevMesh& D, const VarSet& vs, const ibis_fv_desc& d, ibis_residual* r,
                      int nlog, const double* prim, hipStream_t st, const BcFold* bf) {
    const int sel = r->opt_layout * 2 + (r->opt_offsets ? 1 : 0);
    switch (sel) {
        case 0: launch_cell_fs_t<DIM, LIM, 0, false>(D, vs, d, r, nlog, prim, st, bf); break;
        case 1: launch_cell_fs_t<DIM, LIM, 0, true>(D, vs, d, r, nlog, prim, st, bf); break;
        case 2: launch_cell_fs_t<DIM, LIM, 1, false>(D, vs, d, r, nlog, prim, st, bf); break;
        case 3: launch_cell_fs_t<DIM, LIM, 1, true>(D, vs, d, r, nlog, prim, st, bf); break;
        case 4: launch_cell_fs_t<DIM, LIM, 2, false>(D, vs, d, r, nlog, prim, st, bf); break;
        default: launch_cell_fs_t<DIM, LIM, 2, true>(D, vs, d, r, nlog, prim, st, bf); break;
    }
}

template <int DIM>
void launch_cell_fs(const DevMesh& D, const VarSet& vs, const ibis_fv_desc& d, ibis_residual* r,
                    int nlog, const double* prim, hipStream_t st, const BcFold* bf = nullptr) {
    if (d.limiter == 0)
        launch_cell_fs_l<DIM, 0>(D, vs, d, r, nlog, prim, st, bf);
    else if (d.limiter == 1)
        launch_cell_fs_l<DIM, 1>(D, vs, d, r, nlog, prim, st, bf);
    else
        launch_cell_fs_l<DIM, 2>(D, vs, d, r, nlog, prim, st, bf);
}

template <int DIM, int FLUX, int TS, class R>
void launch_face_slot(const DevMesh& D, const FaceParams& fp, const R* FL, double* dudt,
                      hipStream_t st) {
    const int nlog = (D.N + TS - 1) / TS;
    const size_t shm = (size_t)(DIM + 2) * D.maxf * TS * sizeof(double);
    k_face_slot_fs<DIM, FLUX, TS, R><<<xcd_grid(nlog), TS, shm, st>>>(D, fp, nlog, FL, dudt);
}

template <int DIM, int FLUX, class R>
void launch_face_slot_ts(const DevMesh& D, const FaceParams& fp, int ts, const R* FL, double* dudt,
                         hipStream_t st) {
    if (ts == 64)
        launch_face_slot<DIM, FLUX, 64, R>(D, fp, FL, dudt, st);
    else if (ts == 128)
        launch_face_slot<DIM, FLUX, 128, R>(D, fp, FL, dudt, st);
    else
        launch_face_slot<DIM, FLUX, 256, R>(D, fp, FL, dudt, st);
}

// f32: float face values in layout 2 (reduced-precision modes);
// face_kernel 0 = tile face lists, 1 = the same with per-entry face data,
// otherwise the slot-major kernel with that many cells per block
template <int DIM, int FLUX>
void launch_face_fs_f(const DevMesh& D, const FaceParams& fp, int layout, bool f32, int face_kernel,
                      const int2* eidx, const double* egeom, const void* eg32, const double* FL,
                      const double* FR, double* dudt, hipStream_t st) {
    const unsigned g = xcd_grid(D.ntiles);
    const size_t shm = (size_t)(DIM + 2) * D.maxtf * sizeof(double);
    if (f32) {
        const float* fl = reinterpret_cast<const float*>(FL);
        if (face_kernel == 1 && D.n_from_t && DIM == 2 && D.planar2d)
            k_face_entry_fs<DIM, FLUX, false, float, true><<<g, D.tile, shm, st>>>(D, fp, eidx, egeom, fl, dudt,
                                                                                 ibis_stage_epilogue{}, ViscFace{});
        else if (face_kernel == 1 && D.n_from_t && eg32)
            k_face_entry_fs<DIM, FLUX, false, float, false, true><<<g, D.tile, shm, st>>>(
                D, fp, eidx, reinterpret_cast<const double*>(eg32), fl, dudt, ibis_stage_epilogue{}, ViscFace{});
        else if (face_kernel == 1 && D.n_from_t)
            k_face_entry_fs<DIM, FLUX, false, float><<<g, D.tile, shm, st>>>(D, fp, eidx, egeom, fl, dudt,
                                                                          ibis_stage_epilogue{}, ViscFace{});
        else if (face_kernel == 1)
            k_face_entry_fs<DIM, FLUX, true, float><<<g, D.tile, shm, st>>>(D, fp, eidx, egeom, fl, dudt,
                                                                         ibis_stage_epilogue{}, ViscFace{});
        else if (face_kernel)
            launch_face_slot_ts<DIM, FLUX, float>(D, fp, face_kernel, fl, dudt, st);
        else
            k_face_tile_fs<DIM, FLUX, 2, float><<<g, D.tile, shm, st>>>(D, fp, fl, fl, dudt);
        return;
    }
    if (layout == 2 && face_kernel == 1) {
        const ibis_stage_epilogue none{};
        if (D.n_from_t && DIM == 2 && D.planar2d)
            k_face_entry_fs<DIM, FLUX, false, double, true><<<g, D.tile, shm, st>>>(D, fp, eidx, egeom, FL, dudt,
                                                                                  none, ViscFace{});
        else if (D.n_from_t)
            k_face_entry_fs<DIM, FLUX, false><<<g, D.tile, shm, st>>>(D, fp, eidx, egeom, FL, dudt, none, ViscFace{});
        else
            k_face_entry_fs<DIM, FLUX, true><<<g, D.tile, shm, st>>>(D, fp, eidx, egeom, FL, dudt, none, ViscFace{});
    } else if (layout == 2 && face_kernel) {
        launch_face_slot_ts<DIM, FLUX, double>(D, fp, face_kernel, FL, dudt, st);
    } else if (layout == 0) {
        k_face_tile_fs<DIM, FLUX, 0><<<g, D.tile, shm, st>>>(D, fp, FL, FR, dudt);
    } else if (layout == 1) {
        k_face_tile_fs<DIM, FLUX, 1><<<g, D.tile, shm, st>>>(D, fp, FL, FR, dudt);
    } else {
        k_face_tile_fs<DIM, FLUX, 2><<<g, D.tile, shm, st>>>(D, fp, FL, FR, dudt);
    }
}

// the per-entry face kernel with the Runge-Kutta stage epilogue (FP64,
// normals equal to cross(t1, t2))
template <int DIM, int FLUX>
void launch_face_epi_f(const DevMesh& D, const FaceParams& fp, const int2* eidx,
                       const double* egeom, const double* FL, double* dudt,
                       const ibis_stage_epilogue& epi, hipStream_t st) {
    const unsigned g = xcd_grid(D.ntiles);
    // fluxes, then (for the local time steps) normals and areas
    const size_t shm = (size_t)(DIM + 2 + (epi.dt_cells_out ? 4 : 0)) * D.maxtf * sizeof(double);
    if (DIM == 2 && D.planar2d)
        k_face_entry_fs<DIM, FLUX, false, double, true, false, true><<<g, D.tile, shm, st>>>(
            D, fp, eidx, egeom, FL, dudt, epi, ViscFace{});
    else
        k_face_entry_fs<DIM, FLUX, false, double, false, false, true><<<g, D.tile, shm, st>>>(
            D, fp, eidx, egeom, FL, dudt, epi, ViscFace{});
}

template <int DIM>
void launch_face_epi(const DevMesh& D, const FaceParams& fp, int flux, const int2* eidx,
                     const double* egeom, const double* FL, double* dudt,
                     const ibis_stage_epilogue& epi, hipStream_t st) {
    switch (flux) {
        case IBIS_FLUX_HANEL:
            launch_face_epi_f<DIM, IBIS_FLUX_HANEL>(D, fp, eidx, egeom, FL, dudt, epi, st);
            break;
        case IBIS_FLUX_AUSMDV:
            launch_face_epi_f<DIM, IBIS_FLUX_AUSMDV>(D, fp, eidx, egeom, FL, dudt, epi, st);
            break;
        case IBIS_FLUX_LDFSS2:
            launch_face_epi_f<DIM, IBIS_FLUX_LDFSS2>(D, fp, eidx, egeom, FL, dudt, epi, st);
            break;
        default:
            launch_face_epi_f<DIM, IBIS_FLUX_RUSANOV>(D, fp, eidx, egeom, FL, dudt, epi, st);
            break;
    }
}

template <int DIM>
void launch_face_fs(const DevMesh& D, const FaceParams& fp, int flux, int layout, bool f32,
                    int face_kernel, const int2* eidx, const double* egeom, const void* eg32,
                    const double* FL,
                    const double* FR, double* dudt, hipStream_t st) {
#define IBIS_FF(K) \
    launch_face_fs_f<DIM, K>(D, fp, layout, f32, face_kernel, eidx, egeom, eg32, FL, FR, dudt, st)
    switch (flux) {
        case IBIS_FLUX_HANEL: IBIS_FF(IBIS_FLUX_HANEL); break;
        case IBIS_FLUX_AUSMDV: IBIS_FF(IBIS_FLUX_AUSMDV); break;
        case IBIS_FLUX_LDFSS2: IBIS_FF(IBIS_FLUX_LDFSS2); break;
        default: IBIS_FF(IBIS_FLUX_RUSANOV); break;
    }
#undef IBIS_FF
}

// the fused viscous path's cell kernel (12-plane stencil in LDS)
size_t visc_stencil_bytes(const DevMesh& D) {
    return (size_t)12 * (D.tile + D.maxhalo) * sizeof(double);
}

// vgrad: where the viscous gradients go (r->grad for the residual; the
// adjoint keeps them apart); write_slots: also the convective gradients
// and limiters
template <int DIM, int LIM>
void launch_cell_visc_l(const DevMesh& D, const VarSet& vs, const ibis_fv_desc& d,
                        ibis_residual* r, const double* prim, double* vgrad, int write_slots,
                        hipStream_t st, const BcFold* bf) {
    const size_t shm = cell_stage_bytes(D, LIM, r->opt_stage);
    const int stage = shm ? 1 : 0;
    const size_t shs = std::max(visc_stencil_bytes(D), shm);
    if (r->prec == IBIS_PREC_FP32_FLUX) {  // float face values
        float* fl = reinterpret_cast<float*>(r->FL);
        k_cell_fs<DIM, LIM, 2, true, float, false, false, true, true>
            <<<xcd_grid(D.ntiles), D.tile, shs, st>>>(D, vs, r->pack, r->grad, r->lim, d.lim_eps,
                                                      d.venkat_K, D.ntiles, write_slots, fl, fl,
                                                      r->offs, prim, stage, r->gconv, vgrad);
        return;
    }
    if (bf)
        k_cell_fs<DIM, LIM, 2, true, double, false, false, true, true, true>
            <<<xcd_grid(D.ntiles), D.tile, shs, st>>>(D, vs, r->pack, r->grad, r->lim, d.lim_eps,
                                                      d.venkat_K, D.ntiles, write_slots, r->FL,
                                                      r->FR, r->offs, prim, stage, r->gconv,
                                                      vgrad, *bf);
    else
        k_cell_fs<DIM, LIM, 2, true, double, false, false, true, true>
            <<<xcd_grid(D.ntiles), D.tile, shs, st>>>(D, vs, r->pack, r->grad, r->lim, d.lim_eps,
                                                      d.venkat_K, D.ntiles, write_slots, r->FL,
                                                      r->FR, r->offs, prim, stage, r->gconv,
                                                      vgrad);
}

// bf: the boundary pass folded into the staging (FP64 only), or null
template <int DIM>
void launch_cell_visc(const DevMesh& D, const VarSet& vs, const ibis_fv_desc& d, ibis_residual* r,
                      const double* prim, double* vgrad, int write_slots, hipStream_t st,
                      const BcFold* bf = nullptr) {
    if (d.limiter == 0)
        launch_cell_visc_l<DIM, 0>(D, vs, d, r, prim, vgrad, write_slots, st, bf);
    else if (d.limiter == 1)
        launch_cell_visc_l<DIM, 1>(D, vs, d, r, prim, vgrad, write_slots, st, bf);
    else
        launch_cell_visc_l<DIM, 2>(D, vs, d, r, prim, vgrad, write_slots, st, bf);
}

// the fused viscous path can run this residual (second order, the options
// its kernels need, a stencil that fits the LDS)
bool visc_fusable(const ibis_residual* r, const VarSet& vs) {
    const DevMesh& D = r->d;
    const ibis_fv_desc& d = r->desc;
    return d.viscous && d.order == 2 && !r->has_post && r->prec != IBIS_PREC_FP32_RECON &&
           r->opt_fuse_visc && r->opt_layout == 2 && r->opt_face_kernel == 1 && r->ent_idx &&
           r->vterm && r->offs && r->opt_offsets && !r->opt_pack && r->opt_stage_cells &&
           D.n_from_t && visc_stencil_bytes(D) <= 65536 && vs.row[2] == P_VX &&
           vs.row[3] == P_VY && vs.row[4] == P_VZ;
}

// the fused viscous path's face kernel (per-entry face data, FP64)
template <int DIM, int FLUX>
void launch_face_visc_f(const DevMesh& D, const FaceParams& fp, const int2* eidx,
                        const double* egeom, const void* eg32, const double* FL, bool f32,
                        double* dudt, const ViscFace& vf, hipStream_t st) {
    const unsigned g = xcd_grid(D.ntiles);
    const size_t shm = (size_t)(DIM + 2) * D.maxtf * sizeof(double);
    const ibis_stage_epilogue none{};
    if (f32) {  // float face values and flux arithmetic, FP64 viscous terms and sums
        const float* fl = reinterpret_cast<const float*>(FL);
        if (DIM == 2 && D.planar2d)
            k_face_entry_fs<DIM, FLUX, false, float, true, false, false, true><<<g, D.tile, shm, st>>>(
                D, fp, eidx, egeom, fl, dudt, none, vf);
        else if (eg32)
            k_face_entry_fs<DIM, FLUX, false, float, false, true, false, true><<<g, D.tile, shm, st>>>(
                D, fp, eidx, reinterpret_cast<const double*>(eg32), fl, dudt, none, vf);
        else
            k_face_entry_fs<DIM, FLUX, false, float, false, false, false, true><<<g, D.tile, shm, st>>>(
                D, fp, eidx, egeom, fl, dudt, none, vf);
        return;
    }
    if (DIM == 2 && D.planar2d)
        k_face_entry_fs<DIM, FLUX, false, double, true, false, false, true><<<g, D.tile, shm, st>>>(
            D, fp, eidx, egeom, FL, dudt, ibis_stage_epilogue{}, vf);
    else
        k_face_entry_fs<DIM, FLUX, false, double, false, false, false, true><<<g, D.tile, shm, st>>>(
            D, fp, eidx, egeom, FL, dudt, ibis_stage_epilogue{}, vf);
}

template <int DIM>
void launch_face_visc(const DevMesh& D, const FaceParams& fp, int flux, const int2* eidx,
                      const double* egeom, const void* eg32, const double* FL, bool f32,
                      double* dudt, const ViscFace& vf, hipStream_t st) {
#define IBIS_FV(K) launch_face_visc_f<DIM, K>(D, fp, eidx, egeom, eg32, FL, f32, dudt, vf, st)
    switch (flux) {
        case IBIS_FLUX_HANEL: IBIS_FV(IBIS_FLUX_HANEL); break;
        case IBIS_FLUX_AUSMDV: IBIS_FV(IBIS_FLUX_AUSMDV); break;
        case IBIS_FLUX_LDFSS2: IBIS_FV(IBIS_FLUX_LDFSS2); break;
        default: IBIS_FV(IBIS_FLUX_RUSANOV); break;
    }
#undef IBIS_FV
}

// thermo reconstruction pair: (slot, prim row) of the two variables,
// FV/convective_flux.cpp:181-209
void thermo_pair(int thermo, int* sA, int* rA, int* sB, int* rB) {
    switch (thermo) {
        case IBIS_THERMO_RHO_P: *sA = G_P; *rA = P_P; *sB = G_RHO; *rB = P_RHO; break;
        case IBIS_THERMO_RHO_T: *sA = G_RHO; *rA = P_RHO; *sB = G_T; *rB = P_T; break;
        case IBIS_THERMO_RHO_U: *sA = G_RHO; *rA = P_RHO; *sB = G_U; *rB = P_E; break;
        default: *sA = G_P; *rA = P_P; *sB = G_T; *rB = P_T; break;
    }
}

}  // namespace

static hipEvent_t take_event(ibis_residual* r) {
    if (!r->pool.empty()) {
        hipEvent_t e = r->pool.back();
        r->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

namespace {

// per-kernel event timing of one residual call
struct Prof {
    ibis_residual* r;
    hipStream_t st;
    hipEvent_t a = nullptr;
    void beg() {
        if (!r->profiling) return;
        a = take_event(r);
        (void)hipEventRecord(a, st);
    }
    void end(int slot) {
        if (!r->profiling) return;
        hipEvent_t b = take_event(r);
        (void)hipEventRecord(b, st);
        r->pending.push_back({slot, a, b});
    }
};

template <class S>
struct GenBufs {
    S* grad;
    S* lim;
    S* flux;
    double* face_rho;  // viscous face states for estimate_dt (null: not kept)
    double* face_T;
};

void residual_params(const ibis_residual* r, FaceParams* fp, VarSet* vs) {
    const ibis_fv_desc& d = r->desc;
    fp->thermo = d.thermo;
    fp->limited = d.limiter != IBIS_LIM_UNLIMITED;
    thermo_pair(d.thermo, &fp->slotA, &fp->rowA, &fp->slotB, &fp->rowB);
    fp->R = d.R;
    fp->Cv = d.Cv;
    fp->gamma = d.gamma;
    fp->delta = d.ldfss_delta;
    vs->n = 5;
    vs->ngrad = r->d.dim == 3 ? 5 : 4;
    vs->row[0] = fp->rowA;
    vs->slot[0] = fp->slotA;
    vs->row[1] = fp->rowB;
    vs->slot[1] = fp->slotB;
    vs->row[2] = P_VX;
    vs->slot[2] = G_VX;
    vs->row[3] = P_VY;
    vs->slot[3] = G_VY;
    vs->row[4] = P_VZ;
    vs->slot[4] = G_VZ;
}

// The residual as the reference's sequence of passes, on scalar S: ghost
// fill, gradients + limiters, face flux, flux override, viscous ghost fill,
// viscous gradients and flux, surface integral (FV/finite_volume.cpp:57-96).
template <class S>
void run_general(ibis_residual* r, const GenBufs<S>& B, S* prim, S* dudt, bool recon,
                 hipStream_t st, Prof& pf) {
    const DevMesh& D = r->d;
    const ibis_fv_desc& d = r->desc;
    const size_t ntab = r->ntab;
    const GasParams gp{d.R, d.Cv, d.Cp, d.gamma};
    FaceParams fp;
    VarSet vs;
    residual_params(r, &fp, &vs);
    const bool has_post = r->has_post;
    if (r->has_pre) {
        pf.beg();
        k_bc<S><<<nblocks(D.nb), BLOCK, 0, st>>>(D, r->kinds, r->pars, gp, prim);
        pf.end(0);
    }
    if (recon) {
        pf.beg();
        if (D.dim == 3)
            launch_cell_grad<S, 3>(D, vs, d.limiter, prim, B.grad, B.lim, d.lim_eps, d.venkat_K, st);
        else
            launch_cell_grad<S, 2>(D, vs, d.limiter, prim, B.grad, B.lim, d.lim_eps, d.venkat_K, st);
        if constexpr (std::is_same<S, ibis::Dual>::value) {
            if (r->freeze_lim && d.limiter != IBIS_LIM_UNLIMITED) {
                const long nl = (long)NGRAD * D.N;
                k_freeze_tangent<<<nblocks(nl), BLOCK, 0, st>>>(nl, B.lim);
            }
        }
        pf.end(1);
    }
    pf.beg();
    if (D.dim == 3)
        launch_face_flux<S, 3>(D, d.flux, fp, recon, r->premul, prim, B.grad, B.lim, B.flux, st);
    else
        launch_face_flux<S, 2>(D, d.flux, fp, recon, r->premul, prim, B.grad, B.lim, B.flux, st);
    pf.end(2);
    if (has_post) {
        pf.beg();
        k_post_flux<S><<<nblocks(D.nb), BLOCK, 0, st>>>(D, r->kinds + 2 * ntab,
                                                        r->pars + 2 * ntab * IBIS_NPAR, gp,
                                                        r->premul ? 1 : 0, B.flux);
        pf.end(3);
    }
    if (d.viscous) {
        if (r->has_visc) {
            pf.beg();
            k_bc<S><<<nblocks(D.nb), BLOCK, 0, st>>>(D, r->kinds + ntab,
                                                     r->pars + ntab * IBIS_NPAR, gp, prim);
            pf.end(4);
        }
        VarSet vv;
        vv.n = 4;
        vv.ngrad = 4;
        const int rows[5] = {P_T, P_VX, P_VY, P_VZ, P_VZ};
        const int slots[5] = {G_T, G_VX, G_VY, G_VZ, G_VZ};
        for (int q = 0; q < 5; q++) {
            vv.row[q] = rows[q];
            vv.slot[q] = slots[q];
        }
        ViscParams vp{d.R, d.Cv, d.Cp, d.mu0, d.T0, d.Ts, d.Pr};
        pf.beg();
        if (D.dim == 3)
            launch_cell_grad<S, 3>(D, vv, 0, prim, B.grad, B.lim, 0.0, 0.0, st);
        else
            launch_cell_grad<S, 2>(D, vv, 0, prim, B.grad, B.lim, 0.0, 0.0, st);
        pf.end(5);
        pf.beg();
        if (D.dim == 3)
            k_visc_flux<S, 3><<<nblocks(D.F), BLOCK, 0, st>>>(D, vp, prim, B.grad, B.flux,
                                                              B.face_rho, B.face_T);
        else
            k_visc_flux<S, 2><<<nblocks(D.F), BLOCK, 0, st>>>(D, vp, prim, B.grad, B.flux,
                                                              B.face_rho, B.face_T);
        pf.end(6);
    }
    const unsigned nbc = nblocks(D.N);
    pf.beg();
    if (D.dim == 3) {
        if (r->premul)
            k_integral<S, 3, true><<<nbc, BLOCK, 0, st>>>(D, B.flux, dudt);
        else
            k_integral<S, 3, false><<<nbc, BLOCK, 0, st>>>(D, B.flux, dudt);
    } else {
        if (r->premul)
            k_integral<S, 2, true><<<nbc, BLOCK, 0, st>>>(D, B.flux, dudt);
        else
            k_integral<S, 2, false><<<nbc, BLOCK, 0, st>>>(D, B.flux, dudt);
    }
    pf.end(7);
}

// the residual; with epi (only when stage_fusable) the face kernel runs
// the Runge-Kutta stage epilogue and dudt may be null
ibis_status compute_dudt_impl(ibis_residual* r, double* prim, double* dudt,
                              int allow_reconstruction, ibis_stream_t stream,
                              const ibis_stage_epilogue* epi) {
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const DevMesh& D = r->d;
    const ibis_fv_desc& d = r->desc;
    const bool recon = allow_reconstruction && d.order == 2;
    Prof pf{r, st};
    FaceParams fp;
    VarSet vs;
    residual_params(r, &fp, &vs);
    const bool fused = !d.viscous && !r->has_post;
    const bool visc_fused = recon && !epi && visc_fusable(r, vs);
    if (r->prec != IBIS_PREC_FP64 && !(fused && recon) &&
        !(visc_fused && r->prec == IBIS_PREC_FP32_FLUX))
        return ibis::fail(IBIS_ERR_INVALID,
                          "the reduced-precision modes need the second-order residual without "
                          "flux-override boundaries (FP32 reconstruction: inviscid only; "
                          "viscous: with the options of the fused viscous path)");
    // viscous, second order: boundary actions (both passes), cell kernel
    // (viscous gradients, convective gradients, limiter, face values), face
    // kernel (convective and viscous flux, surface integral)
    if (visc_fused) {
        const bool f32 = r->prec == IBIS_PREC_FP32_FLUX;
        const GasParams gp{d.R, d.Cv, d.Cp, d.gamma};
        // option fold_boundary (FP64): both boundary passes run in the cell
        // kernel's staging of each ghost
        const bool fold = !f32 && r->opt_fold;
        const BcFold bf{r->kinds, r->pars, r->ntab, r->has_pre ? 1 : 0, r->has_visc ? 1 : 0,
                        gp, prim, r->FL, nullptr};
        pf.beg();
        if (D.nb > 0 && f32)
            k_bc_visc_fs<float><<<nblocks(D.nb), BLOCK, 0, st>>>(
                D, r->kinds, r->pars, r->ntab, r->has_pre ? 1 : 0, r->has_visc ? 1 : 0, gp, vs,
                prim, reinterpret_cast<float*>(r->FL), r->gconv, nullptr);
        else if (D.nb > 0 && !fold)
            k_bc_visc_fs<double><<<nblocks(D.nb), BLOCK, 0, st>>>(
                D, r->kinds, r->pars, r->ntab, r->has_pre ? 1 : 0, r->has_visc ? 1 : 0, gp, vs,
                prim, r->FL, r->gconv, nullptr);
        pf.end(0);
        pf.beg();
        if (D.dim == 3)
            launch_cell_visc<3>(D, vs, d, r, prim, r->grad, r->opt_keep_grads, st,
                                fold ? &bf : nullptr);
        else
            launch_cell_visc<2>(D, vs, d, r, prim, r->grad, r->opt_keep_grads, st,
                                fold ? &bf : nullptr);
        pf.end(1);
        pf.beg();
        const ViscParams vp{d.R, d.Cv, d.Cp, d.mu0, d.T0, d.Ts, d.Pr};
        if (D.dim == 3)
            k_visc_terms<3><<<nblocks(D.F), BLOCK, 0, st>>>(D, vp, prim, r->grad, r->vterm,
                                                           r->face_rho, r->face_T);
        else
            k_visc_terms<2><<<nblocks(D.F), BLOCK, 0, st>>>(D, vp, prim, r->grad, r->vterm,
                                                           r->face_rho, r->face_T);
        pf.end(6);
        pf.beg();
        const ViscFace vf{r->vterm};
        if (D.dim == 3)
            launch_face_visc<3>(D, fp, d.flux, r->ent_idx, r->ent_geom, r->ent_geom32, r->FL, f32,
                                dudt, vf, st);
        else
            launch_face_visc<2>(D, fp, d.flux, r->ent_idx, r->ent_geom, r->ent_geom32, r->FL, f32,
                                dudt, vf, st);
        pf.end(2);
        HIP_TRY(hipGetLastError());
        return IBIS_OK;
    }
    if (fused) {
        const GasParams gp{d.R, d.Cv, d.Cp, d.gamma};
        if (r->has_pre && !recon) {
            pf.beg();
            k_bc<double><<<nblocks(D.nb), BLOCK, 0, st>>>(D, r->kinds, r->pars, gp, prim);
            pf.end(0);
        }
        if (recon) {
            // face-state path: boundary actions + ghost face states, pack,
            // cell reconstruction, fused face flux + integral; with the option
            // fold_boundary the boundary pass runs in the cell kernel's staging
            const bool fold = fold_on(r);
            const BcFold bf{r->kinds, r->pars, r->ntab, r->has_pre ? 1 : 0, 0, gp, prim, r->FL,
                            nullptr};
            pf.beg();
            if (D.nb > 0 && !fold) {
                const unsigned g = nblocks(D.nb);
                const int* kd = r->kinds;
                const double* pr = r->pars;
                if (r->prec != IBIS_PREC_FP64) {
                    float* fl = reinterpret_cast<float*>(r->FL);
                    if (r->has_pre)
                        k_bc_ghost_fs<2, float><<<g, BLOCK, 0, st>>>(D, kd, pr, gp, vs, prim, fl, fl);
                    else
                        k_ghost_fs<2, float><<<g, BLOCK, 0, st>>>(D, vs, prim, fl, fl);
                } else if (r->opt_layout == 0) {
                    if (r->has_pre)
                        k_bc_ghost_fs<0><<<g, BLOCK, 0, st>>>(D, kd, pr, gp, vs, prim, r->FL, r->FR);
                    else
                        k_ghost_fs<0><<<g, BLOCK, 0, st>>>(D, vs, prim, r->FL, r->FR);
                } else if (r->opt_layout == 1) {
                    if (r->has_pre)
                        k_bc_ghost_fs<1><<<g, BLOCK, 0, st>>>(D, kd, pr, gp, vs, prim, r->FL, r->FR);
                    else
                        k_ghost_fs<1><<<g, BLOCK, 0, st>>>(D, vs, prim, r->FL, r->FR);
                } else {
                    if (r->has_pre)
                        k_bc_ghost_fs<2><<<g, BLOCK, 0, st>>>(D, kd, pr, gp, vs, prim, r->FL, r->FR);
                    else
                        k_ghost_fs<2><<<g, BLOCK, 0, st>>>(D, vs, prim, r->FL, r->FR);
                }
            }
            const bool packed =
                r->prec == IBIS_PREC_FP64 && (r->opt_pack || r->opt_layout != 2);
            if (packed) k_pack<<<nblocks(D.Nt), BLOCK, 0, st>>>(D.Nt, vs, prim, r->pack);
            pf.end(0);
            const int nlog = D.ntiles;  // one block per tile
            pf.beg();
            if (D.dim == 3)
                launch_cell_fs<3>(D, vs, d, r, nlog, prim, st, fold ? &bf : nullptr);
            else
                launch_cell_fs<2>(D, vs, d, r, nlog, prim, st, fold ? &bf : nullptr);
            pf.end(1);
            pf.beg();
            const bool f32 = r->prec != IBIS_PREC_FP64;
            const int lay = f32 ? 2 : r->opt_layout;
            const int fk = lay == 2 ? r->opt_face_kernel : 0;
            if (epi && D.dim == 3)
                launch_face_epi<3>(D, fp, d.flux, r->ent_idx, r->ent_geom, r->FL, dudt, *epi, st);
            else if (epi)
                launch_face_epi<2>(D, fp, d.flux, r->ent_idx, r->ent_geom, r->FL, dudt, *epi, st);
            else if (D.dim == 3)
                launch_face_fs<3>(D, fp, d.flux, lay, f32, fk, r->ent_idx, r->ent_geom,
                                  r->ent_geom32, r->FL, r->FR, dudt, st);
            else
                launch_face_fs<2>(D, fp, d.flux, lay, f32, fk, r->ent_idx, r->ent_geom,
                                  r->ent_geom32, r->FL, r->FR, dudt, st);
            pf.end(2);
        } else {
            // first order: face flux and surface integral as one tiled kernel
            pf.beg();
            if (D.dim == 3)
                launch_face_tile<3>(D, d.flux, fp, false, prim, r->grad, r->lim, dudt, st);
            else
                launch_face_tile<2>(D, d.flux, fp, false, prim, r->grad, r->lim, dudt, st);
            pf.end(2);
        }
        HIP_TRY(hipGetLastError());
        return IBIS_OK;
    }
    GenBufs<double> B{r->grad, r->lim, r->flux, r->face_rho, r->face_T};
    run_general<double>(r, B, prim, dudt, recon, st, pf);
    HIP_TRY(hipGetLastError());
    return IBIS_OK;
}

}  // namespace

ibis_status ibis_compute_dudt(ibis_residual* r, double* prim, double* dudt,
                              int allow_reconstruction, ibis_stream_t stream) {
    if (!r || !prim || !dudt) return ibis::fail(IBIS_ERR_INVALID, "ibis_compute_dudt: null argument");
    return compute_dudt_impl(r, prim, dudt, allow_reconstruction, stream, nullptr);
}

int ibis_residual_stage_fusable(const ibis_residual* r) {
    if (!r) return 0;
    const DevMesh& D = r->d;
    const ibis_fv_desc& d = r->desc;
    return (r->prec == IBIS_PREC_FP64 && !d.viscous && !r->has_post && d.order == 2 &&
            r->opt_layout == 2 && r->opt_face_kernel == 1 && r->ent_idx && D.n_from_t)
               ? 1
               : 0;
}

int ibis_residual_stage_dt_fusable(const ibis_residual* r) {
    if (!ibis_residual_stage_fusable(r)) return 0;
    const DevMesh& D = r->d;
    return (size_t)(D.dim + 6) * D.maxtf * sizeof(double) <= MAX_LDS_BYTES ? 1 : 0;
}

ibis_status ibis_compute_dudt_stage(ibis_residual* r, double* prim, double* dudt,
                                    const ibis_stage_epilogue* epi, ibis_stream_t stream) {
    if (!r || !prim || !epi || !epi->u || !epi->inst_of_cell || !epi->dt || !epi->prim_out ||
        epi->nk < 0 || epi->nk > IBIS_EPI_MAXK)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_compute_dudt_stage: bad argument");
    for (int j = 0; j < epi->nk; j++)
        if (!epi->k[j]) return ibis::fail(IBIS_ERR_INVALID, "ibis_compute_dudt_stage: null stage");
    if (!ibis_residual_stage_fusable(r))
        return ibis::fail(IBIS_ERR_INVALID,
                          "ibis_compute_dudt_stage: the residual's path has no stage epilogue");
    if (epi->dt_cells_out && !ibis_residual_stage_dt_fusable(r))
        return ibis::fail(IBIS_ERR_INVALID,
                          "ibis_compute_dudt_stage: the residual's epilogue cannot write time steps");
    // the time-step constants are the residual's own (as ibis_residual_cell_dt)
    ibis_stage_epilogue e = *epi;
    const ibis_fv_desc& d = r->desc;
    e.dt_gamma = d.gamma;
    e.dt_R = d.R;
    e.dt_sf = d.signal_factor;
    const DevMesh& D = r->d;
    // at least a tile's worth of rows per tile (one per thread): the first
    // few tiles take them all and the others pay nothing on their way
    e.ghost_per_tile =
        e.ghost_rows && D.Nt > D.N ? std::max(D.tile, (D.Nt - D.N + D.ntiles - 1) / D.ntiles) : 0;
    return compute_dudt_impl(r, prim, dudt, 1, stream, &e);
}

// tangent state and scratch, allocated (and zeroed) on first use
static hipError_t ensure_dual(ibis_residual* r) {
    if (r->dual_prim) return hipSuccess;
    const DevMesh& D = r->d;
    const size_t N = D.N, F = D.F, Nt = D.Nt;
    const int nc = D.dim + 2;
    const size_t sz = sizeof(ibis::Dual);
    ibis::Dual** ptrs[5] = {&r->dual_prim, &r->dual_dudt, &r->dual_grad, &r->dual_lim,
                            &r->dual_flux};
    const size_t bytes[5] = {sz * NPRIM * Nt, sz * nc * Nt, sz * NGRAD * 3 * N, sz * NGRAD * N,
                             sz * nc * F};
    for (int q = 0; q < 5; q++) {
        const size_t b = std::max<size_t>(bytes[q], 16);
        hipError_t e = hipMalloc((void**)ptrs[q], b);
        if (e == hipSuccess) e = hipMemset(*ptrs[q], 0, b);
        if (e != hipSuccess) {
            for (int k = 0; k <= q; k++) {
                if (*ptrs[k]) (void)hipFree(*ptrs[k]);
                *ptrs[k] = nullptr;
            }
            return e;
        }
    }
    return hipSuccess;
}

// Tangent of the residual: the same passes on (value, tangent) pairs with
// the reference's Dual arithmetic (UT/dual.h), replacing FiniteVolume<dual>.
ibis_status ibis_compute_dudt_jvp(ibis_residual* r, double* prim, double* prim_dot, double* dudt,
                                  double* dudt_dot, int allow_reconstruction,
                                  ibis_stream_t stream) {
    if (!r || !prim || !prim_dot || !dudt || !dudt_dot)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_compute_dudt_jvp: null argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const DevMesh& D = r->d;
    const ibis_fv_desc& d = r->desc;
    HIP_TRY(ensure_dual(r));
    const long np = (long)NPRIM * D.Nt;
    k_to_dual<<<nblocks(np), BLOCK, 0, st>>>(np, prim, prim_dot, r->dual_prim);
    const bool recon = allow_reconstruction && d.order == 2;
    Prof pf{r, st};
    GenBufs<ibis::Dual> B{r->dual_grad, r->dual_lim, r->dual_flux, nullptr, nullptr};
    run_general<ibis::Dual>(r, B, r->dual_prim, r->dual_dudt, recon, st, pf);
    // ghost rows of the state are written by the boundary actions, as in the
    // reference; the residual rows [0, N) of dudt
    k_from_dual<<<nblocks(D.Nt), BLOCK, 0, st>>>(NPRIM, D.Nt, D.Nt, r->dual_prim, prim, prim_dot);
    k_from_dual<<<nblocks(D.N), BLOCK, 0, st>>>(D.dim + 2, D.N, D.Nt, r->dual_dudt, dudt,
                                                dudt_dot);
    HIP_TRY(hipGetLastError());
    return IBIS_OK;
}

ibis_status ibis_steady_matvec(ibis_residual* r, const double* cons, const double* vec,
                               double* result, double dt_star, int allow_reconstruction,
                               ibis_stream_t stream) {
    if (!r || !cons || !vec || !result)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_steady_matvec: null argument");
    if (!(dt_star > 0.0)) return ibis::fail(IBIS_ERR_INVALID, "ibis_steady_matvec: dt_star <= 0");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const DevMesh& D = r->d;
    const ibis_fv_desc& d = r->desc;
    HIP_TRY(ensure_dual(r));
    const int nc = D.dim + 2;
    // the conserved tangent goes through dual_dudt, which the residual
    // overwrites only after the conversion has consumed it
    k_seed_cons<<<nblocks(D.Nt), BLOCK, 0, st>>>(D.N, D.Nt, nc, cons, vec, r->dual_dudt);
    if (D.dim == 3)
        k_cons_to_prim<ibis::Dual, 3><<<nblocks(D.Nt), BLOCK, 0, st>>>(D.Nt, d.R, d.Cv,
                                                                        r->dual_dudt, r->dual_prim);
    else
        k_cons_to_prim<ibis::Dual, 2><<<nblocks(D.Nt), BLOCK, 0, st>>>(D.Nt, d.R, d.Cv,
                                                                        r->dual_dudt, r->dual_prim);
    const bool recon = allow_reconstruction && d.order == 2;
    Prof pf{r, st};
    GenBufs<ibis::Dual> B{r->dual_grad, r->dual_lim, r->dual_flux, nullptr, nullptr};
    run_general<ibis::Dual>(r, B, r->dual_prim, r->dual_dudt, recon, st, pf);
    k_jv_result<<<nblocks(D.N), BLOCK, 0, st>>>(D.N, D.Nt, nc, dt_star, vec, r->dual_dudt, result);
    HIP_TRY(hipGetLastError());
    return IBIS_OK;
}

ibis_status ibis_steady_rhs(ibis_residual* r, double* prim, double* dudt, double* rhs,
                            int allow_reconstruction, ibis_stream_t stream) {
    if (!r || !prim || !dudt || !rhs) return ibis::fail(IBIS_ERR_INVALID, "ibis_steady_rhs: null argument");
    ibis_status s = ibis_compute_dudt(r, prim, dudt, allow_reconstruction, stream);
    if (s != IBIS_OK) return s;
    const DevMesh& D = r->d;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    k_rhs_vector<<<nblocks(D.N), BLOCK, 0, st>>>(D.N, D.Nt, D.dim + 2, dudt, rhs);
    HIP_TRY(hipGetLastError());
    return IBIS_OK;
}

ibis_status ibis_cons_to_prim_jvp(ibis_residual* r, const double* cons, const double* cons_dot,
                                  double* prim, double* prim_dot, ibis_stream_t stream) {
    if (!r || !cons || !cons_dot || !prim || !prim_dot)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_cons_to_prim_jvp: null argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const DevMesh& D = r->d;
    HIP_TRY(ensure_dual(r));
    const int nc = D.dim + 2;
    const long ncv = (long)nc * D.Nt;
    k_to_dual<<<nblocks(ncv), BLOCK, 0, st>>>(ncv, cons, cons_dot, r->dual_dudt);
    if (D.dim == 3)
        k_cons_to_prim<ibis::Dual, 3><<<nblocks(D.Nt), BLOCK, 0, st>>>(D.Nt, r->desc.R, r->desc.Cv,
                                                                        r->dual_dudt, r->dual_prim);
    else
        k_cons_to_prim<ibis::Dual, 2><<<nblocks(D.Nt), BLOCK, 0, st>>>(D.Nt, r->desc.R, r->desc.Cv,
                                                                        r->dual_dudt, r->dual_prim);
    k_from_dual<<<nblocks(D.Nt), BLOCK, 0, st>>>(NPRIM, D.Nt, D.Nt, r->dual_prim, prim, prim_dot);
    HIP_TRY(hipGetLastError());
    return IBIS_OK;
}

ibis_status ibis_residual_set_precision(ibis_residual* r, int mode) {
    if (!r) return ibis::fail(IBIS_ERR_INVALID, "ibis_residual_set_precision: null residual");
    if (mode != IBIS_PREC_FP64 && mode != IBIS_PREC_FP32_FLUX && mode != IBIS_PREC_FP32_RECON)
        return ibis::fail(IBIS_ERR_INVALID, "Unknown precision mode");
    r->generation++;
    const DevMesh& D = r->d;
    if (mode != IBIS_PREC_FP64 && r->ent_idx && !r->ent_geom32 && D.ntf > 0 && D.n_from_t) {
        HIP_TRY(hipMalloc(&r->ent_geom32, (size_t)32 * D.ntf));
        k_entries32<<<nblocks(D.ntf), BLOCK>>>(D, reinterpret_cast<char*>(r->ent_geom32));
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipDeviceSynchronize());
    }
    if (mode == IBIS_PREC_FP32_RECON && r->offs && !r->offs32) {
        const size_t n = (size_t)3 * D.maxf * D.N;
        HIP_TRY(hipMalloc(&r->offs32, sizeof(float) * n));
        k_to_float<<<(unsigned)((n + BLOCK - 1) / BLOCK), BLOCK>>>(n, r->offs, r->offs32);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipDeviceSynchronize());
    }
    r->prec = mode;
    return IBIS_OK;
}

ibis_status ibis_residual_sizes(const ibis_residual* r, int64_t* sizes) {
    if (!r || !sizes) return ibis::fail(IBIS_ERR_INVALID, "ibis_residual_sizes: null argument");
    const DevMesh& D = r->d;
    sizes[0] = D.dim;
    sizes[1] = D.N;
    sizes[2] = D.Nt;
    sizes[3] = D.F;
    sizes[4] = D.G;
    return IBIS_OK;
}

// The per-entry face data (face_kernel 1) of the tile lists in r->d, and its
// float copies when the residual has them (reduced-precision modes).
static ibis_status build_entries(ibis_residual* r) {
    const DevMesh& D = r->d;
    const bool had32 = r->ent_geom32 != nullptr;
    for (void** p : {(void**)&r->ent_idx, (void**)&r->ent_geom, (void**)&r->ent_geom32})
        if (*p) {
            HIP_TRY(hipFree(*p));
            *p = nullptr;
        }
    const size_t ne = std::max(D.ntf, 1);
    // entry geometry: 1 = with normals, 2 = planar 2D, 0 = t1, t2, area
    const int with_n = !D.n_from_t ? 1 : (D.dim == 2 && D.planar2d ? 2 : 0);
    HIP_TRY(hipMalloc(&r->ent_idx, sizeof(int2) * ne));
    HIP_TRY(hipMalloc(&r->ent_geom, sizeof(double) * (with_n == 1 ? 10 : 7) * ne));
    if (D.ntf > 0) k_entries<<<nblocks(D.ntf), BLOCK>>>(D, with_n, r->ent_idx, r->ent_geom);
    HIP_TRY(hipGetLastError());
    if (had32 && D.ntf > 0 && D.n_from_t) {
        HIP_TRY(hipMalloc(&r->ent_geom32, (size_t)32 * D.ntf));
        k_entries32<<<nblocks(D.ntf), BLOCK>>>(D, reinterpret_cast<char*>(r->ent_geom32));
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipDeviceSynchronize());
    return IBIS_OK;
}

ibis_status ibis_residual_set_option(ibis_residual* r, const char* key, int64_t value) {
    if (!r || !key) return ibis::fail(IBIS_ERR_INVALID, "ibis_residual_set_option: null argument");
    const std::string k(key);
    r->generation++;
    if (k == "face_state_layout") {
        if (value < 0 || value > 2)
            return ibis::fail(IBIS_ERR_INVALID, "face_state_layout must be 0, 1 or 2");
        r->opt_layout = (int)value;
        return IBIS_OK;
    }
    if (k == "precomputed_offsets") {
        if (value && !r->offs) {
            const DevMesh& D = r->d;
            HIP_TRY(hipMalloc(&r->offs, sizeof(double) * 3 * (size_t)D.maxf * D.N));
            k_offsets<<<nblocks(D.N), BLOCK>>>(D, r->offs);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipDeviceSynchronize());
        }
        r->opt_offsets = value ? 1 : 0;
        return IBIS_OK;
    }
    if (k == "face_kernel") {
        if (value != 0 && value != 1 && value != 64 && value != 128 && value != 256)
            return ibis::fail(IBIS_ERR_INVALID, "face_kernel must be 0, 1, 64, 128 or 256");
        const DevMesh& D = r->d;
        if (value > 1 && (size_t)(D.dim + 2) * D.maxf * value * sizeof(double) > 65536)
            return ibis::fail(IBIS_ERR_INVALID, "face_kernel: block does not fit in LDS");
        if (value == 1 && !r->ent_idx) {
            if (!ibis_entry_indices_fit(D.maxf, D.N, D.G))
                return ibis::fail(IBIS_ERR_INVALID, "face_kernel 1: mesh too large for 32-bit value indices");
            IBIS_TRY(build_entries(r));
        }
        r->opt_face_kernel = (int)value;
        return IBIS_OK;
    }
    if (k == "tile") {
        if (value != 128 && value != 256 && value != 512)
            return ibis::fail(IBIS_ERR_INVALID, "tile must be 128, 256 or 512");
        if (value == r->d.tile) return IBIS_OK;
        DevMesh D = r->d;
        IBIS_TRY(mesh_tiles(r->m, (int)value, &D));
        // the tile kernels keep (dim+2) doubles per tile-list entry in LDS
        if ((size_t)(D.dim + 2) * D.maxtf * sizeof(double) > MAX_LDS_BYTES)
            return ibis::fail(IBIS_ERR_INVALID, "tile: a tile's face list does not fit in LDS");
        r->d = D;
        if (r->ent_idx) IBIS_TRY(build_entries(r));
        return IBIS_OK;
    }
    if (k == "pack_cells") {
        r->opt_pack = value ? 1 : 0;
        return IBIS_OK;
    }
    if (k == "freeze_limiter") {
        r->freeze_lim = value ? 1 : 0;
        return IBIS_OK;
    }
    if (k == "stage_offsets") {
        r->opt_stage = value ? 1 : 0;
        return IBIS_OK;
    }
    if (k == "cell_prefetch") {
        r->opt_cell_prefetch = value ? 1 : 0;
        return IBIS_OK;
    }
    if (k == "stage_cells") {
        r->opt_stage_cells = value ? 1 : 0;
        return IBIS_OK;
    }
    if (k == "keep_gradients") {
        r->opt_keep_grads = value ? 1 : 0;
        return IBIS_OK;
    }
    if (k == "fuse_viscous") {
        r->opt_fuse_visc = value ? 1 : 0;
        return IBIS_OK;
    }
    if (k == "fold_boundary") {
        r->opt_fold = value ? 1 : 0;
        return IBIS_OK;
    }
    return ibis::fail(IBIS_ERR_INVALID, ("unknown option " + k).c_str());
}

// ---------------------------------------------------------------------------
// Adjoint (vector-Jacobian product) of the residual
// ---------------------------------------------------------------------------
namespace {

ibis_status ensure_adjoint(ibis_residual* r) {
    if (r->adj_prim_c) return IBIS_OK;
    const DevMesh& D = r->d;
    const size_t N = D.N, Nt = D.Nt, F = D.F, nb = std::max(D.nb, 1);
    const int nc = D.dim + 2;
    auto al = [&](void** p, size_t bytes) -> hipError_t {
        hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 8));
        if (e == hipSuccess) e = hipMemset(*p, 0, std::max<size_t>(bytes, 8));
        if (e == hipSuccess) r->adj.push_back(*p);
        return e;
    };
    HIP_TRY(al((void**)&r->adj_prim_c, sizeof(double) * NPRIM * Nt));
    HIP_TRY(al((void**)&r->adj_vgrad, sizeof(double) * NGRAD * 3 * N));
    // face and cell cotangents: 7 rows each for the reconstructed values
    // (5 used by the second-order path, 7 prim rows by the first-order one)
    HIP_TRY(al((void**)&r->adj_face, sizeof(double) * (7 + 7 + 4 + 4 + 12) * F));
    HIP_TRY(al((void**)&r->adj_cell, sizeof(double) * (7 + 4 + 9 * (size_t)D.maxf) * N));
    HIP_TRY(al((void**)&r->adj_bcb, sizeof(double) * NPRIM * nb));
    HIP_TRY(al((void**)&r->adj_zero, sizeof(double) * nc * Nt));
    HIP_TRY(al((void**)&r->adj_dudt, sizeof(double) * nc * Nt));
    HIP_TRY(al((void**)&r->adj_force, sizeof(double) * 3 * nb));
    // geometry tables of the reverse kernels: LSQ weights per (cell, slot)
    // and the face offsets (the same expressions as the forward kernels)
    HIP_TRY(al((void**)&r->adj_lsqw, sizeof(double) * 3 * (size_t)D.maxf * N));
    if (D.dim == 3)
        k_lsq_table<3><<<nblocks(D.N), BLOCK>>>(D, r->adj_lsqw);
    else
        k_lsq_table<2><<<nblocks(D.N), BLOCK>>>(D, r->adj_lsqw);
    if (!r->offs) {
        HIP_TRY(hipMalloc(&r->offs, sizeof(double) * 3 * (size_t)D.maxf * D.N));
        k_offsets<<<nblocks(D.N), BLOCK>>>(D, r->offs);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    return IBIS_OK;
}

ibis_status adjoint_supported(const ibis_residual* r) {
    const ibis_fv_desc& d = r->desc;
    if (d.order != 1 && d.order != 2)
        return ibis::fail(IBIS_ERR_INVALID, "the adjoint needs the first- or second-order residual");
    // LDFSS2: the reference's own tangent of it (FiniteVolume<Dual>) is NaN
    // on subsonic faces (the Dual pow of a negative base, dual.h:239-245),
    // so there is no reference derivative for a reverse mode to match
    if (d.flux == IBIS_FLUX_LDFSS2)
        return ibis::fail(IBIS_ERR_INVALID, "the adjoint supports the AUSMDV, Hanel and Rusanov fluxes");
    if (r->prec != IBIS_PREC_FP64)
        return ibis::fail(IBIS_ERR_INVALID, "the adjoint is FP64 only");
    return IBIS_OK;
}

template <int DIM>
void launch_vjp(ibis_residual* r, const VarSet& vs, const FaceParams& fp, const AdjBufs& B,
                bool active, hipStream_t st) {
    const DevMesh& D = r->d;
    const ibis_fv_desc& d = r->desc;
    const bool visc = d.viscous != 0;
    const bool ord1 = d.order == 1;
    const unsigned nf = nblocks(D.F), nc = nblocks(D.N);
    const GasParams gp{d.R, d.Cv, d.Cp, d.gamma};
    if (ord1) {
        switch (d.flux) {
            case IBIS_FLUX_AUSMDV: k_vjp_conv_face1<DIM, IBIS_FLUX_AUSMDV><<<nf, BLOCK, 0, st>>>(D, fp, B); break;
            case IBIS_FLUX_HANEL: k_vjp_conv_face1<DIM, IBIS_FLUX_HANEL><<<nf, BLOCK, 0, st>>>(D, fp, B); break;
            default: k_vjp_conv_face1<DIM, IBIS_FLUX_RUSANOV><<<nf, BLOCK, 0, st>>>(D, fp, B); break;
        }
    }
    ViscParams vp{d.R, d.Cv, d.Cp, d.mu0, d.T0, d.Ts, d.Pr};
    if (!ord1 && visc) {  // both face reverses in one kernel
        switch (d.flux) {
            case IBIS_FLUX_AUSMDV: k_vjp_face_cv<DIM, IBIS_FLUX_AUSMDV><<<nf, BLOCK, 0, st>>>(D, fp, vs, vp, B); break;
            case IBIS_FLUX_HANEL: k_vjp_face_cv<DIM, IBIS_FLUX_HANEL><<<nf, BLOCK, 0, st>>>(D, fp, vs, vp, B); break;
            default: k_vjp_face_cv<DIM, IBIS_FLUX_RUSANOV><<<nf, BLOCK, 0, st>>>(D, fp, vs, vp, B); break;
        }
    } else if (!ord1) {
        switch (d.flux) {
            case IBIS_FLUX_AUSMDV: k_vjp_conv_face<DIM, IBIS_FLUX_AUSMDV><<<nf, BLOCK, 0, st>>>(D, fp, vs, B); break;
            case IBIS_FLUX_HANEL: k_vjp_conv_face<DIM, IBIS_FLUX_HANEL><<<nf, BLOCK, 0, st>>>(D, fp, vs, B); break;
            default: k_vjp_conv_face<DIM, IBIS_FLUX_RUSANOV><<<nf, BLOCK, 0, st>>>(D, fp, vs, B); break;
        }
    } else if (visc) {
        k_vjp_visc_face<DIM><<<nf, BLOCK, 0, st>>>(D, vp, B);
    }
    // LDS of k_vjp_cell: maxf x 3 doubles and 2 x maxf ints per thread
    const size_t vshm = (size_t)D.maxf * BLOCK * (3 * sizeof(double) + 2 * sizeof(int));
#define CELL(LIM, ACT, VIS) \
    k_vjp_cell<DIM, LIM, ACT, VIS><<<nc, BLOCK, vshm, st>>>(D, vs, d.lim_eps, d.venkat_K, B)
    const int lim = d.limiter;
    if (ord1) {
        if (visc)
            k_vjp_cell<DIM, 0, false, true, true><<<nc, BLOCK, vshm, st>>>(D, vs, d.lim_eps, d.venkat_K, B);
        else
            k_vjp_cell<DIM, 0, false, false, true><<<nc, BLOCK, vshm, st>>>(D, vs, d.lim_eps, d.venkat_K, B);
    } else if (visc) {
        if (lim == 0) CELL(0, false, true);
        else if (lim == 1) { if (active) CELL(1, true, true); else CELL(1, false, true); }
        else { if (active) CELL(2, true, true); else CELL(2, false, true); }
    } else {
        if (lim == 0) CELL(0, false, false);
        else if (lim == 1) { if (active) CELL(1, true, false); else CELL(1, false, false); }
        else { if (active) CELL(2, true, false); else CELL(2, false, false); }
    }
#undef CELL
    if (D.nb > 0) {
        const unsigned ng = nblocks(D.nb);
        const int nt = r->ntab;
        if (visc && ord1)
            k_vjp_ghost<DIM, true, true><<<ng, BLOCK, 0, st>>>(D, vs, r->kinds, r->pars, nt, gp, B);
        else if (visc)
            k_vjp_ghost<DIM, true><<<ng, BLOCK, 0, st>>>(D, vs, r->kinds, r->pars, nt, gp, B);
        else if (ord1)
            k_vjp_ghost<DIM, false, true><<<ng, BLOCK, 0, st>>>(D, vs, r->kinds, r->pars, nt, gp, B);
        else
            k_vjp_ghost<DIM, false><<<ng, BLOCK, 0, st>>>(D, vs, r->kinds, r->pars, nt, gp, B);
    }
    if (visc && ord1)
        k_vjp_gather<DIM, true, true><<<nc, BLOCK, 0, st>>>(D, vs, B);
    else if (visc)
        k_vjp_gather<DIM, true><<<nc, BLOCK, 0, st>>>(D, vs, B);
    else if (ord1)
        k_vjp_gather<DIM, false, true><<<nc, BLOCK, 0, st>>>(D, vs, B);
    else
        k_vjp_gather<DIM, false><<<nc, BLOCK, 0, st>>>(D, vs, B);
    // the ghost kernel zeroes the output's ghost rows of the ghosts on a
    // boundary face; any others are zeroed here
    if (D.G > D.nb) k_zero_rows<<<nblocks(D.G), BLOCK, 0, st>>>(NPRIM, D.N, D.Nt, B.pb);
}

}  // namespace

ibis_status ibis_compute_dudt_vjp(ibis_residual* r, double* prim, const double* dudt_bar,
                                  uint64_t force_tags, const double* force_dir, double* prim_bar,
                                  int limiter_mode, ibis_stream_t stream) {
    if (!r || !prim || !prim_bar || (force_tags && !force_dir))
        return ibis::fail(IBIS_ERR_INVALID, "ibis_compute_dudt_vjp: null argument");
    if (limiter_mode != IBIS_LIMITER_FROZEN && limiter_mode != IBIS_LIMITER_ACTIVE)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_compute_dudt_vjp: unknown limiter mode");
    ibis_status s = adjoint_supported(r);
    if (s != IBIS_OK) return s;
    s = ensure_adjoint(r);
    if (s != IBIS_OK) return s;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const DevMesh& D = r->d;
    const ibis_fv_desc& d = r->desc;
    const size_t N = D.N, F = D.F, Nt = D.Nt;
    const GasParams gp{d.R, d.Cv, d.Cp, d.gamma};
    FaceParams fp;
    VarSet vs;
    residual_params(r, &fp, &vs);
    // forward: the residual's state, gradients and limiters (recomputed);
    // on the fused viscous path with its own kernels: the boundary kernel
    // stores the ghosts' rows between the two passes in adj_prim_c (7 x G;
    // the valid rows are prim's), the cell kernel writes the convective
    // gradients and limiters and puts the viscous gradients in adj_vgrad;
    // on the fused inviscid path the state the reverse reads is prim itself
    const bool vis_fused = visc_fusable(r, vs);
    const bool inv_fused = !d.viscous && d.order == 2 && !r->has_post &&
                           r->prec == IBIS_PREC_FP64 && r->opt_layout == 2 && r->offs &&
                           r->opt_offsets && !r->opt_pack && r->opt_stage_cells;
    // the option fold_boundary: the boundary passes run in the cell kernel's
    // staging, which also records the ghosts between the viscous passes
    if (vis_fused) {
        const bool fold = r->opt_fold;
        const BcFold bf{r->kinds, r->pars, r->ntab, r->has_pre ? 1 : 0, r->has_visc ? 1 : 0,
                        gp, prim, r->FL, r->adj_prim_c};
        if (D.nb > 0 && !fold)
            k_bc_visc_fs<double><<<nblocks(D.nb), BLOCK, 0, st>>>(
                D, r->kinds, r->pars, r->ntab, r->has_pre ? 1 : 0, r->has_visc ? 1 : 0, gp, vs,
                prim, r->FL, r->gconv, r->adj_prim_c);
        if (D.dim == 3)
            launch_cell_visc<3>(D, vs, d, r, prim, r->adj_vgrad, 1, st, fold ? &bf : nullptr);
        else
            launch_cell_visc<2>(D, vs, d, r, prim, r->adj_vgrad, 1, st, fold ? &bf : nullptr);
    } else if (inv_fused) {
        // the inviscid second-order path's own kernels: boundary actions and
        // ghost face values, then the cell kernel writing the slots too
        const bool fold = fold_on(r);
        const BcFold bf{r->kinds, r->pars, r->ntab, r->has_pre ? 1 : 0, 0, gp, prim, r->FL,
                        nullptr};
        if (D.nb > 0 && !fold) {
            if (r->has_pre)
                k_bc_ghost_fs<2><<<nblocks(D.nb), BLOCK, 0, st>>>(D, r->kinds, r->pars, gp, vs,
                                                                  prim, r->FL, r->FR);
            else
                k_ghost_fs<2><<<nblocks(D.nb), BLOCK, 0, st>>>(D, vs, prim, r->FL, r->FR);
        }
        r->force_slots = 1;
        if (D.dim == 3)
            launch_cell_fs<3>(D, vs, d, r, D.ntiles, prim, st, fold ? &bf : nullptr);
        else
            launch_cell_fs<2>(D, vs, d, r, D.ntiles, prim, st, fold ? &bf : nullptr);
        r->force_slots = 0;
    } else {
    if (r->has_pre) k_bc<double><<<nblocks(D.nb), BLOCK, 0, st>>>(D, r->kinds, r->pars, gp, prim);
    HIP_TRY(hipMemcpyAsync(r->adj_prim_c, prim, sizeof(double) * NPRIM * Nt,
                           hipMemcpyDeviceToDevice, st));
    if (d.order == 2) {
        if (D.dim == 3)
            launch_cell_grad<double, 3>(D, vs, d.limiter, r->adj_prim_c, r->grad, r->lim,
                                        d.lim_eps, d.venkat_K, st);
        else
            launch_cell_grad<double, 2>(D, vs, d.limiter, r->adj_prim_c, r->grad, r->lim,
                                        d.lim_eps, d.venkat_K, st);
    }
    if (d.viscous) {
        if (r->has_visc)
            k_bc<double><<<nblocks(D.nb), BLOCK, 0, st>>>(D, r->kinds + r->ntab,
                                                          r->pars + (size_t)r->ntab * IBIS_NPAR, gp, prim);
        VarSet vv;
        vv.n = 4;
        vv.ngrad = 4;
        const int rows[5] = {P_T, P_VX, P_VY, P_VZ, P_VZ};
        const int slots[5] = {G_T, G_VX, G_VY, G_VZ, G_VZ};
        for (int q = 0; q < 5; q++) {
            vv.row[q] = rows[q];
            vv.slot[q] = slots[q];
        }
        if (D.dim == 3)
            launch_cell_grad<double, 3>(D, vv, 0, prim, r->adj_vgrad, r->lim, 0.0, 0.0, st);
        else
            launch_cell_grad<double, 2>(D, vv, 0, prim, r->adj_vgrad, r->lim, 0.0, 0.0, st);
    }
    }
    AdjBufs B;
    B.prim_c = (vis_fused || inv_fused) ? prim : r->adj_prim_c;
    B.prim_cg = vis_fused ? r->adj_prim_c : nullptr;
    B.prim_v = prim;
    B.cgrad = r->grad;
    B.clim = r->lim;
    B.vgrad = r->adj_vgrad;
    B.post_kinds = r->kinds + 2 * r->ntab;
    B.db = dudt_bar ? dudt_bar : r->adj_zero;
    B.force_tags = force_tags;
    for (int k = 0; k < 3; k++) B.force_dir[k] = force_tags ? force_dir[k] : 0.0;
    double* fa = r->adj_face;
    B.QLb = fa;
    B.QRb = fa + 7 * F;
    B.VLb = fa + 14 * F;
    B.VRb = fa + 18 * F;
    B.Gb = fa + 22 * F;
    double* ca = r->adj_cell;
    B.cself = ca;
    B.vself = ca + 7 * N;
    B.send = ca + 11 * N;
    B.bcb = r->adj_bcb;
    B.pb = prim_bar;
    B.lsqw = r->adj_lsqw;
    B.FLv = (vis_fused || inv_fused) ? r->FL : nullptr;
    B.offs = r->offs;
    const bool active = limiter_mode == IBIS_LIMITER_ACTIVE;
    if (D.dim == 3)
        launch_vjp<3>(r, vs, fp, B, active, st);
    else
        launch_vjp<2>(r, vs, fp, B, active, st);
    HIP_TRY(hipGetLastError());
    return IBIS_OK;
}

ibis_status ibis_cons_to_prim_vjp(ibis_residual* r, const double* cons, const double* prim_bar,
                                  double* cons_bar, ibis_stream_t stream) {
    if (!r || !cons || !prim_bar || !cons_bar)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_cons_to_prim_vjp: null argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const DevMesh& D = r->d;
    if (D.dim == 3)
        k_cons_to_prim_vjp<3><<<nblocks(D.Nt), BLOCK, 0, st>>>(D.Nt, r->desc.R, r->desc.Cv, cons,
                                                                prim_bar, cons_bar);
    else
        k_cons_to_prim_vjp<2><<<nblocks(D.Nt), BLOCK, 0, st>>>(D.Nt, r->desc.R, r->desc.Cv, cons,
                                                                prim_bar, cons_bar);
    HIP_TRY(hipGetLastError());
    return IBIS_OK;
}

ibis_status ibis_wall_force(ibis_residual* r, double* prim, uint64_t force_tags, double* force,
                            ibis_stream_t stream) {
    if (!r || !prim || !force) return ibis::fail(IBIS_ERR_INVALID, "ibis_wall_force: null argument");
    ibis_status s = ensure_adjoint(r);
    if (s != IBIS_OK) return s;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const DevMesh& D = r->d;
    if (r->prec != IBIS_PREC_FP64)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_wall_force is FP64 only");
    Prof pf{r, st};
    GenBufs<double> B{r->grad, r->lim, r->flux, r->face_rho, r->face_T};
    run_general<double>(r, B, prim, r->adj_dudt, r->desc.order == 2, st, pf);
    for (int k = 0; k < 3; k++) force[k] = 0.0;
    if (D.nb == 0) return IBIS_OK;
    k_force_terms<<<nblocks(D.nb), BLOCK, 0, st>>>(D, r->flux, force_tags, r->adj_force);
    HIP_TRY(hipGetLastError());
    std::vector<double> h(3 * (size_t)D.nb);
    HIP_TRY(hipMemcpyAsync(h.data(), r->adj_force, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    const std::vector<int>& order = r->m->b_in_tag_order;
    for (int k = 0; k < 3; k++)
        for (int b = 0; b < D.nb; b++) force[k] += h[(size_t)k * D.nb + order[b]];
    return IBIS_OK;
}

ibis_status ibis_residual_profile(ibis_residual* r, int enable) {
    if (!r) return ibis::fail(IBIS_ERR_INVALID, "ibis_residual_profile: null residual");
    r->profiling = enable != 0;
    return IBIS_OK;
}

ibis_status ibis_residual_kernel_times(ibis_residual* r, double* total_ms, int64_t* counts,
                                       int reset) {
    if (!r) return ibis::fail(IBIS_ERR_INVALID, "ibis_residual_kernel_times: null residual");
    for (auto& e : r->pending) {
        HIP_TRY(hipEventSynchronize(e.b));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, e.a, e.b));
        r->acc_ms[e.slot] += ms;
        r->acc_n[e.slot] += 1;
        r->pool.push_back(e.a);
        r->pool.push_back(e.b);
    }
    r->pending.clear();
    for (int q = 0; q < IBIS_NKERNELS; q++) {
        if (total_ms) total_ms[q] = r->acc_ms[q];
        if (counts) counts[q] = r->acc_n[q];
        if (reset) {
            r->acc_ms[q] = 0.0;
            r->acc_n[q] = 0;
        }
    }
    return IBIS_OK;
}

ibis_status ibis_estimate_dt(ibis_residual* r, const double* prim, double* dt_host,
                             ibis_stream_t stream) {
    if (!r || !prim || !dt_host) return ibis::fail(IBIS_ERR_INVALID, "ibis_estimate_dt: null argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const DevMesh& D = r->d;
    const ibis_fv_desc& d = r->desc;
    ViscParams vp{d.R, d.Cv, d.Cp, d.mu0, d.T0, d.Ts, d.Pr};
    const unsigned nb = nblocks(D.N);
    k_dt_partial<<<nb, BLOCK, 0, st>>>(D, prim, r->face_rho, r->face_T, d.viscous ? 1 : 0,
                                       d.signal_factor, vp, d.gamma, r->partial);
    k_min_final<<<1, BLOCK, 0, st>>>(r->partial, (int)nb, r->dt_dev);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(r->host_buf, r->dt_dev, sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    *dt_host = r->host_buf[0];
    return IBIS_OK;
}

ibis_status ibis_residual_cell_dt(ibis_residual* r, const double* prim, double* dt_cells,
                                  ibis_stream_t stream) {
    if (!r || !prim || !dt_cells) return ibis::fail(IBIS_ERR_INVALID, "ibis_residual_cell_dt: null argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const DevMesh& D = r->d;
    const ibis_fv_desc& d = r->desc;
    ViscParams vp{d.R, d.Cv, d.Cp, d.mu0, d.T0, d.Ts, d.Pr};
    k_dt_cells<<<nblocks(D.N), BLOCK, 0, st>>>(D, prim, r->face_rho, r->face_T, d.viscous ? 1 : 0,
                                               d.signal_factor, vp, d.gamma, dt_cells);
    HIP_TRY(hipGetLastError());
    return IBIS_OK;
}

ibis_status ibis_count_bad_cells(ibis_residual* r, const double* prim, uint64_t* n_host,
                                 ibis_stream_t stream) {
    if (!r || !prim || !n_host) return ibis::fail(IBIS_ERR_INVALID, "ibis_count_bad_cells: null argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const DevMesh& D = r->d;
    k_zero_u64<<<1, 1, 0, st>>>(r->count_dev);
    k_count_bad<<<nblocks(D.N), BLOCK, 0, st>>>(D.N, D.Nt, prim, r->count_dev);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(r->host_buf + 1, r->count_dev, sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    uint64_t v;
    std::memcpy(&v, r->host_buf + 1, 8);
    *n_host = v;
    return IBIS_OK;
}

ibis_status ibis_cons_to_prim(ibis_residual* r, const double* cons, double* prim,
                              ibis_stream_t stream) {
    if (!r || !cons || !prim) return ibis::fail(IBIS_ERR_INVALID, "ibis_cons_to_prim: null argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const DevMesh& D = r->d;
    if (D.dim == 3)
        k_cons_to_prim<double, 3><<<nblocks(D.Nt), BLOCK, 0, st>>>(D.Nt, r->desc.R, r->desc.Cv, cons, prim);
    else
        k_cons_to_prim<double, 2><<<nblocks(D.Nt), BLOCK, 0, st>>>(D.Nt, r->desc.R, r->desc.Cv, cons, prim);
    HIP_TRY(hipGetLastError());
    return IBIS_OK;
}

ibis_status ibis_prim_to_cons(ibis_residual* r, const double* prim, double* cons,
                              ibis_stream_t stream) {
    if (!r || !cons || !prim) return ibis::fail(IBIS_ERR_INVALID, "ibis_prim_to_cons: null argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const DevMesh& D = r->d;
    if (D.dim == 3)
        k_prim_to_cons<3><<<nblocks(D.Nt), BLOCK, 0, st>>>(D.Nt, prim, cons);
    else
        k_prim_to_cons<2><<<nblocks(D.Nt), BLOCK, 0, st>>>(D.Nt, prim, cons);
    HIP_TRY(hipGetLastError());
    return IBIS_OK;
}

ibis_status ibis_apply_time_derivative(ibis_residual* r, const double* u0, double* u1,
                                       const double* k, double dt, ibis_stream_t stream) {
    if (!r || !u0 || !u1 || !k) return ibis::fail(IBIS_ERR_INVALID, "ibis_apply_time_derivative: null argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const DevMesh& D = r->d;
    const long n = (long)(D.dim + 2) * D.Nt;
    k_axpy<<<nblocks(n), BLOCK, 0, st>>>(n, u0, u1, k, dt);
    HIP_TRY(hipGetLastError());
    return IBIS_OK;
}

// One RK step on caller-supplied stage buffers: k[0..s) and k_tmp of
// (dim+2) x n_total doubles, flow_tmp of 7 x n_total.
static ibis_status rk_step_on(ibis_residual* r, int s, const double* a, const double* b,
                              double cfl, double* cons, double* prim, double* dt_prev, double t,
                              double max_time, double plot_remaining, double* dt_taken,
                              double* const* k, double* k_tmp, double* flow_tmp,
                              ibis_stream_t stream) {
    ibis_status rc;
    // SO/runge_kutta.cpp:143-147
    if ((rc = ibis_compute_dudt(r, prim, k[0], 1, stream)) != IBIS_OK) return rc;
    double stable_dt;
    if ((rc = ibis_estimate_dt(r, prim, &stable_dt, stream)) != IBIS_OK) return rc;
    // RungeKutta::estimate_dt, SO/runge_kutta.cpp:108-119 (Kokkos::min ordering)
    auto kmin = [](double x, double y) { return (y < x) ? y : x; };
    double dt = kmin(cfl * stable_dt, 1.5 * (*dt_prev));
    dt = kmin(dt, max_time - t);
    if (plot_remaining > 0.0) dt = kmin(dt, plot_remaining);
    for (int i = 1; i < s; i++) {  // SO/runge_kutta.cpp:151-182
        if ((rc = ibis_apply_time_derivative(r, cons, k_tmp, k[0], a[i * s + 0] * dt, stream)) !=
            IBIS_OK)
            return rc;
        for (int j = 1; j < i; j++) {
            if (a[i * s + j] < 1e-14) continue;
            if ((rc = ibis_apply_time_derivative(r, k_tmp, k_tmp, k[j], a[i * s + j] * dt,
                                                 stream)) != IBIS_OK)
                return rc;
        }
        if ((rc = ibis_cons_to_prim(r, k_tmp, flow_tmp, stream)) != IBIS_OK) return rc;
        if ((rc = ibis_compute_dudt(r, flow_tmp, k[i], 1, stream)) != IBIS_OK) return rc;
    }
    for (int i = 0; i < s; i++)  // SO/runge_kutta.cpp:185-191
        if ((rc = ibis_apply_time_derivative(r, cons, cons, k[i], b[i] * dt, stream)) != IBIS_OK)
            return rc;
    if ((rc = ibis_cons_to_prim(r, cons, prim, stream)) != IBIS_OK) return rc;
    *dt_prev = dt;
    if (dt_taken) *dt_taken = dt;
    return IBIS_OK;
}

ibis_status ibis_rk_step(ibis_residual* r, int s, const double* a, const double* b, double cfl,
                         double* cons, double* prim, double* dt_prev, double t, double max_time,
                         double plot_remaining, double* dt_taken, ibis_stream_t stream) {
    if (!r || !a || !b || !cons || !prim || !dt_prev || s < 1 || s > 8)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_rk_step: bad argument");
    const DevMesh& D = r->d;
    const size_t ncv = (size_t)(D.dim + 2) * D.Nt;
    while ((int)r->k.size() < s) {
        double* p = nullptr;
        HIP_TRY(hipMalloc(&p, ncv * sizeof(double)));
        HIP_TRY(hipMemset(p, 0, ncv * sizeof(double)));
        r->k.push_back(p);
    }
    if (!r->k_tmp) {
        HIP_TRY(hipMalloc(&r->k_tmp, ncv * sizeof(double)));
        HIP_TRY(hipMemset(r->k_tmp, 0, ncv * sizeof(double)));
        HIP_TRY(hipMalloc(&r->flow_tmp, (size_t)NPRIM * D.Nt * sizeof(double)));
        HIP_TRY(hipMemset(r->flow_tmp, 0, (size_t)NPRIM * D.Nt * sizeof(double)));
    }
    return rk_step_on(r, s, a, b, cfl, cons, prim, dt_prev, t, max_time, plot_remaining,
                      dt_taken, r->k.data(), r->k_tmp, r->flow_tmp, stream);
}

// ---------------------------------------------------------------------------
// ibis_rk_step_noalloc: the RK step with its stage buffers in `scratch`, a
// fixed device array of the library, instead of buffers allocated on the
// first call.  Each GPU holds its own copy of the array (its address is
// looked up once per device with hipGetSymbolAddress); every call on a GPU
// shares it and finds the previous call's contents in it.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[IBIS_RK_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices];  // per device: the device address of `scratch`

// ghost rows [n_valid, n_total) of `nrows` rows of length n_total
__global__ void k_zero_ghost_rows(int nrows, long n_valid, long n_total, double* a) {
    const long ng = n_total - n_valid;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)nrows * ng) return;
    a[(i / ng) * n_total + n_valid + i % ng] = 0.0;
}
}  // namespace

static ibis_status scratch_address(double** out) {
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_rk_step_noalloc: device id out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_dev[dev];
    return IBIS_OK;
}

ibis_status ibis_rk_step_noalloc(ibis_residual* r, int s, const double* a, const double* b,
                                 double cfl, double* cons, double* prim, double* dt_prev,
                                 double t, double max_time, double plot_remaining,
                                 double* dt_taken, ibis_stream_t stream) {
    if (!r || !a || !b || !cons || !prim || !dt_prev || s < 1 || s > 8)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_rk_step_noalloc: bad argument");
    const DevMesh& D = r->d;
    const size_t nc = (size_t)(D.dim + 2);
    const size_t ncv = nc * D.Nt;
    const size_t need = ((size_t)s + 1) * ncv + (size_t)NPRIM * D.Nt;  // doubles
    if (need * sizeof(double) > (size_t)IBIS_RK_NOALLOC_SCRATCH_BYTES) {
        char msg[200];
        snprintf(msg, sizeof(msg),
                 "ibis_rk_step_noalloc: %d stages on %lld cells need %zu bytes of stage "
                 "buffers, above the fixed %zu-byte scratch",
                 s, (long long)D.Nt, need * sizeof(double), (size_t)IBIS_RK_NOALLOC_SCRATCH_BYTES);
        return ibis::fail(IBIS_ERR_INVALID, msg);
    }
    double* base = nullptr;
    ibis_status rc = scratch_address(&base);
    if (rc != IBIS_OK) return rc;
    double* k[8];
    for (int i = 0; i < s; i++) k[i] = base + (size_t)i * ncv;
    double* k_tmp = base + (size_t)s * ncv;
    double* flow_tmp = base + ((size_t)s + 1) * ncv;
    // the residual writes only the valid rows of each k_j; its ghost rows
    // must read 0, as in the zero-initialised buffers of ibis_rk_step (the
    // scratch holds whatever the previous call left)
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const long ng = (long)D.Nt - (long)D.N;
    if (ng > 0) {
        const long n = (long)s * (long)nc * ng;
        k_zero_ghost_rows<<<(unsigned)((n + BLOCK - 1) / BLOCK), BLOCK, 0, st>>>(
            (int)(s * nc), D.N, D.Nt, base);
        HIP_TRY(hipGetLastError());
    }
    return rk_step_on(r, s, a, b, cfl, cons, prim, dt_prev, t, max_time, plot_remaining,
                      dt_taken, k, k_tmp, flow_tmp, stream);
}

// FiniteVolume::compute_convective_gradient / compute_viscous_gradient
// (FV/finite_volume.cpp:273-288): the boundary actions write the ghost rows,
// then the LSQ gradients of the reconstruction variables (convective:
// FV/convective_flux.cpp:181-209) or of T, vx, vy, vz (viscous:
// FV/viscous_flux.cpp:122-130) overwrite their slots of the gradient
// storage; the other slots keep their values.
ibis_status ibis_compute_convective_gradient(ibis_residual* r, double* prim,
                                             ibis_stream_t stream) {
    if (!r || !prim)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_compute_convective_gradient: null argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const DevMesh& D = r->d;
    const ibis_fv_desc& d = r->desc;
    const GasParams gp{d.R, d.Cv, d.Cp, d.gamma};
    if (r->has_pre) k_bc<double><<<nblocks(D.nb), BLOCK, 0, st>>>(D, r->kinds, r->pars, gp, prim);
    FaceParams fp;
    VarSet vs;
    residual_params(r, &fp, &vs);
    vs.n = vs.ngrad = D.dim == 3 ? 5 : 4;  // vz only in 3D (convective_flux.cpp:206-208)
    if (D.dim == 3)
        launch_cell_grad<double, 3>(D, vs, 0, prim, r->grad, r->lim, 0.0, 0.0, st);
    else
        launch_cell_grad<double, 2>(D, vs, 0, prim, r->grad, r->lim, 0.0, 0.0, st);
    HIP_TRY(hipGetLastError());
    return IBIS_OK;
}

ibis_status ibis_compute_viscous_gradient(ibis_residual* r, double* prim, ibis_stream_t stream) {
    if (!r || !prim)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_compute_viscous_gradient: null argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const DevMesh& D = r->d;
    const ibis_fv_desc& d = r->desc;
    const GasParams gp{d.R, d.Cv, d.Cp, d.gamma};
    if (r->has_pre) k_bc<double><<<nblocks(D.nb), BLOCK, 0, st>>>(D, r->kinds, r->pars, gp, prim);
    if (r->has_visc)
        k_bc<double><<<nblocks(D.nb), BLOCK, 0, st>>>(D, r->kinds + r->ntab,
                                                      r->pars + (size_t)r->ntab * IBIS_NPAR, gp,
                                                      prim);
    VarSet vv;
    vv.n = vv.ngrad = 4;
    const int rows[5] = {P_T, P_VX, P_VY, P_VZ, P_VZ};
    const int slots[5] = {G_T, G_VX, G_VY, G_VZ, G_VZ};
    for (int q = 0; q < 5; q++) {
        vv.row[q] = rows[q];
        vv.slot[q] = slots[q];
    }
    if (D.dim == 3)
        launch_cell_grad<double, 3>(D, vv, 0, prim, r->grad, r->lim, 0.0, 0.0, st);
    else
        launch_cell_grad<double, 2>(D, vv, 0, prim, r->grad, r->lim, 0.0, 0.0, st);
    HIP_TRY(hipGetLastError());
    return IBIS_OK;
}

ibis_status ibis_cell_gradients(const ibis_residual* r, const double** grads) {
    if (!r || !grads) return ibis::fail(IBIS_ERR_INVALID, "ibis_cell_gradients: null argument");
    *grads = r->grad;
    return IBIS_OK;
}

ibis_status ibis_residual_scratch(ibis_residual* r, double** grads, double** limiters,
                                  double** flux, int* flux_has_area) {
    if (!r) return ibis::fail(IBIS_ERR_INVALID, "ibis_residual_scratch: null residual");
    if (grads) *grads = r->grad;
    if (limiters) *limiters = r->lim;
    if (flux) *flux = r->flux;
    if (flux_has_area) *flux_has_area = r->premul ? 1 : 0;
    return IBIS_OK;
}
