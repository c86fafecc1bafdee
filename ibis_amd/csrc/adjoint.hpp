// adjoint.hpp -- discrete adjoint (vector-Jacobian product) of the residual.
//
// Included by residual.hip after the forward kernels, inside its anonymous
// namespace (it uses DevMesh, VarSet, FaceParams, ViscParams).
//
// Given a cotangent db of dU/dt (rows [0, N)) and, optionally, a cotangent
// on the face momentum fluxes of selected boundary tags (a wall force
// functional), the kernels compute pb = (dR/dprim)^T db on the valid cells,
// following FiniteVolume::compute_dudt (finite_volume.cpp:57-96) backwards:
//
//   integral (finite_volume.cpp:221-255)      flux bar of every face
//   k_vjp_conv_face  flux calculator, frame rotation, thermo closure
//                    -> bars of the reconstructed values of both sides
//   k_vjp_visc_face  viscous flux (viscous_flux.cpp:133-184), Hasselbacher
//                    face gradients and face averages -> bars of T, v of
//                    both sides and of the two cells' viscous gradients
//   k_vjp_face_cv    the two above in one pass (second-order viscous)
//   k_vjp_cell       MUSCL reconstruction, limiter (active or frozen) and
//                    the cell's own LSQ stencil
//   k_vjp_ghost      ghost cells: their stencil contributions, then the
//                    viscous and pre-reconstruction boundary actions
//                    backwards onto the valid cell
//   k_vjp_gather     per valid cell: its own terms plus what its
//                    neighbours' stencils and its ghosts send back
//
// The forward pass keeps the gradients, the limiters, the ghosts' state
// before and after the viscous ghost fill and, on the fused paths, the face
// values its cell kernel wrote; fluxes are recomputed.  All sums are
// gathers in a fixed order, so the result is deterministic.
#ifndef IBIS_ADJOINT_HPP
#define IBIS_ADJOINT_HPP

// needs flux_adj.hpp, included by residual.hip at global scope

struct AdjBufs {
    // forward state
    const double* prim_c;  // 7 x Nt, ghosts after the pre-reconstruction actions
    // or, when set, the ghosts' rows after the pre-reconstruction actions
    // (7 x G), prim_c then holding only the valid rows (prim's own)
    const double* prim_cg;
    const double* prim_v;  // 7 x Nt, ghosts after the viscous actions
    const double* cgrad;   // NGRAD x 3 x N convective gradients
    const double* clim;    // NGRAD x N limiters
    const double* vgrad;   // NGRAD x 3 x N viscous gradients
    const int* post_kinds; // post-convective action table (ntags x IBIS_MAX_ACT)
    // seeds
    const double* db;      // (dim+2) x Nt
    unsigned long long force_tags;
    double force_dir[3];
    // face cotangents
    double* QLb;           // 5 x F: reconstructed values, left side
    double* QRb;           // 5 x F
    double* VLb;           // 4 x F: T, vx, vy, vz of the left cell (viscous)
    double* VRb;           // 4 x F
    double* Gb;            // 12 x F: viscous gradients of T, vx, vy, vz of either cell
                           // (equal for the two cells of an interior face; a
                           // boundary face's goes to its valid cell)
    // cell cotangents
    double* cself;         // 5 x N
    double* vself;         // 4 x N
    // what cell c's stencil sends to the neighbour in its slot j, as the
    // value that neighbour adds: send[(j*9 + q)*N + c], q < 5 the
    // reconstruction variables (LSQ weight . gradient bar, plus the
    // limiter's extreme-value bar when that neighbour is the extreme),
    // q = 5 + k the viscous T, vx, vy, vz (LSQ weight . gradient bar)
    double* send;
    double* bcb;           // 7 x nb: what each ghost sends to its valid cell
    double* pb;            // 7 x Nt output
    // geometry tables, built once: the LSQ weights of each cell's stencil
    // for the neighbour in slot j, w[(j*3 + k)*N + c], and the face offsets
    // x_face - x_cell, offs[(j*3 + k)*N + c]
    const double* lsqw;
    const double* offs;
    // the forward pass's face values (layout 2, as the fused cell kernel
    // wrote them), or null: then the reverse recomputes them
    const double* FLv;
};

// LSQ weights of cell c for a neighbour at centroid offset (dx, dy, dz)
// (GR/gradient.h:66-116, the same expression as the forward kernel)
template <int DIM>
__device__ __forceinline__ void lsq_weights(const DevMesh& m, int c, double dx, double dy,
                                            double dz, double* w) {
    const int N = m.N;
    const double r11 = m.lsq[c], r12 = m.lsq[N + c], r22 = m.lsq[2 * (size_t)N + c];
    double r13 = 0.0, r23 = 0.0, r33 = 0.0;
    if (DIM == 3) {
        r13 = m.lsq[3 * (size_t)N + c];
        r23 = m.lsq[4 * (size_t)N + c];
        r33 = m.lsq[5 * (size_t)N + c];
    }
    const double beta = (r12 * r23 - r13 * r23) / (r11 * r22);
    const double q12 = r12 / r11, q23 = r23 / r22;
    const double a1 = dx / (r11 * r11);
    const double a2 = 1.0 / (r22 * r22) * (dy - q12 * dx);
    double a3 = 0.0;
    if (DIM == 3) a3 = 1.0 / (r33 * r33) * (dz - q23 * dy + beta * dx);
    w[0] = a1 - q12 * a2 + beta * a3;
    w[1] = a2 - q23 * a3;
    w[2] = a3;
}

// w . gb for a weight triple stored at stride N (the table below); in 2D the
// z-weight is 0 (lsq_weights), so the z terms are left out
template <int DIM>
__device__ __forceinline__ double lsq_dot(const double* w, int N, const double* gb) {
    if (DIM == 3) return w[0] * gb[0] + w[N] * gb[1] + w[2 * (size_t)N] * gb[2];
    return w[0] * gb[0] + w[N] * gb[1];
}

// the weight table: lsq_weights of cell i for the neighbour in each of its
// slots, with the centroid differences written as the kernels write them
template <int DIM>
__global__ void __launch_bounds__(BLOCK) k_lsq_table(DevMesh m, double* __restrict__ w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = m.N, Nt = m.Nt;
    if (i >= N) return;
    for (int j = 0; j < m.nfaces[i]; j++) {
        const int nb = m.ell_nbr[(size_t)j * N + i];
        double wj[3];
        lsq_weights<DIM>(m, i, m.cen[nb] - m.cen[i], m.cen[Nt + nb] - m.cen[Nt + i],
                         m.cen[2 * (size_t)Nt + nb] - m.cen[2 * (size_t)Nt + i], wj);
        for (int k = 0; k < 3; k++) w[(size_t)(j * 3 + k) * N + i] = wj[k];
    }
}

// cotangent of the face flux (global frame, conserved rows) from dU/dt and
// the force functional
template <int DIM>
__device__ __forceinline__ void face_flux_bar(const DevMesh& m, const AdjBufs& B, int f,
                                              double* fb) {
    constexpr int NC = DIM + 2;
    const int N = m.N, Nt = m.Nt, F = m.F;
    const int l = m.L[f], r = m.R[f];
    const double A = m.geom[12 * (size_t)F + f];
#pragma unroll
    for (int q = 0; q < NC; q++) fb[q] = 0.0;
    if (l < N) {
        const double V = m.vol[l];
#pragma unroll
        for (int q = 0; q < NC; q++) fb[q] += B.db[(size_t)q * Nt + l] * (-A) / V;
    }
    if (r < N) {
        const double V = m.vol[r];
#pragma unroll
        for (int q = 0; q < NC; q++) fb[q] += B.db[(size_t)q * Nt + r] * A / V;
    }
    if (B.force_tags && (l >= N || r >= N)) {
        const int g = l >= N ? l : r;
        const int b = m.b_of_ghost[g - N];
        if ((B.force_tags >> m.b_tag[b]) & 1ull) {
            // force on the boundary: momentum flux * area, oriented out of
            // the valid cell
            const double s = l < N ? A : -A;
            for (int k = 0; k < DIM; k++) fb[1 + k] += B.force_dir[k] * s;
        }
    }
}

// 1 when face f is a boundary face whose tag overrides its convective flux
// with a constant (ConstantFlux, boundary.cpp:405-443): that flux does not
// depend on the state, so no cotangent flows through it
__device__ __forceinline__ bool constant_flux_face(const DevMesh& m, const AdjBufs& B, int f) {
    const int l = m.L[f], r = m.R[f], N = m.N;
    if (l < N && r < N) return false;
    const int g = l >= N ? l : r;
    const int t = m.b_tag[m.b_of_ghost[g - N]];
    for (int a = 0; a < IBIS_MAX_ACT; a++)
        if (B.post_kinds[t * IBIS_MAX_ACT + a] == IBIS_BC_CONSTANT_FLUX) return true;
    return false;
}

// Convective face flux of face f backwards, from its cotangent fb: bars of
// the reconstructed values
template <int DIM, int FLUX>
__device__ __forceinline__ void conv_face_rev(const DevMesh& m, const FaceParams& fp,
                                              const VarSet& vs, const AdjBufs& B, int f,
                                              const double (&fb)[DIM + 2]) {
    const int F = m.F, N = m.N, Nt = m.Nt;
    constexpr int NC = DIM + 2;
    constexpr bool T3 = DIM == 3;
    // forward: reconstruction of both sides (FV/convective_flux.cpp:243-420)
    double q[2][5], st[2][NPRIM];
    for (int side = 0; side < 2; side++) {
        const int c = side ? m.R[f] : m.L[f];
        const bool valid = c < N;
        if (B.FLv) {  // the same values, as the forward cell kernel stored them
            const int slot = side ? m.slotR[f] : m.slotL[f];
            for (int v = 0; v < 5; v++) q[side][v] = B.FLv[soa_index(m, c, slot, v)];
        } else {
            const double dx = m.geom[9 * (size_t)F + f] - m.cen[c];
            const double dy = m.geom[10 * (size_t)F + f] - m.cen[Nt + c];
            const double dz = m.geom[11 * (size_t)F + f] - m.cen[2 * (size_t)Nt + c];
            for (int v = 0; v < 5; v++) {
                const double u = (B.prim_cg && c >= N)
                                     ? B.prim_cg[(size_t)vs.row[v] * m.G + (c - N)]
                                     : B.prim_c[(size_t)vs.row[v] * Nt + c];
                double gx = 0.0, gy = 0.0, gz = 0.0, L = 1.0;
                if (valid) {
                    const double* g = B.cgrad + (size_t)vs.slot[v] * 3 * N;
                    gx = g[c];
                    gy = g[N + c];
                    gz = g[2 * (size_t)N + c];
                    if (fp.limited) L = B.clim[(size_t)vs.slot[v] * N + c];
                }
                q[side][v] = u + L * (gx * dx + gy * dy + gz * dz);
            }
        }
        double rho = 0.0, p = 0.0, T = 0.0, e = 0.0;
        const double a = q[side][0], b = q[side][1];
        switch (fp.thermo) {
            case IBIS_THERMO_RHO_P: p = a; rho = b; break;
            case IBIS_THERMO_RHO_T: rho = a; T = b; break;
            case IBIS_THERMO_RHO_U: rho = a; e = b; break;
            default: p = a; T = b; break;
        }
        ibis::thermo_update<double>(fp.thermo, fp.R, fp.Cv, rho, p, T, e);
        st[side][ibis::S_RHO] = rho;
        st[side][ibis::S_P] = p;
        st[side][ibis::S_T] = T;
        st[side][ibis::S_E] = e;
    }
    const double nx = m.geom[f], ny = m.geom[(size_t)F + f], nz = m.geom[2 * (size_t)F + f];
    const double t1x = m.geom[3 * (size_t)F + f], t1y = m.geom[4 * (size_t)F + f],
                 t1z = m.geom[5 * (size_t)F + f];
    const double t2x = m.geom[6 * (size_t)F + f], t2y = m.geom[7 * (size_t)F + f],
                 t2z = m.geom[8 * (size_t)F + f];
    for (int side = 0; side < 2; side++) {
        const double a = q[side][2], b = q[side][3], c = q[side][4];
        st[side][4] = a * nx + b * ny + c * nz;
        st[side][5] = a * t1x + b * t1y + c * t1z;
        st[side][6] = a * t2x + b * t2y + c * t2z;
    }
    // backward: rotation to the global frame (transpose), flux, rotation to
    // the face frame (transpose), thermo closure
    double ob[5];
    ob[0] = fb[0];
    ob[1] = nx * fb[1] + ny * fb[2] + (T3 ? nz * fb[3] : 0.0);
    ob[2] = t1x * fb[1] + t1y * fb[2] + (T3 ? t1z * fb[3] : 0.0);
    ob[3] = T3 ? t2x * fb[1] + t2y * fb[2] + t2z * fb[3] : 0.0;
    ob[4] = fb[NC - 1];
    double sb[2][NPRIM];
    for (int side = 0; side < 2; side++)
        for (int k = 0; k < NPRIM; k++) sb[side][k] = 0.0;
    if (FLUX == IBIS_FLUX_AUSMDV)
        ibis::flux_ausmdv_vjp<T3>(st[0], st[1], fp.R, fp.Cv, fp.gamma, ob, sb[0], sb[1]);
    else if (FLUX == IBIS_FLUX_HANEL)
        ibis::flux_hanel_vjp<T3>(st[0], st[1], fp.R, fp.Cv, fp.gamma, ob, sb[0], sb[1]);
    else
        ibis::flux_rusanov_vjp<T3>(st[0], st[1], fp.R, fp.Cv, fp.gamma, ob, sb[0], sb[1]);
    for (int side = 0; side < 2; side++) {
        const double s4 = sb[side][4], s5 = sb[side][5], s6 = sb[side][6];
        double qb[5];
        qb[2] = nx * s4 + t1x * s5 + t2x * s6;
        qb[3] = ny * s4 + t1y * s5 + t2y * s6;
        qb[4] = nz * s4 + t1z * s5 + t2z * s6;
        ibis::thermo_vjp(fp.thermo, fp.R, fp.Cv, st[side][ibis::S_RHO], st[side][ibis::S_P],
                         st[side][ibis::S_T], st[side][ibis::S_E], sb[side], &qb[0], &qb[1]);
        double* out = side ? B.QRb : B.QLb;
        for (int v = 0; v < 5; v++) out[(size_t)v * F + f] = qb[v];
    }
}

// Convective face flux backwards, one thread per face
template <int DIM, int FLUX>
__global__ void __launch_bounds__(BLOCK)
    k_vjp_conv_face(DevMesh m, FaceParams fp, VarSet vs, AdjBufs B) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    double fb[DIM + 2];
    face_flux_bar<DIM>(m, B, f, fb);
    if (constant_flux_face(m, B, f))
        for (int q = 0; q < DIM + 2; q++) fb[q] = 0.0;
    conv_face_rev<DIM, FLUX>(m, fp, vs, B, f, fb);
}

// First-order convective face flux backwards (copy_reconstruct,
// FV/convective_flux.cpp:212-240): the face states are the two cells'
// whole prim rows, rotated to the face frame, so the flux's cotangents go
// straight to rho, p, T, e and (rotated back) vx, vy, vz of each side:
// QLb / QRb hold 7 rows here, in prim row order.
template <int DIM, int FLUX>
__global__ void __launch_bounds__(BLOCK) k_vjp_conv_face1(DevMesh m, FaceParams fp, AdjBufs B) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    const int F = m.F, Nt = m.Nt;
    if (f >= F) return;
    constexpr int NC = DIM + 2;
    constexpr bool T3 = DIM == 3;
    double fb[NC];
    face_flux_bar<DIM>(m, B, f, fb);
    if (constant_flux_face(m, B, f))
        for (int q = 0; q < NC; q++) fb[q] = 0.0;
    const double nx = m.geom[f], ny = m.geom[(size_t)F + f], nz = m.geom[2 * (size_t)F + f];
    const double t1x = m.geom[3 * (size_t)F + f], t1y = m.geom[4 * (size_t)F + f],
                 t1z = m.geom[5 * (size_t)F + f];
    const double t2x = m.geom[6 * (size_t)F + f], t2y = m.geom[7 * (size_t)F + f],
                 t2z = m.geom[8 * (size_t)F + f];
    double st[2][NPRIM];
    for (int side = 0; side < 2; side++) {
        const int c = side ? m.R[f] : m.L[f];
        for (int k = 0; k < NPRIM; k++) st[side][k] = B.prim_c[(size_t)k * Nt + c];
        const double a = st[side][4], b = st[side][5], cc = st[side][6];
        st[side][4] = a * nx + b * ny + cc * nz;
        st[side][5] = a * t1x + b * t1y + cc * t1z;
        st[side][6] = a * t2x + b * t2y + cc * t2z;
    }
    double ob[5];
    ob[0] = fb[0];
    ob[1] = nx * fb[1] + ny * fb[2] + (T3 ? nz * fb[3] : 0.0);
    ob[2] = t1x * fb[1] + t1y * fb[2] + (T3 ? t1z * fb[3] : 0.0);
    ob[3] = T3 ? t2x * fb[1] + t2y * fb[2] + t2z * fb[3] : 0.0;
    ob[4] = fb[NC - 1];
    double sb[2][NPRIM];
    for (int side = 0; side < 2; side++)
        for (int k = 0; k < NPRIM; k++) sb[side][k] = 0.0;
    if (FLUX == IBIS_FLUX_AUSMDV)
        ibis::flux_ausmdv_vjp<T3>(st[0], st[1], fp.R, fp.Cv, fp.gamma, ob, sb[0], sb[1]);
    else if (FLUX == IBIS_FLUX_HANEL)
        ibis::flux_hanel_vjp<T3>(st[0], st[1], fp.R, fp.Cv, fp.gamma, ob, sb[0], sb[1]);
    else
        ibis::flux_rusanov_vjp<T3>(st[0], st[1], fp.R, fp.Cv, fp.gamma, ob, sb[0], sb[1]);
    for (int side = 0; side < 2; side++) {
        const double s4 = sb[side][4], s5 = sb[side][5], s6 = sb[side][6];
        double* out = side ? B.QRb : B.QLb;
        out[(size_t)P_RHO * F + f] = sb[side][P_RHO];
        out[(size_t)P_P * F + f] = sb[side][P_P];
        out[(size_t)P_T * F + f] = sb[side][P_T];
        out[(size_t)P_E * F + f] = sb[side][P_E];
        out[(size_t)P_VX * F + f] = nx * s4 + t1x * s5 + t2x * s6;
        out[(size_t)P_VY * F + f] = ny * s4 + t1y * s5 + t2y * s6;
        out[(size_t)P_VZ * F + f] = nz * s4 + t1z * s5 + t2z * s6;
    }
}

// GAS/transport_properties.h:27-42: d mu / d T of Sutherland's law
__device__ __forceinline__ double sutherland_dT(double T, double mu, const ViscParams& vp) {
    return mu * (1.5 / T - 1.0 / (T + vp.Ts));
}

// Viscous face flux of face f backwards (viscous_flux.cpp:16-184), from its
// cotangent fb
template <int DIM>
__device__ __forceinline__ void visc_face_rev(const DevMesh& m, const ViscParams& vp,
                                              const AdjBufs& B, int f,
                                              const double (&fb)[DIM + 2]) {
    const int F = m.F, N = m.N, Nt = m.Nt;
    constexpr int NC = DIM + 2;
    const int l = m.L[f], r = m.R[f];
    const bool lv = l < N, rv = r < N;
    const double n[3] = {m.geom[f], m.geom[(size_t)F + f], m.geom[2 * (size_t)F + f]};
    const double* P = B.prim_v;
    // forward: face gradients of T, vx, vy, vz (var k at g[k][0..2])
    const int slots[4] = {G_T, G_VX, G_VY, G_VZ};
    const int rows[4] = {P_T, P_VX, P_VY, P_VZ};
    double g[4][3];
    double ehat[3] = {0.0, 0.0, 0.0}, len_e = 1.0, edn = 1.0;
    const bool interior = lv && rv;
    if (!interior) {
        const int c = lv ? l : r;
        for (int k = 0; k < 4; k++)
            for (int d = 0; d < 3; d++)
                g[k][d] = d < 2 ? B.vgrad[((size_t)slots[k] * 3 + d) * N + c]
                                : grad_z<double, DIM>(B.vgrad + (size_t)slots[k] * 3 * N, N, c);
    } else {
        const double ex = m.cen[r] - m.cen[l];
        const double ey = m.cen[Nt + r] - m.cen[Nt + l];
        const double ez =
            DIM == 2 && m.flat_z ? 0.0 : m.cen[2 * (size_t)Nt + r] - m.cen[2 * (size_t)Nt + l];
        len_e = sqrt(ex * ex + ey * ey + ez * ez);
        ehat[0] = ex / len_e;
        ehat[1] = ey / len_e;
        ehat[2] = ez / len_e;
        edn = ehat[0] * n[0] + ehat[1] * n[1] + ehat[2] * n[2];
        for (int k = 0; k < 4; k++) {
            const double* gg = B.vgrad + (size_t)slots[k] * 3 * N;
            const double ax = 0.5 * (gg[l] + gg[r]);
            const double ay = 0.5 * (gg[N + l] + gg[N + r]);
            const double az = 0.5 * (grad_z<double, DIM>(gg, N, l) + grad_z<double, DIM>(gg, N, r));
            const double vl = P[(size_t)rows[k] * Nt + l], vr = P[(size_t)rows[k] * Nt + r];
            const double corr = ax * ehat[0] + ay * ehat[1] + az * ehat[2] - (vr - vl) / len_e;
            g[k][0] = ax - corr * n[0] / edn;
            g[k][1] = ay - corr * n[1] / edn;
            g[k][2] = az - corr * n[2] / edn;
        }
    }
    const double temp = 0.5 * (P[(size_t)P_T * Nt + l] + P[(size_t)P_T * Nt + r]);
    const double vx = 0.5 * (P[(size_t)P_VX * Nt + l] + P[(size_t)P_VX * Nt + r]);
    const double vy = 0.5 * (P[(size_t)P_VY * Nt + l] + P[(size_t)P_VY * Nt + r]);
    const double vz = 0.5 * (P[(size_t)P_VZ * Nt + l] + P[(size_t)P_VZ * Nt + r]);
    const double mu = sutherland<double>(temp, vp);
    const double kc = vp.Cp * mu / vp.Pr;
    const double lambda = -2.0 / 3.0 * mu;
    const double (&gT)[3] = g[0];
    const double (&gx)[3] = g[1];
    const double (&gy)[3] = g[2];
    const double (&gz)[3] = g[3];
    const double div = gx[0] + gy[1] + gz[2];
    const double bulk = lambda * div;
    const double txx = 2.0 * mu * gx[0] + bulk, tyy = 2.0 * mu * gy[1] + bulk,
                 tzz = 2.0 * mu * gz[2] + bulk;
    const double txy = mu * (gx[1] + gy[0]), txz = mu * (gx[2] + gz[0]), tyz = mu * (gy[2] + gz[1]);
    // backward: flux -= (n.tau_x, n.tau_y, [n.tau_z], n.theta)
    const double Fxb = -fb[1], Fyb = -fb[2], Fzb = DIM == 3 ? -fb[3] : 0.0, Feb = -fb[NC - 1];
    const double thxb = n[0] * Feb, thyb = n[1] * Feb, thzb = n[2] * Feb;
    const double txxb = n[0] * Fxb + vx * thxb;
    const double tyyb = n[1] * Fyb + vy * thyb;
    const double tzzb = n[2] * Fzb + vz * thzb;
    const double txyb = n[1] * Fxb + n[0] * Fyb + vy * thxb + vx * thyb;
    const double txzb = n[2] * Fxb + n[0] * Fzb + vz * thxb + vx * thzb;
    const double tyzb = n[2] * Fyb + n[1] * Fzb + vz * thyb + vy * thzb;
    const double vxb = thxb * txx + thyb * txy + thzb * txz;
    const double vyb = thxb * txy + thyb * tyy + thzb * tyz;
    const double vzb = thxb * txz + thyb * tyz + thzb * tzz;
    const double kcb = thxb * gT[0] + thyb * gT[1] + thzb * gT[2];
    double gb[4][3];
    gb[0][0] = thxb * kc;
    gb[0][1] = thyb * kc;
    gb[0][2] = thzb * kc;
    const double bulkb = txxb + tyyb + tzzb;
    const double divb = bulkb * lambda;
    double mub = 2.0 * gx[0] * txxb + 2.0 * gy[1] * tyyb + 2.0 * gz[2] * tzzb +
                 (gx[1] + gy[0]) * txyb + (gx[2] + gz[0]) * txzb + (gy[2] + gz[1]) * tyzb;
    mub += bulkb * div * (-2.0 / 3.0) + kcb * vp.Cp / vp.Pr;
    gb[1][0] = 2.0 * mu * txxb + divb;
    gb[2][1] = 2.0 * mu * tyyb + divb;
    gb[3][2] = 2.0 * mu * tzzb + divb;
    gb[1][1] = mu * txyb;
    gb[2][0] = mu * txyb;
    gb[1][2] = mu * txzb;
    gb[3][0] = mu * txzb;
    gb[2][2] = mu * tyzb;
    gb[3][1] = mu * tyzb;
    const double tempb = mub * sutherland_dT(temp, mu, vp);
    // face averages: half to each side
    double Lv[4] = {0.5 * tempb, 0.5 * vxb, 0.5 * vyb, 0.5 * vzb};
    double Rv[4] = {0.5 * tempb, 0.5 * vxb, 0.5 * vyb, 0.5 * vzb};
    // the gradient bars of the cells: the valid cell's gradient is the face
    // gradient at a boundary face (the ghost side gets none, and no kernel
    // reads it); both cells get half the Hasselbacher average's bar
    double G[4][3];
    if (!interior) {
        for (int k = 0; k < 4; k++)
            for (int d = 0; d < 3; d++) G[k][d] = gb[k][d];
    } else {  // Hasselbacher average
        for (int k = 0; k < 4; k++) {
            const double corrb = -(gb[k][0] * n[0] + gb[k][1] * n[1] + gb[k][2] * n[2]) / edn;
            for (int d = 0; d < 3; d++) {
                const double ab = gb[k][d] + corrb * ehat[d];
                G[k][d] = 0.5 * ab;
            }
            Rv[k] += -corrb / len_e;
            Lv[k] += corrb / len_e;
        }
    }
    for (int k = 0; k < 4; k++) {
        B.VLb[(size_t)k * F + f] = Lv[k];
        B.VRb[(size_t)k * F + f] = Rv[k];
        // in 2D the z component is not stored: its one reader (k_vjp_cell)
        // multiplies it by an LSQ z-weight, which is 0 in 2D
        for (int d = 0; d < DIM; d++) B.Gb[(size_t)(k * 3 + d) * F + f] = G[k][d];
    }
}

// Viscous face flux backwards, one thread per face
template <int DIM>
__global__ void __launch_bounds__(BLOCK) k_vjp_visc_face(DevMesh m, ViscParams vp, AdjBufs B) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    double fb[DIM + 2];
    face_flux_bar<DIM>(m, B, f, fb);
    visc_face_rev<DIM>(m, vp, B, f, fb);
}

// The convective and viscous face reverses of a second-order viscous
// residual in one pass over the faces: each face's flux cotangent is formed
// once (the convective one then zeroed on a constant-flux boundary face)
template <int DIM, int FLUX>
__global__ void __launch_bounds__(BLOCK)
    k_vjp_face_cv(DevMesh m, FaceParams fp, VarSet vs, ViscParams vp, AdjBufs B) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    double fb[DIM + 2];
    face_flux_bar<DIM>(m, B, f, fb);
    visc_face_rev<DIM>(m, vp, B, f, fb);
    if (constant_flux_face(m, B, f))
        for (int q = 0; q < DIM + 2; q++) fb[q] = 0.0;
    conv_face_rev<DIM, FLUX>(m, fp, vs, B, f, fb);
}

// Per valid cell: reconstruction, limiter and own-stencil terms
// ORD1: first-order residual -- the cell's own cotangent is the sum of its
// faces' (7 prim rows); there is no gradient or limiter to go through.
template <int DIM, int LIM, bool ACTIVE, bool VISC, bool ORD1 = false>
__global__ void __launch_bounds__(BLOCK)
    k_vjp_cell(DevMesh m, VarSet vs, double eps, double venkat_K, AdjBufs B) {
    // each thread's face offsets (maxf x 3 doubles), face ids (one's
    // complement when the cell is the face's right cell) and neighbours
    // (maxf ints each) are copied to LDS once, in planes of BLOCK, and read
    // from there by every variable's loops; a thread reads only its own slots
    extern __shared__ double vsh[];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = m.N, Nt = m.Nt, F = m.F, MF = m.maxf;
    if (i >= N) return;
    const int nf = m.nfaces[i];
    const int ngrad = DIM == 3 ? 5 : 4;
    double* sd = vsh;  // (j * 3 + k) * BLOCK + tid
    int* sf = reinterpret_cast<int*>(vsh + (size_t)MF * 3 * BLOCK);  // j * BLOCK + tid
    int* sn = sf + (size_t)MF * BLOCK;
    const int tid = threadIdx.x;
    for (int j = 0; j < nf; j++) {
        const int f = m.ell_face[(size_t)j * N + i];
        sf[j * BLOCK + tid] = m.ell_sign[(size_t)j * N + i] > 0 ? f : ~f;
        if (ACTIVE) sn[j * BLOCK + tid] = m.ell_nbr[(size_t)j * N + i];
        for (int k = 0; k < 3; k++)  // z: +0 on a flat 2D mesh
            sd[(j * 3 + k) * BLOCK + tid] =
                (k == 2 && DIM == 2 && m.flat_z) ? 0.0 : B.offs[(size_t)(j * 3 + k) * N + i];
    }
    if (ORD1) {
        for (int k = 0; k < NPRIM; k++) {
            double self = 0.0;
            for (int j = 0; j < nf; j++) {
                const int fe = sf[j * BLOCK + tid];
                self += (fe >= 0 ? B.QLb : B.QRb)[(size_t)k * F + (fe >= 0 ? fe : ~fe)];
            }
            B.cself[(size_t)k * N + i] = self;
        }
    }
    for (int v = 0; v < (ORD1 ? 0 : 5); v++) {
        const int row = vs.row[v], slot = vs.slot[v];
        const double* u = B.prim_c + (size_t)row * Nt;
        // ghost c's value at ug[c - N]
        const double* ug = B.prim_cg ? B.prim_cg + (size_t)row * m.G : u + N;
        const double ui = u[i];
        const double* gg = B.cgrad + (size_t)slot * 3 * N;
        const double g[3] = {gg[i], gg[N + i], grad_z<double, DIM>(gg, N, i)};
        const double phi = LIM != 0 ? B.clim[(size_t)slot * N + i] : 1.0;
        double self = 0.0, phib = 0.0, gb[3] = {0.0, 0.0, 0.0};
        // MUSCL: q = u + phi * (g . d) on each face of the cell
        for (int j = 0; j < nf; j++) {
            const int fe = sf[j * BLOCK + tid];
            const double qb = (fe >= 0 ? B.QLb : B.QRb)[(size_t)v * F + (fe >= 0 ? fe : ~fe)];
            const double d[3] = {sd[(j * 3 + 0) * BLOCK + tid], sd[(j * 3 + 1) * BLOCK + tid],
                                 sd[(j * 3 + 2) * BLOCK + tid]};
            self += qb;
            phib += qb * (g[0] * d[0] + g[1] * d[1] + g[2] * d[2]);
            for (int k = 0; k < DIM; k++) gb[k] += qb * phi * d[k];
        }
        double ext = 0.0;
        int extj = -1;
        if (LIM != 0 && ACTIVE && phib != 0.0) {
            // re-run the limiter with the Dual branch choices (UT/dual.h):
            // max/min keep the running value only when strictly better
            double umax = ui, umin = ui;
            int jmax = -1, jmin = -1;
            for (int j = 0; j < nf; j++) {
                const int c = sn[j * BLOCK + tid];
                const double un = c < N ? u[c] : ug[c - N];
                if (!(umax > un)) {
                    umax = un;
                    jmax = j;
                }
                if (!(umin < un)) {
                    umin = un;
                    jmin = j;
                }
            }
            double eps2 = 0.0;
            if (LIM == 2) {
                const double V = fabs(m.vol[i]);  // cell size h^dim = |V|
                const double K3 = venkat_K * venkat_K * venkat_K;
                eps2 = (DIM == 3) ? K3 * V : K3 * (V * sqrt(V));
            }
            // Venkatakrishnan: the running minimum as the fraction
            // bn / bd, compared by cross-multiplication as the forward
            // residual does (ibis::venkat_take on Dual: new on ties)
            double cur = 1.0, bn = 1.0, bd = 1.0;
            int jstar = -1;
            double d1s = 0.0, d2s = 0.0, dstar[3] = {0.0, 0.0, 0.0};
            int sgs = 0;
            for (int j = 0; j < nf; j++) {
                const double d[3] = {sd[(j * 3 + 0) * BLOCK + tid], sd[(j * 3 + 1) * BLOCK + tid],
                                     sd[(j * 3 + 2) * BLOCK + tid]};
                const double d2 = g[0] * d[0] + g[1] * d[1] + g[2] * d[2];
                const int sg = (d2 > 0.0) - (d2 < 0.0);
                if (sg == 0) continue;
                const double d1 = sg > 0 ? umax - ui : umin - ui;
                bool take;
                double val = 0.0, num = 0.0, den = 1.0;
                if (LIM == 1) {
                    val = d1 / (sg * (fabs(d2) + eps));
                    take = !(cur < val);  // Dual min: the new value on ties
                } else {
                    num = d1 * d1 + eps2 + 2.0 * (d1 * d2);
                    den = d1 * d1 + 2.0 * (d2 * d2) + d1 * d2 + eps2;
                    take = !(bn * den < num * bd);
                }
                if (take) {
                    cur = val;
                    bn = num;
                    bd = den;
                    jstar = j;
                    d1s = d1;
                    d2s = d2;
                    sgs = sg;
                    dstar[0] = d[0];
                    dstar[1] = d[1];
                    dstar[2] = d[2];
                }
            }
            if (jstar >= 0) {
                double d1b, d2b;
                if (LIM == 1) {
                    const double den = sgs * (fabs(d2s) + eps);
                    d1b = phib / den;
                    d2b = -phib * cur / den;  // d den / d d2 = 1 for either sign
                } else {
                    const double num = d1s * d1s + eps2 + 2.0 * (d1s * d2s);
                    const double den = d1s * d1s + 2.0 * (d2s * d2s) + d1s * d2s + eps2;
                    d1b = phib * ((2.0 * d1s + 2.0 * d2s) - num / den * (2.0 * d1s + d2s)) / den;
                    d2b = phib * (2.0 * d1s - num / den * (4.0 * d2s + d1s)) / den;
                }
                // d1 = U_ext - u_i, U_ext the running max / min
                const int je = sgs > 0 ? jmax : jmin;
                if (je >= 0) {
                    ext = d1b;
                    extj = je;
                } else {
                    self += d1b;  // the extreme is u_i itself: d1 = 0 identically
                }
                self -= d1b;
                for (int k = 0; k < DIM; k++) gb[k] += d2b * dstar[k];
            }
        }
        // own LSQ stencil g = sum_j w_j (u_j - u_i): the cell's own term and
        // what slot j's neighbour receives (2D vz: no gradient is computed)
        for (int j = 0; j < nf; j++) {
            const double* w = B.lsqw + (size_t)j * 3 * N + i;
            double x = 0.0;
            if (v < ngrad) {
                x = lsq_dot<DIM>(w, N, gb);
                self -= x;
            }
            if (extj == j) x += ext;
            B.send[((size_t)j * 9 + v) * N + i] = x;
        }
        B.cself[(size_t)v * N + i] = self;
    }
    if (VISC) {
        for (int k = 0; k < 4; k++) {
            double self = 0.0, gb[3] = {0.0, 0.0, 0.0};
            for (int j = 0; j < nf; j++) {
                const int fe = sf[j * BLOCK + tid];
                const bool left = fe >= 0;
                const int f = left ? fe : ~fe;
                self += (left ? B.VLb : B.VRb)[(size_t)k * F + f];
                for (int d = 0; d < DIM; d++) gb[d] += B.Gb[(size_t)(k * 3 + d) * F + f];
            }
            for (int j = 0; j < nf; j++) {
                const double* w = B.lsqw + (size_t)j * 3 * N + i;
                const double x = lsq_dot<DIM>(w, N, gb);
                self -= x;
                B.send[((size_t)j * 9 + 5 + k) * N + i] = x;
            }
            B.vself[(size_t)k * N + i] = self;
        }
    }
}

// slot of face f in the list of cell c
__device__ __forceinline__ int slot_of(const DevMesh& m, int f, int c) {
    return m.L[f] == c ? m.slotL[f] : m.slotR[f];
}

// reverse of one boundary action on the ghost bar gb (7 prim rows) and
// the valid cell's bar vb (boundary.cpp:15-399); par: the action's
// parameters, vst: the valid cell's prim values (forward)
__device__ __forceinline__ void bc_action_vjp(const DevMesh& m, int kind, int f,
                                              const double* par, const double* vst,
                                              const GasParams& gp, double* gb, double* vb) {
    const int F = m.F;
    switch (kind) {
        case IBIS_BC_SUBSONIC_INFLOW:    // boundary.cpp:286-342
        case IBIS_BC_SUBSONIC_OUTFLOW: {  // boundary.cpp:348-399
            const double n[3] = {m.geom[f], m.geom[(size_t)F + f], m.geom[2 * (size_t)F + f]};
            const double gR = gp.gamma * gp.R;
            const double a = sqrt(gR * vst[P_T]);
            const double rr = vst[P_RHO];
            const double D = rr * a;
            const bool in = kind == IBIS_BC_SUBSONIC_INFLOW;
            // forward: face pressure and density, ghost pressure/density/T
            double pf, rf, vc = 0.0, X;
            if (in) {
                for (int k = 0; k < 3; k++) vc += n[k] * (par[4 + k] - vst[P_VX + k]);
                pf = 0.5 * (par[1] + vst[P_P] - rr * a * vc);
                rf = par[2] + (pf - par[1]) / a / a;
                X = (par[1] - pf) / D;  // v_face = v_in - n X
            } else {
                pf = par[0];
                rf = rr + (pf - vst[P_P]) / a / a;
                X = (vst[P_P] - pf) / D;  // v_face = v_valid + n X
            }
            const double pg = 2.0 * pf - vst[P_P];
            const double rg = 2.0 * rf - rr;
            const double Tg = pg / (rg * gp.R);
            // backward: e = Cv T, T = p / (rho R), ghost = 2 face - valid
            const double Tgb = gb[P_T] + gp.Cv * gb[P_E];
            const double pgb = gb[P_P] + Tgb / (rg * gp.R);
            const double rgb = gb[P_RHO] - Tgb * Tg / rg;
            double vfb[3];
            for (int k = 0; k < 3; k++) {
                vfb[k] = 2.0 * gb[P_VX + k];
                vb[P_VX + k] -= gb[P_VX + k];
            }
            const double rfb = 2.0 * rgb;
            vb[P_RHO] -= rgb;
            vb[P_P] -= pgb;
            double pfb = in ? 2.0 * pgb : 0.0;  // the outflow's face pressure is given
            const double nv = n[0] * vfb[0] + n[1] * vfb[1] + n[2] * vfb[2];
            double Xb, prb = 0.0, rrb = 0.0, ab = 0.0;
            if (in) {
                Xb = -nv;
                pfb += -Xb / D;
            } else {
                for (int k = 0; k < 3; k++) vb[P_VX + k] += vfb[k];
                Xb = nv;
                prb += Xb / D;
            }
            const double Db = -Xb * X / D;
            rrb += Db * a;
            ab += Db * rr;
            const double dpa = in ? pf - par[1] : pf - vst[P_P];  // rho_face's pressure term
            if (in) {
                pfb += rfb / (a * a);
            } else {
                vb[P_RHO] += rfb;
                prb += -rfb / (a * a);
            }
            ab += rfb * (-2.0) * dpa / (a * a * a);
            if (in) {  // p_face = (p_in + p - rho a vc) / 2
                prb += 0.5 * pfb;
                rrb += -0.5 * pfb * a * vc;
                ab += -0.5 * pfb * rr * vc;
                const double vcb = -0.5 * pfb * rr * a;
                for (int k = 0; k < 3; k++) vb[P_VX + k] += -n[k] * vcb;
            }
            vb[P_P] += prb;
            vb[P_RHO] += rrb;
            vb[P_T] += ab * gR / (2.0 * a);
            for (int k = 0; k < NPRIM; k++) gb[k] = 0.0;
            break;
        }
        case IBIS_BC_FLOW_STATE_COPY:
            for (int k = 0; k < NPRIM; k++) gb[k] = 0.0;
            break;
        case IBIS_BC_INTERNAL_COPY:
            for (int k = 0; k < NPRIM; k++) {
                vb[k] += gb[k];
                gb[k] = 0.0;
            }
            break;
        case IBIS_BC_REFLECT_NORMAL: {
            for (int k = 0; k < 4; k++) {  // rho, p, T, e
                vb[k] += gb[k];
                gb[k] = 0.0;
            }
            const double nx = m.geom[0 * F + f], ny = m.geom[1 * F + f], nz = m.geom[2 * F + f];
            const double t1x = m.geom[3 * F + f], t1y = m.geom[4 * F + f], t1z = m.geom[5 * F + f];
            const double t2x = m.geom[6 * F + f], t2y = m.geom[7 * F + f], t2z = m.geom[8 * F + f];
            const double bx = gb[P_VX], by = gb[P_VY], bz = gb[P_VZ];
            // v_g = xs n + ys t1 + zs t2, xs = -(v.n), ys = v.t1, zs = v.t2
            const double xsb = bx * nx + by * ny + bz * nz;
            const double ysb = bx * t1x + by * t1y + bz * t1z;
            const double zsb = bx * t2x + by * t2y + bz * t2z;
            vb[P_VX] += -xsb * nx + ysb * t1x + zsb * t2x;
            vb[P_VY] += -xsb * ny + ysb * t1y + zsb * t2y;
            vb[P_VZ] += -xsb * nz + ysb * t1z + zsb * t2z;
            gb[P_VX] = gb[P_VY] = gb[P_VZ] = 0.0;
            break;
        }
        case IBIS_BC_VEL_REFLECT:
            for (int k = P_VX; k <= P_VZ; k++) {
                vb[k] -= gb[k];
                gb[k] = 0.0;
            }
            break;
        case IBIS_BC_FIX_TEMPERATURE:
            vb[P_T] -= gb[P_T];
            gb[P_T] = 0.0;
            break;
        default:
            break;
    }
}

// Per boundary face: the ghost's bars, the boundary actions backwards
template <int DIM, bool VISC, bool ORD1 = false>
__global__ void __launch_bounds__(BLOCK)
    k_vjp_ghost(DevMesh m, VarSet vs, const int* __restrict__ kinds,
                const double* __restrict__ pars, int ntab, GasParams gp, AdjBufs B) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m.nb) return;
    const int N = m.N, F = m.F, Nt = m.Nt;
    const int f = m.b_face[b], g = m.b_ghost[b], va = m.b_valid[b], t = m.b_tag[b];
    double vst[NPRIM];  // the valid cell's state, which the actions read
    for (int k = 0; k < NPRIM; k++) vst[k] = B.prim_c[(size_t)k * Nt + va];
    const int j = slot_of(m, f, va);
    double gb[NPRIM], vb[NPRIM];
    for (int k = 0; k < NPRIM; k++) gb[k] = vb[k] = 0.0;
    if (VISC) {  // viscous path reads the ghost after the viscous actions
        const bool gl = m.L[f] == g;
        const int rows[4] = {P_T, P_VX, P_VY, P_VZ};
        for (int k = 0; k < 4; k++) {
            double x = (gl ? B.VLb : B.VRb)[(size_t)k * F + f];
            x += B.send[((size_t)j * 9 + 5 + k) * N + va];
            gb[rows[k]] += x;
        }
        for (int a = IBIS_MAX_ACT - 1; a >= 0; a--)
            bc_action_vjp(m, kinds[ntab + t * IBIS_MAX_ACT + a], f,
                          pars + ((size_t)ntab + t * IBIS_MAX_ACT + a) * IBIS_NPAR, vst, gp, gb, vb);
    }
    if (ORD1) {  // first order: the ghost's whole row is the face state
        const bool gl = m.L[f] == g;
        for (int k = 0; k < NPRIM; k++) gb[k] += (gl ? B.QLb : B.QRb)[(size_t)k * F + f];
    } else {  // convective path reads the ghost after the pre-reconstruction actions
        const bool gl = m.L[f] == g;
        for (int v = 0; v < 5; v++) {
            double x = (gl ? B.QLb : B.QRb)[(size_t)v * F + f];  // ghost side: plain copy
            x += B.send[((size_t)j * 9 + v) * N + va];
            gb[vs.row[v]] += x;
        }
    }
    for (int a = IBIS_MAX_ACT - 1; a >= 0; a--)
        bc_action_vjp(m, kinds[t * IBIS_MAX_ACT + a], f, pars + ((size_t)t * IBIS_MAX_ACT + a) * IBIS_NPAR,
                      vst, gp, gb, vb);
    for (int k = 0; k < NPRIM; k++) B.bcb[(size_t)k * m.nb + b] = vb[k];
    // the output has no ghost rows: this ghost's are zero
    for (int k = 0; k < NPRIM; k++) B.pb[(size_t)k * Nt + g] = 0.0;
}

// Per valid cell: own terms + neighbours' stencils + ghosts' actions
template <int DIM, bool VISC, bool ORD1 = false>
__global__ void __launch_bounds__(BLOCK) k_vjp_gather(DevMesh m, VarSet vs, AdjBufs B) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = m.N, Nt = m.Nt;
    if (i >= N) return;
    double out[NPRIM];
    for (int k = 0; k < NPRIM; k++) out[k] = 0.0;
    if (ORD1)
        for (int k = 0; k < NPRIM; k++) out[k] += B.cself[(size_t)k * N + i];
    else
        for (int v = 0; v < 5; v++) out[vs.row[v]] += B.cself[(size_t)v * N + i];
    const int vrows[4] = {P_T, P_VX, P_VY, P_VZ};
    if (VISC)
        for (int k = 0; k < 4; k++) out[vrows[k]] += B.vself[(size_t)k * N + i];
    const int nf = m.nfaces[i];
    for (int j = 0; j < nf; j++) {
        const int nb = m.ell_nbr[(size_t)j * N + i];
        const int f = m.ell_face[(size_t)j * N + i];
        if (nb >= N) {  // ghost: its actions' share for this cell
            const int b = m.b_of_ghost[nb - N];
            for (int k = 0; k < NPRIM; k++) out[k] += B.bcb[(size_t)k * m.nb + b];
            continue;
        }
        // cell nb's stencil holds this cell at slot s
        const int s = slot_of(m, f, nb);
        for (int v = 0; v < (ORD1 ? 0 : 5); v++) out[vs.row[v]] += B.send[((size_t)s * 9 + v) * N + nb];
        if (VISC)
            for (int k = 0; k < 4; k++) out[vrows[k]] += B.send[((size_t)s * 9 + 5 + k) * N + nb];
    }
    for (int k = 0; k < NPRIM; k++) B.pb[(size_t)k * Nt + i] = out[k];
}

__global__ void __launch_bounds__(BLOCK) k_zero_rows(int rows, int N, int Nt, double* __restrict__ a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nt - N) return;
    for (int k = 0; k < rows; k++) a[(size_t)k * Nt + N + i] = 0.0;
}

// conserved_to_primatives backwards (primative_conserved_conversion.cpp:
// 5-30): prim bar -> cons bar, all n_total rows
template <int DIM>
__global__ void __launch_bounds__(BLOCK)
    k_cons_to_prim_vjp(int Nt, double R, double Cv, const double* __restrict__ cq,
                       const double* __restrict__ pb, double* __restrict__ cb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nt) return;
    const double rho = cq[i];
    const double mx = cq[(size_t)Nt + i], my = cq[2 * (size_t)Nt + i];
    const double mz = DIM == 3 ? cq[3 * (size_t)Nt + i] : 0.0;
    const double E = cq[(size_t)(DIM + 1) * Nt + i];
    const double vx = mx / rho, vy = my / rho, vz = DIM == 3 ? mz / rho : 0.0;
    const double ke = 0.5 * (vx * vx + vy * vy + vz * vz);
    const double u = E / rho - ke;
    const double T = u / Cv;
    // p = rho R T, T = u / Cv, u = E/rho - ke
    double rhob = pb[(size_t)P_RHO * Nt + i] + pb[(size_t)P_P * Nt + i] * R * T;
    const double Tb = pb[(size_t)P_T * Nt + i] + pb[(size_t)P_P * Nt + i] * rho * R;
    const double ub = pb[(size_t)P_E * Nt + i] + Tb / Cv;
    double vxb = pb[(size_t)P_VX * Nt + i] - ub * vx;
    double vyb = pb[(size_t)P_VY * Nt + i] - ub * vy;
    double vzb = DIM == 3 ? pb[(size_t)P_VZ * Nt + i] - ub * vz : 0.0;
    const double Eb = ub / rho;
    rhob -= ub * E / (rho * rho);
    rhob -= (vxb * vx + vyb * vy + vzb * vz) / rho;
    cb[i] = rhob;
    cb[(size_t)Nt + i] = vxb / rho;
    cb[2 * (size_t)Nt + i] = vyb / rho;
    if (DIM == 3) cb[3 * (size_t)Nt + i] = vzb / rho;
    cb[(size_t)(DIM + 1) * Nt + i] = Eb;
}

#endif
