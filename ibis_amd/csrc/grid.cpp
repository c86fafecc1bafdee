// grid.cpp -- host-side GridBlock construction for the MI355X residual.
//
// Builds faces, cell/face connectivity, geometry, ghost cells and the LSQ
// factors from an SU2-like description with the reference's numbering and
// arithmetic (src/grid/grid/grid.h:89-443, interface.h:180-323,
// cell.h:228-300, geom.h:12-85, gradient.h:119-162), then optionally
// renumbers cells and faces for locality so that the device kernels'
// gathers are coalesced.  Renumbering keeps every per-cell list in the
// reference order, so results are bit-identical under the permutation.
//
// Face identification uses a sort of packed vertex-set keys instead of the
// reference's string hash map; both give faces in order of first
// appearance, which is what fixes the reference's face numbering.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

#include "ibis_internal.h"

namespace {

enum { LINE = 3, TRI = 5, QUAD = 9, TETRA = 10, HEX = 12, WEDGE = 13, PYRAMID = 14 };

int nverts(int t) {
    switch (t) {
        case LINE: return 2;
        case TRI: return 3;
        case QUAD: return 4;
        case TETRA: return 4;
        case HEX: return 8;
        case WEDGE: return 6;
        case PYRAMID: return 5;
    }
    return -1;
}

// Local faces of an element in VTK order (src/grid/grid/grid_io.cpp:298-348).
// Each entry: face shape, then up to 4 local vertex indices (-1 padded).
struct LocalFace {
    int shape;
    int v[4];
};
int local_faces(int t, LocalFace* out) {
    static const LocalFace tri[] = {{LINE, {0, 1, -1, -1}}, {LINE, {1, 2, -1, -1}},
                                    {LINE, {2, 0, -1, -1}}};
    static const LocalFace quad[] = {{LINE, {0, 1, -1, -1}}, {LINE, {1, 2, -1, -1}},
                                     {LINE, {2, 3, -1, -1}}, {LINE, {3, 0, -1, -1}}};
    static const LocalFace tet[] = {{TRI, {0, 1, 2, -1}}, {TRI, {0, 1, 3, -1}},
                                    {TRI, {1, 2, 3, -1}}, {TRI, {0, 2, 3, -1}}};
    static const LocalFace hex[] = {{QUAD, {0, 1, 2, 3}}, {QUAD, {0, 1, 5, 4}},
                                    {QUAD, {4, 5, 6, 7}}, {QUAD, {2, 3, 7, 6}},
                                    {QUAD, {0, 4, 7, 3}}, {QUAD, {1, 5, 6, 2}}};
    static const LocalFace wedge[] = {{TRI, {0, 1, 2, -1}}, {TRI, {3, 5, 4, -1}},
                                      {QUAD, {1, 4, 5, 2}}, {QUAD, {0, 2, 5, 3}},
                                      {QUAD, {0, 3, 4, 1}}};
    static const LocalFace pyr[] = {{QUAD, {0, 3, 2, 1}}, {TRI, {2, 3, 4, -1}},
                                    {TRI, {0, 4, 3, -1}}, {TRI, {0, 1, 4, -1}},
                                    {TRI, {1, 2, 4, -1}}};
    const LocalFace* src = nullptr;
    int n = 0;
    switch (t) {
        case TRI: src = tri; n = 3; break;
        case QUAD: src = quad; n = 4; break;
        case TETRA: src = tet; n = 4; break;
        case HEX: src = hex; n = 6; break;
        case WEDGE: src = wedge; n = 5; break;
        case PYRAMID: src = pyr; n = 5; break;
        default: return -1;
    }
    std::copy(src, src + n, out);
    return n;
}

// vertex-set key: ids sorted descending, packed two per word
struct Key {
    uint64_t a, b;
    bool operator<(const Key& o) const { return a < o.a || (a == o.a && b < o.b); }
    bool operator==(const Key& o) const { return a == o.a && b == o.b; }
};
Key make_key(const int64_t* ids, int n) {
    uint64_t s[4] = {0xffffffffull, 0xffffffffull, 0xffffffffull, 0xffffffffull};
    for (int i = 0; i < n; i++) s[i] = (uint64_t)ids[i];
    std::sort(s, s + n, [](uint64_t x, uint64_t y) { return x > y; });
    return Key{(s[0] << 32) | s[1], (s[2] << 32) | s[3]};
}

struct Builder {
    int dim;
    int64_t nv, N, G = 0, F = 0;
    const double* xyz;
    std::vector<double> vx, vy, vz;
    std::vector<int64_t> cv_off, cv;
    std::vector<int> cshape;
    std::vector<int64_t> fv_off, fv;
    std::vector<int> fshape;
    std::vector<int64_t> cf_off, cf;
    std::vector<Key> ukeys;          // sorted unique face keys
    std::vector<int64_t> ukey_face;  // face id of each unique key
};

double tri_area(const Builder& b, int64_t p, int64_t q, int64_t r) {
    // Ibis::area_of_triangle, GR/geom.h:29-49
    double ab_x = b.vx[q] - b.vx[p], ab_y = b.vy[q] - b.vy[p], ab_z = b.vz[q] - b.vz[p];
    double ac_x = b.vx[r] - b.vx[p], ac_y = b.vy[r] - b.vy[p], ac_z = b.vz[r] - b.vz[p];
    double cx = ab_y * ac_z - ab_z * ac_y;
    double cy = ab_z * ac_x - ab_x * ac_z;
    double cz = ab_x * ac_y - ab_y * ac_x;
    return 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
}

double quad_area(const Builder& b, int64_t p, int64_t q, int64_t r, int64_t s) {
    double a1 = tri_area(b, p, q, r);
    double a2 = tri_area(b, p, r, s);
    return a1 + a2;
}

uint64_t spread_bits3(uint64_t x) {  // 21 bits -> every third bit
    x &= 0x1fffffull;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}
uint64_t spread_bits2(uint64_t x) {  // 32 bits -> every second bit
    x &= 0xffffffffull;
    x = (x | x << 16) & 0x0000ffff0000ffffull;
    x = (x | x << 8) & 0x00ff00ff00ff00ffull;
    x = (x | x << 4) & 0x0f0f0f0f0f0f0f0full;
    x = (x | x << 2) & 0x3333333333333333ull;
    x = (x | x << 1) & 0x5555555555555555ull;
    return x;
}

}  // namespace

ibis_status ibis_grid_build(const ibis_grid_io* io, int reorder, ibis_grid** out) {
    if (!io || !out) return ibis::fail(IBIS_ERR_INVALID, "ibis_grid_build: null argument");
    if (io->dim != 2 && io->dim != 3)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_grid_build: dim must be 2 or 3");
    if (io->ntags < 0 || io->ntags > IBIS_MAX_TAGS)
        return ibis::fail(IBIS_ERR_INVALID, "ibis_grid_build: too many boundary tags");
    Builder b;
    b.dim = io->dim;
    b.nv = io->nv;
    b.N = io->nc;
    b.vx.resize(b.nv);
    b.vy.resize(b.nv);
    b.vz.resize(b.nv);
    for (int64_t i = 0; i < b.nv; i++) {
        b.vx[i] = io->xyz[3 * i];
        b.vy[i] = io->xyz[3 * i + 1];
        b.vz[i] = io->dim == 3 ? io->xyz[3 * i + 2] : 0.0;  // read_vertex, grid_io.cpp:175-199
    }
    if (b.nv >= (int64_t)0xffffffff)
        return ibis::fail(IBIS_ERR_GRID, "ibis_grid_build: too many vertices");

    // ---- faces in order of first appearance (GR/grid.h:109-133) ----
    struct Entry {
        Key k;
        int64_t e;
    };
    std::vector<Entry> entries;
    std::vector<int> entry_shape;
    std::vector<int64_t> entry_ids;  // 4 per entry
    b.cv_off.assign(1, 0);
    b.cf_off.assign(1, 0);
    for (int64_t c = 0; c < b.N; c++) {
        int t = io->ctype[c];
        int n = nverts(t);
        if (n < 0 || t == LINE) return ibis::fail(IBIS_ERR_GRID, "unsupported cell type");
        const int64_t* ids = io->cvtx + 8 * c;
        for (int j = 0; j < n; j++) {
            if (ids[j] < 0 || ids[j] >= b.nv)
                return ibis::fail(IBIS_ERR_GRID, "cell vertex id out of range");
            b.cv.push_back(ids[j]);
        }
        b.cv_off.push_back((int64_t)b.cv.size());
        b.cshape.push_back(t);
        LocalFace lf[6];
        int nf = local_faces(t, lf);
        for (int j = 0; j < nf; j++) {
            int64_t v[4] = {-1, -1, -1, -1};
            int m = 0;
            for (int q = 0; q < 4; q++)
                if (lf[j].v[q] >= 0) v[m++] = ids[lf[j].v[q]];
            entries.push_back({make_key(v, m), (int64_t)entries.size()});
            entry_shape.push_back(lf[j].shape);
            for (int q = 0; q < 4; q++) entry_ids.push_back(v[q]);
        }
        b.cf_off.push_back((int64_t)entries.size());
    }
    const int64_t M = (int64_t)entries.size();
    std::vector<Entry> sorted(entries);
    std::sort(sorted.begin(), sorted.end(), [](const Entry& x, const Entry& y) {
        return x.k < y.k || (x.k == y.k && x.e < y.e);
    });
    // groups of equal keys; the first entry of a group is its first appearance
    std::vector<int64_t> firsts;
    std::vector<int64_t> group_of_sorted(M);
    for (int64_t s = 0; s < M; s++) {
        if (s == 0 || !(sorted[s].k == sorted[s - 1].k)) {
            firsts.push_back(sorted[s].e);
            b.ukeys.push_back(sorted[s].k);
        }
        group_of_sorted[s] = (int64_t)firsts.size() - 1;
    }
    const int64_t U = (int64_t)firsts.size();
    std::vector<int64_t> order(U);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(),
              [&](int64_t x, int64_t y) { return firsts[x] < firsts[y]; });
    b.ukey_face.assign(U, 0);
    for (int64_t r = 0; r < U; r++) b.ukey_face[order[r]] = r;
    b.F = U;
    b.cf.assign(M, 0);
    for (int64_t s = 0; s < M; s++) b.cf[sorted[s].e] = b.ukey_face[group_of_sorted[s]];
    b.fv_off.assign(b.F + 1, 0);
    b.fshape.assign(b.F, 0);
    std::vector<int64_t> fv4((size_t)b.F * 4, -1);
    for (int64_t r = 0; r < U; r++) {
        int64_t e = firsts[order[r]];  // vertices as listed by the first cell
        b.fshape[r] = entry_shape[e];
        for (int q = 0; q < 4; q++) fv4[4 * r + q] = entry_ids[4 * e + q];
    }
    std::vector<Entry>().swap(entries);
    std::vector<Entry>().swap(sorted);

    auto find_face = [&](const int64_t* ids, int n) -> int64_t {
        Key k = make_key(ids, n);
        auto it = std::lower_bound(b.ukeys.begin(), b.ukeys.end(), k);
        if (it == b.ukeys.end() || !(*it == k)) return -1;
        return b.ukey_face[it - b.ukeys.begin()];
    };

    // ---- boundaries and ghost ids (GR/grid.h:406-443), tags in sorted order ----
    const int ntags = io->ntags;
    std::vector<std::vector<int64_t>> tag_faces(ntags), tag_ghosts(ntags);
    std::vector<std::pair<int64_t, int64_t>> ghost_map;
    std::vector<char> has_ghost(b.F, 0);
    for (int t = 0; t < ntags; t++) {
        for (int64_t e = 0; e < io->nb; e++) {
            if (io->btag[e] != t) continue;
            int n = nverts(io->btype[e]);
            if (n < 0) return ibis::fail(IBIS_ERR_GRID, "unsupported boundary element");
            int64_t f = find_face(io->bvtx + 4 * e, n);
            if (f < 0) return ibis::fail(IBIS_ERR_GRID, "boundary element is not a face of the grid");
            tag_faces[t].push_back(f);
            if (io->tag_ghost[t]) {
                int64_t gid = b.N + b.G;
                b.G++;
                tag_ghosts[t].push_back(gid);
                if (!has_ghost[f]) {
                    has_ghost[f] = 1;
                    ghost_map.push_back({f, gid});
                }
            }
        }
    }
    std::sort(ghost_map.begin(), ghost_map.end());
    const int64_t N = b.N, G = b.G, F = b.F, Nt = N + G;

    // ---- geometry (reference numbering) ----
    std::vector<double> geom((size_t)13 * F, 0.0);
    double* nx = &geom[0 * F];
    double* ny = &geom[1 * F];
    double* nz = &geom[2 * F];
    double* t1x = &geom[3 * F];
    double* t1y = &geom[4 * F];
    double* t1z = &geom[5 * F];
    double* t2x = &geom[6 * F];
    double* t2y = &geom[7 * F];
    double* t2z = &geom[8 * F];
    double* fcx = &geom[9 * F];
    double* fcy = &geom[10 * F];
    double* fcz = &geom[11 * F];
    double* area = &geom[12 * F];
    for (int64_t f = 0; f < F; f++) {
        const int64_t* v = &fv4[4 * f];
        int n = nverts(b.fshape[f]);
        // Interfaces::compute_centres, GR/interface.h:301-323
        double x = 0.0, y = 0.0, z = 0.0;
        for (int q = 0; q < n; q++) {
            x += b.vx[v[q]];
            y += b.vy[v[q]];
            z += b.vz[v[q]];
        }
        fcx[f] = x / (double)n;
        fcy[f] = y / (double)n;
        fcz[f] = z / (double)n;
        // Interfaces::compute_areas, GR/interface.h:253-299
        if (b.fshape[f] == LINE) {
            double dx = b.vx[v[1]] - b.vx[v[0]], dy = b.vy[v[1]] - b.vy[v[0]],
                   dz = b.vz[v[1]] - b.vz[v[0]];
            area[f] = std::sqrt(dx * dx + dy * dy + dz * dz);
        } else if (b.fshape[f] == TRI) {
            area[f] = tri_area(b, v[0], v[1], v[2]);
        } else {
            area[f] = quad_area(b, v[0], v[1], v[2], v[3]);
        }
        // Interfaces::compute_orientations, GR/interface.h:180-251
        double x0 = b.vx[v[0]], x1 = b.vx[v[1]], y0 = b.vy[v[0]], y1 = b.vy[v[1]];
        double z0 = b.vz[v[0]], z1 = b.vz[v[1]];
        double il = 1. / std::sqrt((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0) +
                                   (z1 - z0) * (z1 - z0));
        t1x[f] = il * (x1 - x0);
        t1y[f] = il * (y1 - y0);
        t1z[f] = il * (z1 - z0);
        if (b.fshape[f] == LINE) {
            t2x[f] = 0.0;
            t2y[f] = 0.0;
            t2z[f] = 1.0;
        } else {
            double sx = b.vx[v[2]] - b.vx[v[0]], sy = b.vy[v[2]] - b.vy[v[0]],
                   sz = b.vz[v[2]] - b.vz[v[0]];
            double proj = t1x[f] * sx + t1y[f] * sy + t1z[f] * sz;
            double ux = sx - proj * t1x[f], uy = sy - proj * t1y[f], uz = sz - proj * t1z[f];
            double il2 = 1.0 / std::sqrt(ux * ux + uy * uy + uz * uz);
            t2x[f] = il2 * ux;
            t2y[f] = il2 * uy;
            t2z[f] = il2 * uz;
        }
        nx[f] = t1y[f] * t2z[f] - t1z[f] * t2y[f];  // cross, UT/vector3.h:153-160
        ny[f] = t1z[f] * t2x[f] - t1x[f] * t2z[f];
        nz[f] = t1x[f] * t2y[f] - t1y[f] * t2x[f];
    }
    std::vector<double> cen((size_t)3 * Nt, 0.0), vol(Nt, 0.0);
    double* cx = &cen[0];
    double* cy = &cen[Nt];
    double* cz = &cen[2 * Nt];
    for (int64_t c = 0; c < N; c++) {  // Cells::compute_centroids, GR/cell.h:228-255
        double x = 0.0, y = 0.0, z = 0.0;
        int64_t n = b.cv_off[c + 1] - b.cv_off[c];
        for (int64_t q = b.cv_off[c]; q < b.cv_off[c + 1]; q++) {
            x += b.vx[b.cv[q]];
            y += b.vy[b.cv[q]];
            z += b.vz[b.cv[q]];
        }
        cx[c] = x / (double)n;
        cy[c] = y / (double)n;
        cz[c] = z / (double)n;
    }
    const int64_t NONE = -1;
    std::vector<int64_t> L(F, NONE), R(F, NONE);
    std::vector<int32_t> outsign(M, 0);
    for (int64_t c = 0; c < N; c++) {  // compute_interface_connectivity, GR/grid.h:180-232
        for (int64_t j = b.cf_off[c]; j < b.cf_off[c + 1]; j++) {
            int64_t f = b.cf[j];
            double dx = fcx[f] - cx[c], dy = fcy[f] - cy[c], dz = fcz[f] - cz[c];
            double dot = dx * nx[f] + dy * ny[f] + dz * nz[f];
            if (dot > 0.0) {
                L[f] = c;
                outsign[j] = 1;
            } else {
                R[f] = c;
                outsign[j] = -1;
            }
        }
    }
    for (auto& fg : ghost_map) {
        if (L[fg.first] == NONE)
            L[fg.first] = fg.second;
        else
            R[fg.first] = fg.second;
    }
    for (int64_t f = 0; f < F; f++)
        if (L[f] == NONE || R[f] == NONE)
            return ibis::fail(IBIS_ERR_GRID,
                              "a face has no cell on one side (boundary without a ghost tag?)");
    for (int64_t c = 0; c < N; c++) {  // Cells::compute_volumes, GR/cell.h:257-300
        const int64_t* v = &b.cv[b.cv_off[c]];
        if (b.cshape[c] == TRI) {
            vol[c] = tri_area(b, v[0], v[1], v[2]);
        } else if (b.cshape[c] == QUAD) {
            vol[c] = quad_area(b, v[0], v[1], v[2], v[3]);
        } else {  // GR/geom.h:67-83
            double volume = 0.0;
            for (int64_t j = b.cf_off[c]; j < b.cf_off[c + 1]; j++) {
                int64_t f = b.cf[j];
                int out = outsign[j];
                double dot = fcx[f] * out * nx[f] + fcy[f] * out * ny[f] + fcz[f] * out * nz[f];
                volume += dot * area[f];
            }
            vol[c] = volume / 3.0;
        }
    }
    std::vector<int64_t> nbr(M);
    for (int64_t c = 0; c < N; c++)  // compute_cell_neighbours, GR/grid.h:340-358
        for (int64_t j = b.cf_off[c]; j < b.cf_off[c + 1]; j++) {
            int64_t f = b.cf[j];
            nbr[j] = (L[f] == c) ? R[f] : L[f];
        }
    for (int t = 0; t < ntags; t++)  // compute_ghost_cell_centres, GR/grid.h:364-403
        for (int64_t f : tag_faces[t]) {
            int64_t valid = L[f] < N ? L[f] : R[f];
            int64_t ghost = L[f] < N ? R[f] : L[f];
            if (ghost < N) continue;
            double fx = fcx[f], fy = fcy[f], fz = fcz[f];
            double dx = fx - cx[valid], dy = fy - cy[valid], dz = fz - cz[valid];
            cx[ghost] = fx + dx;
            cy[ghost] = fy + dy;
            cz[ghost] = fz + dz;
        }
    const int nl = b.dim == 2 ? 3 : 6;
    std::vector<double> lsq((size_t)nl * N, 0.0);
    for (int64_t i = 0; i < N; i++) {  // WLSGradient::compute_weights, GR/gradient.h:119-162
        double sxx = 0.0, sxy = 0.0, sxz = 0.0, syy = 0.0, syz = 0.0, szz = 0.0;
        for (int64_t j = b.cf_off[i]; j < b.cf_off[i + 1]; j++) {
            int64_t n = nbr[j];
            double dx = cx[n] - cx[i], dy = cy[n] - cy[i], dz = cz[n] - cz[i];
            sxx += dx * dx;
            sxy += dx * dy;
            sxz += dx * dz;
            syy += dy * dy;
            syz += dy * dz;
            szz += dz * dz;
        }
        double r11 = std::sqrt(sxx);
        double r12 = 1.0 / r11 * sxy;
        double r22 = std::sqrt(syy - r12 * r12);
        double r13 = 1.0 / r11 * sxz;
        double r23 = 1.0 / r22 * (syz - r12 / r11 * sxz);
        double r33 = std::sqrt(szz - (r13 * r13 + r23 * r23));
        lsq[0 * N + i] = r11;
        lsq[1 * N + i] = r12;
        lsq[2 * N + i] = r22;
        if (nl == 6) {
            lsq[3 * N + i] = r13;
            lsq[4 * N + i] = r23;
            lsq[5 * N + i] = r33;
        }
    }

    // ---- locality renumbering ----
    std::vector<int64_t> cnew(Nt), fnew(F);
    std::iota(cnew.begin(), cnew.end(), 0);
    std::iota(fnew.begin(), fnew.end(), 0);
    if (reorder) {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int64_t c = 0; c < N; c++) {
            const double p[3] = {cx[c], cy[c], cz[c]};
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], p[a]);
                hi[a] = std::max(hi[a], p[a]);
            }
        }
        std::vector<std::pair<uint64_t, int64_t>> code(N);
        const int bits = b.dim == 3 ? 21 : 32;
        const double scale = (double)((1ull << bits) - 1);
        for (int64_t c = 0; c < N; c++) {
            const double p[3] = {cx[c], cy[c], cz[c]};
            uint64_t q[3];
            for (int a = 0; a < 3; a++) {
                double w = hi[a] > lo[a] ? (p[a] - lo[a]) / (hi[a] - lo[a]) : 0.0;
                q[a] = (uint64_t)std::llround(std::min(1.0, std::max(0.0, w)) * scale);
            }
            uint64_t m = b.dim == 3
                             ? (spread_bits3(q[0]) | spread_bits3(q[1]) << 1 | spread_bits3(q[2]) << 2)
                             : (spread_bits2(q[0]) | spread_bits2(q[1]) << 1);
            code[c] = {m, c};
        }
        std::sort(code.begin(), code.end());
        for (int64_t k = 0; k < N; k++) cnew[code[k].second] = k;
        // ghosts: after valid cells, in order of their valid neighbour
        std::vector<std::pair<int64_t, int64_t>> gkey;
        for (int64_t f = 0; f < F; f++) {
            if (L[f] >= N) gkey.push_back({cnew[R[f]], L[f]});
            if (R[f] >= N) gkey.push_back({cnew[L[f]], R[f]});
        }
        std::sort(gkey.begin(), gkey.end());
        int64_t next = N;
        std::vector<char> seen(G, 0);
        for (auto& k : gkey)
            if (!seen[k.second - N]) {
                seen[k.second - N] = 1;
                cnew[k.second] = next++;
            }
        for (int64_t gcell = N; gcell < Nt; gcell++)
            if (!seen[gcell - N]) cnew[gcell] = next++;
        // faces: by (owner, other side) with owner the smaller renumbered side
        std::vector<std::pair<std::pair<int64_t, int64_t>, int64_t>> fkey(F);
        for (int64_t f = 0; f < F; f++) {
            int64_t a = cnew[L[f]], c2 = cnew[R[f]];
            fkey[f] = {{std::min(a, c2), std::max(a, c2)}, f};
        }
        std::sort(fkey.begin(), fkey.end());
        for (int64_t k = 0; k < F; k++) fnew[fkey[k].second] = k;
    }

    ibis_grid* g = new ibis_grid();
    g->dim = b.dim;
    g->N = N;
    g->G = G;
    g->F = F;
    g->ntags = ntags;
    g->n_lsq = nl;
    g->cell_new_of_old = cnew;
    g->face_new_of_old = fnew;
    g->geom.assign((size_t)13 * F, 0.0);
    for (int k = 0; k < 13; k++)
        for (int64_t f = 0; f < F; f++) g->geom[(size_t)k * F + fnew[f]] = geom[(size_t)k * F + f];
    g->left.assign(F, 0);
    g->right.assign(F, 0);
    for (int64_t f = 0; f < F; f++) {
        g->left[fnew[f]] = cnew[L[f]];
        g->right[fnew[f]] = cnew[R[f]];
    }
    std::vector<int64_t> cold(Nt);
    for (int64_t c = 0; c < Nt; c++) cold[cnew[c]] = c;
    g->cf_off.assign(N + 1, 0);
    g->cf.resize(M);
    g->nbr.resize(M);
    g->outsign.resize(M);
    for (int64_t k = 0; k < N; k++) {
        int64_t c = cold[k];
        int64_t n = b.cf_off[c + 1] - b.cf_off[c];
        g->cf_off[k + 1] = g->cf_off[k] + n;
        for (int64_t q = 0; q < n; q++) {
            int64_t j = b.cf_off[c] + q;
            g->cf[g->cf_off[k] + q] = fnew[b.cf[j]];
            g->nbr[g->cf_off[k] + q] = cnew[nbr[j]];
            g->outsign[g->cf_off[k] + q] = outsign[j];
        }
    }
    g->centroid.assign((size_t)3 * Nt, 0.0);
    g->volume.assign(Nt, 0.0);
    for (int64_t c = 0; c < Nt; c++) {
        for (int a = 0; a < 3; a++) g->centroid[(size_t)a * Nt + cnew[c]] = cen[(size_t)a * Nt + c];
        g->volume[cnew[c]] = vol[c];
    }
    g->lsq.assign((size_t)nl * N, 0.0);
    for (int a = 0; a < nl; a++)
        for (int64_t c = 0; c < N; c++) g->lsq[(size_t)a * N + cnew[c]] = lsq[(size_t)a * N + c];
    g->tag_off.assign(ntags + 1, 0);
    for (int t = 0; t < ntags; t++) {
        g->tag_off[t + 1] = g->tag_off[t] + (int64_t)tag_faces[t].size();
        for (size_t i = 0; i < tag_faces[t].size(); i++) {
            g->bfaces.push_back(fnew[tag_faces[t][i]]);
            g->ghosts.push_back(i < tag_ghosts[t].size() ? cnew[tag_ghosts[t][i]] : -1);
        }
    }
    *out = g;
    return IBIS_OK;
}

void ibis_grid_destroy(ibis_grid* g) { delete g; }

ibis_status ibis_grid_sizes(const ibis_grid* g, int64_t* out) {
    if (!g || !out) return ibis::fail(IBIS_ERR_INVALID, "ibis_grid_sizes: null argument");
    out[0] = g->dim;
    out[1] = g->N;
    out[2] = g->G;
    out[3] = g->F;
    out[4] = (int64_t)g->cf.size();
    out[5] = g->n_lsq;
    out[6] = g->ntags;
    out[7] = (int64_t)g->bfaces.size();
    return IBIS_OK;
}

ibis_status ibis_grid_permutation(const ibis_grid* g, int64_t* cell_new_of_old,
                                  int64_t* face_new_of_old) {
    if (!g) return ibis::fail(IBIS_ERR_INVALID, "ibis_grid_permutation: null grid");
    if (cell_new_of_old)
        std::memcpy(cell_new_of_old, g->cell_new_of_old.data(), g->cell_new_of_old.size() * 8);
    if (face_new_of_old)
        std::memcpy(face_new_of_old, g->face_new_of_old.data(), g->face_new_of_old.size() * 8);
    return IBIS_OK;
}

ibis_status ibis_grid_export(const ibis_grid* g, double* face_geom, int64_t* face_lr,
                             int64_t* cell_off, int64_t* cell_faces, int32_t* outsigns,
                             int64_t* neighbours, double* centroid, double* volume, double* lsq,
                             int64_t* tag_off, int64_t* bfaces, int64_t* ghosts) {
    if (!g) return ibis::fail(IBIS_ERR_INVALID, "ibis_grid_export: null grid");
    auto cp = [](void* dst, const void* src, size_t bytes) {
        if (dst && bytes) std::memcpy(dst, src, bytes);
    };
    cp(face_geom, g->geom.data(), g->geom.size() * 8);
    if (face_lr) {
        cp(face_lr, g->left.data(), g->F * 8);
        cp(face_lr + g->F, g->right.data(), g->F * 8);
    }
    cp(cell_off, g->cf_off.data(), g->cf_off.size() * 8);
    cp(cell_faces, g->cf.data(), g->cf.size() * 8);
    cp(outsigns, g->outsign.data(), g->outsign.size() * 4);
    cp(neighbours, g->nbr.data(), g->nbr.size() * 8);
    cp(centroid, g->centroid.data(), g->centroid.size() * 8);
    cp(volume, g->volume.data(), g->volume.size() * 8);
    cp(lsq, g->lsq.data(), g->lsq.size() * 8);
    cp(tag_off, g->tag_off.data(), g->tag_off.size() * 8);
    cp(bfaces, g->bfaces.data(), g->bfaces.size() * 8);
    cp(ghosts, g->ghosts.data(), g->ghosts.size() * 8);
    return IBIS_OK;
}

void ibis_grid_fill_desc(const ibis_grid* g, ibis_mesh_desc* d) {
    d->dim = g->dim;
    d->ntags = g->ntags;
    d->n_cells = g->N;
    d->n_ghost = g->G;
    d->n_faces = g->F;
    d->face_left = g->left.data();
    d->face_right = g->right.data();
    d->face_geom = g->geom.data();
    d->cell_face_off = g->cf_off.data();
    d->cell_faces = g->cf.data();
    d->cell_outsigns = g->outsign.data();
    d->cell_neighbours = g->nbr.data();
    d->centroid = g->centroid.data();
    d->volume = g->volume.data();
    d->lsq = g->lsq.data();
    d->tag_off = g->tag_off.data();
    d->bfaces = g->bfaces.data();
}
