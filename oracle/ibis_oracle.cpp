// ibis_oracle.cpp -- CPU restatement of ibis's cell residual (dU/dt).
//
// TEST INFRASTRUCTURE ONLY: the checker for the MI355X path and the timed
// CPU baseline.  Nothing in the product links or loads this file.
//
// Each function restates the reference loop by loop and keeps the
// reference's operation order, so that on identical inputs it reproduces the
// reference's IEEE double arithmetic.  The reference itself cannot be built
// here (its Kokkos, nlohmann::json and spdlog submodules are empty), so this
// restatement is pinned by the reference's own known-answer tests and by
// invariants (see tests/test_oracle_*.py).
//
// Citations are relative to the reference repository root.
//   FV/  = src/finite_volume/finite_volume/
//   GR/  = src/grid/grid/
//   GAS/ = src/gas/gas/
//   UT/  = src/util/util/
//   SO/  = src/solvers/solvers/

#include "ibis_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// Forward-mode dual number, UT/dual.h:10-346 (Ibis::Dual<double>).
// Comparisons, min, max and abs branch on the real part.
// ---------------------------------------------------------------------------
struct Dual {
    double r, d;
    Dual() : r(0.0), d(0.0) {}
    Dual(double re) : r(re), d(0.0) {}
    Dual(double re, double du) : r(re), d(du) {}
    Dual& operator=(double re) {
        r = re;
        d = 0.0;
        return *this;
    }
    friend bool operator<(const Dual& a, const Dual& b) { return a.r < b.r; }
    friend bool operator>(const Dual& a, const Dual& b) { return a.r > b.r; }
    friend bool operator<=(const Dual& a, const Dual& b) { return a.r <= b.r; }
    friend bool operator>=(const Dual& a, const Dual& b) { return a.r >= b.r; }
    friend bool operator==(const Dual& a, const Dual& b) { return a.r == b.r; }
    friend bool operator!=(const Dual& a, const Dual& b) { return a.r != b.r; }
    friend Dual operator+(const Dual& a, const Dual& b) { return Dual(a.r + b.r, a.d + b.d); }
    friend Dual operator+(const Dual& a, double b) { return Dual(a.r + b, a.d); }
    friend Dual operator+(double a, const Dual& b) { return Dual(a + b.r, b.d); }
    Dual& operator+=(const Dual& o) {
        r += o.r;
        d += o.d;
        return *this;
    }
    friend Dual operator-(const Dual& a, const Dual& b) { return Dual(a.r - b.r, a.d - b.d); }
    friend Dual operator-(const Dual& a, double b) { return Dual(a.r - b, a.d); }
    friend Dual operator-(double a, const Dual& b) { return Dual(a - b.r, -b.d); }
    Dual& operator-=(const Dual& o) {
        r -= o.r;
        d -= o.d;
        return *this;
    }
    Dual operator-() const { return Dual(-r, -d); }
    friend Dual operator*(const Dual& a, const Dual& b) {
        return Dual(a.r * b.r, a.r * b.d + a.d * b.r);
    }
    friend Dual operator*(const Dual& a, double b) { return Dual(a.r * b, a.d * b); }
    friend Dual operator*(double a, const Dual& b) { return Dual(b.r * a, b.d * a); }
    friend Dual operator/(const Dual& n, const Dual& m) {
        return Dual(n.r / m.r, (n.d * m.r - n.r * m.d) / (m.r * m.r));
    }
    friend Dual operator/(const Dual& n, double m) { return Dual(n.r / m, n.d / m); }
    friend Dual operator/(double n, const Dual& m) {
        return Dual(n / m.r, -m.d * n / (m.r * m.r));
    }
};

// Kokkos::min/max on arithmetic types: (b < a) ? b : a and (a < b) ? b : a.
inline double kmin(double a, double b) { return (b < a) ? b : a; }
inline double kmax(double a, double b) { return (a < b) ? b : a; }
inline Dual kmin(const Dual& a, const Dual& b) { return a.r < b.r ? a : b; }
inline Dual kmax(const Dual& a, const Dual& b) { return a.r > b.r ? a : b; }
inline Dual kmin(double a, const Dual& b) { return a < b.r ? Dual(a) : b; }
inline Dual kmax(double a, const Dual& b) { return a > b.r ? Dual(a) : b; }
inline Dual kmin(const Dual& a, double b) { return a.r < b ? a : Dual(b); }
inline Dual kmax(const Dual& a, double b) { return a.r > b ? a : Dual(b); }
// Venkatakrishnan running minimum as a fraction (bn, bd): a face ratio
// num / den (den > 0) replaces it when num * bd < bn * den; a Dual candidate
// also on a tie, as kmin does.
inline bool venkat_take(double num, double den, double bn, double bd) { return num * bd < bn * den; }
inline bool venkat_take(const Dual& num, const Dual& den, const Dual& bn, const Dual& bd) {
    return !(bn.r * den.r < num.r * bd.r);
}
inline double kabs(double x) { return std::fabs(x); }
inline Dual kabs(const Dual& x) { return (x.r < 0.0) ? -x : x; }
inline double ksqrt(double x) { return std::sqrt(x); }
inline Dual ksqrt(const Dual& x) {
    double re = std::sqrt(x.r);
    return Dual(re, x.d / (2.0 * re));
}
inline double kpow(double b, double p) { return std::pow(b, p); }
inline Dual kpow(const Dual& b, const Dual& p) {
    double re = std::pow(b.r, p.r);
    double du = b.d * p.r * std::pow(b.r, p.r - 1.0) + p.d * re * std::log(b.r);
    return Dual(re, du);
}
inline double kfloor(double x) { return std::floor(x); }
// UT/dual.h:299-302: floor is applied to the dual part as well.
inline Dual kfloor(const Dual& x) { return Dual(std::floor(x.r), std::floor(x.d)); }
inline double kcopysign(double m, double s) { return std::copysign(m, s); }
// UT/dual.h:309-314
inline Dual kcopysign(const Dual& m, const Dual& s) {
    double re = std::copysign(m.r, s.r);
    double du = (re - m.r) < 1e-15 ? m.d : -m.d;
    return Dual(re, du);
}
inline double real_part(double x) { return x; }
inline double real_part(const Dual& x) { return x.r; }
inline bool kisnan(double x) { return std::isnan(x); }
inline bool kisinf(double x) { return std::isinf(x); }

// rows of the primitive array, GAS/gas_state.h:40-46 and the velocity vector
enum { P_RHO = 0, P_P = 1, P_T = 2, P_E = 3, P_VX = 4, P_VY = 5, P_VZ = 6, NPRIM = 7 };
// gradient / limiter slots (GR/gradient.h:36-42, FV/limiter.h:15-44)
enum { G_P = 0, G_RHO = 1, G_T = 2, G_U = 3, G_VX = 4, G_VY = 5, G_VZ = 6, NGRAD = 7 };

// ---------------------------------------------------------------------------
// Grid (GR/grid.h:89-167 and helpers)
// ---------------------------------------------------------------------------
enum ElemType { LINE = 3, TRI = 5, QUAD = 9, TETRA = 10, HEX = 12, WEDGE = 13, PYRAMID = 14 };

struct FaceKey {
    int64_t v[4];
    int n;
    bool operator==(const FaceKey& o) const {
        if (n != o.n) return false;
        for (int i = 0; i < n; i++)
            if (v[i] != o.v[i]) return false;
        return true;
    }
};
struct FaceKeyHash {
    size_t operator()(const FaceKey& k) const {
        uint64_t h = 1469598103934665603ull ^ (uint64_t)k.n;
        for (int i = 0; i < k.n; i++) {
            h ^= (uint64_t)k.v[i] + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
        }
        return (size_t)h;
    }
};

// InterfaceLookup::hash_vertex_ids sorts the ids in descending order
// (GR/interface.cpp:38-46); the lookup only needs set identity.
FaceKey make_key(const int64_t* ids, int n) {
    FaceKey k;
    k.n = n;
    for (int i = 0; i < n; i++) k.v[i] = ids[i];
    std::sort(k.v, k.v + n, [](int64_t a, int64_t b) { return a > b; });
    return k;
}

// vtk_face_order, GR/grid_io.cpp:298-348
int vtk_faces(int type, const int64_t* id, int64_t faces[6][4], int nfv[6], int ftype[6]) {
    auto set = [&](int f, int t, std::initializer_list<int64_t> l) {
        int j = 0;
        for (int64_t x : l) faces[f][j++] = x;
        nfv[f] = j;
        ftype[f] = t;
    };
    switch (type) {
        case LINE:
            set(0, LINE, {id[0], id[1]});
            return 1;
        case TRI:
            set(0, LINE, {id[0], id[1]});
            set(1, LINE, {id[1], id[2]});
            set(2, LINE, {id[2], id[0]});
            return 3;
        case QUAD:
            set(0, LINE, {id[0], id[1]});
            set(1, LINE, {id[1], id[2]});
            set(2, LINE, {id[2], id[3]});
            set(3, LINE, {id[3], id[0]});
            return 4;
        case TETRA:
            set(0, TRI, {id[0], id[1], id[2]});
            set(1, TRI, {id[0], id[1], id[3]});
            set(2, TRI, {id[1], id[2], id[3]});
            set(3, TRI, {id[0], id[2], id[3]});
            return 4;
        case HEX:
            set(0, QUAD, {id[0], id[1], id[2], id[3]});
            set(1, QUAD, {id[0], id[1], id[5], id[4]});
            set(2, QUAD, {id[4], id[5], id[6], id[7]});
            set(3, QUAD, {id[2], id[3], id[7], id[6]});
            set(4, QUAD, {id[0], id[4], id[7], id[3]});
            set(5, QUAD, {id[1], id[5], id[6], id[2]});
            return 6;
        case WEDGE:
            set(0, TRI, {id[0], id[1], id[2]});
            set(1, TRI, {id[3], id[5], id[4]});
            set(2, QUAD, {id[1], id[4], id[5], id[2]});
            set(3, QUAD, {id[0], id[2], id[5], id[3]});
            set(4, QUAD, {id[0], id[3], id[4], id[1]});
            return 5;
        case PYRAMID:
            set(0, QUAD, {id[0], id[3], id[2], id[1]});
            set(1, TRI, {id[2], id[3], id[4]});
            set(2, TRI, {id[0], id[4], id[3]});
            set(3, TRI, {id[0], id[1], id[4]});
            set(4, TRI, {id[1], id[2], id[4]});
            return 5;
    }
    return -1;
}

int n_vertices_of(int type) {
    switch (type) {
        case LINE: return 2;
        case TRI: return 3;
        case QUAD: return 4;
        case TETRA: return 4;
        case HEX: return 8;
        case WEDGE: return 6;
        case PYRAMID: return 5;
    }
    return 0;
}

const int64_t NONE = std::numeric_limits<int64_t>::max();

struct Grid {
    int dim = 2;
    int64_t N = 0, G = 0, F = 0;
    std::vector<double> vx, vy, vz;
    // faces
    std::vector<int64_t> fv_off, fv;  // face vertex ids
    std::vector<int> fshape;
    std::vector<int64_t> left, right;
    std::vector<double> area, nx, ny, nz, t1x, t1y, t1z, t2x, t2y, t2z, cx, cy, cz;
    // cells
    std::vector<int64_t> cv_off, cv;
    std::vector<int> cshape;
    std::vector<int64_t> cf_off, cf, nbr;
    std::vector<int> outsign;
    std::vector<double> x, y, z, vol;
    std::vector<double> r;  // n_lsq x N
    int n_lsq = 3;
    // boundaries, one entry per tag in sorted order
    int ntags = 0;
    std::vector<std::vector<int64_t>> bfaces, ghosts;
};

// Ibis::area_of_triangle, GR/geom.h:29-49
double area_of_triangle(const Grid& g, int64_t a, int64_t b, int64_t c) {
    double ab_x = g.vx[b] - g.vx[a];
    double ab_y = g.vy[b] - g.vy[a];
    double ab_z = g.vz[b] - g.vz[a];
    double ac_x = g.vx[c] - g.vx[a];
    double ac_y = g.vy[c] - g.vy[a];
    double ac_z = g.vz[c] - g.vz[a];
    double cross_x = ab_y * ac_z - ab_z * ac_y;
    double cross_y = ab_z * ac_x - ab_x * ac_z;
    double cross_z = ab_x * ac_y - ab_y * ac_x;
    return 0.5 * std::sqrt(cross_x * cross_x + cross_y * cross_y + cross_z * cross_z);
}

// GR/geom.h:56-62
double area_of_quadrilateral(const Grid& g, int64_t a, int64_t b, int64_t c, int64_t d) {
    double a1 = area_of_triangle(g, a, b, c);
    double a2 = area_of_triangle(g, a, c, d);
    return a1 + a2;
}

Grid* build_grid(const or_grid_io* io) {
    Grid* gp = new Grid();
    Grid& g = *gp;
    g.dim = io->dim;
    g.vx.resize(io->nv);
    g.vy.resize(io->nv);
    g.vz.resize(io->nv);
    for (int64_t i = 0; i < io->nv; i++) {
        g.vx[i] = io->xyz[3 * i];
        g.vy[i] = io->xyz[3 * i + 1];
        g.vz[i] = (io->dim == 3) ? io->xyz[3 * i + 2] : 0.0;  // GR/grid_io.cpp:175-199
    }

    // GR/grid.h:109-133: faces numbered in order of first appearance
    std::unordered_map<FaceKey, int64_t, FaceKeyHash> lookup;
    lookup.reserve((size_t)(io->nc * (io->dim == 3 ? 3.2 : 2.2)) + 16);
    g.N = io->nc;
    g.cv_off.push_back(0);
    g.cf_off.push_back(0);
    g.fv_off.push_back(0);
    for (int64_t c = 0; c < io->nc; c++) {
        int type = io->ctype[c];
        const int64_t* ids = io->cvtx + 8 * c;
        int nv = n_vertices_of(type);
        for (int j = 0; j < nv; j++) g.cv.push_back(ids[j]);
        g.cv_off.push_back((int64_t)g.cv.size());
        g.cshape.push_back(type);
        int64_t faces[6][4];
        int nfv[6], ftype[6];
        int nf = vtk_faces(type, ids, faces, nfv, ftype);
        for (int f = 0; f < nf; f++) {
            FaceKey k = make_key(faces[f], nfv[f]);
            auto it = lookup.find(k);
            int64_t fid;
            if (it == lookup.end()) {
                fid = (int64_t)lookup.size();
                lookup.emplace(k, fid);
                for (int j = 0; j < nfv[f]; j++) g.fv.push_back(faces[f][j]);
                g.fv_off.push_back((int64_t)g.fv.size());
                g.fshape.push_back(ftype[f]);
            } else {
                fid = it->second;
            }
            g.cf.push_back(fid);
        }
        g.cf_off.push_back((int64_t)g.cf.size());
    }
    g.F = (int64_t)g.fshape.size();

    // setup_boundaries, GR/grid.h:406-443 (tags iterated in sorted order)
    g.ntags = io->ntags;
    g.bfaces.assign(io->ntags, {});
    g.ghosts.assign(io->ntags, {});
    std::vector<std::pair<int64_t, int64_t>> ghost_map;  // face -> ghost, std::map semantics
    std::unordered_map<int64_t, int64_t> ghost_seen;
    g.G = 0;
    for (int t = 0; t < io->ntags; t++) {
        for (int64_t b = 0; b < io->nb; b++) {
            if (io->btag[b] != t) continue;
            int nfv = n_vertices_of(io->btype[b]);
            FaceKey k = make_key(io->bvtx + 4 * b, nfv);
            auto it = lookup.find(k);
            int64_t fid = (it == lookup.end()) ? NONE : it->second;
            g.bfaces[t].push_back(fid);
            if (io->tag_ghost[t]) {
                int64_t gid = g.N + g.G;
                g.G++;
                g.ghosts[t].push_back(gid);
                if (ghost_seen.find(fid) == ghost_seen.end()) {
                    ghost_seen.emplace(fid, gid);
                    ghost_map.push_back({fid, gid});
                }
            }
        }
    }
    std::sort(ghost_map.begin(), ghost_map.end());

    int64_t Nt = g.N + g.G;
    g.left.assign(g.F, NONE);
    g.right.assign(g.F, NONE);
    g.area.assign(g.F, 0.0);
    for (auto* v : {&g.nx, &g.ny, &g.nz, &g.t1x, &g.t1y, &g.t1z, &g.t2x, &g.t2y, &g.t2z,
                    &g.cx, &g.cy, &g.cz})
        v->assign(g.F, 0.0);
    g.x.assign(Nt, 0.0);
    g.y.assign(Nt, 0.0);
    g.z.assign(Nt, 0.0);
    g.vol.assign(Nt, 0.0);
    g.outsign.assign(g.cf.size(), 0);
    g.nbr.assign(g.cf.size(), 0);

    // Interfaces::compute_centres, GR/interface.h:301-323
    for (int64_t f = 0; f < g.F; f++) {
        double x = 0.0, y = 0.0, z = 0.0;
        int64_t nv = g.fv_off[f + 1] - g.fv_off[f];
        for (int64_t j = 0; j < nv; j++) {
            int64_t v = g.fv[g.fv_off[f] + j];
            x += g.vx[v];
            y += g.vy[v];
            z += g.vz[v];
        }
        g.cx[f] = x / (double)nv;
        g.cy[f] = y / (double)nv;
        g.cz[f] = z / (double)nv;
    }
    // Interfaces::compute_areas, GR/interface.h:253-299
    for (int64_t f = 0; f < g.F; f++) {
        const int64_t* v = &g.fv[g.fv_off[f]];
        switch (g.fshape[f]) {
            case LINE: {
                double dx = g.vx[v[1]] - g.vx[v[0]];
                double dy = g.vy[v[1]] - g.vy[v[0]];
                double dz = g.vz[v[1]] - g.vz[v[0]];
                g.area[f] = std::sqrt(dx * dx + dy * dy + dz * dz);  // GR/geom.h:12-26
                break;
            }
            case TRI: g.area[f] = area_of_triangle(g, v[0], v[1], v[2]); break;
            case QUAD: g.area[f] = area_of_quadrilateral(g, v[0], v[1], v[2], v[3]); break;
        }
    }
    // Interfaces::compute_orientations, GR/interface.h:180-251
    for (int64_t f = 0; f < g.F; f++) {
        const int64_t* v = &g.fv[g.fv_off[f]];
        double x0 = g.vx[v[0]], x1 = g.vx[v[1]];
        double y0 = g.vy[v[0]], y1 = g.vy[v[1]];
        double z0 = g.vz[v[0]], z1 = g.vz[v[1]];
        double il = 1. / std::sqrt((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0) +
                                   (z1 - z0) * (z1 - z0));
        g.t1x[f] = il * (x1 - x0);
        g.t1y[f] = il * (y1 - y0);
        g.t1z[f] = il * (z1 - z0);
        if (g.fshape[f] == LINE) {
            g.t2x[f] = 0.0;
            g.t2y[f] = 0.0;
            g.t2z[f] = 1.0;
        } else {
            double t1x = g.t1x[f], t1y = g.t1y[f], t1z = g.t1z[f];
            double t2xs = g.vx[v[2]] - g.vx[v[0]];
            double t2ys = g.vy[v[2]] - g.vy[v[0]];
            double t2zs = g.vz[v[2]] - g.vz[v[0]];
            double proj = t1x * t2xs + t1y * t2ys + t1z * t2zs;
            double t2x = t2xs - proj * t1x;
            double t2y = t2ys - proj * t1y;
            double t2z = t2zs - proj * t1z;
            double il2 = 1.0 / std::sqrt(t2x * t2x + t2y * t2y + t2z * t2z);
            g.t2x[f] = il2 * t2x;
            g.t2y[f] = il2 * t2y;
            g.t2z[f] = il2 * t2z;
        }
        // cross(tan1, tan2, norm), UT/vector3.h:153-160
        g.nx[f] = g.t1y[f] * g.t2z[f] - g.t1z[f] * g.t2y[f];
        g.ny[f] = g.t1z[f] * g.t2x[f] - g.t1x[f] * g.t2z[f];
        g.nz[f] = g.t1x[f] * g.t2y[f] - g.t1y[f] * g.t2x[f];
    }
    // Cells::compute_centroids, GR/cell.h:228-255 (vertex average)
    for (int64_t c = 0; c < g.N; c++) {
        double x = 0.0, y = 0.0, z = 0.0;
        int64_t nv = g.cv_off[c + 1] - g.cv_off[c];
        for (int64_t j = 0; j < nv; j++) {
            int64_t v = g.cv[g.cv_off[c] + j];
            x += g.vx[v];
            y += g.vy[v];
            z += g.vz[v];
        }
        g.x[c] = x / (double)nv;
        g.y[c] = y / (double)nv;
        g.z[c] = z / (double)nv;
    }
    // compute_interface_connectivity, GR/grid.h:180-232
    for (int64_t c = 0; c < g.N; c++) {
        for (int64_t j = g.cf_off[c]; j < g.cf_off[c + 1]; j++) {
            int64_t f = g.cf[j];
            double dx = g.cx[f] - g.x[c];
            double dy = g.cy[f] - g.y[c];
            double dz = g.cz[f] - g.z[c];
            double dot = dx * g.nx[f] + dy * g.ny[f] + dz * g.nz[f];
            if (dot > 0.0) {
                g.left[f] = c;
                g.outsign[j] = 1;
            } else {
                g.right[f] = c;
                g.outsign[j] = -1;
            }
        }
    }
    for (auto& fg : ghost_map) {
        int64_t f = fg.first;
        if (f == NONE) continue;
        if (g.left[f] == NONE)
            g.left[f] = fg.second;
        else
            g.right[f] = fg.second;
    }
    // Cells::compute_volumes, GR/cell.h:257-300 and GR/geom.h:67-83
    for (int64_t c = 0; c < g.N; c++) {
        const int64_t* v = &g.cv[g.cv_off[c]];
        switch (g.cshape[c]) {
            case TRI: g.vol[c] = area_of_triangle(g, v[0], v[1], v[2]); break;
            case QUAD: g.vol[c] = area_of_quadrilateral(g, v[0], v[1], v[2], v[3]); break;
            default: {
                double volume = 0.0;
                for (int64_t j = g.cf_off[c]; j < g.cf_off[c + 1]; j++) {
                    int64_t f = g.cf[j];
                    int out = g.outsign[j];
                    double dot = g.cx[f] * out * g.nx[f] + g.cy[f] * out * g.ny[f] +
                                 g.cz[f] * out * g.nz[f];
                    volume += dot * g.area[f];
                }
                g.vol[c] = volume / 3.0;
            }
        }
    }
    // compute_cell_neighbours, GR/grid.h:340-358
    for (int64_t c = 0; c < g.N; c++) {
        for (int64_t j = g.cf_off[c]; j < g.cf_off[c + 1]; j++) {
            int64_t f = g.cf[j];
            g.nbr[j] = (g.left[f] == c) ? g.right[f] : g.left[f];
        }
    }
    // compute_ghost_cell_centres, GR/grid.h:364-403
    for (int t = 0; t < g.ntags; t++) {
        for (int64_t f : g.bfaces[t]) {
            if (f == NONE) continue;
            int64_t l = g.left[f], r = g.right[f];
            int64_t valid, ghost;
            if (l < g.N) {
                valid = l;
                ghost = r;
            } else {
                valid = r;
                ghost = l;
            }
            if (ghost == NONE || ghost >= Nt) continue;
            double fx = g.cx[f], fy = g.cy[f], fz = g.cz[f];
            double dx = fx - g.x[valid];
            double dy = fy - g.y[valid];
            double dz = fz - g.z[valid];
            g.x[ghost] = fx + dx;
            g.y[ghost] = fy + dy;
            g.z[ghost] = fz + dz;
        }
    }
    // WLSGradient::compute_weights, GR/gradient.h:119-162
    g.n_lsq = (g.dim == 2) ? 3 : 6;
    g.r.assign((size_t)g.n_lsq * g.N, 0.0);
    for (int64_t i = 0; i < g.N; i++) {
        double sxx = 0.0, sxy = 0.0, sxz = 0.0, syy = 0.0, syz = 0.0, szz = 0.0;
        double xi = g.x[i], yi = g.y[i], zi = g.z[i];
        for (int64_t j = g.cf_off[i]; j < g.cf_off[i + 1]; j++) {
            int64_t n = g.nbr[j];
            double dx = g.x[n] - xi;
            double dy = g.y[n] - yi;
            double dz = g.z[n] - zi;
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
        g.r[0 * g.N + i] = r11;
        g.r[1 * g.N + i] = r12;
        g.r[2 * g.N + i] = r22;
        if (g.dim == 3) {
            g.r[3 * g.N + i] = r13;
            g.r[4 * g.N + i] = r23;
            g.r[5 * g.N + i] = r33;
        }
    }
    return gp;
}

// WLSGradient::compute_gradients, GR/gradient.h:66-116.
template <class T>
void lsq_gradient(const Grid& g, const T* values, T* gx, T* gy, T* gz) {
    const int64_t N = g.N;
    const int dim = g.dim;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; i++) {
        T grad_x = 0.0, grad_y = 0.0, grad_z = 0.0;
        T u_i = values[i];
        double r11 = g.r[0 * N + i], r12 = g.r[1 * N + i], r22 = g.r[2 * N + i];
        double r13 = 0.0, r23 = 0.0, r33 = 0.0;
        if (dim == 3) {
            r13 = g.r[3 * N + i];
            r23 = g.r[4 * N + i];
            r33 = g.r[5 * N + i];
        }
        double beta = (r12 * r23 - r13 * r23) / (r11 * r22);  // GR/gradient.h:89 (sic)
        double xi = g.x[i], yi = g.y[i], zi = g.z[i];
        for (int64_t j = g.cf_off[i]; j < g.cf_off[i + 1]; j++) {
            int64_t n = g.nbr[j];
            T diff_u = values[n] - u_i;
            double dx = g.x[n] - xi;
            double dy = g.y[n] - yi;
            double dz = g.z[n] - zi;
            double alpha_1 = dx / (r11 * r11);
            double alpha_2 = 1.0 / (r22 * r22) * (dy - r12 / r11 * dx);
            double alpha_3 = 0.0;
            if (dim == 3) alpha_3 = 1.0 / (r33 * r33) * (dz - r23 / r22 * dy + beta * dx);
            double w_1 = alpha_1 - r12 / r11 * alpha_2 + beta * alpha_3;
            double w_2 = alpha_2 - r23 / r22 * alpha_3;
            double w_3 = alpha_3;
            grad_x += w_1 * diff_u;
            grad_y += w_2 * diff_u;
            grad_z += w_3 * diff_u;
        }
        gx[i] = grad_x;
        gy[i] = grad_y;
        gz[i] = grad_z;
    }
}

// ---------------------------------------------------------------------------
// Finite volume (FV/finite_volume.cpp and friends), templated on the scalar.
// ---------------------------------------------------------------------------
template <class T>
struct FV {
    Grid* g;
    or_fv_config cfg;
    int64_t N, Nt, F;
    int ncons;
    // scratch, FV/finite_volume.h:114-141 and FV/convective_flux.h:68-87
    std::vector<T> flux;   // ncons x F
    std::vector<T> left, right;  // 7 x F
    std::vector<T> grad;   // NGRAD x 3 x N
    std::vector<T> lim;    // NGRAD x N
    std::vector<T> face_fs;  // 7 x F (viscous face states)
    std::vector<T> face_grad;  // 4 x 3 x F: face gradients of T, vx, vy, vz (viscous)

    FV(Grid* grid, const or_fv_config* c) : g(grid), cfg(*c) {
        N = g->N;
        Nt = g->N + g->G;
        F = g->F;
        ncons = g->dim + 2;
        flux.assign((size_t)ncons * F, T(0.0));
        left.assign((size_t)NPRIM * F, T(0.0));
        right.assign((size_t)NPRIM * F, T(0.0));
        grad.assign((size_t)NGRAD * 3 * N, T(0.0));
        lim.assign((size_t)NGRAD * N, T(0.0));
        face_fs.assign((size_t)NPRIM * F, T(0.0));
        face_grad.assign((size_t)12 * F, T(0.0));
    }
    T* gcomp(int var, int comp) { return &grad[((size_t)var * 3 + comp) * N]; }

    // cons index helpers, FV/conserved_quantities.cpp:16-23
    int i_mass() const { return 0; }
    int i_mx() const { return 1; }
    int i_my() const { return 2; }
    int i_mz() const { return 3; }
    int i_en() const { return 1 + g->dim; }

    // --- ghost cell actions, FV/boundaries/boundary.cpp ---------------------
    void ghost_valid(int64_t f, int64_t& ghost, int64_t& valid) const {
        int64_t l = g->left[f], r = g->right[f];
        if (l < N) {
            ghost = r;
            valid = l;
        } else {
            ghost = l;
            valid = r;
        }
    }
    void apply_action(int kind, const double* par, const std::vector<int64_t>& faces, T* fs) {
        T* rho = fs + P_RHO * Nt;
        T* p = fs + P_P * Nt;
        T* temp = fs + P_T * Nt;
        T* e = fs + P_E * Nt;
        T* vx = fs + P_VX * Nt;
        T* vy = fs + P_VY * Nt;
        T* vz = fs + P_VZ * Nt;
        const double R = cfg.R, Cv = cfg.Cv, gamma = cfg.gamma;
        for (size_t i = 0; i < faces.size(); i++) {
            int64_t f = faces[i], gh, va;
            ghost_valid(f, gh, va);
            switch (kind) {
                case OR_BC_FLOW_STATE_COPY:  // boundary.cpp:15-45
                    temp[gh] = par[0];
                    p[gh] = par[1];
                    rho[gh] = par[2];
                    e[gh] = par[3];
                    vx[gh] = par[4];
                    vy[gh] = par[5];
                    vz[gh] = par[6];
                    break;
                case OR_BC_INTERNAL_COPY:  // boundary.cpp:100-136
                    temp[gh] = temp[va];
                    p[gh] = p[va];
                    rho[gh] = rho[va];
                    e[gh] = e[va];
                    vx[gh] = vx[va];
                    vy[gh] = vy[va];
                    vz[gh] = vz[va];
                    break;
                case OR_BC_REFLECT_NORMAL: {  // boundary.cpp:142-206
                    temp[gh] = temp[va];
                    p[gh] = p[va];
                    rho[gh] = rho[va];
                    e[gh] = e[va];
                    T x = vx[va], y = vy[va], z = vz[va];
                    double nx = g->nx[f], ny = g->ny[f], nz = g->nz[f];
                    double t1x = g->t1x[f], t1y = g->t1y[f], t1z = g->t1z[f];
                    double t2x = g->t2x[f], t2y = g->t2y[f], t2z = g->t2z[f];
                    T x_star = -(x * nx + y * ny + z * nz);
                    T y_star = x * t1x + y * t1y + z * t1z;
                    T z_star = x * t2x + y * t2y + z * t2z;
                    vx[gh] = x_star * nx + y_star * t1x + z_star * t2x;
                    vy[gh] = x_star * ny + y_star * t1y + z_star * t2y;
                    vz[gh] = x_star * nz + y_star * t1z + z_star * t2z;
                    break;
                }
                case OR_BC_VEL_REFLECT:  // boundary.cpp:209-241
                    vx[gh] = -vx[va];
                    vy[gh] = -vy[va];
                    vz[gh] = -vz[va];
                    break;
                case OR_BC_FIX_TEMPERATURE:  // boundary.cpp:244-277
                    temp[gh] = 2 * T(par[0]) - temp[va];
                    break;
                case OR_BC_SUBSONIC_INFLOW: {  // boundary.cpp:286-342
                    double nx = g->nx[f], ny = g->ny[f], nz = g->nz[f];
                    T in_vx = par[4], in_vy = par[5], in_vz = par[6];
                    T vel_correct = nx * (in_vx - vx[va]) + ny * (in_vy - vy[va]) +
                                    nz * (in_vz - vz[va]);
                    T a = ksqrt(T(gamma * R) * temp[va]);
                    T rho_ref = rho[va];
                    T p_in = par[1];
                    T p_face = 0.5 * (p_in + p[va] - rho_ref * a * vel_correct);
                    T rho_face = T(par[2]) + (p_face - p_in) / a / a;
                    T vx_face = in_vx - nx * (p_in - p_face) / (rho_ref * a);
                    T vy_face = in_vy - ny * (p_in - p_face) / (rho_ref * a);
                    T vz_face = in_vz - nz * (p_in - p_face) / (rho_ref * a);
                    p[gh] = 2 * p_face - p[va];
                    rho[gh] = 2 * rho_face - rho[va];
                    temp[gh] = p[gh] / (rho[gh] * R);  // update_thermo_from_rhop
                    e[gh] = Cv * temp[gh];
                    vx[gh] = 2 * vx_face - vx[va];
                    vy[gh] = 2 * vy_face - vy[va];
                    vz[gh] = 2 * vz_face - vz[va];
                    break;
                }
                case OR_BC_SUBSONIC_OUTFLOW: {  // boundary.cpp:348-399
                    double nx = g->nx[f], ny = g->ny[f], nz = g->nz[f];
                    T p_face = par[0];
                    T rho_ref = rho[va];
                    T p_ref = p[va];
                    T a = ksqrt(T(gamma * R) * temp[va]);
                    T rho_face = rho[va] + (p_face - p_ref) / a / a;
                    T vx_face = vx[va] + nx * (p_ref - p_face) / (rho_ref * a);
                    T vy_face = vy[va] + ny * (p_ref - p_face) / (rho_ref * a);
                    T vz_face = vz[va] + nz * (p_ref - p_face) / (rho_ref * a);
                    p[gh] = 2 * p_face - p[va];
                    rho[gh] = 2 * rho_face - rho[va];
                    temp[gh] = p[gh] / (rho[gh] * R);
                    e[gh] = Cv * temp[gh];
                    vx[gh] = 2 * vx_face - vx[va];
                    vy[gh] = 2 * vy_face - vy[va];
                    vz[gh] = 2 * vz_face - vz[va];
                    break;
                }
                default: break;
            }
        }
    }
    void apply_bc_list(const int32_t* kinds, const double* pars, T* fs) {
        for (int t = 0; t < g->ntags; t++)
            for (int a = 0; a < OR_MAX_ACT; a++) {
                int k = kinds[t * OR_MAX_ACT + a];
                if (k == OR_BC_NONE) continue;
                apply_action(k, pars + (t * OR_MAX_ACT + a) * OR_NPAR, g->bfaces[t], fs);
            }
    }

    // --- thermo, GAS/gas_model.h:11-116 ---------------------------------------
    void thermo_update(int mode, T& rho, T& p, T& temp, T& e) const {
        const double R = cfg.R, Cv = cfg.Cv;
        switch (mode) {
            case OR_THERMO_RHO_P:  // update_thermo_from_rhop
                temp = p / (rho * R);
                e = Cv * temp;
                break;
            case OR_THERMO_RHO_U:  // update_thermo_from_rhou
                temp = e / Cv;
                p = rho * R * temp;
                break;
            case OR_THERMO_RHO_T:  // update_thermo_from_rhoT
                p = rho * R * temp;
                e = Cv * temp;
                break;
            case OR_THERMO_P_T:  // update_thermo_from_pT
                rho = p / (R * temp);
                e = Cv * temp;
                break;
        }
    }

    // --- gradients, FV/convective_flux.cpp:181-209 ----------------------------
    void convective_gradients(const T* fs) {
        auto gg = [&](int prow, int slot) {
            lsq_gradient<T>(*g, fs + prow * Nt, gcomp(slot, 0), gcomp(slot, 1), gcomp(slot, 2));
        };
        switch (cfg.thermo) {
            case OR_THERMO_RHO_P: gg(P_P, G_P); gg(P_RHO, G_RHO); break;
            case OR_THERMO_RHO_U: gg(P_RHO, G_RHO); gg(P_E, G_U); break;
            case OR_THERMO_RHO_T: gg(P_RHO, G_RHO); gg(P_T, G_T); break;
            case OR_THERMO_P_T: gg(P_P, G_P); gg(P_T, G_T); break;
        }
        gg(P_VX, G_VX);
        gg(P_VY, G_VY);
        if (g->dim == 3) gg(P_VZ, G_VZ);
    }

    // --- limiters, FV/convective_flux.cpp:423-462 and FV/limiter.cpp:24-62 -----
    void limiter_one(const T* values, int slot) {
        const Grid& G = *g;
        T* out = &lim[(size_t)slot * N];
        const T* gx = gcomp(slot, 0);
        const T* gy = gcomp(slot, 1);
        const T* gz = gcomp(slot, 2);
        const int kind = cfg.limiter;
        const double eps = cfg.lim_eps;
        const double K = cfg.venkat_K;
        const int dim = G.dim;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; i++) {
            T Ui = values[i];
            T U_min = Ui, U_max = Ui;
            for (int64_t j = G.cf_off[i]; j < G.cf_off[i + 1]; j++) {
                int64_t n = G.nbr[j];
                U_min = kmin(U_min, values[n]);
                U_max = kmax(U_max, values[n]);
            }
            T phi = 1.0;
            double x = G.x[i], y = G.y[i], z = G.z[i];
            if (kind == OR_LIM_BARTH_JESPERSEN) {
                for (int64_t j = G.cf_off[i]; j < G.cf_off[i + 1]; j++) {
                    int64_t f = G.cf[j];
                    double dx = G.cx[f] - x;
                    double dy = G.cy[f] - y;
                    double dz = G.cz[f] - z;
                    T delta_2 = gx[i] * dx + gy[i] * dy + gz[i] * dz;
                    int s = (delta_2 > T(0.0)) - (delta_2 < T(0.0));
                    delta_2 = (double)s * (kabs(delta_2) + eps);
                    if (s > 0)
                        phi = kmin(phi, (U_max - Ui) / delta_2);
                    else if (s < 0)
                        phi = kmin(phi, (U_min - Ui) / delta_2);
                }
            } else {
                // Venkatakrishnan (1995) limiter.  Not in the reference; the
                // formulation is this project's: eps^2 = K^3 * h^3 with the
                // cell size h^dim = |V| (the reference's volume of a cell with
                // non-planar faces can be negative), i.e. (K*K*K)*|V| in 3D and
                // (K*K*K)*(|V|*sqrt(|V|)) in 2D, per-face phi_f =
                // (d1^2 + eps^2 + 2 d1 d2) / (d1^2 + 2 d2^2 + d1 d2 + eps^2),
                // phi = min(1, min_f phi_f), the minimum taken over the
                // fractions by cross-multiplication (venkat_take) and divided
                // out once.
                double V = std::fabs(G.vol[i]);
                double K3 = K * K * K;
                double eps2 = (dim == 3) ? K3 * V : K3 * (V * std::sqrt(V));
                T bn = 1.0, bd = 1.0;
                for (int64_t j = G.cf_off[i]; j < G.cf_off[i + 1]; j++) {
                    int64_t f = G.cf[j];
                    double dx = G.cx[f] - x;
                    double dy = G.cy[f] - y;
                    double dz = G.cz[f] - z;
                    T d2 = gx[i] * dx + gy[i] * dy + gz[i] * dz;
                    T d1;
                    if (d2 > T(0.0))
                        d1 = U_max - Ui;
                    else if (d2 < T(0.0))
                        d1 = U_min - Ui;
                    else
                        continue;
                    T d1s = d1 * d1;
                    T d2s = d2 * d2;
                    T num = d1s + eps2 + 2.0 * (d1 * d2);
                    T den = d1s + 2.0 * d2s + d1 * d2 + eps2;
                    if (venkat_take(num, den, bn, bd)) {
                        bn = num;
                        bd = den;
                    }
                }
                phi = bn / bd;
            }
            out[i] = phi;
        }
    }
    void compute_limiters(const T* fs) {
        if (cfg.limiter == OR_LIM_UNLIMITED) return;
        switch (cfg.thermo) {
            case OR_THERMO_RHO_P: limiter_one(fs + P_P * Nt, G_P); limiter_one(fs + P_RHO * Nt, G_RHO); break;
            case OR_THERMO_RHO_U: limiter_one(fs + P_RHO * Nt, G_RHO); limiter_one(fs + P_E * Nt, G_U); break;
            case OR_THERMO_RHO_T: limiter_one(fs + P_RHO * Nt, G_RHO); limiter_one(fs + P_T * Nt, G_T); break;
            case OR_THERMO_P_T: limiter_one(fs + P_P * Nt, G_P); limiter_one(fs + P_T * Nt, G_T); break;
        }
        limiter_one(fs + P_VX * Nt, G_VX);
        limiter_one(fs + P_VY * Nt, G_VY);
        limiter_one(fs + P_VZ * Nt, G_VZ);  // FV/convective_flux.cpp:459-460, also in 2D
    }

    // --- reconstruction, FV/convective_flux.cpp:212-420 -----------------------
    // linear_interpolate, FV/convective_flux.cpp:243-255
    T lin(T value, int slot, double dx, double dy, double dz, int64_t i, T limiter,
          bool valid) {
        T gx = 0.0, gy = 0.0, gz = 0.0;
        if (valid) {
            gx = gcomp(slot, 0)[i];
            gy = gcomp(slot, 1)[i];
            gz = gcomp(slot, 2)[i];
        }
        return value + limiter * (gx * dx + gy * dy + gz * dz);
    }
    void copy_reconstruct(const T* fs) {
        for (int64_t f = 0; f < F; f++) {
            int64_t l = g->left[f], r = g->right[f];
            for (int v = 0; v < NPRIM; v++) {
                left[v * F + f] = fs[v * Nt + l];
                right[v * F + f] = fs[v * Nt + r];
            }
        }
    }
    void reconstruct_side(const T* fs, int64_t f, int64_t c, T* st) {
        const bool valid = c < N;
        const bool limit = cfg.limiter != OR_LIM_UNLIMITED && valid;
        double dx = g->cx[f] - g->x[c];
        double dy = g->cy[f] - g->y[c];
        double dz = g->cz[f] - g->z[c];
        auto L = [&](int slot) -> T { return limit ? lim[(size_t)slot * N + c] : T(1.0); };
        T rho = 0.0, p = 0.0, temp = 0.0, e = 0.0;
        switch (cfg.thermo) {
            case OR_THERMO_RHO_P:
                p = lin(fs[P_P * Nt + c], G_P, dx, dy, dz, c, L(G_P), valid);
                rho = lin(fs[P_RHO * Nt + c], G_RHO, dx, dy, dz, c, L(G_RHO), valid);
                break;
            case OR_THERMO_RHO_U:
                rho = lin(fs[P_RHO * Nt + c], G_RHO, dx, dy, dz, c, L(G_RHO), valid);
                e = lin(fs[P_E * Nt + c], G_U, dx, dy, dz, c, L(G_U), valid);
                break;
            case OR_THERMO_RHO_T:
                rho = lin(fs[P_RHO * Nt + c], G_RHO, dx, dy, dz, c, L(G_RHO), valid);
                temp = lin(fs[P_T * Nt + c], G_T, dx, dy, dz, c, L(G_T), valid);
                break;
            case OR_THERMO_P_T:
                p = lin(fs[P_P * Nt + c], G_P, dx, dy, dz, c, L(G_P), valid);
                temp = lin(fs[P_T * Nt + c], G_T, dx, dy, dz, c, L(G_T), valid);
                break;
        }
        // the face arrays keep whatever the previous call left in the slots
        // the thermo update does not touch; mirror that by starting from them
        T rho0 = st[P_RHO], p0 = st[P_P], t0 = st[P_T], e0 = st[P_E];
        switch (cfg.thermo) {
            case OR_THERMO_RHO_P: temp = t0; e = e0; break;
            case OR_THERMO_RHO_U: p = p0; temp = t0; break;
            case OR_THERMO_RHO_T: p = p0; e = e0; break;
            case OR_THERMO_P_T: rho = rho0; e = e0; break;
        }
        thermo_update(cfg.thermo, rho, p, temp, e);
        st[P_RHO] = rho;
        st[P_P] = p;
        st[P_T] = temp;
        st[P_E] = e;
        st[P_VX] = lin(fs[P_VX * Nt + c], G_VX, dx, dy, dz, c, L(G_VX), valid);
        st[P_VY] = lin(fs[P_VY * Nt + c], G_VY, dx, dy, dz, c, L(G_VY), valid);
        st[P_VZ] = lin(fs[P_VZ * Nt + c], G_VZ, dx, dy, dz, c, L(G_VZ), valid);
    }
    void linear_reconstruct(const T* fs) {
        convective_gradients(fs);
        compute_limiters(fs);
#pragma omp parallel for schedule(static)
        for (int64_t f = 0; f < F; f++) {
            T st[NPRIM];
            for (int v = 0; v < NPRIM; v++) st[v] = left[v * F + f];
            reconstruct_side(fs, f, g->left[f], st);
            for (int v = 0; v < NPRIM; v++) left[v * F + f] = st[v];
            for (int v = 0; v < NPRIM; v++) st[v] = right[v * F + f];
            reconstruct_side(fs, f, g->right[f], st);
            for (int v = 0; v < NPRIM; v++) right[v * F + f] = st[v];
        }
    }

    // --- flux calculators, FV/flux_calculators/*.cpp ----------------------------
    // st: rho, p, T, e, u, v, w (velocity already in the face frame)
    void flux_ausmdv(const T* L, const T* Rs, T* out) const {  // ausmdv.cpp:9-131
        const double R = cfg.R, Cv = cfg.Cv, gamma = cfg.gamma;
        const bool three_d = g->dim == 3;
        T rL = L[P_RHO], pL = L[P_P];
        T pLrL = pL / rL;
        T uL = L[P_VX], vL = L[P_VY], wL = L[P_VZ];
        T eL = Cv * L[P_T];
        T aL = ksqrt(gamma * R * L[P_T]);
        T keL = 0.5 * (uL * uL + vL * vL + wL * wL);
        T HL = eL + pLrL + keL;
        T rR = Rs[P_RHO], pR = Rs[P_P];
        T pRrR = pR / rR;
        T uR = Rs[P_VX], vR = Rs[P_VY], wR = Rs[P_VZ];
        T eR = Cv * Rs[P_T];
        T aR = ksqrt(gamma * R * Rs[P_T]);
        T keR = 0.5 * (uR * uR + vR * vR + wR * wR);
        T HR = eR + pRrR + keR;

        T alphaL = 2.0 * pLrL / (pLrL + pRrR);
        T alphaR = 2.0 * pRrR / (pLrL + pRrR);
        T am = kmax(aL, aR);
        T ML = uL / am;
        T MR = uR / am;
        T pLplus, uLplus;
        T duL = 0.5 * (uL + kabs(uL));
        if (kabs(ML) <= T(1.0)) {
            pLplus = pL * (ML + 1.0) * (ML + 1.0) * (2.0 - ML) * 0.25;
            uLplus = alphaL * ((uL + am) * (uL + am) / (4.0 * am) - duL) + duL;
        } else {
            pLplus = pL * duL / uL;
            uLplus = duL;
        }
        T pRminus, uRminus;
        T duR = 0.5 * (uR - kabs(uR));
        if (kabs(MR) <= T(1.0)) {
            pRminus = pR * (MR - 1.0) * (MR - 1.0) * (2.0 + MR) * 0.25;
            uRminus = alphaR * (-(uR - am) * (uR - am) / (4.0 * am) - duR) + duR;
        } else {
            pRminus = pR * duR / uR;
            uRminus = duR;
        }
        T ru_half = uLplus * rL + uRminus * rR;
        T p_half = pLplus + pRminus;
        T dp = pL - pR;
        const double K_SWITCH = 10.0;
        dp = K_SWITCH * kabs(dp) / kmin(pL, pR);
        T s = 0.5 * kmin(1.0, dp);
        T ru2_AUSMV = uLplus * rL * uL + uRminus * rR * uR;
        T ru2_AUSMD = 0.5 * (ru_half * (uL + uR) - kabs(ru_half) * (uR - uL));
        T ru2_half = (0.5 + s) * ru2_AUSMV + (0.5 - s) * ru2_AUSMD;
        T m = ru_half, mx, my, mz = 0.0, en;
        if (ru_half >= T(0.0)) {
            mx = ru2_half + p_half;
            my = ru_half * vL;
            if (three_d) mz = ru_half * wL;
            en = ru_half * HL;
        } else {
            mx = ru2_half + p_half;
            my = ru_half * vR;
            if (three_d) mz = ru_half * wR;
            en = ru_half * HR;
        }
        const double C_EFIX = 0.125;
        bool caseA = ((uL - aL) < T(0.0)) && ((uR - aR) > T(0.0));
        bool caseB = ((uL + aL) < T(0.0)) && ((uR + aR) > T(0.0));
        T d_ua = 0.0;
        if (caseA && !caseB) d_ua = C_EFIX * ((uR - aR) - (uL - aL));
        if (caseB && !caseA) d_ua = C_EFIX * ((uR + aR) - (uL + aL));
        if (d_ua != T(0.0)) {
            m -= d_ua * (rR - rL);
            mx -= d_ua * (rR * uR - rL * uL);
            my -= d_ua * (rR * vR - rL * vL);
            if (three_d) mz -= d_ua * (rR * wR - rL * wL);
            en -= d_ua * (rR * HR - rL * HL);
        }
        out[0] = m;
        out[1] = mx;
        out[2] = my;
        out[3] = mz;
        out[4] = en;
    }
    void flux_hanel(const T* L, const T* Rs, T* out) const {  // hanel.cpp:8-64
        const double R = cfg.R, Cv = cfg.Cv, gamma = cfg.gamma;
        T rL = L[P_RHO], pL = L[P_P];
        T eL = Cv * L[P_T];
        T aL = ksqrt(gamma * R * L[P_T]);
        T uL = L[P_VX], vL = L[P_VY], wL = L[P_VZ];
        T keL = 0.5 * (uL * uL + vL * vL + wL * wL);
        T pLrL = pL / rL;
        T HL = eL + pLrL + keL;
        T rR = Rs[P_RHO], pR = Rs[P_P];
        T eR = Cv * Rs[P_T];
        T aR = ksqrt(gamma * R * Rs[P_T]);
        T uR = Rs[P_VX], vR = Rs[P_VY], wR = Rs[P_VZ];
        T keR = 0.5 * (uR * uR + vR * vR + wR * wR);
        T pRrR = pR / rR;
        T HR = eR + pRrR + keR;
        T pLplus, uLplus;
        if (kabs(uL) <= aL) {
            uLplus = 1.0 / (4.0 * aL) * (uL + aL) * (uL + aL);
            pLplus = pL * uLplus * (1.0 / aL * (2.0 - uL / aL));
        } else {
            uLplus = 0.5 * (uL + kabs(uL));
            pLplus = pL * uLplus * (1.0 / uL);
        }
        T pRminus, uRminus;
        if (kabs(uR) <= aR) {
            uRminus = -1.0 / (4.0 * aR) * (uR - aR) * (uR - aR);
            pRminus = pR * uRminus * (1.0 / aR * (-2.0 - uR / aR));
        } else {
            uRminus = 0.5 * (uR - kabs(uR));
            pRminus = pR * uRminus * (1.0 / uR);
        }
        T p_half = pLplus + pRminus;
        out[0] = uLplus * rL + uRminus * rR;
        out[1] = uLplus * rL * uL + uRminus * rR * uR + p_half;
        out[2] = uLplus * rL * vL + uRminus * rR * vR;
        out[3] = (g->dim == 3) ? (uLplus * rL * wL + uRminus * rR * wR) : T(0.0);
        out[4] = uLplus * rL * HL + uRminus * rR * HR;
    }
    void flux_ldfss2(const T* L, const T* Rs, T* out) const {  // ldfss2.cpp:10-89
        const double R = cfg.R, Cv = cfg.Cv, gamma = cfg.gamma;
        const double delta = cfg.ldfss_delta;
        T rL = L[P_RHO], pL = L[P_P];
        T pLrL = pL / rL;
        T uL = L[P_VX], vL = L[P_VY], wL = L[P_VZ];
        T eL = Cv * L[P_T];
        T aL = ksqrt(gamma * R * L[P_T]);
        T keL = 0.5 * (uL * uL + vL * vL + wL * wL);
        T HL = eL + pLrL + keL;
        T rR = Rs[P_RHO], pR = Rs[P_P];
        T pRrR = pR / rR;
        T uR = Rs[P_VX], vR = Rs[P_VY], wR = Rs[P_VZ];
        T eR = Cv * Rs[P_T];
        T aR = ksqrt(gamma * R * Rs[P_T]);
        T keR = 0.5 * (uR * uR + vR * vR + wR * wR);
        T HR = eR + pRrR + keR;
        T am = 0.5 * (aL + aR);
        T ML = uL / am;
        T MR = uR / am;
        T MpL = 0.25 * (ML + 1.0) * (ML + 1.0);
        T MmR = -0.25 * (MR - 1.0) * (MR - 1.0);
        T alphaL = 0.5 * (1.0 + kcopysign(T(1.0), ML));
        T alphaR = 0.5 * (1.0 - kcopysign(T(1.0), MR));
        T betaL = -kmax(0.0, 1.0 - kfloor(kabs(ML)));
        T betaR = -kmax(0.0, 1.0 - kfloor(kabs(MR)));
        T PL = 0.25 * (ML + 1.0) * (ML + 1.0) * (2.0 - ML);
        T PR = 0.25 * (MR - 1.0) * (MR - 1.0) * (2.0 + MR);
        T DL = alphaL * (1.0 + betaL) - betaL * PL;
        T DR = alphaR * (1.0 + betaR) - betaR * PR;
        T Mhalf = 0.25 * betaL * betaR * kpow((ksqrt(0.5 * (ML * ML + MR * MR)) - 1.0), T(2.0));
        T MhalfL = Mhalf * (1.0 - ((pL - pR) / (pL + pR) + delta * (kabs(pL - pR) / pL)));
        T MhalfR = Mhalf * (1.0 + ((pL - pR) / (pL + pR) - delta * (kabs(pL - pR) / pR)));
        T CL = alphaL * (1.0 + betaL) * ML - betaL * MpL - MhalfL;
        T CR = alphaR * (1.0 + betaR) * MR - betaR * MmR + MhalfR;
        T ru_half = am * rL * CL + am * rR * CR;
        T ru2_half = am * rL * CL * uL + am * rR * CR * uR;
        T p_half = DL * pL + DR * pR;
        out[0] = ru_half;
        out[1] = ru2_half + p_half;
        out[2] = am * rL * CL * vL + am * rR * CR * vR;
        out[3] = (g->dim == 3) ? (am * rL * CL * wL + am * rR * CR * wR) : T(0.0);
        out[4] = am * rL * CL * HL + am * rR * CR * HR;
    }
    void flux_rusanov(const T* L, const T* Rs, T* out) const {  // rusanov.cpp:10-52
        const double R = cfg.R, Cv = cfg.Cv, gamma = cfg.gamma;
        T rL = L[P_RHO];
        T eL = Cv * L[P_T];
        T pL = L[P_P];
        T uL = L[P_VX], vL = L[P_VY], wL = L[P_VZ];
        T ruL = rL * uL;
        T keL = 0.5 * (uL * uL + vL * vL + wL * wL);
        T HL = eL + pL / rL + keL;
        T aL = ksqrt(gamma * R * L[P_T]);
        T rR = Rs[P_RHO];
        T eR = Cv * Rs[P_T];
        T pR = Rs[P_P];
        T uR = Rs[P_VX], vR = Rs[P_VY], wR = Rs[P_VZ];
        T ruR = rR * uR;
        T keR = 0.5 * (uR * uR + vR * vR + wR * wR);
        T HR = eR + pR / rR + keR;
        T aR = ksqrt(gamma * R * Rs[P_T]);
        T S_plus = kmax(kabs(uL - aL), kabs(uR - aR));
        S_plus = kmax(S_plus, kabs(uL + aL));
        S_plus = kmax(S_plus, kabs(uR + aR));
        out[0] = 0.5 * (ruL + ruR) - 0.5 * S_plus * (rR - rL);
        out[1] = 0.5 * (ruL * uL + pL + ruR * uR + pR) - 0.5 * S_plus * (rR * uR - rL * uL);
        out[2] = 0.5 * (ruL * vL + ruR * vR) - 0.5 * S_plus * (rR * vR - rL * vL);
        out[3] = (g->dim == 3)
                     ? (0.5 * (ruL * wL + ruR * wR) - 0.5 * S_plus * (rR * wR - rL * wL))
                     : T(0.0);
        out[4] = 0.5 * (ruL * HL + ruR * HR) - 0.5 * S_plus * (rR * HR - rL * HL);
    }

    // ConvectiveFlux::compute_convective_flux, FV/convective_flux.cpp:85-178
    void convective_flux(const T* fs, bool allow_reconstruction) {
        int order = allow_reconstruction ? cfg.order : 1;
        if (order == 1)
            copy_reconstruct(fs);
        else
            linear_reconstruct(fs);
        const Grid& G = *g;
        const bool three_d = G.dim == 3;
#pragma omp parallel for schedule(static)
        for (int64_t f = 0; f < F; f++) {
            T Ls[NPRIM], Rs[NPRIM];
            for (int v = 0; v < NPRIM; v++) {
                Ls[v] = left[v * F + f];
                Rs[v] = right[v * F + f];
            }
            // transform_to_local_frame, UT/vector3.cpp:142-153
            double nx = G.nx[f], ny = G.ny[f], nz = G.nz[f];
            double t1x = G.t1x[f], t1y = G.t1y[f], t1z = G.t1z[f];
            double t2x = G.t2x[f], t2y = G.t2y[f], t2z = G.t2z[f];
            for (T* s : {Ls, Rs}) {
                T a = s[P_VX], b = s[P_VY], c = s[P_VZ];
                s[P_VX] = a * nx + b * ny + c * nz;
                s[P_VY] = a * t1x + b * t1y + c * t1z;
                s[P_VZ] = a * t2x + b * t2y + c * t2z;
            }
            for (int v = 0; v < NPRIM; v++) {
                left[v * F + f] = Ls[v];
                right[v * F + f] = Rs[v];
            }
            T out[5];
            switch (cfg.flux) {
                case OR_FLUX_HANEL: flux_hanel(Ls, Rs, out); break;
                case OR_FLUX_AUSMDV: flux_ausmdv(Ls, Rs, out); break;
                case OR_FLUX_LDFSS2: flux_ldfss2(Ls, Rs, out); break;
                default: flux_rusanov(Ls, Rs, out); break;
            }
            // flux::transform_to_global, FV/convective_flux.cpp:148-177
            T px = out[1], py = out[2], pz = 0.0;
            if (three_d) pz = out[3];
            T x = px * nx + py * t1x + pz * t2x;
            T y = px * ny + py * t1y + pz * t2y;
            T z = px * nz + py * t1z + pz * t2z;
            flux[0 * F + f] = out[0];
            flux[1 * F + f] = x;
            flux[2 * F + f] = y;
            if (three_d) {
                flux[3 * F + f] = z;
                flux[4 * F + f] = out[4];
            } else {
                flux[3 * F + f] = out[4];
            }
        }
    }

    // ConstantFlux::apply, FV/boundaries/boundary.cpp:405-443 (static grid)
    void post_convective_flux_bc() {
        for (int t = 0; t < g->ntags; t++)
            for (int a = 0; a < OR_MAX_ACT; a++) {
                if (cfg.post_kind[t * OR_MAX_ACT + a] != OR_BC_CONSTANT_FLUX) continue;
                const double* par = cfg.post_par + (t * OR_MAX_ACT + a) * OR_NPAR;
                for (int64_t f : g->bfaces[t]) {
                    double nx = g->nx[f], ny = g->ny[f], nz = g->nz[f];
                    T u = cfg.Cv * T(par[0]);
                    T rel_vx = T(par[4]) - 0.0, rel_vy = T(par[5]) - 0.0, rel_vz = T(par[6]) - 0.0;
                    T mass_flux = T(par[2]) * (rel_vx * nx + rel_vy * ny + rel_vz * nz);
                    T p = par[1];
                    T fvx = par[4], fvy = par[5], fvz = par[6];
                    flux[i_mass() * F + f] = mass_flux;
                    flux[i_mx() * F + f] = p * nx + fvx * mass_flux;
                    flux[i_my() * F + f] = p * ny + fvy * mass_flux;
                    if (g->dim == 3) flux[i_mz() * F + f] = p * nz + fvz * mass_flux;
                    flux[i_en() * F + f] =
                        mass_flux * (u + 0.5 * (fvx * fvx + fvy * fvy + fvz * fvz)) +
                        p * (fvx * nx + fvy * nx + fvz * nz);  // v_y * n_x (sic)
                }
            }
    }

    // --- viscous flux, FV/viscous_flux.cpp:16-184 ---------------------------------
    T sutherland(T temp) const {  // GAS/transport_properties.h:27-42
        T mu0 = cfg.mu0, T0 = cfg.T0, Ts = cfg.Ts;
        return mu0 * kpow(temp / T0, T(3.0 / 2.0)) * (T0 + Ts) / (temp + Ts);
    }
    T conductivity(T temp) const {  // GAS/transport_properties.h:64-75
        T mu = sutherland(temp);
        return cfg.Cp * mu / T(cfg.Pr);
    }
    void viscous_gradients(const T* fs) {  // FV/viscous_flux.cpp:122-130
        lsq_gradient<T>(*g, fs + P_T * Nt, gcomp(G_T, 0), gcomp(G_T, 1), gcomp(G_T, 2));
        lsq_gradient<T>(*g, fs + P_VX * Nt, gcomp(G_VX, 0), gcomp(G_VX, 1), gcomp(G_VX, 2));
        lsq_gradient<T>(*g, fs + P_VY * Nt, gcomp(G_VY, 0), gcomp(G_VY, 1), gcomp(G_VY, 2));
        lsq_gradient<T>(*g, fs + P_VZ * Nt, gcomp(G_VZ, 0), gcomp(G_VZ, 1), gcomp(G_VZ, 2));
    }
    void hasselbacher(int slot, T vl, T vr, int64_t l, int64_t r, const double* ehat,
                      const double* n, double len_e, double edn, T* out) {
        const T* gx = gcomp(slot, 0);
        const T* gy = gcomp(slot, 1);
        const T* gz = gcomp(slot, 2);
        T ax = 0.5 * (gx[l] + gx[r]);
        T ay = 0.5 * (gy[l] + gy[r]);
        T az = 0.5 * (gz[l] + gz[r]);
        T avg_dot_ehat = ax * ehat[0] + ay * ehat[1] + az * ehat[2];
        T correction = avg_dot_ehat - (vr - vl) / len_e;
        out[0] = ax - correction * n[0] / edn;
        out[1] = ay - correction * n[1] / edn;
        out[2] = az - correction * n[2] / edn;
    }
    void viscous_flux(const T* fs) {
        viscous_gradients(fs);
        const Grid& G = *g;
        const int dim = G.dim;
#pragma omp parallel for schedule(static)
        for (int64_t f = 0; f < F; f++) {
            int64_t l = G.left[f], r = G.right[f];
            bool lv = l < N, rv = r < N;
            T gT[3], gvx[3], gvy[3], gvz[3];
            if (!lv || !rv) {  // copy_gradients_to_face
                int64_t c = lv ? l : r;
                for (int k = 0; k < 3; k++) {
                    gT[k] = gcomp(G_T, k)[c];
                    gvx[k] = gcomp(G_VX, k)[c];
                    gvy[k] = gcomp(G_VY, k)[c];
                    gvz[k] = gcomp(G_VZ, k)[c];
                }
            } else {  // hasselbacher_average
                double ex = G.x[r] - G.x[l];
                double ey = G.y[r] - G.y[l];
                double ez = G.z[r] - G.z[l];
                double len_e = std::sqrt(ex * ex + ey * ey + ez * ez);
                double ehat[3] = {ex / len_e, ey / len_e, ez / len_e};
                double n[3] = {G.nx[f], G.ny[f], G.nz[f]};
                double edn = ehat[0] * n[0] + ehat[1] * n[1] + ehat[2] * n[2];
                hasselbacher(G_VX, fs[P_VX * Nt + l], fs[P_VX * Nt + r], l, r, ehat, n, len_e, edn, gvx);
                hasselbacher(G_VY, fs[P_VY * Nt + l], fs[P_VY * Nt + r], l, r, ehat, n, len_e, edn, gvy);
                hasselbacher(G_VZ, fs[P_VZ * Nt + l], fs[P_VZ * Nt + r], l, r, ehat, n, len_e, edn, gvz);
                hasselbacher(G_T, fs[P_T * Nt + l], fs[P_T * Nt + r], l, r, ehat, n, len_e, edn, gT);
            }
            for (int k = 0; k < 3; k++) {
                face_grad[(size_t)(0 * 3 + k) * F + f] = gT[k];
                face_grad[(size_t)(1 * 3 + k) * F + f] = gvx[k];
                face_grad[(size_t)(2 * 3 + k) * F + f] = gvy[k];
                face_grad[(size_t)(3 * 3 + k) * F + f] = gvz[k];
            }
            // average_flow_states_pT (GAS/flow_state.h:89-91, GAS/gas_state.h:108-114)
            T p = 0.5 * (fs[P_P * Nt + l] + fs[P_P * Nt + r]);
            T temp = 0.5 * (fs[P_T * Nt + l] + fs[P_T * Nt + r]);
            T vx = 0.5 * (fs[P_VX * Nt + l] + fs[P_VX * Nt + r]);
            T vy = 0.5 * (fs[P_VY * Nt + l] + fs[P_VY * Nt + r]);
            T vz = 0.5 * (fs[P_VZ * Nt + l] + fs[P_VZ * Nt + r]);
            T rho = p / (cfg.R * temp);  // update_thermo_from_pT
            T e = cfg.Cv * temp;
            face_fs[P_RHO * F + f] = rho;
            face_fs[P_P * F + f] = p;
            face_fs[P_T * F + f] = temp;
            face_fs[P_E * F + f] = e;
            face_fs[P_VX * F + f] = vx;
            face_fs[P_VY * F + f] = vy;
            face_fs[P_VZ * F + f] = vz;
            T mu = sutherland(temp);
            T k = conductivity(temp);
            T lambda = -2.0 / 3.0 * mu;
            T bulk = lambda * (gvx[0] + gvy[1] + gvz[2]);
            T tau_xx = 2.0 * mu * gvx[0] + bulk;
            T tau_yy = 2.0 * mu * gvy[1] + bulk;
            T tau_zz = 2.0 * mu * gvz[2] + bulk;
            T tau_xy = mu * (gvx[1] + gvy[0]);
            T tau_xz = mu * (gvx[2] + gvz[0]);
            T tau_yz = mu * (gvy[2] + gvz[1]);
            T theta_x = vx * tau_xx + vy * tau_xy + vz * tau_xz + k * gT[0];
            T theta_y = vx * tau_xy + vy * tau_yy + vz * tau_yz + k * gT[1];
            T theta_z = vx * tau_xz + vy * tau_yz + vz * tau_zz + k * gT[2];
            double nx = G.nx[f], ny = G.ny[f], nz = G.nz[f];
            flux[i_mx() * F + f] -= nx * tau_xx + ny * tau_xy + nz * tau_xz;
            flux[i_my() * F + f] -= nx * tau_xy + ny * tau_yy + nz * tau_yz;
            if (dim == 3) flux[i_mz() * F + f] -= nx * tau_xz + ny * tau_yz + nz * tau_zz;
            flux[i_en() * F + f] -= nx * theta_x + ny * theta_y + nz * theta_z;
        }
    }

    // --- FiniteVolume::flux_surface_integral, FV/finite_volume.cpp:221-255 -------
    void surface_integral(T* dudt) {
        const Grid& G = *g;
        const bool three_d = G.dim == 3;
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < N; c++) {
            T dm = 0.0, dmx = 0.0, dmy = 0.0, dmz = 0.0, de = 0.0;
            for (int64_t j = G.cf_off[c]; j < G.cf_off[c + 1]; j++) {
                int64_t f = G.cf[j];
                double area = -G.area[f] * G.outsign[j];
                dm += flux[i_mass() * F + f] * area;
                dmx += flux[i_mx() * F + f] * area;
                dmy += flux[i_my() * F + f] * area;
                if (three_d) dmz += flux[i_mz() * F + f] * area;
                de += flux[i_en() * F + f] * area;
            }
            double V = G.vol[c];
            dudt[i_mass() * Nt + c] = dm / V;
            dudt[i_mx() * Nt + c] = dmx / V;
            dudt[i_my() * Nt + c] = dmy / V;
            if (three_d) dudt[i_mz() * Nt + c] = dmz / V;
            dudt[i_en() * Nt + c] = de / V;
        }
    }

    // --- FiniteVolume::compute_dudt, FV/finite_volume.cpp:57-96 ------------------
    // FiniteVolume::compute_convective_gradient, FV/finite_volume.cpp:283-288:
    // pre-reconstruction actions, then the convective gradients
    void compute_convective_gradient(T* fs) {
        apply_bc_list(cfg.pre_kind, cfg.pre_par, fs);
        convective_gradients(fs);
    }
    // FiniteVolume::compute_viscous_gradient, FV/finite_volume.cpp:273-280:
    // pre-reconstruction and pre-viscous-gradient actions, then T, vx, vy, vz
    void compute_viscous_gradient(T* fs) {
        apply_bc_list(cfg.pre_kind, cfg.pre_par, fs);
        apply_bc_list(cfg.visc_kind, cfg.visc_par, fs);
        viscous_gradients(fs);
    }

    int compute_dudt(T* fs, T* dudt, bool allow_reconstruction) {
        apply_bc_list(cfg.pre_kind, cfg.pre_par, fs);
        convective_flux(fs, allow_reconstruction);
        post_convective_flux_bc();
        if (cfg.viscous) {
            apply_bc_list(cfg.visc_kind, cfg.visc_par, fs);
            viscous_flux(fs);
        }
        surface_integral(dudt);
        return 0;
    }

    // --- FiniteVolume::estimate_dt, FV/finite_volume.cpp:131-184 -----------------
    double estimate_dt(const T* fs) {
        const Grid& G = *g;
        double dt = std::numeric_limits<double>::max();  // Kokkos::Min identity
        const bool viscous = cfg.viscous != 0;
        const double sf = cfg.signal_factor;
#pragma omp parallel for reduction(min : dt) schedule(static)
        for (int64_t c = 0; c < N; c++) {
            T vx = fs[P_VX * Nt + c], vy = fs[P_VY * Nt + c], vz = fs[P_VZ * Nt + c];
            T src = 0.0, srv = 0.0;
            double volume = G.vol[c];
            for (int64_t j = G.cf_off[c]; j < G.cf_off[c + 1]; j++) {
                int64_t f = G.cf[j];
                double area = G.area[f];
                T dot = vx * G.nx[f] + vy * G.ny[f] + vz * G.nz[f];
                T sig_vel = kabs(dot) + ksqrt(cfg.gamma * cfg.R * fs[P_T * Nt + c]);
                src += sig_vel * area;
                if (viscous) {
                    double gamma = cfg.gamma;
                    T tf = face_fs[P_T * F + f];
                    T mu = sutherland(tf);
                    T k = conductivity(tf);
                    T rho = face_fs[P_RHO * F + f];
                    T Pr = mu * cfg.Cp / k;
                    T tmp = (gamma / rho) * (mu / Pr) * area * area;
                    srv += tmp / volume;
                }
            }
            T local_dt = volume / (src + sf * srv);
            dt = kmin(real_part(local_dt), dt);
        }
        return dt;
    }

    // --- FiniteVolume::count_bad_cells, FV/finite_volume.cpp:258-270 -------------
    int64_t count_bad_cells(const T* fs) {
        int64_t n = 0;
        for (int64_t c = 0; c < N; c++) {
            T temp = fs[P_T * Nt + c], rho = fs[P_RHO * Nt + c];
            if (temp < T(0.0) || rho < T(0.0) || kisnan(real_part(rho)) ||
                kisinf(real_part(rho)))
                n++;
        }
        return n;
    }

    // --- conversions, FV/primative_conserved_conversion.cpp:5-51 -----------------
    void cons_to_prim(const T* cq, T* fs) {
        const bool three_d = g->dim == 3;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < Nt; i++) {
            T rho = cq[i_mass() * Nt + i];
            T vx = cq[i_mx() * Nt + i] / rho;
            T vy = cq[i_my() * Nt + i] / rho;
            T vz = three_d ? cq[i_mz() * Nt + i] / rho : T(0.0);
            T ke = 0.5 * (vx * vx + vy * vy + vz * vz);
            T u = cq[i_en() * Nt + i] / rho - ke;
            fs[P_RHO * Nt + i] = rho;
            fs[P_VX * Nt + i] = vx;
            fs[P_VY * Nt + i] = vy;
            fs[P_VZ * Nt + i] = vz;
            fs[P_E * Nt + i] = u;
            T temp = u / cfg.Cv;  // update_thermo_from_rhou
            fs[P_T * Nt + i] = temp;
            fs[P_P * Nt + i] = rho * cfg.R * temp;
        }
    }
    void prim_to_cons(const T* fs, T* cq) {
        const bool three_d = g->dim == 3;
        for (int64_t i = 0; i < Nt; i++) {
            T vx = fs[P_VX * Nt + i], vy = fs[P_VY * Nt + i], vz = fs[P_VZ * Nt + i];
            T rho = fs[P_RHO * Nt + i];
            cq[i_mass() * Nt + i] = fs[P_RHO * Nt + i];
            cq[i_mx() * Nt + i] = rho * vx;
            cq[i_my() * Nt + i] = rho * vy;
            if (three_d) cq[i_mz() * Nt + i] = rho * vz;
            T ke = 0.5 * (vx * vx + vy * vy + vz * vz);
            cq[i_en() * Nt + i] = rho * (ke + fs[P_E * Nt + i]);
        }
    }
};

struct FVHandle {
    Grid* g;
    FV<double>* fv;
    FV<Dual>* fvd;
    // RK scratch
    std::vector<std::vector<double>> k;
    std::vector<double> k_tmp, flow_tmp;
};

}  // namespace

extern "C" {

void* or_grid_build(const or_grid_io* io) { return build_grid(io); }
void or_grid_free(void* g) { delete static_cast<Grid*>(g); }

void or_grid_sizes(const void* gp, int64_t* out) {
    const Grid& g = *static_cast<const Grid*>(gp);
    int64_t nb = 0;
    for (auto& v : g.bfaces) nb += (int64_t)v.size();
    out[0] = g.dim;
    out[1] = g.N;
    out[2] = g.G;
    out[3] = g.F;
    out[4] = (int64_t)g.cf.size();
    out[5] = g.n_lsq;
    out[6] = g.ntags;
    out[7] = nb;
}

void or_grid_export(const void* gp, double* face_geom, int64_t* face_lr, int64_t* cell_off,
                    int64_t* cell_faces, int32_t* outsigns, int64_t* neighbours,
                    double* centroid, double* volume, double* lsq, int64_t* tag_off,
                    int64_t* bfaces, int64_t* ghosts) {
    const Grid& g = *static_cast<const Grid*>(gp);
    const int64_t F = g.F, Nt = g.N + g.G;
    if (face_geom) {
        const std::vector<double>* rows[13] = {&g.nx, &g.ny, &g.nz, &g.t1x, &g.t1y,
                                               &g.t1z, &g.t2x, &g.t2y, &g.t2z, &g.cx,
                                               &g.cy, &g.cz, &g.area};
        for (int k = 0; k < 13; k++) std::memcpy(face_geom + k * F, rows[k]->data(), F * 8);
    }
    if (face_lr) {
        for (int64_t f = 0; f < F; f++) {
            face_lr[f] = g.left[f] == NONE ? -1 : g.left[f];
            face_lr[F + f] = g.right[f] == NONE ? -1 : g.right[f];
        }
    }
    if (cell_off) std::memcpy(cell_off, g.cf_off.data(), (g.N + 1) * 8);
    if (cell_faces) std::memcpy(cell_faces, g.cf.data(), g.cf.size() * 8);
    if (outsigns) std::memcpy(outsigns, g.outsign.data(), g.outsign.size() * 4);
    if (neighbours)
        for (size_t j = 0; j < g.nbr.size(); j++)
            neighbours[j] = g.nbr[j] == NONE ? -1 : g.nbr[j];
    if (centroid) {
        std::memcpy(centroid, g.x.data(), Nt * 8);
        std::memcpy(centroid + Nt, g.y.data(), Nt * 8);
        std::memcpy(centroid + 2 * Nt, g.z.data(), Nt * 8);
    }
    if (volume) std::memcpy(volume, g.vol.data(), Nt * 8);
    if (lsq) std::memcpy(lsq, g.r.data(), g.r.size() * 8);
    if (tag_off) {
        int64_t o = 0;
        for (int t = 0; t < g.ntags; t++) {
            tag_off[t] = o;
            for (size_t i = 0; i < g.bfaces[t].size(); i++) {
                if (bfaces) bfaces[o + i] = g.bfaces[t][i] == NONE ? -1 : g.bfaces[t][i];
                if (ghosts) ghosts[o + i] = i < g.ghosts[t].size() ? g.ghosts[t][i] : -1;
            }
            o += (int64_t)g.bfaces[t].size();
        }
        tag_off[g.ntags] = o;
    }
}

void or_grid_gradient(const void* gp, const double* values, double* grad) {
    const Grid& g = *static_cast<const Grid*>(gp);
    lsq_gradient<double>(g, values, grad, grad + g.N, grad + 2 * g.N);
}

void* or_fv_create(void* g, const or_fv_config* cfg) {
    FVHandle* h = new FVHandle();
    h->g = static_cast<Grid*>(g);
    h->fv = new FV<double>(h->g, cfg);
    h->fvd = nullptr;
    return h;
}

void or_fv_free(void* p) {
    FVHandle* h = static_cast<FVHandle*>(p);
    delete h->fv;
    delete h->fvd;
    delete h;
}

int or_compute_dudt(void* p, double* prim, double* dudt, int allow_reconstruction) {
    return static_cast<FVHandle*>(p)->fv->compute_dudt(prim, dudt, allow_reconstruction != 0);
}

int or_compute_dudt_jvp(void* p, double* prim, double* prim_dot, double* dudt,
                        double* dudt_dot, int allow_reconstruction) {
    FVHandle* h = static_cast<FVHandle*>(p);
    if (!h->fvd) h->fvd = new FV<Dual>(h->g, &h->fv->cfg);
    FV<Dual>& fv = *h->fvd;
    const int64_t Nt = fv.Nt;
    std::vector<Dual> fs((size_t)NPRIM * Nt), du((size_t)fv.ncons * Nt);
    for (size_t i = 0; i < fs.size(); i++) fs[i] = Dual(prim[i], prim_dot[i]);
    for (size_t i = 0; i < du.size(); i++) du[i] = Dual(dudt[i], dudt_dot[i]);
    int rc = fv.compute_dudt(fs.data(), du.data(), allow_reconstruction != 0);
    for (size_t i = 0; i < fs.size(); i++) {
        prim[i] = fs[i].r;
        prim_dot[i] = fs[i].d;
    }
    for (size_t i = 0; i < du.size(); i++) {
        dudt[i] = du[i].r;
        dudt_dot[i] = du[i].d;
    }
    return rc;
}

double or_estimate_dt(void* p, const double* prim) {
    return static_cast<FVHandle*>(p)->fv->estimate_dt(prim);
}

int64_t or_count_bad_cells(void* p, const double* prim) {
    return static_cast<FVHandle*>(p)->fv->count_bad_cells(prim);
}

void or_cons_to_prim(void* p, const double* cons, double* prim) {
    static_cast<FVHandle*>(p)->fv->cons_to_prim(cons, prim);
}

void or_cons_to_prim_jvp(void* p, const double* cons, const double* cons_dot, double* prim,
                         double* prim_dot) {
    FVHandle* h = static_cast<FVHandle*>(p);
    if (!h->fvd) h->fvd = new FV<Dual>(h->g, &h->fv->cfg);
    FV<Dual>& fv = *h->fvd;
    const int64_t Nt = fv.Nt;
    std::vector<Dual> cq((size_t)fv.ncons * Nt), fs((size_t)NPRIM * Nt);
    for (size_t i = 0; i < cq.size(); i++) cq[i] = Dual(cons[i], cons_dot[i]);
    fv.cons_to_prim(cq.data(), fs.data());
    for (size_t i = 0; i < fs.size(); i++) {
        prim[i] = fs[i].r;
        prim_dot[i] = fs[i].d;
    }
}

void or_prim_to_cons(void* p, const double* prim, double* cons) {
    static_cast<FVHandle*>(p)->fv->prim_to_cons(prim, cons);
}

void or_compute_convective_gradient(void* p, double* prim) {
    static_cast<FVHandle*>(p)->fv->compute_convective_gradient(prim);
}

void or_compute_viscous_gradient(void* p, double* prim) {
    static_cast<FVHandle*>(p)->fv->compute_viscous_gradient(prim);
}

void or_fv_intermediates(void* p, double* grads, double* lim, double* flux, double* left,
                         double* right) {
    FV<double>& fv = *static_cast<FVHandle*>(p)->fv;
    if (grads) std::memcpy(grads, fv.grad.data(), fv.grad.size() * 8);
    if (lim) std::memcpy(lim, fv.lim.data(), fv.lim.size() * 8);
    if (flux) std::memcpy(flux, fv.flux.data(), fv.flux.size() * 8);
    if (left) std::memcpy(left, fv.left.data(), fv.left.size() * 8);
    if (right) std::memcpy(right, fv.right.data(), fv.right.size() * 8);
}

void or_fv_face_gradients(void* p, double* out) {
    FV<double>& fv = *static_cast<FVHandle*>(p)->fv;
    std::memcpy(out, fv.face_grad.data(), fv.face_grad.size() * 8);
}

// One flux calculator on one face, states already in the face frame
// (rho, p, T, e, u_n, u_t1, u_t2); out: mass, n-, t1-, t2-momentum, energy.
// The calculators read only the gas constants, the LDFSS delta and dim.
void or_flux_face(const or_fv_config* cfg, int dim, const double* L, const double* R,
                  double* out) {
    Grid g;
    g.dim = dim;
    FV<double> fv(&g, cfg);
    double l[NPRIM], r[NPRIM];
    for (int v = 0; v < NPRIM; v++) {
        l[v] = L[v];
        r[v] = R[v];
    }
    switch (cfg->flux) {
        case OR_FLUX_HANEL: fv.flux_hanel(l, r, out); break;
        case OR_FLUX_AUSMDV: fv.flux_ausmdv(l, r, out); break;
        case OR_FLUX_LDFSS2: fv.flux_ldfss2(l, r, out); break;
        default: fv.flux_rusanov(l, r, out); break;
    }
}

// Dual<double> arithmetic (UT/dual.h), for checking against known answers.
// a, b: (real, dual) pairs; out: (real, dual).  op: 0 a+b, 1 a-b, 2 a*b,
// 3 a/b, 4 sqrt(a), 5 abs(a), 6 pow(a, b), 7 min(a, b), 8 max(a, b),
// 9 floor(a), 10 copysign(a, b), 11 a+b.r, 12 a-b.r, 13 b.r-a, 14 a*b.r,
// 15 b.r*a, 16 a/b.r, 17 b.r/a, 18 a+=b, 19 a-=b
void or_dual_eval(int op, const double* a, const double* b, double* out) {
    Dual x(a[0], a[1]), y(b[0], b[1]), z;
    double yr = b[0];
    switch (op) {
        case 0: z = x + y; break;
        case 1: z = x - y; break;
        case 2: z = x * y; break;
        case 3: z = x / y; break;
        case 4: z = ksqrt(x); break;
        case 5: z = kabs(x); break;
        case 6: z = kpow(x, y); break;
        case 7: z = kmin(x, y); break;
        case 8: z = kmax(x, y); break;
        case 9: z = kfloor(x); break;
        case 10: z = kcopysign(x, y); break;
        case 11: z = x + yr; break;
        case 12: z = x - yr; break;
        case 13: z = yr - x; break;
        case 14: z = x * yr; break;
        case 15: z = yr * x; break;
        case 16: z = x / yr; break;
        case 17: z = yr / x; break;
        case 18: z = x; z += y; break;
        case 19: z = x; z -= y; break;
        default: z = Dual(std::nan(""), std::nan("")); break;
    }
    out[0] = z.r;
    out[1] = z.d;
}

// RungeKutta::take_step + estimate_dt, SO/runge_kutta.cpp:108-203 (static grid)
double or_rk_step(void* p, int s, const double* a, const double* b, double cfl, double* cons,
                  double* prim, double* dt_prev, double t, double max_time,
                  double plot_remaining) {
    FVHandle* h = static_cast<FVHandle*>(p);
    FV<double>& fv = *h->fv;
    const size_t ncv = (size_t)fv.ncons * fv.Nt;
    if (h->k.size() != (size_t)s) h->k.assign(s, std::vector<double>(ncv, 0.0));
    if (h->k_tmp.size() != ncv) h->k_tmp.assign(ncv, 0.0);
    if (h->flow_tmp.size() != (size_t)NPRIM * fv.Nt) h->flow_tmp.assign((size_t)NPRIM * fv.Nt, 0.0);
    fv.compute_dudt(prim, h->k[0].data(), true);
    double stable_dt = fv.estimate_dt(prim);
    double dt_startup = kmin(cfl * stable_dt, 1.5 * (*dt_prev));
    double dt = kmin(dt_startup, max_time - t);
    if (plot_remaining > 0.0) dt = kmin(dt, plot_remaining);
    for (int i = 1; i < s; i++) {
        // apply_time_derivative(U0, U1, dUdt, dt), FV/conserved_quantities.cpp:80-94
        double ai0 = a[i * s + 0] * dt;
        const std::vector<double>& k0 = h->k[0];
        for (size_t q = 0; q < ncv; q++) h->k_tmp[q] = cons[q] + k0[q] * ai0;
        for (int j = 1; j < i; j++) {
            if (a[i * s + j] < 1e-14) continue;  // SO/runge_kutta.cpp:166
            double aij = a[i * s + j] * dt;
            const std::vector<double>& kj = h->k[j];
            for (size_t q = 0; q < ncv; q++) h->k_tmp[q] += kj[q] * aij;
        }
        fv.cons_to_prim(h->k_tmp.data(), h->flow_tmp.data());
        fv.compute_dudt(h->flow_tmp.data(), h->k[i].data(), true);
    }
    for (int i = 0; i < s; i++) {
        double bi = b[i] * dt;
        const std::vector<double>& ki = h->k[i];
        for (size_t q = 0; q < ncv; q++) cons[q] += ki[q] * bi;
    }
    fv.cons_to_prim(cons, prim);
    *dt_prev = dt;
    return dt;
}

void or_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

}  // extern "C"
