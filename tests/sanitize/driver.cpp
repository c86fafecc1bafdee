// Sanitizer driver (CPU test): runs the host grid builder of libibis_gpu
// (ibis_amd/csrc/grid.cpp) and the CPU oracle under AddressSanitizer and
// UndefinedBehaviorSanitizer on a case file written by
// tests/test_sanitize.py, and cross-checks the two grid builds.
//
// Case file (little-endian): int32 dim, ntags; int64 nv; nv*3 f64 xyz;
// int64 nc; nc i32 ctype; nc*8 i64 cvtx; ntags i32 tag_ghost; int64 nb;
// nb i32 btag; nb i32 btype; nb*4 i64 bvtx; int64 cfg_bytes; the
// or_fv_config bytes; int64 n_prim; n_prim f64 prim (7 x N_total).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/ibis_gpu.h"
#include "../../oracle/ibis_oracle.h"

namespace {

struct Reader {
    FILE* f;
    template <class T>
    std::vector<T> vec(size_t n) {
        std::vector<T> v(n);
        if (n && fread(v.data(), sizeof(T), n, f) != n) {
            fprintf(stderr, "short read\n");
            exit(2);
        }
        return v;
    }
    template <class T>
    T one() {
        return vec<T>(1)[0];
    }
};

int fail(const char* what) {
    fprintf(stderr, "FAIL: %s\n", what);
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) return fail("usage: driver CASE_FILE");
    Reader rd{fopen(argv[1], "rb")};
    if (!rd.f) return fail("cannot open the case file");
    const int32_t dim = rd.one<int32_t>(), ntags = rd.one<int32_t>();
    const int64_t nv = rd.one<int64_t>();
    auto xyz = rd.vec<double>(nv * 3);
    const int64_t nc = rd.one<int64_t>();
    auto ctype = rd.vec<int32_t>(nc);
    auto cvtx = rd.vec<int64_t>(nc * 8);
    auto tag_ghost = rd.vec<int32_t>(ntags);
    const int64_t nb = rd.one<int64_t>();
    auto btag = rd.vec<int32_t>(nb);
    auto btype = rd.vec<int32_t>(nb);
    auto bvtx = rd.vec<int64_t>(nb * 4);
    const int64_t cfg_bytes = rd.one<int64_t>();
    if (cfg_bytes != (int64_t)sizeof(or_fv_config)) return fail("or_fv_config size mismatch");
    or_fv_config cfg;
    if (fread(&cfg, sizeof(cfg), 1, rd.f) != 1) return fail("short read (config)");
    const int64_t n_prim = rd.one<int64_t>();
    auto prim = rd.vec<double>(n_prim);
    fclose(rd.f);

    // the library's grid builder, reference numbering and renumbered
    ibis_grid_io io{dim, ntags, nv, xyz.data(), nc, ctype.data(), cvtx.data(),
                    tag_ghost.data(), nb, btag.data(), btype.data(), bvtx.data()};
    ibis_grid* g0 = nullptr;
    ibis_grid* g1 = nullptr;
    if (ibis_grid_build(&io, 0, &g0) != IBIS_OK) return fail(ibis_last_error());
    if (ibis_grid_build(&io, 1, &g1) != IBIS_OK) return fail(ibis_last_error());
    int64_t s0[8], s1[8];
    ibis_grid_sizes(g0, s0);
    ibis_grid_sizes(g1, s1);
    if (std::memcmp(s0, s1, sizeof(s0)) != 0) return fail("renumbered grid sizes differ");
    const int64_t N = s0[1], G = s0[2], F = s0[3], nnz = s0[4], n_lsq = s0[5], nbf = s0[7];
    const int64_t Nt = N + G;
    if (n_prim != 7 * Nt) return fail("prim size does not match the grid");
    std::vector<double> geom(13 * F), cen(3 * Nt), vol(Nt), lsq(n_lsq * N);
    std::vector<int64_t> lr(2 * F), off(N + 1), cf(nnz), nbr(nnz), toff(ntags + 1), bf(nbf),
        gh(nbf), cnew(Nt), fnew(F);
    std::vector<int32_t> os(nnz);
    ibis_grid_export(g0, geom.data(), lr.data(), off.data(), cf.data(), os.data(), nbr.data(),
                     cen.data(), vol.data(), lsq.data(), toff.data(), bf.data(), gh.data());
    ibis_grid_permutation(g1, cnew.data(), fnew.data());

    // the oracle's grid: same sizes and the same arrays bit for bit
    or_grid_io oio{dim, ntags, nv, xyz.data(), nc, ctype.data(), cvtx.data(),
                   tag_ghost.data(), nb, btag.data(), btype.data(), bvtx.data()};
    void* og = or_grid_build(&oio);
    if (!og) return fail("oracle grid build");
    int64_t so[8];
    or_grid_sizes(og, so);
    if (so[1] != N || so[2] != G || so[3] != F || so[4] != nnz) return fail("oracle sizes differ");
    std::vector<double> ogeom(13 * F), ocen(3 * Nt), ovol(Nt), olsq(n_lsq * N);
    std::vector<int64_t> olr(2 * F), ooff(N + 1), ocf(nnz), onbr(nnz), otoff(ntags + 1),
        obf(nbf), ogh(nbf);
    std::vector<int32_t> oos(nnz);
    or_grid_export(og, ogeom.data(), olr.data(), ooff.data(), ocf.data(), oos.data(),
                   onbr.data(), ocen.data(), ovol.data(), olsq.data(), otoff.data(), obf.data(),
                   ogh.data());
    if (geom != ogeom || lr != olr || off != ooff || cf != ocf || os != oos || nbr != onbr ||
        cen != ocen || vol != ovol || lsq != olsq || bf != obf || gh != ogh)
        return fail("grid builder differs from the oracle");

    // the oracle's residual, estimate_dt, tangent residual and one RK step
    void* fv = or_fv_create(og, &cfg);
    const int ncons = dim + 2;
    std::vector<double> p(prim), dudt(ncons * Nt, 0.0), dot(7 * Nt), ddot(ncons * Nt, 0.0),
        d2(ncons * Nt, 0.0);
    for (int64_t i = 0; i < 7 * Nt; i++) dot[i] = std::sin(0.37 * (double)i) * 1e-3 * p[i];
    or_compute_dudt(fv, p.data(), dudt.data(), 1);
    const double dt = or_estimate_dt(fv, p.data());
    // dt may be negative: the reference's 3D volume of a cell with non-planar
    // faces can be (jittered hexes, pyramids); it must still be a number
    if (!std::isfinite(dt)) return fail("estimate_dt");
    std::vector<double> pj(prim);
    or_compute_dudt_jvp(fv, pj.data(), dot.data(), d2.data(), ddot.data(), 1);
    for (int64_t q = 0; q < ncons; q++)
        for (int64_t c = 0; c < N; c++)
            if (d2[q * Nt + c] != dudt[q * Nt + c]) return fail("JVP value differs from dU/dt");
    if (or_count_bad_cells(fv, p.data()) != 0) return fail("bad cells");
    std::vector<double> cons(ncons * Nt, 0.0);
    or_prim_to_cons(fv, p.data(), cons.data());
    const double a[9] = {0, 0, 0, 1, 0, 0, 0.25, 0.25, 0};
    const double b[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};
    double dt_prev = 1e300;
    or_rk_step(fv, 3, a, b, 0.5, cons.data(), p.data(), &dt_prev, 0.0, 1e30, -1.0);
    double sum = 0.0;
    for (int64_t q = 0; q < ncons; q++)
        for (int64_t c = 0; c < N; c++) sum += std::fabs(dudt[q * Nt + c]);
    or_fv_free(fv);
    or_grid_free(og);
    ibis_grid_destroy(g0);
    ibis_grid_destroy(g1);
    printf("ok N=%lld F=%lld dt=%.17g sum|dudt|=%.17g\n", (long long)N, (long long)F, dt, sum);
    return std::isfinite(sum) ? 0 : fail("non-finite residual");
}
