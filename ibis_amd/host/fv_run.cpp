// fv_run.cpp -- command-line driver of the C++ host layer: a reference job's
// grid (SU2) and config (config.json) in, the residual and its companions
// out, all through FiniteVolume / RungeKutta (ibis_fv.hpp) on the GPU.
//
//   ibis_fv_run GRID.su2 CONFIG.json PRIM.bin OUT_PREFIX [STEPS [METHOD [CFL]]]
//
// PRIM.bin: 7 x n_total FP64 primitive rows (rho, p, T, e, vx, vy, vz) in
// the reference's cell numbering (ghost rows are filled by the boundary
// actions).  Writes, in the reference numbering:
//   OUT_PREFIX.dudt.bin  (dim+2) x n_total  dU/dt (valid rows)
//   OUT_PREFIX.prim.bin  7 x n_total        the state after compute_dudt (ghost rows)
//   OUT_PREFIX.cgrad.bin 7 x 3 x n_cells    cell_gradients() after compute_convective_gradient
//   OUT_PREFIX.vgrad.bin 7 x 3 x n_cells    cell_gradients() after compute_viscous_gradient
//   OUT_PREFIX.cons.bin  (dim+2) x n_total  after STEPS RungeKutta steps (METHOD, CFL)
//   OUT_PREFIX.txt       dt, bad cells, then one line per step: dt taken
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "ibis_fv.hpp"

namespace {

std::vector<double> read_doubles(const std::string& path, size_t n) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::vector<double> v(n);
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(double)));
    if ((size_t)f.gcount() != n * sizeof(double))
        throw std::runtime_error(path + ": expected " + std::to_string(n) + " doubles");
    return v;
}

void write_doubles(const std::string& path, const std::vector<double>& v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
    if (!f) throw std::runtime_error("cannot write " + path);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 5) {
        std::cerr << "usage: ibis_fv_run GRID.su2 CONFIG.json PRIM.bin OUT_PREFIX "
                     "[STEPS [METHOD [CFL]]]\n";
        return 2;
    }
    try {
        const std::string out = argv[4];
        const int steps = argc > 5 ? std::stoi(argv[5]) : 0;
        const std::string method = argc > 6 ? argv[6] : "ssp-rk3";
        const double cfl = argc > 7 ? std::stod(argv[7]) : 0.5;

        const ibis::Json config = ibis::Json::parse_file(argv[2]);
        ibis::GridBlock grid(argv[1], config.at("grid"));
        ibis::FiniteVolume fv(grid, config);
        ibis::IdealGas gas = fv.gas_model();
        ibis::TransportProperties tp = fv.trans_prop();
        const size_t nt = grid.num_total_cells(), nc = grid.dim() + 2;
        std::cout << "grid: dim " << grid.dim() << ", " << grid.num_cells() << " cells, "
                  << grid.num_ghost_cells() << " ghosts, " << grid.num_interfaces() << " faces"
                  << std::endl;
        const std::vector<double> prim0 = grid.to_grid_order(read_doubles(argv[3], 7 * nt), 7);

        ibis::FlowStates fs(nt);
        ibis::ConservedQuantities dudt(nt, grid.dim());
        fs.copy_from_host(prim0);
        fv.compute_dudt(fs, grid, dudt, gas, tp);
        const double dt = fv.estimate_dt(fs, grid, gas, tp);
        const size_t bad = fv.count_bad_cells(fs, grid.num_cells());
        write_doubles(out + ".dudt.bin", grid.to_reference_order(dudt.copy_to_host(), nc));
        write_doubles(out + ".prim.bin", grid.to_reference_order(fs.copy_to_host(), 7));

        fs.copy_from_host(prim0);
        fv.compute_convective_gradient(fs, grid, gas, tp);
        write_doubles(out + ".cgrad.bin",
                      grid.cells_to_reference_order(fv.cell_gradients().copy_to_host(), 21));
        fs.copy_from_host(prim0);
        fv.compute_viscous_gradient(fs, grid, gas, tp);
        write_doubles(out + ".vgrad.bin",
                      grid.cells_to_reference_order(fv.cell_gradients().copy_to_host(), 21));

        std::ofstream txt(out + ".txt");
        txt.precision(17);
        txt << dt << "\n" << bad << "\n";
        fs.copy_from_host(prim0);
        ibis::ConservedQuantities cq(nt, grid.dim());
        fv.primatives_to_conserved(cq, fs);
        ibis::RungeKutta rk(fv, ibis::ButcherTableau::named(method), cfl, 1e30);
        for (int s = 0; s < steps; s++) txt << rk.take_step(cq, fs) << "\n";
        write_doubles(out + ".cons.bin", grid.to_reference_order(cq.copy_to_host(), nc));
        std::cout << "residual, dt " << dt << ", " << bad << " bad cells, " << steps
                  << " " << method << " steps" << std::endl;
    } catch (const std::exception& e) {
        std::cerr << "ibis_fv_run: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
