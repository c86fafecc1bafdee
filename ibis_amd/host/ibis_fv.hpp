// ibis_fv.hpp -- C++ host layer over the C ABI of libibis_gpu
// (include/ibis_gpu.h), with the reference's class surface for the cell
// residual:
//
//   GridIO            src/grid/grid/grid_io.h:55 (SU2 ingest, grid_io.cpp:218-268)
//   GridBlock         src/grid/grid/grid.h:36-42 (GridBlock(file, json&))
//   IdealGas          src/gas/gas/gas_model.h:47-62
//   TransportProperties src/gas/gas/transport_properties.h:85-110
//   FlowStates, ConservedQuantities  gas/flow_state.h:51, finite_volume/conserved_quantities.h
//   FiniteVolume      src/finite_volume/finite_volume/finite_volume.h:28-142
//   ButcherTableau, RungeKutta::take_step  src/solvers/solvers/runge_kutta.h:14-41,
//                     runge_kutta.cpp:108-203
//
// State lives in device memory (HBM) in the grid's numbering: 7 x n_total
// primitive rows (rho, p, T, e, vx, vy, vz) and (dim+2) x n_total conserved
// rows (mass, momentum, energy).  GridBlock converts host arrays between
// the reference numbering and the grid's (the grid may be renumbered for
// locality; results do not depend on it).
//
// Errors: configuration errors throw std::runtime_error, as the reference's
// factories do; a failing library call throws ibis::Error with the
// library's message.
#ifndef IBIS_HOST_FV_HPP
#define IBIS_HOST_FV_HPP

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ibis_gpu.h"
#include "json.hpp"

namespace ibis {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// An unstructured grid as read from an SU2 file.
struct GridIO {
    int dim = 2;
    std::vector<double> xyz;      // n_vertices x 3
    std::vector<int32_t> ctype;   // SU2 element code per cell
    std::vector<int64_t> cvtx;    // n_cells x 8, -1 padded
    struct Marker {
        std::vector<int32_t> type;
        std::vector<int64_t> vtx;  // x 4, -1 padded
    };
    std::map<std::string, Marker> markers;

    GridIO() = default;
    explicit GridIO(const std::string& su2_file);
    size_t num_cells() const { return ctype.size(); }
    size_t num_vertices() const { return xyz.size() / 3; }
};

struct IdealGas {
    double R = 287.0, Cv = 2.5 * 287.0, Cp = 3.5 * 287.0, gamma = 3.5 / 2.5;
    IdealGas() = default;
    explicit IdealGas(double R_);
    explicit IdealGas(const Json& config);  // {"R", "Cv", "Cp", "gamma"}
};

// Sutherland viscosity and constant-Prandtl conductivity.
struct TransportProperties {
    double mu0 = 1.716e-5, T0 = 273.0, Ts = 110.0, Pr = 0.72;
    TransportProperties() = default;
    explicit TransportProperties(const Json& config);
};

class GridBlock {
public:
    // config: the job's "grid" object; its "boundaries" fix the tags (sorted)
    // and which of them carry ghost cells.  reorder: renumber for locality.
    GridBlock(const GridIO& grid_io, const Json& config, bool reorder = true);
    GridBlock(const std::string& file_name, const Json& config, bool reorder = true);
    ~GridBlock();
    GridBlock(const GridBlock&) = delete;
    GridBlock& operator=(const GridBlock&) = delete;

    size_t dim() const { return (size_t)dim_; }
    size_t num_cells() const { return (size_t)N_; }
    size_t num_ghost_cells() const { return (size_t)G_; }
    size_t num_total_cells() const { return (size_t)(N_ + G_); }
    size_t num_interfaces() const { return (size_t)F_; }
    const std::vector<std::string>& boundary_tags() const { return tags_; }
    // grid index of each reference cell (valid cells, then ghosts)
    const std::vector<int64_t>& cell_new_of_old() const { return cell_new_of_old_; }

    // rows x n_total arrays between the reference numbering and the grid's
    std::vector<double> to_grid_order(const std::vector<double>& ref, size_t rows) const;
    std::vector<double> to_reference_order(const std::vector<double>& grid, size_t rows) const;
    // rows x n_cells arrays (cell-only quantities such as gradients)
    std::vector<double> cells_to_reference_order(const std::vector<double>& grid,
                                                 size_t rows) const;

    ibis_mesh* mesh() const { return mesh_; }

private:
    void build(const GridIO& io, const Json& config, bool reorder);
    int dim_ = 2;
    int64_t N_ = 0, G_ = 0, F_ = 0;
    std::vector<std::string> tags_;
    std::vector<int64_t> cell_new_of_old_;
    ibis_grid* grid_ = nullptr;
    ibis_mesh* mesh_ = nullptr;
};

// rows x n_total doubles in device memory.
class DeviceRows {
public:
    DeviceRows(size_t rows, size_t n_total);
    ~DeviceRows();
    DeviceRows(const DeviceRows&) = delete;
    DeviceRows& operator=(const DeviceRows&) = delete;
    double* data() { return p_; }
    const double* data() const { return p_; }
    size_t rows() const { return rows_; }
    size_t size() const { return n_; }
    void copy_from_host(const std::vector<double>& h);  // synchronous
    std::vector<double> copy_to_host() const;           // synchronous

private:
    double* p_ = nullptr;
    size_t rows_, n_;
};

struct FlowStates : DeviceRows {
    explicit FlowStates(size_t n_total) : DeviceRows(7, n_total) {}
};

struct ConservedQuantities : DeviceRows {
    ConservedQuantities(size_t n_total, size_t dim) : DeviceRows(dim + 2, n_total) {}
};

// The cell gradients (device): 7 slots (p, rho, T, u, vx, vy, vz) x 3
// components x n_cells, slot-major.
struct Gradients {
    const double* data = nullptr;
    size_t num_cells = 0;
    std::vector<double> copy_to_host() const;
};

class FiniteVolume {
public:
    // config: the job config (convective_flux, viscous_flux, gas_model,
    // transport_properties, grid.boundaries)
    FiniteVolume(GridBlock& grid, const Json& config);
    ~FiniteVolume();
    FiniteVolume(const FiniteVolume&) = delete;
    FiniteVolume& operator=(const FiniteVolume&) = delete;

    // FiniteVolume::compute_dudt (finite_volume.cpp:57-96, static grid).
    // Writes the ghost rows of flow_state.  Returns 0, as the reference does.
    size_t compute_dudt(FlowStates& flow_state, GridBlock& grid, ConservedQuantities& dudt,
                        IdealGas& gas_model, TransportProperties& trans_prop,
                        bool allow_reconstruction = true);
    double estimate_dt(const FlowStates& flow_state, GridBlock& grid, IdealGas& gas_model,
                       TransportProperties& trans_prop);
    size_t count_bad_cells(const FlowStates& fs, size_t num_cells);

    const Gradients& cell_gradients() const { return grads_; }
    void compute_viscous_gradient(FlowStates& fs, const GridBlock& grid,
                                  const IdealGas& gas_model, const TransportProperties& trans_prop);
    void compute_convective_gradient(FlowStates& fs, const GridBlock& grid,
                                     const IdealGas& gas_model,
                                     const TransportProperties& trans_prop);

    void primatives_to_conserved(ConservedQuantities& cq, const FlowStates& fs);
    void conserved_to_primatives(const ConservedQuantities& cq, FlowStates& fs);

    ibis_residual* handle() { return r_; }
    const IdealGas& gas_model() const { return gas_; }
    const TransportProperties& trans_prop() const { return tp_; }

private:
    void check_models(const IdealGas& gm, const TransportProperties& tp) const;
    ibis_residual* r_ = nullptr;
    size_t n_total_ = 0, dim_ = 2;
    IdealGas gas_;
    TransportProperties tp_;
    Gradients grads_;
};

class ButcherTableau {
public:
    // a: stages-1 rows (row i-1 holds a(i, 0..i-1)), b, c -- the reference's
    // json form {"a", "b", "c"} (runge_kutta.h:21-25)
    ButcherTableau(std::vector<std::vector<double>> a, std::vector<double> b,
                   std::vector<double> c);
    explicit ButcherTableau(const Json& config);
    static ButcherTableau named(const std::string& method);  // euler, midpoint, rk3, ssp-rk3, rk4
    size_t num_stages() const { return b_.size(); }
    // s x s, row i = the stage-i weights of k_0..k_{i-1}
    std::vector<double> full_a() const;
    const std::vector<double>& b() const { return b_; }

private:
    std::vector<std::vector<double>> a_;
    std::vector<double> b_, c_;
};

// The explicit step of RungeKutta::take_step (runge_kutta.cpp:108-203) on
// device state: dt = min(cfl dt_stable, 1.5 dt_prev, max_time - t[, plot]).
class RungeKutta {
public:
    RungeKutta(FiniteVolume& fv, ButcherTableau tableau, double cfl, double max_time,
               double plot_frequency = -1.0, double dt_init = -1.0);
    double take_step(ConservedQuantities& cq, FlowStates& fs);
    double t() const { return t_; }
    double dt() const { return dt_; }

private:
    FiniteVolume& fv_;
    ButcherTableau tab_;
    std::vector<double> a_;
    double cfl_, max_time_, plot_frequency_;
    double t_ = 0.0, dt_, since_plot_ = 0.0;
};

}  // namespace ibis

#endif
