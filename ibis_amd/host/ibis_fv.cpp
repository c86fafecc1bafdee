// ibis_fv.cpp -- the C++ host layer of ibis_fv.hpp.
#include "ibis_fv.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>

namespace ibis {

namespace {

void check(ibis_status s, const char* what) {
    if (s != IBIS_OK) throw Error(std::string(what) + ": " + ibis_last_error());
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

// vertices of an SU2 element (src/grid/grid/grid_io.cpp:11-40)
int su2_vertices(int code) {
    switch (code) {
        case 3: return 2;   // line
        case 5: return 3;   // triangle
        case 9: return 4;   // quadrilateral
        case 10: return 4;  // tetrahedron
        case 12: return 8;  // hexahedron
        case 13: return 6;  // wedge
        case 14: return 5;  // pyramid
    }
    throw std::runtime_error("Unknown SU2 element type " + std::to_string(code));
}

std::vector<std::string> tokens(const std::string& line) {
    std::string s = line;
    std::replace(s.begin(), s.end(), '=', ' ');
    std::istringstream in(s);
    std::vector<std::string> out;
    std::string t;
    while (in >> t) out.push_back(t);
    return out;
}

int flux_kind(const std::string& s) {
    if (s == "hanel") return IBIS_FLUX_HANEL;
    if (s == "ausmdv") return IBIS_FLUX_AUSMDV;
    if (s == "ldfss2") return IBIS_FLUX_LDFSS2;
    if (s == "rusanov") return IBIS_FLUX_RUSANOV;
    throw std::runtime_error("Unknown flux calculator " + s);
}

int thermo_kind(const std::string& s) {
    if (s == "rho_p") return IBIS_THERMO_RHO_P;
    if (s == "rho_T") return IBIS_THERMO_RHO_T;
    if (s == "rho_u") return IBIS_THERMO_RHO_U;
    if (s == "p_T") return IBIS_THERMO_P_T;
    throw std::runtime_error("Unknown thermo interpolator " + s);
}

int limiter_kind(const std::string& s) {
    if (s == "unlimited") return IBIS_LIM_UNLIMITED;
    if (s == "barth_jespersen") return IBIS_LIM_BARTH_JESPERSEN;
    if (s == "venkatakrishnan") return IBIS_LIM_VENKATAKRISHNAN;
    throw std::runtime_error("Unknown limiter " + s);
}

// one boundary action (boundaries/boundary.cpp:449-492) as kind + parameters
int action(const Json& a, double* par) {
    const std::string& t = a.at("type").string();
    auto flow_state = [&](int kind) {
        const Json& f = a.at("flow_state");
        const char* keys[7] = {"T", "p", "rho", "energy", "vx", "vy", "vz"};
        for (int k = 0; k < 7; k++) par[k] = f.at(keys[k]).number();
        return kind;
    };
    if (t == "flow_state_copy") return flow_state(IBIS_BC_FLOW_STATE_COPY);
    if (t == "internal_copy") return IBIS_BC_INTERNAL_COPY;
    if (t == "internal_copy_reflect_normal") return IBIS_BC_REFLECT_NORMAL;
    if (t == "internal_vel_copy_reflect") return IBIS_BC_VEL_REFLECT;
    if (t == "fix_temperature") {
        par[0] = a.at("temperature").number();
        return IBIS_BC_FIX_TEMPERATURE;
    }
    if (t == "subsonic_inflow") return flow_state(IBIS_BC_SUBSONIC_INFLOW);
    if (t == "subsonic_outflow") {
        par[0] = a.at("pressure").number();
        return IBIS_BC_SUBSONIC_OUTFLOW;
    }
    if (t == "constant_flux") return flow_state(IBIS_BC_CONSTANT_FLUX);
    throw std::runtime_error("Unknown boundary action " + t);
}

// FiniteVolume(GridBlock&, json) configuration (finite_volume.cpp:13-54,
// convective_flux.cpp:61-82) as the C ABI's descriptor
ibis_fv_desc make_desc(const Json& config, IdealGas& gas, TransportProperties& tp) {
    ibis_fv_desc d;
    std::memset(&d, 0, sizeof(d));
    const Json& cf = config.at("convective_flux");
    const Json& vf = config.at("viscous_flux");
    const Json& fc = cf.at("flux_calculator");
    d.flux = flux_kind(fc.at("type").string());
    d.ldfss_delta = fc.value("delta", 2.0);
    d.order = (int32_t)cf.at("reconstruction_order").number();
    if (d.order != 1 && d.order != 2)
        throw std::runtime_error("Invalid reconstruction order " + std::to_string(d.order));
    d.thermo = IBIS_THERMO_RHO_U;
    d.limiter = IBIS_LIM_BARTH_JESPERSEN;
    d.lim_eps = 1e-25;
    d.venkat_K = 5.0;
    if (d.order > 1) {
        d.thermo = thermo_kind(cf.at("thermo_interpolator").string());
        const Json& lim = cf.at("limiter");
        d.limiter = limiter_kind(lim.at("type").string());
        d.lim_eps = lim.value("epsilon", 1e-25);
        d.venkat_K = lim.value("K", 5.0);
    }
    d.viscous = vf.at("enabled").boolean() ? 1 : 0;
    d.signal_factor = vf.at("signal_factor").number();
    if (config.contains("gas_model")) gas = IdealGas(config.at("gas_model"));
    if (config.contains("transport_properties"))
        tp = TransportProperties(config.at("transport_properties"));
    d.R = gas.R;
    d.Cv = gas.Cv;
    d.Cp = gas.Cp;
    d.gamma = gas.gamma;
    d.mu0 = tp.mu0;
    d.T0 = tp.T0;
    d.Ts = tp.Ts;
    d.Pr = tp.Pr;
    const auto& bcs = config.at("grid").at("boundaries").object();
    if (bcs.size() > IBIS_MAX_TAGS)
        throw std::runtime_error("at most " + std::to_string(IBIS_MAX_TAGS) + " boundary tags");
    d.ntags = (int32_t)bcs.size();
    int t = 0;
    for (const auto& kv : bcs) {  // sorted tags
        struct List {
            const char* key;
            int32_t* kind;
            double* par;
        } lists[3] = {{"pre_reconstruction", d.pre_kind, d.pre_par},
                      {"pre_viscous_grad", d.visc_kind, d.visc_par},
                      {"post_convective_flux", d.post_kind, d.post_par}};
        for (const List& l : lists) {
            if (!kv.second.contains(l.key)) continue;
            const auto& acts = kv.second.at(l.key).array();
            if (acts.size() > IBIS_MAX_ACT)
                throw std::runtime_error("at most " + std::to_string(IBIS_MAX_ACT) +
                                         " actions per boundary list");
            for (size_t a = 0; a < acts.size(); a++) {
                const size_t slot = (size_t)t * IBIS_MAX_ACT + a;
                l.kind[slot] = action(acts[a], l.par + slot * IBIS_NPAR);
            }
        }
        t++;
    }
    return d;
}

}  // namespace

// ---- GridIO: SU2 reader (grid_io.cpp:218-268) ------------------------------

GridIO::GridIO(const std::string& su2_file) {
    std::ifstream f(su2_file);
    if (!f) throw std::runtime_error("cannot open " + su2_file);
    std::vector<std::string> lines;
    for (std::string ln; std::getline(f, ln);) {
        const size_t a = ln.find_first_not_of(" \t\r");
        if (a == std::string::npos) continue;
        const size_t b = ln.find_last_not_of(" \t\r");
        ln = ln.substr(a, b - a + 1);
        if (ln[0] == '%') continue;
        lines.push_back(ln);
    }
    size_t i = 0;
    auto ints = [](const std::string& ln) {
        std::istringstream in(ln);
        std::vector<int64_t> v;
        for (int64_t x; in >> x;) v.push_back(x);
        return v;
    };
    while (i < lines.size()) {
        const std::string& ln = lines[i];
        if (ln.rfind("NDIME", 0) == 0) {
            dim = std::stoi(tokens(ln).at(1));
            i++;
        } else if (ln.rfind("NPOIN", 0) == 0) {
            const size_t n = std::stoul(tokens(ln).at(1));
            xyz.assign(3 * n, 0.0);
            for (size_t k = 0; k < n; k++) {
                std::istringstream in(lines.at(i + 1 + k));
                in >> xyz[3 * k] >> xyz[3 * k + 1];
                if (dim == 3) in >> xyz[3 * k + 2];
            }
            i += n + 1;
        } else if (ln.rfind("NELEM", 0) == 0) {
            const size_t n = std::stoul(tokens(ln).at(1));
            ctype.assign(n, 0);
            cvtx.assign(8 * n, -1);
            for (size_t k = 0; k < n; k++) {
                const std::vector<int64_t> v = ints(lines.at(i + 1 + k));
                ctype[k] = (int32_t)v.at(0);
                const int nv = su2_vertices(ctype[k]);
                for (int j = 0; j < nv; j++) cvtx[8 * k + j] = v.at(1 + j);
            }
            i += n + 1;
        } else if (ln.rfind("NMARK", 0) == 0) {
            const int nm = std::stoi(tokens(ln).at(1));
            i++;
            for (int m = 0; m < nm; m++) {
                const std::string tag = tokens(lines.at(i)).at(1);
                const size_t ne = std::stoul(tokens(lines.at(i + 1)).at(1));
                Marker& mk = markers[tag];  // a repeated tag appends (grid_io.cpp:254-264)
                for (size_t k = 0; k < ne; k++) {
                    const std::vector<int64_t> v = ints(lines.at(i + 2 + k));
                    mk.type.push_back((int32_t)v.at(0));
                    const int nv = su2_vertices((int)v.at(0));
                    for (int j = 0; j < 4; j++) mk.vtx.push_back(j < nv ? v.at(1 + j) : -1);
                }
                i += ne + 2;
            }
        } else {
            i++;
        }
    }
}

IdealGas::IdealGas(double R_) : R(R_), Cv(5.0 / 2.0 * R_), Cp(7.0 / 2.0 * R_) {
    gamma = Cp / Cv;
}

IdealGas::IdealGas(const Json& c)
    : R(c.at("R").number()), Cv(c.at("Cv").number()), Cp(c.at("Cp").number()),
      gamma(c.at("gamma").number()) {}

TransportProperties::TransportProperties(const Json& c) {
    const Json& v = c.at("viscosity");
    const Json& k = c.at("thermal_conductivity");
    if (v.at("type").string() != "sutherland")
        throw std::runtime_error("Unknown viscosity model " + v.at("type").string());
    if (k.at("type").string() != "constant_prandtl_number")
        throw std::runtime_error("Unknown thermal conductivity model " + k.at("type").string());
    mu0 = v.at("mu_0").number();
    T0 = v.at("T_0").number();
    Ts = v.at("T_s").number();
    Pr = k.at("Pr").number();
}

// ---- GridBlock ----------------------------------------------------------------

GridBlock::GridBlock(const GridIO& grid_io, const Json& config, bool reorder) {
    build(grid_io, config, reorder);
}

GridBlock::GridBlock(const std::string& file_name, const Json& config, bool reorder) {
    build(GridIO(file_name), config, reorder);
}

void GridBlock::build(const GridIO& io, const Json& config, bool reorder) {
    const auto& bcs = config.at("boundaries").object();
    std::vector<int32_t> ghost, btag, btype;
    std::vector<int64_t> bvtx;
    int t = 0;
    for (const auto& kv : bcs) {  // tags in sorted order (grid.h:414)
        tags_.push_back(kv.first);
        ghost.push_back(kv.second.value("ghost_cells", true) ? 1 : 0);
        auto it = io.markers.find(kv.first);
        if (it != io.markers.end()) {
            for (size_t k = 0; k < it->second.type.size(); k++) {
                btag.push_back(t);
                btype.push_back(it->second.type[k]);
                for (int j = 0; j < 4; j++) bvtx.push_back(it->second.vtx[4 * k + j]);
            }
        }
        t++;
    }
    ibis_grid_io g{io.dim,          (int32_t)tags_.size(), (int64_t)io.num_vertices(),
                   io.xyz.data(),   (int64_t)io.num_cells(), io.ctype.data(),
                   io.cvtx.data(),  ghost.data(),          (int64_t)btag.size(),
                   btag.data(),     btype.data(),          bvtx.data()};
    check(ibis_grid_build(&g, reorder ? 1 : 0, &grid_), "ibis_grid_build");
    int64_t sz[8];
    check(ibis_grid_sizes(grid_, sz), "ibis_grid_sizes");
    dim_ = (int)sz[0];
    N_ = sz[1];
    G_ = sz[2];
    F_ = sz[3];
    cell_new_of_old_.assign((size_t)(N_ + G_), 0);
    check(ibis_grid_permutation(grid_, cell_new_of_old_.data(), nullptr), "ibis_grid_permutation");
    check(ibis_mesh_from_grid(grid_, &mesh_), "ibis_mesh_from_grid");
}

GridBlock::~GridBlock() {
    if (mesh_) ibis_mesh_destroy(mesh_);
    if (grid_) ibis_grid_destroy(grid_);
}

std::vector<double> GridBlock::to_grid_order(const std::vector<double>& ref, size_t rows) const {
    const size_t nt = num_total_cells();
    if (ref.size() != rows * nt) throw std::invalid_argument("to_grid_order: size mismatch");
    std::vector<double> out(ref.size());
    for (size_t r = 0; r < rows; r++)
        for (size_t c = 0; c < nt; c++) out[r * nt + (size_t)cell_new_of_old_[c]] = ref[r * nt + c];
    return out;
}

std::vector<double> GridBlock::to_reference_order(const std::vector<double>& grid,
                                                  size_t rows) const {
    const size_t nt = num_total_cells();
    if (grid.size() != rows * nt) throw std::invalid_argument("to_reference_order: size mismatch");
    std::vector<double> out(grid.size());
    for (size_t r = 0; r < rows; r++)
        for (size_t c = 0; c < nt; c++) out[r * nt + c] = grid[r * nt + (size_t)cell_new_of_old_[c]];
    return out;
}

std::vector<double> GridBlock::cells_to_reference_order(const std::vector<double>& grid,
                                                        size_t rows) const {
    const size_t n = num_cells();
    if (grid.size() != rows * n)
        throw std::invalid_argument("cells_to_reference_order: size mismatch");
    std::vector<double> out(grid.size());
    for (size_t r = 0; r < rows; r++)
        for (size_t c = 0; c < n; c++) out[r * n + c] = grid[r * n + (size_t)cell_new_of_old_[c]];
    return out;
}

// ---- device state -------------------------------------------------------------

DeviceRows::DeviceRows(size_t rows, size_t n_total) : rows_(rows), n_(rows * n_total) {
    hip_check(hipMalloc((void**)&p_, std::max<size_t>(n_, 1) * sizeof(double)), "hipMalloc");
    hip_check(hipMemset(p_, 0, std::max<size_t>(n_, 1) * sizeof(double)), "hipMemset");
}

DeviceRows::~DeviceRows() {
    if (p_) (void)hipFree(p_);
}

void DeviceRows::copy_from_host(const std::vector<double>& h) {
    if (h.size() != n_) throw std::invalid_argument("copy_from_host: size mismatch");
    hip_check(hipMemcpy(p_, h.data(), n_ * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
}

std::vector<double> DeviceRows::copy_to_host() const {
    std::vector<double> h(n_);
    hip_check(hipMemcpy(h.data(), p_, n_ * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    return h;
}

std::vector<double> Gradients::copy_to_host() const {
    std::vector<double> h(7 * 3 * num_cells);
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    hip_check(hipMemcpy(h.data(), data, h.size() * sizeof(double), hipMemcpyDeviceToHost),
              "hipMemcpy");
    return h;
}

// ---- FiniteVolume ---------------------------------------------------------------

FiniteVolume::FiniteVolume(GridBlock& grid, const Json& config)
    : n_total_(grid.num_total_cells()), dim_(grid.dim()) {
    ibis_fv_desc d = make_desc(config, gas_, tp_);
    check(ibis_residual_create(grid.mesh(), &d, &r_), "ibis_residual_create");
    check(ibis_cell_gradients(r_, &grads_.data), "ibis_cell_gradients");
    grads_.num_cells = grid.num_cells();
}

FiniteVolume::~FiniteVolume() {
    if (r_) ibis_residual_destroy(r_);
}

void FiniteVolume::check_models(const IdealGas& gm, const TransportProperties& tp) const {
    // the device residual holds the configuration's gas and transport models
    if (gm.R != gas_.R || gm.Cv != gas_.Cv || gm.Cp != gas_.Cp || gm.gamma != gas_.gamma)
        throw std::invalid_argument("gas model differs from the one in the configuration");
    if (tp.mu0 != tp_.mu0 || tp.T0 != tp_.T0 || tp.Ts != tp_.Ts || tp.Pr != tp_.Pr)
        throw std::invalid_argument(
            "transport properties differ from the ones in the configuration");
}

size_t FiniteVolume::compute_dudt(FlowStates& fs, GridBlock& grid, ConservedQuantities& dudt,
                                  IdealGas& gm, TransportProperties& tp,
                                  bool allow_reconstruction) {
    check_models(gm, tp);
    if (fs.size() != 7 * grid.num_total_cells() || dudt.size() != (dim_ + 2) * n_total_)
        throw std::invalid_argument("compute_dudt: state sizes do not match the grid");
    check(ibis_compute_dudt(r_, fs.data(), dudt.data(), allow_reconstruction ? 1 : 0, nullptr),
          "ibis_compute_dudt");
    return 0;
}

double FiniteVolume::estimate_dt(const FlowStates& fs, GridBlock& grid, IdealGas& gm,
                                 TransportProperties& tp) {
    check_models(gm, tp);
    if (fs.size() != 7 * grid.num_total_cells())
        throw std::invalid_argument("estimate_dt: state size does not match the grid");
    double dt = 0.0;
    check(ibis_estimate_dt(r_, fs.data(), &dt, nullptr), "ibis_estimate_dt");
    return dt;
}

size_t FiniteVolume::count_bad_cells(const FlowStates& fs, size_t num_cells) {
    (void)num_cells;  // the residual's mesh fixes the valid cells
    uint64_t n = 0;
    check(ibis_count_bad_cells(r_, fs.data(), &n, nullptr), "ibis_count_bad_cells");
    return (size_t)n;
}

void FiniteVolume::compute_viscous_gradient(FlowStates& fs, const GridBlock&,
                                            const IdealGas& gm, const TransportProperties& tp) {
    check_models(gm, tp);
    check(ibis_compute_viscous_gradient(r_, fs.data(), nullptr), "ibis_compute_viscous_gradient");
}

void FiniteVolume::compute_convective_gradient(FlowStates& fs, const GridBlock&,
                                               const IdealGas& gm,
                                               const TransportProperties& tp) {
    check_models(gm, tp);
    check(ibis_compute_convective_gradient(r_, fs.data(), nullptr),
          "ibis_compute_convective_gradient");
}

void FiniteVolume::primatives_to_conserved(ConservedQuantities& cq, const FlowStates& fs) {
    check(ibis_prim_to_cons(r_, fs.data(), cq.data(), nullptr), "ibis_prim_to_cons");
}

void FiniteVolume::conserved_to_primatives(const ConservedQuantities& cq, FlowStates& fs) {
    check(ibis_cons_to_prim(r_, cq.data(), fs.data(), nullptr), "ibis_cons_to_prim");
}

// ---- Runge-Kutta ----------------------------------------------------------------

ButcherTableau::ButcherTableau(std::vector<std::vector<double>> a, std::vector<double> b,
                               std::vector<double> c)
    : a_(std::move(a)), b_(std::move(b)), c_(std::move(c)) {
    if (b_.empty() || b_.size() > 8) throw std::runtime_error("Butcher tableau: 1 to 8 stages");
    if (a_.size() + 1 != b_.size())
        throw std::runtime_error("Butcher tableau: a needs one row per stage after the first");
}

ButcherTableau::ButcherTableau(const Json& config)
    : ButcherTableau(
          [&] {
              std::vector<std::vector<double>> a;
              for (const Json& row : config.at("a").array()) {
                  a.emplace_back();
                  for (const Json& x : row.array()) a.back().push_back(x.number());
              }
              return a;
          }(),
          [&] {
              std::vector<double> b;
              for (const Json& x : config.at("b").array()) b.push_back(x.number());
              return b;
          }(),
          [&] {
              std::vector<double> c;
              for (const Json& x : config.at("c").array()) c.push_back(x.number());
              return c;
          }()) {}

ButcherTableau ButcherTableau::named(const std::string& m) {
    // the job DSL's tableaux (src/ibis/ibis/commands/prep/prep.py:845-859)
    if (m == "euler") return ButcherTableau({}, {1.0}, {});
    if (m == "midpoint") return ButcherTableau({{0.5}}, {0.0, 1.0}, {0.5});
    if (m == "rk3") return ButcherTableau({{0.5}, {-1.0, 2.0}}, {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0}, {0.5, 1.0});
    if (m == "ssp-rk3") return ButcherTableau({{1.0}, {0.25, 0.25}}, {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0}, {1.0, 0.5});
    if (m == "rk4")
        return ButcherTableau({{0.5}, {0.0, 0.5}, {0.0, 0.0, 1.0}},
                              {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0}, {0.5, 0.5, 1.0});
    throw std::runtime_error("Unknown method " + m);
}

std::vector<double> ButcherTableau::full_a() const {
    const size_t s = num_stages();
    std::vector<double> a(s * s, 0.0);
    for (size_t i = 1; i < s; i++)
        for (size_t j = 0; j < a_[i - 1].size() && j < i; j++) a[i * s + j] = a_[i - 1][j];
    return a;
}

RungeKutta::RungeKutta(FiniteVolume& fv, ButcherTableau tableau, double cfl, double max_time,
                       double plot_frequency, double dt_init)
    : fv_(fv), tab_(std::move(tableau)), a_(tab_.full_a()), cfl_(cfl), max_time_(max_time),
      plot_frequency_(plot_frequency),
      dt_(dt_init > 0 ? dt_init : std::numeric_limits<double>::max()) {}

double RungeKutta::take_step(ConservedQuantities& cq, FlowStates& fs) {
    double plot_remaining = -1.0;
    if (plot_frequency_ > 0.0 && since_plot_ < plot_frequency_)
        plot_remaining = plot_frequency_ - since_plot_;
    double taken = 0.0;
    check(ibis_rk_step(fv_.handle(), (int)tab_.num_stages(), a_.data(), tab_.b().data(), cfl_,
                       cq.data(), fs.data(), &dt_, t_, max_time_, plot_remaining, &taken, nullptr),
          "ibis_rk_step");
    t_ += taken;
    since_plot_ += taken;
    return taken;
}

}  // namespace ibis
