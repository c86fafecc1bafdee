// python_api.cpp -- the pybind11 module of the library (python_api), with
// the reference's binding surface (src/python/python/python_api.cpp:14-70:
// GasState, PyIdealGas, Vector3, the flux calculator classes) plus the GPU
// residual: GridBlock and FiniteVolume of the C++ host layer, whose
// methods take device pointers (torch's tensor.data_ptr()) and a stream
// handle, and the adjoint entry points the torch.autograd.Function wrappers
// (ibis_amd/autograd.py) call.
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <array>
#include <cmath>
#include <string>
#include <vector>

#include "ibis_fv.hpp"

namespace py = pybind11;

namespace {

// GasState (src/gas/gas/gas_state.h): one state's thermodynamic variables
struct GasState {
    double pressure = 0.0, temp = 0.0, rho = 0.0, energy = 0.0;
};

struct Vector3 {
    double x = 0.0, y = 0.0, z = 0.0;
};

// IdealGas closures, src/gas/gas/gas_model.h:11-40,91-116, same operation order
struct PyIdealGas {
    ibis::IdealGas g;
    explicit PyIdealGas(double R) : g(R) {}
    double speed_of_sound(const GasState& gs) const { return std::sqrt(g.gamma * g.R * gs.temp); }
    void update_thermo_from_pT(GasState& gs) const {
        gs.rho = gs.pressure / (g.R * gs.temp);
        gs.energy = g.Cv * gs.temp;
    }
    void update_thermo_from_rhoT(GasState& gs) const {
        gs.pressure = gs.rho * g.R * gs.temp;
        gs.energy = g.Cv * gs.temp;
    }
    void update_thermo_from_rhop(GasState& gs) const {
        gs.temp = gs.pressure / (gs.rho * g.R);
        gs.energy = g.Cv * gs.temp;
    }
    void update_thermo_from_rhou(GasState& gs) const {
        gs.temp = gs.energy / g.Cv;
        gs.pressure = gs.rho * g.R * gs.temp;
    }
};

// flux calculators (FV/flux_calc.h:27-76): the reference binds their names
struct FluxCalculator {
    std::string n;
    std::string name() const { return n; }
};
struct Hanel : FluxCalculator {
    Hanel() { n = "hanel"; }
};
struct Ausmdv : FluxCalculator {
    Ausmdv() { n = "ausmdv"; }
};
struct Ldfss2 : FluxCalculator {
    Ldfss2() { n = "ldfss2"; }
};
struct Rusanov : FluxCalculator {
    Rusanov() { n = "rusanov"; }
};

void check(ibis_status s, const char* what) {
    if (s != IBIS_OK) throw ibis::Error(std::string(what) + ": " + ibis_last_error());
}

double* dptr(uintptr_t p) { return reinterpret_cast<double*>(p); }
ibis_stream_t sptr(uintptr_t s) { return reinterpret_cast<ibis_stream_t>(s); }

// GridBlock(grid_file, grid config as JSON text)
struct PyGridBlock {
    ibis::GridBlock g;
    PyGridBlock(const std::string& file, const std::string& config, bool reorder)
        : g(file, ibis::Json::parse(config), reorder) {}
};

// FiniteVolume(grid, job config as JSON text) on device pointers
struct PyFiniteVolume {
    ibis::FiniteVolume fv;
    const ibis::GridBlock& grid;
    PyFiniteVolume(PyGridBlock& gb, const std::string& config)
        : fv(gb.g, ibis::Json::parse(config)), grid(gb.g) {}

    void compute_dudt(uintptr_t prim, uintptr_t dudt, bool allow_reconstruction,
                      uintptr_t stream) {
        check(ibis_compute_dudt(fv.handle(), dptr(prim), dptr(dudt), allow_reconstruction ? 1 : 0,
                                sptr(stream)),
              "ibis_compute_dudt");
    }
    double estimate_dt(uintptr_t prim, uintptr_t stream) {
        double dt = 0.0;
        check(ibis_estimate_dt(fv.handle(), dptr(prim), &dt, sptr(stream)), "ibis_estimate_dt");
        return dt;
    }
    uint64_t count_bad_cells(uintptr_t prim, uintptr_t stream) {
        uint64_t n = 0;
        check(ibis_count_bad_cells(fv.handle(), dptr(prim), &n, sptr(stream)),
              "ibis_count_bad_cells");
        return n;
    }
    void compute_dudt_vjp(uintptr_t prim, uintptr_t dudt_bar, uint64_t force_tags,
                          std::array<double, 3> force_dir, uintptr_t prim_bar, int limiter_mode,
                          uintptr_t stream) {
        check(ibis_compute_dudt_vjp(fv.handle(), dptr(prim), dptr(dudt_bar), force_tags,
                                    force_dir.data(), dptr(prim_bar), limiter_mode, sptr(stream)),
              "ibis_compute_dudt_vjp");
    }
    std::array<double, 3> wall_force(uintptr_t prim, uint64_t force_tags, uintptr_t stream) {
        std::array<double, 3> f{0.0, 0.0, 0.0};
        check(ibis_wall_force(fv.handle(), dptr(prim), force_tags, f.data(), sptr(stream)),
              "ibis_wall_force");
        return f;
    }
    uint64_t tag_mask(const std::vector<std::string>& names) const {
        const auto& tags = grid.boundary_tags();
        uint64_t m = 0;
        for (const std::string& n : names) {
            size_t k = 0;
            while (k < tags.size() && tags[k] != n) k++;
            if (k == tags.size()) throw std::invalid_argument("unknown boundary tag " + n);
            m |= uint64_t(1) << k;
        }
        return m;
    }
};

}  // namespace

PYBIND11_MODULE(python_api, m) {
    m.doc() = "ibis python module (MI355X residual)";
    py::register_exception<ibis::Error>(m, "IbisError", PyExc_RuntimeError);

    py::class_<GasState>(m, "GasState")
        .def(py::init<>())
        .def_readwrite("p", &GasState::pressure, "pressure (Pa)")
        .def_readwrite("T", &GasState::temp, "temperature (K)")
        .def_readwrite("rho", &GasState::rho, "density (kg/m^3)")
        .def_readwrite("energy", &GasState::energy, "energy (J/kg)");

    py::class_<PyIdealGas>(m, "PyIdealGas")
        .def(py::init<double>())
        .def("R", [](const PyIdealGas& g) { return g.g.R; })
        .def("Cv", [](const PyIdealGas& g) { return g.g.Cv; })
        .def("Cp", [](const PyIdealGas& g) { return g.g.Cp; })
        .def("gamma", [](const PyIdealGas& g) { return g.g.gamma; })
        .def("speed_of_sound", &PyIdealGas::speed_of_sound)
        .def("update_thermo_from_pT", &PyIdealGas::update_thermo_from_pT)
        .def("update_thermo_from_rhoT", &PyIdealGas::update_thermo_from_rhoT)
        .def("update_thermo_from_rhop", &PyIdealGas::update_thermo_from_rhop)
        .def("update_thermo_from_rhou", &PyIdealGas::update_thermo_from_rhou);

    py::class_<Vector3>(m, "Vector3")
        .def(py::init<>())
        .def(py::init([](double x) { return Vector3{x, 0.0, 0.0}; }))
        .def(py::init([](double x, double y) { return Vector3{x, y, 0.0}; }))
        .def(py::init([](double x, double y, double z) { return Vector3{x, y, z}; }))
        .def_readwrite("x", &Vector3::x, "x")
        .def_readwrite("y", &Vector3::y, "y")
        .def_readwrite("z", &Vector3::z, "z");

    // flux calculators (FV/flux_calc.h:40-76)
    py::class_<FluxCalculator>(m, "FluxCalculator").def("name", &FluxCalculator::name);
    py::class_<Hanel, FluxCalculator>(m, "PyHanel").def(py::init<>());
    py::class_<Ausmdv, FluxCalculator>(m, "PyAusmdv").def(py::init<>());
    py::class_<Ldfss2, FluxCalculator>(m, "PyLdfss2").def(py::init<>());
    py::class_<Rusanov, FluxCalculator>(m, "PyRusanov").def(py::init<>());

    py::class_<PyGridBlock>(m, "GridBlock")
        .def(py::init<const std::string&, const std::string&, bool>(), py::arg("grid_file"),
             py::arg("config"), py::arg("reorder") = true)
        .def("dim", [](const PyGridBlock& b) { return b.g.dim(); })
        .def("num_cells", [](const PyGridBlock& b) { return b.g.num_cells(); })
        .def("num_ghost_cells", [](const PyGridBlock& b) { return b.g.num_ghost_cells(); })
        .def("num_total_cells", [](const PyGridBlock& b) { return b.g.num_total_cells(); })
        .def("num_interfaces", [](const PyGridBlock& b) { return b.g.num_interfaces(); })
        .def("boundary_tags", [](const PyGridBlock& b) { return b.g.boundary_tags(); })
        .def("cell_new_of_old", [](const PyGridBlock& b) { return b.g.cell_new_of_old(); })
        .def("to_grid_order",
             [](const PyGridBlock& b, const std::vector<double>& a, size_t rows) {
                 return b.g.to_grid_order(a, rows);
             })
        .def("to_reference_order", [](const PyGridBlock& b, const std::vector<double>& a,
                                      size_t rows) { return b.g.to_reference_order(a, rows); });

    py::class_<PyFiniteVolume>(m, "FiniteVolume")
        .def(py::init<PyGridBlock&, const std::string&>(), py::keep_alive<1, 2>(),
             py::arg("grid"), py::arg("config"))
        .def("compute_dudt", &PyFiniteVolume::compute_dudt, py::arg("prim"), py::arg("dudt"),
             py::arg("allow_reconstruction") = true, py::arg("stream") = 0)
        .def("estimate_dt", &PyFiniteVolume::estimate_dt, py::arg("prim"), py::arg("stream") = 0)
        .def("count_bad_cells", &PyFiniteVolume::count_bad_cells, py::arg("prim"),
             py::arg("stream") = 0)
        .def("compute_dudt_vjp", &PyFiniteVolume::compute_dudt_vjp, py::arg("prim"),
             py::arg("dudt_bar"), py::arg("force_tags"), py::arg("force_dir"),
             py::arg("prim_bar"), py::arg("limiter_mode"), py::arg("stream") = 0)
        .def("wall_force", &PyFiniteVolume::wall_force, py::arg("prim"), py::arg("force_tags"),
             py::arg("stream") = 0)
        .def("tag_mask", &PyFiniteVolume::tag_mask);
}
