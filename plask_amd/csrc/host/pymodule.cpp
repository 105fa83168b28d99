// pybind11 module plask_amd._core: the C++ solver cores for the Python solver classes
// (the role Boost.Python plays for therm_python.cpp / electr_python.cpp in the reference).
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "fem_solvers.hpp"

namespace py = pybind11;
using namespace plask_amd;

template <typename T> static py::array_t<T> to_np(const std::vector<T>& v) {
    py::array_t<T> a(v.size());
    std::copy(v.begin(), v.end(), a.mutable_data());
    return a;
}

PYBIND11_MODULE(_core, m) {
    m.doc() = "C++ solver cores over the gfx950 FEM C ABI";
    py::register_exception<ComputationError>(m, "ComputationError", PyExc_ArithmeticError);

    py::class_<IterParams>(m, "IterParams")
        .def(py::init<>())
        .def_readwrite("maxit", &IterParams::maxit)
        .def_readwrite("maxerr", &IterParams::maxerr)
        .def_readwrite("precond", &IterParams::precond)
        .def_readwrite("precision", &IterParams::precision)
        .def_readwrite("inner_rtol", &IterParams::inner_rtol)
        .def_readwrite("noconv", &IterParams::noconv)
        .def_readonly("converged", &IterParams::converged)
        .def_readonly("iters", &IterParams::iters)
        .def_readonly("err", &IterParams::err)
        .def_readonly("seconds", &IterParams::seconds);

    py::class_<FemCore>(m, "FemCore")
        .def("set_mesh", &FemCore::set_mesh)
        .def("set_dirichlet", &FemCore::set_dirichlet)
        .def("solution", [](FemCore& s) { return to_np(s.solution()); })
        .def("invalidate", &FemCore::invalidate)
        .def_property_readonly("initialized", &FemCore::initialized)
        .def_property_readonly("num_nodes", &FemCore::num_nodes)
        .def_property_readonly("num_elements", &FemCore::num_elements)
        .def_readwrite("iter", &FemCore::iter)
        .def_readwrite("logger", &FemCore::logger)
        .def_readwrite("id", &FemCore::id);

    py::class_<ThermalCore, FemCore>(m, "ThermalCore")
        .def(py::init<int, bool, int, std::string>(), py::arg("dim"), py::arg("cylindrical") = false,
             py::arg("device") = 0, py::arg("id") = "thermal")
        .def("set_conductivity", &ThermalCore::set_conductivity)
        .def("set_conductivity_callback",
             [](ThermalCore& s, py::function f) {
                 s.set_conductivity_callback([f](const double* t, double* kh, double* kv, int64_t n) {
                     py::gil_scoped_acquire g;
                     py::array_t<double> ta(n, t);
                     py::tuple r = f(ta);
                     auto a = r[0].cast<py::array_t<double, py::array::c_style | py::array::forcecast>>();
                     auto b = r[1].cast<py::array_t<double, py::array::c_style | py::array::forcecast>>();
                     if (a.size() != n || b.size() != n) throw std::runtime_error("conductivity callback size mismatch");
                     std::copy(a.data(), a.data() + n, kh);
                     std::copy(b.data(), b.data() + n, kv);
                 });
             })
        .def("set_heat", &ThermalCore::set_heat)
        .def("initialize", &ThermalCore::initialize)
        .def("compute", &ThermalCore::compute, py::arg("loops") = 0)
        .def_readwrite("inittemp", &ThermalCore::inittemp)
        .def_readwrite("maxerr", &ThermalCore::maxerr)
        .def_readwrite("loopno", &ThermalCore::loopno)
        .def_readonly("toterr", &ThermalCore::toterr)
        .def_readonly("maxT", &ThermalCore::maxT);

    py::class_<ElectricalCore, FemCore>(m, "ElectricalCore")
        .def(py::init<int, bool, int, std::string>(), py::arg("dim"), py::arg("cylindrical") = false,
             py::arg("device") = 0, py::arg("id") = "electrical")
        .def("set_conductivity", &ElectricalCore::set_conductivity)
        .def("set_junctions", &ElectricalCore::set_junctions)
        .def("set_junction_params", &ElectricalCore::set_junction_params)
        .def("set_flags", &ElectricalCore::set_flags)
        .def("set_start_cond", &ElectricalCore::set_start_cond)
        .def("initialize", &ElectricalCore::initialize)
        .def("compute", &ElectricalCore::compute, py::arg("loops") = 0)
        .def("currents", [](ElectricalCore& s) { return to_np(s.currents()); })
        .def("heats", [](ElectricalCore& s) { return to_np(s.heats()); })
        .def("conductivity_h", [](ElectricalCore& s) { return to_np(s.conductivity_h()); })
        .def("conductivity_v", [](ElectricalCore& s) { return to_np(s.conductivity_v()); })
        .def_readwrite("maxerr", &ElectricalCore::maxerr)
        .def_readwrite("stable", &ElectricalCore::stable)
        .def_readwrite("loopno", &ElectricalCore::loopno)
        .def_readonly("toterr", &ElectricalCore::toterr)
        .def_readonly("maxcur", &ElectricalCore::maxcur);
}
