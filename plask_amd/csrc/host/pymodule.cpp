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

// Array arguments: one contiguous copy (numpy arrays, or any sequence numpy converts)
// instead of pybind11's element-by-element sequence conversion, which costs ~20 ns per
// value (~70 ms for two 1.7M-element conductivity arrays).
template <typename T> using arr = py::array_t<T, py::array::c_style | py::array::forcecast>;
template <typename T> static std::vector<T> vec(const arr<T>& a) {
    return std::vector<T>(a.data(), a.data() + a.size());
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
        .def("set_dirichlet",
             [](FemCore& s, const arr<int64_t>& nodes, const arr<double>& values) {
                 s.set_dirichlet(vec(nodes), vec(values));
             })
        .def("solution", [](FemCore& s) { return to_np(s.solution()); })
        .def("interp",
             [](FemCore& s, const arr<double>& pts, double fill) {
                 if (pts.ndim() != 2 || pts.shape(1) != s.dim())
                     throw std::invalid_argument("points must be an (n, dim) array");
                 return to_np(s.interp(pts.data(), (int64_t)pts.shape(0), fill));
             })
        .def("postprocess",
             [](FemCore& s, int what, py::object kh, py::object kv, bool fetch) {
                 std::vector<double> a, b;
                 if (!kh.is_none()) {
                     a = vec(kh.cast<arr<double>>());
                     b = vec(kv.cast<arr<double>>());
                 }
                 return to_np(s.postprocess(what, a, b, fetch));
             },
             py::arg("what"), py::arg("kh") = py::none(), py::arg("kv") = py::none(), py::arg("fetch") = false)
        .def("interp_elements",
             [](FemCore& s, int what, const arr<double>& pts, double fill) {
                 if (pts.ndim() != 2 || pts.shape(1) != s.dim())
                     throw std::invalid_argument("points must be an (n, dim) array");
                 return to_np(s.interp_elements(what, pts.data(), (int64_t)pts.shape(0), fill));
             })
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
        .def("set_conductivity",
             [](ThermalCore& s, const arr<double>& k0h, const arr<double>& k0v, const arr<double>& ah,
                const arr<double>& av) { s.set_conductivity(vec(k0h), vec(k0v), vec(ah), vec(av)); })
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
        .def("set_heat", [](ThermalCore& s, const arr<double>& q) { s.set_heat(vec(q)); })
        .def("set_convection",
             [](ThermalCore& s, const arr<int64_t>& nodes, const arr<double>& coeff, const arr<double>& ambient) {
                 s.set_convection(vec(nodes), vec(coeff), vec(ambient));
             })
        .def("set_boundary_loads",
             [](ThermalCore& s, const arr<int64_t>& flux_nodes, const arr<double>& flux, const arr<int64_t>& rad_nodes,
                const arr<double>& emissivity, const arr<double>& ambient) {
                 s.set_boundary_loads(vec(flux_nodes), vec(flux), vec(rad_nodes), vec(emissivity), vec(ambient));
             })
        .def("initialize", &ThermalCore::initialize)
        .def("compute", &ThermalCore::compute, py::arg("loops") = 0)
        // the linked core must outlive the link (the caller unlinks with None)
        .def("link_heat", [](ThermalCore& s, FemCore* src) { s.link_heat(src); }, py::arg("src").none(true))
        .def_readwrite("inittemp", &ThermalCore::inittemp)
        .def_readwrite("maxerr", &ThermalCore::maxerr)
        .def_readwrite("loopno", &ThermalCore::loopno)
        .def_readonly("toterr", &ThermalCore::toterr)
        .def_readonly("maxT", &ThermalCore::maxT);

    py::class_<DynamicCore, ThermalCore>(m, "DynamicCore")
        .def(py::init<int, bool, int, std::string>(), py::arg("dim"), py::arg("cylindrical") = false,
             py::arg("device") = 0, py::arg("id") = "dynamic")
        .def("set_capacity_callback",
             [](DynamicCore& s, py::function f) {
                 s.set_capacity_callback([f](const double* t, double* c, int64_t n) {
                     py::gil_scoped_acquire g;
                     py::array_t<double> ta(n, t);
                     auto a = f(ta).cast<py::array_t<double, py::array::c_style | py::array::forcecast>>();
                     if (a.size() != n) throw std::runtime_error("capacity callback size mismatch");
                     std::copy(a.data(), a.data() + n, c);
                 });
             })
        .def("initialize", &DynamicCore::initialize)
        .def("compute", &DynamicCore::compute, py::arg("time"))
        .def_readwrite("methodparam", &DynamicCore::methodparam)
        .def_readwrite("timestep", &DynamicCore::timestep)
        .def_readwrite("rebuildfreq", &DynamicCore::rebuildfreq)
        .def_readwrite("logfreq", &DynamicCore::logfreq)
        .def_readwrite("elapstime", &DynamicCore::elapstime)
        .def_readwrite("held", &DynamicCore::held);

    py::class_<ElectricalCore, FemCore>(m, "ElectricalCore")
        .def(py::init<int, bool, int, std::string>(), py::arg("dim"), py::arg("cylindrical") = false,
             py::arg("device") = 0, py::arg("id") = "electrical")
        .def("set_conductivity",
             [](ElectricalCore& s, const arr<double>& sh, const arr<double>& sv) { s.set_conductivity(vec(sh), vec(sv)); })
        .def("set_junctions",
             [](ElectricalCore& s, const arr<int64_t>& elems, int ncorner, const arr<int64_t>& bottom,
                const arr<int64_t>& top, const arr<int>& junction, const arr<double>& heights) {
                 s.set_junctions(vec(elems), ncorner, vec(bottom), vec(top), vec(junction), vec(heights));
             })
        .def("set_junction_params",
             [](ElectricalCore& s, const arr<double>& beta, const arr<double>& js) {
                 s.set_junction_params(vec(beta), vec(js));
             })
        .def("set_flags",
             [](ElectricalCore& s, const arr<unsigned char>& select, const arr<unsigned char>& noheat) {
                 s.set_flags(vec(select), vec(noheat));
             })
        .def("set_start_cond", &ElectricalCore::set_start_cond, py::arg("cond_v"), py::arg("cond_h") = 0.)
        .def("set_junction_law",
             [](ElectricalCore& s, py::object f) {
                 if (f.is_none()) {
                     s.set_junction_law(nullptr);
                     return;
                 }
                 s.set_junction_law([f](const double* U, const double* jy, double* ch, double* cv, int64_t n) {
                     py::gil_scoped_acquire g;
                     py::array_t<double> ua(n, U), ja(n, jy);
                     py::tuple r = f(ua, ja).cast<py::tuple>();
                     if (r.size() != 2) throw std::runtime_error("junction law must return (cond_h, cond_v)");
                     auto a = r[0].cast<py::array_t<double, py::array::c_style | py::array::forcecast>>();
                     auto b = r[1].cast<py::array_t<double, py::array::c_style | py::array::forcecast>>();
                     if (a.size() != n || b.size() != n) throw std::runtime_error("junction law size mismatch");
                     std::copy(a.data(), a.data() + n, ch);
                     std::copy(b.data(), b.data() + n, cv);
                 });
             })
        .def("initialize", &ElectricalCore::initialize)
        .def("compute", &ElectricalCore::compute, py::arg("loops") = 0)
        .def("link_temperature",
             [](ElectricalCore& s, FemCore* src, const arr<double>& s0h, const arr<double>& s0v, const arr<double>& ah,
                const arr<double>& av) { s.link_temperature(src, vec(s0h), vec(s0v), vec(ah), vec(av)); },
             py::arg("src").none(true), py::arg("s0h"), py::arg("s0v"), py::arg("ah"), py::arg("av"))
        .def_property_readonly("device_coeffs", &ElectricalCore::device_coeffs)
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
