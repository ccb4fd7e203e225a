// abdtk_pybind.cpp - the reference's `abdtk` Python module (python/bindings.cpp:15-83)
// as a pybind11 extension over the C ABI of libadb.so: same module name,
// function names, argument names and enum values.  The bodies call the gfx950
// kernels through include/adb_capi.h instead of ipc-toolkit; `ij` is accepted
// and unused, as in the reference.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <stdexcept>
#include <string>

#include "../../include/adb_capi.h"

namespace py = pybind11;
using arr = py::array_t<double, py::array::c_style | py::array::forcecast>;

namespace {

// 4 points (a list of 4 vec3, a 4x3 array or 12 numbers) -> 12 doubles
void points12(const py::object& pts, double x[12]) {
    arr a = arr::ensure(py::module_::import("numpy").attr("asarray")(pts, "float64"));
    if (!a || a.size() != 12) throw py::value_error("expected 4 points (12 coordinates)");
    for (int i = 0; i < 12; i++) x[i] = a.data()[i];
}

void check(int rc, const char* what) {
    if (rc != ADB_OK) throw std::runtime_error(std::string(what) + " failed with status " + std::to_string(rc));
}

// H comes back column-major from the C ABI (Eigen's storage of mat12);
// numpy gets the mathematical matrix, row-major
arr to_numpy12x12(const double* Hc) {
    arr H({12, 12});
    double* h = H.mutable_data();
    for (int r = 0; r < 12; r++)
        for (int c = 0; c < 12; c++) h[12 * r + c] = Hc[12 * c + r];
    return H;
}

py::tuple hess_pt(const py::object& pt, const py::object& /*ij*/, int type, double dist) {
    double x[12], H[144];
    arr g(12);
    points12(pt, x);
    check(adb_ipc_hess_pt_12x12(nullptr, x, type, dist, H, g.mutable_data()), "ipc_hess_pt_12x12");
    return py::make_tuple(to_numpy12x12(H), g);
}

py::tuple hess_ee(const py::object& ee, const py::object& /*ij*/, int type, double dist) {
    double x[12], H[144], m = 1.0;
    arr g(12);
    points12(ee, x);
    check(adb_ipc_hess_ee_12x12(nullptr, x, type, dist, H, g.mutable_data(), &m), "ipc_hess_ee_12x12");
    return py::make_tuple(to_numpy12x12(H), g, m);
}

double collision_time(bool ee, const arr* v[8]) {
    double x[24];
    for (int i = 0; i < 8; i++) {
        if (v[i]->size() != 3) throw py::value_error("expected 3-vectors");
        for (int k = 0; k < 3; k++) x[3 * i + k] = v[i]->data()[k];
    }
    const double t = ee ? adb_ee_collision_time(x) : adb_pt_collision_time(x);
    if (t < 0.0) throw std::runtime_error("collision time: device call failed");
    return t;
}

}  // namespace

PYBIND11_MODULE(abdtk, m) {
    m.doc() = "ADB contact kernels (IPC barrier Hessians, collision times) on gfx950 through libadb.so";
    py::enum_<adb_pt_type>(m, "PointTriangleDistanceType")
        .value("P_T0", ADB_P_T0).value("P_T1", ADB_P_T1).value("P_T2", ADB_P_T2)
        .value("P_E0", ADB_P_E0).value("P_E1", ADB_P_E1).value("P_E2", ADB_P_E2)
        .value("P_T", ADB_P_T).value("AUTO", ADB_PT_AUTO)
        .export_values();
    py::enum_<adb_ee_type>(m, "EdgeEdgeDistanceType")
        .value("EA0_EB0", ADB_EA0_EB0).value("EA0_EB1", ADB_EA0_EB1).value("EA1_EB0", ADB_EA1_EB0)
        .value("EA1_EB1", ADB_EA1_EB1).value("EA_EB0", ADB_EA_EB0).value("EA_EB1", ADB_EA_EB1)
        .value("EA0_EB", ADB_EA0_EB).value("EA1_EB", ADB_EA1_EB).value("EA_EB", ADB_EA_EB)
        .value("AUTO", ADB_EE_AUTO)
        .export_values();
    m.def("ipc_hess_pt_12x12", [](const py::object& pt, const py::object& ij, adb_pt_type t, double d) {
              return hess_pt(pt, ij, static_cast<int>(t), d); },
          "Computes the hessian and gradient of the point-triangle distance",
          py::arg("pt"), py::arg("ij"), py::arg("pt_type"), py::arg("dist"));
    m.def("ipc_hess_ee_12x12", [](const py::object& ee, const py::object& ij, adb_ee_type t, double d) {
              return hess_ee(ee, ij, static_cast<int>(t), d); },
          "Computes the hessian and gradient of the edge-edge distance",
          py::arg("ee"), py::arg("ij"), py::arg("ee_type"), py::arg("dist"));
    m.def("ee_collision_time",
          [](arr a0, arr a1, arr a2, arr a3, arr b0, arr b1, arr b2, arr b3) {
              const arr* v[8] = {&a0, &a1, &a2, &a3, &b0, &b1, &b2, &b3};
              return collision_time(true, v); },
          "Computes the collision time between two edge-edge pairs",
          py::arg("ei0_t0"), py::arg("ei1_t0"), py::arg("ej0_t0"), py::arg("ej1_t0"),
          py::arg("ei0_t1"), py::arg("ei1_t1"), py::arg("ej0_t1"), py::arg("ej1_t1"));
    m.def("pt_collision_time",
          [](arr a0, arr a1, arr a2, arr a3, arr b0, arr b1, arr b2, arr b3) {
              const arr* v[8] = {&a0, &a1, &a2, &a3, &b0, &b1, &b2, &b3};
              return collision_time(false, v); },
          "Computes the collision time between a point and a triangle",
          py::arg("p0_t0"), py::arg("t0_t0"), py::arg("t1_t0"), py::arg("t2_t0"),
          py::arg("p0_t1"), py::arg("t0_t1"), py::arg("t1_t1"), py::arg("t2_t1"));
}
