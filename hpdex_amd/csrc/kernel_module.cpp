// kernel_module.cpp -- Python extension `hpdex_amd.backend.bin.kernel`, the
// counterpart of the reference's pybind11 module (src/cpp/module.cpp:193-243):
// the same two functions, keyword names and defaults, returning read-only
// float64 arrays (:102-106, :138-139).  It only converts arguments, releases
// the GIL and calls the C-ABI in include/hpdex_mi355x.h; a nonzero return
// becomes RuntimeError(message), as pybind11 translates std::runtime_error.
//
// layout="csr": the reference swaps rows and columns here (:52-53), which
// indexes indptr out of bounds unless the matrix is square.  The wrapper's
// stated meaning is a CSR matrix whose columns are the samples
// (src/hpdex/backend/mannwhitneyu_.py:47-51); that is the CSC layout of its
// transpose, so it is passed through unchanged.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "hpdex_mi355x.h"

namespace py = pybind11;

namespace {

int dtype_code(const py::array& a) {
    const py::dtype d = a.dtype();
    if (d.is(py::dtype::of<int8_t>())) return HPDEX_INT8;
    if (d.is(py::dtype::of<uint8_t>())) return HPDEX_UINT8;
    if (d.is(py::dtype::of<int16_t>())) return HPDEX_INT16;
    if (d.is(py::dtype::of<uint16_t>())) return HPDEX_UINT16;
    if (d.is(py::dtype::of<int32_t>())) return HPDEX_INT32;
    if (d.is(py::dtype::of<uint32_t>())) return HPDEX_UINT32;
    if (d.is(py::dtype::of<int64_t>())) return HPDEX_INT64;
    if (d.is(py::dtype::of<uint64_t>())) return HPDEX_UINT64;
    if (d.is(py::dtype::of<float>())) return HPDEX_FLOAT32;
    if (d.is(py::dtype::of<double>())) return HPDEX_FLOAT64;
    return -1;
}

using I64 = py::array_t<int64_t, py::array::c_style | py::array::forcecast>;
using I32 = py::array_t<int32_t, py::array::c_style | py::array::forcecast>;

py::array_t<double> readonly(py::array_t<double> a) {
    a.attr("flags").attr("writeable") = py::bool_(false);
    return a;
}

struct Csc {
    py::array data;
    I64 indices, indptr;
    I32 gid;
    size_t R = 0, C = 0, nnz = 0;
    int dtype = -1;
};

Csc prepare(py::array data, py::array indices, py::array indptr, py::array group_id,
            const std::string& layout) {
    Csc m;
    m.dtype = dtype_code(data);
    if (m.dtype < 0) throw std::runtime_error("Unsupported dtype for data");
    if (layout != "csc" && layout != "csr")
        throw std::runtime_error("layout must be 'csc' or 'csr'");
    m.data = py::array::ensure(data, py::array::c_style);
    m.indices = I64::ensure(indices);
    m.indptr = I64::ensure(indptr);
    m.gid = I32::ensure(group_id);
    if (!m.data || !m.indices || !m.indptr || !m.gid)
        throw std::runtime_error("could not convert the CSC arrays");
    if (m.indptr.size() < 1) throw std::runtime_error("indptr must hold at least one entry");
    m.R = (size_t)m.gid.size();
    m.C = (size_t)m.indptr.size() - 1;
    m.nnz = (size_t)m.data.size();
    if ((size_t)m.indices.size() < m.nnz)
        throw std::runtime_error("indices holds fewer entries than data");
    return m;
}

}  // namespace

PYBIND11_MODULE(kernel, mod) {
    mod.doc() = "hpdex MI355X backend: Mann-Whitney U and group means on gfx950";

    mod.def(
        "mannwhitneyu",
        [](py::array data, py::array indices, py::array indptr, py::array group_id,
           size_t n_targets, bool ref_sorted, bool tar_sorted, bool tie_correction,
           bool use_continuity, bool fast_norm, int zero_handling, int alternative, int method,
           int threads, std::string layout, bool show_progress) {
            (void)show_progress;
            Csc m = prepare(data, indices, indptr, group_id, layout);
            hpdex_mwu_opts o{};
            o.ref_sorted = ref_sorted;
            o.tar_sorted = tar_sorted;
            o.tie_correction = tie_correction;
            o.use_continuity = use_continuity;
            o.fast_norm = fast_norm;
            o.zero_handling = zero_handling;
            o.alternative = alternative;
            o.method = method;
            const size_t np = n_targets >= 1 ? m.C * n_targets : 0;
            py::array_t<double> U1(np), P(np);
            double* u = U1.mutable_data();
            double* p = P.mutable_data();
            char err[1024] = {0};
            int rc;
            {
                py::gil_scoped_release nogil;
                rc = hpdex_mwu_csc(m.dtype, m.data.data(), m.indices.data(), m.indptr.data(), m.R,
                                   m.C, m.nnz, m.gid.data(), n_targets, &o, threads, u, p, err,
                                   sizeof err);
            }
            if (rc != 0) throw std::runtime_error(err);
            return py::make_tuple(readonly(U1), readonly(P));
        },
        py::arg("data"), py::arg("indices"), py::arg("indptr"), py::arg("group_id"),
        py::arg("n_targets"), py::arg("ref_sorted") = false, py::arg("tar_sorted") = false,
        py::arg("tie_correction") = true, py::arg("use_continuity") = false,
        py::arg("fast_norm") = true, py::arg("zero_handling") = 0, py::arg("alternative") = 2,
        py::arg("method") = 2, py::arg("threads") = -1, py::arg("layout") = "csc",
        py::arg("show_progress") = false);

    mod.def(
        "group_mean",
        [](py::array data, py::array indices, py::array indptr, py::array group_id,
           size_t n_groups, bool include_zeros, int threads, bool use_kahan, std::string layout) {
            Csc m = prepare(data, indices, indptr, group_id, layout);
            py::array_t<double> out(m.C * n_groups);
            double* o = out.mutable_data();
            char err[1024] = {0};
            int rc;
            {
                py::gil_scoped_release nogil;
                rc = hpdex_group_mean_csc(m.dtype, m.data.data(), m.indices.data(),
                                          m.indptr.data(), m.R, m.C, m.nnz, m.gid.data(),
                                          n_groups, include_zeros ? 1 : 0, use_kahan ? 1 : 0,
                                          threads, o, err, sizeof err);
            }
            if (rc != 0) throw std::runtime_error(err);
            return readonly(out);
        },
        py::arg("data"), py::arg("indices"), py::arg("indptr"), py::arg("group_id"),
        py::arg("n_groups"), py::arg("include_zeros") = true, py::arg("threads") = -1,
        py::arg("use_kahan") = false, py::arg("layout") = "csc");
}
