// pybind11 binding of the gko:: host library (module ginkgo_amd._gko).
// Python reaches the solvers through the same C++ classes a C++ user
// would: Cg<V>::build().with_criteria(...).on(exec)->generate(A)->apply(b, x).
// Device memory handed in from Python (torch tensors) is wrapped in
// non-owning views; the caller keeps it alive.
#include <sstream>

#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "ginkgo_amd/ginkgo.hpp"

namespace py = pybind11;
using namespace gko;

namespace {

using crit_param = deferred_factory_parameter<const stop::CriterionFactory>;
using op_param = deferred_factory_parameter<const LinOpFactory>;

crit_param to_crit(const py::handle& o)
{
    if (py::isinstance<stop::Iteration::parameters_type>(o))
        return crit_param(o.cast<stop::Iteration::parameters_type>());
    if (py::isinstance<stop::ResidualNorm<double>::parameters_type>(o))
        return crit_param(o.cast<stop::ResidualNorm<double>::parameters_type>());
    if (py::isinstance<stop::ResidualNorm<float>::parameters_type>(o))
        return crit_param(o.cast<stop::ResidualNorm<float>::parameters_type>());
    if (py::isinstance<stop::ImplicitResidualNorm<double>::parameters_type>(o))
        return crit_param(o.cast<stop::ImplicitResidualNorm<double>::parameters_type>());
    if (py::isinstance<stop::ImplicitResidualNorm<float>::parameters_type>(o))
        return crit_param(o.cast<stop::ImplicitResidualNorm<float>::parameters_type>());
    if (py::isinstance<stop::CriterionFactory>(o))
        return crit_param(o.cast<std::shared_ptr<const stop::CriterionFactory>>());
    throw py::type_error("not a stopping criterion");
}

op_param to_op(const py::handle& o)
{
    if (o.is_none()) return op_param();
    if (py::isinstance<preconditioner::Jacobi<double>::parameters_type>(o))
        return op_param(o.cast<preconditioner::Jacobi<double>::parameters_type>());
    if (py::isinstance<preconditioner::Jacobi<float>::parameters_type>(o))
        return op_param(o.cast<preconditioner::Jacobi<float>::parameters_type>());
    if (py::isinstance<solver::Cg<double>::parameters_type>(o))
        return op_param(o.cast<solver::Cg<double>::parameters_type>());
    if (py::isinstance<solver::Cg<float>::parameters_type>(o))
        return op_param(o.cast<solver::Cg<float>::parameters_type>());
    if (py::isinstance<solver::Bicgstab<double>::parameters_type>(o))
        return op_param(o.cast<solver::Bicgstab<double>::parameters_type>());
    if (py::isinstance<solver::Bicgstab<float>::parameters_type>(o))
        return op_param(o.cast<solver::Bicgstab<float>::parameters_type>());
    if (py::isinstance<LinOpFactory>(o)) return op_param(o.cast<std::shared_ptr<const LinOpFactory>>());
    throw py::type_error("not a LinOp factory or its parameters");
}

template <typename V>
std::shared_ptr<typename matrix::Csr<V>::strategy_type> strategy(const std::string& s, int sw)
{
    using C = matrix::Csr<V>;
    if (s == "classical") return std::make_shared<typename C::classical>(sw);
    if (s == "load_balance") return std::make_shared<typename C::load_balance>();
    // load_balance:<nwarps>[:<warp_size>[:cuda|hip]] -- the reference's
    // load_balance(nwarps, warp_size = 32, cuda_strategy = true) constructor
    if (s.rfind("load_balance:", 0) == 0) {
        const auto a = s.find(':');
        const auto b = s.find(':', a + 1);
        const int64 nw = std::stoll(s.substr(a + 1, b - a - 1));
        const auto c = b == std::string::npos ? b : s.find(':', b + 1);
        const int ws = b == std::string::npos ? 32 : std::stoi(s.substr(b + 1, c - b - 1));
        const bool cuda = c == std::string::npos || s.substr(c + 1) != "hip";
        return std::make_shared<typename C::load_balance>(nw, ws, cuda);
    }
    if (s == "automatical") return std::make_shared<typename C::automatical>();
    if (s == "merge_path") return std::make_shared<typename C::merge_path>();
    if (s == "sparselib") return std::make_shared<typename C::sparselib>();
    throw py::value_error(
        "strategy must be classical, load_balance, merge_path, sparselib or automatical");
}

template <typename T>
std::vector<T> vec(py::array_t<T, py::array::c_style | py::array::forcecast> a)
{
    return std::vector<T>(a.data(), a.data() + a.size());
}

template <typename T>
py::array_t<T> arr(const std::vector<T>& v)
{
    py::array_t<T> out(v.size());
    std::copy(v.begin(), v.end(), out.mutable_data());
    return out;
}

// One batch solver class with its factory and parameters (batch::solver::
// Bicgstab and Cg share the surface).
template <typename BS>
void bind_batch_solver(py::module_& m, const std::string& name, const std::string& sfx)
{
    using V = typename BS::value_type;
    using MV = batch::MultiVector<V>;
    using BJ = batch::preconditioner::Jacobi<V>;
    using BL = batch::log::BatchConvergence<V>;
    py::class_<BS, std::shared_ptr<BS>>(m, (name + sfx).c_str())
        .def("apply", [](const BS& s, const MV* b, MV* x) { s.apply(b, x); })
        .def("tune", [](BS& s, const MV* b, const MV* x) { s.tune(b, x); })
        .def(
            "capture",
            [](const BS& s, const MV* b, MV* x, bool zero) {
                return std::shared_ptr<Graph>(s.capture(b, x, zero));
            },
            py::arg("b"), py::arg("x"), py::arg("zero_initial_guess") = true,
            // the graph keeps the solver, b and x alive
            py::keep_alive<0, 1>(), py::keep_alive<0, 2>(), py::keep_alive<0, 3>())
        .def("record_logs", &BS::record_logs)
        .def("add_logger", [](BS& s, std::shared_ptr<const BL> l) { s.add_logger(l); })
        .def_property_readonly("block_sizes", &BS::get_block_sizes)
        .def_property_readonly("matrix_in_lds", &BS::get_matrix_in_lds);
    py::class_<typename BS::Factory, std::shared_ptr<typename BS::Factory>>(
        m, (name + "Factory" + sfx).c_str())
        .def("generate", [](const typename BS::Factory& f, std::shared_ptr<const batch::BatchLinOp> a) {
            return std::shared_ptr<BS>(f.generate(std::move(a)));
        });
    using BP = typename BS::parameters_type;
    py::class_<BP>(m, (name + "Params" + sfx).c_str())
        .def(py::init<>())
        .def("with_max_iterations", [](BP p, int v) { return p.with_max_iterations(v); })
        .def("with_tolerance", [](BP p, double v) { return p.with_tolerance(v); })
        .def("with_tolerance_type",
             [](BP p, batch::stop::tolerance_type t) { return p.with_tolerance_type(t); })
        .def("with_preconditioner",
             [](BP p, const typename BJ::parameters_type& j) { return p.with_preconditioner(j); })
        .def("on", [](const BP& p, std::shared_ptr<const Executor> e) { return p.on(e); });
}

// The batch module (batch::MultiVector / matrix::Csr / matrix::BlockDiagCsr /
// solver::Bicgstab / solver::Cg / log::BatchConvergence).
template <typename V>
void bind_batch(py::module_& m, const std::string& sfx)
{
    using MV = batch::MultiVector<V>;
    using BM = batch::matrix::BatchMatrix<V>;
    using BCsr = batch::matrix::Csr<V, int32>;
    using BEll = batch::matrix::Ell<V, int32>;
    using BDense = batch::matrix::Dense<V>;
    using BDiag = batch::matrix::BlockDiagCsr<V, int32>;
    using BJ = batch::preconditioner::Jacobi<V>;
    using BL = batch::log::BatchConvergence<V>;
    py::class_<MV, batch::BatchLinOp, std::shared_ptr<MV>>(m, ("BatchMultiVector" + sfx).c_str())
        .def_static("create",
                    [](std::shared_ptr<const Executor> e, size_type num, size_type n) {
                        return std::shared_ptr<MV>(
                            MV::create(e, batch_dim<2>{num, dim<2>{n, 1}}));
                    })
        .def_static("from_numpy",
                    [](std::shared_ptr<const Executor> e, size_type num, size_type n,
                       py::array_t<V, py::array::c_style | py::array::forcecast> a) {
                        return std::shared_ptr<MV>(MV::create_from_host(e, num, n, vec<V>(a)));
                    })
        .def_static("uneven",
                    [](std::shared_ptr<const Executor> e,
                       py::array_t<int64, py::array::c_style | py::array::forcecast> off,
                       py::object values) {
                        const auto o = vec<int64>(off);
                        if (values.is_none()) {
                            return std::shared_ptr<MV>(MV::create_uneven(e, o));
                        }
                        const auto v = vec<V>(values.cast<py::array_t<
                            V, py::array::c_style | py::array::forcecast>>());
                        return std::shared_ptr<MV>(MV::create_uneven(e, o, &v));
                    },
                    py::arg("exec"), py::arg("row_offsets"), py::arg("values") = py::none())
        .def("to_numpy", [](const MV& x) { return arr(x.to_host()); })
        .def("fill", &MV::fill)
        .def_property_readonly("total_rows", &MV::get_total_rows);
    py::class_<BM, batch::BatchLinOp, std::shared_ptr<BM>>(m, ("BatchMatrix" + sfx).c_str())
        .def("apply", [](const BM& a, const MV* b, MV* x) { a.apply(b, x); })
        .def("apply", [](const BM& a, const MV* alpha, const MV* b, const MV* beta,
                         MV* x) { a.apply(alpha, b, beta, x); })
        .def_property_readonly("size", [](const BM& a) {
            return py::make_tuple(a.get_num_batch_items(), a.get_common_size().rows,
                                  a.get_common_size().cols);
        });
    py::class_<BCsr, BM, std::shared_ptr<BCsr>>(m, ("BatchCsr" + sfx).c_str())
        .def_static("from_numpy",
                    [](std::shared_ptr<const Executor> e, size_type num, size_type n,
                       py::array_t<int32, py::array::c_style | py::array::forcecast> rp,
                       py::array_t<int32, py::array::c_style | py::array::forcecast> ci,
                       py::array_t<V, py::array::c_style | py::array::forcecast> v,
                       py::object cols) {
                        const size_type nc = cols.is_none() ? n : cols.cast<size_type>();
                        return std::shared_ptr<BCsr>(BCsr::create_from_host(
                            e, num, dim<2>{n, nc}, vec<int32>(rp), vec<int32>(ci), vec<V>(v)));
                    },
                    py::arg("exec"), py::arg("num_items"), py::arg("rows"), py::arg("row_ptrs"),
                    py::arg("col_idxs"), py::arg("values"), py::arg("cols") = py::none());
    py::class_<BEll, BM, std::shared_ptr<BEll>>(m, ("BatchEll" + sfx).c_str())
        .def_static("from_numpy",
                    [](std::shared_ptr<const Executor> e, size_type num, size_type rows,
                       size_type cols, size_type width,
                       py::array_t<int32, py::array::c_style | py::array::forcecast> ci,
                       py::array_t<V, py::array::c_style | py::array::forcecast> v) {
                        return std::shared_ptr<BEll>(BEll::create_from_host(
                            e, num, dim<2>{rows, cols}, width, vec<int32>(ci), vec<V>(v)));
                    })
        .def_static("from_dense",
                    [](std::shared_ptr<const Executor> e,
                       py::array_t<V, py::array::c_style | py::array::forcecast> d) {
                        if (d.ndim() != 3) {
                            throw DimensionMismatch(__FILE__, __LINE__,
                                                    "from_dense: (items, rows, cols) array");
                        }
                        return std::shared_ptr<BEll>(BEll::create_from_dense_host(
                            e, d.shape(0), dim<2>{static_cast<size_type>(d.shape(1)),
                                                  static_cast<size_type>(d.shape(2))},
                            vec<V>(d)));
                    })
        .def_property_readonly("num_stored_elements_per_row",
                               &BEll::get_num_stored_elements_per_row)
        .def("col_idxs", [](const BEll& a) { return arr(a.col_idxs_to_host()); })
        .def("values", [](const BEll& a) { return arr(a.values_to_host()); });
    py::class_<BDense, BM, std::shared_ptr<BDense>>(m, ("BatchDense" + sfx).c_str())
        .def_static("from_numpy",
                    [](std::shared_ptr<const Executor> e,
                       py::array_t<V, py::array::c_style | py::array::forcecast> d) {
                        if (d.ndim() != 3) {
                            throw DimensionMismatch(__FILE__, __LINE__,
                                                    "from_numpy: (items, rows, cols) array");
                        }
                        return std::shared_ptr<BDense>(BDense::create_from_host(
                            e, d.shape(0), dim<2>{static_cast<size_type>(d.shape(1)),
                                                  static_cast<size_type>(d.shape(2))},
                            vec<V>(d)));
                    })
        .def("values", [](const BDense& a) { return arr(a.values_to_host()); });
    py::class_<BDiag, BM, std::shared_ptr<BDiag>>(
        m, ("BatchBlockDiagCsr" + sfx).c_str())
        .def_static("from_numpy",
                    [](std::shared_ptr<const Executor> e,
                       py::array_t<int64, py::array::c_style | py::array::forcecast> off,
                       py::array_t<int32, py::array::c_style | py::array::forcecast> rp,
                       py::array_t<int32, py::array::c_style | py::array::forcecast> ci,
                       py::array_t<V, py::array::c_style | py::array::forcecast> v) {
                        return std::shared_ptr<BDiag>(BDiag::create_from_host(
                            e, vec<int64>(off), vec<int32>(rp), vec<int32>(ci), vec<V>(v)));
                    });
    using BDEll = batch::matrix::BlockDiagEll<V, int32>;
    using BDDense = batch::matrix::BlockDiagDense<V>;
    py::class_<BDEll, BM, std::shared_ptr<BDEll>>(m, ("BatchBlockDiagEll" + sfx).c_str())
        .def_static("from_numpy",
                    [](std::shared_ptr<const Executor> e,
                       py::array_t<int64, py::array::c_style | py::array::forcecast> off,
                       py::array_t<int32, py::array::c_style | py::array::forcecast> rp,
                       py::array_t<int32, py::array::c_style | py::array::forcecast> ci,
                       py::array_t<V, py::array::c_style | py::array::forcecast> v) {
                        return std::shared_ptr<BDEll>(BDEll::create_from_block_diag_csr_host(
                            e, vec<int64>(off), vec<int32>(rp), vec<int32>(ci), vec<V>(v)));
                    })
        .def_property_readonly("item_offsets",
                               [](const BDEll& a) { return arr(a.get_item_offsets()); })
        .def_property_readonly("item_widths", [](const BDEll& a) { return arr(a.get_item_widths()); })
        .def("col_idxs", [](const BDEll& a) { return arr(a.col_idxs_to_host()); })
        .def("values", [](const BDEll& a) { return arr(a.values_to_host()); });
    py::class_<BDDense, BM, std::shared_ptr<BDDense>>(m, ("BatchBlockDiagDense" + sfx).c_str())
        .def_static("from_numpy",
                    [](std::shared_ptr<const Executor> e,
                       py::array_t<int64, py::array::c_style | py::array::forcecast> off,
                       py::array_t<V, py::array::c_style | py::array::forcecast> v) {
                        return std::shared_ptr<BDDense>(
                            BDDense::create_from_host(e, vec<int64>(off), vec<V>(v)));
                    })
        .def_static("from_block_diag_csr",
                    [](std::shared_ptr<const Executor> e,
                       py::array_t<int64, py::array::c_style | py::array::forcecast> off,
                       py::array_t<int32, py::array::c_style | py::array::forcecast> rp,
                       py::array_t<int32, py::array::c_style | py::array::forcecast> ci,
                       py::array_t<V, py::array::c_style | py::array::forcecast> v) {
                        return std::shared_ptr<BDDense>(
                            BDDense::template create_from_block_diag_csr_host<int32>(
                                e, vec<int64>(off), vec<int32>(rp), vec<int32>(ci), vec<V>(v)));
                    })
        .def_property_readonly("item_offsets",
                               [](const BDDense& a) { return arr(a.get_item_offsets()); })
        .def("values", [](const BDDense& a) { return arr(a.values_to_host()); });
    py::class_<typename BJ::parameters_type>(m, ("BatchJacobiParams" + sfx).c_str())
        .def(py::init<>())
        .def("with_max_block_size", [](typename BJ::parameters_type p, unsigned b) {
            return p.with_max_block_size(b);
        });
    py::class_<BL, std::shared_ptr<BL>>(m, ("BatchConvergence" + sfx).c_str())
        .def_static("create", &BL::create)
        .def_property_readonly("num_iterations",
                               [](const BL& l) { return arr(l.get_num_iterations()); })
        .def_property_readonly("residual_norm",
                               [](const BL& l) { return arr(l.get_residual_norm()); });
    bind_batch_solver<batch::solver::Bicgstab<V>>(m, "BatchBicgstab", sfx);
    bind_batch_solver<batch::solver::Cg<V>>(m, "BatchCg", sfx);
}

template <typename V>
void bind_value_type(py::module_& m, const std::string& sfx)
{
    using Dense = matrix::Dense<V>;
    using Csr = matrix::Csr<V>;
    using Sellp = matrix::Sellp<V>;
    using Ell = matrix::Ell<V>;
    using Coo = matrix::Coo<V>;
    using Hyb = matrix::Hybrid<V>;

    py::class_<Dense, LinOp, std::shared_ptr<Dense>>(m, ("Dense" + sfx).c_str())
        .def_static("create",
                    [](std::shared_ptr<const Executor> e, size_type r, size_type c) {
                        return std::shared_ptr<Dense>(Dense::create(e, dim<2>{r, c}));
                    })
        .def_static("create_view",
                    [](std::shared_ptr<const Executor> e, size_type r, size_type c,
                       std::uintptr_t ptr, size_type stride) {
                        return std::shared_ptr<Dense>(Dense::create_view(
                            e, dim<2>{r, c}, reinterpret_cast<V*>(ptr), stride));
                    })
        .def_static("from_numpy",
                    [](std::shared_ptr<const Executor> e,
                       py::array_t<V, py::array::c_style | py::array::forcecast> a) {
                        const size_type r = a.ndim() > 0 ? a.shape(0) : 1;
                        const size_type c = a.ndim() > 1 ? a.shape(1) : 1;
                        return std::shared_ptr<Dense>(
                            Dense::create_from_host(e, dim<2>{r, c}, vec<V>(a)));
                    })
        .def("to_numpy",
             [](const Dense& d) {
                 auto h = d.to_host();
                 py::array_t<V> out({d.get_size().rows, d.get_size().cols});
                 std::copy(h.begin(), h.end(), out.mutable_data());
                 return out;
             })
        .def("fill", &Dense::fill)
        .def("scale", &Dense::scale)
        .def("add_scaled", &Dense::add_scaled)
        .def("compute_dot", &Dense::compute_dot)
        .def("compute_norm2", &Dense::compute_norm2)
        .def("copy_from", [](Dense& d, const Dense& o) { d.copy_from(&o); })
        .def("copy_from", [](Dense& d, const matrix::Dense<std::conditional_t<
                                          std::is_same<V, double>::value, float, double>>& o) {
            d.copy_from(&o);
        })
        .def_property_readonly("stride", &Dense::get_stride)
        .def_property_readonly("data_ptr", [](Dense& d) {
            return reinterpret_cast<std::uintptr_t>(d.get_values());
        });

    py::class_<Csr, LinOp, std::shared_ptr<Csr>>(m, ("Csr" + sfx).c_str())
        .def_static(
            "create_from_host",
            [](std::shared_ptr<const Executor> e, size_type r, size_type c,
               py::array_t<int32> rp, py::array_t<int32> ci, py::array_t<V> v,
               const std::string& s, int sw) {
                return std::shared_ptr<Csr>(Csr::create_from_host(
                    e, dim<2>{r, c}, vec<int32>(rp), vec<int32>(ci), vec<V>(v), strategy<V>(s, sw)));
            },
            py::arg("exec"), py::arg("rows"), py::arg("cols"), py::arg("row_ptrs"),
            py::arg("col_idxs"), py::arg("values"), py::arg("strategy") = "automatical",
            py::arg("subwave") = 0)
        .def_static(
            "create_view",
            [](std::shared_ptr<const Executor> e, size_type r, size_type c, size_type nnz,
               std::uintptr_t v, std::uintptr_t ci, std::uintptr_t rp, const std::string& s,
               int sw) {
                return std::shared_ptr<Csr>(Csr::create_view(
                    e, dim<2>{r, c}, nnz, reinterpret_cast<V*>(v), reinterpret_cast<int32*>(ci),
                    reinterpret_cast<int32*>(rp), strategy<V>(s, sw)));
            },
            py::arg("exec"), py::arg("rows"), py::arg("cols"), py::arg("nnz"), py::arg("values"),
            py::arg("col_idxs"), py::arg("row_ptrs"), py::arg("strategy") = "automatical",
            py::arg("subwave") = 0)
        .def("transpose", [](const Csr& a) { return std::shared_ptr<Csr>(a.transpose()); })
        .def("to_sellp",
             [](const Csr& a, size_type ss, size_type sf) {
                 return std::shared_ptr<Sellp>(a.to_sellp(ss, sf));
             },
             py::arg("slice_size") = 64, py::arg("stride_factor") = 1)
        .def("to_ell", [](const Csr& a) { return std::shared_ptr<Ell>(a.to_ell()); })
        .def("to_coo", [](const Csr& a) { return std::shared_ptr<Coo>(a.to_coo()); })
        .def("to_hybrid",
             [](const Csr& a, const std::string& s, double p0, double p1) {
                 std::shared_ptr<typename Hyb::strategy_type> st;
                 if (s == "automatic") st = std::make_shared<typename Hyb::automatic>();
                 else if (s == "column_limit")
                     st = std::make_shared<typename Hyb::column_limit>(static_cast<size_type>(p0));
                 else if (s == "imbalance_limit")
                     st = std::make_shared<typename Hyb::imbalance_limit>(p0);
                 else if (s == "imbalance_bounded_limit")
                     st = std::make_shared<typename Hyb::imbalance_bounded_limit>(p0, p1);
                 else if (s == "minimal_storage_limit")
                     st = std::make_shared<typename Hyb::minimal_storage_limit>();
                 else
                     throw py::value_error("unknown Hybrid strategy " + s);
                 return std::shared_ptr<Hyb>(Hyb::create_from_csr(&a, st));
             },
             py::arg("strategy") = "automatic", py::arg("p0") = 0.0, py::arg("p1") = 0.0)
        .def("to_f32", [](const Csr& a) {
            return std::shared_ptr<matrix::Csr<float>>(a.template convert_precision<float>());
        })
        .def("to_f64", [](const Csr& a) {
            return std::shared_ptr<matrix::Csr<double>>(a.template convert_precision<double>());
        })
        .def("max_row_nnz", &Csr::max_row_nnz)
        .def("compact_indices", &Csr::compact_indices)
        .def("build_column_panels", &Csr::build_column_panels, py::arg("num_panels") = 4)
        .def("build_slices", &Csr::build_slices, py::arg("compact") = true)
        .def_property_readonly("has_slices", &Csr::has_slices)
        .def_property_readonly("num_column_panels", &Csr::get_num_column_panels)
        .def_property_readonly("has_compact_indices", &Csr::has_compact_indices)
        .def_property_readonly("effective_strategy", &Csr::get_effective_strategy)
        .def_property_readonly("nnz", &Csr::get_num_stored_elements)
        .def("row_ptrs", [](const Csr& a) { return arr(a.row_ptrs_to_host()); })
        .def("srow", [](const Csr& a) { return arr(a.srow_to_host()); })
        .def_property_readonly("num_srow_elements", &Csr::get_num_srow_elements)
        .def("col_idxs", [](const Csr& a) { return arr(a.col_idxs_to_host()); })
        .def("values", [](const Csr& a) { return arr(a.values_to_host()); })
        // device address of the values for writing in place (get_values():
        // drops the column panels)
        .def_property_readonly("values_data_ptr", [](Csr& a) {
            return reinterpret_cast<std::uintptr_t>(a.get_values());
        });

    py::class_<Sellp, LinOp, std::shared_ptr<Sellp>>(m, ("Sellp" + sfx).c_str())
        .def_property_readonly("slice_size", &Sellp::get_slice_size)
        .def_property_readonly("total_cols", &Sellp::get_total_cols)
        .def("slice_sets", [](const Sellp& a) { return arr(a.slice_sets_to_host()); })
        .def("col_idxs", [](const Sellp& a) { return arr(a.col_idxs_to_host()); })
        .def("values", [](const Sellp& a) { return arr(a.values_to_host()); });

    py::class_<Ell, LinOp, std::shared_ptr<Ell>>(m, ("Ell" + sfx).c_str())
        .def_property_readonly("width", &Ell::get_num_stored_elements_per_row)
        .def_property_readonly("stride", &Ell::get_stride)
        .def("col_idxs", [](const Ell& a) { return arr(a.col_idxs_to_host()); })
        .def("values", [](const Ell& a) { return arr(a.values_to_host()); });

    py::class_<Coo, LinOp, std::shared_ptr<Coo>>(m, ("Coo" + sfx).c_str())
        .def_static("from_numpy",
                    [](std::shared_ptr<const Executor> e, size_type rows, size_type cols,
                       py::array_t<int32, py::array::c_style | py::array::forcecast> ri,
                       py::array_t<int32, py::array::c_style | py::array::forcecast> ci,
                       py::array_t<V, py::array::c_style | py::array::forcecast> v) {
                        return std::shared_ptr<Coo>(Coo::create_from_host(
                            e, dim<2>{rows, cols}, vec<int32>(ri), vec<int32>(ci), vec<V>(v)));
                    })
        .def("apply2", [](const Coo& a, const LinOp* b, LinOp* x) { a.apply2(b, x); })
        .def("apply2", [](const Coo& a, const LinOp* alpha, const LinOp* b, LinOp* x) {
            a.apply2(alpha, b, x);
        })
        .def("to_csr", [](const Coo& a) { return std::shared_ptr<Csr>(a.to_csr()); })
        .def_property_readonly("nnz", &Coo::get_num_stored_elements)
        .def_property_readonly("index_strategy",
                               [](const Coo& a) { return a.get_index()->get_effective_strategy(); })
        .def("row_idxs", [](const Coo& a) { return arr(a.row_idxs_to_host()); })
        .def("col_idxs", [](const Coo& a) { return arr(a.col_idxs_to_host()); })
        .def("values", [](const Coo& a) { return arr(a.values_to_host()); });

    py::class_<Hyb, LinOp, std::shared_ptr<Hyb>>(m, ("Hybrid" + sfx).c_str())
        .def_property_readonly("ell_width", &Hyb::get_ell_num_stored_elements_per_row)
        .def_property_readonly("ell_stride", &Hyb::get_ell_stride)
        .def_property_readonly("coo_nnz", &Hyb::get_coo_num_stored_elements)
        .def_property_readonly("strategy", [](const Hyb& a) { return a.get_strategy()->get_name(); })
        .def("ell_col_idxs", [](const Hyb& a) { return arr(a.get_ell()->col_idxs_to_host()); })
        .def("ell_values", [](const Hyb& a) { return arr(a.get_ell()->values_to_host()); })
        .def("coo_row_idxs", [](const Hyb& a) { return arr(a.get_coo()->row_idxs_to_host()); })
        .def("coo_col_idxs", [](const Hyb& a) { return arr(a.get_coo()->col_idxs_to_host()); })
        .def("coo_values", [](const Hyb& a) { return arr(a.get_coo()->values_to_host()); });

    // criteria parameters
    using RN = stop::ResidualNorm<V>;
    py::class_<typename RN::parameters_type>(m, ("ResidualNorm" + sfx).c_str())
        .def(py::init<>())
        .def("with_reduction_factor",
             [](typename RN::parameters_type p, double r) { return p.with_reduction_factor(r); })
        .def("with_baseline",
             [](typename RN::parameters_type p, stop::mode b) { return p.with_baseline(b); })
        .def_readonly("reduction_factor", &RN::parameters_type::reduction_factor);
    using IRN = stop::ImplicitResidualNorm<V>;
    py::class_<typename IRN::parameters_type>(m, ("ImplicitResidualNorm" + sfx).c_str())
        .def(py::init<>())
        .def("with_reduction_factor",
             [](typename IRN::parameters_type p, double r) { return p.with_reduction_factor(r); })
        .def("with_baseline",
             [](typename IRN::parameters_type p, stop::mode b) { return p.with_baseline(b); });

    // Jacobi
    using J = preconditioner::Jacobi<V>;
    py::class_<typename J::parameters_type>(m, ("JacobiParams" + sfx).c_str())
        .def(py::init<>())
        .def("with_max_block_size",
             [](typename J::parameters_type p, uint32_t b) { return p.with_max_block_size(b); })
        .def("on", [](const typename J::parameters_type& p, std::shared_ptr<const Executor> e) {
            return std::shared_ptr<LinOpFactory>(p.on(e));
        });
    py::class_<J, LinOp, std::shared_ptr<J>>(m, ("Jacobi" + sfx).c_str())
        .def("inverse_diagonal", [](const J& j) { return arr(j.inverse_diagonal_to_host()); });

    // Krylov solvers
    auto bind_krylov = [&](auto tag, const std::string& name) {
        using S = typename decltype(tag)::type;
        using P = typename S::parameters_type;
        py::class_<P>(m, (name + "Params" + sfx).c_str())
            .def(py::init<>())
            .def("with_criteria",
                 [](P p, py::args a) {
                     std::vector<crit_param> l;
                     for (auto o : a) l.push_back(to_crit(o));
                     p.with_criteria_list(l);
                     return p;
                 })
            .def("with_preconditioner",
                 [](P p, py::object o) { return p.with_preconditioner(to_op(o)); })
            .def("with_generated_preconditioner",
                 [](P p, std::shared_ptr<const LinOp> o) {
                     return p.with_generated_preconditioner(o);
                 })
            .def("on", [](const P& p, std::shared_ptr<const Executor> e) {
                return std::shared_ptr<LinOpFactory>(p.on(e));
            });
        py::class_<S, LinOp, std::shared_ptr<S>>(m, (name + sfx).c_str())
            .def("set_poll_chunk", &S::set_poll_chunk)
            .def_property_readonly("last_iterations",
                                   [](const S& s) { return s.get_last_stats().iterations; })
            .def_property_readonly("last_used_fused_path",
                                   [](const S& s) { return s.get_last_stats().used_fused_path; })
            .def_property_readonly("last_status", [](const S& s) {
                std::vector<int> out;
                for (auto st : s.get_last_stats().status) out.push_back(st.raw());
                return out;
            });
    };
    struct cg_tag { using type = solver::Cg<V>; };
    struct bicg_tag { using type = solver::Bicgstab<V>; };
    bind_krylov(cg_tag{}, "Cg");
    bind_krylov(bicg_tag{}, "Bicgstab");

    using I = solver::Ir<V>;
    using IP = typename I::parameters_type;
    py::class_<IP>(m, ("IrParams" + sfx).c_str())
        .def(py::init<>())
        .def("with_criteria",
             [](IP p, py::args a) {
                 p.criteria.list.clear();
                 for (auto o : a) p.criteria.list.push_back(to_crit(o));
                 return p;
             })
        .def("with_solver", [](IP p, py::object o) { return p.with_solver(to_op(o)); })
        .def("with_relaxation_factor", [](IP p, double w) { return p.with_relaxation_factor(w); })
        .def("on", [](const IP& p, std::shared_ptr<const Executor> e) {
            return std::shared_ptr<LinOpFactory>(p.on(e));
        });
    py::class_<I, LinOp, std::shared_ptr<I>>(m, ("Ir" + sfx).c_str())
        .def_property_readonly("last_iterations",
                               [](const I& s) { return s.get_last_stats().iterations; })
        .def_property_readonly("last_inner_iterations",
                               [](const I& s) { return s.get_last_stats().inner_iterations; })
        .def_property_readonly("last_status", [](const I& s) {
            std::vector<int> out;
            for (auto st : s.get_last_stats().status) out.push_back(st.raw());
            return out;
        });
}


// ---------------------------------------------------- profiler bindings ----

using PH = log::ProfilerHook;

int64 ns_of(std::chrono::nanoseconds t) { return static_cast<int64>(t.count()); }

// A Python callable kept by C++ code that may release it without the GIL.
std::shared_ptr<py::object> keep_py(py::object o)
{
    return std::shared_ptr<py::object>(new py::object(std::move(o)), [](py::object* p) {
        py::gil_scoped_acquire g;
        delete p;
    });
}

PH::hook_function py_hook(py::object fn)
{
    auto f = keep_py(std::move(fn));
    return [f](const char* name, log::profile_event_category c) {
        py::gil_scoped_acquire g;
        (*f)(name, c);
    };
}

py::list py_entries(const std::vector<PH::summary_entry>& e)
{
    py::list out;
    for (const auto& x : e) out.append(x);
    return out;
}

// Hands the summary to a Python object's write(entries, overhead_ns) /
// write_nested(root, overhead_ns).  It runs when the hook is destroyed,
// which may be inside a destructor: an exception is reported, not thrown.
struct PyWriter : PH::SummaryWriter, PH::NestedSummaryWriter {
    explicit PyWriter(py::object o) : obj{keep_py(std::move(o))} {}
    void write(const std::vector<PH::summary_entry>& e, std::chrono::nanoseconds ov) override
    {
        py::gil_scoped_acquire g;
        try {
            obj->attr("write")(py_entries(e), ns_of(ov));
        } catch (py::error_already_set& err) {
            err.discard_as_unraisable("ProfilerHook summary writer");
        }
    }
    void write_nested(const PH::nested_summary_entry& root, std::chrono::nanoseconds ov) override
    {
        py::gil_scoped_acquire g;
        try {
            obj->attr("write_nested")(root, ns_of(ov));
        } catch (py::error_already_set& err) {
            err.discard_as_unraisable("ProfilerHook summary writer");
        }
    }
    std::shared_ptr<py::object> obj;
};

// TableSummaryWriter for Python: the table goes to stream.write (stream
// None: sys.stderr)
struct PyTableWriter {
    py::object stream;
    std::string header;
    void emit(const std::string& text) const
    {
        py::object s = stream.is_none() ? py::module_::import("sys").attr("stderr") : stream;
        s.attr("write")(text);
    }
    void write(const std::vector<PH::summary_entry>& e, int64 overhead) const
    {
        std::ostringstream os;
        PH::TableSummaryWriter(os, header).write(e, std::chrono::nanoseconds{overhead});
        emit(os.str());
    }
    void write_nested(const PH::nested_summary_entry& root, int64 overhead) const
    {
        std::ostringstream os;
        PH::TableSummaryWriter(os, header).write_nested(root, std::chrono::nanoseconds{overhead});
        emit(os.str());
    }
};

// hook.user_range(name) as a context manager
struct PyRange {
    std::shared_ptr<const PH> hook;
    std::string name;
    std::unique_ptr<log::profiling_scope_guard> guard;
};

void bind_profiler(py::module_& m)
{
    py::enum_<log::profile_event_category>(m, "profile_event_category")
        .value("memory", log::profile_event_category::memory)
        .value("operation", log::profile_event_category::operation)
        .value("object", log::profile_event_category::object)
        .value("linop", log::profile_event_category::linop)
        .value("factory", log::profile_event_category::factory)
        .value("solver", log::profile_event_category::solver)
        .value("criterion", log::profile_event_category::criterion)
        .value("user", log::profile_event_category::user)
        .value("internal", log::profile_event_category::internal);

    py::class_<Timer, std::shared_ptr<Timer>>(m, "Timer")
        .def_static("create_for_executor", [](std::shared_ptr<const Executor> e) {
            return std::shared_ptr<Timer>(Timer::create_for_executor(std::move(e)));
        });
    py::class_<CpuTimer, Timer, std::shared_ptr<CpuTimer>>(m, "CpuTimer").def(py::init<>());
    py::class_<Mi355xTimer, Timer, std::shared_ptr<Mi355xTimer>>(m, "Mi355xTimer")
        .def(py::init<std::shared_ptr<const Executor>>());

    py::class_<PH::summary_entry>(m, "summary_entry")
        .def(py::init([](std::string name, int64 inclusive, int64 exclusive, int64 count) {
                 return PH::summary_entry{std::move(name), std::chrono::nanoseconds{inclusive},
                                          std::chrono::nanoseconds{exclusive}, count};
             }),
             py::arg("name"), py::arg("inclusive_ns"), py::arg("exclusive_ns"), py::arg("count"))
        .def_readonly("name", &PH::summary_entry::name)
        .def_property_readonly("inclusive_ns",
                               [](const PH::summary_entry& e) { return ns_of(e.inclusive); })
        .def_property_readonly("exclusive_ns",
                               [](const PH::summary_entry& e) { return ns_of(e.exclusive); })
        .def_readonly("count", &PH::summary_entry::count);
    py::class_<PH::nested_summary_entry>(m, "nested_summary_entry")
        .def(py::init([](std::string name, int64 elapsed, int64 count,
                         std::vector<PH::nested_summary_entry> children) {
                 return PH::nested_summary_entry{std::move(name), std::chrono::nanoseconds{elapsed},
                                                 count, std::move(children)};
             }),
             py::arg("name"), py::arg("elapsed_ns"), py::arg("count"),
             py::arg("children") = std::vector<PH::nested_summary_entry>{})
        .def_readonly("name", &PH::nested_summary_entry::name)
        .def_property_readonly("elapsed_ns",
                               [](const PH::nested_summary_entry& e) { return ns_of(e.elapsed); })
        .def_readonly("count", &PH::nested_summary_entry::count)
        .def_readonly("children", &PH::nested_summary_entry::children);
    py::class_<PyTableWriter>(m, "TableSummaryWriter")
        .def(py::init([](py::object stream, std::string header) {
                 return PyTableWriter{std::move(stream), std::move(header)};
             }),
             py::arg("stream") = py::none(), py::arg("header") = "Runtime summary")
        .def("write", &PyTableWriter::write, py::arg("entries"), py::arg("overhead_ns"))
        .def("write_nested", &PyTableWriter::write_nested, py::arg("root"),
             py::arg("overhead_ns"));
    py::class_<PyRange>(m, "profiling_scope_guard")
        .def("__enter__",
             [](PyRange& r) -> PyRange& {
                 if (!r.guard) {
                     r.guard = std::make_unique<log::profiling_scope_guard>(
                         r.hook->user_range(r.name.c_str()));
                 }
                 return r;
             },
             py::return_value_policy::reference)
        .def("__exit__", [](PyRange& r, py::args) {
            r.guard.reset();
            return false;
        });

    auto timer_or_cpu = [](py::object timer) -> std::shared_ptr<Timer> {
        if (timer.is_none()) return std::make_shared<CpuTimer>();
        return timer.cast<std::shared_ptr<Timer>>();
    };
    py::class_<PH, log::Logger, std::shared_ptr<PH>>(m, "ProfilerHook")
        .def_static("create_custom",
                    [](py::object begin, py::object end) {
                        return PH::create_custom(py_hook(std::move(begin)), py_hook(std::move(end)));
                    },
                    py::arg("begin"), py::arg("end"))
        .def_static("create_roctx", &PH::create_roctx)
        .def_static("create_for_executor",
                    [](std::shared_ptr<const Executor> e) { return PH::create_for_executor(e); })
        .def_static(
            "create_summary",
            [timer_or_cpu](py::object timer, py::object writer, bool check) {
                std::unique_ptr<PH::SummaryWriter> w;
                if (writer.is_none()) {
                    w = std::make_unique<PH::TableSummaryWriter>();
                } else {
                    w = std::make_unique<PyWriter>(std::move(writer));
                }
                return PH::create_summary(timer_or_cpu(std::move(timer)), std::move(w), check);
            },
            py::arg("timer") = py::none(), py::arg("writer") = py::none(),
            py::arg("debug_check_nesting") = false)
        .def_static(
            "create_nested_summary",
            [timer_or_cpu](py::object timer, py::object writer, bool check) {
                std::unique_ptr<PH::NestedSummaryWriter> w;
                if (writer.is_none()) {
                    w = std::make_unique<PH::TableSummaryWriter>();
                } else {
                    w = std::make_unique<PyWriter>(std::move(writer));
                }
                return PH::create_nested_summary(timer_or_cpu(std::move(timer)), std::move(w),
                                                 check);
            },
            py::arg("timer") = py::none(), py::arg("writer") = py::none(),
            py::arg("debug_check_nesting") = false)
        .def("set_object_name",
             [](PH& h, const LinOp* o, std::string n) { h.set_object_name(o, std::move(n)); })
        .def("set_object_name", [](PH& h, const LinOpFactory* o,
                                   std::string n) { h.set_object_name(o, std::move(n)); })
        .def("set_synchronization", &PH::set_synchronization)
        .def("user_range", [](std::shared_ptr<const PH> h, std::string name) {
            return PyRange{std::move(h), std::move(name), nullptr};
        });
}

}  // namespace

PYBIND11_MODULE(_gko, m)
{
    m.doc() = "gko:: host library (MI355X executor) -- pybind11 binding";

    // translators are tried last-registered first: the base class goes first
    auto base = py::register_exception<Error>(m, "GkoError");
    py::register_exception<NotSupported>(m, "NotSupported", base.ptr());
    py::register_exception<DimensionMismatch>(m, "DimensionMismatch", base.ptr());
    py::register_exception<NotImplemented>(m, "NotImplemented", base.ptr());
    py::register_exception<AllocationError>(m, "AllocationError", base.ptr());
    py::register_exception<HipError>(m, "HipError", base.ptr());

    py::class_<Executor, std::shared_ptr<Executor>>(m, "Executor")
        .def("synchronize", &Executor::synchronize)
        .def_property_readonly("is_host", &Executor::is_host)
        .def("add_logger",
             [](const Executor& e, std::shared_ptr<const log::Logger> l) { e.add_logger(l); })
        .def("remove_logger",
             [](const Executor& e, const log::Logger* l) { e.remove_logger(l); })
        .def("clear_loggers", &Executor::clear_loggers)
        .def_property_readonly("num_loggers",
                               [](const Executor& e) { return e.get_loggers().size(); })
        .def("set_log_propagation", &Executor::set_log_propagation)
        .def("should_propagate_log", &Executor::should_propagate_log);
    py::class_<HostExecutor, Executor, std::shared_ptr<HostExecutor>>(m, "HostExecutor")
        .def_static("create", &HostExecutor::create);
    py::class_<Graph, std::shared_ptr<Graph>>(m, "Graph")
        .def("replay", &Graph::replay)
        .def_property_readonly("executor", &Graph::get_executor);
    py::class_<Mi355xExecutor, Executor, std::shared_ptr<Mi355xExecutor>>(m, "Mi355xExecutor")
        .def_static(
            "create",
            // stream None: an own stream; an integer handle (0 = the null
            // stream, where torch runs by default): exactly that stream
            [](int dev, py::object stream, bool caching) {
                if (stream.is_none()) return Mi355xExecutor::create(dev, nullptr, caching);
                return Mi355xExecutor::create_on_stream(
                    dev, reinterpret_cast<void*>(stream.cast<std::uintptr_t>()), caching);
            },
            py::arg("device") = 0, py::arg("stream") = py::none(), py::arg("caching") = true)
        .def_static("get_num_devices", &Mi355xExecutor::get_num_devices)
        .def_property_readonly("device_id", &Mi355xExecutor::get_device_id)
        .def_property_readonly("num_multiprocessor", &Mi355xExecutor::get_num_multiprocessor)
        .def_property_readonly("stream", [](const Mi355xExecutor& e) {
            return reinterpret_cast<std::uintptr_t>(e.get_stream());
        })
        .def_property_readonly("pool_bytes", &Mi355xExecutor::get_pool_bytes)
        .def("release_pool", &Mi355xExecutor::release_pool);

    py::class_<LinOp, std::shared_ptr<LinOp>>(m, "LinOp")
        .def("apply", [](const LinOp& op, const LinOp* b, LinOp* x) { op.apply(b, x); })
        .def("apply", [](const LinOp& op, const LinOp* a, const LinOp* b, const LinOp* beta,
                         LinOp* x) { op.apply(a, b, beta, x); })
        .def_property_readonly("size",
                               [](const LinOp& op) {
                                   return py::make_tuple(op.get_size().rows, op.get_size().cols);
                               })
        .def("add_logger", [](const LinOp& op, std::shared_ptr<const log::Logger> l) {
            op.add_logger(l);
        })
        .def("remove_logger", [](const LinOp& op, const log::Logger* l) { op.remove_logger(l); })
        .def("clear_loggers", &LinOp::clear_loggers);
    py::class_<LinOpFactory, std::shared_ptr<LinOpFactory>>(m, "LinOpFactory")
        .def("generate", [](const LinOpFactory& f, std::shared_ptr<const LinOp> a) {
            return std::shared_ptr<LinOp>(f.generate(a));
        })
        .def("add_logger", [](const LinOpFactory& f, std::shared_ptr<const log::Logger> l) {
            f.add_logger(l);
        })
        .def("clear_loggers", &LinOpFactory::clear_loggers);

    py::enum_<stop::mode>(m, "mode")
        .value("absolute", stop::mode::absolute)
        .value("initial_resnorm", stop::mode::initial_resnorm)
        .value("rhs_norm", stop::mode::rhs_norm);
    py::class_<stop::CriterionFactory, std::shared_ptr<stop::CriterionFactory>>(m,
                                                                                "CriterionFactory");
    py::class_<stop::Iteration::parameters_type>(m, "Iteration")
        .def(py::init<>())
        .def("with_max_iters",
             [](stop::Iteration::parameters_type p, size_type n) { return p.with_max_iters(n); });

    py::class_<log::Logger, std::shared_ptr<log::Logger>>(m, "Logger");
    py::class_<log::Convergence, log::Logger, std::shared_ptr<log::Convergence>>(m, "Convergence")
        .def_static("create", &log::Convergence::create)
        .def_property_readonly("num_iterations", &log::Convergence::get_num_iterations)
        .def_property_readonly("has_converged", &log::Convergence::has_converged)
        .def_property_readonly("residual_norm", &log::Convergence::get_residual_norm);
    bind_profiler(m);

    bind_value_type<double>(m, "_f64");
    bind_value_type<float>(m, "_f32");
    py::enum_<batch::stop::tolerance_type>(m, "batch_tolerance_type")
        .value("absolute", batch::stop::tolerance_type::absolute)
        .value("relative", batch::stop::tolerance_type::relative);
    py::class_<batch::BatchLinOp, std::shared_ptr<batch::BatchLinOp>>(m, "BatchLinOp")
        .def_property_readonly("num_batch_items", &batch::BatchLinOp::get_num_batch_items);
    bind_batch<double>(m, "_f64");
    bind_batch<float>(m, "_f32");
}
