// C++17 front-end of libtinyad_amd.so with the shape of TinyAD's
// ScalarFunction API (include/TinyAD/ScalarFunction.hh, Utils/NewtonDirection.hh,
// Utils/LineSearch.hh, Utils/Helpers.hh).  Header-only; every call goes
// through the C ABI of include/tinyad_amd.h and non-zero statuses become
// std::runtime_error, as the reference's TINYAD_ERROR_throw / TINYAD_ASSERT do
// (Utils/Out.hh:64-157).
//
// Differences from the reference, by design:
//   * element functions are entries of the library's energy catalog, or
//     device functors compiled by the caller (DeviceElements.hh), plus
//     per-element parameter rows (the reference instantiates a user lambda
//     for three scalar types, ScalarObjectiveTerm.hh:54-115);
//   * Eigen is optional: vectors are std::vector<double>, the Hessian is a
//     CSR SparseMatrix with the same pattern Eigen's setFromTriplets yields.
//
//   auto func = tinyad_amd::scalar_function<2>(tinyad_amd::range(n_vertices));
//   func.add_elements<3>(tinyad_amd::range(n_faces), TAD_E_SYMDIR_TRI,
//       [&](int f) { return std::array<int, 3>{F[f][0], F[f][1], F[f][2]}; },
//       [&](int f) { return rest_params[f]; });                  // 5 doubles
//   std::vector<double> x = func.x_from_data([&](int v) { return std::array<double, 2>{...}; });
//   auto [f, g, H_proj] = func.eval_with_hessian_proj(x);
#pragma once

#include <tinyad_amd.h>

#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <cstring>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

namespace tinyad_amd {

inline void check(tad_status s)
{
    if (s != TAD_OK)
        throw std::runtime_error(std::string("[tinyad_amd] ") + tad_last_error());
}

constexpr double default_hessian_projection_eps = 1e-9;  // HessianProjection.hh:16

using VectorXd = std::vector<double>;

// range(n) = {0, ..., n-1} (Helpers.hh:17-27)
inline std::vector<int> range(int n)
{
    if (n < 0) throw std::runtime_error("[tinyad_amd] range: negative size");
    std::vector<int> r(n);
    for (int i = 0; i < n; ++i) r[i] = i;
    return r;
}

// Handle types (Support/Common.hh, Support/*.hh): functions are built over a
// range of variable handles, and every handle maps to an index through
// idx_from_handle().  Integers map to themselves; handles of OpenMesh and PMP
// (h.idx()), geometry-central (h.getIndex()) and polymesh (h.idx.value) are
// recognised by those members, so no mesh-library header is needed here.
// For another handle type, overload idx_from_handle(const MyHandle&) in
// namespace tinyad_amd before including this header.
namespace detail {
template <class H, class = void> struct has_idx_fn : std::false_type {};
template <class H> struct has_idx_fn<H, std::void_t<decltype(std::declval<const H&>().idx() + 0)>> : std::true_type {};
template <class H, class = void> struct has_get_index : std::false_type {};
template <class H> struct has_get_index<H, std::void_t<decltype(std::declval<const H&>().getIndex() + 0)>> : std::true_type {};
template <class H, class = void> struct has_idx_value : std::false_type {};
template <class H> struct has_idx_value<H, std::void_t<decltype(std::declval<const H&>().idx.value + 0)>> : std::true_type {};
} // namespace detail

template <class H>
int64_t idx_from_handle(const H& h)
{
    if constexpr (std::is_integral<H>::value) return (int64_t)h;
    else if constexpr (detail::has_idx_fn<H>::value) return (int64_t)h.idx();
    else if constexpr (detail::has_get_index<H>::value) return (int64_t)h.getIndex();
    else if constexpr (detail::has_idx_value<H>::value) return (int64_t)h.idx.value;
    else
    {
        static_assert(sizeof(H) == 0, "Handle type not supported: overload tinyad_amd::idx_from_handle() for it "
                                      "(TinyAD Support/Common.hh)");
        return -1;
    }
}

// CSR matrix (row-major).  For the symmetric Hessians returned here it equals
// the column-major Eigen::SparseMatrix the reference returns.
struct SparseMatrix
{
    int64_t n_rows = 0, n_cols = 0;
    std::vector<int32_t> outer;  // row offsets, n_rows + 1
    std::vector<int32_t> inner;  // column indices, sorted within each row
    std::vector<double> values;

    int64_t rows() const { return n_rows; }
    int64_t cols() const { return n_cols; }
    int64_t nonZeros() const { return (int64_t)values.size(); }
    double coeff(int64_t r, int64_t c) const
    {
        for (int32_t k = outer[r]; k < outer[r + 1]; ++k)
            if (inner[k] == c) return values[k];
        return 0.0;
    }
    VectorXd operator*(const VectorXd& v) const
    {
        VectorXd y(n_rows, 0.0);
        for (int64_t r = 0; r < n_rows; ++r)
            for (int32_t k = outer[r]; k < outer[r + 1]; ++k) y[r] += values[k] * v[inner[k]];
        return y;
    }
};

struct EvalSettings
{
    int device = 0;
    tad_precision precision = TAD_PREC_FP64;
};

// One per device, shared by the functions created on it.
class Context
{
public:
    explicit Context(int device = 0) { check(tad_ctx_create(device, &h_)); }
    ~Context() { tad_ctx_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    tad_ctx handle() const { return h_; }

    static std::shared_ptr<Context> on_device(int device)
    {
        static std::vector<std::weak_ptr<Context>> cache;
        if ((int)cache.size() <= device) cache.resize(device + 1);
        std::shared_ptr<Context> c = cache[device].lock();
        if (!c)
        {
            c = std::make_shared<Context>(device);
            cache[device] = c;
        }
        return c;
    }

private:
    tad_ctx h_ = nullptr;
};

// ScalarFunction<d, double, VariableHandleT> (ScalarFunction.hh:36-236).
// Movable, not copyable.
template <int variable_dimension, class VariableHandleT = int>
class ScalarFunction
{
    static_assert(variable_dimension >= 1 && variable_dimension <= 3, "variable dimension must be 1, 2 or 3");

public:
    static constexpr int d = variable_dimension;
    using PassiveVariableVectorType = std::array<double, d>;

    ScalarFunction() = default;
    ScalarFunction(std::vector<VariableHandleT> variable_handles, const EvalSettings& settings_ = {})
        : settings(settings_), variable_handles(std::move(variable_handles))
    {
        // handle indices must be a permutation of 0..n-1 (variable_indices_compact,
        // ScalarFunctionImpl.hh:47-61, Element.hh:176-194)
        std::vector<char> used(this->variable_handles.size(), 0);
        for (const VariableHandleT& h : this->variable_handles)
        {
            const int64_t v = idx_from_handle(h);
            if (v < 0 || v >= (int64_t)used.size() || used[v])
                throw std::runtime_error("[tinyad_amd] variable handles are not a permutation of 0..n-1");
            used[v] = 1;
        }
        if (this->variable_handles.empty()) throw std::runtime_error("[tinyad_amd] no variable handles");
        ctx_ = Context::on_device(settings.device);
        n_vars = d * (int64_t)this->variable_handles.size();
        tad_fn h = nullptr;
        check(tad_fn_create(ctx_->handle(), d, (int64_t)this->variable_handles.size(), settings.precision, &h));
        fn_.reset(h, [](tad_fn p) { tad_fn_destroy(p); });
    }
    ScalarFunction(const ScalarFunction&) = delete;
    ScalarFunction& operator=(const ScalarFunction&) = delete;
    ScalarFunction(ScalarFunction&&) = default;
    ScalarFunction& operator=(ScalarFunction&&) = default;

    // add_elements<N>(range, ...) (ScalarFunctionImpl.hh:63-100): conn_of(e)
    // returns the N handles element e accesses, in access order; params_of(e)
    // returns its parameter row (any container with size() and operator[]).
    template <int element_valence, class Range, class ConnFn, class ParamFn>
    void add_elements(const Range& element_range, tad_energy energy, ConnFn conn_of, ParamFn params_of)
    {
        require_fn();
        std::vector<int32_t> conn;
        std::vector<double> params;
        int64_t n = 0;
        int np = -1;
        for (auto e : element_range)
        {
            const auto c = conn_of(e);
            static_assert(std::tuple_size<std::decay_t<decltype(c)>>::value == element_valence,
                          "conn_of must return std::array<handle, N>");
            for (int a = 0; a < element_valence; ++a) conn.push_back((int32_t)idx_from_handle(c[a]));
            const auto p = params_of(e);
            if (np < 0) np = (int)p.size();
            if ((int)p.size() != np) throw std::runtime_error("[tinyad_amd] parameter rows differ in length");
            for (int k = 0; k < np; ++k) params.push_back(p[k]);
            ++n;
        }
        check(tad_fn_add_elements(fn_.get(), energy, element_valence, n, conn.data(), params.data(),
                                  np < 0 ? 0 : np));
        n_elements += n;
    }

    // add_elements<N>(range, lambda) for an energy compiled on the device:
    // ops from tinyad_amd::device_element<F, N, d, NP>() (DeviceElements.hh),
    // conn_of / params_of as above (params_of rows must have NP entries).
    template <int element_valence, class Range, class ConnFn, class ParamFn>
    void add_elements(const Range& element_range, const tad_device_ops& ops, ConnFn conn_of, ParamFn params_of)
    {
        require_fn();
        if (ops.valence != element_valence)
            throw std::runtime_error("[tinyad_amd] device energy valence differs from add_elements<N>");
        std::vector<int32_t> conn;
        std::vector<double> params;
        int64_t n = 0;
        for (auto e : element_range)
        {
            const auto c = conn_of(e);
            static_assert(std::tuple_size<std::decay_t<decltype(c)>>::value == element_valence,
                          "conn_of must return std::array<handle, N>");
            for (int a = 0; a < element_valence; ++a) conn.push_back((int32_t)idx_from_handle(c[a]));
            const auto p = params_of(e);
            if ((int)p.size() != ops.n_params)
                throw std::runtime_error("[tinyad_amd] parameter row length differs from the device energy's");
            for (int k = 0; k < ops.n_params; ++k) params.push_back(p[k]);
            ++n;
        }
        check(tad_fn_add_elements_device(fn_.get(), &ops, n, conn.data(), params.data()));
        n_elements += n;
    }

    // add_elements_dynamic<V1, V2, ...>(range, lambda) (ScalarFunctionImpl.hh:102-214)
    // for a device energy compiled at each valence of ops: handles_of(e)
    // returns element e's handles (any container, at most the largest valence).
    template <class Range, class HandlesFn, class ParamFn>
    void add_elements_dynamic(const Range& element_range, const std::vector<tad_device_ops>& ops,
                              HandlesFn handles_of, ParamFn params_of)
    {
        require_fn();
        if (ops.empty()) throw std::runtime_error("[tinyad_amd] add_elements_dynamic: no valences");
        std::vector<int64_t> ptr{ 0 };
        std::vector<int32_t> handles;
        std::vector<double> params;
        for (auto e : element_range)
        {
            const auto h = handles_of(e);
            for (const auto& v : h) handles.push_back((int32_t)idx_from_handle(v));
            ptr.push_back((int64_t)handles.size());
            const auto p = params_of(e);
            if ((int)p.size() != ops[0].n_params)
                throw std::runtime_error("[tinyad_amd] parameter row length differs from the device energy's");
            for (int k = 0; k < ops[0].n_params; ++k) params.push_back(p[k]);
        }
        std::vector<const tad_device_ops*> op_ptrs;
        for (const auto& o : ops) op_ptrs.push_back(&o);
        const int64_t n = (int64_t)ptr.size() - 1;
        check(tad_fn_add_elements_dynamic_device(fn_.get(), op_ptrs.data(), (int)op_ptrs.size(), n, ptr.data(),
                                                 handles.data(), params.data()));
        n_elements += n;
    }

    // objective terms: number, and (energy, valence, element count) per term
    int n_terms() const
    {
        int n = 0;
        check(tad_fn_n_terms(fn_.get(), &n));
        return n;
    }
    int64_t term_elements(int term) const
    {
        int64_t n = 0;
        check(tad_fn_term_info(fn_.get(), term, nullptr, nullptr, &n));
        return n;
    }

    // x_from_data (ScalarFunctionImpl.hh:216-235)
    VectorXd x_from_data(std::function<PassiveVariableVectorType(VariableHandleT)> read_user_data) const
    {
        VectorXd x(n_vars);
        for (const VariableHandleT& h : variable_handles)
        {
            const int64_t v = idx_from_handle(h);
            const PassiveVariableVectorType p = read_user_data(h);
            for (int i = 0; i < d; ++i) x[d * v + i] = p[i];
        }
        for (double xi : x)
            if (!(xi == xi) || xi == 1.0 / 0.0 || xi == -1.0 / 0.0)
                throw std::runtime_error("[tinyad_amd] x_from_data: non-finite value");
        return x;
    }

    // x_to_data (ScalarFunctionImpl.hh:237-254)
    void x_to_data(const VectorXd& x,
                   std::function<void(VariableHandleT, const PassiveVariableVectorType&)> write_user_data) const
    {
        if ((int64_t)x.size() != n_vars) throw std::runtime_error("[tinyad_amd] x_to_data: size mismatch");
        for (const VariableHandleT& h : variable_handles)
        {
            const int64_t v = idx_from_handle(h);
            PassiveVariableVectorType p;
            for (int i = 0; i < d; ++i) p[i] = x[d * v + i];
            write_user_data(h, p);
        }
    }

    double eval(const VectorXd& x) const
    {
        require_fn();
        check_size(x);
        double f = 0.0;
        check(tad_fn_eval(fn_.get(), x.data(), TAD_MEM_HOST, &f));
        return f;
    }
    double operator()(const VectorXd& x) const { return eval(x); }

    std::tuple<double, VectorXd> eval_with_gradient(const VectorXd& x) const
    {
        require_fn();
        check_size(x);
        double f = 0.0;
        VectorXd g(n_vars);
        check(tad_fn_eval_with_gradient(fn_.get(), x.data(), &f, g.data(), TAD_MEM_HOST));
        return { f, std::move(g) };
    }

    std::tuple<double, VectorXd, SparseMatrix> eval_with_derivatives(const VectorXd& x) const
    {
        require_fn();
        check_size(x);
        double f = 0.0;
        VectorXd g(n_vars);
        SparseMatrix H = empty_hessian();
        check(tad_fn_eval_with_derivatives(fn_.get(), x.data(), &f, g.data(), H.values.data(), TAD_MEM_HOST));
        return { f, std::move(g), std::move(H) };
    }

    SparseMatrix eval_hessian(const VectorXd& x) const { return std::get<2>(eval_with_derivatives(x)); }
    SparseMatrix eval_hessian_of_quadratic() const { return eval_hessian(VectorXd(n_vars, 0.0)); }

    std::tuple<double, VectorXd, SparseMatrix> eval_with_hessian_proj(
        const VectorXd& x, double projection_eps = default_hessian_projection_eps) const
    {
        require_fn();
        check_size(x);
        double f = 0.0;
        VectorXd g(n_vars);
        SparseMatrix H = empty_hessian();
        check(tad_fn_eval_with_hessian_proj(fn_.get(), x.data(), projection_eps, &f, g.data(), H.values.data(),
                                            TAD_MEM_HOST));
        return { f, std::move(g), std::move(H) };
    }

    // --- beyond the reference's call shape (no counterpart in TinyAD) -------

    // Handle ranges [offsets[s], offsets[s+1]) are independent meshes evaluated
    // together; segment_energies() returns their energies from the last eval.
    void set_segments(const std::vector<int64_t>& handle_offsets)
    {
        require_fn();
        if (handle_offsets.size() < 2) throw std::runtime_error("[tinyad_amd] set_segments: need two offsets");
        check(tad_fn_set_segments(fn_.get(), (int64_t)handle_offsets.size() - 1, handle_offsets.data()));
        n_segments_ = (int64_t)handle_offsets.size() - 1;
    }
    std::vector<double> segment_energies() const
    {
        require_fn();
        std::vector<double> f((size_t)n_segments_);
        check(tad_fn_segment_energies(fn_.get(), f.data(), TAD_MEM_HOST));
        return f;
    }

    // Times the element-kernel and assembly launch shapes on x and keeps the
    // fastest; results are bit-identical for every shape.  Returns the report.
    std::string autotune(const VectorXd& x, double projection_eps = default_hessian_projection_eps)
    {
        require_fn();
        check_size(x);
        std::string report(16384, '\0');
        check(tad_fn_autotune(fn_.get(), x.data(), TAD_MEM_HOST, projection_eps, &report[0], report.size()));
        report.resize(std::strlen(report.c_str()));
        return report;
    }

    // Launch shapes (as chosen by autotune) to store and restore without timing.
    struct LaunchShape
    {
        int block = 0, grid = 0, map = 0;
    };
    LaunchShape launch_shape(int term) const
    {
        require_fn();
        LaunchShape ls;
        check(tad_fn_get_launch_shape(fn_.get(), term, &ls.block, &ls.grid, &ls.map));
        return ls;
    }
    void set_launch_shape(int term, const LaunchShape& ls)
    {
        require_fn();
        check(tad_fn_set_launch_shape(fn_.get(), term, ls.block, ls.grid, ls.map));
    }
    int assembly_shape() const
    {
        require_fn();
        int v = 0;
        check(tad_fn_get_assembly_shape(fn_.get(), &v));
        return v;
    }
    void set_assembly_shape(int shape)
    {
        require_fn();
        check(tad_fn_set_assembly_shape(fn_.get(), shape));
    }

    // Replaces the parameter rows of term `term` (the terms in add_elements
    // order), n_elem x params_per_elem row-major; the pattern is kept.
    void set_params(int term, const std::vector<double>& params)
    {
        require_fn();
        check(tad_fn_set_params(fn_.get(), term, params.data(), TAD_MEM_HOST));
    }

    tad_fn handle() const { return fn_.get(); }

    // Non-zeros of the Hessian pattern (explicit zeros included).
    int64_t nnz() const
    {
        require_fn();
        int64_t nz = 0;
        check(tad_fn_pattern(fn_.get(), &nz, nullptr, nullptr, nullptr, nullptr));
        return nz;
    }

    EvalSettings settings;
    int64_t n_vars = 0;
    int64_t n_elements = 0;
    std::vector<VariableHandleT> variable_handles;

private:
    int64_t n_segments_ = 0;
    void require_fn() const
    {
        if (!fn_) throw std::runtime_error("[tinyad_amd] empty (default-constructed or moved-from) ScalarFunction");
    }
    void check_size(const VectorXd& x) const
    {
        if ((int64_t)x.size() != n_vars)
            throw std::runtime_error("[tinyad_amd] x has " + std::to_string(x.size()) + " entries, expected " +
                                     std::to_string(n_vars));
    }
    SparseMatrix empty_hessian() const
    {
        SparseMatrix H;
        int64_t nnz = 0;
        check(tad_fn_pattern(fn_.get(), &nnz, nullptr, nullptr, nullptr, nullptr));
        H.n_rows = H.n_cols = n_vars;
        H.outer.resize(n_vars + 1);
        H.inner.resize(nnz);
        H.values.resize(nnz > 0 ? nnz : 1);
        check(tad_fn_pattern(fn_.get(), nullptr, nullptr, nullptr, H.outer.data(), H.inner.data()));
        H.values.resize(nnz);
        return H;
    }

    std::shared_ptr<Context> ctx_;
    std::shared_ptr<tad_fn_s> fn_;
};

// scalar_function<d>(range) (ScalarFunctionImpl.hh:418-442)
// over any range of handles (a std::vector, mesh.vertices(), ...; Support/*.hh)
template <int variable_dimension, class HandleRange>
auto scalar_function(const HandleRange& variable_range, const EvalSettings& settings = {})
{
    using H = std::decay_t<decltype(*std::begin(variable_range))>;
    std::vector<H> handles;
    for (const auto& h : variable_range) handles.push_back(h);
    return ScalarFunction<variable_dimension, H>(std::move(handles), settings);
}

// VectorFunction<d, double, VariableHandleT> (VectorFunction.hh:19-192): a
// function R^n -> R^m whose elements produce fixed-size residual segments,
// compiled on the device as residual functors (device_residual<F, N, d, M, NP>()
// in DeviceElements.hh).  Movable, not copyable.  J is returned as a CSR
// SparseMatrix (n_outputs x n_vars) with the pattern the reference's triplets
// give (each residual row holds its element's variables, ascending).
template <int variable_dimension, class VariableHandleT = int>
class VectorFunction
{
    static_assert(variable_dimension >= 1 && variable_dimension <= 3, "variable dimension must be 1, 2 or 3");

public:
    static constexpr int d = variable_dimension;
    using PassiveVariableVectorType = std::array<double, d>;

    VectorFunction() = default;
    VectorFunction(std::vector<VariableHandleT> variable_handles, const EvalSettings& settings_ = {})
        : settings(settings_), variable_handles(std::move(variable_handles))
    {
        // handle indices must be a permutation of 0..n-1 (VectorFunctionImpl.hh:40-56)
        std::vector<char> used(this->variable_handles.size(), 0);
        for (const VariableHandleT& h : this->variable_handles)
        {
            const int64_t v = idx_from_handle(h);
            if (v < 0 || v >= (int64_t)used.size() || used[v])
                throw std::runtime_error("[tinyad_amd] variable handles are not a permutation of 0..n-1");
            used[v] = 1;
        }
        if (this->variable_handles.empty()) throw std::runtime_error("[tinyad_amd] no variable handles");
        ctx_ = Context::on_device(settings.device);
        n_vars = d * (int64_t)this->variable_handles.size();
        tad_fn h = nullptr;
        check(tad_fn_create_vector(ctx_->handle(), d, (int64_t)this->variable_handles.size(), settings.precision,
                                   &h));
        fn_.reset(h, [](tad_fn p) { tad_fn_destroy(p); });
    }
    VectorFunction(const VectorFunction&) = delete;
    VectorFunction& operator=(const VectorFunction&) = delete;
    VectorFunction(VectorFunction&&) = default;
    VectorFunction& operator=(VectorFunction&&) = default;

    // add_elements<N, M>(range, lambda) (VectorFunctionImpl.hh:58-90) for a
    // residual functor: ops from device_residual<F, N, d, M, NP>(), conn_of(e)
    // the N handles element e accesses, params_of(e) its NP parameters.
    template <int element_valence, int outputs_per_element, class Range, class ConnFn, class ParamFn>
    void add_elements(const Range& element_range, const tad_device_residual_ops& ops, ConnFn conn_of,
                      ParamFn params_of)
    {
        require_fn();
        if (ops.valence != element_valence || ops.n_outputs != outputs_per_element)
            throw std::runtime_error("[tinyad_amd] residual ops differ from add_elements<N, M>");
        std::vector<int32_t> conn;
        std::vector<double> params;
        int64_t n = 0;
        for (auto e : element_range)
        {
            const auto c = conn_of(e);
            static_assert(std::tuple_size<std::decay_t<decltype(c)>>::value == element_valence,
                          "conn_of must return std::array<handle, N>");
            for (int a = 0; a < element_valence; ++a) conn.push_back((int32_t)idx_from_handle(c[a]));
            const auto p = params_of(e);
            if ((int)p.size() != ops.n_params)
                throw std::runtime_error("[tinyad_amd] parameter row length differs from the residual's");
            for (int k = 0; k < ops.n_params; ++k) params.push_back(p[k]);
            ++n;
        }
        check(tad_fn_add_residuals_device(fn_.get(), &ops, n, conn.data(), params.data()));
        n_elements += n;
        n_outputs += n * outputs_per_element;
        jac_.outer.clear();
    }

    VectorXd x_from_data(std::function<PassiveVariableVectorType(VariableHandleT)> read_user_data) const
    {
        VectorXd x((size_t)n_vars, std::nan(""));
        for (const VariableHandleT& h : variable_handles)
        {
            const PassiveVariableVectorType v = read_user_data(h);
            for (int i = 0; i < d; ++i) x[d * idx_from_handle(h) + i] = v[i];
        }
        for (double xi : x)
            if (!std::isfinite(xi)) throw std::runtime_error("[tinyad_amd] x_from_data: non-finite entry");
        return x;
    }

    void x_to_data(const VectorXd& x, std::function<void(VariableHandleT, PassiveVariableVectorType)> write_user_data)
        const
    {
        check_size(x);
        for (const VariableHandleT& h : variable_handles)
        {
            PassiveVariableVectorType v;
            for (int i = 0; i < d; ++i) v[i] = x[d * idx_from_handle(h) + i];
            write_user_data(h, v);
        }
    }

    // eval(x) / operator() (VectorFunctionImpl.hh:130-147): the residuals.
    VectorXd eval(const VectorXd& x) const
    {
        require_fn();
        check_size(x);
        VectorXd r((size_t)n_outputs);
        check(tad_fn_eval_residuals(fn_.get(), x.data(), r.data(), TAD_MEM_HOST));
        return r;
    }
    VectorXd operator()(const VectorXd& x) const { return eval(x); }

    // eval_with_jacobian (VectorFunctionImpl.hh:157-190)
    void eval_with_jacobian(const VectorXd& x, VectorXd& r, SparseMatrix& J) const
    {
        require_fn();
        check_size(x);
        r.assign((size_t)n_outputs, 0.0);
        J = empty_jacobian();
        check(tad_fn_eval_with_jacobian(fn_.get(), x.data(), r.data(), J.values.data(), TAD_MEM_HOST));
    }
    std::tuple<VectorXd, SparseMatrix> eval_with_jacobian(const VectorXd& x) const
    {
        VectorXd r;
        SparseMatrix J;
        eval_with_jacobian(x, r, J);
        return { std::move(r), std::move(J) };
    }

    // eval_sum_of_squares (VectorFunctionImpl.hh:228-239)
    double eval_sum_of_squares(const VectorXd& x) const
    {
        require_fn();
        check_size(x);
        double f = 0.0;
        check(tad_fn_eval(fn_.get(), x.data(), TAD_MEM_HOST, &f));
        return f;
    }

    // eval_sum_of_squares_with_derivatives (VectorFunctionImpl.hh:241-297):
    // f = r . r, g = 2 J^T r
    void eval_sum_of_squares_with_derivatives(const VectorXd& x, double& f, VectorXd& g, VectorXd& r,
                                              SparseMatrix& J) const
    {
        require_fn();
        check_size(x);
        g.assign((size_t)n_vars, 0.0);
        r.assign((size_t)n_outputs, 0.0);
        J = empty_jacobian();
        check(tad_fn_eval_sum_of_squares_with_derivatives(fn_.get(), x.data(), &f, g.data(), r.data(),
                                                          J.values.data(), TAD_MEM_HOST));
    }
    std::tuple<double, VectorXd, VectorXd, SparseMatrix> eval_sum_of_squares_with_derivatives(const VectorXd& x) const
    {
        double f = 0.0;
        VectorXd g, r;
        SparseMatrix J;
        eval_sum_of_squares_with_derivatives(x, f, g, r, J);
        return { f, std::move(g), std::move(r), std::move(J) };
    }

    tad_fn handle() const { return fn_.get(); }

    // Non-zeros of the Gauss-Newton matrix pattern (the scalar Hessian pattern).
    int64_t nnz() const
    {
        require_fn();
        int64_t nz = 0;
        check(tad_fn_pattern(fn_.get(), &nz, nullptr, nullptr, nullptr, nullptr));
        return nz;
    }

    EvalSettings settings;
    int64_t n_vars = 0;
    int64_t n_elements = 0;
    int64_t n_outputs = 0;
    std::vector<VariableHandleT> variable_handles;

private:
    void require_fn() const
    {
        if (!fn_) throw std::runtime_error("[tinyad_amd] empty (default-constructed or moved-from) VectorFunction");
    }
    void check_size(const VectorXd& x) const
    {
        if ((int64_t)x.size() != n_vars)
            throw std::runtime_error("[tinyad_amd] x has " + std::to_string(x.size()) + " entries, expected " +
                                     std::to_string(n_vars));
    }
    SparseMatrix empty_jacobian() const
    {
        if (jac_.outer.empty())
        {
            int64_t nnz = 0;
            check(tad_fn_jacobian_pattern(fn_.get(), &nnz, nullptr, nullptr, nullptr, nullptr));
            jac_.n_rows = n_outputs;
            jac_.n_cols = n_vars;
            jac_.outer.resize((size_t)n_outputs + 1);
            jac_.inner.resize((size_t)(nnz > 0 ? nnz : 1));
            check(tad_fn_jacobian_pattern(fn_.get(), nullptr, nullptr, nullptr, jac_.outer.data(), jac_.inner.data()));
            jac_.inner.resize((size_t)nnz);
        }
        SparseMatrix J = jac_;
        J.values.assign(J.inner.size() > 0 ? J.inner.size() : 1, 0.0);
        J.values.resize(J.inner.size());
        return J;
    }

    mutable SparseMatrix jac_;  // pattern cache (values unused)
    std::shared_ptr<Context> ctx_;
    std::shared_ptr<tad_fn_s> fn_;
};

// vector_function<d>(range) (VectorFunctionImpl.hh:300-325)
template <int variable_dimension, class HandleRange>
auto vector_function(const HandleRange& variable_range, const EvalSettings& settings = {})
{
    using H = std::decay_t<decltype(*std::begin(variable_range))>;
    std::vector<H> handles;
    for (const auto& h : variable_range) handles.push_back(h);
    return VectorFunction<variable_dimension, H>(std::move(handles), settings);
}

// LinearSolver (LinearSolver.hh:13-19): solver options kept across Newton
// iterations (the pattern is fixed per function, so nothing is re-analyzed).
// coarse_refresh = k > 1: newton_direction rebuilds the coarse space of the
// two-level preconditioner on every k-th call and reuses it in between (the
// role the reference's cached symbolic analysis plays across Newton steps).
struct LinearSolver
{
    tad_solve_opts opts = []() { tad_solve_opts o; tad_solve_default_opts(&o); return o; }();
    tad_solve_stats last_stats {};
    int coarse_refresh = 1;
    int reused_in_row = 0;  // calls that reused the coarse space since it was built
    // set once this solver's own call has built the coarse space: the first
    // call of a solver always builds it, whatever an earlier solver on the
    // same function left behind
    bool coarse_built = false;
};

// identity(n) (Helpers.hh:82-89): the n x n sparse identity.
inline SparseMatrix identity(int64_t n)
{
    SparseMatrix I;
    I.n_rows = I.n_cols = n;
    I.outer.resize((size_t)n + 1);
    I.inner.resize((size_t)n);
    I.values.assign((size_t)n, 1.0);
    for (int64_t i = 0; i <= n; ++i) I.outer[i] = (int32_t)i;
    for (int64_t i = 0; i < n; ++i) I.inner[i] = (int32_t)i;
    return I;
}

// newton_direction(g, H_proj, solver, w) (NewtonDirection.hh:25-48):
// solves (w I + H) d = -g; throws if the solve fails.
template <int dd, class H>
VectorXd newton_direction(const ScalarFunction<dd, H>& func, const VectorXd& g, const SparseMatrix& H_proj,
                          LinearSolver& solver, double w_identity = 0.0)
{
    const int64_t n = func.n_vars;
    if ((int64_t)g.size() != n || H_proj.rows() != n || H_proj.cols() != n ||
        (int64_t)H_proj.values.size() != func.nnz())
        throw std::runtime_error("[tinyad_amd] newton_direction: size mismatch");
    VectorXd minus_g(g.size());
    for (size_t i = 0; i < g.size(); ++i) minus_g[i] = -g[i];
    VectorXd dvec(g.size());
    tad_solve_opts o = solver.opts;
    o.w_identity = w_identity;
    if (!solver.coarse_built)
        o.coarse_reuse = 0;
    else if (solver.coarse_refresh > 1 && solver.reused_in_row + 1 < solver.coarse_refresh)
        o.coarse_reuse = 1;
    check(tad_fn_solve(func.handle(), H_proj.values.data(), minus_g.data(), dvec.data(), TAD_MEM_HOST, &o,
                       &solver.last_stats));
    solver.reused_in_row = solver.last_stats.coarse_reused ? solver.reused_in_row + 1 : 0;
    if (solver.last_stats.coarse_unknowns > 0 && !solver.last_stats.coarse_reused) solver.coarse_built = true;
    if (!solver.last_stats.converged) throw std::runtime_error("[tinyad_amd] Linear solve failed.");
    return dvec;
}

// gauss_newton_direction(r, J, solver, w) (GaussNewtonDirection.hh:24-47):
// d with (w I + J^T J) d = -J^T r.  The GPU form takes the function and the
// point x instead of (r, J): J^T J and J^T r are formed on the device from
// the element Jacobians at x (J never leaves the device) and solved by the
// library's PCG as 2 (w I + J^T J) d = -2 J^T r; throws if the solve fails.
template <int dd, class H>
VectorXd gauss_newton_direction(const VectorFunction<dd, H>& func, const VectorXd& x, LinearSolver& solver,
                                double w_identity = 0.0)
{
    const int64_t n = func.n_vars;
    if ((int64_t)x.size() != n) throw std::runtime_error("[tinyad_amd] gauss_newton_direction: size mismatch");
    double f = 0.0;
    VectorXd g((size_t)n), JtJ((size_t)func.nnz()), minus_g((size_t)n), dvec((size_t)n);
    check(tad_fn_eval_with_derivatives(func.handle(), x.data(), &f, g.data(), JtJ.data(), TAD_MEM_HOST));
    for (int64_t i = 0; i < n; ++i) minus_g[i] = -g[i];
    tad_solve_opts o = solver.opts;
    o.w_identity = 2.0 * w_identity;
    if (!solver.coarse_built)
        o.coarse_reuse = 0;
    else if (solver.coarse_refresh > 1 && solver.reused_in_row + 1 < solver.coarse_refresh)
        o.coarse_reuse = 1;
    check(tad_fn_solve(func.handle(), JtJ.data(), minus_g.data(), dvec.data(), TAD_MEM_HOST, &o, &solver.last_stats));
    solver.reused_in_row = solver.last_stats.coarse_reused ? solver.reused_in_row + 1 : 0;
    if (solver.last_stats.coarse_unknowns > 0 && !solver.last_stats.coarse_reused) solver.coarse_built = true;
    if (!solver.last_stats.converged) throw std::runtime_error("[tinyad_amd] Linear solve failed.");
    return dvec;
}

// gauss_newton_direction with a fresh solver (GaussNewtonDirection.hh:57-64).
template <int dd, class H>
VectorXd gauss_newton_direction(const VectorFunction<dd, H>& func, const VectorXd& x, double w_identity = 0.0)
{
    LinearSolver solver;
    return gauss_newton_direction(func, x, solver, w_identity);
}

// newton_direction_reduced_basis(g, H_proj, B, solver, w) (NewtonDirection.hh:76-105):
// d = B (B^T H B + w I)^-1 (-B^T g) for an n x m basis B.  A basis that
// selects variables (every column one entry 1, distinct rows: the pinning
// case) runs the two-level preconditioned solve on the free variables (d is 0
// on the variables B omits); any other basis runs conjugate gradients on the
// reduced system with products by B, H and B^T (tad_fn_solve_basis).
template <int dd, class H>
VectorXd newton_direction_reduced_basis(const ScalarFunction<dd, H>& func, const VectorXd& g,
                                        const SparseMatrix& H_proj, const SparseMatrix& B, LinearSolver& solver,
                                        double w_identity = 0.0)
{
    const int64_t n = (int64_t)g.size();
    if (B.rows() != n || H_proj.rows() != n || H_proj.cols() != n || B.cols() < 1)
        throw std::runtime_error("[tinyad_amd] newton_direction_reduced_basis: size mismatch");
    std::vector<unsigned char> free_vars((size_t)n, 0);
    std::vector<int> per_col((size_t)B.cols(), 0);
    bool selects = true;
    for (int64_t r = 0; r < B.rows() && selects; ++r)
        for (int32_t k = B.outer[r]; k < B.outer[r + 1]; ++k)
        {
            if (B.values[k] == 0.0) continue;
            if (B.values[k] != 1.0 || free_vars[r])
            {
                selects = false;
                break;
            }
            free_vars[r] = 1;
            per_col[B.inner[k]]++;
        }
    for (int c : per_col) selects = selects && c == 1;
    VectorXd minus_g(g.size());
    for (size_t i = 0; i < g.size(); ++i) minus_g[i] = -g[i];
    VectorXd dvec(g.size());
    tad_solve_opts o = solver.opts;
    o.w_identity = w_identity;
    if (selects)
        check(tad_fn_solve_reduced(func.handle(), H_proj.values.data(), minus_g.data(), dvec.data(), TAD_MEM_HOST,
                                   free_vars.data(), &o, &solver.last_stats));
    else
        check(tad_fn_solve_basis(func.handle(), H_proj.values.data(), B.cols(), B.outer.data(), B.inner.data(),
                                 B.values.data(), minus_g.data(), dvec.data(), TAD_MEM_HOST, &o, &solver.last_stats));
    if (!solver.last_stats.converged) throw std::runtime_error("[tinyad_amd] Linear solve failed.");
    return dvec;
}

// newton_direction_reduced_basis(g, H_proj, B, w) with a fresh solver (NewtonDirection.hh:114-123).
template <int dd, class H>
VectorXd newton_direction_reduced_basis(const ScalarFunction<dd, H>& func, const VectorXd& g,
                                        const SparseMatrix& H_proj, const SparseMatrix& B, double w_identity = 0.0)
{
    LinearSolver solver;
    return newton_direction_reduced_basis(func, g, H_proj, B, solver, w_identity);
}

// line_search(x0, d, f, g, eval, s_max, shrink, max_iters, armijo) (LineSearch.hh:26-65),
// evaluated on the host through func.eval (one energy launch per trial).
template <int dd, class H>
VectorXd line_search(const VectorXd& x0, const VectorXd& dvec, double f, const VectorXd& g,
                     const ScalarFunction<dd, H>& func, double s_max = 1.0, double shrink = 0.8, int max_iters = 64,
                     double armijo_const = 1e-4)
{
    if (x0.size() != g.size()) throw std::runtime_error("[tinyad_amd] line_search: size mismatch");
    if (s_max <= 0.0) throw std::runtime_error("[tinyad_amd] Max step size not positive.");
    const bool try_one = s_max > 1.0;
    double dg = 0.0;
    for (size_t i = 0; i < g.size(); ++i) dg += dvec[i] * g[i];
    VectorXd x_new = x0;
    double s = s_max;
    for (int i = 0; i < max_iters; ++i)
    {
        for (size_t j = 0; j < x0.size(); ++j) x_new[j] = x0[j] + s * dvec[j];
        const double f_new = func.eval(x_new);
        if (f_new != f_new) throw std::runtime_error("[tinyad_amd] line_search: NaN energy");
        if (f_new <= f + armijo_const * s * dg) return x_new;
        if (try_one && s > 1.0 && s * shrink < 1.0) s = 1.0;
        else s *= shrink;
    }
    return x0;
}

// line_search(x0, d, f, g, eval, ...) (LineSearch.hh:26-65) with any energy
// callable double(const VectorXd&), e.g. a VectorFunction's
// eval_sum_of_squares as in GaussNewtonTest.cc:146-149.
template <class EvalFn, class = std::enable_if_t<std::is_invocable_r<double, EvalFn, const VectorXd&>::value>>
VectorXd line_search(const VectorXd& x0, const VectorXd& dvec, double f, const VectorXd& g, EvalFn eval,
                     double s_max = 1.0, double shrink = 0.8, int max_iters = 64, double armijo_const = 1e-4)
{
    if (x0.size() != g.size() || x0.size() != dvec.size())
        throw std::runtime_error("[tinyad_amd] line_search: size mismatch");
    if (s_max <= 0.0) throw std::runtime_error("[tinyad_amd] Max step size not positive.");
    const bool try_one = s_max > 1.0;
    double dg = 0.0;
    for (size_t i = 0; i < g.size(); ++i) dg += dvec[i] * g[i];
    VectorXd x_new = x0;
    double s = s_max;
    for (int i = 0; i < max_iters; ++i)
    {
        for (size_t j = 0; j < x0.size(); ++j) x_new[j] = x0[j] + s * dvec[j];
        const double f_new = eval(x_new);
        if (f_new != f_new) throw std::runtime_error("[tinyad_amd] line_search: NaN energy");
        if (f_new <= f + armijo_const * s * dg) return x_new;
        if (try_one && s > 1.0 && s * shrink < 1.0) s = 1.0;
        else s *= shrink;
    }
    return x0;
}

// newton_decrement (NewtonDecrement.hh:20-26)
inline double newton_decrement(const VectorXd& dvec, const VectorXd& g)
{
    double s = 0.0;
    for (size_t i = 0; i < g.size(); ++i) s += dvec[i] * g[i];
    return -0.5 * s;
}

// Whole projected-Newton iteration on the device (eval -> PCG -> batched line search).
template <int dd, class H>
tad_newton_stats newton_step(const ScalarFunction<dd, H>& func, VectorXd& x, const tad_newton_opts* opts = nullptr)
{
    tad_newton_stats st {};
    check(tad_newton_step(func.handle(), x.data(), TAD_MEM_HOST, opts, &st));
    return st;
}

} // namespace tinyad_amd
