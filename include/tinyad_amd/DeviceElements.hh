// Element energies written once as generic device functors, the MI355X form of
// the reference's generic element lambdas (ScalarFunction.hh:72-78,
// tests/NewtonTest.cc:28-44).  Compile the including file with hipcc for
// gfx950 (with -Xclang -target-feature -Xclang -packed-fp32-ops, like the
// library).
//
//   struct MyEnergy
//   {
//       // x: the element's valence * d coordinates, vertex-major (x[a * d + i]),
//       // p: its parameter row; S is double (passive value), tad::Dual<K, T, double>
//       // or tad::HyperDualM<T, M, double> (derivatives)
//       template <class S>
//       __device__ S operator()(const S* x, const double* p) const { ... }
//   };
//   func.add_elements<3>(range(n_faces), tinyad_amd::device_element<MyEnergy, 3, 2, 5>(),
//                        conn_of, params_of);
//
// S is tad::Dual<K, T, double> (one pass) for K = valence * d <= 6 and
// tad::HyperDualM<T, M, double> above (one pass per M = 2 (FP64) or 1 (FP32)
// Hessian entries of a row).  Values are FP64 in both precisions (T is the
// derivative type: double, or float for FP32 contributions), so coordinate
// differences and every value the functor forms keep FP64 accuracy and only
// the derivative parts are FP32; an element far from the origin (|x| >> its
// size) keeps the derivative accuracy of one at the origin.
// Inside the functor use +, -, *, / (with doubles or ints), comparisons and
// tad::sqr, tad::sqrt, tad::exp, tad::log, tad::sin, tad::cos, tad::pow(S, double),
// tad::pow(S, int).
// With add_elements_dynamic the same functor is compiled for several
// valences; an element with fewer handles than its group's valence sees 0 in
// the unused trailing coordinates (pass the handle count in p when the energy
// must know it).
// A functor whose energy depends on the coordinates only through their
// differences may say so with `static constexpr bool translation_invariant =
// true;`: its projected Hessians then take the smaller eigensolve without the
// d translations (for tets 9 x 9 instead of 12 x 12).
// Return S(INFINITY) (or INFINITY) for an infeasible element, which then
// contributes a zero gradient and, projected, eps I (as an infinite Double<k>
// does in the reference).  The functor must be stateless (it is
// default-constructed on the device); per-element data comes in p.
//
// For each energy, device_element<> instantiates the library's element kernel
// (forward mode over the element's K = valence * d variables, one Dual<K, T>
// pass or hyper-dual passes, the dense K x K eigen-projection, scatter through the
// library's slot maps) for FP64 and FP32 contributions and the three
// derivative modes, and the line-search energy kernel, all in the caller's
// code object.  The value returned by every mode is the functor evaluated in
// FP64 with S = double, the same function the line search evaluates, so both
// give the same bits.
#pragma once

#include <tinyad_amd.h>
#include <tinyad_amd/device/HyperDual.hpp>
#include <tinyad_amd/device/Kernels.hpp>

#include <hip/hip_runtime.h>

namespace tinyad_amd {

// F::translation_invariant (optional, default false): the energy depends on
// the element's coordinates only through their differences.  The projected
// Hessian then goes through the eigensolve of its (K - d) x (K - d) part
// without the d translations (as the library's own simplices do: 9 x 9
// instead of 12 x 12 for tets); an element whose Hessian is not translation
// invariant to rounding is reported as non-finite (TAD_ERR_NONFINITE).
template <class F, class = void>
struct declared_translation_invariant : std::false_type
{
};
template <class F>
struct declared_translation_invariant<F, std::void_t<decltype(F::translation_invariant)>>
    : std::bool_constant<F::translation_invariant>
{
};

template <class F, int NN, int DD, int NPAR>
struct DeviceElement
{
    static constexpr int N = NN, D = DD, NP = NPAR, K = NN * DD;
    // elements of dynamic-valence terms end in padding slots (handle -1):
    // their coordinates read as 0 and their derivatives are not stored
    static constexpr bool padded = true;
    // an element may list a handle more than once: its entries are merged
    // into the first occurrence (one set of variables per handle, as the
    // reference's Element::variables gives) before the projection
    static constexpr bool merges = true;
    static_assert(NN >= 1 && DD >= 1 && DD <= 3 && K <= 12, "valence * d must be in 1..12, d in 1..3");
    static_assert(NPAR >= 0, "negative parameter count");

    // n_active: handles before the padding; the padded coordinates are
    // passive constants, so their Hessian rows and columns are zero, as for
    // the local variables a reference element never accesses
    // up to this many element variables one pass with a full Dual<K> per
    // intermediate; above, hyper-dual passes
    static constexpr int kFullDualMaxK = 6;
    // second directions per hyper-dual pass (Hessian entries per evaluation):
    // 2 in FP64 (8.3 instead of 9.5 ms for 1.3M tets; 4: 10.0 ms), 1 in FP32
    // (4.5 ms; 2: 5.4 ms), measured on symmetric Dirichlet tets
    template <class T>
    static constexpr int second_dirs() { return sizeof(T) == 8 ? 2 : 1; }

    // Pass (i, j0) seeds e1 on x_i and e2m on x_{j0+m}: .a = g_i and
    // .ab[m] = H(i, j0+m).  Results enter g and Hp through selects on
    // compile-time indices, so both stay in registers while the pass loops
    // are not unrolled.  The gradient alone needs the passes j0 = i.
    template <class T, bool HESS, int M>
    __device__ static void hyper_passes(const double* X, const double* p, int na, T* g, T* Hp)
    {
#pragma unroll 1
        for (int i = 0; i < na; ++i)
        {
            const int j_end = HESS ? na : i + 1;
#pragma unroll 1
            for (int j0 = i; j0 < j_end; j0 += M)
            {
                tad::HyperDualM<T, M, double> hx[K];
#pragma unroll
                for (int k = 0; k < K; ++k)
                {
                    hx[k] = tad::HyperDualM<T, M, double>(X[k]);
                    hx[k].a = (k == i) ? T(1) : T(0);
#pragma unroll
                    for (int m = 0; m < M; ++m) hx[k].b[m] = (k == j0 + m) ? T(1) : T(0);
                }
                const tad::HyperDualM<T, M, double> r =
                    F{}(static_cast<const tad::HyperDualM<T, M, double>*>(hx), p);
                if (j0 == i)
                {
#pragma unroll
                    for (int k = 0; k < K; ++k) g[k] = (k == i) ? r.a : g[k];
                }
                if constexpr (HESS)
                {
#pragma unroll
                    for (int m = 0; m < M; ++m)
                    {
                        const int j = j0 + m;
                        const int at = j < j_end ? tad::Sym<K>::idx(i, j) : -1;
#pragma unroll
                        for (int s = 0; s < tad::Sym<K>::n; ++s) Hp[s] = (s == at) ? r.ab[m] : Hp[s];
                    }
                }
            }
        }
    }

    template <class T, int MODE, class Sink>
    __device__ static double eval(const double* X, const double* p, T eps, const Sink& out, bool& nonfinite,
                                  int n_active)
    {
        const double f = F{}(X, p);
        const int na = n_active * DD;  // active variables (the rest is padding)
        T g[K], Hp[tad::Sym<K>::n];
        if constexpr (K <= kFullDualMaxK)
        {
            using S = tad::Dual<K, T, double>;
            S v[K];
#pragma unroll
            for (int i = 0; i < K; ++i) v[i] = (i < na) ? tad::dual_var<K, T, double>(X[i], i) : S(X[i]);
            const S E = F{}(static_cast<const S*>(v), p);
#pragma unroll
            for (int i = 0; i < K; ++i) g[i] = E.g[i];
#pragma unroll
            for (int i = 0; i < tad::Sym<K>::n; ++i) Hp[i] = E.h[i];
        }
        else
        {
#pragma unroll
            for (int k = 0; k < K; ++k) g[k] = T(0);
#pragma unroll
            for (int s = 0; s < tad::Sym<K>::n; ++s) Hp[s] = T(0);
            if constexpr (MODE >= tad::MODE_HESS) hyper_passes<T, true, second_dirs<T>()>(X, p, na, g, Hp);
            else hyper_passes<T, false, 1>(X, p, na, g, Hp);
        }
        tad::write_dense_record<NN, DD, T, MODE, declared_translation_invariant<F>::value>(g, Hp, eps, out,
                                                                                          nonfinite);
        return f;
    }

    __device__ static double energy(const double* X, const double* p) { return F{}(X, p); }
};

// Residual elements of vector functions (VectorObjectiveTerm.hh:46-134): the
// functor writes the element's M residuals,
//   template <class S>
//   __device__ void operator()(const S* x, const double* p, S* r) const;
// with S = double (values) or tad::HyperDualM<T, M2, double> (derivatives;
// only the first-order parts .a and .b[] are read).  The element kernel
// evaluates the sum of squares f_e = r_e . r_e (FP64), g_e = 2 J_e^T r_e and
// the Gauss-Newton Hessian 2 J_e^T J_e (positive semi-definite: written
// without projection in modes 2 and 3).  Pass (i, j0) seeds .a on x_i and
// .b[m] on x_{j0+m}, so one evaluation of the M residuals yields column i of
// J_e and columns j0 .. j0+M2-1: (J_e^T J_e)(i, j) = sum_o J_oi J_oj.
template <class F, int NN, int DD, int MM, int NPAR>
struct DeviceResidual
{
    static constexpr int N = NN, D = DD, NP = NPAR, M = MM, K = NN * DD;
    static_assert(NN >= 1 && DD >= 1 && DD <= 3 && K <= 12, "valence * d must be in 1..12, d in 1..3");
    static_assert(MM >= 1 && MM <= 64, "residuals per element must be in 1..64");
    static_assert(NPAR >= 0, "negative parameter count");

    template <class T>
    static constexpr int second_dirs() { return sizeof(T) == 8 ? 2 : 1; }

    template <class T, bool HESS, int M2>
    __device__ static void gn_passes(const double* X, const double* p, T* g, T* Hp)
    {
#pragma unroll 1
        for (int i = 0; i < K; ++i)
        {
            const int j_end = HESS ? K : i + 1;
#pragma unroll 1
            for (int j0 = i; j0 < j_end; j0 += M2)
            {
                tad::HyperDualM<T, M2, double> hx[K];
#pragma unroll
                for (int k = 0; k < K; ++k)
                {
                    hx[k] = tad::HyperDualM<T, M2, double>(X[k]);
                    hx[k].a = (k == i) ? T(1) : T(0);
#pragma unroll
                    for (int m = 0; m < M2; ++m) hx[k].b[m] = (k == j0 + m) ? T(1) : T(0);
                }
                tad::HyperDualM<T, M2, double> r[M];
                F{}(static_cast<const tad::HyperDualM<T, M2, double>*>(hx), p, r);
                if (j0 == i)
                {
                    T gi = T(0);
#pragma unroll
                    for (int o = 0; o < M; ++o) gi += T(r[o].v) * r[o].a;
#pragma unroll
                    for (int k = 0; k < K; ++k) g[k] = (k == i) ? T(2) * gi : g[k];
                }
                if constexpr (HESS)
                {
#pragma unroll
                    for (int m = 0; m < M2; ++m)
                    {
                        const int j = j0 + m;
                        T h = T(0);
#pragma unroll
                        for (int o = 0; o < M; ++o) h += r[o].a * r[o].b[m];
                        const int at = j < j_end ? tad::Sym<K>::idx(i, j) : -1;
#pragma unroll
                        for (int s = 0; s < tad::Sym<K>::n; ++s) Hp[s] = (s == at) ? T(2) * h : Hp[s];
                    }
                }
            }
        }
    }

    template <class T, int MODE, class Sink>
    __device__ static double eval(const double* X, const double* p, T eps, const Sink& out, bool& nonfinite)
    {
        const double f = energy(X, p);
        T g[K], Hp[tad::Sym<K>::n];
#pragma unroll
        for (int k = 0; k < K; ++k) g[k] = T(0);
#pragma unroll
        for (int s = 0; s < tad::Sym<K>::n; ++s) Hp[s] = T(0);
        if constexpr (MODE >= tad::MODE_HESS) gn_passes<T, true, second_dirs<T>()>(X, p, g, Hp);
        else gn_passes<T, false, 1>(X, p, g, Hp);
        // 2 J^T J is positive semi-definite: modes 2 and 3 both write it as is
        constexpr int WMODE = MODE >= tad::MODE_HESS ? tad::MODE_HESS : MODE;
        tad::write_dense_record<NN, DD, T, WMODE>(g, Hp, eps, out, nonfinite);
        return f;
    }

    // r_e . r_e in output order (the reference's res.dot(res),
    // VectorObjectiveTerm.hh:336-338)
    __device__ static double energy(const double* X, const double* p)
    {
        double r[M];
        F{}(X, p, r);
        double f = 0.0;
#pragma unroll
        for (int o = 0; o < M; ++o) f += r[o] * r[o];
        return f;
    }
};

// Residuals and Jacobian rows of residual elements, one lane per element.
// Row (e, o) holds dr_o/dx for the element's K variables in ascending global
// order (handle rank among the element's handles, then coordinate).
template <class E, class F>
__global__ __launch_bounds__(tad::kBlock) void k_residuals(int64_t n, const int32_t* __restrict__ conn,
                                                           const double* __restrict__ params,
                                                           const double* __restrict__ x, double* __restrict__ r,
                                                           double* __restrict__ J)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    constexpr int N = E::N, D = E::D, K = E::K, M = E::M;
    double X[K];
    int32_t h[N];
#pragma unroll
    for (int a = 0; a < N; ++a)
    {
        h[a] = conn[a * n + e];
#pragma unroll
        for (int i = 0; i < D; ++i) X[a * D + i] = x[(int64_t)h[a] * D + i];
    }
    double p[E::NP > 0 ? E::NP : 1];
#pragma unroll
    for (int k = 0; k < E::NP; ++k) p[k] = params[k * n + e];
    double rv[M];
    F{}(static_cast<const double*>(X), p, rv);
#pragma unroll
    for (int o = 0; o < M; ++o) r[e * M + o] = rv[o];
    if (!J) return;
    int rank[N];
#pragma unroll
    for (int a = 0; a < N; ++a)
    {
        int c = 0;
#pragma unroll
        for (int b = 0; b < N; ++b) c += (h[b] < h[a]) ? 1 : 0;
        rank[a] = c;
    }
    double* Je = J + e * (int64_t)(M * K);
#pragma unroll 1
    for (int i = 0; i < K; ++i)
    {
        tad::HyperDualM<double, 1, double> hx[K];
#pragma unroll
        for (int k = 0; k < K; ++k)
        {
            hx[k] = tad::HyperDualM<double, 1, double>(X[k]);
            hx[k].a = (k == i) ? 1.0 : 0.0;
            hx[k].b[0] = 0.0;
        }
        tad::HyperDualM<double, 1, double> ro[M];
        F{}(static_cast<const tad::HyperDualM<double, 1, double>*>(hx), p, ro);
        int ra = 0;
#pragma unroll
        for (int a = 0; a < N; ++a) ra = (a == i / D) ? rank[a] : ra;
        const int col = ra * D + i % D;
#pragma unroll
        for (int o = 0; o < M; ++o) Je[o * K + col] = ro[o].a;
    }
}

namespace detail {

template <class E, class T>
inline hipError_t launch_device_element_t(const tad_device_element_args* a, dim3 grid, hipStream_t s)
{
    const tad::ElemOut<T> o{ static_cast<T*>(a->G), static_cast<T*>(a->Hc), a->gpos, a->hpos };
    switch (a->mode)
    {
        case tad::MODE_GRAD:
            tad::k_element<E, T, tad::MODE_GRAD><<<grid, tad::kBlock, 0, s>>>(a->n_elem, a->conn, a->params, a->x,
                                                                             a->eps, a->elem_f, o, a->err);
            break;
        case tad::MODE_HESS:
            tad::k_element<E, T, tad::MODE_HESS><<<grid, tad::kBlock, 0, s>>>(a->n_elem, a->conn, a->params, a->x,
                                                                             a->eps, a->elem_f, o, a->err);
            break;
        case tad::MODE_HESS_PROJ:
            tad::k_element<E, T, tad::MODE_HESS_PROJ><<<grid, tad::kBlock, 0, s>>>(
                a->n_elem, a->conn, a->params, a->x, a->eps, a->elem_f, o, a->err);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class E>
int launch_device_element(const tad_device_element_args* a)
{
    if (!a || a->n_elem < 0) return (int)hipErrorInvalidValue;
    if (a->n_elem == 0) return (int)hipSuccess;
    const int64_t tiles = (a->n_elem + tad::kBlock - 1) / tad::kBlock;
    if (tiles > 0x7fffffff) return (int)hipErrorInvalidValue;
    const dim3 grid((unsigned)tiles);
    hipStream_t s = static_cast<hipStream_t>(a->stream);
    if (a->precision == 0) return (int)launch_device_element_t<E, double>(a, grid, s);
    if (a->precision == 1) return (int)launch_device_element_t<E, float>(a, grid, s);
    return (int)hipErrorInvalidValue;
}

template <class E>
int launch_device_energy(const tad_device_energy_args* a)
{
    if (!a || a->n_elem < 0 || a->ncand < 1 || a->ncand > tad::kMaxCand) return (int)hipErrorInvalidValue;
    if (a->n_elem == 0) return (int)hipSuccess;
    const int64_t tiles = (a->n_elem + tad::kBlock - 1) / tad::kBlock;
    if (tiles > 0x7fffffff) return (int)hipErrorInvalidValue;
    tad::k_energy<E><<<dim3((unsigned)tiles), tad::kBlock, 0, static_cast<hipStream_t>(a->stream)>>>(
        a->n_elem, a->conn, a->params, a->x, a->dir, a->steps, a->ncand, a->out, a->stride);
    return (int)hipGetLastError();
}

template <class E, class F>
int launch_device_residuals(const tad_device_residual_args* a)
{
    if (!a || a->n_elem < 0 || !a->r) return (int)hipErrorInvalidValue;
    if (a->n_elem == 0) return (int)hipSuccess;
    const int64_t tiles = (a->n_elem + tad::kBlock - 1) / tad::kBlock;
    if (tiles > 0x7fffffff) return (int)hipErrorInvalidValue;
    k_residuals<E, F><<<dim3((unsigned)tiles), tad::kBlock, 0, static_cast<hipStream_t>(a->stream)>>>(
        a->n_elem, a->conn, a->params, a->x, a->r, a->J);
    return (int)hipGetLastError();
}

} // namespace detail

// Launch table of energy F on elements of `valence` handles in dimension d
// with NP parameters per element, for tad_fn_add_elements_device.
template <class F, int valence, int d, int NP>
tad_device_ops device_element()
{
    using E = DeviceElement<F, valence, d, NP>;
    tad_device_ops o{};
    o.version = TAD_DEVICE_OPS_VERSION;
    o.valence = valence;
    o.dim = d;
    o.n_params = NP;
    o.element = &detail::launch_device_element<E>;
    o.energy = &detail::launch_device_energy<E>;
    return o;
}

// Launch table of residual functor F on elements of `valence` handles in
// dimension d with M residuals and NP parameters per element, for
// tad_fn_add_residuals_device.
template <class F, int valence, int d, int M, int NP>
tad_device_residual_ops device_residual()
{
    using E = DeviceResidual<F, valence, d, M, NP>;
    tad_device_residual_ops o{};
    o.version = TAD_DEVICE_OPS_VERSION;
    o.valence = valence;
    o.dim = d;
    o.n_params = NP;
    o.n_outputs = M;
    o.element = &detail::launch_device_element<E>;
    o.energy = &detail::launch_device_energy<E>;
    o.residuals = &detail::launch_device_residuals<E, F>;
    return o;
}

} // namespace tinyad_amd
