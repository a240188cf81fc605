// TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of TinyAD's sparse second-order path, used as the parity
// oracle by tests/, by __graft_entry__.smoke() and by bench.py's CPU baseline
// leg.  Nothing in the shipped library links, loads or calls this file.
//
// It restates, in plain C++17 + OpenMP with no Eigen, the reference semantics:
//   * forward-mode scalar with full k x k Hessian and the reference operator
//     formulas (include/TinyAD/Scalar.hh:62-78 ctors, :220-234 unary minus,
//     :249-264 sqr, :579-736 +/-, :738-821 *, :823-891 /);
//   * per-element energies written the way the reference's tests write them
//     (symmetric Dirichlet: tests/NewtonTest.cc:28-44 and
//     tests/ScalarTestHessianBlock.cc:50-64; position penalty:
//     tests/NewtonTest.cc:48-55; quadratic forms: tests/ScalarFunctionTest.cc:10-151,
//     tests/DynamicElementsTest.cc:100-111);
//   * local variable order = first access = connectivity order
//     (include/TinyAD/Detail/Element.hh:208-260);
//   * per-element projection: positive-diagonal-dominance early out, dense
//     symmetric eigensolve, clamp to eps (eps >= 0) or absolute value
//     (eps < 0), unchanged when nothing was clamped
//     (include/TinyAD/Utils/HessianProjection.hh:16-101);
//   * element-order serial accumulation of f, g and i-major triplets
//     (include/TinyAD/Detail/ScalarObjectiveTerm.hh:224-278), duplicate
//     triplets summed in insertion order (ScalarFunctionImpl.hh:378-399);
//   * passive eval with element-order sum (ScalarObjectiveTerm.hh:162-186)
//     and the early INFINITY between terms (ScalarFunctionImpl.hh:256-273);
//   * Armijo backtracking line search (include/TinyAD/Utils/LineSearch.hh:14-65);
//   * Newton direction on w*I + H (include/TinyAD/Utils/NewtonDirection.hh:25-48);
//     the reference factorizes with SimplicialLDLT; this oracle offers the same
//     preconditioned CG the GPU path runs (for like-for-like CPU timing) and
//     leaves the direct solve to scipy in the tests;
//   * newton_decrement (include/TinyAD/Utils/NewtonDecrement.hh:20-26).
//
// Parity pinning (see tests/test_oracle_*.py):
//   Dual ops          -> reference operator known answers (ScalarTestBinaryOperators.cc,
//                        ScalarTestUnaryOperators.cc) + finite differences.
//   energies          -> closed forms (isometry, uniform scale), translation
//                        invariance, finite differences, torch.func.hessian.
//   projection        -> numpy.linalg.eigh spectral projection, PD-unchanged case,
//                        ScalarFunctionTest.cc:112-151 known answer.
//   assembly          -> scipy coo->csr, nnz formulas (NewtonTest.cc:62-66),
//                        Laplacian known answer (DynamicElementsTest.cc:60-91).
//   line search/Newton-> NewtonTest.cc:12-95 convergence (f = 4, |g| = 0),
//                        scipy.sparse.linalg.spsolve for the direction.
//   PCG               -> scipy spsolve on small SPD systems.
// Bit-level agreement with Eigen's own rounding (eigen solver, determinant
// order, setFromTriplets order) cannot be pinned without Eigen: parity there
// is normwise, not bitwise.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <numeric>
#include <string>
#include <stdexcept>
#include <memory>
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

// ---------------------------------------------------------------------------
// Forward-mode scalar with value, gradient and full K x K Hessian.
// ---------------------------------------------------------------------------
template <int K>
struct Dual
{
    double val = 0.0;
    double grad[K];
    double hess[K][K];

    Dual() { std::memset(grad, 0, sizeof(grad)); std::memset(hess, 0, sizeof(hess)); }
    Dual(double v) : Dual() { val = v; }                 // passive constant (Scalar.hh:62-68)
    Dual(double v, int idx) : Dual() { val = v; grad[idx] = 1.0; } // active variable (Scalar.hh:70-78)
};

template <int K> inline Dual<K> operator-(const Dual<K>& a)
{
    Dual<K> r; r.val = -a.val;
    for (int i = 0; i < K; ++i) r.grad[i] = -a.grad[i];
    for (int i = 0; i < K; ++i) for (int j = 0; j < K; ++j) r.hess[i][j] = -a.hess[i][j];
    return r;
}
template <int K> inline Dual<K> operator+(const Dual<K>& a, const Dual<K>& b)
{
    Dual<K> r; r.val = a.val + b.val;
    for (int i = 0; i < K; ++i) r.grad[i] = a.grad[i] + b.grad[i];
    for (int i = 0; i < K; ++i) for (int j = 0; j < K; ++j) r.hess[i][j] = a.hess[i][j] + b.hess[i][j];
    return r;
}
template <int K> inline Dual<K> operator-(const Dual<K>& a, const Dual<K>& b)
{
    Dual<K> r; r.val = a.val - b.val;
    for (int i = 0; i < K; ++i) r.grad[i] = a.grad[i] - b.grad[i];
    for (int i = 0; i < K; ++i) for (int j = 0; j < K; ++j) r.hess[i][j] = a.hess[i][j] - b.hess[i][j];
    return r;
}
template <int K> inline Dual<K> operator+(const Dual<K>& a, double b) { Dual<K> r = a; r.val += b; return r; }
template <int K> inline Dual<K> operator+(double a, const Dual<K>& b) { Dual<K> r = b; r.val += a; return r; }
template <int K> inline Dual<K> operator-(const Dual<K>& a, double b) { Dual<K> r = a; r.val -= b; return r; }
template <int K> inline Dual<K> operator-(double a, const Dual<K>& b)
{
    Dual<K> r = -b; r.val = a - b.val; return r;
}
// a*b: val ab, grad b ga + a gb, Hess b Ha + ga gb^T + gb ga^T + a Hb  (Scalar.hh:738-765)
template <int K> inline Dual<K> operator*(const Dual<K>& a, const Dual<K>& b)
{
    Dual<K> r; r.val = a.val * b.val;
    for (int i = 0; i < K; ++i) r.grad[i] = b.val * a.grad[i] + a.val * b.grad[i];
    for (int i = 0; i < K; ++i)
        for (int j = 0; j < K; ++j)
            r.hess[i][j] = b.val * a.hess[i][j] + a.grad[i] * b.grad[j] + b.grad[i] * a.grad[j] + a.val * b.hess[i][j];
    return r;
}
template <int K> inline Dual<K> operator*(const Dual<K>& a, double b)
{
    Dual<K> r = a; r.val *= b;
    for (int i = 0; i < K; ++i) r.grad[i] *= b;
    for (int i = 0; i < K; ++i) for (int j = 0; j < K; ++j) r.hess[i][j] *= b;
    return r;
}
template <int K> inline Dual<K> operator*(double a, const Dual<K>& b) { return b * a; }
// a/b: grad (b ga - a gb)/b^2, Hess (Ha - g gb^T - gb g^T - (a/b) Hb)/b  (Scalar.hh:823-846)
template <int K> inline Dual<K> operator/(const Dual<K>& a, const Dual<K>& b)
{
    Dual<K> r; r.val = a.val / b.val;
    for (int i = 0; i < K; ++i) r.grad[i] = (b.val * a.grad[i] - a.val * b.grad[i]) / (b.val * b.val);
    for (int i = 0; i < K; ++i)
        for (int j = 0; j < K; ++j)
            r.hess[i][j] = (a.hess[i][j] - r.grad[i] * b.grad[j] - b.grad[i] * r.grad[j] - r.val * b.hess[i][j]) / b.val;
    return r;
}
template <int K> inline Dual<K> operator/(const Dual<K>& a, double b)
{
    Dual<K> r = a; r.val /= b;
    for (int i = 0; i < K; ++i) r.grad[i] /= b;
    for (int i = 0; i < K; ++i) for (int j = 0; j < K; ++j) r.hess[i][j] /= b;
    return r;
}
// a/b with passive a (Scalar.hh:868-884)
template <int K> inline Dual<K> operator/(double a, const Dual<K>& b)
{
    Dual<K> r; r.val = a / b.val;
    const double c = -a / (b.val * b.val);
    for (int i = 0; i < K; ++i) r.grad[i] = c * b.grad[i];
    for (int i = 0; i < K; ++i)
        for (int j = 0; j < K; ++j)
            r.hess[i][j] = (-r.grad[i] * b.grad[j] - b.grad[i] * r.grad[j] - r.val * b.hess[i][j]) / b.val;
    return r;
}
// sqr: val a^2, grad 2 a ga, Hess 2 (a Ha + ga ga^T)  (Scalar.hh:249-264)
template <int K> inline Dual<K> sqr(const Dual<K>& a)
{
    Dual<K> r; r.val = a.val * a.val;
    for (int i = 0; i < K; ++i) r.grad[i] = 2.0 * a.val * a.grad[i];
    for (int i = 0; i < K; ++i)
        for (int j = 0; j < K; ++j)
            r.hess[i][j] = 2.0 * (a.val * a.hess[i][j] + a.grad[i] * a.grad[j]);
    return r;
}
inline double sqr(double a) { return a * a; }
template <int K> inline double value(const Dual<K>& a) { return a.val; }
inline double value(double a) { return a; }

// ---------------------------------------------------------------------------
// Energy catalog (shared parameter layouts with the HIP library).
// ---------------------------------------------------------------------------
enum Energy
{
    E_SYMDIR_TRI = 1,   // d=2, N=3: params [Mr^-1 (2x2 row-major), w]
    E_SYMDIR_TET = 2,   // d=3, N=4: params [Mr^-1 (3x3 row-major), w]
    E_PENALTY = 3,      // N=1:      params [target (d), w]
    E_QUADRATIC = 4,    // any N, d: params [A (KxK row-major), b (K), c]
};

inline int energy_params(int energy, int d, int N)
{
    const int K = d * N;
    switch (energy)
    {
        case E_SYMDIR_TRI: return 5;
        case E_SYMDIR_TET: return 10;
        case E_PENALTY: return d + 1;
        case E_QUADRATIC: return K * K + K + 1;
        default: return -1;
    }
}

// Symmetric Dirichlet on a triangle mapped to the plane, written as in the
// reference's test lambdas: M = col_mat(b - a, c - a), INFINITY if det(M) <= 0,
// else w * (||J||_F^2 + ||J^-1||_F^2) with J = M * Mr^-1.
template <class S>
S symdir_tri(const S* v, const double* p)
{
    const S M00 = v[2] - v[0], M10 = v[3] - v[1];
    const S M01 = v[4] - v[0], M11 = v[5] - v[1];
    const S detM = M00 * M11 - M01 * M10;
    if (value(detM) <= 0.0)
        return S((double)INFINITY);
    const double R00 = p[0], R01 = p[1], R10 = p[2], R11 = p[3], w = p[4];
    const S J00 = M00 * R00 + M01 * R10, J01 = M00 * R01 + M01 * R11;
    const S J10 = M10 * R00 + M11 * R10, J11 = M10 * R01 + M11 * R11;
    // 2x2 inverse: adjugate scaled by 1/det
    const S detJ = J00 * J11 - J01 * J10;
    const S inv = 1.0 / detJ;
    const S I00 = J11 * inv, I01 = -J01 * inv, I10 = -J10 * inv, I11 = J00 * inv;
    const S nJ = sqr(J00) + sqr(J01) + sqr(J10) + sqr(J11);
    const S nI = sqr(I00) + sqr(I01) + sqr(I10) + sqr(I11);
    return w * (nJ + nI);
}

// Symmetric Dirichlet on a tetrahedron: M = col_mat(b - a, c - a, d - a),
// INFINITY if det(M) <= 0, else w * (||J||^2 + ||J^-1||^2), J = M * Mr^-1.
template <class S>
S symdir_tet(const S* v, const double* p)
{
    S M[3][3];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k)
            M[i][k] = v[3 * (k + 1) + i] - v[i];
    const S detM = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1])
                 - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0])
                 + M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    if (value(detM) <= 0.0)
        return S((double)INFINITY);
    S J[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            J[i][j] = M[i][0] * p[0 * 3 + j] + M[i][1] * p[1 * 3 + j] + M[i][2] * p[2 * 3 + j];
    const double w = p[9];
    // 3x3 inverse via cofactors
    const S C00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const S C01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const S C02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    const S detJ = J[0][0] * C00 + J[0][1] * C01 + J[0][2] * C02;
    const S inv = 1.0 / detJ;
    const S C10 = J[0][2] * J[2][1] - J[0][1] * J[2][2];
    const S C11 = J[0][0] * J[2][2] - J[0][2] * J[2][0];
    const S C12 = J[0][1] * J[2][0] - J[0][0] * J[2][1];
    const S C20 = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    const S C21 = J[0][2] * J[1][0] - J[0][0] * J[1][2];
    const S C22 = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    // J^-1 = C^T / det; its squared Frobenius norm
    const S Ci[9] = { C00, C01, C02, C10, C11, C12, C20, C21, C22 };
    S nI = sqr(Ci[0] * inv);
    for (int t = 1; t < 9; ++t) nI = nI + sqr(Ci[t] * inv);
    S nJ = sqr(J[0][0]);
    for (int t = 1; t < 9; ++t) nJ = nJ + sqr(J[t / 3][t % 3]);
    return w * (nJ + nI);
}

// Quadratic position penalty w * ||target - p||^2 (NewtonTest.cc:48-55 with weight).
template <class S>
S penalty(const S* v, const double* p, int d)
{
    S acc = sqr(p[0] - v[0]);
    for (int i = 1; i < d; ++i) acc = acc + sqr(p[i] - v[i]);
    return p[d] * acc;
}

// Dense quadratic 0.5 x^T A x + b^T x + c over the element's local variables.
template <class S>
S quadratic(const S* v, const double* p, int K)
{
    const double* A = p;
    const double* b = p + K * K;
    const double c = p[K * K + K];
    S acc = S(c);
    for (int i = 0; i < K; ++i)
    {
        if (b[i] != 0.0) acc = acc + b[i] * v[i];
        for (int j = 0; j < K; ++j)
            if (A[i * K + j] != 0.0)
                acc = acc + (0.5 * A[i * K + j]) * (v[i] * v[j]);
    }
    return acc;
}

template <class S>
S eval_energy(int energy, const S* v, const double* p, int d, int N)
{
    switch (energy)
    {
        case E_SYMDIR_TRI: return symdir_tri<S>(v, p);
        case E_SYMDIR_TET: return symdir_tet<S>(v, p);
        case E_PENALTY: return penalty<S>(v, p, d);
        case E_QUADRATIC: return quadratic<S>(v, p, d * N);
        default: throw std::runtime_error("unknown energy");
    }
}

// ---------------------------------------------------------------------------
// Projection (HessianProjection.hh:23-101), dense symmetric eigensolve by
// cyclic Jacobi (Eigen's SelfAdjointEigenSolver is a different algorithm for
// the same spectral decomposition).
// ---------------------------------------------------------------------------
inline bool positive_diagonally_dominant(const double* H, int k, double eps)
{
    for (int i = 0; i < k; ++i)
    {
        double off = 0.0;
        for (int j = 0; j < k; ++j)
            if (i != j) off += std::fabs(H[i * k + j]);
        if (H[i * k + i] < off + eps)
            return false;
    }
    return true;
}

// Cyclic Jacobi with Rutishauser's termination (Rutishauser 1971; the form in
// Press et al., Numerical Recipes, sec. 11.1): the rotated pair is set to an
// exact zero, the first three sweeps skip entries below a fifth of the mean
// off-diagonal magnitude, later sweeps zero an entry that no longer changes
// either diagonal entry it couples, and the iteration ends when every
// off-diagonal entry is exactly zero.
inline void jacobi_eigen(const double* H, int k, double* lambda, double* V)
{
    std::vector<double> A(H, H + k * k);
    for (int i = 0; i < k * k; ++i) V[i] = 0.0;
    for (int i = 0; i < k; ++i) V[i * k + i] = 1.0;
    for (int sweep = 1; sweep <= 64; ++sweep)
    {
        double sm = 0.0;
        for (int p = 0; p < k - 1; ++p)
            for (int q = p + 1; q < k; ++q) sm += std::fabs(A[p * k + q]);
        if (sm == 0.0)
            break;
        const double tresh = sweep < 4 ? 0.2 * sm / (k * k) : 0.0;
        for (int p = 0; p < k - 1; ++p)
            for (int q = p + 1; q < k; ++q)
            {
                const double apq = A[p * k + q];
                const double app = A[p * k + p], aqq = A[q * k + q];
                const double g = 100.0 * std::fabs(apq);
                if (sweep > 4 && std::fabs(app) + g == std::fabs(app) && std::fabs(aqq) + g == std::fabs(aqq))
                {
                    A[p * k + q] = A[q * k + p] = 0.0;
                    continue;
                }
                if (std::fabs(apq) <= tresh || apq == 0.0)
                    continue;
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int r = 0; r < k; ++r)
                {
                    const double arp = A[r * k + p], arq = A[r * k + q];
                    A[r * k + p] = c * arp - s * arq;
                    A[r * k + q] = s * arp + c * arq;
                }
                for (int r = 0; r < k; ++r)
                {
                    const double apr = A[p * k + r], aqr = A[q * k + r];
                    A[p * k + r] = c * apr - s * aqr;
                    A[q * k + r] = s * apr + c * aqr;
                }
                A[p * k + q] = A[q * k + p] = 0.0;  // annihilated in exact arithmetic
                for (int r = 0; r < k; ++r)
                {
                    const double vrp = V[r * k + p], vrq = V[r * k + q];
                    V[r * k + p] = c * vrp - s * vrq;
                    V[r * k + q] = s * vrp + c * vrq;
                }
            }
    }
    for (int i = 0; i < k; ++i) lambda[i] = A[i * k + i];
}

inline void project_positive_definite(double* H, int k, double eps)
{
    if (k == 0) return;
    if (positive_diagonally_dominant(H, k, eps))
        return;
    std::vector<double> lam(k), V(k * k);
    jacobi_eigen(H, k, lam.data(), V.data());
    bool all_positive = true;
    for (int i = 0; i < k; ++i)
    {
        if (eps < 0.0)
        {
            if (lam[i] < 0.0) { lam[i] = -lam[i]; all_positive = false; }
        }
        else if (lam[i] < eps)
        {
            lam[i] = eps; all_positive = false;
        }
    }
    if (all_positive)
        return;
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j)
        {
            double acc = 0.0;
            for (int m = 0; m < k; ++m) acc += V[i * k + m] * lam[m] * V[j * k + m];
            H[i * k + j] = acc;
        }
}

// ---------------------------------------------------------------------------
// Objective function (ScalarFunction restated).
// ---------------------------------------------------------------------------
struct Term
{
    int energy = 0, N = 0, K = 0, n_params = 0;
    long n_elem = 0;
    std::vector<int> conn;      // n_elem x N
    std::vector<double> params; // n_elem x n_params
};

struct Function
{
    int d = 0;
    long n_handles = 0, n_vars = 0;
    int n_threads = 0; // 0: reference default max(1, omp_get_max_threads() - 1) (Parallel.hh:20-31)
    std::vector<Term> terms;

    // Output of the last second-order evaluation.
    std::vector<int> row_ptr, col_idx;
    std::vector<double> values;

    int threads() const
    {
#ifdef _OPENMP
        if (n_threads >= 1) return n_threads;
        return std::max(1, omp_get_max_threads() - 1);
#else
        return 1;
#endif
    }
};

struct ElementResult
{
    double val;
    double grad[12];
    double hess[144];
    int idx[12];
    int n_loc;  // local variables in use (idx_local_to_global.size())
};

// Element::variables (Element.hh:208-260): the element's handles in access
// (= connectivity) order; a handle met again reuses the local variables of
// its first access, so its derivatives add up there, and local variables
// past the last new handle stay unused (zero rows and columns of the k x k
// Hessian, which the projection sees, ScalarObjectiveTerm.hh:242-249; only
// the used ones are scattered, :259-276).
template <int K>
void element_second_order(const Function& fn, const Term& t, long e, const double* x,
                          bool project, double eps, ElementResult& out)
{
    const int d = fn.d;
    Dual<K> v[K];
    int n_loc = 0;
    for (int a = 0; a < t.N; ++a)
    {
        const long g0 = (long)d * t.conn[e * t.N + a];
        int start = -1;  // local variables already given to this handle
        for (int c = 0; c < n_loc && start < 0; c += d)
            if (out.idx[c] == g0) start = c;
        if (start < 0)
        {
            start = n_loc;
            for (int i = 0; i < d; ++i) out.idx[n_loc + i] = (int)(g0 + i);
            n_loc += d;
        }
        for (int i = 0; i < d; ++i) v[a * d + i] = Dual<K>(x[g0 + i], start + i);
    }
    out.n_loc = n_loc;
    for (int i = n_loc; i < K; ++i) out.idx[i] = -1;  // unused local variables
    const Dual<K> r = eval_energy<Dual<K>>(t.energy, v, &t.params[e * t.n_params], std::min(d, K), t.N);
    out.val = r.val;
    for (int i = 0; i < K; ++i) out.grad[i] = r.grad[i];
    for (int i = 0; i < K; ++i) for (int j = 0; j < K; ++j) out.hess[i * K + j] = r.hess[i][j];
    if (project)
        project_positive_definite(out.hess, K, eps);
}

typedef void (*ElemFn)(const Function&, const Term&, long, const double*, bool, double, ElementResult&);

inline ElemFn element_fn(int K)
{
    switch (K)
    {
        case 1: return &element_second_order<1>;
        case 2: return &element_second_order<2>;
        case 3: return &element_second_order<3>;
        case 4: return &element_second_order<4>;
        case 5: return &element_second_order<5>;
        case 6: return &element_second_order<6>;
        case 7: return &element_second_order<7>;
        case 8: return &element_second_order<8>;
        case 9: return &element_second_order<9>;
        case 10: return &element_second_order<10>;
        case 11: return &element_second_order<11>;
        case 12: return &element_second_order<12>;
        default: throw std::runtime_error("element size K must be in 1..12");
    }
}

inline double element_energy(const Function& fn, const Term& t, long e, const double* x)
{
    double v[12];
    for (int a = 0; a < t.N; ++a)
        for (int i = 0; i < fn.d; ++i)
            v[a * fn.d + i] = x[(long)fn.d * t.conn[e * t.N + a] + i];
    return eval_energy<double>(t.energy, v, &t.params[e * t.n_params], fn.d, t.N);
}

// Passive eval (ScalarFunctionImpl.hh:256-273, ScalarObjectiveTerm.hh:162-186).
inline double eval(const Function& fn, const double* x)
{
    double f = 0.0;
    for (const Term& t : fn.terms)
    {
        if (f == (double)INFINITY)
            return (double)INFINITY;
        std::vector<double> res(t.n_elem);
        #pragma omp parallel for schedule(static) num_threads(fn.threads())
        for (long e = 0; e < t.n_elem; ++e)
            res[e] = element_energy(fn, t, e, x);
        double s = 0.0;
        for (long e = 0; e < t.n_elem; ++e) s += res[e];
        f += s;
    }
    return f;
}

struct Triplet { int r, c; double v; };

// Second-order eval: f, g, and the triplet list in reference order.
inline void eval_second_order(const Function& fn, const double* x, bool project, double eps,
                              double& f, double* g, std::vector<Triplet>& trip, bool want_hessian)
{
    f = 0.0;
    for (long i = 0; i < fn.n_vars; ++i) g[i] = 0.0;
    trip.clear();
    for (const Term& t : fn.terms)
    {
        std::vector<ElementResult> res(t.n_elem);
        ElemFn ef = element_fn(t.K);
        bool nonfinite = false;
        #pragma omp parallel for schedule(static) num_threads(fn.threads()) reduction(||:nonfinite)
        for (long e = 0; e < t.n_elem; ++e)
        {
            ef(fn, t, e, x, project, eps, res[e]);
            for (int i = 0; i < t.K; ++i) if (!std::isfinite(res[e].grad[i])) nonfinite = true;
            for (int i = 0; i < t.K * t.K; ++i) if (!std::isfinite(res[e].hess[i])) nonfinite = true;
        }
        if (nonfinite)
            throw std::runtime_error("non-finite element gradient or Hessian");
        for (long e = 0; e < t.n_elem; ++e)
        {
            const ElementResult& r = res[e];
            f += r.val;
            for (int i = 0; i < r.n_loc; ++i) g[r.idx[i]] += r.grad[i];
            if (want_hessian)
                for (int i = 0; i < r.n_loc; ++i)
                    for (int j = 0; j < r.n_loc; ++j)
                        trip.push_back({ r.idx[i], r.idx[j], r.hess[i * t.K + j] });
        }
    }
}

// Triplets -> CSR with duplicates summed in insertion order.
// Linear-time triplet -> CSR with duplicates summed in triplet order (the
// semantics of Eigen's setFromTriplets, ScalarFunctionImpl.hh:398): bucket the
// triplets per row in input order, merge equal columns of a row through a
// column marker, then sort each row's columns.
inline void triplets_to_csr(long n, const std::vector<Triplet>& trip,
                            std::vector<int>& row_ptr, std::vector<int>& col_idx, std::vector<double>& values)
{
    std::vector<long> rstart(n + 1, 0);
    for (const Triplet& t : trip) rstart[t.r + 1]++;
    for (long r = 0; r < n; ++r) rstart[r + 1] += rstart[r];
    std::vector<long> order(trip.size());
    {
        std::vector<long> fill(rstart.begin(), rstart.end() - 1);
        for (long i = 0; i < (long)trip.size(); ++i) order[fill[trip[i].r]++] = i;
    }
    row_ptr.assign(n + 1, 0);
    col_idx.clear();
    values.clear();
    std::vector<long> marker(n, -1);
    std::vector<std::pair<int, double>> row;
    for (long r = 0; r < n; ++r)
    {
        row.clear();
        for (long i = rstart[r]; i < rstart[r + 1]; ++i)
        {
            const Triplet& t = trip[order[i]];
            if (marker[t.c] < 0)
            {
                marker[t.c] = (long)row.size();
                row.emplace_back(t.c, t.v);
            }
            else
                row[marker[t.c]].second += t.v;
        }
        for (const auto& e : row) marker[e.first] = -1;
        std::sort(row.begin(), row.end(), [](const std::pair<int, double>& a, const std::pair<int, double>& b) {
            return a.first < b.first;
        });
        for (const auto& e : row)
        {
            col_idx.push_back(e.first);
            values.push_back(e.second);
        }
        row_ptr[r + 1] = (int)col_idx.size();
    }
}

// ---------------------------------------------------------------------------
// Solver utilities.
// ---------------------------------------------------------------------------
inline void spmv(long n, const int* rp, const int* ci, const double* v, const double* x, double* y)
{
    #pragma omp parallel for schedule(static)
    for (long r = 0; r < n; ++r)
    {
        double s = 0.0;
        for (int k = rp[r]; k < rp[r + 1]; ++k) s += v[k] * x[ci[k]];
        y[r] = s;
    }
}

inline double dot(long n, const double* a, const double* b)
{
    double s = 0.0;
    for (long i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

// The same sum taken in chunks of 8192 terms (each chunk in order, the chunk
// sums in order): the result does not depend on the thread count.
inline double dot_chunked(long n, const double* a, const double* b)
{
    constexpr long kChunk = 8192;
    const long nc = (n + kChunk - 1) / kChunk;
    std::vector<double> part(nc);
    #pragma omp parallel for schedule(static)
    for (long c = 0; c < nc; ++c)
    {
        double s = 0.0;
        const long e = std::min(n, (c + 1) * kChunk);
        for (long i = c * kChunk; i < e; ++i) s += a[i] * b[i];
        part[c] = s;
    }
    double s = 0.0;
    for (long c = 0; c < nc; ++c) s += part[c];
    return s;
}

// Invert the bs x bs diagonal blocks of (w I + A) (bs = 1, 2 or 3).
inline void block_jacobi_inverse(long n, int bs, const int* rp, const int* ci, const double* v,
                                 double w, std::vector<double>& Dinv)
{
    const long nb = n / bs;
    Dinv.assign(nb * bs * bs, 0.0);
    for (long b = 0; b < nb; ++b)
    {
        double B[9] = { 0 };
        for (int i = 0; i < bs; ++i)
        {
            const long r = b * bs + i;
            for (int k = rp[r]; k < rp[r + 1]; ++k)
            {
                const long c = ci[k];
                if (c / bs == b) B[i * bs + (c - b * bs)] += v[k];
            }
            B[i * bs + i] += w;
        }
        double* o = &Dinv[b * bs * bs];
        if (bs == 1) o[0] = 1.0 / B[0];
        else if (bs == 2)
        {
            const double det = B[0] * B[3] - B[1] * B[2];
            o[0] = B[3] / det; o[1] = -B[1] / det; o[2] = -B[2] / det; o[3] = B[0] / det;
        }
        else
        {
            const double c00 = B[4] * B[8] - B[5] * B[7];
            const double c01 = B[5] * B[6] - B[3] * B[8];
            const double c02 = B[3] * B[7] - B[4] * B[6];
            const double det = B[0] * c00 + B[1] * c01 + B[2] * c02;
            o[0] = c00 / det; o[3] = c01 / det; o[6] = c02 / det;
            o[1] = (B[2] * B[7] - B[1] * B[8]) / det;
            o[4] = (B[0] * B[8] - B[2] * B[6]) / det;
            o[7] = (B[1] * B[6] - B[0] * B[7]) / det;
            o[2] = (B[1] * B[5] - B[2] * B[4]) / det;
            o[5] = (B[2] * B[3] - B[0] * B[5]) / det;
            o[8] = (B[0] * B[4] - B[1] * B[3]) / det;
        }
    }
}

// Preconditioned CG on (w I + A) x = b, x0 = 0, block-Jacobi preconditioner,
// stop when ||r||_2 <= tol ||b||_2.  Returns iterations, -1 on breakdown.
inline int pcg(long n, int bs, const int* rp, const int* ci, const double* v, double w,
               const double* b, double* x, double tol, int max_iters, double* rel_res)
{
    std::vector<double> Dinv;
    block_jacobi_inverse(n, bs, rp, ci, v, w, Dinv);
    std::vector<double> r(b, b + n), z(n), p(n), q(n);
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) x[i] = 0.0;
    auto apply_prec = [&](const double* in, double* out) {
        #pragma omp parallel for schedule(static)
        for (long blk = 0; blk < n / bs; ++blk)
            for (int i = 0; i < bs; ++i)
            {
                double s = 0.0;
                for (int j = 0; j < bs; ++j) s += Dinv[blk * bs * bs + i * bs + j] * in[blk * bs + j];
                out[blk * bs + i] = s;
            }
    };
    const double bnorm = std::sqrt(dot_chunked(n, b, b));
    if (bnorm == 0.0) { if (rel_res) *rel_res = 0.0; return 0; }
    apply_prec(r.data(), z.data());
    p = z;
    double rho = dot_chunked(n, r.data(), z.data());
    int it = 0;
    double rr = bnorm * bnorm;
    while (it < max_iters && std::sqrt(rr) > tol * bnorm)
    {
        spmv(n, rp, ci, v, p.data(), q.data());
        if (w != 0.0)
        {
            #pragma omp parallel for schedule(static)
            for (long i = 0; i < n; ++i) q[i] += w * p[i];
        }
        const double pq = dot_chunked(n, p.data(), q.data());
        if (!(pq > 0.0)) return -1;
        const double alpha = rho / pq;
        #pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
        apply_prec(r.data(), z.data());
        const double rho_new = dot_chunked(n, r.data(), z.data());
        rr = dot_chunked(n, r.data(), r.data());
        const double beta = rho_new / rho;
        rho = rho_new;
        #pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        ++it;
    }
    if (rel_res) *rel_res = std::sqrt(rr) / bnorm;
    return it;
}

// ---------------------------------------------------------------------------
// Vector functions (VectorFunction.hh:19-192): elements produce fixed-size
// segments of a residual vector r(x).  Residuals are written as the reference's
// GaussNewtonTest writes its element lambdas (tests/GaussNewtonTest.cc:30-74):
//   R_SYMDIR_TRI (d=2, N=3, 8 outputs): params [Mr^-1 (2x2 row-major), s]:
//     s * (J00, J01, J10, J11, Ji00, Ji01, Ji10, Ji11), J = M Mr^-1,
//     Ji = Mr M^-1 (Mr formed from Mr^-1 by the 2x2 inverse), every output
//     INFINITY if det(M) <= 0;
//   R_PIN (N=1, d outputs): params [target (d), s]: s * (target - x).
// The Jacobian follows VectorObjectiveTerm.hh:188-246: per residual, one
// triplet per local variable in local (connectivity) order, rows numbered in
// term then element then output order; the sparse matrix sums duplicates.
// ---------------------------------------------------------------------------
enum Residual
{
    R_SYMDIR_TRI = 101,
    R_PIN = 102,
};

inline int residual_outputs(int rid, int d, int N)
{
    if (rid == R_SYMDIR_TRI) return (d == 2 && N == 3) ? 8 : -1;
    if (rid == R_PIN) return N == 1 ? d : -1;
    return -1;
}

inline int residual_params(int rid, int d)
{
    if (rid == R_SYMDIR_TRI) return 5;
    if (rid == R_PIN) return d + 1;
    return -1;
}

template <class S>
void residual_symdir_tri(const S* v, const double* p, S* r)
{
    const S M00 = v[2] - v[0], M10 = v[3] - v[1];
    const S M01 = v[4] - v[0], M11 = v[5] - v[1];
    const S detM = M00 * M11 - M01 * M10;
    if (value(detM) <= 0.0)
    {
        for (int o = 0; o < 8; ++o) r[o] = S((double)INFINITY);
        return;
    }
    const double R00 = p[0], R01 = p[1], R10 = p[2], R11 = p[3], s = p[4];
    // Mr = (Mr^-1)^-1
    const double dR = R00 * R11 - R01 * R10;
    const double Q00 = R11 / dR, Q01 = -R01 / dR, Q10 = -R10 / dR, Q11 = R00 / dR;
    const S J00 = M00 * R00 + M01 * R10, J01 = M00 * R01 + M01 * R11;
    const S J10 = M10 * R00 + M11 * R10, J11 = M10 * R01 + M11 * R11;
    // M^-1 = adj(M) / det(M); Ji = Mr M^-1
    const S inv = 1.0 / detM;
    const S N00 = M11 * inv, N01 = -M01 * inv, N10 = -M10 * inv, N11 = M00 * inv;
    const S I00 = Q00 * N00 + Q01 * N10, I01 = Q00 * N01 + Q01 * N11;
    const S I10 = Q10 * N00 + Q11 * N10, I11 = Q10 * N01 + Q11 * N11;
    r[0] = s * J00; r[1] = s * J01; r[2] = s * J10; r[3] = s * J11;
    r[4] = s * I00; r[5] = s * I01; r[6] = s * I10; r[7] = s * I11;
}

template <class S>
void residual_pin(const S* v, const double* p, int d, S* r)
{
    for (int i = 0; i < d; ++i) r[i] = p[d] * (p[i] - v[i]);
}

template <class S>
void eval_residual(int rid, const S* v, const double* p, int d, S* r)
{
    switch (rid)
    {
        case R_SYMDIR_TRI: residual_symdir_tri<S>(v, p, r); break;
        case R_PIN: residual_pin<S>(v, p, d, r); break;
        default: throw std::runtime_error("unknown residual");
    }
}

struct RTerm
{
    int rid = 0, N = 0, K = 0, M = 0, n_params = 0;
    long n_elem = 0;
    std::vector<int> conn;
    std::vector<double> params;
};

struct VFunction
{
    int d = 0;
    long n_handles = 0, n_vars = 0, n_outputs = 0;
    std::vector<RTerm> terms;
    std::vector<int> row_ptr, col_idx;  // last Jacobian (CSR)
    std::vector<double> values;
};

// r (VectorObjectiveTerm.hh:162-183): element segments in term order.
inline void veval(const VFunction& fn, const double* x, double* r)
{
    long base = 0;
    for (const RTerm& t : fn.terms)
    {
        #pragma omp parallel for schedule(static)
        for (long e = 0; e < t.n_elem; ++e)
        {
            double v[12];
            for (int a = 0; a < t.N; ++a)
                for (int i = 0; i < fn.d; ++i) v[a * fn.d + i] = x[(long)fn.d * t.conn[e * t.N + a] + i];
            eval_residual<double>(t.rid, v, &t.params[e * t.n_params], fn.d, r + base + e * t.M);
        }
        base += t.n_elem * t.M;
    }
}

// sum of squares (VectorObjectiveTerm.hh:326-348, VectorFunctionImpl.hh:228-239):
// per element res.dot(res), summed in element order, terms in order.
inline double vsum_of_squares(const VFunction& fn, const double* x)
{
    double f = 0.0;
    for (const RTerm& t : fn.terms)
    {
        std::vector<double> sq(t.n_elem);
        #pragma omp parallel for schedule(static)
        for (long e = 0; e < t.n_elem; ++e)
        {
            double v[12], r[32];
            for (int a = 0; a < t.N; ++a)
                for (int i = 0; i < fn.d; ++i) v[a * fn.d + i] = x[(long)fn.d * t.conn[e * t.N + a] + i];
            eval_residual<double>(t.rid, v, &t.params[e * t.n_params], fn.d, r);
            double q = 0.0;
            for (int o = 0; o < t.M; ++o) q += r[o] * r[o];
            sq[e] = q;
        }
        double s = 0.0;
        for (long e = 0; e < t.n_elem; ++e) s += sq[e];
        f += s;
    }
    return f;
}

template <int K>
void residual_first_order(const VFunction& fn, const RTerm& t, long e, const double* x, double* rv, double* jac,
                          int* idx)
{
    Dual<K> v[K];
    for (int a = 0; a < t.N; ++a)
        for (int i = 0; i < fn.d; ++i)
        {
            const long gi = (long)fn.d * t.conn[e * t.N + a] + i;
            v[a * fn.d + i] = Dual<K>(x[gi], a * fn.d + i);
            idx[a * fn.d + i] = (int)gi;
        }
    std::vector<Dual<K>> r(t.M);
    // d <= K always (K = d * N); the bound lets the compiler see it
    eval_residual<Dual<K>>(t.rid, v, &t.params[e * t.n_params], std::min(fn.d, K), r.data());
    for (int o = 0; o < t.M; ++o)
    {
        rv[o] = r[o].val;
        for (int i = 0; i < K; ++i) jac[o * K + i] = r[o].grad[i];
    }
}

// r and the Jacobian triplets (VectorObjectiveTerm.hh:188-246).
inline void veval_with_jacobian(const VFunction& fn, const double* x, double* r, std::vector<Triplet>& trip)
{
    trip.clear();
    long base = 0;
    for (const RTerm& t : fn.terms)
    {
        std::vector<double> jac((size_t)t.n_elem * t.M * t.K);
        std::vector<int> idx((size_t)t.n_elem * t.K);
        bool nonfinite = false;
        #pragma omp parallel for schedule(static) reduction(||:nonfinite)
        for (long e = 0; e < t.n_elem; ++e)
        {
            double* jr = &jac[(size_t)e * t.M * t.K];
            int* ix = &idx[(size_t)e * t.K];
            switch (t.K)
            {
                case 1: residual_first_order<1>(fn, t, e, x, r + base + e * t.M, jr, ix); break;
                case 2: residual_first_order<2>(fn, t, e, x, r + base + e * t.M, jr, ix); break;
                case 3: residual_first_order<3>(fn, t, e, x, r + base + e * t.M, jr, ix); break;
                case 6: residual_first_order<6>(fn, t, e, x, r + base + e * t.M, jr, ix); break;
                default: break;
            }
            for (int i = 0; i < t.M * t.K; ++i) if (!std::isfinite(jr[i])) nonfinite = true;
        }
        if (nonfinite) throw std::runtime_error("non-finite residual gradient");
        for (long e = 0; e < t.n_elem; ++e)
            for (int o = 0; o < t.M; ++o)
                for (int i = 0; i < t.K; ++i)
                    trip.push_back({ (int)(base + e * t.M + o), idx[(size_t)e * t.K + i],
                                     jac[((size_t)e * t.M + o) * t.K + i] });
        base += t.n_elem * t.M;
    }
}

} // namespace oracle

// ---------------------------------------------------------------------------
// C API for ctypes (tests and the CPU baseline).
// ---------------------------------------------------------------------------
using namespace oracle;

static thread_local std::string g_err;

extern "C" {

const char* or_last_error() { return g_err.c_str(); }

void* or_fn_create(int d, long n_handles)
{
    if (d < 1 || d > 3 || n_handles < 1) { g_err = "invalid dimension or handle count"; return nullptr; }
    Function* f = new Function();
    f->d = d; f->n_handles = n_handles; f->n_vars = (long)d * n_handles;
    return f;
}

void or_fn_destroy(void* fn) { delete (Function*)fn; }

void or_fn_set_threads(void* fn, int n) { ((Function*)fn)->n_threads = n; }

int or_fn_threads(void* fn) { return ((Function*)fn)->threads(); }

int or_fn_add_elements(void* fnp, int energy, int N, long n_elem, const int* conn,
                       const double* params, int n_params)
{
    Function& fn = *(Function*)fnp;
    const int K = fn.d * N;
    if (K < 1 || K > 12) { g_err = "element size d*N must be in 1..12"; return 1; }
    if (energy_params(energy, fn.d, N) != n_params) { g_err = "parameter count mismatch"; return 1; }
    if ((energy == E_SYMDIR_TRI && (fn.d != 2 || N != 3)) || (energy == E_SYMDIR_TET && (fn.d != 3 || N != 4))
        || (energy == E_PENALTY && N != 1))
    { g_err = "energy does not match dimension/valence"; return 1; }
    Term t;
    t.energy = energy; t.N = N; t.K = K; t.n_params = n_params; t.n_elem = n_elem;
    t.conn.assign(conn, conn + n_elem * N);
    t.params.assign(params, params + n_elem * n_params);
    for (long i = 0; i < n_elem * N; ++i)
        if (t.conn[i] < 0 || t.conn[i] >= fn.n_handles) { g_err = "handle out of range"; return 1; }
    fn.terms.push_back(std::move(t));
    return 0;
}

double or_eval(void* fnp, const double* x) { return eval(*(Function*)fnp, x); }

int or_eval_with_gradient(void* fnp, const double* x, double* f, double* g)
{
    try
    {
        std::vector<Triplet> trip;
        eval_second_order(*(Function*)fnp, x, false, 0.0, *f, g, trip, false);
        return 0;
    }
    catch (const std::exception& ex) { g_err = ex.what(); return 3; }
}

// project = 0: eval_with_derivatives; 1: eval_with_hessian_proj(eps)
int or_eval_second_order(void* fnp, const double* x, int project, double eps, double* f, double* g, long* nnz)
{
    Function& fn = *(Function*)fnp;
    try
    {
        std::vector<Triplet> trip;
        eval_second_order(fn, x, project != 0, eps, *f, g, trip, true);
        triplets_to_csr(fn.n_vars, trip, fn.row_ptr, fn.col_idx, fn.values);
        *nnz = (long)fn.values.size();
        return 0;
    }
    catch (const std::exception& ex) { g_err = ex.what(); return 3; }
}

void or_get_csr(void* fnp, int* row_ptr, int* col_idx, double* values)
{
    Function& fn = *(Function*)fnp;
    std::copy(fn.row_ptr.begin(), fn.row_ptr.end(), row_ptr);
    std::copy(fn.col_idx.begin(), fn.col_idx.end(), col_idx);
    std::copy(fn.values.begin(), fn.values.end(), values);
}

// Dense per-element result: val, grad[K], hess[K*K], global idx[K] (-1 for
// local variables the element leaves unused).
int or_element(void* fnp, int term, long e, const double* x, int project, double eps,
               double* val, double* grad, double* hess, int* idx)
{
    Function& fn = *(Function*)fnp;
    if (term < 0 || term >= (int)fn.terms.size()) { g_err = "bad term"; return 1; }
    const Term& t = fn.terms[term];
    if (e < 0 || e >= t.n_elem) { g_err = "bad element"; return 1; }
    ElementResult r;
    element_fn(t.K)(fn, t, e, x, project != 0, eps, r);
    *val = r.val;
    std::copy(r.grad, r.grad + t.K, grad);
    std::copy(r.hess, r.hess + t.K * t.K, hess);
    std::copy(r.idx, r.idx + t.K, idx);
    return 0;
}

// Time the reference-style eval_with_hessian_proj; returns seconds per call
// (median of reps after one warm-up).
double or_time_eval_with_hessian_proj(void* fnp, const double* x, double eps, int reps)
{
    Function& fn = *(Function*)fnp;
    std::vector<double> g(fn.n_vars);
    std::vector<double> ts;
    for (int r = 0; r < reps + 1; ++r)
    {
        auto t0 = std::chrono::steady_clock::now();
        double f;
        std::vector<Triplet> trip;
        eval_second_order(fn, x, true, eps, f, g.data(), trip, true);
        triplets_to_csr(fn.n_vars, trip, fn.row_ptr, fn.col_idx, fn.values);
        auto t1 = std::chrono::steady_clock::now();
        if (r > 0) ts.push_back(std::chrono::duration<double>(t1 - t0).count());
    }
    std::sort(ts.begin(), ts.end());
    return ts[ts.size() / 2];
}

// Block-Jacobi PCG on (w I + A) x = b.  n_threads: OpenMP threads of its
// products and vector updates (0: the reference's default, max(1,
// omp_get_max_threads() - 1), Parallel.hh:20-31); the iterates do not depend
// on it (chunked, ordered dot products).
int or_pcg(long n, int bs, const int* rp, const int* ci, const double* v, double w,
           const double* b, double* x, double tol, int max_iters, double* rel_res, int n_threads)
{
#ifdef _OPENMP
    const int prev = omp_get_max_threads();
    omp_set_num_threads(n_threads >= 1 ? n_threads : std::max(1, prev - 1));
    const int it = pcg(n, bs, rp, ci, v, w, b, x, tol, max_iters, rel_res);
    omp_set_num_threads(prev);
    return it;
#else
    (void)n_threads;
    return pcg(n, bs, rp, ci, v, w, b, x, tol, max_iters, rel_res);
#endif
}

// Armijo backtracking line search (LineSearch.hh:26-65).  Writes the accepted
// point to x_out; returns the number of trials, or -(trials) when no step was
// accepted (x_out = x0, the reference's warning path).
int or_line_search(void* fnp, const double* x0, const double* d, double f, const double* g,
                   double s_max, double shrink, int max_iters, double armijo_const,
                   double* x_out, double* f_out, double* s_out)
{
    Function& fn = *(Function*)fnp;
    const long n = fn.n_vars;
    if (s_max <= 0.0) { g_err = "Max step size not positive."; return 0; }
    const bool try_one = s_max > 1.0;
    const double dg = dot(n, d, g);
    double s = s_max;
    for (int i = 0; i < max_iters; ++i)
    {
        for (long j = 0; j < n; ++j) x_out[j] = x0[j] + s * d[j];
        const double f_new = eval(fn, x_out);
        if (f_new != f_new) { g_err = "energy is NaN"; return 0; }
        if (f_new <= f + armijo_const * s * dg)
        {
            *f_out = f_new; *s_out = s;
            return i + 1;
        }
        if (try_one && s > 1.0 && s * shrink < 1.0) s = 1.0;
        else s *= shrink;
    }
    for (long j = 0; j < n; ++j) x_out[j] = x0[j];
    *f_out = f; *s_out = 0.0;
    return -max_iters;
}

// Scalar-operator known answers: evaluates reference test expressions on a
// 2-variable Dual and writes val, grad[2], hess[4].
//   which 0: sqr(x) / y                          (ScalarTestBinaryOperators.cc:623-642)
//   which 1: (0.5x^2 - y^2 + 2x - y) / (-(x-2)^2 - (y-3)^2 + 1)   (:655-699)
//   which 2: (sqr(x) + x) * (sqr(y) - y) / (y - 1)                (:711-727)
//   which 3: -(x * y) + x - y + 2 * x * 3 - 1                      (unary/±/× mix)
void or_dual_known_answer(int which, double xv, double yv, double* out)
{
    Dual<2> x(xv, 0), y(yv, 1), f;
    switch (which)
    {
        case 0: f = sqr(x) / y; break;
        case 1:
        {
            Dual<2> a = 0.5 * sqr(x) - sqr(y) + 2.0 * x - y;
            Dual<2> b = -sqr(x - 2.0) - sqr(y - 3.0) + 1.0;
            f = a / b; break;
        }
        case 2: f = (sqr(x) + x) * (sqr(y) - y) / (y - 1.0); break;
        default: f = -(x * y) + x - y + 2.0 * x * 3.0 - 1.0; break;
    }
    out[0] = f.val; out[1] = f.grad[0]; out[2] = f.grad[1];
    out[3] = f.hess[0][0]; out[4] = f.hess[0][1]; out[5] = f.hess[1][0]; out[6] = f.hess[1][1];
}

// Mixed-derivative VJP of the symmetric Dirichlet tet energy with respect to
// the rest positions theta (per vertex, same AoS layout as x), as TinyAD would
// compute it: a second-order scalar over the 24 variables (x_e, theta_e) whose
// Hessian block rows 0..11 / columns 12..23 is d2E / dx dtheta (the truncated
// Hessian block of Scalar<24, T, true, 0, 12, 12, 12>, Scalar.hh:24-47,
// :188-196, pinned to the full block by ScalarTestHessianBlock.cc:50-100).
// E_e = w_scale det(Mr(theta)) / 6 (||J||^2 + ||J^-1||^2), J = M(x) Mr(theta)^-1.
// out[3 v + i] = sum_e sum_k (d2E_e / dx_k dtheta_{v,i}) lambda_k, elements in order.
int or_tet_rest_mixed_vjp(long n_tets, const int* conn, const double* x, const double* rest, const double* lambda,
                          double w_scale, long n_vars, double* out)
{
    for (long i = 0; i < n_vars; ++i) out[i] = 0.0;
    std::vector<double> contrib((size_t)n_tets * 12);
    #pragma omp parallel for schedule(static)
    for (long e = 0; e < n_tets; ++e)
    {
        using D = Dual<24>;
        D X[12], R[12];
        for (int a = 0; a < 4; ++a)
            for (int i = 0; i < 3; ++i)
            {
                X[3 * a + i] = D(x[3 * conn[4 * e + a] + i], 3 * a + i);
                R[3 * a + i] = D(rest[3 * conn[4 * e + a] + i], 12 + 3 * a + i);
            }
        D M[3][3], Mr[3][3];
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k)
            {
                M[i][k] = X[3 * (k + 1) + i] - X[i];
                Mr[i][k] = R[3 * (k + 1) + i] - R[i];
            }
        const D detM = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1])
                     - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0])
                     + M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
        double* c = &contrib[(size_t)e * 12];
        if (detM.val <= 0.0)
        {
            for (int k = 0; k < 12; ++k) c[k] = 0.0;
            continue;
        }
        // Mr^-1 via cofactors
        const D r00 = Mr[1][1] * Mr[2][2] - Mr[1][2] * Mr[2][1];
        const D r01 = Mr[1][2] * Mr[2][0] - Mr[1][0] * Mr[2][2];
        const D r02 = Mr[1][0] * Mr[2][1] - Mr[1][1] * Mr[2][0];
        const D detR = Mr[0][0] * r00 + Mr[0][1] * r01 + Mr[0][2] * r02;
        const D idR = 1.0 / detR;
        D Ri[3][3];
        Ri[0][0] = r00 * idR; Ri[1][0] = r01 * idR; Ri[2][0] = r02 * idR;
        Ri[0][1] = (Mr[0][2] * Mr[2][1] - Mr[0][1] * Mr[2][2]) * idR;
        Ri[1][1] = (Mr[0][0] * Mr[2][2] - Mr[0][2] * Mr[2][0]) * idR;
        Ri[2][1] = (Mr[0][1] * Mr[2][0] - Mr[0][0] * Mr[2][1]) * idR;
        Ri[0][2] = (Mr[0][1] * Mr[1][2] - Mr[0][2] * Mr[1][1]) * idR;
        Ri[1][2] = (Mr[0][2] * Mr[1][0] - Mr[0][0] * Mr[1][2]) * idR;
        Ri[2][2] = (Mr[0][0] * Mr[1][1] - Mr[0][1] * Mr[1][0]) * idR;
        D J[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) J[i][j] = M[i][0] * Ri[0][j] + M[i][1] * Ri[1][j] + M[i][2] * Ri[2][j];
        const D C00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        const D C01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        const D C02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        const D detJ = J[0][0] * C00 + J[0][1] * C01 + J[0][2] * C02;
        const D inv = 1.0 / detJ;
        const D Ci[9] = { C00, C01, C02,
                          J[0][2] * J[2][1] - J[0][1] * J[2][2], J[0][0] * J[2][2] - J[0][2] * J[2][0],
                          J[0][1] * J[2][0] - J[0][0] * J[2][1], J[0][1] * J[1][2] - J[0][2] * J[1][1],
                          J[0][2] * J[1][0] - J[0][0] * J[1][2], J[0][0] * J[1][1] - J[0][1] * J[1][0] };
        D nI = sqr(Ci[0] * inv);
        for (int t = 1; t < 9; ++t) nI = nI + sqr(Ci[t] * inv);
        D nJ = sqr(J[0][0]);
        for (int t = 1; t < 9; ++t) nJ = nJ + sqr(J[t / 3][t % 3]);
        const D E = (w_scale / 6.0) * detR * (nJ + nI);
        for (int k = 0; k < 12; ++k)
        {
            double acc = 0.0;
            for (int r = 0; r < 12; ++r) acc += E.hess[r][12 + k] * lambda[3 * conn[4 * e + r / 3] + r % 3];
            c[k] = acc;
        }
    }
    for (long e = 0; e < n_tets; ++e)
        for (int k = 0; k < 12; ++k) out[3 * conn[4 * e + k / 3] + k % 3] += contrib[(size_t)e * 12 + k];
    return 0;
}

// --- vector functions ---------------------------------------------------------
void* or_vfn_create(int d, long n_handles)
{
    if (d < 1 || d > 3 || n_handles < 1) { g_err = "invalid dimension or handle count"; return nullptr; }
    VFunction* f = new VFunction();
    f->d = d; f->n_handles = n_handles; f->n_vars = (long)d * n_handles;
    return f;
}

void or_vfn_destroy(void* fn) { delete (VFunction*)fn; }

int or_vfn_add_elements(void* fnp, int rid, int N, long n_elem, const int* conn, const double* params, int n_params)
{
    VFunction& fn = *(VFunction*)fnp;
    const int M = residual_outputs(rid, fn.d, N);
    if (M < 0) { g_err = "residual does not match dimension/valence"; return 1; }
    if (residual_params(rid, fn.d) != n_params) { g_err = "parameter count mismatch"; return 1; }
    RTerm t;
    t.rid = rid; t.N = N; t.K = fn.d * N; t.M = M; t.n_params = n_params; t.n_elem = n_elem;
    if (t.K != 1 && t.K != 2 && t.K != 3 && t.K != 6) { g_err = "unsupported element size"; return 1; }
    t.conn.assign(conn, conn + n_elem * N);
    t.params.assign(params, params + n_elem * n_params);
    for (long i = 0; i < n_elem * N; ++i)
        if (t.conn[i] < 0 || t.conn[i] >= fn.n_handles) { g_err = "handle out of range"; return 1; }
    fn.n_outputs += n_elem * M;
    fn.terms.push_back(std::move(t));
    return 0;
}

long or_vfn_n_outputs(void* fnp) { return ((VFunction*)fnp)->n_outputs; }

void or_vfn_eval(void* fnp, const double* x, double* r) { veval(*(VFunction*)fnp, x, r); }

double or_vfn_sum_of_squares(void* fnp, const double* x) { return vsum_of_squares(*(VFunction*)fnp, x); }

// eval_with_jacobian (VectorFunctionImpl.hh:157-176): r, then J (n_outputs x
// n_vars) from the triplets; fetch it with or_vfn_get_jacobian.
int or_vfn_eval_with_jacobian(void* fnp, const double* x, double* r, long* nnz)
{
    VFunction& fn = *(VFunction*)fnp;
    try
    {
        std::vector<Triplet> trip;
        veval_with_jacobian(fn, x, r, trip);
        triplets_to_csr(fn.n_outputs, trip, fn.row_ptr, fn.col_idx, fn.values);
        *nnz = (long)fn.values.size();
        return 0;
    }
    catch (const std::exception& ex) { g_err = ex.what(); return 3; }
}

void or_vfn_get_jacobian(void* fnp, int* row_ptr, int* col_idx, double* values)
{
    VFunction& fn = *(VFunction*)fnp;
    std::copy(fn.row_ptr.begin(), fn.row_ptr.end(), row_ptr);
    std::copy(fn.col_idx.begin(), fn.col_idx.end(), col_idx);
    std::copy(fn.values.begin(), fn.values.end(), values);
}

} // extern "C"
