/*
 * tinyad_amd.h -- C ABI of the MI355X (gfx950) sparse second-order path.
 *
 * One shared library, libtinyad_amd.so, exports exactly the functions below.
 * Signatures use plain pointers, sizes and enums only.  Every function returns
 * a tad_status; on failure tad_last_error() holds a message for the calling
 * thread (the reference's TINYAD_ERROR_throw / TINYAD_ASSERT, Utils/Out.hh:64-157,
 * throws std::runtime_error instead; the C++ front-end in
 * include/tinyad_amd/ScalarFunction.hh turns non-zero statuses back into
 * std::runtime_error).
 *
 * Memory: every array argument is paired with the tad_mem of the call.  With
 * TAD_MEM_DEVICE the pointers are HIP device pointers on the context's device
 * and nothing crosses PCIe; with TAD_MEM_HOST they are host pointers.
 *
 * Variable layout matches the reference: x has n_vars = d * n_handles entries,
 * entry d * handle + i is coordinate i of that handle (Detail/Element.hh:159-170).
 *
 * Hessians are returned as CSR values over a fixed pattern (tad_fn_pattern).
 * The pattern is the union of all element (i, j) pairs in full symmetric
 * storage, explicit zeros included, i.e. the pattern Eigen's setFromTriplets
 * produces for the reference's triplets (ScalarFunctionImpl.hh:378-399).  For a
 * symmetric matrix the CSR arrays equal the CSC arrays of Eigen's
 * column-major SparseMatrix.
 *
 * Determinism: no floating-point atomics are used.  Element contributions are
 * summed per output entry in element order (terms in insertion order), so
 * repeated calls return identical bits.
 */
#ifndef TINYAD_AMD_H
#define TINYAD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum tad_status
{
    TAD_OK = 0,
    TAD_ERR_INVALID = 1,     /* bad argument / size mismatch (TINYAD_ASSERT_*)            */
    TAD_ERR_HIP = 2,         /* HIP runtime error                                         */
    TAD_ERR_NONFINITE = 3,   /* non-finite element gradient or Hessian                   */
                             /* (ScalarObjectiveTerm.hh:252-253)                          */
    TAD_ERR_SOLVE = 4,       /* linear solve failed (NewtonDirection.hh:43-44)           */
    TAD_ERR_NAN_ENERGY = 5,  /* NaN energy in the line search (LineSearch.hh:52)          */
    TAD_ERR_UNSUPPORTED = 6
} tad_status;

typedef enum tad_mem
{
    TAD_MEM_HOST = 0,
    TAD_MEM_DEVICE = 1
} tad_mem;

typedef enum tad_precision
{
    TAD_PREC_FP64 = 0,  /* everything in FP64                                              */
    TAD_PREC_MIXED = 1  /* element gradient/Hessian/projection in FP32; energies, CSR      */
                        /* reduction, gradient reduction and the solve in FP64            */
} tad_precision;

/* Energy catalog (the per-element functions add_elements can register).
 * Parameter rows, one per element (row-major, params_per_elem doubles):
 *   TAD_E_SYMDIR_TRI  d=2, N=3: [Mr^-1 (2x2 row-major), w]     5 doubles
 *   TAD_E_SYMDIR_TET  d=3, N=4: [Mr^-1 (3x3 row-major), w]    10 doubles
 *     E = +inf if det(M) <= 0 else w (||M Mr^-1||_F^2 + ||(M Mr^-1)^-1||_F^2),
 *     M = col_mat(v1 - v0, v2 - v0[, v3 - v0])   (tests/NewtonTest.cc:28-44)
 *   TAD_E_PENALTY     N=1, any d: [target (d), w]     E = w ||target - x_v||^2
 *                                                       (tests/NewtonTest.cc:48-55)
 *   TAD_E_QUADRATIC   any N, d with N*d <= 8: [A (KxK), b (K), c], K = N*d
 *     E = 0.5 x_e^T A x_e + b^T x_e + c over the element's local variables
 *     (tests/ScalarFunctionTest.cc, tests/DynamicElementsTest.cc:100-111)      */
typedef enum tad_energy
{
    TAD_E_SYMDIR_TRI = 1,
    TAD_E_SYMDIR_TET = 2,
    TAD_E_PENALTY = 3,
    TAD_E_QUADRATIC = 4,
    TAD_E_DEVICE = 100      /* terms added by tad_fn_add_elements_device */
} tad_energy;

typedef struct tad_ctx_s* tad_ctx;
typedef struct tad_fn_s* tad_fn;

/* Message of the last failed call on this thread ("" if none). */
const char* tad_last_error(void);

/* Library version string. */
const char* tad_version(void);

/* --- context: one device, one HIP stream --------------------------------- */

/* Binds device `device`; all work of the context's functions is queued on one
 * HIP stream (created here, default priority) unless tad_ctx_set_stream is
 * given an existing one.  Replaces the reference's OpenMP team
 * (Detail/Parallel.hh:20-69, EvalSettings.hh:14-21). */
tad_status tad_ctx_create(int device, tad_ctx* out);
/* Releases the caller's reference; the context lives on until the last
 * function created on it is destroyed. */
tad_status tad_ctx_destroy(tad_ctx ctx);
/* hip_stream: a hipStream_t (NULL = the context's own stream). */
tad_status tad_ctx_set_stream(tad_ctx ctx, void* hip_stream);
tad_status tad_ctx_get_stream(tad_ctx ctx, void** hip_stream);
tad_status tad_ctx_synchronize(tad_ctx ctx);
/* When enabled, every kernel launch is bracketed by HIP events on the
 * context's stream; tad_ctx_kernel_stats reports per-kernel-name totals. */
tad_status tad_ctx_enable_timing(tad_ctx ctx, int enable);
/* name: kernel family ("element", "assemble_hessian", "spmv", ...);
 * returns launch count and summed milliseconds since the last reset. */
tad_status tad_ctx_kernel_stats(tad_ctx ctx, const char* name, int64_t* launches, double* total_ms);
tad_status tad_ctx_reset_stats(tad_ctx ctx);

/* --- objective function ---------------------------------------------------- */

/* scalar_function<d>(range(n_handles)) (ScalarFunctionImpl.hh:418-442, ctor
 * :47-61).  d in 1..3; handles are 0..n_handles-1. */
tad_status tad_fn_create(tad_ctx ctx, int d, int64_t n_handles, tad_precision precision, tad_fn* out);
tad_status tad_fn_destroy(tad_fn fn);

/* add_elements<N>(range(n_elem), energy) (ScalarFunctionImpl.hh:63-100).
 * conn: n_elem x N handle indices (row-major, host memory), local variable order
 * = connectivity order (first-access order of Element::variables, Element.hh:208-260).
 * params: n_elem x params_per_elem (row-major, host memory).  A quadratic
 * element may list a handle more than once: as Element::variables gives a
 * handle met again the local variables of its first access (Element.hh:221-243),
 * the repeated entries share that handle's variables, their derivatives add up
 * and the merged element Hessian is what gets projected.  Symmetric Dirichlet
 * elements with a repeated vertex (degenerate simplices) are rejected.
 * Invalidates the Hessian pattern. */
tad_status tad_fn_add_elements(tad_fn fn, tad_energy energy, int valence, int64_t n_elem,
                               const int32_t* conn, const double* params, int params_per_elem);

/* --- elements with caller-compiled device energies ---------------------------
 * The reference's add_elements takes any generic lambda over the element's
 * active scalars (ScalarFunction.hh:72-78, type-erased in
 * ScalarObjectiveTerm.hh:75-134).  Here such an energy is a functor compiled
 * by the caller with hipcc into its own code object; the header
 * include/tinyad_amd/DeviceElements.hh instantiates its element kernels
 * (value + gradient, + Hessian, + projected Hessian, each with FP64 and FP32
 * contributions) and its line-search energy kernel, and fills a
 * tad_device_ops with host functions that launch them.  The library owns the
 * pattern, the scatter maps, assembly, solve and line search as for its own
 * energies, and calls these functions with the arguments below. */
#define TAD_DEVICE_OPS_VERSION 2

typedef struct tad_device_element_args
{
    int64_t n_elem;
    const int32_t* conn;    /* device, SoA: conn[a * n_elem + e]                       */
    const double* params;   /* device, SoA: params[k * n_elem + e]                     */
    const double* x;        /* device, n_vars                                           */
    double eps;             /* projection eps (mode 3)                                  */
    double* elem_f;         /* device: energy of element e at elem_f[e]                 */
    void* G;                /* gradient-part slots, double or float by precision       */
    void* Hc;               /* Hessian-block slots, double or float by precision       */
    const uint32_t* gpos;   /* slot of (vertex a, element e): gpos[a * n_elem + e]      */
    const uint32_t* hpos;   /* slot of upper block (a <= b) of element e; top bit set:  */
                            /* the slot holds the transposed block                      */
    int* err;               /* set to 1 when a derivative is not finite                 */
    int mode;               /* 1 gradient, 2 Hessian, 3 projected Hessian               */
    int precision;          /* 0: FP64 contributions, 1: FP32                           */
    void* stream;           /* hipStream_t to launch on                                 */
} tad_device_element_args;

typedef struct tad_device_energy_args
{
    int64_t n_elem;
    const int32_t* conn;
    const double* params;
    const double* x;
    const double* dir;      /* NULL: energies at x; else at x + steps[c] * dir          */
    const double* steps;    /* device, ncand step sizes                                 */
    int ncand;              /* 1..8                                                     */
    double* out;            /* out[c * stride + e]                                      */
    int64_t stride;
    void* stream;
} tad_device_energy_args;

typedef struct tad_device_ops
{
    int version;            /* TAD_DEVICE_OPS_VERSION of the header that built them    */
    int valence;            /* handles per element                                      */
    int dim;                /* d: must equal the function's                             */
    int n_params;           /* parameters per element (>= 0)                            */
    /* both return a hipError_t value (0: launched) */
    int (*element)(const tad_device_element_args* args);
    int (*energy)(const tad_device_energy_args* args);
} tad_device_ops;

/* add_elements<N>(range(n_elem), lambda) for a device energy: conn (n_elem x
 * valence) and params (n_elem x n_params), row-major host memory, as for
 * tad_fn_add_elements.  ops is copied; its functions must stay loaded while the
 * function exists.  Terms added this way take the plain one-lane-per-element
 * launch (no autotuned shapes, no rest-shape products).  An element may list a
 * handle more than once (merged as for quadratic elements): the slot maps give
 * such a repeated entry no slot (tad::kNoSlot), and the element kernels of
 * DeviceElements.hh fold its derivatives into the first occurrence. */
tad_status tad_fn_add_elements_device(tad_fn fn, const tad_device_ops* ops, int64_t n_elem, const int32_t* conn,
                                      const double* params);

/* add_elements_dynamic<V1, V2, ...>(range(n_elem), lambda)
 * (ScalarFunctionImpl.hh:102-214): elements with different numbers of
 * handles.  Element e accesses handles[elem_ptr[e] .. elem_ptr[e+1]) (host
 * memory, access order); it joins the group of the smallest valence among
 * ops[0..n_ops) that is >= its handle count, its unused trailing slots become
 * padding (coordinates read as 0, no contributions), and one term per
 * non-empty group is added in the order of ops.  ops: n_ops distinct
 * valences of one energy (same dimension and parameter count); params:
 * n_elem x n_params row-major.  Fails when an element has more handles than
 * the largest valence. */
tad_status tad_fn_add_elements_dynamic_device(tad_fn fn, const tad_device_ops* const* ops, int n_ops, int64_t n_elem,
                                              const int64_t* elem_ptr, const int32_t* handles, const double* params);

/* --- vector functions: residuals, Jacobians, sums of squares -------------------
 *
 * VectorFunction (VectorFunction.hh:19-192): each element produces a fixed
 * number of residuals; r(x) stacks them term by term, element by element
 * (VectorObjectiveTerm.hh:162-183).  The element functions are device functors
 * compiled by the caller (include/tinyad_amd/DeviceElements.hh,
 * tinyad_amd::device_residual<>), the GPU form of the reference's
 * `(auto& element) -> TINYAD_VECTOR_TYPE(element)` lambdas. */

typedef struct tad_device_residual_args
{
    int64_t n_elem;
    const int32_t* conn;    /* device, SoA: conn[a * n_elem + e]                       */
    const double* params;   /* device, SoA: params[k * n_elem + e]                     */
    const double* x;        /* device, n_vars                                           */
    double* r;              /* device: residual o of element e at r[e * n_outputs + o]  */
    double* J;              /* device or NULL: row (e, o) of the Jacobian at            */
                            /* J[(e * n_outputs + o) * K ..], K = valence * d values in  */
                            /* ascending variable order                                 */
    void* stream;           /* hipStream_t to launch on                                 */
} tad_device_residual_args;

typedef struct tad_device_residual_ops
{
    int version;            /* TAD_DEVICE_OPS_VERSION                                   */
    int valence;            /* handles per element                                      */
    int dim;                /* d: must equal the function's                             */
    int n_params;           /* parameters per element (>= 0)                            */
    int n_outputs;          /* residuals per element (1..64)                            */
    /* sum of squares f_e = r_e . r_e with g_e = 2 J_e^T r_e and, modes 2 and 3,
     * the Gauss-Newton Hessian 2 J_e^T J_e (no projection: it is positive
     * semi-definite), through the library's slot maps */
    int (*element)(const tad_device_element_args* args);
    /* f_e = r_e . r_e (passive FP64) for eval and the line search */
    int (*energy)(const tad_device_energy_args* args);
    /* r_e and, when args->J is set, the rows of J_e */
    int (*residuals)(const tad_device_residual_args* args);
} tad_device_residual_ops;

/* vector_function<d>(range(n_handles)) (VectorFunctionImpl.hh:300-325, ctor
 * :40-56): a function whose terms are residual elements.  Every tad_fn entry
 * point applies to the sum of squares f(x) = r(x) . r(x): tad_fn_eval is
 * eval_sum_of_squares, tad_fn_eval_with_gradient gives g = 2 J^T r, and the
 * second-order entry points, tad_fn_solve, tad_line_search and
 * tad_newton_step see the Gauss-Newton Hessian 2 J^T J (eps is ignored), so
 * tad_newton_step is a Gauss-Newton step (solve w = 2 w_identity for the
 * reference's gauss_newton_direction(r, J, solver, w_identity)). */
tad_status tad_fn_create_vector(tad_ctx ctx, int d, int64_t n_handles, tad_precision precision, tad_fn* out);

/* add_elements<N, M>(range(n_elem), lambda) (VectorFunctionImpl.hh:58-90):
 * conn (n_elem x valence) and params (n_elem x n_params), row-major host
 * memory; the element's residuals follow those of the terms added before.
 * Residual elements must list distinct handles (one Jacobian column per entry). */
tad_status tad_fn_add_residuals_device(tad_fn fn, const tad_device_residual_ops* ops, int64_t n_elem,
                                       const int32_t* conn, const double* params);

/* n_outputs (VectorFunction.hh:183): residuals of all terms. */
tad_status tad_fn_n_outputs(tad_fn fn, int64_t* n_outputs);

/* Builds (once) and returns the Jacobian pattern: CSR of the n_outputs x n_vars
 * matrix, row (term, element, output) holding the element's valence * d
 * variables in ascending order (the reference builds it from triplets,
 * VectorObjectiveTerm.hh:230-240).  Device pointers owned by fn; host copies
 * when row_ptr_host / col_idx_host are non-NULL. */
tad_status tad_fn_jacobian_pattern(tad_fn fn, int64_t* nnz, const int32_t** row_ptr_dev, const int32_t** col_idx_dev,
                                   int32_t* row_ptr_host, int32_t* col_idx_host);

/* eval(x) of a vector function (VectorFunctionImpl.hh:130-147): r, n_outputs values. */
tad_status tad_fn_eval_residuals(tad_fn fn, const double* x, double* r, tad_mem mem);

/* eval_with_jacobian(x, r, J) (VectorFunctionImpl.hh:157-176): r and the J
 * values in the order of tad_fn_jacobian_pattern (J_values may be NULL). */
tad_status tad_fn_eval_with_jacobian(tad_fn fn, const double* x, double* r, double* J_values, tad_mem mem);

/* eval_sum_of_squares_with_derivatives(x, f, g, r, J)
 * (VectorFunctionImpl.hh:258-275): f = sum_e r_e . r_e, g = 2 J^T r, r and
 * J (r, J_values may be NULL). */
tad_status tad_fn_eval_sum_of_squares_with_derivatives(tad_fn fn, const double* x, double* f, double* g, double* r,
                                                       double* J_values, tad_mem mem);

/* The objective terms (the reference's objective_terms, ScalarFunction.hh:229):
 * their number, and per term its energy (TAD_E_DEVICE for device energies),
 * valence and element count. */
tad_status tad_fn_n_terms(tad_fn fn, int* n_terms);
tad_status tad_fn_term_info(tad_fn fn, int term, int* energy, int* valence, int64_t* n_elem);

tad_status tad_fn_n_vars(tad_fn fn, int64_t* n_vars);
tad_status tad_fn_n_elements(tad_fn fn, int64_t* n_elements);

/* Builds (once) and returns the Hessian CSR pattern: n_vars + 1 row offsets and
 * nnz column indices, device pointers owned by fn, valid until the next
 * add_elements or tad_fn_destroy.  Host copies are written to row_ptr_host /
 * col_idx_host when those are non-NULL. */
tad_status tad_fn_pattern(tad_fn fn, int64_t* nnz, const int32_t** row_ptr_dev, const int32_t** col_idx_dev,
                          int32_t* row_ptr_host, int32_t* col_idx_host);

/* eval(x) (ScalarFunctionImpl.hh:256-282): passive energy, FP64 sum. */
tad_status tad_fn_eval(tad_fn fn, const double* x, tad_mem mem, double* f);

/* Like tad_fn_eval with x and f in device memory, but the library allocates
 * nothing, not even on the first call: no counterpart in the reference; for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops).  The per-workgroup partial sums go to a static device array of
 * 65,536 doubles (512 KiB per device, held until the process ends), so the
 * function's catalog terms may hold at most 16,777,216 elements in total
 * (larger functions fail with TAD_ERR_UNSUPPORTED, as do functions with
 * device-functor terms).  Asynchronous: f_dev[0] is written in the order of
 * the context's stream and the call does not synchronize.  The sum order
 * differs from tad_fn_eval's (equal to rounding).  Calls that overlap on one
 * device share the array and race. */
tad_status tad_fn_eval_noalloc(tad_fn fn, const double* x_dev, double* f_dev);

/* eval_with_gradient(x, f, g) (ScalarFunctionImpl.hh:284-314). g: n_vars. */
tad_status tad_fn_eval_with_gradient(tad_fn fn, const double* x, double* f, double* g, tad_mem mem);

/* eval_with_derivatives(x, f, g, H) (ScalarFunctionImpl.hh:316-352): unprojected
 * Hessian.  H_values: nnz values in the pattern order (may be NULL: the values
 * stay in the function's device buffer, see tad_fn_device_buffers). */
tad_status tad_fn_eval_with_derivatives(tad_fn fn, const double* x, double* f, double* g, double* H_values,
                                        tad_mem mem);

/* eval_with_hessian_proj(x, f, g, H_proj, eps) (ScalarFunctionImpl.hh:378-416,
 * per-element projection HessianProjection.hh:48-101; eps >= 0 clamps
 * eigenvalues to eps, eps < 0 takes absolute values; default 1e-9). */
tad_status tad_fn_eval_with_hessian_proj(tad_fn fn, const double* x, double eps, double* f, double* g,
                                         double* H_values, tad_mem mem);

/* --- batches of independent problems ---------------------------------------- */

/* Declares consecutive handle ranges [handle_offsets[s], handle_offsets[s+1])
 * as independent problems (e.g. many meshes evaluated in one launch).  Every
 * element must lie inside one segment and, within each term, elements must be
 * ordered by segment.  The reference has no counterpart: it evaluates one
 * ScalarFunction per mesh. */
tad_status tad_fn_set_segments(tad_fn fn, int64_t n_segments, const int64_t* handle_offsets);

/* Per-segment energies of the last evaluation (n_segments values). */
tad_status tad_fn_segment_energies(tad_fn fn, double* f, tad_mem mem);

/* Replaces the parameter rows of term `term` (n_elem x params_per_elem,
 * row-major, same layout as add_elements); the sparsity pattern is kept.
 * Counterpart of changing the data an add_elements lambda captures by
 * reference (ScalarFunction.hh:60-77: the element lambda reads it on every
 * evaluation). */
tad_status tad_fn_set_params(tad_fn fn, int term, const double* params, tad_mem mem);

/* Rest-shape mixed products for implicit differentiation.  For a
 * TAD_E_SYMDIR_TET term whose parameters are derived from per-vertex rest
 * positions theta (same layout as x) as Mr(theta)^-1 and
 * w_e = w_scale det(Mr_e(theta)) / 6, writes
 *   out = sum_e (d2 E_e / dtheta dx)^T lambda        (n_vars values)
 * at x.  Inverted elements (det M <= 0, energy +inf) contribute nothing.
 * No kernel counterpart in the reference; its semantics are those of the
 * truncated mixed Hessian block Scalar<24, T, true, 0, 12, 12, 12>
 * (Scalar.hh:24-47, :188-196) contracted with lambda. */
tad_status tad_fn_rest_vjp(tad_fn fn, int term, const double* x, const double* rest, const double* lambda,
                           double w_scale, double* out, tad_mem mem);

/* Times launch shapes of the projected-Hessian element kernel on x for every
 * symmetric-Dirichlet term and keeps the fastest: the plain grid (one tile of
 * 256 elements per workgroup) or a persistent grid of 64 / 128 / 256 lanes
 * with 1, 2, 4 or 8 workgroups per CU and a strided or per-XCD tile order.
 * Results are bit-identical for every shape.  report (optional) receives a
 * text table of the timings.  No counterpart in the reference. */
tad_status tad_fn_autotune(tad_fn fn, const double* x, tad_mem mem, double eps, char* report, size_t report_len);

/* The launch shapes tad_fn_autotune chose, to store and restore without
 * timing again.  Element launch of term `term` (projected Hessian path of a
 * symmetric-Dirichlet term): block 0 is the plain grid; otherwise a
 * persistent grid of `grid` workgroups (1 .. 2^20) of 64, 128 or 256 lanes,
 * map 0 strided or 1 per-XCD contiguous (grid a multiple of 8).  Hessian
 * assembly shape 0 .. 3 (d*d lanes or one thread per CSR block, d*d lanes or
 * one thread per upper block writing it and its mirror).  Every shape gives
 * bit-identical results.  No counterpart in the reference. */
tad_status tad_fn_get_launch_shape(tad_fn fn, int term, int* block, int* grid, int* map);
tad_status tad_fn_set_launch_shape(tad_fn fn, int term, int block, int grid, int map);
tad_status tad_fn_get_assembly_shape(tad_fn fn, int* shape);
tad_status tad_fn_set_assembly_shape(tad_fn fn, int shape);

/* Pipelined second-order evaluation (no counterpart in the reference): the
 * largest catalog term's elements run in `chunks` launches and the Hessian and
 * gradient rows each launch completes are assembled on a second stream while
 * the next one runs.  Results are identical for every chunk count (each entry
 * is summed in the same fixed order).  chunks: 0 automatic (off), 1 off, at
 * most 64 (fewer when the term has fewer 256-element tiles);
 * tad_fn_autotune picks among 1, 2, 4 and 8.  get: the request and the chunk
 * count in use. */
tad_status tad_fn_get_pipeline(tad_fn fn, int* requested, int* used);
tad_status tad_fn_set_pipeline(tad_fn fn, int chunks);

/* CG iterations of the direction solves replayed from a HIP graph (no
 * counterpart in the reference): the launches of 8 iterations (and a kernel
 * that advances the iteration index, which the kernels then read on the
 * device) are recorded once per operand set and replayed.  Identical results
 * either way.  mode: 1 on, 0 off (the default: on MI355X direct launches are
 * as fast at 3K handles and faster above).  Solves with kernel timing enabled
 * always launch directly. */
tad_status tad_fn_set_solve_graphs(tad_fn fn, int mode);

/* Device pointers of the function's own buffers holding the gradient and CSR
 * values of the last second-order evaluation (valid until the next call). */
tad_status tad_fn_device_buffers(tad_fn fn, const double** g_dev, const double** H_values_dev);

/* --- solver ---------------------------------------------------------------- */

typedef struct tad_solve_opts
{
    double w_identity;  /* regularization w in (w I + H) d = rhs (NewtonDirection.hh:32)  */
    double tol;         /* relative residual ||r|| <= tol ||rhs||                        */
    int max_iters;
    int check_every;    /* host polls convergence every this many iterations             */
    int coarse_aggregate; /* two-level preconditioner: handles per aggregate of the      */
                        /* coarse space; 0: block-Jacobi only; < 0 (default): automatic, */
                        /* about 1,400 coarse unknowns up to 30K handles, growing with   */
                        /* the cube root of the handle count to at most 2,800, with at   */
                        /* least 32 handles per aggregate                                */
    int coarse_reuse;   /* 1: keep the coarse space and inverse coarse matrix of the     */
                        /* function's previous solve when it is compatible (same         */
                        /* aggregates, unknowns per aggregate and w, neither solve       */
                        /* reduced) instead of rebuilding them from this H; 0 (default): */
                        /* rebuild every solve.  A lagged coarse operator is still       */
                        /* symmetric positive definite, so CG converges to the same      */
                        /* tolerance; only the iteration count changes.                 */
} tad_solve_opts;

typedef struct tad_solve_stats
{
    int iters;
    int converged;
    double rel_residual;
    int coarse_unknowns;    /* size of the coarse space used (0: block-Jacobi only)      */
    int coarse_reused;      /* 1: the coarse inverse of an earlier solve was used        */
} tad_solve_stats;

tad_status tad_solve_default_opts(tad_solve_opts* opts);

/* Solves (w I + H) sol = rhs by preconditioned CG (FP64) on the function's
 * pattern; replaces the SimplicialLDLT factorize/solve of newton_direction
 * (NewtonDirection.hh:25-48, LinearSolver.hh:13-19).  The preconditioner is
 * block-Jacobi (d x d diagonal blocks) plus, unless coarse_aggregate is 0, an
 * additive coarse correction on aggregates of about that many handles
 * (translations, and rotations about the positions of the last second-order
 * evaluation when there was one; Galerkin coarse matrix inverted in FP64; the
 * iterations read an FP32 copy of the inverse, shifted to stay positive
 * definite, when its condition estimate bounds the change of the correction
 * by 10 %).  Like SimplicialLDLT, the solve reads H's lower triangle (and
 * diagonal) only: the matrix solved with is that triangle mirrored.
 * H_values == NULL uses the function's buffer from the last second-order eval. */
tad_status tad_fn_solve(tad_fn fn, const double* H_values, const double* rhs, double* sol, tad_mem mem,
                        const tad_solve_opts* opts, tad_solve_stats* stats);

/* The solve restricted to a subspace whose basis B selects variables
 * (newton_direction_reduced_basis, NewtonDirection.hh:76-105, for B made of
 * columns of the identity, the pinning case): sol = B (B^T H B + w I)^-1 B^T
 * rhs, i.e. CG on the free variables, and sol = 0 on the fixed ones.
 * free_vars: n_vars bytes, nonzero = free, in the memory space `mem` like
 * the vectors.  Same preconditioner as tad_fn_solve, restricted to the free
 * variables. */
tad_status tad_fn_solve_reduced(tad_fn fn, const double* H_values, const double* rhs, double* sol, tad_mem mem,
                                const unsigned char* free_vars, const tad_solve_opts* opts, tad_solve_stats* stats);

/* The solve in a general subspace (newton_direction_reduced_basis,
 * NewtonDirection.hh:76-105, for any sparse basis): sol = B y with
 * (B^T H B + w I_m) y = B^T rhs, B an n_vars x m CSR matrix (b_row_ptr:
 * n_vars + 1 offsets from 0, b_col: column indices in 0 .. m-1, b_val:
 * finite values; all in the memory space `mem` like the vectors).  Conjugate
 * gradients on y with the operator applied as products with B, H and B^T
 * and the diagonal of B^T H B + w I as preconditioner (no coarse space);
 * tol, max_iters and w_identity of `opts` apply.  TAD_ERR_SOLVE when a
 * diagonal entry or a CG step shows B^T H B + w I not positive definite
 * (the reference's factorization fails there); stats->converged = 0 when
 * max_iters ran out. */
tad_status tad_fn_solve_basis(tad_fn fn, const double* H_values, int64_t m, const int32_t* b_row_ptr,
                              const int32_t* b_col, const double* b_val, const double* rhs, double* sol, tad_mem mem,
                              const tad_solve_opts* opts, tad_solve_stats* stats);

typedef struct tad_newton_opts
{
    double projection_eps;  /* eval_with_hessian_proj eps (default 1e-9); NaN: unprojected */
                            /* Hessian (eval_with_derivatives, ScalarFunctionImpl.hh:333)  */
    tad_solve_opts solve;   /* Newton direction solve                                   */
    double ls_s_max;        /* line_search defaults (LineSearch.hh:33-36):               */
    double ls_shrink;       /*   s_max 1, shrink 0.8, 64 trials, Armijo c 1e-4          */
    int ls_max_iters;
    double ls_armijo;
    int ls_batch;           /* step sizes evaluated per energy launch (1..8)            */
    int coarse_refresh;     /* k > 1: the direction solve reuses the function's coarse   */
                            /* space (solve.coarse_reuse) until k solves have used it,  */
                            /* i.e. rebuilds it on every k-th Newton step.  0 or 1      */
                            /* (default): as solve.coarse_reuse says.                   */
} tad_newton_opts;

typedef struct tad_newton_stats
{
    double f;               /* energy at the input x                                   */
    double f_new;           /* energy at the returned x                                */
    double step;            /* accepted step size (0 if the line search failed)         */
    double newton_decrement;/* -0.5 d^T g (NewtonDecrement.hh:20-26)                     */
    double grad_norm;       /* ||g||_2 at the input x                                  */
    int cg_iters;
    int ls_trials;          /* energies evaluated in the line search                    */
    int ls_failed;          /* 1: no step accepted, x unchanged (LineSearch.hh:62-64)   */
    int solve_converged;
    int coarse_unknowns;    /* coarse space of the direction solve (0: none)            */
    int coarse_reused;      /* 1: that solve reused an earlier coarse inverse           */
} tad_newton_stats;

tad_status tad_newton_default_opts(tad_newton_opts* opts);

/* One projected-Newton iteration on x (n_vars, updated in place):
 * eval_with_hessian_proj -> newton_direction -> line_search
 * (the loop body of tests/NewtonTest.cc:69-75). */
tad_status tad_newton_step(tad_fn fn, double* x, tad_mem mem, const tad_newton_opts* opts, tad_newton_stats* stats);

/* line_search(x0, d, f, g, eval) (LineSearch.hh:26-65) on device data. */
tad_status tad_line_search(tad_fn fn, const double* x0, const double* dir, double f, const double* g,
                           double* x_out, tad_mem mem, const tad_newton_opts* opts, double* f_out,
                           double* step_out, int* trials_out);

/* newton_decrement(d, g) = -0.5 d^T g (NewtonDecrement.hh:20-26): the
 * stopping criterion, with the fixed-order FP64 device dot product of
 * tad_newton_step (whose tad_newton_stats.newton_decrement it reproduces for
 * the step's direction).  dir, g: n_vars values in memory space `mem`. */
tad_status tad_newton_decrement(tad_fn fn, const double* dir, const double* g, tad_mem mem, double* out);

/* --- self test -------------------------------------------------------------- */

/* Runs the device Dual<2> operators on a reference known-answer expression
 * (the same list as the CPU oracle's or_dual_known_answer) and writes
 * val, grad[2], Hess[2x2] to out[7] (host memory). */
tad_status tad_selftest_dual(tad_ctx ctx, int which, double x, double y, double* out);

#ifdef __cplusplus
}
#endif

#endif /* TINYAD_AMD_H */
