// Internal interface between the host library (tad_api.cpp) and the kernel
// launchers (tad_kernels.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include "tinyad_amd.h"
#include "tinyad_amd/device/Modes.hpp"

namespace tad {

struct TermDev
{
    int energy = 0;
    int N = 0;          // valence
    int n_params = 0;
    int64_t n_elem = 0;
    int32_t* conn = nullptr;   // device SoA: conn[a * n_elem + e]
    double* params = nullptr;  // device SoA: params[k * n_elem + e]
    uint32_t* gpos = nullptr;  // SoA [N][n_elem]: gradient-part slot of (vertex a, element e)
    uint32_t* hpos = nullptr;  // SoA [N(N+1)/2][n_elem]: slot of upper block (a <= b) of element e;
                               // top bit: the slot's block is (b, a), write the transpose
    int64_t f_base = 0;        // first entry of this term in the element-energy array
    // Hessian blocks summed per group of kElemGroup consecutive elements (one
    // wave of the element kernel) in LDS before they are written (grouped
    // catalog terms): hpos then holds, per (block a <= b, element), the
    // group slot (bits 0-15), the pass of the sum it joins (bits 16-23) and
    // the transpose flag (bit 31); gpass[q * N(N+1)/2 + block] the passes of
    // group q for that block; group q's slots are gout[gout_ptr[q] ..
    // gout_ptr[q+1]) (contribution slots in Hc); group_max the most slots of
    // one group.  Gradient parts keep one slot per element (gpos).
    int grouped = 0;
    int group_max = 0;
    uint8_t* gpass = nullptr;
    uint32_t* gout_ptr = nullptr;
    uint32_t* gout = nullptr;
    // launch shape of the projected-Hessian element kernel chosen by the
    // autotuner: tune_block = 0 is the plain one-tile-per-workgroup grid,
    // otherwise a persistent grid of tune_grid workgroups of tune_block lanes
    // with tile map tune_map (0 strided, 1 per-XCD contiguous)
    int tune_block = 0;
    int tune_grid = 0;
    int tune_map = 0;
    // caller-compiled element and energy kernels (energy == TAD_E_DEVICE)
    tad_device_ops dev_ops{};
    // residual terms of vector functions: residuals per element, first
    // residual row of the term, first Jacobian value of the term, and the
    // caller-compiled residual / Jacobian launcher
    int n_out = 0;
    int64_t r_base = 0;
    int64_t jv_base = 0;
    int (*residuals)(const tad_device_residual_args*) = nullptr;
};

// Elements per group of a grouped term (one wave), and the most distinct
// upper blocks a group may have (its LDS accumulators); terms with a larger
// group keep one slot per element block.
constexpr int kElemGroup = 64;
constexpr int kGroupMaxBlocks = 224;  // 4 waves x 224 x 9 doubles < 64 KiB of LDS

// Solver bookkeeping written by block 0 of the PCG kernels (see tad_linalg.hip).
struct PcgScalars
{
    double bb, rr, tol2;
    double rho_hist[2], alpha_hist[2];  // rho_k and alpha_k, slot k & 1
    int iters, max_iters, done, breakdown;
    int k;  // iteration index for launches that read it from here (graph replays)
};

// Kernel timing hook (set by the host library when timing is enabled).
struct Timer
{
    virtual void begin(const char* name, hipStream_t s) = 0;
    virtual void end(const char* name, hipStream_t s) = 0;
    virtual ~Timer() {}
};

constexpr int kReduceBlocks = 256;

// MODE: 1 grad, 2 hessian, 3 projected hessian.  precision: 0 fp64, 1 fp32
// contributions.  G / Hc: destination-sorted gradient-part and Hessian-block
// contribution arrays (d and d*d values per slot).
hipError_t launch_element(const TermDev& t, int d, int precision, int mode, const double* x, double eps,
                          double* elem_f, void* G, void* Hc, int* err, hipStream_t s);

// The elements [e0, e1) of a catalog term (launch shape as for the whole term).
hipError_t launch_element_range(const TermDev& t, int d, int precision, int mode, const double* x, double eps,
                                double* elem_f, void* G, void* Hc, int* err, hipStream_t s, int64_t e0, int64_t e1);

// Energies at x + steps[c] * dir (dir may be null: plain x) for c < ncand,
// written to elem_f[c * stride + f_base + e].
hipError_t launch_energy(const TermDev& t, int d, const double* x, const double* dir, const double* steps,
                         int ncand, double* elem_f, int64_t stride, hipStream_t s);

// Energy without library allocations: per-workgroup partial sums of the
// passive energies of every (catalog) term into the static scratch array of
// the device (at most kNoAllocPartials of them), then their fixed-order sum
// into f_out[0] (device).  noalloc_scratch gives the array's address on
// `device` (which must be current).
constexpr int64_t kNoAllocPartials = 65536;
constexpr int kMaxDevices = 64;
hipError_t noalloc_scratch(int device, double** ptr);
int64_t noalloc_partials(const TermDev* terms, int n_terms);
hipError_t launch_energy_noalloc(const TermDev* terms, int n_terms, int d, const double* x, double* scratch_ptr,
                                 double* f_out, hipStream_t s);

// Deterministic sum of ncand rows of n doubles (row stride `stride`) into out[c].
hipError_t launch_sum(const double* in, int64_t n, int ncand, int64_t stride, double* partials, double* out,
                      hipStream_t s);

// Device pattern of the Hessian assembly.
//   Per CSR block b (shapes 0, 1): its row block_row[b], the run
//   [hslot_ptr[m], hslot_ptr[m+1]) of m = mirror[b] (b itself for column >=
//   row, else the upper block (col, row), read transposed).
//   Per upper block i (shapes 2, 3; upper blocks in CSR order): its run
//   [uslot_ptr[i], uslot_ptr[i+1]) and uplace[i] = (offset of entry (0,0) in
//   H, row stride of its block row, the same two for the mirrored lower block
//   or (-1, 0) on the diagonal).
struct HessAsm
{
    int64_t n_blocks = 0;
    const int32_t* block_row = nullptr;
    const int32_t* brow_ptr = nullptr;
    const int64_t* hslot_ptr = nullptr;
    const int32_t* mirror = nullptr;
    const int32_t* row_ptr = nullptr;
    int64_t n_upper = 0;
    const int64_t* uslot_ptr = nullptr;
    const int4* uplace = nullptr;
    // host copies for row ranges: first CSR block of each row (n_rows + 1) and
    // first upper block of each row (n_rows + 1)
    const int32_t* brow_ptr_h = nullptr;
    const int64_t* urow_ptr_h = nullptr;
};

constexpr int kAsmShapes = 4;

// CSR values of H from the Hessian contributions Hc; gradient of handle u =
// sum of slots [gslot_ptr[u], ..) of G.  shape 0: d*d lanes per block, 1: one
// thread per block, 2: d*d lanes per upper block writing it and its mirror,
// 3: one thread per upper block writing both.  Every shape sums each run in
// slot order (identical results).
// r0 >= 0: only the block rows (handles) [r0, r1).
hipError_t launch_assemble_hessian(int d, int precision, int shape, const HessAsm& a, const void* Hc, double* H,
                                   hipStream_t s, int64_t r0 = -1, int64_t r1 = -1);

hipError_t launch_assemble_gradient(int d, int precision, int64_t n_handles, const int64_t* gslot_ptr,
                                    const void* G, double* g, hipStream_t s, int64_t v0 = -1, int64_t v1 = -1);

hipError_t launch_fill(double* p, int64_t n, double v, hipStream_t s);

// x_out = x0 + step * dir  (same rounding as the line-search energy kernel)
hipError_t launch_axpy_point(const double* x0, const double* dir, const double* step, int64_t n, double* x_out,
                             hipStream_t s);

// dot(a, b) -> out (deterministic)
hipError_t launch_dot(const double* a, const double* b, int64_t n, double* partials, double* out, hipStream_t s);

// --- PCG -------------------------------------------------------------------
// Coarse space of the two-level preconditioner (additive correction
// z += P A_c^-1 P^T r on top of block-Jacobi).  P maps kc unknowns per
// aggregate to its handles: d translations and, when kc > d, the rotations
// about the aggregate's first handle (lever arms dvec[v] = x_v - x_root).
struct CoarseDev
{
    int kc = 0;                      // unknowns per aggregate; 0: no coarse correction
    int64_t n_agg = 0, nc = 0;       // aggregates, coarse unknowns (kc * n_agg)
    const int32_t* agg_of = nullptr; // [n_handles]
    const int32_t* agg_ptr = nullptr;  // [n_agg + 1] members agg_vert[agg_ptr[a] .. agg_ptr[a+1]), ascending
    const int32_t* agg_vert = nullptr; // [n_handles]
    // member lists cut into chunks of at most kCoarseChunk handles, chunks of
    // one aggregate consecutive: chunk ch covers agg_vert[chunk_ptr[ch] ..
    // chunk_ptr[ch+1]); aggregate a owns chunks agg_chunk[a] .. agg_chunk[a+1]
    int64_t n_chunks = 0;
    const int32_t* chunk_ptr = nullptr;  // [n_chunks + 1]
    const int32_t* agg_chunk = nullptr;  // [n_agg + 1]
    int max_agg_chunks = 0;              // most chunks of one aggregate
    double* chunk_part = nullptr;        // [n_chunks * kc] restriction partials
    const double* dvec = nullptr;    // [n_handles * d]
    const double* Ainv = nullptr;    // [nc * nc] inverse Galerkin matrix (symmetric, both triangles)
    const float* Ainv32 = nullptr;   // [nc * nc] its FP32 copy for the CG iterations (launch_coarse_to32), or null
    double* rc = nullptr;            // [nc] restricted residual
    double* cparts = nullptr;        // [2 * Gc] partials of r_c . A_c^-1 r_c, slot k & 1
    int Gc = 0;                      // workgroups of the coarse product
    const unsigned char* free_vars = nullptr;  // [n_vars] reduced basis (null: every variable free)
};

constexpr int kCoarseChunk = 256;  // handles per restriction chunk, fine blocks per Galerkin chunk

// Workgroups of the coarse product (one per aggregate), i.e. the r_c . y_c
// partials it writes per CG iteration.
int coarse_apply_groups(int64_t n_agg);

// Lever arms dvec[v] = x_v - x_(first member of v's aggregate).
hipError_t launch_coarse_dvec(int d, int64_t nh, const double* x, const CoarseDev& c, double* dvec, hipStream_t s);

// A_c (column-major nc x nc, zeroed by the caller) = P^T (w I + H) P.  The
// fine CSR blocks of each coarse block pair (a, b) (pair_blocks, CSR order)
// are cut into chunks of at most kCoarseChunk blocks (gchunk_ptr; the chunks
// of pair p are pair_chunk[p] .. pair_chunk[p+1]); one wave sums a chunk,
// then each pair sums its chunk partials (gpart, n_gchunks * kc^2) in order.
hipError_t launch_coarse_galerkin(int d, int64_t n_pairs, int64_t n_gchunks, const int64_t* gchunk_ptr,
                                  const int32_t* pair_chunk, const int32_t* pair_blocks, const int32_t* pair_ab,
                                  const int32_t* block_row, const int32_t* brow_ptr, const int32_t* bcol,
                                  const int32_t* row_ptr, const double* H, double w, const CoarseDev& c,
                                  double* gpart, double* Ac, hipStream_t s);

// A (n x n, column-major, symmetric positive definite) <- A^-1 (symmetric),
// by blocked Gauss-Jordan sweeps with FP64 MFMA products.  work:
// spd_inverse_work(n) doubles; info[0] = 1 when a pivot block is not
// positive definite at working precision (A is then garbage).
int64_t spd_inverse_work(int64_t n);
hipError_t launch_spd_inverse(int64_t n, double* A, double* work, int* info, hipStream_t s);

// A_jj <- A_jj (1 + rel), 1 where A_jj = 0 (column-major n x n).
hipError_t launch_shift_diagonal(int64_t n, double* A, double rel, hipStream_t s);
// *out = ||A||_inf = max_i sum_j |A_ij| of the n x n matrix A (rowsum: n
// doubles of work space); fixed-order sums.
hipError_t launch_abs_rowmax(int64_t n, const double* A, double* rowsum, double* out, hipStream_t s);
// A32 = fl32(A + delta I).  With delta = 2^-23 ||A||_inf the rounding error E
// has ||E||_2 <= ||E||_inf <= 2^-24 (||A||_inf + delta) < delta, so for a
// symmetric positive definite A the difference A32 - A = delta I + E is
// positive definite and so is A32.
hipError_t launch_coarse_to32(int64_t n, const double* A, double delta, float* A32, hipStream_t s);

// Copies the lower triangle of the column-major n x n matrix A to the upper.
hipError_t launch_symmetrize_lower(int64_t n, double* A, hipStream_t s);

// z += P A_c^-1 P^T r (the restriction r_c = P^T r, then the coarse product
// and the prolongation in one launch, with the partials of r_c . A_c^-1 r_c
// into slot `slot`); every kernel returns at once once the solve has
// terminated.
// kdev (optional): the slot is (*kdev) & 1, read on the device.
hipError_t launch_coarse_correction(int d, int64_t nh, const CoarseDev& c, const double* r, double* z, int slot,
                                    PcgScalars* sc, hipStream_t s, Timer* timer, const int* kdev = nullptr);
// free_vars (optional, n bytes, 0 = fixed): CG on the free variables only, the
// reduced basis of newton_direction_reduced_basis (NewtonDirection.hh:76-105)
// for a basis B that selects variables; fixed variables of the solution are 0.
// Dinv: the block-Jacobi inverse, per handle the upper triangle of its
// symmetric d x d block (d (d + 1) / 2 doubles, row by row).
hipError_t launch_pcg_setup(int d, int64_t n_handles, const int32_t* brow_ptr, const int32_t* diag_block,
                            const int32_t* row_ptr, const double* H, double w, const unsigned char* free_vars,
                            double* Dinv, hipStream_t s);

hipError_t launch_pcg_init(int d, int64_t n_handles, const double* b, const unsigned char* free_vars,
                           const double* Dinv, double* x, double* r, double* z, double* partials, PcgScalars* sc,
                           const CoarseDev* coarse, hipStream_t s);

// The solver's copy of H: the n_sym blocks of its lower triangle (diagonal
// included) in CSR order, d*d contiguous row-major values each (sym_full: CSR
// block of each stored block), diagonal blocks symmetrized from their lower
// entries: the solve uses H's scalar lower triangle, as SimplicialLDLT does.
hipError_t launch_sym_pack(int d, int64_t n_sym, const int32_t* sym_full, const int32_t* block_row,
                           const int32_t* brow_ptr, const int32_t* bcol, const int32_t* row_ptr, const double* H,
                           double* sym_H, hipStream_t s);

// One CG iteration k: SpMV on z with the termination test, then (with_update)
// the scalar recurrences and the p / q / x / r / z update, and the coarse
// correction of z when `coarse` has kc > 0.  pbuf holds 2 * n values (p, then
// s = (w I + H) z).  read_k: the kernels take k from sc->k on the device
// instead of the argument (launches recorded once and replayed, with
// launch_pcg_advance after each iteration).  The SpMV reads H from sym_H
// (packed by launch_sym_pack): CSR block b reads stored block sym_src[b], or
// the transpose of ~sym_src[b] when negative.
hipError_t launch_pcg_iteration(int d, int64_t n_handles, const int32_t* brow_ptr, const int32_t* bcol,
                                const int32_t* sym_src, const double* sym_H, double w, const double* Dinv, double* x,
                                double* r, double* z, double* pbuf, double* q, double* partials, PcgScalars* sc, int k,
                                bool with_update, const unsigned char* free_vars, const CoarseDev* coarse,
                                hipStream_t s, Timer* timer, bool read_k = false);
// sc->k += 1 (one lane)
hipError_t launch_pcg_advance(PcgScalars* sc, hipStream_t s);

// General reduced-basis solves (tad_basis.hip): y = A x (+ w add) for a CSR
// matrix of nrows rows; dinv = 1 / diag(B^T H B + w I) with B^T given as CSR
// (flag[0] = 1 where that diagonal is not positive); z = dinv r; y += alpha p,
// r -= alpha q; p = z + beta p.
hipError_t launch_csr_spmv(int64_t nrows, const int32_t* row_ptr, const int32_t* col, const double* val,
                           const double* x, double w, const double* add, double* y, hipStream_t s);
hipError_t launch_basis_diag(int64_t m, const int32_t* bt_ptr, const int32_t* bt_col, const double* bt_val,
                             const int32_t* h_row_ptr, const int32_t* h_col, const double* H, double w, double* dinv,
                             int* flag, hipStream_t s);
hipError_t launch_basis_scale(int64_t m, const double* r, const double* dinv, double* z, hipStream_t s);
hipError_t launch_basis_step(int64_t m, double alpha, const double* p, const double* q, double* y, double* r,
                             hipStream_t s);
hipError_t launch_basis_direction(int64_t m, const double* z, double beta, double* p, hipStream_t s);

// Size (doubles) of the partial-sum buffer the PCG kernels need.
int64_t pcg_parts_size(int d, int64_t n_handles);

// Per-segment energies: ranges[(t * n_seg + s) * 2 + {0,1}] is the element
// range of segment s in term t (absolute indices into elem_f).
hipError_t launch_segment_sum(const double* elem_f, const int64_t* ranges, int n_terms, int64_t n_seg, double* out,
                              hipStream_t s);

// Rest-shape mixed products of a symmetric-Dirichlet tet term, written to
// the term's gradient slots of G (FP64).
hipError_t launch_tet_rest_vjp(const TermDev& t, const double* x, const double* rest, const double* lam,
                               double w_scale, double* G, hipStream_t s);
hipError_t launch_selftest_dual(int which, double x, double y, double* out_dev, hipStream_t s);

} // namespace tad
