// Host side of libtinyad_amd.so: the C ABI declared in include/tinyad_amd.h.
//
// Object model (mirrors include/TinyAD/ScalarFunction.hh:36-236):
//   tad_ctx  -- a device and one HIP stream (the reference's OpenMP team,
//               Detail/Parallel.hh:20-69).
//   tad_fn   -- ScalarFunction: variable dimension d, handle count, objective
//               terms added by tad_fn_add_elements (ScalarObjectiveTerm), plus
//               the device-resident state: connectivity and parameters (SoA),
//               the CSR pattern, the scatter maps (slot of every element
//               block / gradient part), the destination-sorted contribution
//               arrays and the solver work vectors.
//
// Evaluation = element kernel per term (scatters its blocks to their slots)
// -> contiguous per-destination sums, in element order, into g and the CSR
// values -> deterministic FP64 energy sum.
#include "tinyad_amd.h"
#include "tad_internal.hpp"

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

using namespace tad;

namespace {

thread_local std::string g_last_error;

tad_status fail(tad_status st, const std::string& msg)
{
    g_last_error = msg;
    return st;
}

#define TAD_HIP(call)                                                                              \
    do                                                                                             \
    {                                                                                              \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(TAD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));           \
    } while (0)

#define TAD_TRY(call)                                                                              \
    do                                                                                             \
    {                                                                                              \
        tad_status s_ = (call);                                                                    \
        if (s_ != TAD_OK) return s_;                                                               \
    } while (0)

constexpr uint32_t kNoSlotHost = 0xffffffffu;  // slot of a padding or repeated entry (tad::kNoSlot)

int energy_param_count(int energy, int d, int N)
{
    const int K = d * N;
    switch (energy)
    {
        case TAD_E_SYMDIR_TRI: return 5;
        case TAD_E_SYMDIR_TET: return 10;
        case TAD_E_PENALTY: return d + 1;
        case TAD_E_QUADRATIC: return K * K + K + 1;
        default: return -1;
    }
}

template <class T>
struct DevBuf
{
    T* p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t count)
    {
        if (count <= n && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e == hipSuccess) n = count;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

struct TimedPair
{
    std::string name;
    hipEvent_t a, b;
};

struct KernelTimer : Timer
{
    bool enabled = false;
    std::vector<TimedPair> pending;
    std::vector<hipEvent_t> pool;
    std::map<std::string, std::pair<int64_t, double>> stats;
    std::map<std::string, hipEvent_t> open;

    hipEvent_t get()
    {
        if (!pool.empty())
        {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        (void)hipEventCreate(&e);
        return e;
    }
    void begin(const char* name, hipStream_t s) override
    {
        if (!enabled) return;
        hipEvent_t e = get();
        (void)hipEventRecord(e, s);
        auto it = open.find(name);
        if (it != open.end())
        {
            // a begin left open by an error path: recycle its event
            pool.push_back(it->second);
            it->second = e;
        }
        else
            open[name] = e;
    }
    void end(const char* name, hipStream_t s) override
    {
        if (!enabled) return;
        auto it = open.find(name);
        if (it == open.end()) return;
        hipEvent_t e = get();
        (void)hipEventRecord(e, s);
        pending.push_back({ name, it->second, e });
        open.erase(it);
        if (pending.size() > 4096) collect();
    }
    void collect()
    {
        for (auto& p : pending)
        {
            (void)hipEventSynchronize(p.b);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, p.a, p.b);
            auto& st = stats[p.name];
            st.first += 1;
            st.second += ms;
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    ~KernelTimer() override
    {
        collect();
        for (auto& kv : open) (void)hipEventDestroy(kv.second);
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

} // namespace

struct tad_ctx_s
{
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    KernelTimer timer;
    double* pinned = nullptr;  // small pinned staging buffer (256 doubles)
    double* noalloc = nullptr;  // this device's static scratch array (tad_fn_eval_noalloc)
    // second stream and chunk events of pipelined evaluations (created on first use)
    hipStream_t aux = nullptr;
    std::vector<hipEvent_t> events;
    // one reference held by the creator plus one per live function: the
    // context is released when the last of them is gone, whatever the order
    // of tad_ctx_destroy and tad_fn_destroy calls
    int refs = 1;
};

static void ctx_release(tad_ctx ctx)
{
    if (--ctx->refs > 0) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->timer.collect();
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
    if (ctx->aux) (void)hipStreamDestroy(ctx->aux);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
}

struct HostTerm
{
    int energy, N, n_params;
    int64_t n_elem;
    std::vector<int32_t> conn;  // row-major n_elem x N
    bool repeats = false;       // some element lists a handle more than once

    // entry a of element e repeats the handle of an earlier entry (it gets no
    // slot: the element kernel merges it into the first occurrence)
    bool repeated(int64_t e, int a) const
    {
        if (!repeats) return false;
        const int32_t* c = conn.data() + e * N;
        for (int b = 0; b < a; ++b)
            if (c[b] == c[a]) return true;
        return false;
    }
};

struct tad_fn_s
{
    tad_ctx ctx = nullptr;
    int d = 0;
    int64_t n_handles = 0, n_vars = 0, n_elem_total = 0;
    int precision = 0;
    // vector function (residual terms only): residual count and the Jacobian
    // pattern (CSR, n_outputs x n_vars), built on first use
    bool vec = false;
    int64_t n_outputs = 0, nnz_J = 0;
    bool jpattern_ready = false;
    DevBuf<int32_t> jrow_ptr, jcol;
    DevBuf<double> r_buf, J_buf;
    std::vector<TermDev> terms;
    std::vector<HostTerm> hterms;

    // pattern (block rows = handles)
    bool pattern_ready = false;
    int64_t n_blocks = 0, nnz = 0;
    DevBuf<int32_t> brow_ptr, bcol, block_row, diag_block, row_ptr, col_idx;
    // destination-sorted contribution slots: CSR block b owns Hessian slots
    // [hslot_ptr[b], hslot_ptr[b+1]), handle u gradient slots [gslot_ptr[u], ..)
    int64_t n_hslots = 0, n_gslots = 0;
    DevBuf<int64_t> hslot_ptr, gslot_ptr;
    DevBuf<int32_t> block_mirror;  // block whose run assembles this block (itself, or the transposed upper one)
    // per upper block: its run and (H offset, row stride) of it and its mirror
    int64_t n_upper = 0;
    DevBuf<int64_t> uslot_ptr;
    DevBuf<int32_t> uplace;        // 4 per upper block
    // the solver's copy of H: the lower triangle's blocks (diagonal included)
    // in CSR order, D*D contiguous values each, the triangle the reference's
    // SimplicialLDLT reads (LinearSolver.hh:14); per CSR block the stored
    // block it reads (its own ordinal, or ~ordinal of its mirrored lower block)
    DevBuf<int32_t> sym_full;      // per stored block: its CSR block
    DevBuf<int32_t> sym_src;       // per CSR block
    DevBuf<double> sym_H;          // n_sym * D * D
    int64_t n_sym = 0;
    int asm_shape = 0;             // Hessian assembly launch shape (tad_fn_autotune)
    std::vector<int32_t> h_brow_ptr, h_bcol;  // host copy of the block pattern (coarse aggregation)
    std::vector<int64_t> h_urow_ptr;          // first upper block of each block row (host)
    // pipelined second-order evaluation: the largest catalog term's elements
    // in pipe_used chunks on the context's stream; after chunk k the rows
    // [pipe_rows[k], pipe_rows[k+1]) have all their contributions and are
    // assembled on a second stream while the next chunk runs
    int pipe_chunks = 0;       // requested chunks (0: automatic, 1: off)
    bool pipe_ready = false;   // plan below valid for the pattern and request
    int pipe_term = -1;
    int pipe_used = 1;
    std::vector<int64_t> pipe_e, pipe_rows;

    // evaluation buffers
    DevBuf<unsigned char> gc_buf, hc_buf;  // contributions (FP64 or FP32 by precision)
    DevBuf<double> elem_f, x_buf, g_buf, H_buf, partials, results, steps;
    DevBuf<int> err;
    // line search
    DevBuf<double> ls_f, x_tmp;
    // solver
    DevBuf<double> Dinv, r, z, p, q, sol;
    DevBuf<PcgScalars> scal;
    DevBuf<double> pcg_parts;
    bool h_valid = false;
    // segments (independent problems)
    int64_t n_seg = 0;
    std::vector<int64_t> seg_offsets;
    DevBuf<int64_t> seg_ranges;
    DevBuf<double> seg_f;
    // adjoint products (FP64 gradient slots when the contributions are FP32)
    DevBuf<double> adj_g, adj_rest, adj_lam;
    // two-level preconditioner: handle aggregates, coarse block pairs with
    // their runs of fine CSR blocks, the coarse matrix and its inverse
    struct Coarse
    {
        int requested = 0;            // aggregate size asked for when the pattern was built (0: none)
        int target = 0;               // aggregate size the pattern was built with
        int64_t n_agg = 0, n_pairs = 0, n_chunks = 0, n_gchunks = 0;
        int max_agg_chunks = 0;
        DevBuf<int32_t> agg_of, agg_ptr, agg_vert, pair_blocks, pair_ab, chunk_ptr, agg_chunk, pair_chunk;
        DevBuf<int64_t> gchunk_ptr;
        DevBuf<double> dvec, Ac, rc, cparts, chunk_part, gpart, inv_work, rowsum;
        DevBuf<float> Ainv32;  // the inverse in FP32 with a diagonal shift, read by every CG iteration
        bool use32 = false;    // Ainv32 belongs to the current factor
        DevBuf<int> info;
        // the inverse coarse matrix in Ac and the lever arms in dvec belong
        // together and stay usable by a later solve that asks for reuse when
        // it needs the same aggregates, kc and w and neither solve is reduced
        bool factor_valid = false;
        int f_kc = 0;
        double f_w = 0.0;
        void release()
        {
            agg_of.release(); agg_ptr.release(); agg_vert.release(); pair_blocks.release(); pair_ab.release();
            chunk_ptr.release(); agg_chunk.release(); pair_chunk.release(); gchunk_ptr.release();
            dvec.release(); Ac.release(); rc.release(); cparts.release(); chunk_part.release();
            gpart.release(); inv_work.release(); rowsum.release(); Ainv32.release(); info.release();
            requested = target = 0;
            n_agg = n_pairs = n_chunks = n_gchunks = 0;
            max_agg_chunks = 0;
            factor_valid = false;
        }
    } co;
    // positions of the last second-order evaluation (rotation modes of the coarse space)
    DevBuf<double> x_hess;
    DevBuf<unsigned char> free_buf;  // staged free-variable mask of a reduced-basis solve
    // general reduced basis (tad_fn_solve_basis): B (n x m CSR), B^T, and the
    // vectors y r z p q dinv (m each) and t s rhs out (n each)
    DevBuf<int32_t> bas_rp, bas_ci, bas_tp, bas_tc;
    DevBuf<double> bas_v, bas_tv, bas_vec;
    bool x_hess_valid = false;
    int last_solve_iters = 0;  // iterations of the previous solve (first poll of the next one)
    // CG iterations recorded as a HIP graph (tad_fn_set_solve_graphs): the
    // executable graph and the operands it was recorded with
    int solve_graphs = 0;      // 1 on, 0 off
    hipGraphExec_t pcg_graph = nullptr;
    std::vector<uint64_t> pcg_graph_key;
    bool pcg_graph_failed = false;  // recording failed once: launch directly
    int solves_since_rebuild = 0;  // solves that reused the coarse space since it was built

    hipStream_t s() const { return ctx->stream; }
    HessAsm hess_asm() const
    {
        HessAsm a;
        a.n_blocks = n_blocks;
        a.block_row = block_row.p;
        a.brow_ptr = brow_ptr.p;
        a.hslot_ptr = hslot_ptr.p;
        a.mirror = block_mirror.p;
        a.row_ptr = row_ptr.p;
        a.n_upper = n_upper;
        a.uslot_ptr = uslot_ptr.p;
        a.uplace = reinterpret_cast<const int4*>(uplace.p);
        a.brow_ptr_h = h_brow_ptr.data();
        a.urow_ptr_h = h_urow_ptr.data();
        return a;
    }
};

namespace {

tad_status check_fn(tad_fn fn)
{
    if (!fn) return fail(TAD_ERR_INVALID, "null function handle");
    return TAD_OK;
}

void free_pattern(tad_fn fn)
{
    fn->brow_ptr.release(); fn->bcol.release(); fn->block_row.release(); fn->diag_block.release();
    fn->row_ptr.release(); fn->col_idx.release(); fn->hslot_ptr.release(); fn->gslot_ptr.release();
    fn->block_mirror.release(); fn->uslot_ptr.release(); fn->uplace.release();
    fn->sym_full.release(); fn->sym_src.release(); fn->sym_H.release();
    fn->co.release();
    fn->h_brow_ptr.clear(); fn->h_bcol.clear(); fn->h_urow_ptr.clear();
    fn->pipe_ready = false;
    for (TermDev& t : fn->terms)
    {
        if (t.gpos) (void)hipFree(t.gpos);
        if (t.hpos) (void)hipFree(t.hpos);
        t.gpos = t.hpos = nullptr;
        if (t.gpass) (void)hipFree(t.gpass);
        if (t.gout_ptr) (void)hipFree(t.gout_ptr);
        if (t.gout) (void)hipFree(t.gout);
        t.gpass = nullptr;
        t.gout_ptr = t.gout = nullptr;
        t.grouped = t.group_max = 0;
    }
    fn->pattern_ready = false;
    fn->jrow_ptr.release(); fn->jcol.release();
    fn->jpattern_ready = false;
}

// Builds the block CSR pattern, the scalar CSR pattern and the element-order
// gather maps.  Runs once per function (after the last add_elements).
tad_status build_pattern(tad_fn fn)
{
    const int d = fn->d;
    const int64_t nh = fn->n_handles;
    // 1. adjacency (with duplicates) per handle
    std::vector<int64_t> cnt(nh + 1, 0);
    for (const HostTerm& t : fn->hterms)
        for (int64_t e = 0; e < t.n_elem; ++e)
            for (int a = 0; a < t.N; ++a)
                if (t.conn[e * t.N + a] >= 0) cnt[t.conn[e * t.N + a] + 1] += t.N;
    for (int64_t u = 0; u < nh; ++u) cnt[u + 1] += cnt[u];
    std::vector<int32_t> adj(cnt[nh]);
    {
        std::vector<int64_t> cur(cnt.begin(), cnt.end() - 1);
        for (const HostTerm& t : fn->hterms)
            for (int64_t e = 0; e < t.n_elem; ++e)
                for (int a = 0; a < t.N; ++a)
                {
                    const int32_t u = t.conn[e * t.N + a];
                    if (u < 0) continue;
                    for (int b = 0; b < t.N; ++b) adj[cur[u]++] = t.conn[e * t.N + b];
                }
    }
    // handles used by no element keep empty rows, as in the reference's
    // triplet assembly; their diagonal block index is -1
    std::vector<int32_t> brow_ptr(nh + 1, 0), bcol, diag(nh, -1);
    bcol.reserve(adj.size() / 3 + nh);
    for (int64_t u = 0; u < nh; ++u)
    {
        auto b = adj.begin() + cnt[u], e = adj.begin() + cnt[u + 1];
        std::sort(b, e);
        e = std::unique(b, e);
        while (b != e && *b < 0) ++b;  // padding entries of dynamic-valence elements
        for (auto it = b; it != e; ++it) bcol.push_back(*it);
        brow_ptr[u + 1] = (int32_t)bcol.size();
        for (int32_t k = brow_ptr[u]; k < brow_ptr[u + 1]; ++k)
            if (bcol[k] == u) diag[u] = k - brow_ptr[u];
    }
    const int64_t nb = (int64_t)bcol.size();
    if ((double)nb * d * d >= 2147483647.0)
        return fail(TAD_ERR_UNSUPPORTED, "Hessian has more than 2^31 non-zeros");
    std::vector<int32_t> block_row(nb);
    for (int64_t u = 0; u < nh; ++u)
        for (int32_t k = brow_ptr[u]; k < brow_ptr[u + 1]; ++k) block_row[k] = (int32_t)u;

    auto find_block = [&](int32_t u, int32_t v) -> int64_t {
        const int32_t* b = bcol.data() + brow_ptr[u];
        const int32_t* e = bcol.data() + brow_ptr[u + 1];
        return (int64_t)(std::lower_bound(b, e, v) - bcol.data());
    };

    // 2. contribution slots: per destination a contiguous run, ordered by
    //    term and element.  Hessian runs exist for the upper CSR blocks
    //    (column >= row); an element writes each of its N(N+1)/2 blocks once,
    //    in the orientation of that upper block (top bit of its slot: write
    //    the transpose), and a lower block is assembled from its mirror's run.
    //    Gradient runs per handle.  Per element the slot of each block and
    //    gradient part (SoA, coalesced in the element kernel).
    //    Catalog terms are summed per group of kElemGroup consecutive
    //    elements first (in LDS, by the element kernel): such a term has one
    //    slot per distinct upper block of each group.  Within a group the
    //    blocks are added round by round (local block (a, b) in order) and
    //    within a round in passes, so that no two lanes of the wave add to
    //    the same accumulator at once; the plan below numbers each group's
    //    blocks and each contribution's pass.  A term whose groups touch more
    //    than kGroupMaxBlocks blocks keeps one slot per element block.
    const size_t nterms = fn->hterms.size();
    std::vector<char> grouped(nterms, 0);
    std::vector<int> gmax(nterms, 0);
    std::vector<std::vector<uint32_t>> hcode(nterms), gofs(nterms);
    std::vector<std::vector<uint8_t>> gpass(nterms);
    std::vector<std::vector<int32_t>> gblock(nterms);  // per group slot: its upper CSR block
    {
        std::vector<int64_t> stamp(nb, -1), rstamp(nb, -1);
        std::vector<int32_t> gsof(nb, 0), rcnt(nb, 0);
        int64_t gid = 0;
        for (size_t ti = 0; ti < nterms; ++ti)
        {
            const HostTerm& t = fn->hterms[ti];
            if (fn->terms[ti].dev_ops.element || t.N < 2 || t.n_elem == 0 || t.repeats) continue;
            const int64_t ne = t.n_elem;
            const int NB = t.N * (t.N + 1) / 2;
            const int64_t ng = (ne + kElemGroup - 1) / kElemGroup;
            std::vector<uint32_t> code((size_t)ne * NB);
            std::vector<uint8_t> np_((size_t)ng * NB, 0);
            std::vector<uint32_t> go((size_t)ng + 1, 0);
            std::vector<int32_t> gb;
            bool ok = true;
            int gm = 0;
            for (int64_t q = 0; q < ng && ok; ++q, ++gid)
            {
                const int64_t e0 = q * kElemGroup, e1 = std::min<int64_t>(ne, e0 + kElemGroup);
                const size_t g0 = gb.size();
                int bi = 0;
                for (int a = 0; a < t.N; ++a)
                    for (int b = a; b < t.N; ++b, ++bi)
                    {
                        const int64_t rkey = gid * 16 + bi;
                        int maxp = -1;
                        for (int64_t e = e0; e < e1; ++e)
                        {
                            const int32_t u = t.conn[e * t.N + a], v = t.conn[e * t.N + b];
                            const int64_t kb = find_block(std::min(u, v), std::max(u, v));
                            if (stamp[kb] != gid)
                            {
                                stamp[kb] = gid;
                                gsof[kb] = (int32_t)(gb.size() - g0);
                                gb.push_back((int32_t)kb);
                            }
                            if (rstamp[kb] != rkey)
                            {
                                rstamp[kb] = rkey;
                                rcnt[kb] = 0;
                            }
                            const int pass = rcnt[kb]++;
                            maxp = std::max(maxp, pass);
                            code[(size_t)bi * ne + e] =
                                (uint32_t)gsof[kb] | ((uint32_t)(pass & 0xff) << 16) | (u > v ? 0x80000000u : 0u);
                        }
                        if (maxp > 255) ok = false;
                        np_[(size_t)q * NB + bi] = (uint8_t)(maxp + 1);
                    }
                const int G = (int)(gb.size() - g0);
                if (G > kGroupMaxBlocks) ok = false;
                gm = std::max(gm, G);
                go[q + 1] = (uint32_t)gb.size();
            }
            if (!ok) continue;
            grouped[ti] = 1;
            gmax[ti] = gm;
            hcode[ti] = std::move(code);
            gpass[ti] = std::move(np_);
            gofs[ti] = std::move(go);
            gblock[ti] = std::move(gb);
        }
    }
    std::vector<int64_t> hptr(nb + 1, 0);
    std::vector<int64_t> gptr(nh + 1, 0);
    for (size_t ti = 0; ti < nterms; ++ti)
    {
        const HostTerm& t = fn->hterms[ti];
        if (grouped[ti])
            for (int32_t kb : gblock[ti]) hptr[kb + 1] += 1;
        for (int64_t e = 0; e < t.n_elem; ++e)
            for (int a = 0; a < t.N; ++a)
            {
                const int32_t u = t.conn[e * t.N + a];
                if (u < 0 || t.repeated(e, a)) continue;
                gptr[u + 1] += 1;
                if (grouped[ti]) continue;
                for (int b = a; b < t.N; ++b)
                {
                    const int32_t v = t.conn[e * t.N + b];
                    if (v < 0 || t.repeated(e, b)) continue;
                    hptr[find_block(std::min(u, v), std::max(u, v)) + 1] += 1;
                }
            }
    }
    for (int64_t k = 0; k < nb; ++k) hptr[k + 1] += hptr[k];
    for (int64_t u = 0; u < nh; ++u) gptr[u + 1] += gptr[u];
    // slot 2^31 - 1 (with the transpose flag: kNoSlotHost) stays unused
    if (hptr[nb] >= 2147483647LL || gptr[nh] >= 4294967295LL)
        return fail(TAD_ERR_UNSUPPORTED, "more than 2^31 element contributions");
    std::vector<int32_t> mirror(nb);
    for (int64_t u = 0; u < nh; ++u)
        for (int32_t k = brow_ptr[u]; k < brow_ptr[u + 1]; ++k)
            mirror[k] = bcol[k] >= u ? k : (int32_t)find_block(bcol[k], (int32_t)u);
    // the solver's lower-triangle copy: ordinals of the lower blocks, and per
    // block the stored block it reads, checked against the pattern (mirror
    // maps each lower block to its upper block; lower_of inverts it)
    std::vector<int32_t> sym_full, sym_src(nb);
    {
        std::vector<int32_t> ord(nb, -1), lower_of(nb, -1);
        for (int64_t u = 0; u < nh; ++u)
            for (int32_t k = brow_ptr[u]; k < brow_ptr[u + 1]; ++k)
                if (bcol[k] <= u)
                {
                    ord[k] = (int32_t)sym_full.size();
                    sym_full.push_back(k);
                    const int32_t m = mirror[k];
                    const bool ok = bcol[k] == u ? (m == k)
                                                 : (m >= 0 && m < nb && block_row[m] == bcol[k] &&
                                                    bcol[m] == (int32_t)u && lower_of[m] < 0);
                    if (!ok) return fail(TAD_ERR_UNSUPPORTED, "Hessian mirror map inconsistent");
                    lower_of[m] = k;
                }
        for (int64_t u = 0; u < nh; ++u)
            for (int32_t k = brow_ptr[u]; k < brow_ptr[u + 1]; ++k)
            {
                if (bcol[k] <= u)
                {
                    sym_src[k] = ord[k];
                    continue;
                }
                const int32_t l = lower_of[k];
                if (l < 0 || ord[l] < 0) return fail(TAD_ERR_UNSUPPORTED, "Hessian pattern not symmetric");
                sym_src[k] = ~ord[l];
            }
    }
    std::vector<std::vector<uint32_t>> gpos(nterms), hpos(nterms), gout(nterms);
    {
        std::vector<int64_t> hc(hptr.begin(), hptr.end() - 1), gc(gptr.begin(), gptr.end() - 1);
        for (size_t ti = 0; ti < nterms; ++ti)
        {
            const HostTerm& t = fn->hterms[ti];
            const int64_t ne = t.n_elem;
            const int NB = t.N * (t.N + 1) / 2;
            gpos[ti].resize((size_t)ne * t.N);
            if (grouped[ti])
            {
                // group slots in group order, each group's in its own order
                gout[ti].resize(gblock[ti].size());
                for (size_t i = 0; i < gblock[ti].size(); ++i) gout[ti][i] = (uint32_t)hc[gblock[ti][i]]++;
                hpos[ti] = std::move(hcode[ti]);
            }
            else
                hpos[ti].resize((size_t)ne * NB);
            for (int64_t e = 0; e < ne; ++e)
                for (int a = 0; a < t.N; ++a)
                {
                    const int32_t u = t.conn[e * t.N + a];
                    const bool skip_a = u < 0 || t.repeated(e, a);
                    gpos[ti][a * ne + e] = skip_a ? kNoSlotHost : (uint32_t)gc[u]++;
                    if (grouped[ti]) continue;
                    for (int b = a; b < t.N; ++b)
                    {
                        const int32_t v = t.conn[e * t.N + b];
                        const int bi = a * t.N - (a * (a - 1)) / 2 + (b - a);
                        if (skip_a || v < 0 || t.repeated(e, b))
                        {
                            hpos[ti][(size_t)bi * ne + e] = kNoSlotHost;
                            continue;
                        }
                        const uint32_t flag = (u > v) ? 0x80000000u : 0u;
                        hpos[ti][(size_t)bi * ne + e] =
                            (uint32_t)hc[find_block(std::min(u, v), std::max(u, v))]++ | flag;
                    }
                }
        }
    }
    // 3. scalar CSR
    const int64_t n = fn->n_vars;
    std::vector<int32_t> row_ptr(n + 1, 0);
    for (int64_t u = 0; u < nh; ++u)
    {
        const int nbu = brow_ptr[u + 1] - brow_ptr[u];
        for (int i = 0; i < d; ++i) row_ptr[d * u + i + 1] = row_ptr[d * u + i] + d * nbu;
    }
    const int64_t nnz = row_ptr[n];
    std::vector<int32_t> col_idx(nnz);
    for (int64_t u = 0; u < nh; ++u)
    {
        const int nbu = brow_ptr[u + 1] - brow_ptr[u];
        for (int i = 0; i < d; ++i)
        {
            int64_t o = row_ptr[d * u + i];
            for (int k = 0; k < nbu; ++k)
                for (int j = 0; j < d; ++j) col_idx[o++] = d * bcol[brow_ptr[u] + k] + j;
        }
    }
    // 4. upper-run placement: run i of upper block (u, v) lands at (u, v) and,
    //    transposed, at (v, u); every H entry must be written exactly once
    std::vector<int64_t> uptr;
    std::vector<int32_t> uplace;
    uptr.reserve((nb + nh) / 2 + 1);
    uplace.reserve(2 * (nb + nh) + 4);
    {
        std::vector<uint8_t> hits(nnz, 0);
        auto place = [&](int64_t off, int64_t stride, bool tr) -> bool {
            for (int i = 0; i < d; ++i)
                for (int j = 0; j < d; ++j)
                {
                    const int64_t q = off + (tr ? (int64_t)j * stride + i : (int64_t)i * stride + j);
                    if (q < 0 || q >= nnz || hits[q]++) return false;
                }
            return true;
        };
        for (int64_t u = 0; u < nh; ++u)
        {
            const int32_t nbu = brow_ptr[u + 1] - brow_ptr[u];
            for (int32_t k = brow_ptr[u]; k < brow_ptr[u + 1]; ++k)
            {
                const int32_t v = bcol[k];
                if (v < u) continue;
                const int64_t off_up = row_ptr[d * u] + (int64_t)d * (k - brow_ptr[u]);
                int64_t off_lo = -1, stride_lo = 0;
                if (v != u)
                {
                    const int64_t lb = find_block(v, (int32_t)u);
                    off_lo = row_ptr[d * v] + (int64_t)d * (lb - brow_ptr[v]);
                    stride_lo = (int64_t)d * (brow_ptr[v + 1] - brow_ptr[v]);
                }
                if (!place(off_up, (int64_t)d * nbu, false) || (off_lo >= 0 && !place(off_lo, stride_lo, true)))
                    return fail(TAD_ERR_UNSUPPORTED, "Hessian placement map overlaps");
                uptr.push_back(hptr[k]);
                uplace.push_back((int32_t)off_up);
                uplace.push_back(d * nbu);
                uplace.push_back((int32_t)off_lo);
                uplace.push_back((int32_t)stride_lo);
            }
        }
        for (int64_t q = 0; q < nnz; ++q)
            if (hits[q] != 1) return fail(TAD_ERR_UNSUPPORTED, "Hessian placement map misses an entry");
        uptr.push_back(hptr[nb]);
        // runs of lower blocks are empty, so consecutive upper runs tile the slots
        for (size_t i = 0; i + 1 < uptr.size(); ++i)
            if (uptr[i] > uptr[i + 1]) return fail(TAD_ERR_UNSUPPORTED, "upper runs out of order");
    }
    hipStream_t s = fn->s();
    auto up = [&](auto& buf, const auto& vec) -> hipError_t {
        hipError_t e = buf.ensure(vec.size());
        if (e != hipSuccess) return e;
        if (vec.empty()) return hipSuccess;
        return hipMemcpyAsync(buf.p, vec.data(), vec.size() * sizeof(vec[0]), hipMemcpyHostToDevice, s);
    };
    TAD_HIP(up(fn->brow_ptr, brow_ptr));
    TAD_HIP(up(fn->bcol, bcol));
    TAD_HIP(up(fn->block_row, block_row));
    TAD_HIP(up(fn->diag_block, diag));
    TAD_HIP(up(fn->row_ptr, row_ptr));
    TAD_HIP(up(fn->col_idx, col_idx));
    TAD_HIP(up(fn->hslot_ptr, hptr));
    TAD_HIP(up(fn->gslot_ptr, gptr));
    TAD_HIP(up(fn->block_mirror, mirror));
    TAD_HIP(up(fn->uslot_ptr, uptr));
    TAD_HIP(up(fn->uplace, uplace));
    TAD_HIP(up(fn->sym_full, sym_full));
    TAD_HIP(up(fn->sym_src, sym_src));
    for (size_t ti = 0; ti < fn->terms.size(); ++ti)
    {
        TermDev& td = fn->terms[ti];
        if (td.n_elem == 0) continue;
        TAD_HIP(hipMalloc((void**)&td.gpos, gpos[ti].size() * sizeof(uint32_t)));
        TAD_HIP(hipMalloc((void**)&td.hpos, hpos[ti].size() * sizeof(uint32_t)));
        TAD_HIP(hipMemcpyAsync(td.gpos, gpos[ti].data(), gpos[ti].size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        TAD_HIP(hipMemcpyAsync(td.hpos, hpos[ti].data(), hpos[ti].size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        td.grouped = grouped[ti];
        td.group_max = gmax[ti];
        if (grouped[ti])
        {
            TAD_HIP(hipMalloc((void**)&td.gpass, gpass[ti].size()));
            TAD_HIP(hipMalloc((void**)&td.gout_ptr, gofs[ti].size() * sizeof(uint32_t)));
            TAD_HIP(hipMalloc((void**)&td.gout, std::max<size_t>(1, gout[ti].size()) * sizeof(uint32_t)));
            TAD_HIP(hipMemcpyAsync(td.gpass, gpass[ti].data(), gpass[ti].size(), hipMemcpyHostToDevice, s));
            TAD_HIP(hipMemcpyAsync(td.gout_ptr, gofs[ti].data(), gofs[ti].size() * sizeof(uint32_t),
                                   hipMemcpyHostToDevice, s));
            TAD_HIP(hipMemcpyAsync(td.gout, gout[ti].data(), gout[ti].size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                                   s));
        }
    }
    TAD_HIP(hipStreamSynchronize(s));
    fn->h_brow_ptr = std::move(brow_ptr);
    fn->h_bcol = std::move(bcol);
    fn->h_urow_ptr.assign((size_t)nh + 1, 0);
    for (int64_t u = 0; u < nh; ++u)
    {
        int64_t c = 0;
        for (int32_t k = fn->h_brow_ptr[u]; k < fn->h_brow_ptr[u + 1]; ++k) c += fn->h_bcol[k] >= u ? 1 : 0;
        fn->h_urow_ptr[u + 1] = fn->h_urow_ptr[u] + c;
    }
    fn->n_hslots = hptr[nb];
    fn->n_gslots = gptr[nh];
    fn->n_blocks = nb;
    fn->n_upper = (int64_t)uptr.size() - 1;
    fn->n_sym = (int64_t)sym_full.size();
    fn->nnz = nnz;
    fn->pattern_ready = true;
    return TAD_OK;
}

// Aggregates of handles for the two-level preconditioner, grown on the block
// graph: the lowest unaggregated handle starts an aggregate, which takes
// unaggregated neighbours breadth first until it holds `target` handles;
// aggregates below target/4 are then merged into the neighbouring aggregate
// they share the most couplings with (lowest index on ties), and aggregates
// are renumbered in order.  Coarse block pairs (a, b) with their runs of fine
// CSR blocks (CSR order) give the Galerkin sums.  Deterministic; once per
// function and target.
tad_status build_coarse(tad_fn fn, int target)
{
    const int64_t nh = fn->n_handles;
    const std::vector<int32_t>& rp = fn->h_brow_ptr;
    const std::vector<int32_t>& cb = fn->h_bcol;
    std::vector<int32_t> agg(nh, -1);
    int32_t na = 0;
    std::vector<int32_t> queue;
    queue.reserve(target);
    for (int64_t root = 0; root < nh; ++root)
    {
        if (agg[root] >= 0) continue;
        queue.clear();
        queue.push_back((int32_t)root);
        agg[root] = na;
        for (size_t head = 0; head < queue.size() && (int)queue.size() < target; ++head)
        {
            const int32_t u = queue[head];
            for (int32_t k = rp[u]; k < rp[u + 1] && (int)queue.size() < target; ++k)
            {
                const int32_t v = cb[k];
                if (agg[v] < 0)
                {
                    agg[v] = na;
                    queue.push_back(v);
                }
            }
        }
        ++na;
    }
    std::vector<int64_t> size(na, 0);
    for (int64_t v = 0; v < nh; ++v) size[agg[v]]++;
    const int64_t small_lim = std::max(1, target / 4);
    std::vector<std::vector<int32_t>> members(na);
    for (int64_t v = 0; v < nh; ++v) members[agg[v]].push_back((int32_t)v);
    std::vector<int32_t> merged(na);
    for (int32_t a = 0; a < na; ++a) merged[a] = a;
    for (int32_t a = 0; a < na; ++a)
    {
        if (size[a] >= small_lim) continue;
        std::map<int32_t, int64_t> links;  // neighbouring large aggregate -> couplings
        for (int32_t u : members[a])
            for (int32_t k = rp[u]; k < rp[u + 1]; ++k)
            {
                const int32_t b = agg[cb[k]];
                if (b != a && size[b] >= small_lim) links[b]++;
            }
        int32_t best = -1;
        int64_t bl = 0;
        for (const auto& kv : links)
            if (kv.second > bl)
            {
                best = kv.first;
                bl = kv.second;
            }
        if (best >= 0) merged[a] = best;  // a small component on its own stays an aggregate
    }
    std::vector<int32_t> renum(na, -1);
    int32_t nn = 0;
    for (int32_t a = 0; a < na; ++a)
        if (merged[a] == a) renum[a] = nn++;
    std::vector<int32_t> agg_of(nh);
    for (int64_t v = 0; v < nh; ++v) agg_of[v] = renum[merged[agg[v]]];
    std::vector<int32_t> agg_ptr(nn + 1, 0), agg_vert(nh);
    for (int64_t v = 0; v < nh; ++v) agg_ptr[agg_of[v] + 1]++;
    for (int32_t a = 0; a < nn; ++a) agg_ptr[a + 1] += agg_ptr[a];
    {
        std::vector<int32_t> fill(agg_ptr.begin(), agg_ptr.end() - 1);
        for (int64_t v = 0; v < nh; ++v) agg_vert[fill[agg_of[v]]++] = (int32_t)v;  // ascending per aggregate
    }
    // coarse block pairs and their runs of fine blocks
    const int64_t nb = (int64_t)cb.size();
    std::vector<std::pair<int64_t, int32_t>> keyed(nb);
    for (int64_t u = 0; u < nh; ++u)
        for (int32_t k = rp[u]; k < rp[u + 1]; ++k)
            keyed[k] = { (int64_t)agg_of[u] * nn + agg_of[cb[k]], k };
    std::sort(keyed.begin(), keyed.end());
    std::vector<int64_t> pair_ptr(1, 0);
    std::vector<int32_t> pair_ab, pair_blocks(nb);
    for (int64_t i = 0; i < nb; ++i)
    {
        if (i == 0 || keyed[i].first != keyed[i - 1].first)
        {
            if (i > 0) pair_ptr.push_back(i);
            pair_ab.push_back((int32_t)(keyed[i].first / nn));
            pair_ab.push_back((int32_t)(keyed[i].first % nn));
        }
        pair_blocks[i] = keyed[i].second;
    }
    pair_ptr.push_back(nb);
    // chunks of at most kCoarseChunk handles / fine blocks (parallel sums)
    std::vector<int32_t> chunk_ptr, agg_chunk(nn + 1, 0);
    for (int32_t a = 0; a < nn; ++a)
    {
        agg_chunk[a] = (int32_t)chunk_ptr.size();
        for (int32_t o = agg_ptr[a]; o < agg_ptr[a + 1]; o += kCoarseChunk) chunk_ptr.push_back(o);
    }
    agg_chunk[nn] = (int32_t)chunk_ptr.size();
    chunk_ptr.push_back((int32_t)nh);
    const int64_t n_pairs = (int64_t)pair_ab.size() / 2;
    std::vector<int64_t> gchunk_ptr;
    std::vector<int32_t> pair_chunk(n_pairs + 1, 0);
    for (int64_t q = 0; q < n_pairs; ++q)
    {
        pair_chunk[q] = (int32_t)gchunk_ptr.size();
        for (int64_t o = pair_ptr[q]; o < pair_ptr[q + 1]; o += kCoarseChunk) gchunk_ptr.push_back(o);
    }
    pair_chunk[n_pairs] = (int32_t)gchunk_ptr.size();
    gchunk_ptr.push_back(nb);
    auto& c = fn->co;
    c.release();
    hipStream_t s = fn->s();
    auto up = [&](auto& buf, const auto& vec) -> hipError_t {
        hipError_t e = buf.ensure(vec.size());
        if (e != hipSuccess || vec.empty()) return e;
        return hipMemcpyAsync(buf.p, vec.data(), vec.size() * sizeof(vec[0]), hipMemcpyHostToDevice, s);
    };
    TAD_HIP(up(c.agg_of, agg_of));
    TAD_HIP(up(c.agg_ptr, agg_ptr));
    TAD_HIP(up(c.agg_vert, agg_vert));
    TAD_HIP(up(c.pair_blocks, pair_blocks));
    TAD_HIP(up(c.pair_ab, pair_ab));
    TAD_HIP(up(c.chunk_ptr, chunk_ptr));
    TAD_HIP(up(c.agg_chunk, agg_chunk));
    TAD_HIP(up(c.gchunk_ptr, gchunk_ptr));
    TAD_HIP(up(c.pair_chunk, pair_chunk));
    TAD_HIP(hipStreamSynchronize(s));
    c.n_agg = nn;
    c.n_pairs = n_pairs;
    c.n_chunks = (int64_t)chunk_ptr.size() - 1;
    c.n_gchunks = (int64_t)gchunk_ptr.size() - 1;
    for (int32_t a = 0; a < nn; ++a) c.max_agg_chunks = std::max(c.max_agg_chunks, agg_chunk[a + 1] - agg_chunk[a]);
    c.target = target;
    return TAD_OK;
}

// Largest coarse matrix the two-level preconditioner builds (dense FP64 inverse).
constexpr int64_t kMaxCoarse = 8192;
// Automatic aggregate size (tad_solve_opts.coarse_aggregate < 0): about
// kAutoCoarse coarse unknowns up to kAutoHandles handles, growing with the
// cube root of the handle count above, to at most 2 kAutoCoarse.
constexpr int64_t kAutoCoarse = 1400;
constexpr double kAutoHandles = 30000.0;
constexpr int64_t kAutoMinAggregate = 32;
// Largest relative change of the coarse correction accepted from its FP32 copy.
constexpr double kCoarse32Change = 0.1;

// Per solve: coarse unknowns per aggregate, lever arms, A_c = P^T (w I + H) P,
// its inverse (blocked Gauss-Jordan sweeps with FP64 MFMA, tad_dense.hip;
// fixed-order sums, so the result is deterministic), symmetrized.
// *use = false when the coarse space is off or A_c is not numerically
// positive definite (block-Jacobi alone then).  With o.coarse_reuse the
// inverse and lever arms of the previous compatible solve are kept (*reused).
tad_status setup_coarse(tad_fn fn, const double* H_dev, const tad_solve_opts& o, const unsigned char* free_dev,
                        CoarseDev* cd, bool* use, bool* reused)
{
    *use = false;
    *reused = false;
    const int d = fn->d;
    const int64_t nh = fn->n_handles;
    if (o.coarse_aggregate == 0 || nh == 0 || fn->n_blocks == 0) return TAD_OK;
    const bool rot = fn->x_hess_valid && d >= 2;
    const int kc = rot ? (d == 3 ? 6 : 3) : d;
    // automatic size (measured optimum, ms per Newton iteration with the
    // coarse inverse rebuilt every 5th: 30^3 cube 1,404 unknowns 3.95 against
    // 2,790 4.47; 40^3 1,380 7.25 against 2,778 7.41; 60^3 1,404 20.7,
    // 2,526 19.4, 2,808 19.0, 3,186 21.0), at least 32 handles per aggregate
    const double nc_auto = std::min(2.0 * kAutoCoarse,
                                    kAutoCoarse * std::cbrt(std::max(1.0, (double)nh / kAutoHandles)));
    const int64_t agg = o.coarse_aggregate > 0
                            ? o.coarse_aggregate
                            : std::max<int64_t>(kAutoMinAggregate, (int64_t)std::ceil((double)(nh * kc) / nc_auto));
    // keep the dense coarse matrix bounded: larger aggregates for large meshes
    // (grown further while breadth-first aggregates come out too small)
    const int want = (int)std::max<int64_t>(agg, (nh * kc + kMaxCoarse - 1) / kMaxCoarse);
    auto& c = fn->co;
    if (c.requested != want)
    {
        c.factor_valid = false;
        int target = want;
        for (;;)
        {
            TAD_TRY(build_coarse(fn, target));
            if (kc * c.n_agg <= kMaxCoarse || target >= nh) break;
            target = (int)std::min<int64_t>(nh, (int64_t)target + target / 2 + 1);
        }
        c.requested = want;
    }
    if (c.n_agg < 2) return TAD_OK;  // one aggregate: nothing a coarse space adds at this size
    const int64_t nc = kc * c.n_agg;
    if (nc > kMaxCoarse) return TAD_OK;
    hipStream_t s = fn->s();
    KernelTimer& tm = fn->ctx->timer;
    TAD_HIP(c.dvec.ensure((size_t)nh * d));
    TAD_HIP(c.Ac.ensure((size_t)(nc * nc)));
    TAD_HIP(c.rc.ensure((size_t)nc));
    const int Gc = coarse_apply_groups(c.n_agg);
    TAD_HIP(c.cparts.ensure((size_t)2 * Gc));
    TAD_HIP(c.info.ensure(2));
    TAD_HIP(c.chunk_part.ensure((size_t)c.n_chunks * kc));
    TAD_HIP(c.gpart.ensure((size_t)c.n_gchunks * kc * kc));
    cd->kc = kc;
    cd->n_agg = c.n_agg;
    cd->nc = nc;
    cd->agg_of = c.agg_of.p;
    cd->agg_ptr = c.agg_ptr.p;
    cd->agg_vert = c.agg_vert.p;
    cd->n_chunks = c.n_chunks;
    cd->chunk_ptr = c.chunk_ptr.p;
    cd->agg_chunk = c.agg_chunk.p;
    cd->max_agg_chunks = c.max_agg_chunks;
    cd->chunk_part = c.chunk_part.p;
    cd->dvec = c.dvec.p;
    cd->Ainv = c.Ac.p;
    cd->Ainv32 = c.use32 ? c.Ainv32.p : nullptr;
    cd->rc = c.rc.p;
    cd->cparts = c.cparts.p;
    cd->Gc = Gc;
    cd->free_vars = free_dev;
    if (o.coarse_reuse && c.factor_valid && !free_dev && c.f_kc == kc && c.f_w == o.w_identity)
    {
        *use = true;
        *reused = true;
        return TAD_OK;
    }
    c.factor_valid = false;
    c.use32 = false;
    cd->Ainv32 = nullptr;
    tm.begin("coarse_setup", s);
    tm.begin("coarse_galerkin", s);
    if (rot) TAD_HIP(launch_coarse_dvec(d, nh, fn->x_hess.p, *cd, c.dvec.p, s));
    else TAD_HIP(launch_fill(c.dvec.p, nh * d, 0.0, s));
    TAD_HIP(launch_fill(c.Ac.p, nc * nc, 0.0, s));
    TAD_HIP(launch_coarse_galerkin(d, c.n_pairs, c.n_gchunks, c.gchunk_ptr.p, c.pair_chunk.p, c.pair_blocks.p,
                                   c.pair_ab.p, fn->block_row.p, fn->brow_ptr.p, fn->bcol.p, fn->row_ptr.p, H_dev,
                                   o.w_identity, *cd, c.gpart.p, c.Ac.p, s));
    // reduced basis: coarse directions without free variables leave A_c only
    // semidefinite; a relative diagonal shift makes it definite (those
    // directions prolong to zero)
    if (free_dev) TAD_HIP(launch_shift_diagonal(nc, c.Ac.p, 1e-10, s));
    tm.end("coarse_galerkin", s);
    TAD_HIP(c.inv_work.ensure((size_t)spd_inverse_work(nc)));
    TAD_HIP(c.rowsum.ensure((size_t)nc + 2));  // row sums, then ||A_c||_inf, ||A_c^-1||_inf
    TAD_HIP(launch_abs_rowmax(nc, c.Ac.p, c.rowsum.p, c.rowsum.p + nc, s));
    tm.begin("coarse_inverse", s);
    TAD_HIP(launch_spd_inverse(nc, c.Ac.p, c.inv_work.p, c.info.p, s));
    tm.end("coarse_inverse", s);
    TAD_HIP(launch_abs_rowmax(nc, c.Ac.p, c.rowsum.p, c.rowsum.p + nc + 1, s));
    int* info = reinterpret_cast<int*>(fn->ctx->pinned + 96);  // past the solver scalars at 56..
    double* norms = reinterpret_cast<double*>(fn->ctx->pinned + 104);
    TAD_HIP(hipMemcpyAsync(info, c.info.p, sizeof(int), hipMemcpyDeviceToHost, s));
    TAD_HIP(hipMemcpyAsync(norms, c.rowsum.p + nc, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    TAD_HIP(hipStreamSynchronize(s));
    if (info[0] != 0)
    {
        tm.end("coarse_setup", s);
        cd->kc = 0;
        return TAD_OK;  // not positive definite at working precision: block-Jacobi alone
    }
    // The CG iterations read an FP32 copy with the diagonal shift
    // delta = 2^-23 ||A_c^-1||_inf that keeps it positive definite
    // (launch_coarse_to32) when the shift is small against the smallest
    // eigenvalue of A_c^-1, 1 / lambda_max(A_c) >= 1 / ||A_c||_inf:
    // 2^-22 kappa_inf <= 0.1 bounds the relative change of the coarse
    // correction by 10 %.  Otherwise (e.g. a Hessian that is singular up to
    // the projection's eps along rigid motions) they read the FP64 inverse.
    const double kappa = norms[0] * norms[1];
    if (std::isfinite(kappa) && std::ldexp(kappa, -22) <= kCoarse32Change)
    {
        TAD_HIP(c.Ainv32.ensure((size_t)(nc * nc)));
        TAD_HIP(launch_coarse_to32(nc, c.Ac.p, std::ldexp(norms[1], -23), c.Ainv32.p, s));
        c.use32 = true;
        cd->Ainv32 = c.Ainv32.p;
    }
    tm.end("coarse_setup", s);
    *use = true;
    c.factor_valid = !free_dev;
    c.f_kc = kc;
    c.f_w = o.w_identity;
    return TAD_OK;
}

tad_status ensure_ready(tad_fn fn)
{
    if (fn->terms.empty() && fn->n_elem_total == 0)
    {
        // no elements: f = 0, g = 0, empty H (ScalarFunctionTest.cc Move1 default case)
    }
    if (!fn->pattern_ready && !fn->terms.empty()) TAD_TRY(build_pattern(fn));
    const size_t tsz = fn->precision == TAD_PREC_FP64 ? sizeof(double) : sizeof(float);
    TAD_HIP(fn->gc_buf.ensure((size_t)fn->n_gslots * fn->d * tsz));
    TAD_HIP(fn->hc_buf.ensure((size_t)fn->n_hslots * fn->d * fn->d * tsz));
    TAD_HIP(fn->elem_f.ensure((size_t)fn->n_elem_total));
    TAD_HIP(fn->x_buf.ensure((size_t)fn->n_vars));
    TAD_HIP(fn->g_buf.ensure((size_t)fn->n_vars));
    TAD_HIP(fn->H_buf.ensure((size_t)fn->nnz));
    TAD_HIP(fn->partials.ensure(8 * kReduceBlocks + 4096));
    TAD_HIP(fn->results.ensure(16));
    TAD_HIP(fn->steps.ensure(16));
    TAD_HIP(fn->err.ensure(1));
    return TAD_OK;
}

const double* stage_x(tad_fn fn, const double* x, tad_mem mem)
{
    if (mem == TAD_MEM_DEVICE) return x;
    if (hipMemcpyAsync(fn->x_buf.p, x, fn->n_vars * sizeof(double), hipMemcpyHostToDevice, fn->s()) != hipSuccess)
        return nullptr;
    return fn->x_buf.p;
}

// Runs element kernels + gathers.  mode: MODE_GRAD / MODE_HESS / MODE_HESS_PROJ.
// g_dev / H_dev: device destinations.  Returns f via *f (synchronizes).
// Most element chunks of a pipelined evaluation.  The automatic setting (0)
// does not pipeline: on the 60^3 twisted cube (1.3 M tets) two to eight
// chunks were 2-10 % slower end to end than one launch in FP64 and in mixed
// precision, since every chunk drains its last waves before the next fills
// the chip; tad_fn_autotune times 1, 2, 4 and 8 chunks and keeps the best.
constexpr int kPipeMaxChunks = 64;

// Plans the pipelined evaluation for the current pattern: the largest catalog
// term is cut into chunks of whole tiles; a handle's stage is the last chunk
// holding one of its elements (-1: only other terms, which are launched
// first); row r is complete after chunk k when every row <= r has stage <= k,
// so the rows become ready as prefixes.
void plan_pipeline(tad_fn fn)
{
    fn->pipe_ready = true;
    fn->pipe_used = 1;
    fn->pipe_term = -1;
    int64_t best = 0;
    for (size_t i = 0; i < fn->terms.size(); ++i)
        if (!fn->terms[i].dev_ops.element && fn->terms[i].n_elem > best)
        {
            best = fn->terms[i].n_elem;
            fn->pipe_term = (int)i;
        }
    int S = fn->pipe_chunks > 0 ? fn->pipe_chunks : 1;
    S = (int)std::min<int64_t>(S, std::max<int64_t>(1, best / 256));
    if (fn->pipe_term < 0 || S <= 1) return;
    const HostTerm& h = fn->hterms[fn->pipe_term];
    const int64_t n = h.n_elem;
    fn->pipe_e.assign((size_t)S + 1, 0);
    for (int k = 1; k < S; ++k) fn->pipe_e[k] = (n * k / S) / 256 * 256;
    fn->pipe_e[S] = n;
    std::vector<int> stage((size_t)fn->n_handles, -1);
    for (int k = 0; k < S; ++k)
        for (int64_t e = fn->pipe_e[k]; e < fn->pipe_e[k + 1]; ++e)
            for (int a = 0; a < h.N; ++a)
            {
                const int32_t v = h.conn[e * h.N + a];
                if (v >= 0) stage[v] = std::max(stage[v], k);
            }
    fn->pipe_rows.assign((size_t)S + 1, 0);
    int run = -1;
    int64_t u = 0;
    for (int k = 0; k < S; ++k)
    {
        while (u < fn->n_handles && std::max(run, stage[u]) <= k)
        {
            run = std::max(run, stage[u]);
            ++u;
        }
        fn->pipe_rows[k + 1] = k + 1 == S ? fn->n_handles : u;
    }
    fn->pipe_used = S;
}

tad_status ensure_aux(tad_ctx ctx, int n_events)
{
    if (!ctx->aux) TAD_HIP(hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking));
    while ((int)ctx->events.size() < n_events)
    {
        hipEvent_t e;
        TAD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        ctx->events.push_back(e);
    }
    return TAD_OK;
}

tad_status eval_core(tad_fn fn, const double* x_dev, int mode, double eps, double* g_dev, double* H_dev, double* f)
{
    hipStream_t s = fn->s();
    KernelTimer& tm = fn->ctx->timer;
    TAD_HIP(hipMemsetAsync(fn->err.p, 0, sizeof(int), s));
    if (!fn->pipe_ready) plan_pipeline(fn);
    const bool pipe = fn->pipe_used > 1 && H_dev && mode >= MODE_HESS && fn->n_blocks > 0;
    if (pipe)
    {
        // other terms first, then the chunks of the pipelined term; the
        // assembly of the rows each chunk completes runs on the second stream.
        // Every entry's sum has the same fixed order whatever the launch order.
        const int S = fn->pipe_used;
        tad_ctx ctx = fn->ctx;
        TAD_TRY(ensure_aux(ctx, S + 2));
        hipStream_t s2 = ctx->aux;
        const HessAsm pat = fn->hess_asm();
        const TermDev& tp = fn->terms[fn->pipe_term];
        TAD_HIP(hipEventRecord(ctx->events[S], s));  // contributions of the previous call are consumed
        TAD_HIP(hipStreamWaitEvent(s2, ctx->events[S], 0));
        tm.begin("element", s);
        for (size_t i = 0; i < fn->terms.size(); ++i)
            if ((int)i != fn->pipe_term)
                TAD_HIP(launch_element(fn->terms[i], fn->d, fn->precision, mode, x_dev, eps, fn->elem_f.p,
                                       fn->gc_buf.p, fn->hc_buf.p, fn->err.p, s));
        for (int k = 0; k < S; ++k)
        {
            TAD_HIP(launch_element_range(tp, fn->d, fn->precision, mode, x_dev, eps, fn->elem_f.p, fn->gc_buf.p,
                                         fn->hc_buf.p, fn->err.p, s, fn->pipe_e[k], fn->pipe_e[k + 1]));
            TAD_HIP(hipEventRecord(ctx->events[k], s));
            const int64_t r0 = fn->pipe_rows[k], r1 = fn->pipe_rows[k + 1];
            if (r1 > r0)
            {
                TAD_HIP(hipStreamWaitEvent(s2, ctx->events[k], 0));
                // one timed span per chunk: gradient and Hessian rows of the chunk
                tm.begin("assemble_hessian", s2);
                TAD_HIP(launch_assemble_hessian(fn->d, fn->precision, fn->asm_shape, pat, fn->hc_buf.p, H_dev, s2, r0,
                                                r1));
                if (g_dev)
                    TAD_HIP(launch_assemble_gradient(fn->d, fn->precision, fn->n_handles, fn->gslot_ptr.p,
                                                     fn->gc_buf.p, g_dev, s2, r0, r1));
                tm.end("assemble_hessian", s2);
            }
        }
        tm.end("element", s);
        TAD_HIP(hipEventRecord(ctx->events[S + 1], s2));
        TAD_HIP(hipStreamWaitEvent(s, ctx->events[S + 1], 0));
    }
    else
    {
        tm.begin("element", s);
        for (const TermDev& t : fn->terms)
            TAD_HIP(launch_element(t, fn->d, fn->precision, mode, x_dev, eps, fn->elem_f.p, fn->gc_buf.p, fn->hc_buf.p,
                                   fn->err.p, s));
        tm.end("element", s);
        if (g_dev)
        {
            tm.begin("assemble_gradient", s);
            TAD_HIP(launch_assemble_gradient(fn->d, fn->precision, fn->n_handles, fn->gslot_ptr.p, fn->gc_buf.p, g_dev,
                                             s));
            tm.end("assemble_gradient", s);
        }
    }
    if (H_dev && mode >= MODE_HESS)
    {
        if (!pipe)
        {
            tm.begin("assemble_hessian", s);
            TAD_HIP(launch_assemble_hessian(fn->d, fn->precision, fn->asm_shape, fn->hess_asm(), fn->hc_buf.p, H_dev,
                                            s));
            tm.end("assemble_hessian", s);
        }
        // positions of this Hessian: lever arms of the coarse space's rotations
        TAD_HIP(fn->x_hess.ensure((size_t)fn->n_vars));
        TAD_HIP(hipMemcpyAsync(fn->x_hess.p, x_dev, fn->n_vars * sizeof(double), hipMemcpyDeviceToDevice, s));
        fn->x_hess_valid = true;
    }
    tm.begin("energy_sum", s);
    TAD_HIP(launch_sum(fn->elem_f.p, fn->n_elem_total, 1, fn->n_elem_total, fn->partials.p, fn->results.p, s));
    if (fn->n_seg > 0)
        TAD_HIP(launch_segment_sum(fn->elem_f.p, fn->seg_ranges.p, (int)fn->terms.size(), fn->n_seg, fn->seg_f.p, s));
    tm.end("energy_sum", s);
    double* pin = fn->ctx->pinned;
    TAD_HIP(hipMemcpyAsync(pin, fn->results.p, sizeof(double), hipMemcpyDeviceToHost, s));
    TAD_HIP(hipMemcpyAsync(pin + 1, fn->err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    TAD_HIP(hipStreamSynchronize(s));
    int errv = 0;
    std::memcpy(&errv, pin + 1, sizeof(int));
    *f = pin[0];
    if (errv) return fail(TAD_ERR_NONFINITE, "non-finite element gradient or Hessian");
    return TAD_OK;
}

tad_status eval_energy_batch(tad_fn fn, const double* x_dev, const double* dir_dev, const double* steps_host, int ncand,
                             double* f_out)
{
    hipStream_t s = fn->s();
    KernelTimer& tm = fn->ctx->timer;
    const int64_t ne = fn->n_elem_total;
    if (dir_dev)
    {
        TAD_HIP(fn->ls_f.ensure((size_t)ne * 8));
        double* pin = fn->ctx->pinned;
        std::memcpy(pin + 16, steps_host, ncand * sizeof(double));
        TAD_HIP(hipMemcpyAsync(fn->steps.p, pin + 16, ncand * sizeof(double), hipMemcpyHostToDevice, s));
    }
    double* dst = dir_dev ? fn->ls_f.p : fn->elem_f.p;
    const int nc = dir_dev ? ncand : 1;
    tm.begin("energy", s);
    for (const TermDev& t : fn->terms)
        TAD_HIP(launch_energy(t, fn->d, x_dev, dir_dev, fn->steps.p, nc, dst, ne, s));
    tm.end("energy", s);
    tm.begin("energy_sum", s);
    TAD_HIP(launch_sum(dst, ne, nc, ne, fn->partials.p, fn->results.p, s));
    if (!dir_dev && fn->n_seg > 0)
        TAD_HIP(launch_segment_sum(fn->elem_f.p, fn->seg_ranges.p, (int)fn->terms.size(), fn->n_seg, fn->seg_f.p, s));
    tm.end("energy_sum", s);
    double* pin = fn->ctx->pinned;
    TAD_HIP(hipMemcpyAsync(pin + 32, fn->results.p, nc * sizeof(double), hipMemcpyDeviceToHost, s));
    TAD_HIP(hipStreamSynchronize(s));
    for (int c = 0; c < nc; ++c) f_out[c] = pin[32 + c];
    return TAD_OK;
}

tad_status dot_dev(tad_fn fn, const double* a, const double* b, double* out)
{
    hipStream_t s = fn->s();
    TAD_HIP(launch_dot(a, b, fn->n_vars, fn->partials.p, fn->results.p + 8, s));
    double* pin = fn->ctx->pinned;
    TAD_HIP(hipMemcpyAsync(pin + 48, fn->results.p + 8, sizeof(double), hipMemcpyDeviceToHost, s));
    TAD_HIP(hipStreamSynchronize(s));
    *out = pin[48];
    return TAD_OK;
}

// a1 . b1 and a2 . b2 with one host synchronization
tad_status dot2_dev(tad_fn fn, const double* a1, const double* b1, const double* a2, const double* b2, double* out1,
                    double* out2)
{
    hipStream_t s = fn->s();
    TAD_HIP(launch_dot(a1, b1, fn->n_vars, fn->partials.p, fn->results.p + 8, s));
    TAD_HIP(launch_dot(a2, b2, fn->n_vars, fn->partials.p, fn->results.p + 9, s));
    double* pin = fn->ctx->pinned;
    TAD_HIP(hipMemcpyAsync(pin + 48, fn->results.p + 8, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    TAD_HIP(hipStreamSynchronize(s));
    *out1 = pin[48];
    *out2 = pin[49];
    return TAD_OK;
}

// Recorded CG launches: iterations per graph.
constexpr int kGraphIters = 8;

// Every operand the recorded CG launches hold: the graph is recorded again
// when any of them changes.
std::vector<uint64_t> pcg_graph_key(tad_fn fn, const double* H_dev, const double* sol_dev,
                                    const unsigned char* free_dev, const CoarseDev* co, double w, hipStream_t s)
{
    auto u = [](const void* p) { return (uint64_t)(uintptr_t)p; };
    uint64_t wbits;
    std::memcpy(&wbits, &w, sizeof(w));
    std::vector<uint64_t> k = { (uint64_t)fn->d, (uint64_t)fn->n_handles, u(fn->brow_ptr.p), u(fn->bcol.p),
                                u(fn->row_ptr.p), u(H_dev), wbits, u(fn->Dinv.p), u(sol_dev), u(fn->r.p),
                                u(fn->z.p), u(fn->p.p), u(fn->q.p), u(fn->pcg_parts.p), u(fn->scal.p), u(free_dev),
                                u(s), u(fn->sym_src.p), u(fn->sym_H.p) };
    if (co)
    {
        const uint64_t more[] = { (uint64_t)co->kc, (uint64_t)co->n_agg, (uint64_t)co->nc, u(co->agg_ptr),
                                  u(co->agg_vert), (uint64_t)co->n_chunks, u(co->chunk_ptr), u(co->agg_chunk),
                                  (uint64_t)co->max_agg_chunks, u(co->chunk_part), u(co->dvec), u(co->Ainv32),
                                  u(co->rc), u(co->cparts), (uint64_t)co->Gc, u(co->free_vars) };
        k.insert(k.end(), std::begin(more), std::end(more));
    }
    return k;
}

tad_status solve_core(tad_fn fn, const double* H_dev, const double* rhs_dev, double* sol_dev,
                      const tad_solve_opts& o, tad_solve_stats* st, const unsigned char* free_dev = nullptr)
{
    const int d = fn->d;
    const int64_t nh = fn->n_handles, n = fn->n_vars;
    hipStream_t s = fn->s();
    KernelTimer& tm = fn->ctx->timer;
    TAD_HIP(fn->Dinv.ensure((size_t)nh * d * (d + 1) / 2));  // packed symmetric blocks
    TAD_HIP(fn->r.ensure(n));
    TAD_HIP(fn->z.ensure(n));
    TAD_HIP(fn->p.ensure(2 * n));
    TAD_HIP(fn->q.ensure(n));
    TAD_HIP(fn->scal.ensure(1));
    TAD_HIP(fn->pcg_parts.ensure((size_t)pcg_parts_size(d, nh)));
    if (o.max_iters <= 0) return fail(TAD_ERR_INVALID, "max_iters must be positive");
    tm.begin("pcg_setup", s);
    TAD_HIP(launch_pcg_setup(d, nh, fn->brow_ptr.p, fn->diag_block.p, fn->row_ptr.p, H_dev, o.w_identity, free_dev,
                             fn->Dinv.p, s));
    tm.end("pcg_setup", s);
    TAD_HIP(fn->sym_H.ensure((size_t)fn->n_sym * d * d));
    tm.begin("sym_pack", s);
    TAD_HIP(launch_sym_pack(d, fn->n_sym, fn->sym_full.p, fn->block_row.p, fn->brow_ptr.p, fn->bcol.p,
                            fn->row_ptr.p, H_dev, fn->sym_H.p, s));
    tm.end("sym_pack", s);
    CoarseDev cd;
    bool use_coarse = false, coarse_reused = false;
    TAD_TRY(setup_coarse(fn, H_dev, o, free_dev, &cd, &use_coarse, &coarse_reused));
    const CoarseDev* co = use_coarse ? &cd : nullptr;
    PcgScalars hs;
    std::memset(&hs, 0, sizeof(hs));
    hs.tol2 = o.tol * o.tol;
    hs.max_iters = o.max_iters;
    PcgScalars* pin = (PcgScalars*)(fn->ctx->pinned + 56);
    std::memcpy(pin, &hs, sizeof(hs));
    TAD_HIP(hipMemcpyAsync(fn->scal.p, pin, sizeof(hs), hipMemcpyHostToDevice, s));
    tm.begin("pcg_init", s);
    TAD_HIP(launch_pcg_init(d, nh, rhs_dev, free_dev, fn->Dinv.p, sol_dev, fn->r.p, fn->z.p, fn->pcg_parts.p,
                            fn->scal.p, co, s));
    tm.end("pcg_init", s);
    const int every = std::max(1, o.check_every);
    Timer* timer = fn->ctx->timer.enabled ? &fn->ctx->timer : nullptr;
    // iterations are launched in chunks; every kernel returns at once after
    // the device has recorded termination, so overshooting a chunk is cheap.
    // Iteration k = max_iters launches only the SpMV, whose termination test
    // then always fires.
    // The first poll waits for 3/4 of the previous solve's iteration count
    // (Newton steps need similar counts), later ones come every `every`:
    // each poll stalls the device for a host round trip.
    int k = 0;
    int chunk = std::max(every, fn->last_solve_iters * 3 / 4);
    // recorded launches: kGraphIters iterations, each followed by the
    // device-side iteration counter's advance, replayed until termination
    // (every kernel returns at once after it, and iteration max_iters stops
    // in the SpMV's termination test)
    const bool graphs = !timer && !fn->pcg_graph_failed && fn->solve_graphs == 1;
    hipGraphExec_t gexec = nullptr;
    if (graphs)
    {
        auto launch_chunk = [&]() -> hipError_t {
            for (int c = 0; c < kGraphIters; ++c)
            {
                hipError_t e = launch_pcg_iteration(d, nh, fn->brow_ptr.p, fn->bcol.p, fn->sym_src.p,
                                                    fn->sym_H.p, o.w_identity, fn->Dinv.p, sol_dev,
                                                    fn->r.p, fn->z.p, fn->p.p, fn->q.p, fn->pcg_parts.p, fn->scal.p,
                                                    0, true, free_dev, co, s, nullptr, true);
                if (e == hipSuccess) e = launch_pcg_advance(fn->scal.p, s);
                if (e != hipSuccess) return e;
            }
            return hipSuccess;
        };
        std::vector<uint64_t> key = pcg_graph_key(fn, H_dev, sol_dev, free_dev, co, o.w_identity, s);
        if (fn->pcg_graph && fn->pcg_graph_key == key)
            gexec = fn->pcg_graph;
        else
        {
            if (fn->pcg_graph) TAD_HIP(hipGraphExecDestroy(fn->pcg_graph));
            fn->pcg_graph = nullptr;
            fn->pcg_graph_key.clear();
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess)
            {
                const hipError_t el = launch_chunk();
                e = hipStreamEndCapture(s, &g);
                if (el != hipSuccess) e = el;
            }
            if (e == hipSuccess) e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
            if (g) (void)hipGraphDestroy(g);
            if (e == hipSuccess)
            {
                fn->pcg_graph = gexec;
                fn->pcg_graph_key = std::move(key);
            }
            else
            {
                // e.g. a caller's stream that cannot be captured: direct launches
                (void)hipGetLastError();
                gexec = nullptr;
                fn->pcg_graph_failed = true;
            }
        }
    }
    while (k <= o.max_iters)
    {
        if (gexec)
        {
            for (int c = 0; c < chunk && k <= o.max_iters; c += kGraphIters, k += kGraphIters)
                TAD_HIP(hipGraphLaunch(gexec, s));
        }
        else
        {
            for (int c = 0; c < chunk && k <= o.max_iters; ++c, ++k)
                TAD_HIP(launch_pcg_iteration(d, nh, fn->brow_ptr.p, fn->bcol.p, fn->sym_src.p, fn->sym_H.p,
                                             o.w_identity,
                                             fn->Dinv.p, sol_dev, fn->r.p, fn->z.p, fn->p.p, fn->q.p, fn->pcg_parts.p,
                                             fn->scal.p, k, k < o.max_iters, free_dev, co, s, timer));
        }
        TAD_HIP(hipMemcpyAsync(pin, fn->scal.p, sizeof(hs), hipMemcpyDeviceToHost, s));
        TAD_HIP(hipStreamSynchronize(s));
        if (pin->done) break;
        chunk = every;
    }
    if (!pin->done) return fail(TAD_ERR_SOLVE, "PCG did not report termination");
    std::memcpy(&hs, pin, sizeof(hs));
    fn->last_solve_iters = hs.iters;
    fn->solves_since_rebuild = coarse_reused ? fn->solves_since_rebuild + 1 : 0;
    const double rel = hs.bb > 0 ? std::sqrt(hs.rr / hs.bb) : 0.0;
    if (st)
    {
        st->iters = hs.iters;
        st->rel_residual = rel;
        st->converged = (!hs.breakdown && hs.rr <= hs.tol2 * hs.bb) ? 1 : 0;
        st->coarse_unknowns = use_coarse ? (int)cd.nc : 0;
        st->coarse_reused = coarse_reused ? 1 : 0;
    }
    if (hs.breakdown) return fail(TAD_ERR_SOLVE, "Linear solve failed: matrix not positive definite");
    if (!std::isfinite(rel)) return fail(TAD_ERR_SOLVE, "Linear solve failed: non-finite residual");
    return TAD_OK;
}

// Armijo backtracking on device data (LineSearch.hh:26-65).  The direction is
// dir_scale * dir (dir_scale = -1 lets the Newton step pass the solution of
// H d' = g without negating it).  Writes the accepted point to x_out.
tad_status line_search_core(tad_fn fn, const double* x0, const double* dir, double dir_scale, double f,
                            double dir_dot_g, double* x_out, const tad_newton_opts& o, double* f_out, double* s_out,
                            int* trials, bool* failed)
{
    if (!(o.ls_s_max > 0.0)) return fail(TAD_ERR_INVALID, "Max step size not positive.");
    const bool try_one = o.ls_s_max > 1.0;
    const int batch = std::min(8, std::max(1, o.ls_batch));
    double s = o.ls_s_max;
    int it = 0;
    *failed = false;
    while (it < o.ls_max_iters)
    {
        const int nc = std::min(batch, o.ls_max_iters - it);
        double svals[8], steps[8], fv[8];
        for (int c = 0; c < nc; ++c)
        {
            svals[c] = s;
            steps[c] = dir_scale * s;
            if (try_one && s > 1.0 && s * o.ls_shrink < 1.0) s = 1.0;
            else s *= o.ls_shrink;
        }
        TAD_TRY(eval_energy_batch(fn, x0, dir, steps, nc, fv));
        for (int c = 0; c < nc; ++c)
        {
            ++it;
            if (fv[c] != fv[c]) return fail(TAD_ERR_NAN_ENERGY, "energy is NaN in line search");
            if (fv[c] <= f + o.ls_armijo * svals[c] * dir_dot_g)
            {
                double* pin = fn->ctx->pinned;
                pin[24] = steps[c];
                TAD_HIP(hipMemcpyAsync(fn->steps.p + 8, pin + 24, sizeof(double), hipMemcpyHostToDevice, fn->s()));
                TAD_HIP(launch_axpy_point(x0, dir, fn->steps.p + 8, fn->n_vars, x_out, fn->s()));
                TAD_HIP(hipStreamSynchronize(fn->s()));
                *f_out = fv[c];
                *s_out = svals[c];
                *trials = it;
                return TAD_OK;
            }
        }
    }
    if (x_out != x0)
        TAD_HIP(hipMemcpyAsync(x_out, x0, fn->n_vars * sizeof(double), hipMemcpyDeviceToDevice, fn->s()));
    TAD_HIP(hipStreamSynchronize(fn->s()));
    *f_out = f;
    *s_out = 0.0;
    *trials = it;
    *failed = true;
    return TAD_OK;
}

} // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* tad_last_error(void) { return g_last_error.c_str(); }

const char* tad_version(void) { return "tinyad_amd 0.1 (gfx950)"; }

tad_status tad_ctx_create(int device, tad_ctx* out)
{
    if (!out) return fail(TAD_ERR_INVALID, "null output pointer");
    int count = 0;
    TAD_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(TAD_ERR_INVALID, "invalid device index");
    TAD_HIP(hipSetDevice(device));
    std::unique_ptr<tad_ctx_s> c(new tad_ctx_s());
    c->device = device;
    TAD_HIP(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
    c->stream = c->own_stream;
    TAD_HIP(hipHostMalloc((void**)&c->pinned, 256 * sizeof(double), hipHostMallocDefault));
    TAD_HIP(noalloc_scratch(device, &c->noalloc));
    *out = c.release();
    return TAD_OK;
}

tad_status tad_ctx_destroy(tad_ctx ctx)
{
    if (!ctx) return TAD_OK;
    // a stream borrowed from the caller may be gone by now: fall back to ours
    ctx->stream = ctx->own_stream;
    ctx_release(ctx);
    return TAD_OK;
}

tad_status tad_ctx_set_stream(tad_ctx ctx, void* hip_stream)
{
    if (!ctx) return fail(TAD_ERR_INVALID, "null context");
    ctx->stream = hip_stream ? (hipStream_t)hip_stream : ctx->own_stream;
    return TAD_OK;
}

tad_status tad_ctx_get_stream(tad_ctx ctx, void** hip_stream)
{
    if (!ctx || !hip_stream) return fail(TAD_ERR_INVALID, "null argument");
    *hip_stream = (void*)ctx->stream;
    return TAD_OK;
}

tad_status tad_ctx_synchronize(tad_ctx ctx)
{
    if (!ctx) return fail(TAD_ERR_INVALID, "null context");
    TAD_HIP(hipStreamSynchronize(ctx->stream));
    return TAD_OK;
}

tad_status tad_ctx_enable_timing(tad_ctx ctx, int enable)
{
    if (!ctx) return fail(TAD_ERR_INVALID, "null context");
    ctx->timer.enabled = enable != 0;
    return TAD_OK;
}

tad_status tad_ctx_kernel_stats(tad_ctx ctx, const char* name, int64_t* launches, double* total_ms)
{
    if (!ctx || !name) return fail(TAD_ERR_INVALID, "null argument");
    ctx->timer.collect();
    auto it = ctx->timer.stats.find(name);
    if (launches) *launches = it == ctx->timer.stats.end() ? 0 : it->second.first;
    if (total_ms) *total_ms = it == ctx->timer.stats.end() ? 0.0 : it->second.second;
    return TAD_OK;
}

tad_status tad_ctx_reset_stats(tad_ctx ctx)
{
    if (!ctx) return fail(TAD_ERR_INVALID, "null context");
    ctx->timer.collect();
    ctx->timer.stats.clear();
    return TAD_OK;
}

tad_status tad_fn_create(tad_ctx ctx, int d, int64_t n_handles, tad_precision precision, tad_fn* out)
{
    if (!ctx || !out) return fail(TAD_ERR_INVALID, "null argument");
    if (d < 1 || d > 3) return fail(TAD_ERR_INVALID, "variable dimension must be 1, 2 or 3");
    if (n_handles < 1) return fail(TAD_ERR_INVALID, "at least one variable handle is required");
    if ((double)d * n_handles >= 2147483647.0) return fail(TAD_ERR_UNSUPPORTED, "too many variables");
    if (precision != TAD_PREC_FP64 && precision != TAD_PREC_MIXED) return fail(TAD_ERR_INVALID, "bad precision");
    tad_fn f = new tad_fn_s();
    f->ctx = ctx;
    ctx->refs += 1;
    f->d = d;
    f->n_handles = n_handles;
    f->n_vars = d * n_handles;
    f->precision = precision;
    *out = f;
    return TAD_OK;
}

tad_status tad_fn_destroy(tad_fn fn)
{
    if (!fn) return TAD_OK;
    (void)hipSetDevice(fn->ctx->device);
    (void)hipStreamSynchronize(fn->s());
    if (fn->pcg_graph) (void)hipGraphExecDestroy(fn->pcg_graph);
    for (TermDev& t : fn->terms)
    {
        if (t.conn) (void)hipFree(t.conn);
        if (t.params) (void)hipFree(t.params);
    }
    free_pattern(fn);
    fn->gc_buf.release(); fn->hc_buf.release(); fn->elem_f.release(); fn->x_buf.release(); fn->g_buf.release(); fn->H_buf.release();
    fn->partials.release(); fn->results.release(); fn->steps.release(); fn->err.release();
    fn->ls_f.release(); fn->x_tmp.release();
    fn->Dinv.release(); fn->r.release(); fn->z.release(); fn->p.release(); fn->q.release(); fn->sol.release();
    fn->scal.release(); fn->pcg_parts.release(); fn->seg_ranges.release(); fn->seg_f.release();
    fn->adj_g.release(); fn->adj_rest.release(); fn->adj_lam.release();
    fn->co.release(); fn->x_hess.release(); fn->free_buf.release();
    fn->bas_rp.release(); fn->bas_ci.release(); fn->bas_tp.release(); fn->bas_tc.release();
    fn->bas_v.release(); fn->bas_tv.release(); fn->bas_vec.release();
    tad_ctx ctx = fn->ctx;
    delete fn;
    ctx_release(ctx);
    return TAD_OK;
}

// Validates and uploads one term (shared by the catalog and device-energy
// entry points): handle range, repeated handles, finite parameters.
// allow_pad: elements may end in padding entries -1 (dynamic valences).
// allow_repeats: an element may list a handle more than once (its entries are
// merged, as the reference's Element::variables does, Element.hh:208-243);
// otherwise that is an error.
static tad_status add_term(tad_fn fn, int energy, int N, int64_t n_elem, const int32_t* conn, const double* params,
                           int np, const tad_device_ops* ops, bool allow_pad, bool allow_repeats)
{
    if (n_elem > 0 && (!conn || (np > 0 && !params)))
        return fail(TAD_ERR_INVALID, "null connectivity or parameters");
    bool repeats = false;
    for (int64_t e = 0; e < n_elem; ++e)
        for (int a = 0; a < N; ++a)
        {
            const int32_t v = conn[e * N + a];
            if (v == -1 && allow_pad && (a + 1 == N || conn[e * N + a + 1] == -1)) continue;
            if (v < 0 || v >= fn->n_handles)
                return fail(TAD_ERR_INVALID, "element " + std::to_string(e) + " references handle out of range");
            for (int b = 0; b < a; ++b)
                if (conn[e * N + b] == v)
                {
                    if (!allow_repeats)
                        return fail(TAD_ERR_INVALID, "element " + std::to_string(e) +
                                                         " references a handle twice (symmetric Dirichlet and "
                                                         "residual elements need distinct handles)");
                    repeats = true;
                }
        }
    for (int64_t i = 0; i < n_elem * (int64_t)np; ++i)
        if (!std::isfinite(params[i])) return fail(TAD_ERR_INVALID, "non-finite element parameter");

    TermDev t;
    t.energy = energy;
    if (ops) t.dev_ops = *ops;
    t.N = N;
    t.n_params = np;
    t.n_elem = n_elem;
    t.f_base = fn->n_elem_total;
    if (n_elem > 0)
    {
        std::vector<int32_t> cs((size_t)n_elem * N);
        for (int64_t e = 0; e < n_elem; ++e)
            for (int a = 0; a < N; ++a) cs[a * n_elem + e] = conn[e * N + a];
        std::vector<double> ps((size_t)n_elem * np);
        for (int64_t e = 0; e < n_elem; ++e)
            for (int k = 0; k < np; ++k) ps[k * n_elem + e] = params[e * np + k];
        TAD_HIP(hipMalloc((void**)&t.conn, cs.size() * sizeof(int32_t)));
        TAD_HIP(hipMemcpy(t.conn, cs.data(), cs.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        if (!ps.empty())
        {
            TAD_HIP(hipMalloc((void**)&t.params, ps.size() * sizeof(double)));
            TAD_HIP(hipMemcpy(t.params, ps.data(), ps.size() * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    HostTerm h;
    h.energy = energy;
    h.N = N;
    h.n_params = np;
    h.n_elem = n_elem;
    h.conn.assign(conn, conn + n_elem * N);
    h.repeats = repeats;
    fn->terms.push_back(t);
    fn->hterms.push_back(std::move(h));
    fn->n_elem_total += n_elem;
    free_pattern(fn);
    fn->h_valid = false;
    fn->n_seg = 0;  // segment ranges must be declared again
    return TAD_OK;
}

tad_status tad_fn_add_elements(tad_fn fn, tad_energy energy, int valence, int64_t n_elem, const int32_t* conn,
                               const double* params, int params_per_elem)
{
    TAD_TRY(check_fn(fn));
    if (fn->vec) return fail(TAD_ERR_INVALID, "vector functions take residual elements (tad_fn_add_residuals_device)");
    const int d = fn->d, N = valence;
    if (N < 1 || d * N > 12) return fail(TAD_ERR_INVALID, "element size d * valence must be in 1..12");
    if (n_elem < 0) return fail(TAD_ERR_INVALID, "negative element count");
    if (energy == TAD_E_SYMDIR_TRI && !(d == 2 && N == 3))
        return fail(TAD_ERR_INVALID, "symmetric Dirichlet triangles need d = 2 and valence 3");
    if (energy == TAD_E_SYMDIR_TET && !(d == 3 && N == 4))
        return fail(TAD_ERR_INVALID, "symmetric Dirichlet tets need d = 3 and valence 4");
    if (energy == TAD_E_PENALTY && N != 1) return fail(TAD_ERR_INVALID, "penalty elements have valence 1");
    // dense K x K quadratics above K = 8 would keep A and its eigenvectors in
    // scratch memory; the catalog stops where every kernel stays in registers
    if (energy == TAD_E_QUADRATIC && d * N > 8)
        return fail(TAD_ERR_UNSUPPORTED, "quadratic elements need d * valence <= 8");
    const int np = energy_param_count(energy, d, N);
    if (np < 0) return fail(TAD_ERR_INVALID, "unknown energy");
    if (np != params_per_elem)
        return fail(TAD_ERR_INVALID, "energy expects " + std::to_string(np) + " parameters per element, got " +
                                         std::to_string(params_per_elem));
    // a symmetric Dirichlet simplex with a repeated vertex is degenerate
    return add_term(fn, energy, N, n_elem, conn, params, np, nullptr, false,
                    energy == TAD_E_QUADRATIC || energy == TAD_E_PENALTY);
}

static tad_status check_device_ops(tad_fn fn, const tad_device_ops* ops)
{
    if (!ops) return fail(TAD_ERR_INVALID, "null device ops");
    if (ops->version != TAD_DEVICE_OPS_VERSION)
        return fail(TAD_ERR_UNSUPPORTED, "device ops built against another TAD_DEVICE_OPS_VERSION (" +
                                             std::to_string(ops->version) + ", library " +
                                             std::to_string(TAD_DEVICE_OPS_VERSION) + ")");
    if (!ops->element || !ops->energy) return fail(TAD_ERR_INVALID, "device ops without element or energy function");
    if (ops->dim != fn->d)
        return fail(TAD_ERR_INVALID, "device energy has dimension " + std::to_string(ops->dim) + ", function has " +
                                         std::to_string(fn->d));
    if (ops->valence < 1 || fn->d * ops->valence > 12)
        return fail(TAD_ERR_INVALID, "element size d * valence must be in 1..12");
    if (ops->n_params < 0) return fail(TAD_ERR_INVALID, "negative parameter count");
    return TAD_OK;
}

tad_status tad_fn_add_elements_dynamic_device(tad_fn fn, const tad_device_ops* const* ops, int n_ops, int64_t n_elem,
                                              const int64_t* elem_ptr, const int32_t* handles, const double* params)
{
    TAD_TRY(check_fn(fn));
    if (fn->vec) return fail(TAD_ERR_INVALID, "vector functions take residual elements (tad_fn_add_residuals_device)");
    if (!ops || n_ops < 1) return fail(TAD_ERR_INVALID, "no device ops");
    if (n_elem < 0) return fail(TAD_ERR_INVALID, "negative element count");
    for (int i = 0; i < n_ops; ++i)
    {
        TAD_TRY(check_device_ops(fn, ops[i]));
        if (ops[i]->n_params != ops[0]->n_params)
            return fail(TAD_ERR_INVALID, "device ops of one dynamic term differ in their parameter count");
        for (int j = 0; j < i; ++j)
            if (ops[j]->valence == ops[i]->valence)
                return fail(TAD_ERR_INVALID, "element valences of add_elements_dynamic are not unique");
    }
    if (n_elem > 0 && !elem_ptr) return fail(TAD_ERR_INVALID, "null element offsets");
    const int np = ops[0]->n_params;
    // group by the smallest valence that holds the element (reference:
    // lower_bound over the sorted static valences)
    std::vector<int> group(n_elem);
    for (int64_t e = 0; e < n_elem; ++e)
    {
        const int64_t c = elem_ptr[e + 1] - elem_ptr[e];
        if (c < 0 || elem_ptr[e] < 0) return fail(TAD_ERR_INVALID, "element offsets must be non-decreasing");
        int best = -1;
        for (int i = 0; i < n_ops; ++i)
            if (ops[i]->valence >= c && (best < 0 || ops[i]->valence < ops[best]->valence)) best = i;
        if (best < 0)
            return fail(TAD_ERR_INVALID, "element valence " + std::to_string(c) +
                                             " exceeds the largest valence passed to add_elements_dynamic");
        group[e] = best;
    }
    if (n_elem > 0 && elem_ptr[n_elem] > 0 && !handles) return fail(TAD_ERR_INVALID, "null handles");
    if (n_elem > 0 && np > 0 && !params) return fail(TAD_ERR_INVALID, "null parameters");
    // validate everything before the first term is added (a handle may
    // appear more than once in an element)
    for (int64_t e = 0; e < n_elem; ++e)
        for (int64_t a = elem_ptr[e]; a < elem_ptr[e + 1]; ++a)
            if (handles[a] < 0 || handles[a] >= fn->n_handles)
                return fail(TAD_ERR_INVALID, "element " + std::to_string(e) + " references handle out of range");
    for (int64_t i = 0; i < n_elem * (int64_t)np; ++i)
        if (!std::isfinite(params[i])) return fail(TAD_ERR_INVALID, "non-finite element parameter");
    for (int i = 0; i < n_ops; ++i)
    {
        const int N = ops[i]->valence;
        std::vector<int32_t> conn;
        std::vector<double> ps;
        int64_t ng = 0;
        for (int64_t e = 0; e < n_elem; ++e)
        {
            if (group[e] != i) continue;
            const int64_t c = elem_ptr[e + 1] - elem_ptr[e];
            for (int a = 0; a < N; ++a) conn.push_back(a < c ? handles[elem_ptr[e] + a] : -1);
            for (int k = 0; k < np; ++k) ps.push_back(params[e * np + k]);
            ++ng;
        }
        if (ng == 0) continue;
        TAD_TRY(add_term(fn, TAD_E_DEVICE, N, ng, conn.data(), ps.data(), np, ops[i], true, true));
    }
    return TAD_OK;
}

tad_status tad_fn_add_elements_device(tad_fn fn, const tad_device_ops* ops, int64_t n_elem, const int32_t* conn,
                                      const double* params)
{
    TAD_TRY(check_fn(fn));
    if (fn->vec) return fail(TAD_ERR_INVALID, "vector functions take residual elements (tad_fn_add_residuals_device)");
    TAD_TRY(check_device_ops(fn, ops));
    if (n_elem < 0) return fail(TAD_ERR_INVALID, "negative element count");
    return add_term(fn, TAD_E_DEVICE, ops->valence, n_elem, conn, params, ops->n_params, ops, false, true);
}

tad_status tad_fn_create_vector(tad_ctx ctx, int d, int64_t n_handles, tad_precision precision, tad_fn* out)
{
    TAD_TRY(tad_fn_create(ctx, d, n_handles, precision, out));
    (*out)->vec = true;
    return TAD_OK;
}

tad_status tad_fn_add_residuals_device(tad_fn fn, const tad_device_residual_ops* ops, int64_t n_elem,
                                       const int32_t* conn, const double* params)
{
    TAD_TRY(check_fn(fn));
    if (!fn->vec) return fail(TAD_ERR_INVALID, "residual elements need a vector function (tad_fn_create_vector)");
    if (!ops) return fail(TAD_ERR_INVALID, "null residual ops");
    if (ops->n_outputs < 1 || ops->n_outputs > 64) return fail(TAD_ERR_INVALID, "residuals per element must be in 1..64");
    if (!ops->residuals) return fail(TAD_ERR_INVALID, "residual ops without a residual function");
    tad_device_ops o{};
    o.version = ops->version;
    o.valence = ops->valence;
    o.dim = ops->dim;
    o.n_params = ops->n_params;
    o.element = ops->element;
    o.energy = ops->energy;
    TAD_TRY(check_device_ops(fn, &o));
    if (n_elem < 0) return fail(TAD_ERR_INVALID, "negative element count");
    const int64_t K = (int64_t)fn->d * ops->valence;
    int64_t nnz_J = 0;
    for (const TermDev& t : fn->terms) nnz_J = std::max(nnz_J, t.jv_base + t.n_elem * t.n_out * (int64_t)t.N * fn->d);
    if ((double)(fn->n_outputs + n_elem * ops->n_outputs) >= 2147483647.0 ||
        (double)(nnz_J + n_elem * ops->n_outputs * K) >= 2147483647.0)
        return fail(TAD_ERR_UNSUPPORTED, "too many residuals or Jacobian entries (int32 CSR indices)");
    // residual Jacobian rows keep one column per element entry
    TAD_TRY(add_term(fn, TAD_E_DEVICE, ops->valence, n_elem, conn, params, ops->n_params, &o, false, false));
    TermDev& t = fn->terms.back();
    t.n_out = ops->n_outputs;
    t.r_base = fn->n_outputs;
    t.jv_base = nnz_J;
    t.residuals = ops->residuals;
    fn->n_outputs += n_elem * ops->n_outputs;
    return TAD_OK;
}

tad_status tad_fn_n_outputs(tad_fn fn, int64_t* n_outputs)
{
    TAD_TRY(check_fn(fn));
    if (!n_outputs) return fail(TAD_ERR_INVALID, "null argument");
    *n_outputs = fn->n_outputs;
    return TAD_OK;
}

static tad_status ensure_jacobian_pattern(tad_fn fn)
{
    if (!fn->vec) return fail(TAD_ERR_INVALID, "not a vector function (tad_fn_create_vector)");
    if (fn->jpattern_ready) return TAD_OK;
    const int d = fn->d;
    int64_t nnz = 0;
    for (const TermDev& t : fn->terms) nnz += t.n_elem * t.n_out * (int64_t)t.N * d;
    std::vector<int32_t> rp((size_t)fn->n_outputs + 1), ci((size_t)nnz);
    rp[0] = 0;
    for (size_t k = 0; k < fn->terms.size(); ++k)
    {
        const TermDev& t = fn->terms[k];
        const HostTerm& h = fn->hterms[k];
        const int K = t.N * d;
        std::vector<int32_t> hs(t.N);
        for (int64_t e = 0; e < t.n_elem; ++e)
        {
            hs.assign(h.conn.begin() + e * t.N, h.conn.begin() + (e + 1) * t.N);
            std::sort(hs.begin(), hs.end());
            for (int o = 0; o < t.n_out; ++o)
            {
                const int64_t row = t.r_base + e * t.n_out + o;
                const int64_t v0 = t.jv_base + (e * t.n_out + o) * (int64_t)K;
                rp[row + 1] = (int32_t)(v0 + K);
                for (int a = 0; a < t.N; ++a)
                    for (int i = 0; i < d; ++i) ci[v0 + a * d + i] = hs[a] * d + i;
            }
        }
    }
    TAD_HIP(fn->jrow_ptr.ensure(rp.size()));
    TAD_HIP(fn->jcol.ensure(ci.size()));
    TAD_HIP(hipMemcpy(fn->jrow_ptr.p, rp.data(), rp.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    if (!ci.empty()) TAD_HIP(hipMemcpy(fn->jcol.p, ci.data(), ci.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    fn->nnz_J = nnz;
    fn->jpattern_ready = true;
    return TAD_OK;
}

tad_status tad_fn_jacobian_pattern(tad_fn fn, int64_t* nnz, const int32_t** row_ptr_dev, const int32_t** col_idx_dev,
                                   int32_t* row_ptr_host, int32_t* col_idx_host)
{
    TAD_TRY(check_fn(fn));
    TAD_HIP(hipSetDevice(fn->ctx->device));
    TAD_TRY(ensure_jacobian_pattern(fn));
    if (nnz) *nnz = fn->nnz_J;
    if (row_ptr_dev) *row_ptr_dev = fn->jrow_ptr.p;
    if (col_idx_dev) *col_idx_dev = fn->jcol.p;
    if (row_ptr_host)
        TAD_HIP(hipMemcpy(row_ptr_host, fn->jrow_ptr.p, (fn->n_outputs + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (col_idx_host && fn->nnz_J)
        TAD_HIP(hipMemcpy(col_idx_host, fn->jcol.p, fn->nnz_J * sizeof(int32_t), hipMemcpyDeviceToHost));
    return TAD_OK;
}

tad_status tad_fn_n_terms(tad_fn fn, int* n_terms)
{
    TAD_TRY(check_fn(fn));
    if (!n_terms) return fail(TAD_ERR_INVALID, "null argument");
    *n_terms = (int)fn->hterms.size();
    return TAD_OK;
}

tad_status tad_fn_term_info(tad_fn fn, int term, int* energy, int* valence, int64_t* n_elem)
{
    TAD_TRY(check_fn(fn));
    if (term < 0 || term >= (int)fn->hterms.size()) return fail(TAD_ERR_INVALID, "term index out of range");
    const HostTerm& t = fn->hterms[term];
    if (energy) *energy = t.energy;
    if (valence) *valence = t.N;
    if (n_elem) *n_elem = t.n_elem;
    return TAD_OK;
}

tad_status tad_fn_n_vars(tad_fn fn, int64_t* n_vars)
{
    TAD_TRY(check_fn(fn));
    *n_vars = fn->n_vars;
    return TAD_OK;
}

tad_status tad_fn_n_elements(tad_fn fn, int64_t* n_elements)
{
    TAD_TRY(check_fn(fn));
    *n_elements = fn->n_elem_total;
    return TAD_OK;
}

tad_status tad_fn_pattern(tad_fn fn, int64_t* nnz, const int32_t** row_ptr_dev, const int32_t** col_idx_dev,
                          int32_t* row_ptr_host, int32_t* col_idx_host)
{
    TAD_TRY(check_fn(fn));
    TAD_TRY(ensure_ready(fn));
    if (nnz) *nnz = fn->nnz;
    if (row_ptr_dev) *row_ptr_dev = fn->row_ptr.p;
    if (col_idx_dev) *col_idx_dev = fn->col_idx.p;
    if (row_ptr_host)
        TAD_HIP(hipMemcpy(row_ptr_host, fn->row_ptr.p, (fn->n_vars + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (col_idx_host && fn->nnz)
        TAD_HIP(hipMemcpy(col_idx_host, fn->col_idx.p, fn->nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
    return TAD_OK;
}

tad_status tad_fn_eval(tad_fn fn, const double* x, tad_mem mem, double* f)
{
    TAD_TRY(check_fn(fn));
    if (!x || !f) return fail(TAD_ERR_INVALID, "null argument");
    TAD_HIP(hipSetDevice(fn->ctx->device));
    TAD_TRY(ensure_ready(fn));
    if (fn->n_elem_total == 0) { *f = 0.0; return TAD_OK; }
    const double* xd = stage_x(fn, x, mem);
    if (!xd) return fail(TAD_ERR_HIP, "copy of x to the device failed");
    double fv = 0.0;
    TAD_TRY(eval_energy_batch(fn, xd, nullptr, nullptr, 1, &fv));
    *f = fv;
    return TAD_OK;
}

tad_status tad_fn_eval_noalloc(tad_fn fn, const double* x_dev, double* f_dev)
{
    TAD_TRY(check_fn(fn));
    if (!x_dev || !f_dev) return fail(TAD_ERR_INVALID, "null argument");
    for (const TermDev& t : fn->terms)
        if (t.dev_ops.element)
            return fail(TAD_ERR_UNSUPPORTED, "tad_fn_eval_noalloc: device-functor terms are not supported");
    const int64_t parts = noalloc_partials(fn->terms.data(), (int)fn->terms.size());
    if (parts > kNoAllocPartials)
        return fail(TAD_ERR_UNSUPPORTED, "tad_fn_eval_noalloc: " + std::to_string(fn->n_elem_total) +
                                             " elements need " + std::to_string(parts) +
                                             " partial sums, the static scratch holds " +
                                             std::to_string(kNoAllocPartials));
    TAD_HIP(hipSetDevice(fn->ctx->device));
    TAD_HIP(launch_energy_noalloc(fn->terms.data(), (int)fn->terms.size(), fn->d, x_dev, fn->ctx->noalloc, f_dev,
                                  fn->s()));
    return TAD_OK;
}

static tad_status eval_second(tad_fn fn, const double* x, int mode, double eps, double* f, double* g, double* H,
                              tad_mem mem)
{
    TAD_TRY(check_fn(fn));
    if (!x || !f) return fail(TAD_ERR_INVALID, "null argument");
    TAD_HIP(hipSetDevice(fn->ctx->device));
    TAD_TRY(ensure_ready(fn));
    hipStream_t s = fn->s();
    if (fn->n_elem_total == 0)
    {
        *f = 0.0;
        if (g)
        {
            if (mem == TAD_MEM_DEVICE) TAD_HIP(launch_fill(g, fn->n_vars, 0.0, s));
            else std::fill(g, g + fn->n_vars, 0.0);
        }
        TAD_HIP(hipStreamSynchronize(s));
        return TAD_OK;
    }
    const double* xd = stage_x(fn, x, mem);
    if (!xd) return fail(TAD_ERR_HIP, "copy of x to the device failed");
    double* g_dev = (mem == TAD_MEM_DEVICE && g) ? g : fn->g_buf.p;
    double* H_dev = (mode >= MODE_HESS) ? ((mem == TAD_MEM_DEVICE && H) ? H : fn->H_buf.p) : nullptr;
    TAD_TRY(eval_core(fn, xd, mode, eps, g_dev, H_dev, f));
    fn->h_valid = (H_dev == fn->H_buf.p);
    if (mem == TAD_MEM_HOST)
    {
        if (g) TAD_HIP(hipMemcpyAsync(g, g_dev, fn->n_vars * sizeof(double), hipMemcpyDeviceToHost, s));
        if (H && H_dev) TAD_HIP(hipMemcpyAsync(H, H_dev, fn->nnz * sizeof(double), hipMemcpyDeviceToHost, s));
        TAD_HIP(hipStreamSynchronize(s));
    }
    return TAD_OK;
}

tad_status tad_fn_eval_with_gradient(tad_fn fn, const double* x, double* f, double* g, tad_mem mem)
{
    return eval_second(fn, x, MODE_GRAD, 0.0, f, g, nullptr, mem);
}

tad_status tad_fn_eval_with_derivatives(tad_fn fn, const double* x, double* f, double* g, double* H_values,
                                        tad_mem mem)
{
    return eval_second(fn, x, MODE_HESS, 0.0, f, g, H_values, mem);
}

tad_status tad_fn_eval_with_hessian_proj(tad_fn fn, const double* x, double eps, double* f, double* g,
                                         double* H_values, tad_mem mem)
{
    if (eps != eps) return fail(TAD_ERR_INVALID, "projection eps is NaN");
    return eval_second(fn, x, MODE_HESS_PROJ, eps, f, g, H_values, mem);
}

// r (and J values) of a vector function at device x into device r_dev / J_dev.
static tad_status eval_residuals_dev(tad_fn fn, const double* xd, double* r_dev, double* J_dev)
{
    hipStream_t s = fn->s();
    KernelTimer& tm = fn->ctx->timer;
    tm.begin("residuals", s);
    for (const TermDev& t : fn->terms)
    {
        if (t.n_elem == 0) continue;
        tad_device_residual_args a{};
        a.n_elem = t.n_elem;
        a.conn = t.conn;
        a.params = t.params;
        a.x = xd;
        a.r = r_dev + t.r_base;
        a.J = J_dev ? J_dev + t.jv_base : nullptr;
        a.stream = (void*)s;
        TAD_HIP((hipError_t)t.residuals(&a));
    }
    tm.end("residuals", s);
    return TAD_OK;
}

// Copies r and J values of a vector function at x to the caller (r, J_values
// may be null).
static tad_status residual_outputs(tad_fn fn, const double* xd, double* r, double* J_values, tad_mem mem)
{
    hipStream_t s = fn->s();
    if (J_values) TAD_TRY(ensure_jacobian_pattern(fn));
    double* r_dev = (mem == TAD_MEM_DEVICE && r) ? r : nullptr;
    if (!r_dev)
    {
        TAD_HIP(fn->r_buf.ensure((size_t)fn->n_outputs));
        r_dev = fn->r_buf.p;
    }
    double* J_dev = nullptr;
    if (J_values)
    {
        J_dev = mem == TAD_MEM_DEVICE ? J_values : nullptr;
        if (!J_dev)
        {
            TAD_HIP(fn->J_buf.ensure((size_t)fn->nnz_J));
            J_dev = fn->J_buf.p;
        }
    }
    TAD_TRY(eval_residuals_dev(fn, xd, r_dev, J_dev));
    if (mem == TAD_MEM_HOST)
    {
        if (r) TAD_HIP(hipMemcpyAsync(r, r_dev, fn->n_outputs * sizeof(double), hipMemcpyDeviceToHost, s));
        if (J_values && fn->nnz_J)
            TAD_HIP(hipMemcpyAsync(J_values, J_dev, fn->nnz_J * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    TAD_HIP(hipStreamSynchronize(s));
    return TAD_OK;
}

tad_status tad_fn_eval_residuals(tad_fn fn, const double* x, double* r, tad_mem mem)
{
    return tad_fn_eval_with_jacobian(fn, x, r, nullptr, mem);
}

tad_status tad_fn_eval_with_jacobian(tad_fn fn, const double* x, double* r, double* J_values, tad_mem mem)
{
    TAD_TRY(check_fn(fn));
    if (!fn->vec) return fail(TAD_ERR_INVALID, "not a vector function (tad_fn_create_vector)");
    if (!x || !r) return fail(TAD_ERR_INVALID, "null argument");
    TAD_HIP(hipSetDevice(fn->ctx->device));
    TAD_TRY(ensure_ready(fn));
    const double* xd = stage_x(fn, x, mem);
    if (!xd) return fail(TAD_ERR_HIP, "copy of x to the device failed");
    return residual_outputs(fn, xd, r, J_values, mem);
}

tad_status tad_fn_eval_sum_of_squares_with_derivatives(tad_fn fn, const double* x, double* f, double* g, double* r,
                                                       double* J_values, tad_mem mem)
{
    TAD_TRY(check_fn(fn));
    if (!fn->vec) return fail(TAD_ERR_INVALID, "not a vector function (tad_fn_create_vector)");
    TAD_TRY(eval_second(fn, x, MODE_GRAD, 0.0, f, g, nullptr, mem));
    if (!r && !J_values) return TAD_OK;
    const double* xd = stage_x(fn, x, mem);
    if (!xd) return fail(TAD_ERR_HIP, "copy of x to the device failed");
    return residual_outputs(fn, xd, r, J_values, mem);
}

tad_status tad_fn_device_buffers(tad_fn fn, const double** g_dev, const double** H_values_dev)
{
    TAD_TRY(check_fn(fn));
    TAD_TRY(ensure_ready(fn));
    if (g_dev) *g_dev = fn->g_buf.p;
    if (H_values_dev) *H_values_dev = fn->H_buf.p;
    return TAD_OK;
}

tad_status tad_solve_default_opts(tad_solve_opts* o)
{
    if (!o) return fail(TAD_ERR_INVALID, "null argument");
    o->w_identity = 0.0;
    o->tol = 1e-10;
    o->max_iters = 20000;
    o->check_every = 16;
    o->coarse_aggregate = -1;
    o->coarse_reuse = 0;
    return TAD_OK;
}

static tad_status solve_entry(tad_fn fn, const double* H_values, const double* rhs, double* sol, tad_mem mem,
                              const unsigned char* free_vars, const tad_solve_opts* opts, tad_solve_stats* stats)
{
    TAD_TRY(check_fn(fn));
    if (!rhs || !sol) return fail(TAD_ERR_INVALID, "null argument");
    TAD_HIP(hipSetDevice(fn->ctx->device));
    TAD_TRY(ensure_ready(fn));
    tad_solve_opts o;
    tad_solve_default_opts(&o);
    if (opts) o = *opts;
    hipStream_t s = fn->s();
    const int64_t n = fn->n_vars;
    const double* Hd;
    if (!H_values)
    {
        if (!fn->h_valid) return fail(TAD_ERR_INVALID, "no Hessian evaluated yet");
        Hd = fn->H_buf.p;
    }
    else if (mem == TAD_MEM_DEVICE)
        Hd = H_values;
    else
    {
        TAD_HIP(hipMemcpyAsync(fn->H_buf.p, H_values, fn->nnz * sizeof(double), hipMemcpyHostToDevice, s));
        fn->h_valid = true;
        Hd = fn->H_buf.p;
    }
    const unsigned char* fd = free_vars;
    if (free_vars && mem == TAD_MEM_HOST)
    {
        TAD_HIP(fn->free_buf.ensure(n));
        TAD_HIP(hipMemcpyAsync(fn->free_buf.p, free_vars, n, hipMemcpyHostToDevice, s));
        fd = fn->free_buf.p;
    }
    TAD_HIP(fn->sol.ensure(n));
    const double* bd = rhs;
    double* xd = sol;
    if (mem == TAD_MEM_HOST)
    {
        TAD_HIP(fn->x_tmp.ensure(n));
        TAD_HIP(hipMemcpyAsync(fn->x_tmp.p, rhs, n * sizeof(double), hipMemcpyHostToDevice, s));
        bd = fn->x_tmp.p;
        xd = fn->sol.p;
    }
    tad_status st = solve_core(fn, Hd, bd, xd, o, stats, fd);
    if (mem == TAD_MEM_HOST)
    {
        TAD_HIP(hipMemcpyAsync(sol, xd, n * sizeof(double), hipMemcpyDeviceToHost, s));
        TAD_HIP(hipStreamSynchronize(s));
    }
    return st;
}

tad_status tad_fn_solve(tad_fn fn, const double* H_values, const double* rhs, double* sol, tad_mem mem,
                        const tad_solve_opts* opts, tad_solve_stats* stats)
{
    return solve_entry(fn, H_values, rhs, sol, mem, nullptr, opts, stats);
}

tad_status tad_fn_solve_reduced(tad_fn fn, const double* H_values, const double* rhs, double* sol, tad_mem mem,
                                const unsigned char* free_vars, const tad_solve_opts* opts, tad_solve_stats* stats)
{
    if (!free_vars) return fail(TAD_ERR_INVALID, "null free-variable mask");
    return solve_entry(fn, H_values, rhs, sol, mem, free_vars, opts, stats);
}

// Length-n dot products a_i . b_i (fixed order, as launch_dot) with one host
// synchronization.
static tad_status dots_n(tad_fn fn, int64_t n, int count, const double* const* a, const double* const* b,
                         double* out)
{
    hipStream_t s = fn->s();
    for (int i = 0; i < count; ++i)
        TAD_HIP(launch_dot(a[i], b[i], n, fn->partials.p, fn->results.p + 10 + i, s));
    double* pin = fn->ctx->pinned;
    TAD_HIP(hipMemcpyAsync(pin + 50, fn->results.p + 10, count * sizeof(double), hipMemcpyDeviceToHost, s));
    TAD_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < count; ++i) out[i] = pin[50 + i];
    return TAD_OK;
}

tad_status tad_fn_solve_basis(tad_fn fn, const double* H_values, int64_t m, const int32_t* b_row_ptr,
                              const int32_t* b_col, const double* b_val, const double* rhs, double* sol, tad_mem mem,
                              const tad_solve_opts* opts, tad_solve_stats* stats)
{
    TAD_TRY(check_fn(fn));
    if (!b_row_ptr || !b_col || !b_val || !rhs || !sol) return fail(TAD_ERR_INVALID, "null argument");
    if (m < 1 || m > INT32_MAX) return fail(TAD_ERR_INVALID, "the basis needs 1 .. 2^31 - 1 columns");
    TAD_HIP(hipSetDevice(fn->ctx->device));
    TAD_TRY(ensure_ready(fn));
    tad_solve_opts o;
    tad_solve_default_opts(&o);
    if (opts) o = *opts;
    if (!(o.tol > 0.0) || o.max_iters < 1) return fail(TAD_ERR_INVALID, "tol must be positive, max_iters >= 1");
    hipStream_t s = fn->s();
    const int64_t n = fn->n_vars;
    const double* Hd;
    if (!H_values)
    {
        if (!fn->h_valid) return fail(TAD_ERR_INVALID, "no Hessian evaluated yet");
        Hd = fn->H_buf.p;
    }
    else if (mem == TAD_MEM_DEVICE)
        Hd = H_values;
    else
    {
        TAD_HIP(hipMemcpyAsync(fn->H_buf.p, H_values, fn->nnz * sizeof(double), hipMemcpyHostToDevice, s));
        fn->h_valid = true;
        Hd = fn->H_buf.p;
    }
    // B on the host, checked, and its transpose (rows of B^T = columns of B,
    // entries in row order of B)
    std::vector<int32_t> rp((size_t)n + 1);
    auto fetch = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
        if (bytes == 0) return hipSuccess;
        if (mem == TAD_MEM_HOST)
        {
            std::memcpy(dst, src, bytes);
            return hipSuccess;
        }
        hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
        return e != hipSuccess ? e : hipStreamSynchronize(s);
    };
    TAD_HIP(fetch(rp.data(), b_row_ptr, rp.size() * sizeof(int32_t)));
    if (rp[0] != 0) return fail(TAD_ERR_INVALID, "basis row offsets must start at 0");
    for (int64_t i = 0; i < n; ++i)
        if (rp[i + 1] < rp[i]) return fail(TAD_ERR_INVALID, "basis row offsets must not decrease");
    const int64_t nb = rp[n];
    std::vector<int32_t> ci((size_t)nb);
    std::vector<double> bv((size_t)nb);
    TAD_HIP(fetch(ci.data(), b_col, ci.size() * sizeof(int32_t)));
    TAD_HIP(fetch(bv.data(), b_val, bv.size() * sizeof(double)));
    std::vector<int32_t> tp((size_t)m + 1, 0);
    for (int64_t k = 0; k < nb; ++k)
    {
        if (ci[k] < 0 || ci[k] >= m) return fail(TAD_ERR_INVALID, "basis column index out of range");
        if (!std::isfinite(bv[k])) return fail(TAD_ERR_INVALID, "basis entries must be finite");
        tp[ci[k] + 1]++;
    }
    for (int64_t j = 0; j < m; ++j) tp[j + 1] += tp[j];
    std::vector<int32_t> tc((size_t)nb);
    std::vector<double> tv((size_t)nb);
    {
        std::vector<int32_t> next(tp.begin(), tp.end() - 1);
        for (int64_t r = 0; r < n; ++r)
            for (int32_t k = rp[r]; k < rp[r + 1]; ++k)
            {
                const int32_t q = next[ci[k]]++;
                tc[q] = (int32_t)r;
                tv[q] = bv[k];
            }
    }
    auto up = [&](auto& buf, const auto& vec) -> hipError_t {
        hipError_t e = buf.ensure(vec.size());
        if (e != hipSuccess || vec.empty()) return e;
        return hipMemcpyAsync(buf.p, vec.data(), vec.size() * sizeof(vec[0]), hipMemcpyHostToDevice, s);
    };
    TAD_HIP(up(fn->bas_rp, rp));
    TAD_HIP(up(fn->bas_ci, ci));
    TAD_HIP(up(fn->bas_v, bv));
    TAD_HIP(up(fn->bas_tp, tp));
    TAD_HIP(up(fn->bas_tc, tc));
    TAD_HIP(up(fn->bas_tv, tv));
    TAD_HIP(fn->bas_vec.ensure((size_t)(6 * m + 4 * n)));
    double* y = fn->bas_vec.p;
    double* r = y + m;
    double* z = r + m;
    double* p = z + m;
    double* q = p + m;
    double* dinv = q + m;
    double* t = dinv + m;
    double* sv = t + n;
    double* bd = sv + n;
    double* out = bd + n;
    if (mem == TAD_MEM_HOST)
        TAD_HIP(hipMemcpyAsync(bd, rhs, n * sizeof(double), hipMemcpyHostToDevice, s));
    else
        TAD_HIP(hipMemcpyAsync(bd, rhs, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    const int32_t* brp = fn->bas_rp.p;
    const int32_t* bci = fn->bas_ci.p;
    const double* bva = fn->bas_v.p;
    const int32_t* btp = fn->bas_tp.p;
    const int32_t* btc = fn->bas_tc.p;
    const double* btv = fn->bas_tv.p;
    // r = B^T rhs, the Jacobi preconditioner of B^T H B + w I, y = 0
    TAD_HIP(launch_csr_spmv(m, btp, btc, btv, bd, 0.0, nullptr, r, s));
    TAD_HIP(hipMemsetAsync(fn->err.p, 0, sizeof(int), s));
    TAD_HIP(launch_basis_diag(m, btp, btc, btv, fn->row_ptr.p, fn->col_idx.p, Hd, o.w_identity, dinv, fn->err.p, s));
    TAD_HIP(launch_fill(y, m, 0.0, s));
    int bad = 0;
    TAD_HIP(hipMemcpyAsync(fn->ctx->pinned + 2, fn->err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    TAD_HIP(hipStreamSynchronize(s));
    std::memcpy(&bad, fn->ctx->pinned + 2, sizeof(int));
    if (bad) return fail(TAD_ERR_SOLVE, "B^T H B + w I has a diagonal entry that is not positive");
    TAD_HIP(launch_basis_scale(m, r, dinv, z, s));
    TAD_HIP(hipMemcpyAsync(p, z, m * sizeof(double), hipMemcpyDeviceToDevice, s));
    double d2[2];
    {
        const double* a[2] = { r, r };
        const double* b[2] = { r, z };
        TAD_TRY(dots_n(fn, m, 2, a, b, d2));
    }
    const double bb = d2[0];
    double rho = d2[1], rr = bb;
    int it = 0;
    bool converged = bb == 0.0;
    const double tol2 = o.tol * o.tol;
    while (!converged && it < o.max_iters)
    {
        // q = B^T H B p + w p
        TAD_HIP(launch_csr_spmv(n, brp, bci, bva, p, 0.0, nullptr, t, s));
        TAD_HIP(launch_csr_spmv(n, fn->row_ptr.p, fn->col_idx.p, Hd, t, 0.0, nullptr, sv, s));
        TAD_HIP(launch_csr_spmv(m, btp, btc, btv, sv, o.w_identity, p, q, s));
        double pq;
        {
            const double* a[1] = { p };
            const double* b[1] = { q };
            TAD_TRY(dots_n(fn, m, 1, a, b, &pq));
        }
        if (!(pq > 0.0)) return fail(TAD_ERR_SOLVE, "B^T H B + w I is not positive definite (CG breakdown)");
        const double alpha = rho / pq;
        TAD_HIP(launch_basis_step(m, alpha, p, q, y, r, s));
        TAD_HIP(launch_basis_scale(m, r, dinv, z, s));
        ++it;
        {
            const double* a[2] = { r, r };
            const double* b[2] = { r, z };
            TAD_TRY(dots_n(fn, m, 2, a, b, d2));
        }
        rr = d2[0];
        if (rr <= tol2 * bb)
        {
            converged = true;
            break;
        }
        const double beta = d2[1] / rho;
        rho = d2[1];
        TAD_HIP(launch_basis_direction(m, z, beta, p, s));
    }
    // sol = B y
    TAD_HIP(launch_csr_spmv(n, brp, bci, bva, y, 0.0, nullptr, out, s));
    if (mem == TAD_MEM_HOST)
    {
        TAD_HIP(hipMemcpyAsync(sol, out, n * sizeof(double), hipMemcpyDeviceToHost, s));
        TAD_HIP(hipStreamSynchronize(s));
    }
    else
        TAD_HIP(hipMemcpyAsync(sol, out, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (stats)
    {
        stats->iters = it;
        stats->converged = converged ? 1 : 0;
        stats->rel_residual = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
        stats->coarse_unknowns = 0;
        stats->coarse_reused = 0;
    }
    return TAD_OK;
}

tad_status tad_newton_default_opts(tad_newton_opts* o)
{
    if (!o) return fail(TAD_ERR_INVALID, "null argument");
    o->projection_eps = 1e-9;
    tad_solve_default_opts(&o->solve);
    o->ls_s_max = 1.0;
    o->ls_shrink = 0.8;
    o->ls_max_iters = 64;
    o->ls_armijo = 1e-4;
    o->ls_batch = 4;
    o->coarse_refresh = 1;
    return TAD_OK;
}

tad_status tad_line_search(tad_fn fn, const double* x0, const double* dir, double f, const double* g, double* x_out,
                           tad_mem mem, const tad_newton_opts* opts, double* f_out, double* step_out, int* trials_out)
{
    TAD_TRY(check_fn(fn));
    if (!x0 || !dir || !g || !x_out) return fail(TAD_ERR_INVALID, "null argument");
    if (mem != TAD_MEM_DEVICE) return fail(TAD_ERR_UNSUPPORTED, "tad_line_search takes device pointers");
    TAD_HIP(hipSetDevice(fn->ctx->device));
    TAD_TRY(ensure_ready(fn));
    tad_newton_opts o;
    tad_newton_default_opts(&o);
    if (opts) o = *opts;
    double dg = 0.0;
    TAD_TRY(dot_dev(fn, dir, g, &dg));
    double fo = 0, so = 0;
    int tr = 0;
    bool failed = false;
    TAD_TRY(line_search_core(fn, x0, dir, 1.0, f, dg, x_out, o, &fo, &so, &tr, &failed));
    if (f_out) *f_out = fo;
    if (step_out) *step_out = so;
    if (trials_out) *trials_out = failed ? -tr : tr;
    return TAD_OK;
}

tad_status tad_newton_decrement(tad_fn fn, const double* dir, const double* g, tad_mem mem, double* out)
{
    TAD_TRY(check_fn(fn));
    if (!dir || !g || !out) return fail(TAD_ERR_INVALID, "null argument");
    TAD_HIP(hipSetDevice(fn->ctx->device));
    TAD_TRY(ensure_ready(fn));
    const int64_t n = fn->n_vars;
    const double* dd = dir;
    const double* gd = g;
    if (mem == TAD_MEM_HOST)
    {
        TAD_HIP(fn->x_tmp.ensure(n));
        TAD_HIP(fn->sol.ensure(n));
        TAD_HIP(hipMemcpyAsync(fn->x_tmp.p, dir, n * sizeof(double), hipMemcpyHostToDevice, fn->s()));
        TAD_HIP(hipMemcpyAsync(fn->sol.p, g, n * sizeof(double), hipMemcpyHostToDevice, fn->s()));
        dd = fn->x_tmp.p;
        gd = fn->sol.p;
    }
    // the same fixed-order device dot product the Newton step uses, so the
    // decrement of tad_newton_stats is reproduced bit for bit
    double dg = 0.0;
    TAD_TRY(dot_dev(fn, dd, gd, &dg));
    *out = -0.5 * dg;
    return TAD_OK;
}

tad_status tad_newton_step(tad_fn fn, double* x, tad_mem mem, const tad_newton_opts* opts, tad_newton_stats* stats)
{
    TAD_TRY(check_fn(fn));
    if (!x) return fail(TAD_ERR_INVALID, "null argument");
    TAD_HIP(hipSetDevice(fn->ctx->device));
    TAD_TRY(ensure_ready(fn));
    tad_newton_opts o;
    tad_newton_default_opts(&o);
    if (opts) o = *opts;
    hipStream_t s = fn->s();
    const int64_t n = fn->n_vars;
    double* xd = x;
    if (mem == TAD_MEM_HOST)
    {
        TAD_HIP(hipMemcpyAsync(fn->x_buf.p, x, n * sizeof(double), hipMemcpyHostToDevice, s));
        xd = fn->x_buf.p;
    }
    tad_newton_stats st;
    std::memset(&st, 0, sizeof(st));
    double f = 0.0;
    // a NaN eps selects the unprojected Hessian, as eval_with_derivatives
    // passes eps = NAN to the shared evaluation (ScalarFunctionImpl.hh:333)
    const int mode = std::isnan(o.projection_eps) ? MODE_HESS : MODE_HESS_PROJ;
    TAD_TRY(eval_core(fn, xd, mode, o.projection_eps, fn->g_buf.p, fn->H_buf.p, &f));
    fn->h_valid = true;
    st.f = f;
    // solve (wI + H) d' = g; the Newton direction is d = -d'
    TAD_HIP(fn->sol.ensure(n));
    tad_solve_stats ss;
    std::memset(&ss, 0, sizeof(ss));
    tad_solve_opts sopt = o.solve;
    if (o.coarse_refresh > 1 && fn->solves_since_rebuild + 1 < o.coarse_refresh) sopt.coarse_reuse = 1;
    TAD_TRY(solve_core(fn, fn->H_buf.p, fn->g_buf.p, fn->sol.p, sopt, &ss));
    st.cg_iters = ss.iters;
    st.solve_converged = ss.converged;
    st.coarse_unknowns = ss.coarse_unknowns;
    st.coarse_reused = ss.coarse_reused;
    double dpg = 0.0, gg = 0.0;
    TAD_TRY(dot2_dev(fn, fn->sol.p, fn->g_buf.p, fn->g_buf.p, fn->g_buf.p, &dpg, &gg));
    const double dg = -dpg;  // d^T g
    st.grad_norm = std::sqrt(gg);
    st.newton_decrement = -0.5 * dg;
    double fo = 0, so = 0;
    int tr = 0;
    bool failed = false;
    TAD_TRY(line_search_core(fn, xd, fn->sol.p, -1.0, f, dg, xd, o, &fo, &so, &tr, &failed));
    st.f_new = fo;
    st.step = so;
    st.ls_trials = tr;
    st.ls_failed = failed ? 1 : 0;
    if (mem == TAD_MEM_HOST)
    {
        TAD_HIP(hipMemcpyAsync(x, xd, n * sizeof(double), hipMemcpyDeviceToHost, s));
        TAD_HIP(hipStreamSynchronize(s));
    }
    if (stats) *stats = st;
    return TAD_OK;
}

tad_status tad_fn_set_segments(tad_fn fn, int64_t n_segments, const int64_t* off)
{
    TAD_TRY(check_fn(fn));
    if (n_segments < 1 || !off) return fail(TAD_ERR_INVALID, "need at least one segment");
    if (off[0] != 0 || off[n_segments] != fn->n_handles)
        return fail(TAD_ERR_INVALID, "segment offsets must start at 0 and end at n_handles");
    for (int64_t k = 0; k < n_segments; ++k)
        if (off[k + 1] < off[k]) return fail(TAD_ERR_INVALID, "segment offsets must be non-decreasing");
    auto seg_of = [&](int32_t v) -> int64_t {
        return (int64_t)(std::upper_bound(off, off + n_segments + 1, (int64_t)v) - off) - 1;
    };
    const size_t nt = fn->hterms.size();
    std::vector<int64_t> ranges(nt * n_segments * 2, 0);
    for (size_t ti = 0; ti < nt; ++ti)
    {
        const HostTerm& t = fn->hterms[ti];
        const int64_t base = fn->terms[ti].f_base;
        for (int64_t k = 0; k < n_segments; ++k)
            ranges[(ti * n_segments + k) * 2] = ranges[(ti * n_segments + k) * 2 + 1] = base;
        int64_t prev = -1;
        for (int64_t e = 0; e < t.n_elem; ++e)
        {
            // padding entries (-1, dynamic-valence terms) belong to no
            // segment; an element without handles stays in the current one
            const int32_t* ce = t.conn.data() + e * t.N;
            const int64_t sg = ce[0] >= 0 ? seg_of(ce[0]) : std::max<int64_t>(prev, 0);
            for (int a = 1; a < t.N; ++a)
                if (ce[a] >= 0 && seg_of(ce[a]) != sg)
                    return fail(TAD_ERR_INVALID, "element " + std::to_string(e) + " spans two segments");
            if (sg < prev) return fail(TAD_ERR_INVALID, "elements of a term must be ordered by segment");
            if (sg != prev)
            {
                ranges[(ti * n_segments + sg) * 2] = base + e;
                prev = sg;
            }
            ranges[(ti * n_segments + sg) * 2 + 1] = base + e + 1;
        }
    }
    TAD_HIP(fn->seg_ranges.ensure(ranges.size()));
    TAD_HIP(fn->seg_f.ensure((size_t)n_segments));
    TAD_HIP(hipMemcpy(fn->seg_ranges.p, ranges.data(), ranges.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    TAD_HIP(hipMemset(fn->seg_f.p, 0, n_segments * sizeof(double)));
    fn->n_seg = n_segments;
    fn->seg_offsets.assign(off, off + n_segments + 1);
    return TAD_OK;
}

tad_status tad_fn_segment_energies(tad_fn fn, double* f, tad_mem mem)
{
    TAD_TRY(check_fn(fn));
    if (!f) return fail(TAD_ERR_INVALID, "null argument");
    if (fn->n_seg == 0) return fail(TAD_ERR_INVALID, "no segments declared");
    TAD_HIP(hipMemcpyAsync(f, fn->seg_f.p, fn->n_seg * sizeof(double),
                           mem == TAD_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, fn->s()));
    TAD_HIP(hipStreamSynchronize(fn->s()));
    return TAD_OK;
}

tad_status tad_fn_get_launch_shape(tad_fn fn, int term, int* block, int* grid, int* map)
{
    TAD_TRY(check_fn(fn));
    if (!block || !grid || !map) return fail(TAD_ERR_INVALID, "null argument");
    if (term < 0 || term >= (int)fn->terms.size()) return fail(TAD_ERR_INVALID, "term index out of range");
    const TermDev& t = fn->terms[term];
    *block = t.tune_block;
    *grid = t.tune_grid;
    *map = t.tune_map;
    return TAD_OK;
}

tad_status tad_fn_set_launch_shape(tad_fn fn, int term, int block, int grid, int map)
{
    TAD_TRY(check_fn(fn));
    if (term < 0 || term >= (int)fn->terms.size()) return fail(TAD_ERR_INVALID, "term index out of range");
    TermDev& t = fn->terms[term];
    if (block == 0)
    {
        t.tune_block = t.tune_grid = t.tune_map = 0;
        return TAD_OK;
    }
    if (!(t.energy == TAD_E_SYMDIR_TRI || t.energy == TAD_E_SYMDIR_TET))
        return fail(TAD_ERR_UNSUPPORTED, "only symmetric-Dirichlet terms have persistent launch shapes");
    if (block != 64 && block != 128 && block != 256) return fail(TAD_ERR_INVALID, "block must be 0, 64, 128 or 256");
    if (grid < 1 || grid > (1 << 20)) return fail(TAD_ERR_INVALID, "grid must be in 1 .. 2^20");
    if (map != 0 && map != 1) return fail(TAD_ERR_INVALID, "map must be 0 or 1");
    if (map == 1 && grid % 8 != 0) return fail(TAD_ERR_INVALID, "per-XCD map needs a grid that is a multiple of 8");
    t.tune_block = block;
    t.tune_grid = grid;
    t.tune_map = map;
    return TAD_OK;
}

tad_status tad_fn_get_assembly_shape(tad_fn fn, int* shape)
{
    TAD_TRY(check_fn(fn));
    if (!shape) return fail(TAD_ERR_INVALID, "null argument");
    *shape = fn->asm_shape;
    return TAD_OK;
}

tad_status tad_fn_set_assembly_shape(tad_fn fn, int shape)
{
    TAD_TRY(check_fn(fn));
    if (shape < 0 || shape >= kAsmShapes) return fail(TAD_ERR_INVALID, "assembly shape must be in 0 .. 3");
    fn->asm_shape = shape;
    return TAD_OK;
}

tad_status tad_fn_get_pipeline(tad_fn fn, int* requested, int* used)
{
    TAD_TRY(check_fn(fn));
    if (!requested && !used) return fail(TAD_ERR_INVALID, "null argument");
    if (requested) *requested = fn->pipe_chunks;
    if (used)
    {
        TAD_HIP(hipSetDevice(fn->ctx->device));
        TAD_TRY(ensure_ready(fn));
        if (!fn->pipe_ready) plan_pipeline(fn);
        *used = fn->pipe_used;
    }
    return TAD_OK;
}

tad_status tad_fn_set_solve_graphs(tad_fn fn, int mode)
{
    TAD_TRY(check_fn(fn));
    if (mode < 0 || mode > 1) return fail(TAD_ERR_INVALID, "solve graph mode must be 0 or 1");
    fn->solve_graphs = mode;
    return TAD_OK;
}

tad_status tad_fn_set_pipeline(tad_fn fn, int chunks)
{
    TAD_TRY(check_fn(fn));
    if (chunks < 0 || chunks > kPipeMaxChunks) return fail(TAD_ERR_INVALID, "pipeline chunks must be in 0..64");
    fn->pipe_chunks = chunks;
    fn->pipe_ready = false;
    return TAD_OK;
}

tad_status tad_fn_autotune(tad_fn fn, const double* x, tad_mem mem, double eps, char* report, size_t report_len)
{
    TAD_TRY(check_fn(fn));
    if (!x) return fail(TAD_ERR_INVALID, "null argument");
    TAD_HIP(hipSetDevice(fn->ctx->device));
    TAD_TRY(ensure_ready(fn));
    hipStream_t s = fn->s();
    const double* xd = stage_x(fn, x, mem);
    if (!xd) return fail(TAD_ERR_HIP, "copy of x to the device failed");
    int n_cu = 0;
    TAD_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, fn->ctx->device));
    // events released and launch shapes restored on every exit path; the new
    // shapes are committed only when every measurement succeeded
    struct Guard
    {
        explicit Guard(tad_fn f) : fn(f) {}
        tad_fn fn;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        std::vector<TermDev> saved;
        int saved_shape = 0, saved_chunks = 0;
        bool commit = false;
        ~Guard()
        {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (!commit)
            {
                for (size_t i = 0; i < fn->terms.size(); ++i)
                {
                    fn->terms[i].tune_block = saved[i].tune_block;
                    fn->terms[i].tune_grid = saved[i].tune_grid;
                    fn->terms[i].tune_map = saved[i].tune_map;
                }
                fn->asm_shape = saved_shape;
                fn->pipe_chunks = saved_chunks;
                fn->pipe_ready = false;
            }
        }
    } guard{fn};
    guard.saved = fn->terms;
    guard.saved_shape = fn->asm_shape;
    guard.saved_chunks = fn->pipe_chunks;
    TAD_HIP(hipEventCreate(&guard.e0));
    TAD_HIP(hipEventCreate(&guard.e1));
    hipEvent_t e0 = guard.e0, e1 = guard.e1;
    std::string rep;
    // whole evaluations, three in a row per timing (as a caller's loop runs
    // them): a launch shape's time in isolation, after an idle device, can
    // rank shapes differently from their time between the other kernels of
    // an evaluation
    auto time_evals = [&](float* out) -> tad_status {
        float best = 1e30f;
        for (int rep_i = 0; rep_i < 4; ++rep_i)
        {
            double fv = 0.0;
            TAD_HIP(hipEventRecord(e0, s));
            for (int b = 0; b < 3; ++b) TAD_TRY(eval_core(fn, xd, MODE_HESS_PROJ, eps, fn->g_buf.p, fn->H_buf.p, &fv));
            TAD_HIP(hipEventRecord(e1, s));
            TAD_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            TAD_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (rep_i > 0) best = std::min(best, ms / 3.f);
        }
        *out = best;
        return TAD_OK;
    };
    struct Cand { int block, grid, map; double ms; };
    std::vector<std::vector<Cand>> shortlist(fn->terms.size());
    for (size_t ti = 0; ti < fn->terms.size(); ++ti)
    {
        TermDev& t = fn->terms[ti];
        if (!(t.energy == TAD_E_SYMDIR_TRI || t.energy == TAD_E_SYMDIR_TET) || t.n_elem == 0) continue;
        std::vector<Cand> cands;
        cands.push_back({ 0, 0, 0, 0.0 });
        for (int b : { 64, 128, 256 })
            for (int wpc : { 1, 2, 4, 8 })
                for (int map : { 0, 1 })
                {
                    int grid = n_cu * wpc;
                    grid = (grid + 7) / 8 * 8;
                    cands.push_back({ b, grid, map, 0.0 });
                }
        for (Cand& c : cands)
        {
            t.tune_block = c.block;
            t.tune_grid = c.grid;
            t.tune_map = c.map;
            TAD_HIP(launch_element(t, fn->d, fn->precision, MODE_HESS_PROJ, xd, eps, fn->elem_f.p, fn->gc_buf.p, fn->hc_buf.p,
                                   fn->err.p, s));
            float best = 1e30f;
            for (int rep_i = 0; rep_i < 3; ++rep_i)
            {
                TAD_HIP(hipEventRecord(e0, s));
                TAD_HIP(launch_element(t, fn->d, fn->precision, MODE_HESS_PROJ, xd, eps, fn->elem_f.p, fn->gc_buf.p, fn->hc_buf.p,
                                       fn->err.p, s));
                TAD_HIP(hipEventRecord(e1, s));
                TAD_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                TAD_HIP(hipEventElapsedTime(&ms, e0, e1));
                best = std::min(best, ms);
            }
            c.ms = best;
        }
        const Cand* bestc = &cands[0];
        for (const Cand& c : cands)
            if (c.ms < bestc->ms) bestc = &c;
        t.tune_block = bestc->block;
        t.tune_grid = bestc->grid;
        t.tune_map = bestc->map;
        // finalists for the whole-evaluation round: the plain grid and the
        // two fastest persistent shapes
        std::vector<Cand> pers(cands.begin() + 1, cands.end());
        std::sort(pers.begin(), pers.end(), [](const Cand& a, const Cand& b) { return a.ms < b.ms; });
        shortlist[ti].push_back(cands[0]);
        for (size_t i = 0; i < pers.size() && i < 2; ++i) shortlist[ti].push_back(pers[i]);
        char line[256];
        std::snprintf(line, sizeof(line), "term energy %d (%lld elements): plain grid %.4f ms; best %s block %d grid %d map %s %.4f ms\n",
                      t.energy, (long long)t.n_elem, cands[0].ms, bestc->block ? "persistent" : "plain", bestc->block,
                      bestc->grid, bestc->map ? "xcd" : "strided", bestc->ms);
        rep += line;
        for (const Cand& c : cands)
        {
            std::snprintf(line, sizeof(line), "  block %3d grid %5d map %-7s %.4f ms\n", c.block, c.grid,
                          c.block == 0 ? "-" : (c.map ? "xcd" : "strided"), c.ms);
            rep += line;
        }
    }
    // Hessian assembly shape (on the contributions the last launch wrote)
    if (fn->n_blocks > 0)
    {
        float ms_shape[kAsmShapes];
        const HessAsm pat = fn->hess_asm();
        for (int shape = 0; shape < kAsmShapes; ++shape)
        {
            float best = 1e30f;
            for (int rep_i = 0; rep_i < 4; ++rep_i)
            {
                TAD_HIP(hipEventRecord(e0, s));
                TAD_HIP(launch_assemble_hessian(fn->d, fn->precision, shape, pat, fn->hc_buf.p, fn->H_buf.p, s));
                TAD_HIP(hipEventRecord(e1, s));
                TAD_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                TAD_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (rep_i > 0) best = std::min(best, ms);
            }
            ms_shape[shape] = best;
        }
        int bs = 0;
        for (int shape = 1; shape < kAsmShapes; ++shape)
            if (ms_shape[shape] < ms_shape[bs]) bs = shape;
        fn->asm_shape = bs;
        char line[256];
        std::snprintf(line, sizeof(line),
                      "Hessian assembly: %d lanes per block %.4f ms, one thread per block %.4f ms, "
                      "%d lanes per upper block %.4f ms, one thread per upper block %.4f ms\n",
                      fn->d * fn->d, ms_shape[0], ms_shape[1], fn->d * fn->d, ms_shape[2], ms_shape[3]);
        rep += line;
        // element launch shape by whole evaluations among the finalists
        fn->pipe_chunks = 1;
        fn->pipe_ready = false;
        for (size_t ti = 0; ti < fn->terms.size(); ++ti)
        {
            if (shortlist[ti].size() < 2) continue;
            TermDev& t = fn->terms[ti];
            const Cand* bestc = nullptr;
            float best_ms = 1e30f;
            std::string fl = "whole evaluations:";
            for (const Cand& c : shortlist[ti])
            {
                t.tune_block = c.block;
                t.tune_grid = c.grid;
                t.tune_map = c.map;
                float ms = 0.f;
                TAD_TRY(time_evals(&ms));
                std::snprintf(line, sizeof(line), " %s block %d grid %d %.4f ms;", c.block ? "persistent" : "plain",
                              c.block, c.grid, ms);
                fl += line;
                if (ms < best_ms)
                {
                    best_ms = ms;
                    bestc = &c;
                }
            }
            t.tune_block = bestc->block;
            t.tune_grid = bestc->grid;
            t.tune_map = bestc->map;
            rep += fl + "\n";
        }
        // pipelined evaluation: whole eval_with_hessian_proj with the chosen
        // shapes for 1, 2, 4 and 8 element chunks (identical results)
        int best_chunks = 1;
        float best_ms = 1e30f;
        std::string pl = "pipelined evaluation:";
        for (int chunks : { 1, 2, 4, 8 })
        {
            fn->pipe_chunks = chunks;
            plan_pipeline(fn);
            if (fn->pipe_used != chunks) continue;
            float best = 1e30f;
            TAD_TRY(time_evals(&best));
            std::snprintf(line, sizeof(line), " %d chunk%s %.4f ms;", chunks, chunks > 1 ? "s" : "", best);
            pl += line;
            if (best < best_ms)
            {
                best_ms = best;
                best_chunks = chunks;
            }
        }
        fn->pipe_chunks = best_chunks;
        fn->pipe_ready = false;
        rep += pl + "\n";
    }
    fn->h_valid = false;
    guard.commit = true;
    if (report && report_len > 0)
    {
        std::strncpy(report, rep.c_str(), report_len - 1);
        report[report_len - 1] = 0;
    }
    return TAD_OK;
}

tad_status tad_fn_set_params(tad_fn fn, int term, const double* params, tad_mem mem)
{
    TAD_TRY(check_fn(fn));
    if (term < 0 || term >= (int)fn->terms.size()) return fail(TAD_ERR_INVALID, "term index out of range");
    if (!params) return fail(TAD_ERR_INVALID, "null argument");
    TAD_HIP(hipSetDevice(fn->ctx->device));
    TermDev& t = fn->terms[term];
    const int64_t ne = t.n_elem;
    const int np = t.n_params;
    if (ne == 0 || np == 0) return TAD_OK;
    std::vector<double> host;
    const double* src = params;
    if (mem == TAD_MEM_DEVICE)
    {
        host.resize((size_t)ne * np);
        TAD_HIP(hipMemcpyAsync(host.data(), params, host.size() * sizeof(double), hipMemcpyDeviceToHost, fn->s()));
        TAD_HIP(hipStreamSynchronize(fn->s()));
        src = host.data();
    }
    std::vector<double> ps((size_t)ne * np);
    for (int64_t e = 0; e < ne; ++e)
        for (int k = 0; k < np; ++k) ps[k * ne + e] = src[e * np + k];
    TAD_HIP(hipMemcpyAsync(t.params, ps.data(), ps.size() * sizeof(double), hipMemcpyHostToDevice, fn->s()));
    TAD_HIP(hipStreamSynchronize(fn->s()));
    fn->h_valid = false;
    return TAD_OK;
}

tad_status tad_fn_rest_vjp(tad_fn fn, int term, const double* x, const double* rest, const double* lam,
                           double w_scale, double* out, tad_mem mem)
{
    TAD_TRY(check_fn(fn));
    if (term < 0 || term >= (int)fn->terms.size()) return fail(TAD_ERR_INVALID, "term index out of range");
    if (!x || !rest || !lam || !out) return fail(TAD_ERR_INVALID, "null argument");
    if (fn->terms[term].energy != TAD_E_SYMDIR_TET || fn->d != 3)
        return fail(TAD_ERR_UNSUPPORTED, "rest-shape products exist for symmetric-Dirichlet tets in 3D only");
    TAD_HIP(hipSetDevice(fn->ctx->device));
    TAD_TRY(ensure_ready(fn));
    hipStream_t s = fn->s();
    const size_t nv = (size_t)fn->n_vars;
    const double *xd = x, *rd = rest, *ld = lam;
    if (mem == TAD_MEM_HOST)
    {
        TAD_HIP(fn->adj_rest.ensure(nv));
        TAD_HIP(fn->adj_lam.ensure(nv));
        TAD_HIP(hipMemcpyAsync(fn->x_buf.p, x, nv * sizeof(double), hipMemcpyHostToDevice, s));
        TAD_HIP(hipMemcpyAsync(fn->adj_rest.p, rest, nv * sizeof(double), hipMemcpyHostToDevice, s));
        TAD_HIP(hipMemcpyAsync(fn->adj_lam.p, lam, nv * sizeof(double), hipMemcpyHostToDevice, s));
        xd = fn->x_buf.p;
        rd = fn->adj_rest.p;
        ld = fn->adj_lam.p;
    }
    // FP64 gradient slots: the evaluation's own slots in FP64 mode (their
    // contents were consumed by the assembly of the evaluation that wrote
    // them), a side buffer of the same layout in mixed mode.  Slots of other
    // terms are zero, so they add nothing.
    double* gs = nullptr;
    if (fn->precision == TAD_PREC_FP64) gs = (double*)fn->gc_buf.p;
    else
    {
        TAD_HIP(fn->adj_g.ensure((size_t)fn->n_gslots * fn->d));
        gs = fn->adj_g.p;
    }
    double* out_dev = (mem == TAD_MEM_DEVICE) ? out : fn->g_buf.p;
    TAD_HIP(launch_fill(gs, fn->n_gslots * fn->d, 0.0, s));
    TAD_HIP(launch_tet_rest_vjp(fn->terms[term], xd, rd, ld, w_scale, gs, s));
    TAD_HIP(launch_assemble_gradient(fn->d, 0, fn->n_handles, fn->gslot_ptr.p, gs, out_dev, s));
    if (mem == TAD_MEM_HOST) TAD_HIP(hipMemcpyAsync(out, out_dev, nv * sizeof(double), hipMemcpyDeviceToHost, s));
    TAD_HIP(hipStreamSynchronize(s));
    return TAD_OK;
}

tad_status tad_selftest_dual(tad_ctx ctx, int which, double x, double y, double* out)
{
    if (!ctx || !out) return fail(TAD_ERR_INVALID, "null argument");
    TAD_HIP(hipSetDevice(ctx->device));
    double* dptr = nullptr;
    TAD_HIP(hipMalloc((void**)&dptr, 8 * sizeof(double)));
    hipError_t e = launch_selftest_dual(which, x, y, dptr, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(ctx->pinned, dptr, 7 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(dptr);
    if (e != hipSuccess) return fail(TAD_ERR_HIP, hipGetErrorString(e));
    std::memcpy(out, ctx->pinned, 7 * sizeof(double));
    return TAD_OK;
}

} // extern "C"
