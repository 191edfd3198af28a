// Batched BiCGStab(2) for many independent structured systems on one GPU (gfx950).
//
// Each system is a 3-D 7-point operator assembled by dare_csr_fill_values (its slot-major
// diagonal view) with its own grid size. All systems' vectors are concatenated; every kernel
// is a persistent launch (a fixed grid looping over a work list) whose items are (system,
// tile) for the operator and (system, 1024-row chunk) for the vector updates. A per-system
// control block in device memory holds that system's BiCGStab2 scalars (the same update rules
// as the single-system solver, BiCGStab2.inl:167-582) and its convergence mask: items of a
// system that has converged, broken down or hit its pass limit return at once, and every
// per-system decision (stop test, limits, breakdown) is taken on the device. Passes are
// captured into a HIP graph and replayed; the host reads one active-system count per replay.
//
// Reductions stay deterministic: each item writes its partial sums to a fixed slot of its
// system, and the finish kernel (one work-group per system) adds them in slot order, whatever
// the execution order of the items.
//
// The mixed form runs FP32 inner solves (FP32 diagonal view, FP32 vectors) inside FP64
// iterative refinement per system, with the residual, the correction and the stop test in FP64.
//
// Tile shape (lanes along k: 16/32/64, the rest along j), planes per tile, persistent grid
// size and the order of the work list are a configuration chosen by dare_batch_autotune,
// which times the candidates on the batch itself.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "krylov_core.h"

namespace dare_hip {
namespace {

constexpr int kChunk = kThreads * kU;   // rows of one vector work item

struct BSys {
    DiaView V;
    const double* dia;     // FP64 slot-major values (the system's matrix)
    float* dia32;          // FP32 copy owned by the batch
    long long row0;        // first row of the system in the concatenated vectors
    int tk_log2;           // tile: 1 << tk_log2 lanes along k, kThreads >> tk_log2 along j
    int tile_i;            // planes per tile
    int tiles_k, tiles_j, tiles_i;
    int op_slot0, op_items;
    int vec_slot0, vec_items;
};

struct BItem {
    int sys;      // -1: a hole of the XCD-slab ordering, skipped
    int local;    // tile or chunk index within the system
};

// Per-system control block: Krylov scalars + masks + results.
struct BCtl {
    KState k;
    double tol_abs;       // threshold of the running Krylov solve
    double outer_tol;     // FP64 threshold of the system
    double bnorm;
    double rnorm;         // last FP64 residual norm (mixed) / true residual (plain, checked)
    int pass;             // pass number of the running Krylov solve
    int max_it;
    int active;           // the running Krylov solve still iterates
    int outer_active;     // refinement still running (mixed), or selected (plain set-up)
    int kconv;            // the last Krylov solve met its threshold
    int iters;            // passes of the last Krylov solve
    int total_iters;
    int refinements;
    int converged;
    int status;           // breakdown step of the last Krylov solve (0: none)
    int outer_status;     // mixed: breakdown that ended the refinement (0: none)
};

struct BParams {
    double tol, inner_tol;
    int tol_kind, max_it, inner_max_it, max_refine, min_passes, mixed;
};

enum BTail : int {
    BT_BNORM = 0,    // ||b||, thresholds, reset of the control block
    BT_INIT = 1,     // prelude of pass 1 + the tests before the first pass
    BT_GAMMA = 2,
    BT_RHO = 3,
    BT_GCR = 4,
    BT_END = 5,      // end of pass: stop test, limits, next-pass prelude
    BT_OUTER = 6,    // mixed: FP64 residual norm, start of the next inner solve
    BT_TRUE = 7,     // plain: true residual of systems that stopped on tau == 0
};

enum BVecOp : int { V_U, V_RX, V_V, V_URS, V_END, V_BNORM, V_ZERO, V_ADD, V_F32 };

template <typename VT>
__device__ inline const VT* dia_of(const BSys& S);
template <>
__device__ inline const double* dia_of<double>(const BSys& S) { return S.dia; }
template <>
__device__ inline const float* dia_of<float>(const BSys& S) { return S.dia32; }

// ------------------------------------------------------------------------------ operator
template <int NC, typename VT, typename XT, class Epi>
__global__ __launch_bounds__(kThreads) void kb_apply(const BSys* __restrict__ sys, const BItem* __restrict__ items,
                                                    int nitems, const XT* __restrict__ x, Epi epi, double* partial,
                                                    long long pstride, const BCtl* __restrict__ ctl, int outer_mask,
                                                    int check_status) {
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const BItem item = items[it];
        if (item.sys < 0) continue;
        const BCtl& c = ctl[item.sys];
        const int on = outer_mask ? c.outer_active : c.active;
        if (!on || (check_status && c.k.status)) continue;     // uniform over the work-group
        const BSys& S = sys[item.sys];
        double acc[Epi::NDOT];
#pragma unroll
        for (int q = 0; q < Epi::NDOT; q++) acc[q] = 0.;
        const int tk = item.local % S.tiles_k;
        const int t2 = item.local / S.tiles_k;
        const int tj = t2 % S.tiles_j;
        const int ti = t2 / S.tiles_j;
        const int k = (tk << S.tk_log2) + (threadIdx.x & ((1 << S.tk_log2) - 1));
        const int j = tj * (kThreads >> S.tk_log2) + (threadIdx.x >> S.tk_log2);
        if (k < S.V.n[2] && j < S.V.n[1]) {
            const int i0 = ti * S.tile_i;
            const int i1 = (i0 + S.tile_i < S.V.n[0]) ? i0 + S.tile_i : S.V.n[0];
            dia_column<NC, VT, XT, Epi>(S.V, dia_of<VT>(S), x + S.row0, S.row0, k, j, i0, i1, epi, acc);
        }
        block_partials_at<Epi::NDOT>(acc, partial, pstride, S.op_slot0 + item.local);
        __syncthreads();
    }
}

struct EpiTrueRes {     // ||b - A x||^2, nothing stored
    static constexpr int NDOT = 1;
    static constexpr bool kNT = false;
    const double* b;
    __device__ double load(long long i) const { return b[i]; }
    __device__ void operator()(long long, double y, double, double bi, double* acc) const {
        const double ri = bi - y;
        acc[0] += ri * ri;
    }
};

// ------------------------------------------------------------------------------ vectors
template <typename XT>
struct BVec {
    XT *x, *r, *rh, *u, *v, *s, *w, *t;
    double* x64;          // FP64 solution (refinement)
    const double* b64;    // FP64 right-hand side
    float* d32;           // FP32 correction (refinement)
};

template <typename XT, int OP>
__global__ __launch_bounds__(kThreads) void kb_vec(const BSys* __restrict__ sys, const BItem* __restrict__ items,
                                                  int nitems, BVec<XT> a, double* partial, long long pstride,
                                                  const BCtl* __restrict__ ctl) {
    constexpr bool kOuter = (OP == V_BNORM || OP == V_ADD || OP == V_F32);
    constexpr int kDots = (OP == V_END) ? 2 : 1;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const BItem item = items[it];
        if (item.sys < 0) continue;
        const BCtl& c = ctl[item.sys];
        const int on = kOuter ? c.outer_active : c.active;
        if (!on || (!kOuter && c.k.status)) continue;
        const BSys& S = sys[item.sys];
        const long long base = S.row0 + static_cast<long long>(item.local) * kChunk;
        const long long end = S.row0 + S.V.nrows;
        double acc[kDots];
#pragma unroll
        for (int q = 0; q < kDots; q++) acc[q] = 0.;
        if (OP == V_F32) {
            // FP32 copy of the system's diagonal view, slot by slot
            const long long nr = S.V.nrows;
            const int slots = S.V.cross ? 8 : 7;
            for (int sl = 0; sl < slots; sl++) {
#pragma unroll
                for (int q = 0; q < kU; q++) {
                    const long long i = base + q * kThreads + threadIdx.x;
                    if (i < end) {
                        const long long lr = i - S.row0;
                        S.dia32[sl * nr + lr] = static_cast<float>(S.dia[sl * nr + lr]);
                    }
                }
            }
            continue;
        }
        const double alpha = c.k.alpha, beta = c.k.beta, o1 = c.k.om1, o2 = c.k.om2;
        if (OP == V_U || OP == V_V) {
            XT* dst = (OP == V_U) ? a.u : a.v;
            const XT* src = (OP == V_U) ? a.r : a.s;
            double sv[kU], dv[kU];
#pragma unroll
            for (int q = 0; q < kU; q++) {
                const long long i0 = base + q * kThreads + threadIdx.x;
                const long long i = i0 < end ? i0 : base;
                sv[q] = static_cast<double>(src[i]);
                dv[q] = static_cast<double>(dst[i]);
            }
#pragma unroll
            for (int q = 0; q < kU; q++) {
                const long long i = base + q * kThreads + threadIdx.x;
                if (i < end) put<false>(dst + i, static_cast<XT>(sv[q] - beta * dv[q]));   // (5) / (12)
            }
        } else if (OP == V_RX) {
            double rv[kU], vv[kU], xv[kU], uv[kU];
#pragma unroll
            for (int q = 0; q < kU; q++) {
                const long long i0 = base + q * kThreads + threadIdx.x;
                const long long i = i0 < end ? i0 : base;
                rv[q] = static_cast<double>(a.r[i]);
                vv[q] = static_cast<double>(a.v[i]);
                xv[q] = static_cast<double>(a.x[i]);
                uv[q] = static_cast<double>(a.u[i]);
            }
#pragma unroll
            for (int q = 0; q < kU; q++) {
                const long long i = base + q * kThreads + threadIdx.x;
                if (i < end) {
                    put<false>(a.r + i, static_cast<XT>(rv[q] - alpha * vv[q]));       // (8)
                    put<false>(a.x + i, static_cast<XT>(1. * xv[q] + alpha * uv[q]));  // (10)
                }
            }
        } else if (OP == V_URS) {
            double uv[kU], rv[kU], vv[kU], sv[kU], wv[kU];
#pragma unroll
            for (int q = 0; q < kU; q++) {
                const long long i0 = base + q * kThreads + threadIdx.x;
                const long long i = i0 < end ? i0 : base;
                uv[q] = static_cast<double>(a.u[i]);
                rv[q] = static_cast<double>(a.r[i]);
                vv[q] = static_cast<double>(a.v[i]);
                sv[q] = static_cast<double>(a.s[i]);
                wv[q] = static_cast<double>(a.w[i]);
            }
#pragma unroll
            for (int q = 0; q < kU; q++) {
                const long long i = base + q * kThreads + threadIdx.x;
                if (i < end) {
                    put<false>(a.u + i, static_cast<XT>(rv[q] - beta * uv[q]));        // (15)
                    put<false>(a.r + i, static_cast<XT>(rv[q] - alpha * vv[q]));       // (16)
                    put<false>(a.s + i, static_cast<XT>(sv[q] - alpha * wv[q]));       // (17)
                }
            }
        } else if (OP == V_END) {
            double xv[kU], rv[kU], sv[kU], tv[kU], uv[kU], vv[kU], wv[kU], hv[kU];
#pragma unroll
            for (int q = 0; q < kU; q++) {
                const long long i0 = base + q * kThreads + threadIdx.x;
                const long long i = i0 < end ? i0 : base;
                xv[q] = static_cast<double>(a.x[i]);
                rv[q] = static_cast<double>(a.r[i]);
                sv[q] = static_cast<double>(a.s[i]);
                tv[q] = static_cast<double>(a.t[i]);
                uv[q] = static_cast<double>(a.u[i]);
                vv[q] = static_cast<double>(a.v[i]);
                wv[q] = static_cast<double>(a.w[i]);
                hv[q] = static_cast<double>(a.rh[i]);
            }
#pragma unroll
            for (int q = 0; q < kU; q++) {
                const long long i = base + q * kThreads + threadIdx.x;
                if (i < end) {
                    double xi = 1. * xv[q] + o1 * rv[q] + o2 * sv[q];                 // (24)
                    xi = 1. * xi + alpha * uv[q];
                    put<false>(a.x + i, static_cast<XT>(xi));
                    const XT rn = static_cast<XT>(1. * rv[q] + (-o1) * sv[q] + (-o2) * tv[q]);   // (25)
                    put<false>(a.r + i, rn);
                    put<false>(a.u + i, static_cast<XT>(1. * uv[q] + (-o1) * vv[q] + (-o2) * wv[q]));
                    const double rd = static_cast<double>(rn);
                    acc[0] += rd * rd;
                    acc[kDots - 1] += hv[q] * rd;
                }
            }
        } else if (OP == V_BNORM) {
            double bv[kU];
#pragma unroll
            for (int q = 0; q < kU; q++) {
                const long long i0 = base + q * kThreads + threadIdx.x;
                bv[q] = a.b64[i0 < end ? i0 : base];
            }
#pragma unroll
            for (int q = 0; q < kU; q++)
                if (base + q * kThreads + threadIdx.x < end) acc[0] += bv[q] * bv[q];
        } else if (OP == V_ZERO) {
#pragma unroll
            for (int q = 0; q < kU; q++) {
                const long long i = base + q * kThreads + threadIdx.x;
                if (i < end) a.x[i] = XT(0);
            }
        } else if (OP == V_ADD) {
#pragma unroll
            for (int q = 0; q < kU; q++) {
                const long long i = base + q * kThreads + threadIdx.x;
                if (i < end) a.x64[i] = a.x64[i] + static_cast<double>(a.d32[i]);
            }
        }
        if (OP == V_END || OP == V_BNORM) {
            block_partials_at<kDots>(acc, partial, pstride, S.vec_slot0 + item.local);
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------------------------ finish
// End of a Krylov solve after `iters` passes; like the single-system driver, the solve counts
// as converged when the last residual norm it saw meets the threshold, whatever stopped it.
__device__ inline void stop_solve(BCtl* c, int iters, int status) {
    c->active = 0;
    c->iters = iters;
    if (status) c->status = status;
    c->kconv = c->k.rnorm <= c->tol_abs ? 1 : 0;
}

// One work-group per system: the system's partials in slot order, then the scalar update and
// the per-system decisions of the host loop of the single-system solver.
template <int NDOT>
__global__ __launch_bounds__(kThreads) void kb_finish(const BSys* __restrict__ sys, const double* __restrict__ partial,
                                                     long long pstride, int vec_slots, BCtl* ctl, int tail,
                                                     int outer_mask, int check_status, BParams P) {
    const int s = blockIdx.x;
    BCtl* c = ctl + s;
    const int on = outer_mask ? c->outer_active : c->active;
    if (!on) return;
    if (check_status && c->k.status) {          // broke down inside the pass
        if (tail == BT_END && threadIdx.x == 0) stop_solve(c, c->pass, c->k.status);
        return;
    }
    const BSys& S = sys[s];
    const int slot0 = vec_slots ? S.vec_slot0 : S.op_slot0;
    const int nsl = vec_slots ? S.vec_items : S.op_items;
    __shared__ double red[kMaxDots][kThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NDOT; j++) {
        double a = 0.;
        for (int q = threadIdx.x; q < nsl; q += kThreads) a += partial[j * pstride + slot0 + q];
        a = wave_sum(a);
        if (lane == 0) red[j][wid] = a;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double d[kMaxDots];
    for (int j = 0; j < kMaxDots; j++) {
        double t = 0.;
        if (j < NDOT)
            for (int w = 0; w < kThreads / 64; w++) t += red[j][w];
        d[j] = t;
    }
    switch (tail) {
    case BT_BNORM:
        c->bnorm = sqrt(d[0]);
        c->outer_tol = P.tol_kind == 1 ? P.tol * c->bnorm : P.tol;
        c->total_iters = 0;
        c->refinements = 0;
        c->converged = 0;
        c->status = 0;
        c->outer_status = 0;
        c->kconv = 0;
        c->iters = 0;
        c->rnorm = 0.;
        if (!P.mixed) {            // plain: the Krylov solve runs on the system itself
            c->active = 1;
            c->tol_abs = c->outer_tol;
            c->max_it = P.max_it;
            c->pass = 1;
        }
        break;
    case BT_INIT: {
        apply_tail(TAIL_INIT, &c->k, d);
        c->pass = 1;
        c->status = 0;
        c->kconv = 0;
        if (c->k.rnorm < c->tol_abs) stop_solve(c, 0, 0);
        else if (1 > c->max_it) stop_solve(c, 1, 0);
        else if (c->k.next_status) stop_solve(c, 1, c->k.next_status);
        break;
    }
    case BT_GAMMA: apply_tail(TAIL_GAMMA, &c->k, d); break;
    case BT_RHO: apply_tail(TAIL_RHO, &c->k, d); break;
    case BT_GCR: apply_tail(TAIL_GCR, &c->k, d); break;
    case BT_END: {
        apply_tail(TAIL_END, &c->k, d);
        const int k = c->pass;
        if (c->k.rnorm <= c->tol_abs && k >= P.min_passes) {
            stop_solve(c, k, 0);
        } else {
            c->pass = k + 1;
            if (k + 1 > c->max_it) stop_solve(c, k + 1, 0);
            else if (c->k.next_status) stop_solve(c, k + 1, c->k.next_status);
        }
        break;
    }
    case BT_OUTER: {
        const double rn = sqrt(d[0]);
        c->rnorm = rn;
        if (rn <= c->outer_tol) {
            c->converged = 1;
            c->outer_active = 0;
        } else if (c->refinements >= P.max_refine || c->total_iters >= P.max_it) {
            c->outer_active = 0;
        } else {                  // next inner FP32 solve on the residual
            c->active = 1;
            c->tol_abs = inner_threshold(P.inner_tol, rn, c->outer_tol);
            c->max_it = P.inner_max_it;
            c->pass = 1;
            c->status = 0;
            c->kconv = 0;
            c->iters = 0;
            c->k.status = 0;
            c->k.next_status = 0;
        }
        break;
    }
    case BT_TRUE:             // the recomputed residual decides (BiCGStab2.inl:514-536)
        c->rnorm = sqrt(d[0]);
        c->converged = c->rnorm <= c->outer_tol ? 1 : 0;
        if (c->converged) c->status = 0;
        c->outer_active = 0;
        break;
    default: break;
    }
}

// Control steps that need no reduction (one thread per system).
enum BCtlOp : int { C_BEGIN, C_AFTER_INNER, C_PLAIN_DONE };

__global__ __launch_bounds__(kThreads) void kb_ctl(BCtl* ctl, int nsys, int op) {
    const int s = blockIdx.x * kThreads + threadIdx.x;
    if (s >= nsys) return;
    BCtl* c = ctl + s;
    if (op == C_BEGIN) {
        c->outer_active = 1;
        c->active = 0;
        c->k.status = 0;
        c->k.next_status = 0;
    } else if (op == C_AFTER_INNER) {
        if (!c->outer_active) return;
        c->total_iters += c->iters;
        c->refinements += 1;
        if (c->status && c->iters <= 1) {      // breakdown at the first pass ends the refinement
            c->outer_active = 0;
            c->outer_status = c->status;
        }
    } else {   // C_PLAIN_DONE: results of the plain form; select tau == 0 stops for a true residual
        c->total_iters = c->iters;
        c->converged = c->kconv;
        c->rnorm = c->k.rnorm;
        c->outer_active = (c->status == 21) ? 1 : 0;
    }
}

__global__ __launch_bounds__(kThreads) void kb_count(const BCtl* __restrict__ ctl, int nsys, int outer_mask,
                                                    int* out) {
    __shared__ int red[kThreads / 64];
    int n = 0;
    for (int s = threadIdx.x; s < nsys; s += kThreads) n += outer_mask ? ctl[s].outer_active : ctl[s].active;
    for (int off = 32; off > 0; off >>= 1) n += __shfl_down(n, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < kThreads / 64; w++) t += red[w];
        out[0] = t;
    }
}

}  // namespace
}  // namespace dare_hip

using namespace dare_hip;

struct dare_batch {
    int device = 0;
    int nsys = 0;
    int ncomp = 1;
    std::vector<dare_csr*> sys;
    std::vector<long long> row0;       // nsys + 1
    long long rows = 0;
    hipStream_t stream = nullptr;      // private stream: graph capture needs a non-null stream
    hipEvent_t ev_in = nullptr, ev_out = nullptr, ev_t0 = nullptr, ev_t1 = nullptr;
    dare_batch_config cfg{};
    // device tables
    BSys* d_sys = nullptr;
    BItem* d_op = nullptr;
    BItem* d_vec = nullptr;
    int n_op = 0, n_vec = 0;           // work-list lengths (holes included)
    long long pstride = 0;
    double* partial = nullptr;         // kMaxDots * pstride
    BCtl* ctl = nullptr;
    int* d_count = nullptr;
    int* h_count = nullptr;            // pinned
    BCtl* h_ctl = nullptr;             // pinned, nsys
    // vectors
    double* b = nullptr;
    double* x = nullptr;
    double* work64 = nullptr;          // 7 * rows (plain form)
    float* work32 = nullptr;           // 9 * rows (mixed form: 7 Krylov + r32 + d32)
    float* dia32 = nullptr;            // 8 * rows
    std::vector<BSys> h_sys;
    // graphs (rebuilt when the configuration changes)
    hipGraphExec_t g64 = nullptr, g32 = nullptr;
    int g64_passes = 0, g32_passes = 0;
    BParams gp64{}, gp32{};            // parameters each graph was captured with
};

namespace {

int tk_auto(int nz) { return nz <= 16 ? 4 : (nz <= 32 ? 5 : 6); }

void drop_graphs(dare_batch* B) {
    if (B->g64) hipGraphExecDestroy(B->g64);
    if (B->g32) hipGraphExecDestroy(B->g32);
    B->g64 = B->g32 = nullptr;
    B->g64_passes = B->g32_passes = 0;
}

// Work list in the configured order. order 0: system-major; 1: system-major split into eight
// contiguous slabs, one per XCD (block b runs on XCD b mod 8 and takes items b, b + G, ...);
// 2: as 1 with the systems sorted by size, largest first.
std::vector<BItem> make_items(const std::vector<int>& counts, const std::vector<int>& sys_order, int order) {
    std::vector<BItem> list;
    for (int s : sys_order)
        for (int l = 0; l < counts[s]; l++) list.push_back(BItem{s, l});
    if (order == 0) return list;
    const long long n = static_cast<long long>(list.size());
    const long long per = (n + kXcd - 1) / kXcd;
    std::vector<BItem> out(static_cast<size_t>(per * kXcd), BItem{-1, 0});
    for (long long p = 0; p < n; p++) out[(p % per) * kXcd + p / per] = list[p];
    return out;
}

int apply_config(dare_ctx* ctx, dare_batch* B, const dare_batch_config& cfg) {
    DARE_CHECK_ARG(cfg.tile_k == 0 || cfg.tile_k == 16 || cfg.tile_k == 32 || cfg.tile_k == 64,
                   "tile_k must be 0 (by nz), 16, 32 or 64");
    DARE_CHECK_ARG(cfg.tile_i >= 1 && cfg.tile_i <= 4096, "tile_i out of range");
    DARE_CHECK_ARG(cfg.grid >= kXcd && cfg.grid <= 65536 && cfg.grid % kXcd == 0,
                   "grid must be a multiple of 8 in [8, 65536]");
    DARE_CHECK_ARG(cfg.order >= 0 && cfg.order <= 2, "order must be 0, 1 or 2");
    std::vector<int> op_counts(B->nsys), vec_counts(B->nsys);
    long long op_total = 0, vec_total = 0;
    for (int s = 0; s < B->nsys; s++) {
        const DiaView& V = B->sys[s]->dia_view;
        BSys& S = B->h_sys[s];
        S.V = V;
        S.dia = B->sys[s]->dia;
        S.dia32 = B->dia32 ? B->dia32 + 8 * B->row0[s] : nullptr;
        S.row0 = B->row0[s];
        S.tk_log2 = cfg.tile_k ? (cfg.tile_k == 16 ? 4 : cfg.tile_k == 32 ? 5 : 6) : tk_auto(V.n[2]);
        S.tile_i = std::min(cfg.tile_i, V.n[0]);
        S.tiles_k = (V.n[2] + (1 << S.tk_log2) - 1) >> S.tk_log2;
        const int tj = kThreads >> S.tk_log2;
        S.tiles_j = (V.n[1] + tj - 1) / tj;
        S.tiles_i = (V.n[0] + S.tile_i - 1) / S.tile_i;
        const long long items = static_cast<long long>(S.tiles_k) * S.tiles_j * S.tiles_i;
        const long long vitems = (V.nrows + kChunk - 1) / kChunk;
        DARE_CHECK_ARG(op_total + items < (1LL << 31) && vec_total + vitems < (1LL << 31), "batch too large");
        S.op_slot0 = static_cast<int>(op_total);
        S.op_items = static_cast<int>(items);
        S.vec_slot0 = static_cast<int>(vec_total);
        S.vec_items = static_cast<int>(vitems);
        op_counts[s] = S.op_items;
        vec_counts[s] = S.vec_items;
        op_total += items;
        vec_total += vitems;
    }
    std::vector<int> order(B->nsys);
    std::iota(order.begin(), order.end(), 0);
    if (cfg.order == 2)
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return B->h_sys[a].V.nrows > B->h_sys[b].V.nrows; });
    const std::vector<BItem> op = make_items(op_counts, order, cfg.order);
    const std::vector<BItem> vec = make_items(vec_counts, order, cfg.order);
    const long long pstride = std::max(op_total, vec_total);
    if (pstride > B->pstride || !B->partial) {
        if (B->partial) hipFree(B->partial);
        B->partial = nullptr;
        B->pstride = 0;
        DARE_HIP_TRY(hipMalloc(&B->partial, sizeof(double) * kMaxDots * pstride));
        B->pstride = pstride;
    }
    if (static_cast<int>(op.size()) > B->n_op || !B->d_op) {
        if (B->d_op) hipFree(B->d_op);
        B->d_op = nullptr;
        DARE_HIP_TRY(hipMalloc(&B->d_op, sizeof(BItem) * op.size()));
    }
    if (static_cast<int>(vec.size()) > B->n_vec || !B->d_vec) {
        if (B->d_vec) hipFree(B->d_vec);
        B->d_vec = nullptr;
        DARE_HIP_TRY(hipMalloc(&B->d_vec, sizeof(BItem) * vec.size()));
    }
    DARE_HIP_TRY(hipStreamSynchronize(B->stream));
    DARE_HIP_TRY(hipMemcpy(B->d_op, op.data(), sizeof(BItem) * op.size(), hipMemcpyHostToDevice));
    DARE_HIP_TRY(hipMemcpy(B->d_vec, vec.data(), sizeof(BItem) * vec.size(), hipMemcpyHostToDevice));
    DARE_HIP_TRY(hipMemcpy(B->d_sys, B->h_sys.data(), sizeof(BSys) * B->nsys, hipMemcpyHostToDevice));
    B->n_op = static_cast<int>(op.size());
    B->n_vec = static_cast<int>(vec.size());
    B->cfg = cfg;
    drop_graphs(B);
    (void)ctx;
    return DARE_OK;
}

template <typename T>
int grow(T** p, size_t n) {
    if (*p) return DARE_OK;
    DARE_HIP_TRY(hipMalloc(p, sizeof(T) * (n ? n : 1)));
    return DARE_OK;
}

// refresh the per-system value pointers (a matrix may have been refilled or regrown)
int refresh_tables(dare_batch* B) {
    bool changed = false;
    for (int s = 0; s < B->nsys; s++) {
        const dare_csr* A = B->sys[s];
        DARE_CHECK_ARG(A->has_dia && A->dia, "a batch system has no assembled values (call dare_csr_fill_values)");
        DARE_CHECK_ARG(!A->ent_lazy, "a batch system's view lacks its E, N, T slots (its structure was rebuilt "
                                     "after the batch was created)");
        float* d32 = B->dia32 ? B->dia32 + 8 * B->row0[s] : nullptr;
        if (B->h_sys[s].dia != A->dia || B->h_sys[s].dia32 != d32) {
            B->h_sys[s].dia = A->dia;
            B->h_sys[s].dia32 = d32;
            changed = true;
        }
    }
    if (changed) {
        DARE_HIP_TRY(hipStreamSynchronize(B->stream));
        DARE_HIP_TRY(hipMemcpy(B->d_sys, B->h_sys.data(), sizeof(BSys) * B->nsys, hipMemcpyHostToDevice));
        drop_graphs(B);
    }
    return DARE_OK;
}

template <int NC, typename VT, typename XT, class Epi>
void launch_op(dare_batch* B, const XT* x, const Epi& epi, int outer_mask, int check) {
    hipLaunchKernelGGL((kb_apply<NC, VT, XT, Epi>), dim3(B->cfg.grid), dim3(kThreads), 0, B->stream, B->d_sys,
                       B->d_op, B->n_op, x, epi, B->partial, B->pstride, B->ctl, outer_mask, check);
}

template <typename XT, int OP>
void launch_vec(dare_batch* B, const BVec<XT>& a) {
    hipLaunchKernelGGL((kb_vec<XT, OP>), dim3(B->cfg.grid), dim3(kThreads), 0, B->stream, B->d_sys, B->d_vec,
                       B->n_vec, a, B->partial, B->pstride, B->ctl);
}

template <int NDOT>
void launch_fin(dare_batch* B, int vec_slots, int tail, int outer_mask, int check, const BParams& P) {
    hipLaunchKernelGGL((kb_finish<NDOT>), dim3(B->nsys), dim3(kThreads), 0, B->stream, B->d_sys, B->partial,
                       B->pstride, vec_slots, B->ctl, tail, outer_mask, check, P);
}

void launch_ctl(dare_batch* B, int op) {
    hipLaunchKernelGGL(kb_ctl, dim3((B->nsys + kThreads - 1) / kThreads), dim3(kThreads), 0, B->stream, B->ctl,
                       B->nsys, op);
}

// One BiCGStab2 pass for every active system (BiCGStab2.inl steps 5-25), 14 launches.
template <int NC, typename VT, typename XT>
void enqueue_pass(dare_batch* B, const BVec<XT>& a, const BParams& P) {
    launch_vec<XT, V_U>(B, a);
    launch_op<NC, VT, XT>(B, a.u, EpiDot<XT>{a.v, a.rh}, 0, 1);
    launch_fin<1>(B, 0, BT_GAMMA, 0, 1, P);
    launch_vec<XT, V_RX>(B, a);
    launch_op<NC, VT, XT>(B, a.r, EpiDot<XT>{a.s, a.rh}, 0, 1);
    launch_fin<1>(B, 0, BT_RHO, 0, 1, P);
    launch_vec<XT, V_V>(B, a);
    launch_op<NC, VT, XT>(B, a.v, EpiDot<XT>{a.w, a.rh}, 0, 1);
    launch_fin<1>(B, 0, BT_GAMMA, 0, 1, P);
    launch_vec<XT, V_URS>(B, a);
    launch_op<NC, VT, XT>(B, a.s, EpiGcr<XT>{a.t, a.r}, 0, 1);
    launch_fin<5>(B, 0, BT_GCR, 0, 1, P);
    launch_vec<XT, V_END>(B, a);
    launch_fin<2>(B, 1, BT_END, 0, 1, P);
}

int read_count(dare_batch* B, int outer_mask, int* out) {
    hipLaunchKernelGGL(kb_count, dim3(1), dim3(kThreads), 0, B->stream, B->ctl, B->nsys, outer_mask, B->d_count);
    DARE_HIP_TRY(hipMemcpyAsync(B->h_count, B->d_count, sizeof(int), hipMemcpyDeviceToHost, B->stream));
    DARE_HIP_TRY(hipStreamSynchronize(B->stream));
    *out = *B->h_count;
    return DARE_OK;
}

// Iterate the running Krylov solves until no system is active: graph replays of `passes`
// passes (or plain launches when passes == 0), one count read-back per replay.
template <int NC, typename VT, typename XT>
int iterate(dare_batch* B, const BVec<XT>& a, const BParams& P, int passes, hipGraphExec_t* gexec, int* gpasses,
            BParams* gparams, int max_rounds) {
    const bool use_graph = passes > 0;
    if (use_graph && (*gexec == nullptr || *gpasses != passes || std::memcmp(gparams, &P, sizeof(BParams)) != 0)) {
        if (*gexec) hipGraphExecDestroy(*gexec);
        *gexec = nullptr;
        hipGraph_t g = nullptr;
        DARE_HIP_TRY(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
        for (int q = 0; q < passes; q++) enqueue_pass<NC, VT, XT>(B, a, P);
        hipLaunchKernelGGL(kb_count, dim3(1), dim3(kThreads), 0, B->stream, B->ctl, B->nsys, 0, B->d_count);
        DARE_HIP_TRY(hipStreamEndCapture(B->stream, &g));
        const hipError_t e = hipGraphInstantiate(gexec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        DARE_HIP_TRY(e);
        *gpasses = passes;
        *gparams = P;
    }
    for (int round = 0; round < max_rounds; round++) {
        int active = 0;
        if (use_graph) {
            DARE_HIP_TRY(hipGraphLaunch(*gexec, B->stream));
            DARE_HIP_TRY(hipMemcpyAsync(B->h_count, B->d_count, sizeof(int), hipMemcpyDeviceToHost, B->stream));
            DARE_HIP_TRY(hipStreamSynchronize(B->stream));
            active = *B->h_count;
        } else {
            enqueue_pass<NC, VT, XT>(B, a, P);
            DARE_LAUNCH_CHECK();
            int rc = read_count(B, 0, &active);
            if (rc) return rc;
        }
        if (active == 0) return DARE_OK;
    }
    set_error("batched solve did not finish within its pass budget");
    return DARE_ERR_HIP;
}

template <int NC>
int solve_impl(dare_batch* B, const BParams& P, int passes, dare_solver_result* res) {
    const long long R = B->rows;
    const int max_rounds_plain = (P.max_it + 2) / std::max(passes, 1) + 2;
    BVec<double> a64{};
    a64.x64 = B->x;
    a64.b64 = B->b;
    launch_ctl(B, C_BEGIN);
    launch_vec<double, V_BNORM>(B, a64);
    launch_fin<1>(B, 1, BT_BNORM, 1, 0, P);
    if (!P.mixed) {
        int rc = grow(&B->work64, static_cast<size_t>(7 * R));
        if (rc) return rc;
        double* w = B->work64;
        BVec<double> a{B->x, w, w + R, w + 2 * R, w + 3 * R, w + 4 * R, w + 5 * R, w + 6 * R, B->x, B->b, nullptr};
        launch_op<NC, double, double>(B, B->x, EpiInit<double>{B->b, a.r, a.rh, a.u, a.v, a.w}, 0, 0);
        launch_fin<2>(B, 0, BT_INIT, 0, 0, P);
        DARE_LAUNCH_CHECK();
        rc = iterate<NC, double, double>(B, a, P, passes, &B->g64, &B->g64_passes, &B->gp64, max_rounds_plain);
        if (rc) return rc;
        launch_ctl(B, C_PLAIN_DONE);
        int need_true = 0;
        rc = read_count(B, 1, &need_true);
        if (rc) return rc;
        if (need_true) {      // tau == 0: true residual check (BiCGStab2.inl:514-536)
            launch_op<NC, double, double>(B, B->x, EpiTrueRes{B->b}, 1, 0);
            launch_fin<1>(B, 0, BT_TRUE, 1, 0, P);
            DARE_LAUNCH_CHECK();
        }
    } else {
        int rc = grow(&B->work32, static_cast<size_t>(9 * R));
        if (rc) return rc;
        rc = grow(&B->dia32, static_cast<size_t>(8 * R));
        if (rc) return rc;
        rc = refresh_tables(B);
        if (rc) return rc;
        float* w = B->work32;
        float* r32 = w + 7 * R;
        float* d32 = w + 8 * R;
        BVec<float> a{d32, w, w + R, w + 2 * R, w + 3 * R, w + 4 * R, w + 5 * R, w + 6 * R, B->x, B->b, d32};
        launch_vec<float, V_F32>(B, a);                   // FP32 copy of every system's values
        const int max_rounds_inner = (P.inner_max_it + 2) / std::max(passes, 1) + 2;
        for (int ref = 0; ref <= P.max_refine; ref++) {
            launch_op<NC, double, double>(B, B->x, EpiResid<>{B->b, r32}, 1, 0);     // r = b - A x (FP64)
            launch_fin<1>(B, 0, BT_OUTER, 1, 0, P);
            DARE_LAUNCH_CHECK();
            int outer = 0;
            rc = read_count(B, 1, &outer);
            if (rc) return rc;
            if (outer == 0) break;
            launch_vec<float, V_ZERO>(B, a);
            launch_op<NC, float, float>(B, d32, EpiInit<float>{r32, a.r, a.rh, a.u, a.v, a.w}, 0, 0);
            launch_fin<2>(B, 0, BT_INIT, 0, 0, P);
            DARE_LAUNCH_CHECK();
            rc = iterate<NC, float, float>(B, a, P, passes, &B->g32, &B->g32_passes, &B->gp32, max_rounds_inner);
            if (rc) return rc;
            launch_vec<float, V_ADD>(B, a);                // x += d (FP64)
            launch_ctl(B, C_AFTER_INNER);
            DARE_LAUNCH_CHECK();
        }
    }
    DARE_HIP_TRY(hipMemcpyAsync(B->h_ctl, B->ctl, sizeof(BCtl) * B->nsys, hipMemcpyDeviceToHost, B->stream));
    DARE_HIP_TRY(hipStreamSynchronize(B->stream));
    for (int s = 0; s < B->nsys; s++) {
        const BCtl& c = B->h_ctl[s];
        dare_solver_result& r = res[s];
        r.converged = c.converged;
        r.iterations = c.total_iters;
        r.status = (P.mixed ? c.outer_status : c.status) ? DARE_ERR_BREAKDOWN : 0;
        r.refinements = c.refinements;
        r.residual = c.rnorm;
        r.bnorm = c.bnorm;
        r.history_len = 0;
    }
    return DARE_OK;
}

int sync_in(dare_ctx* ctx, dare_batch* B) {
    DARE_HIP_TRY(hipEventRecord(B->ev_in, ctx->stream));
    DARE_HIP_TRY(hipStreamWaitEvent(B->stream, B->ev_in, 0));
    return DARE_OK;
}

int sync_out(dare_ctx* ctx, dare_batch* B) {
    DARE_HIP_TRY(hipEventRecord(B->ev_out, B->stream));
    DARE_HIP_TRY(hipStreamWaitEvent(ctx->stream, B->ev_out, 0));
    return DARE_OK;
}

// time `reps` FP32 operator sweeps (SpMV + one fused dot over every system)
template <int NC>
int time_op(dare_batch* B, int reps, double* us) {
    float* w = B->work32;
    const long long R = B->rows;
    BVec<float> a{w + 8 * R, w, w + R, w + 2 * R, w + 3 * R, w + 4 * R, w + 5 * R, w + 6 * R, B->x, B->b, w + 8 * R};
    launch_op<NC, float, float>(B, a.r, EpiDot<float>{a.v, a.rh}, 1, 0);
    DARE_HIP_TRY(hipEventRecord(B->ev_t0, B->stream));
    for (int q = 0; q < reps; q++) launch_op<NC, float, float>(B, a.r, EpiDot<float>{a.v, a.rh}, 1, 0);
    DARE_HIP_TRY(hipEventRecord(B->ev_t1, B->stream));
    DARE_LAUNCH_CHECK();
    DARE_HIP_TRY(hipEventSynchronize(B->ev_t1));
    float ms = 0.f;
    DARE_HIP_TRY(hipEventElapsedTime(&ms, B->ev_t0, B->ev_t1));
    *us = 1e3 * ms / reps;
    return DARE_OK;
}

}  // namespace

extern "C" {

int dare_batch_create(dare_ctx* ctx, int32_t nsys, dare_csr* const* systems, dare_batch** out) {
    DARE_CHECK_ARG(ctx && systems && out && nsys > 0, "invalid batch arguments");
    *out = nullptr;
    int N = -1;
    for (int s = 0; s < nsys; s++) {
        DARE_CHECK_ARG(systems[s] != nullptr, "NULL system");
        DARE_CHECK_ARG(systems[s]->stencil_structured && systems[s]->has_dia && systems[s]->dia,
                       "batch systems must be 3-D structured operators with a diagonal view "
                       "(dare_csr_build_structure on a 3-D grid with 1 or 2 components)");
        if (N < 0) N = systems[s]->dia_view.N;
        DARE_CHECK_ARG(systems[s]->dia_view.N == N, "all systems of a batch need the same number of components");
    }
    DARE_HIP_TRY(hipSetDevice(ctx->device));
    // the batched traversals read every slot: views a fill left without E, N, T get them now,
    // and later fills of these matrices write them
    for (int s = 0; s < nsys; s++) {
        const int erc = dare_hip::dia_ensure_full(systems[s], ctx->stream);
        if (erc) return erc;
        systems[s]->ent_full = 1;
    }
    dare_batch* B = new dare_batch();
    B->device = ctx->device;
    B->nsys = nsys;
    B->ncomp = N;
    B->sys.assign(systems, systems + nsys);
    B->row0.resize(nsys + 1);
    B->row0[0] = 0;
    for (int s = 0; s < nsys; s++) B->row0[s + 1] = B->row0[s] + systems[s]->dia_view.nrows;
    B->rows = B->row0[nsys];
    B->h_sys.resize(nsys);
    auto fail = [&](int rc) {
        dare_batch_destroy(B);
        return rc;
    };
    if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&B->ev_in) != hipSuccess || hipEventCreate(&B->ev_out) != hipSuccess ||
        hipEventCreate(&B->ev_t0) != hipSuccess || hipEventCreate(&B->ev_t1) != hipSuccess) {
        hipGetLastError();
        set_error("batch stream/event creation failed");
        return fail(DARE_ERR_HIP);
    }
    if (hipMalloc(&B->d_sys, sizeof(BSys) * nsys) != hipSuccess || hipMalloc(&B->ctl, sizeof(BCtl) * nsys) != hipSuccess ||
        hipMalloc(&B->d_count, sizeof(int)) != hipSuccess || hipMalloc(&B->b, sizeof(double) * B->rows) != hipSuccess ||
        hipMalloc(&B->x, sizeof(double) * B->rows) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&B->h_count), sizeof(int), 0) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&B->h_ctl), sizeof(BCtl) * nsys, 0) != hipSuccess) {
        hipGetLastError();
        set_error("batch allocation failed");
        return fail(DARE_ERR_OOM);
    }
    if (hipMemsetAsync(B->ctl, 0, sizeof(BCtl) * nsys, B->stream) != hipSuccess) {
        set_error("batch state initialisation failed");
        return fail(DARE_ERR_HIP);
    }
    dare_batch_config cfg{0, 16, 2048, 1};
    const int rc = apply_config(ctx, B, cfg);
    if (rc) return fail(rc);
    *out = B;
    return DARE_OK;
}

int dare_batch_destroy(dare_batch* B) {
    if (!B) return DARE_OK;
    hipSetDevice(B->device);
    if (B->stream) hipStreamSynchronize(B->stream);
    drop_graphs(B);
    for (void* p : {static_cast<void*>(B->d_sys), static_cast<void*>(B->d_op), static_cast<void*>(B->d_vec),
                    static_cast<void*>(B->partial), static_cast<void*>(B->ctl), static_cast<void*>(B->d_count),
                    static_cast<void*>(B->b), static_cast<void*>(B->x), static_cast<void*>(B->work64),
                    static_cast<void*>(B->work32), static_cast<void*>(B->dia32)})
        if (p) hipFree(p);
    if (B->h_count) hipHostFree(B->h_count);
    if (B->h_ctl) hipHostFree(B->h_ctl);
    for (hipEvent_t e : {B->ev_in, B->ev_out, B->ev_t0, B->ev_t1})
        if (e) hipEventDestroy(e);
    if (B->stream) hipStreamDestroy(B->stream);
    delete B;
    return DARE_OK;
}

int dare_batch_vectors(dare_batch* B, double** b, double** x, int64_t* row_offsets) {
    DARE_CHECK_ARG(B != nullptr, "NULL batch");
    if (b) *b = B->b;
    if (x) *x = B->x;
    if (row_offsets)
        for (int s = 0; s <= B->nsys; s++) row_offsets[s] = B->row0[s];
    return DARE_OK;
}

int dare_batch_set_config(dare_ctx* ctx, dare_batch* B, const dare_batch_config* cfg) {
    DARE_CHECK_ARG(ctx && B && cfg, "NULL argument");
    DARE_HIP_TRY(hipSetDevice(B->device));
    return apply_config(ctx, B, *cfg);
}

int dare_batch_get_config(const dare_batch* B, dare_batch_config* cfg) {
    DARE_CHECK_ARG(B && cfg, "NULL argument");
    *cfg = B->cfg;
    return DARE_OK;
}

int dare_batch_autotune(dare_ctx* ctx, dare_batch* B, int32_t reps, dare_batch_config* best, double* best_us) {
    DARE_CHECK_ARG(ctx && B && best, "NULL argument");
    DARE_HIP_TRY(hipSetDevice(B->device));
    const long long R = B->rows;
    int rc = grow(&B->work32, static_cast<size_t>(9 * R));
    if (rc) return rc;
    rc = grow(&B->dia32, static_cast<size_t>(8 * R));
    if (rc) return rc;
    rc = refresh_tables(B);
    if (rc) return rc;
    rc = sync_in(ctx, B);
    if (rc) return rc;
    // every system selected; operands: zero-free FP32 work vectors are fine for timing
    launch_ctl(B, C_BEGIN);
    {
        float* w = B->work32;
        BVec<float> a{w + 8 * R, w, w + R, w + 2 * R, w + 3 * R, w + 4 * R, w + 5 * R, w + 6 * R, B->x, B->b,
                      w + 8 * R};
        launch_vec<float, V_F32>(B, a);
    }
    DARE_LAUNCH_CHECK();
    const dare_batch_config keep = B->cfg;
    const int tks[] = {0, 64, 32};
    const int tis[] = {4, 8, 16, 32};
    const int grids[] = {1024, 2048, 4096, 8192};
    const int orders[] = {0, 1, 2};
    dare_batch_config bc = keep;
    double bt = 1e300;
    reps = reps > 0 ? reps : 5;
    for (int tk : tks)
        for (int ti : tis)
            for (int g : grids)
                for (int o : orders) {
                    const dare_batch_config c{tk, ti, g, o};
                    rc = apply_config(ctx, B, c);
                    if (rc) return rc;
                    // dia32 pointers are part of the tables: keep them valid
                    rc = refresh_tables(B);
                    if (rc) return rc;
                    double us = 0.;
                    rc = B->ncomp == 2 ? time_op<2>(B, reps, &us) : time_op<1>(B, reps, &us);
                    if (rc) return rc;
                    if (us < bt) {
                        bt = us;
                        bc = c;
                    }
                }
    rc = apply_config(ctx, B, bc);
    if (rc) return rc;
    rc = refresh_tables(B);
    if (rc) return rc;
    *best = bc;
    if (best_us) *best_us = bt;
    return sync_out(ctx, B);
}

int dare_batch_solve(dare_ctx* ctx, dare_batch* B, const dare_solver_params* p, int32_t mixed, int32_t graph_passes,
                     dare_solver_result* results) {
    DARE_CHECK_ARG(ctx && B && p && results, "NULL argument");
    DARE_CHECK_ARG(graph_passes >= 0 && graph_passes <= 64, "graph_passes must be in [0, 64]");
    DARE_CHECK_ARG(p->max_it >= 0, "max_it must be >= 0");
    DARE_HIP_TRY(hipSetDevice(B->device));
    int rc = refresh_tables(B);
    if (rc) return rc;
    BParams P{};
    P.tol = p->tol;
    P.tol_kind = p->tol_kind;
    P.max_it = p->max_it;
    P.inner_tol = p->inner_tol > 0. ? p->inner_tol : 1e-4;
    P.inner_max_it = p->inner_max_it > 0 ? p->inner_max_it : p->max_it;
    P.max_refine = p->max_refine > 0 ? p->max_refine : 20;
    P.min_passes = 2;        // the reference stops only for k > 1 (BiCGStab2.inl:553)
    P.mixed = mixed ? 1 : 0;
    rc = sync_in(ctx, B);
    if (rc) return rc;
    rc = B->ncomp == 2 ? solve_impl<2>(B, P, graph_passes, results) : solve_impl<1>(B, P, graph_passes, results);
    if (rc) return rc;
    return sync_out(ctx, B);
}

}  // extern "C"
