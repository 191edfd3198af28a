// Krylov kernels and the BiCGStab(2) driver for gfx950.
//
// BiCGStab2::Solve (MatrixSystem/BiCGStab2.inl:167-582) is restated as nine fused launches per
// pass. All method scalars (alpha, beta, gamma, rho, omega, mu, nu, tau) live in device memory
// and are updated by the one-work-group finish kernel of each reduction, so the host reads
// back only one small status block per pass (the convergence test). Reductions are
// deterministic: fixed grid, fixed per-thread order, block partials summed in block order.
//
// Operators: a generic CSR traversal (grid-stride, one row per lane) and, for matrices
// assembled from the 7-point structure, a tiled traversal of the slot-major diagonal view:
// 64 (k) x 4 (j) x 16 (i) cells per work-group, the i-1/i/i+1 planes of x rolled through
// registers, tiles ordered so that each XCD (blocks b = x mod 8) owns one contiguous slab.
// Matrix values then stream fully coalesced and x is read from HBM about once per SpMV.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <string>
#include <vector>

#include "krylov_core.h"

namespace dare_hip {
namespace {

// An operator the solver can apply: CSR or diagonal view, value type VT.
template <typename VT>
struct SolveOp {
    bool dia;
    CsrOp<VT> csr;
    DiaOp<VT> d;
    long long n;

    template <typename XT, class Epi>
    void launch(hipStream_t sm, const XT* x, const Epi& epi, ReduceWs ws, KState* st, int tail, int check) const {
        unsigned grid;
        if (dia) {
            grid = static_cast<unsigned>(kXcd * d.per_xcd);   // <= kReduceBlocks
            if (d.V.N == 2)
                hipLaunchKernelGGL((k_apply_dia<2, VT, XT, Epi>), dim3(grid), dim3(kThreads), 0, sm, d, x, epi, ws, st,
                                   tail, check);
            else if (d.pair == 4) {   // FP32 operators only (make_dia)
                if constexpr (std::is_same_v<VT, float> && std::is_same_v<XT, float>) {
                    if (d.sym)
                        hipLaunchKernelGGL((k_apply_dia<1, VT, XT, Epi, 4, true>), dim3(grid), dim3(kThreads), 0, sm, d,
                                           x, epi, ws, st, tail, check);
                    else
                        hipLaunchKernelGGL((k_apply_dia<1, VT, XT, Epi, 4>), dim3(grid), dim3(kThreads), 0, sm, d, x,
                                           epi, ws, st, tail, check);
                }
            } else if (d.pair == 2 && d.sym) {   // FP64 operators only (op64)
                if constexpr (std::is_same_v<VT, double>)
                    hipLaunchKernelGGL((k_apply_dia<1, VT, XT, Epi, 2, true>), dim3(grid), dim3(kThreads), 0, sm, d, x,
                                       epi, ws, st, tail, check);
            } else if (d.pair == 2)
                hipLaunchKernelGGL((k_apply_dia<1, VT, XT, Epi, 2>), dim3(grid), dim3(kThreads), 0, sm, d, x, epi, ws,
                                   st, tail, check);
            else
                hipLaunchKernelGGL((k_apply_dia<1, VT, XT, Epi>), dim3(grid), dim3(kThreads), 0, sm, d, x, epi, ws, st,
                                   tail, check);
        } else {
            grid = grid_rows(n);
            hipLaunchKernelGGL((k_apply_csr<VT, XT, Epi>), dim3(grid), dim3(kThreads), 0, sm, csr, x, epi, ws,
                               st, tail, check);
        }
        if (tail >= 0) launch_finish<Epi::NDOT>(sm, ws, grid, st, tail, check);
    }
};

}  // namespace
}  // namespace dare_hip

#include "mg.h"

namespace dare_hip {
namespace {

// ---- per-context solver workspace (device vectors + state), owned by the context
struct Workspace {
    void* buf;
    KState* st;
};

static int ws_reserve(dare_ctx* ctx, size_t bytes, Workspace* out) {
    if (!ctx->solve_state) {
        if (hipMalloc(&ctx->solve_state, sizeof(KState)) != hipSuccess) {
            hipGetLastError();
            set_error("solver state allocation failed");
            return DARE_ERR_OOM;
        }
    }
    if (ctx->solve_bytes < bytes) {
        hipStreamSynchronize(ctx->stream);
        if (ctx->solve_buf) hipFree(ctx->solve_buf);
        ctx->solve_buf = nullptr;
        ctx->solve_bytes = 0;
        if (hipMalloc(&ctx->solve_buf, bytes) != hipSuccess) {
            hipGetLastError();
            set_error("solver workspace allocation failed");
            return DARE_ERR_OOM;
        }
        ctx->solve_bytes = bytes;
    }
    out->buf = ctx->solve_buf;
    out->st = static_cast<KState*>(ctx->solve_state);
    return DARE_OK;
}

static int read_state(dare_ctx* ctx, KState* dev, KState* host) {
    DARE_HIP_TRY(hipMemcpyAsync(ctx->host_pinned, dev, sizeof(KState), hipMemcpyDeviceToHost, ctx->stream));
    DARE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::memcpy(host, ctx->host_pinned, sizeof(KState));
    return DARE_OK;
}

static_assert(sizeof(KState) <= 64 * sizeof(double), "pinned readback block too small");

// Host-driven BiCGStab2 over device-resident scalars; XT is the type of every vector, NT the
// choice of non-temporal stores (vectors larger than the caches).
// With a multigrid hierarchy mg the method runs right-preconditioned (BiCGStab2.inl:353-365):
// every operator application is A M^-1, the iterate y starts at zero and accumulates in y_buf
// (or in x itself when x starts at zero), and x += M^-1 y at the end. The residual r stays the
// residual of A x = b, so the stopping test keeps its meaning.
template <typename VT, typename XT, bool NT>
int bicgstab2_run_nt(dare_ctx* ctx, const SolveOp<VT>& A, const XT* b, XT* x, double tol_abs, int max_it,
                     double* hist, int hist_cap, int* hist_len, int* iters, double* rnorm, int* status,
                     bool* converged, XT* work, int min_passes, bool x_zero, KState* st, ReduceWs ws,
                     const MgHierarchy* mg = nullptr, XT* y_buf = nullptr) {
    const long long n = A.n;
    XT* r = work;
    XT* rh = r + n;
    XT* u = rh + n;
    XT* v = u + n;
    XT* s = v + n;
    XT* w = s + n;
    XT* t = w + n;
    const unsigned gr = grid_for(n);
    hipStream_t sm = ctx->stream;
    *hist_len = 0;
    *status = 0;
    XT* y = (mg && !x_zero) ? y_buf : x;       // the iterate the passes update
    const XT* rhv = x_zero ? b : rh;           // the shadow residual r0: b itself when x0 = 0
    auto pre = [&](const XT* in) -> const XT* { return mg ? mg_apply<VT, XT>(sm, *mg, in) : in; };
    // x = 0: no operator sweep for r = b - A x, and no vector written before the first pass
    // (which writes x and r); a solve that ends before it zeroes x itself (x_written)
    bool x_written = !x_zero;
    if (x_zero) {
        hipLaunchKernelGGL((k_init_zero<XT>), dim3(gr), dim3(kThreads), 0, sm, n, b, ws, st);
        launch_finish<2>(sm, ws, gr, st, TAIL_INIT, 0);
    } else {
        A.launch(sm, x, EpiInit<XT, NT>{b, r, rh, u, v, w}, ws, st, TAIL_INIT, 0);     // (0)-(2)
        if (mg) hipLaunchKernelGGL((k_zero<XT>), dim3(gr), dim3(kThreads), 0, sm, n, y);
    }
    DARE_LAUNCH_CHECK();
    KState h;
    int rc = read_state(ctx, st, &h);
    if (rc) return rc;
    double conv = h.rnorm;
    if (conv < tol_abs) {
        if (!x_written) hipLaunchKernelGGL((k_zero<XT>), dim3(gr), dim3(kThreads), 0, sm, n, x);
        DARE_LAUNCH_CHECK();
        *iters = 0;
        *rnorm = conv;
        *converged = true;
        return DARE_OK;
    }
    int k = 1;
    bool done_ok = false;
    while (true) {
        if (k > max_it) break;
        if (h.next_status) {
            *status = h.next_status;
            break;
        }
        // the first pass from a zero guess reads b for r and 0 for x, u (k_init_zero wrote none)
        const bool first = x_zero && k == 1;
        if (first) hipLaunchKernelGGL((k_upd_u<XT, NT, true>), dim3(gr), dim3(kThreads), 0, sm, n, b, u, st);
        else hipLaunchKernelGGL((k_upd_u<XT, NT>), dim3(gr), dim3(kThreads), 0, sm, n, r, u, st);
        A.launch(sm, pre(u), EpiDot<XT, NT>{v, rhv}, ws, st, TAIL_GAMMA, 1);
        if (first) hipLaunchKernelGGL((k_upd_rx<XT, NT, true>), dim3(gr), dim3(kThreads), 0, sm, n, r, v, y, u, st, b);
        else hipLaunchKernelGGL((k_upd_rx<XT, NT>), dim3(gr), dim3(kThreads), 0, sm, n, r, v, y, u, st, nullptr);
        x_written = true;
        A.launch(sm, pre(r), EpiDot<XT, NT>{s, rhv}, ws, st, TAIL_RHO, 1);
        hipLaunchKernelGGL((k_upd_v<XT, NT>), dim3(gr), dim3(kThreads), 0, sm, n, s, v, st);
        A.launch(sm, pre(v), EpiDot<XT, NT>{w, rhv}, ws, st, TAIL_GAMMA, 1);
        hipLaunchKernelGGL((k_upd_urs<XT, NT>), dim3(gr), dim3(kThreads), 0, sm, n, u, r, v, s, w, st);
        if (mg) A.launch(sm, pre(s), EpiGcrP<XT, NT>{t, r, s}, ws, st, TAIL_GCR, 1);
        else A.launch(sm, s, EpiGcr<XT, NT>{t, r}, ws, st, TAIL_GCR, 1);
        hipLaunchKernelGGL((k_end<XT, NT>), dim3(gr), dim3(kThreads), 0, sm, n, y, r, s, t, u, v, w, rhv, ws, st);
        launch_finish<2>(sm, ws, gr, st, TAIL_END, 1);
        DARE_LAUNCH_CHECK();
        rc = read_state(ctx, st, &h);
        if (rc) return rc;
        if (h.status) {
            *status = h.status;
            if (h.status == 21) {
                // tau == 0: true residual check (BiCGStab2.inl:514-536); preconditioned, the
                // iterate is x + M^-1 y, formed in t (unused from here on)
                const XT* xt = x;
                if (mg) {
                    const XT* z = pre(y);
                    hipLaunchKernelGGL(k_mg_combine<XT>, dim3(gr), dim3(kThreads), 0, sm, n,
                                       static_cast<const XT*>(x_zero ? nullptr : x), z, t);
                    xt = t;
                }
                A.launch(sm, xt, EpiInit<XT, NT>{b, r, rh, u, v, w}, ws, st, TAIL_INIT, 0);
                rc = read_state(ctx, st, &h);
                if (rc) return rc;
                conv = h.rnorm;
                if (conv <= tol_abs) {
                    done_ok = true;
                    *status = 0;
                }
                if (mg) DARE_HIP_TRY(hipMemcpyAsync(x, t, sizeof(XT) * n, hipMemcpyDeviceToDevice, sm));
                *iters = k;
                *rnorm = conv;
                *converged = done_ok;
                return DARE_OK;
            }
            break;
        }
        conv = h.rnorm;
        if (hist && *hist_len < hist_cap) hist[(*hist_len)++] = conv;
        if (conv <= tol_abs && k >= min_passes) break;   // reference: k > 1 (min_passes 2)
        ++k;
    }
    if (!x_written) hipLaunchKernelGGL((k_zero<XT>), dim3(gr), dim3(kThreads), 0, sm, n, x);
    if (mg) {   // x = x0 + M^-1 y (x0 = 0: x holds y)
        const XT* z = pre(y);
        hipLaunchKernelGGL(k_mg_combine<XT>, dim3(gr), dim3(kThreads), 0, sm, n,
                           static_cast<const XT*>(x_zero ? nullptr : x), z, x);
        DARE_LAUNCH_CHECK();
    }
    *iters = k;
    *rnorm = conv;
    *converged = done_ok || conv <= tol_abs;
    return DARE_OK;
}

// st / ws: the device scalars and reduction partials of this solve (the context's own, or the
// allocation-free entry point's static scratch)
template <typename VT, typename XT>
int bicgstab2_run(dare_ctx* ctx, const SolveOp<VT>& A, const XT* b, XT* x, double tol_abs, int max_it,
                  double* hist, int hist_cap, int* hist_len, int* iters, double* rnorm, int* status,
                  bool* converged, XT* work, int min_passes, bool x_zero, KState* st, ReduceWs ws,
                  const MgHierarchy* mg = nullptr, XT* y_buf = nullptr) {
    if (A.n >= kStreamRows)
        return bicgstab2_run_nt<VT, XT, true>(ctx, A, b, x, tol_abs, max_it, hist, hist_cap, hist_len, iters, rnorm,
                                              status, converged, work, min_passes, x_zero, st, ws, mg, y_buf);
    return bicgstab2_run_nt<VT, XT, false>(ctx, A, b, x, tol_abs, max_it, hist, hist_cap, hist_len, iters, rnorm,
                                           status, converged, work, min_passes, x_zero, st, ws, mg, y_buf);
}

static int dot_dev_on(dare_ctx* ctx, long long n, const double* a, const double* b, double* out, KState* st,
                      ReduceWs ws) {
    hipLaunchKernelGGL((k_dot<double>), dim3(grid_for(n)), dim3(kThreads), 0, ctx->stream, n, a, b, ws, st);
    launch_finish<1>(ctx->stream, ws, grid_for(n), st, TAIL_STORE, 0);
    DARE_LAUNCH_CHECK();
    KState h;
    int rc = read_state(ctx, st, &h);
    if (rc) return rc;
    *out = h.dots[0];
    return DARE_OK;
}

static int dot_dev(dare_ctx* ctx, long long n, const double* a, const double* b, double* out) {
    Workspace W;
    int rc = ws_reserve(ctx, 0, &W);
    if (rc) return rc;
    return dot_dev_on(ctx, n, a, b, out, W.st, ctx->ws);
}

// ---- Jacobi right preconditioning (M = diag(A)). The method runs on A M^-1 with y = M x:
// the diagonal view is column-scaled once per solve, x0 -> M x0 before and y -> M^-1 y after,
// so every pass costs what an unpreconditioned pass costs and the residual b - A M^-1 y is the
// true residual b - A x (the stopping test keeps its meaning).
__global__ __launch_bounds__(kThreads) void k_dinv(long long n, const double* __restrict__ diag,
                                                   double* __restrict__ dinv) {
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads)
        dinv[i] = 1.0 / diag[i];
}

// out[s][r] = D[s][r] * dinv[column of slot s in row r]; a removed entry holds 0 and stays 0
template <typename OT>
__global__ __launch_bounds__(kThreads) void k_dia_colscale(DiaView V, const double* __restrict__ D,
                                                           const double* __restrict__ dinv, OT* __restrict__ out) {
    const long long nr = V.nrows;
    const int N = V.N;
    const long long sI = static_cast<long long>(V.n[1]) * V.n[2], sJ = V.n[2];
    for (long long r = blockIdx.x * (long long)kThreads + threadIdx.x; r < nr; r += (long long)gridDim.x * kThreads) {
        const long long node = r / N;
        const int c = static_cast<int>(r % N);
        const int k = static_cast<int>(node % V.n[2]);
        const int j = static_cast<int>((node / V.n[2]) % V.n[1]);
        const int i = static_cast<int>(node / sI);
        const int pos[3] = {i, j, k};
        const long long stride[3] = {sI, sJ, 1};
        long long nb[7];
        // slots 0 W, 1 S, 2 B, 3 C, 4 T, 5 N, 6 E: low neighbour of axis d is slot d, high is 6 - d
        nb[3] = node;
        for (int d = 0; d < 3; d++) {
            const int n = V.n[d];
            long long lo = node - stride[d], hi = node + stride[d];
            if (pos[d] == 0) lo = V.per[d] ? node + static_cast<long long>(n - 1) * stride[d] : node;
            if (pos[d] == n - 1) hi = V.per[d] ? node - static_cast<long long>(n - 1) * stride[d] : node;
            nb[d] = lo;
            nb[6 - d] = hi;
        }
        double dv[8], sc[8];
#pragma unroll
        for (int sl = 0; sl < 7; sl++) {
            dv[sl] = D[sl * nr + r];
            sc[sl] = dinv[nb[sl] * N + c];
        }
        dv[7] = (N == 2 && V.cross) ? D[7 * nr + r] : 0.;
        sc[7] = (N == 2 && V.cross) ? dinv[node * N + (1 - c)] : 0.;
#pragma unroll
        for (int sl = 0; sl < 7; sl++) out[sl * nr + r] = static_cast<OT>(dv[sl] * sc[sl]);
        if (N == 2 && V.cross) out[7 * nr + r] = static_cast<OT>(dv[7] * sc[7]);
    }
}

// 1 / diag(A) from the CSR rows (columns ascending, duplicates merged); a row without a
// diagonal entry gets 1 / 0 = inf, as 1 / diag does on the diagonal view
__global__ __launch_bounds__(kThreads) void k_csr_dinv(long long n, const int32_t* __restrict__ rp,
                                                       const int32_t* __restrict__ col,
                                                       const double* __restrict__ val, double* __restrict__ dinv) {
    for (long long r = blockIdx.x * (long long)kThreads + threadIdx.x; r < n; r += (long long)gridDim.x * kThreads) {
        double d = 0.;
        for (int j = rp[r]; j < rp[r + 1]; j++)
            if (col[j] == r) d = val[j];
        dinv[r] = 1.0 / d;
    }
}

// out[j] = val[j] * dinv[col[j]]: the CSR values of A diag(A)^-1
template <typename OT>
__global__ __launch_bounds__(kThreads) void k_csr_colscale(long long nnz, const int32_t* __restrict__ col,
                                                           const double* __restrict__ val,
                                                           const double* __restrict__ dinv, OT* __restrict__ out) {
    for (long long j = blockIdx.x * (long long)kThreads + threadIdx.x; j < nnz; j += (long long)gridDim.x * kThreads)
        out[j] = static_cast<OT>(val[j] * dinv[col[j]]);
}

// left Jacobi: diag(A)^-1 A, each row scaled by its own 1 / diag (slot-major view, slots slots)
__global__ __launch_bounds__(kThreads) void k_dia_rowscale(long long nr, int slots, const double* __restrict__ D,
                                                           const double* __restrict__ dinv, double* __restrict__ out) {
    for (long long r = blockIdx.x * (long long)kThreads + threadIdx.x; r < nr; r += (long long)gridDim.x * kThreads) {
        const double s = dinv[r];
        for (int sl = 0; sl < slots; sl++) out[sl * nr + r] = D[sl * nr + r] * s;
    }
}
__global__ __launch_bounds__(kThreads) void k_csr_rowscale(long long n, const int32_t* __restrict__ rp,
                                                           const double* __restrict__ val,
                                                           const double* __restrict__ dinv, double* __restrict__ out) {
    for (long long r = blockIdx.x * (long long)kThreads + threadIdx.x; r < n; r += (long long)gridDim.x * kThreads) {
        const double s = dinv[r];
        for (int j = rp[r]; j < rp[r + 1]; j++) out[j] = val[j] * s;
    }
}

__global__ __launch_bounds__(kThreads) void k_scale_by(long long n, double* __restrict__ x, const double* __restrict__ s,
                                                       int invert) {
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads)
        x[i] = invert ? x[i] * s[i] : x[i] / s[i];
}

__global__ __launch_bounds__(kThreads) void k_add_correction_pre(long long n, double* __restrict__ x,
                                                                 const float* __restrict__ d,
                                                                 const double* __restrict__ dinv) {
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads)
        x[i] = x[i] + dinv[i] * static_cast<double>(d[i]);
}

template <typename T>
int grow_buffer(T** p, int64_t* cap, int64_t need, const char* what) {
    if (*cap >= need && *p) return DARE_OK;
    if (*p) hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (hipMalloc(p, sizeof(T) * (need > 0 ? need : 1)) != hipSuccess) {
        (void)hipGetLastError();
        set_error(std::string(what) + " allocation failed");
        return DARE_ERR_OOM;
    }
    *cap = need;
    return DARE_OK;
}

// 1 / diag(A) into A->dinv: slot C of the diagonal view, or the CSR diagonal of a matrix
// without one (its values must be present)
int jacobi_setup(dare_ctx* ctx, dare_csr* A) {
    const long long n = A->nrows;
    int rc = grow_buffer(&A->dinv, &A->cap_dinv, n, "Jacobi diagonal");
    if (rc) return rc;
    if (A->has_dia)
        hipLaunchKernelGGL(k_dinv, dim3(grid_for(n)), dim3(kThreads), 0, ctx->stream, n, A->dia + 3 * n, A->dinv);
    else
        hipLaunchKernelGGL(k_csr_dinv, dim3(grid_for(n)), dim3(kThreads), 0, ctx->stream, n, A->row_ptr, A->col,
                           A->val, A->dinv);
    DARE_LAUNCH_CHECK();
    return DARE_OK;
}

// A diag(A)^-1 as CSR values (FP64 into A->dia_pre, or FP32 into A->val32); dinv must be set
template <typename OT>
int jacobi_csr_values(dare_ctx* ctx, dare_csr* A, OT* out) {
    hipLaunchKernelGGL((k_csr_colscale<OT>), dim3(grid_for(A->nnz)), dim3(kThreads), 0, ctx->stream,
                       (long long)A->nnz, A->col, A->val, A->dinv, out);
    DARE_LAUNCH_CHECK();
    return DARE_OK;
}

// the multigrid hierarchy of A for a solve whose operator is op (slots D0 of type VT)
template <typename VT, typename XT>
int mg_prepare(dare_ctx* ctx, dare_csr* A, const SolveOp<VT>& op, const VT* D0, MgHierarchy** out) {
    if (!op.dia || !A->has_dia) {
        set_error("the multigrid preconditioner needs a device-assembled 3-D operator (diagonal view)");
        return DARE_ERR_UNSUPPORTED;
    }
    const int erc = dia_ensure_full(A, ctx->stream);     // the levels read all seven slots
    if (erc) return erc;
    if (!A->mg) A->mg = new MgHierarchy();
    MgHierarchy* H = static_cast<MgHierarchy*>(A->mg);
    const int rc = mg_setup<VT, XT>(ctx, H, A->dia_view, D0);
    *out = H;
    return rc;
}

// A->sym64 for the last fill: 1 by construction (sym_fill), from the device's slot-by-slot
// check for an operator that may be symmetric (sym_cand; one readback per fill), else 0
static int sym64_prepare(dare_ctx* ctx, dare_csr* A) {
    if (!A->has_dia || A->sym64 >= 0) return DARE_OK;
    const bool one = A->dia_view.N == 1 && !A->dia_view.cross;
    if (one && A->sym_fill) {
        A->sym64 = 1;
    } else if (one && A->sym_cand && A->asym64) {
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, ctx->stream, A->asym64, 0);
        hipLaunchKernelGGL((k_dia_asym<double>), dim3(grid_for(A->nrows)), dim3(kThreads), 0, ctx->stream,
                           A->dia_view, static_cast<const double*>(A->dia), A->asym64);
        DARE_LAUNCH_CHECK();
        DARE_HIP_TRY(hipMemcpyAsync(ctx->host_pinned, A->asym64, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        DARE_HIP_TRY(hipStreamSynchronize(ctx->stream));
        int flag = 1;
        std::memcpy(&flag, ctx->host_pinned, sizeof(int));
        A->sym64 = flag == 0 ? 1 : 0;
    } else {
        A->sym64 = 0;
    }
    return DARE_OK;
}

// The FP64 operator of A: the diagonal view (its symmetric pair traversal when the view is
// exactly symmetric) or the CSR values.
// vecs: the caller's vectors the traversal reads or writes. The paired, quad and two-component
// traversals move 16-byte vectors, so those must start on 16-byte boundaries (any allocation
// does; a slice at an odd row offset does not).
static int op64(dare_ctx* ctx, const dare_csr* A, bool allow_dia, SolveOp<double>* out,
                std::initializer_list<const void*> vecs = {}) {
    SolveOp<double>& op = *out;
    op = SolveOp<double>{};
    op.n = A->nrows;
    op.csr = CsrOp<double>{A->row_ptr, A->col, A->val, A->nrows};
    op.dia = allow_dia && A->has_dia;
    if (op.dia) {
        op.d = make_dia<double>(A->dia_view, A->dia, true);
        op.dia = kXcd * op.d.per_xcd <= kReduceBlocks;   // else the CSR traversal
    }
    if (op.dia && op.d.pair == 2 && A->sym_traversal) {
        const int rc = sym64_prepare(ctx, const_cast<dare_csr*>(A));
        if (rc) return rc;
        op.d.sym = A->sym64 == 1;
    }
    if (op.dia && (op.d.pair > 1 || op.d.V.N == 2))
        for (const void* v : vecs)
            DARE_CHECK_ARG((reinterpret_cast<uintptr_t>(v) & 15) == 0,
                           "vectors must start on a 16-byte boundary (the diagonal-view traversals move 16-byte "
                           "vectors)");
    // any traversal but the symmetric one reads E, N, T (or the CSR values made from them)
    if (!(op.dia && op.d.sym)) return dia_ensure_full(const_cast<dare_csr*>(A), ctx->stream);
    return DARE_OK;
}

}  // namespace
}  // namespace dare_hip

// ---- static scratch of the allocation-free solve (dare_bicgstab2_solve_noalloc) -------------
// Seven FP64 work vectors of at most kNoAllocMaxRows rows, the reduction partials and the
// device scalar block. Each device that loads this library holds its own copy until the
// process ends. Host code never passes `scratch` itself to a kernel: its address on a device
// comes from hipGetSymbolAddress (noalloc_scratch_init, run by dare_ctx_create with that
// device current) and is kept per device.
namespace dare_hip {
namespace {
constexpr long long kNoAllocMaxRows = 1LL << 18;                                   // 262,144 rows (64^3)
constexpr long long kScratchVec = 7 * kNoAllocMaxRows;                            // doubles
constexpr long long kScratchPartial = static_cast<long long>(kReduceBlocks) * kMaxDots;
constexpr long long kScratchState = 64;                                           // KState slot
constexpr long long kScratchCount = 32;                                           // ReduceWs::count slot
constexpr long long kScratchDoubles = kScratchVec + kScratchPartial + kScratchState + kScratchCount;
static_assert(sizeof(KState) <= kScratchState * sizeof(double), "scratch state slot too small");
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};
}  // namespace
}  // namespace dare_hip

static __device__ __align__(256) double scratch[dare_hip::kScratchDoubles];

namespace dare_hip {

int noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("device index beyond the allocation-free scratch table");
        return DARE_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return DARE_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return DARE_ERR_HIP;
    }
    g_scratch[device] = static_cast<double*>(p);
    return DARE_OK;
}

static double* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch[device];
}

}  // namespace dare_hip

namespace dare_hip {
void mg_free(void* p) {
    if (!p) return;
    MgHierarchy* H = static_cast<MgHierarchy*>(p);
    if (H->coarse) hipFree(H->coarse);
    if (H->fine) hipFree(H->fine);
    delete H;
}
}  // namespace dare_hip

using namespace dare_hip;

extern "C" {

int dare_spmv(dare_ctx* ctx, const dare_csr* A, const double* x, double* y) {
    DARE_CHECK_ARG(ctx && A && x && y, "NULL argument");
    if (A->nrows == 0) return DARE_OK;
    // the Tpetra apply replacement: per-row sum in sorted column order. Without a periodic
    // axis no row wraps, and the diagonal view's row order (W,S,B,[X],C,[X],T,N,E) is the
    // sorted column order, so a device-assembled matrix is applied from its diagonal view
    // (80 B/row, no CSR values to materialize) with bit-identical results; otherwise the CSR.
    const bool wrap = A->has_dia && (A->dia_view.per[0] || A->dia_view.per[1] || A->dia_view.per[2]);
    SolveOp<double> op;
    int rc = op64(ctx, A, !wrap, &op, {x, y});
    if (rc) return rc;
    if (!op.dia) {
        rc = csr_ensure_values(const_cast<dare_csr*>(A), ctx->stream);
        if (rc) return rc;
    }
    if (A->nrows >= kStreamRows) op.launch(ctx->stream, x, EpiStore<double, true>{y}, ctx->ws, nullptr, -1, 0);
    else op.launch(ctx->stream, x, EpiStore<double, false>{y}, ctx->ws, nullptr, -1, 0);
    DARE_LAUNCH_CHECK();
    return DARE_OK;
}

int dare_residual(dare_ctx* ctx, const dare_csr* A, const double* b, const double* x, double* r, double* norm_host) {
    DARE_CHECK_ARG(ctx && A && b && x && r, "NULL argument");
    if (A->nrows == 0) {
        if (norm_host) *norm_host = 0.;
        return DARE_OK;
    }
    Workspace W;
    int rc = ws_reserve(ctx, 0, &W);
    if (rc) return rc;
    SolveOp<double> op;
    rc = op64(ctx, A, true, &op, {b, x, r});
    if (rc) return rc;
    if (!op.dia) {
        rc = csr_ensure_values(const_cast<dare_csr*>(A), ctx->stream);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_clear_state, dim3(1), dim3(64), 0, ctx->stream, W.st);
    if (A->nrows >= kStreamRows) op.launch(ctx->stream, x, EpiResid64<true>{b, r}, ctx->ws, W.st, TAIL_STORE, 0);
    else op.launch(ctx->stream, x, EpiResid64<false>{b, r}, ctx->ws, W.st, TAIL_STORE, 0);
    DARE_LAUNCH_CHECK();
    if (norm_host) {
        KState h;
        rc = read_state(ctx, W.st, &h);
        if (rc) return rc;
        *norm_host = std::sqrt(h.dots[0]);
    }
    return DARE_OK;
}

int dare_spmv_dot(dare_ctx* ctx, const dare_csr* A, const double* x, double* y, const double* w,
                  double* out_host) {
    DARE_CHECK_ARG(ctx && A && x && y && w, "NULL argument");
    if (A->nrows == 0) {
        if (out_host) *out_host = 0.;
        return DARE_OK;
    }
    Workspace W;
    int rc = ws_reserve(ctx, 0, &W);
    if (rc) return rc;
    hipLaunchKernelGGL(k_clear_state, dim3(1), dim3(64), 0, ctx->stream, W.st);
    // the operator the solver uses (diagonal view when the matrix has one)
    SolveOp<double> op;
    rc = op64(ctx, A, true, &op, {x, y, w});
    if (rc) return rc;
    if (A->nrows >= kStreamRows) op.launch(ctx->stream, x, EpiDot<double, true>{y, w}, ctx->ws, W.st, TAIL_STORE, 1);
    else op.launch(ctx->stream, x, EpiDot<double, false>{y, w}, ctx->ws, W.st, TAIL_STORE, 1);
    DARE_LAUNCH_CHECK();
    if (out_host) {
        KState h;
        rc = read_state(ctx, W.st, &h);
        if (rc) return rc;
        *out_host = h.dots[0];
    }
    return DARE_OK;
}

int dare_csr_symmetric(dare_ctx* ctx, const dare_csr* A, int32_t* out) {
    DARE_CHECK_ARG(ctx && A && out, "NULL argument");
    *out = 0;
    if (A->nrows == 0) return DARE_OK;
    SolveOp<double> op;
    const int rc = op64(ctx, A, true, &op);
    if (rc) return rc;
    *out = op.dia && op.d.pair == 2 && op.d.sym ? 1 : 0;
    return DARE_OK;
}

int dare_csr_set_symmetric_traversal(dare_csr* A, int32_t enable) {
    DARE_CHECK_ARG(A, "NULL argument");
    A->sym_traversal = enable != 0;
    return DARE_OK;
}

int dare_dot(dare_ctx* ctx, int64_t n, const double* a, const double* b, double* out_host) {
    DARE_CHECK_ARG(ctx && a && b && out_host && n >= 0, "invalid dot arguments");
    return dot_dev(ctx, n, a, b, out_host);
}

int dare_axpby(dare_ctx* ctx, int64_t n, double a, const double* x, double b, double* y) {
    DARE_CHECK_ARG(ctx && x && y && n >= 0, "invalid axpby arguments");
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(kThreads), 0, ctx->stream, (long long)n, a, x, b, y);
    DARE_LAUNCH_CHECK();
    return DARE_OK;
}

int dare_bicgstab2_solve(dare_ctx* ctx, const dare_csr* A, const double* b, double* x,
                         const dare_solver_params* p, dare_solver_result* res) {
    DARE_CHECK_ARG(ctx && A && b && x && p && res, "NULL argument");
    DARE_CHECK_ARG(A->nrows > 0 && A->row_ptr && A->col && A->val, "empty matrix");
    const long long n = A->nrows;
    DARE_CHECK_ARG(p->precond >= 0 && p->precond <= 3,
                   "precond must be 0 (none), 1 (Jacobi), 2 (multigrid) or 3 (Jacobi on the left)");
    const bool use_mg = p->precond == DARE_PRECOND_MG;
    const bool left = p->precond == DARE_PRECOND_JACOBI_LEFT;
    Workspace W;
    // + the preconditioned iterate (multigrid) or the scaled right-hand side (left Jacobi)
    int rc = ws_reserve(ctx, sizeof(double) * (use_mg || left ? 8 : 7) * n, &W);
    if (rc) return rc;
    double bnorm2 = 0.;
    rc = dot_dev(ctx, n, b, b, &bnorm2);
    if (rc) return rc;
    const double bnorm = std::sqrt(bnorm2);
    const double tol_abs = p->tol_kind == 1 ? p->tol * bnorm : p->tol;
    SolveOp<double> op;
    rc = op64(ctx, A, true, &op, {b, x});
    if (rc) return rc;
    const bool pre = p->precond == 1;
    if (pre || left) {          // the scaled views are made from all seven slots
        rc = dia_ensure_full(const_cast<dare_csr*>(A), ctx->stream);
        if (rc) return rc;
    }
    MgHierarchy* mg = nullptr;
    if (use_mg) {
        rc = mg_prepare<double, double>(ctx, const_cast<dare_csr*>(A), op, A->dia, &mg);
        if (rc) return rc;
    }
    if (!op.dia) {
        rc = csr_ensure_values(const_cast<dare_csr*>(A), ctx->stream);
        if (rc) return rc;
    }
    if (pre) {
        dare_csr* Am = const_cast<dare_csr*>(A);
        rc = jacobi_setup(ctx, Am);
        if (rc) return rc;
        if (op.dia) {
            const int slots = A->dia_view.cross ? 8 : 7;
            rc = grow_buffer(&Am->dia_pre, &Am->cap_dia_pre, static_cast<int64_t>(slots) * n, "preconditioned view");
            if (rc) return rc;
            hipLaunchKernelGGL((k_dia_colscale<double>), dim3(grid_for(n)), dim3(kThreads), 0, ctx->stream,
                               A->dia_view, A->dia, A->dinv, Am->dia_pre);
            op.d = make_dia<double>(A->dia_view, Am->dia_pre, true);
        } else {                  // generic CSR: scaled copy of the values
            rc = grow_buffer(&Am->dia_pre, &Am->cap_dia_pre, A->nnz, "preconditioned values");
            if (rc) return rc;
            rc = jacobi_csr_values<double>(ctx, Am, Am->dia_pre);
            if (rc) return rc;
            op.csr.val = Am->dia_pre;
        }
        hipLaunchKernelGGL(k_scale_by, dim3(grid_for(n)), dim3(kThreads), 0, ctx->stream, n, x, A->dinv, 0);  // y0 = M x0
        DARE_LAUNCH_CHECK();
    }
    // left Jacobi (BiCGStab2.inl:276-280 and the M^-1 A applications of steps 6, 9, 13): the
    // method runs on diag(A)^-1 A x = diag(A)^-1 b, so its residual, and the stopping test on
    // it, are the preconditioned ones, as in the reference; the RBNORM normaliser stays ||b||
    const double* bs = b;
    if (left) {
        dare_csr* Am = const_cast<dare_csr*>(A);
        rc = jacobi_setup(ctx, Am);
        if (rc) return rc;
        if (op.dia) {
            const int slots = A->dia_view.cross ? 8 : 7;
            rc = grow_buffer(&Am->dia_pre, &Am->cap_dia_pre, static_cast<int64_t>(slots) * n, "preconditioned view");
            if (rc) return rc;
            hipLaunchKernelGGL(k_dia_rowscale, dim3(grid_for(n)), dim3(kThreads), 0, ctx->stream, n, slots, A->dia,
                               A->dinv, Am->dia_pre);
            op.d = make_dia<double>(A->dia_view, Am->dia_pre, true);
        } else {
            rc = grow_buffer(&Am->dia_pre, &Am->cap_dia_pre, A->nnz, "preconditioned values");
            if (rc) return rc;
            hipLaunchKernelGGL(k_csr_rowscale, dim3(grid_for(n)), dim3(kThreads), 0, ctx->stream, n, A->row_ptr,
                               A->val, A->dinv, Am->dia_pre);
            op.csr.val = Am->dia_pre;
        }
        double* bl = static_cast<double*>(W.buf) + 7 * n;
        DARE_HIP_TRY(hipMemcpyAsync(bl, b, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
        hipLaunchKernelGGL(k_scale_by, dim3(grid_for(n)), dim3(kThreads), 0, ctx->stream, n, bl, A->dinv, 1);
        DARE_LAUNCH_CHECK();
        bs = bl;
    }
    int iters = 0, status = 0, hl = 0;
    double rn = 0.;
    bool conv = false;
    rc = bicgstab2_run<double, double>(ctx, op, bs, x, tol_abs, p->max_it, p->history, p->history_cap, &hl, &iters,
                                       &rn, &status, &conv, static_cast<double*>(W.buf), 2, false, W.st, ctx->ws, mg,
                                       static_cast<double*>(W.buf) + 7 * n);
    if (rc) return rc;
    if (pre) {
        hipLaunchKernelGGL(k_scale_by, dim3(grid_for(n)), dim3(kThreads), 0, ctx->stream, n, x, A->dinv, 1);  // x = M^-1 y
        DARE_LAUNCH_CHECK();
    }
    res->converged = conv ? 1 : 0;
    res->iterations = iters;
    res->status = status ? DARE_ERR_BREAKDOWN : 0;
    res->refinements = 0;
    res->residual = rn;
    res->bnorm = bnorm;
    res->history_len = hl;
    return DARE_OK;
}

int dare_bicgstab2_solve_noalloc(dare_ctx* ctx, const dare_csr* A, const double* b, double* x,
                                 const dare_solver_params* p, dare_solver_result* res) {
    DARE_CHECK_ARG(ctx && A && b && x && p && res, "NULL argument");
    DARE_CHECK_ARG(A->nrows > 0 && A->row_ptr && A->col && A->val, "empty matrix");
    DARE_CHECK_ARG(A->nrows <= kNoAllocMaxRows,
                   "system exceeds the allocation-free maximum of 262144 rows (use dare_bicgstab2_solve)");
    if (p->precond != DARE_PRECOND_NONE) {
        set_error("the allocation-free solve runs unpreconditioned (precond must be 0)");
        return DARE_ERR_UNSUPPORTED;
    }
    double* base = noalloc_scratch(ctx->device);
    if (!base) {
        set_error("allocation-free scratch not set up on this device (create the context first)");
        return DARE_ERR_HIP;
    }
    const long long n = A->nrows;
    double* work = base;
    KState* st = reinterpret_cast<KState*>(base + kScratchVec + kScratchPartial);
    ReduceWs ws{base + kScratchVec,
                reinterpret_cast<unsigned int*>(base + kScratchVec + kScratchPartial + kScratchState)};
    // the scratch keeps the previous call's contents: the scalar block starts from zero, every
    // work vector and partial is written before it is read
    hipLaunchKernelGGL(k_clear_state, dim3(1), dim3(64), 0, ctx->stream, st);
    DARE_LAUNCH_CHECK();
    double bnorm2 = 0.;
    int rc = dot_dev_on(ctx, n, b, b, &bnorm2, st, ws);
    if (rc) return rc;
    const double bnorm = std::sqrt(bnorm2);
    const double tol_abs = p->tol_kind == 1 ? p->tol * bnorm : p->tol;
    SolveOp<double> op;
    rc = op64(ctx, A, true, &op, {b, x});
    if (rc) return rc;
    if (!op.dia) {
        rc = csr_ensure_values(const_cast<dare_csr*>(A), ctx->stream);
        if (rc) return rc;
    }
    int iters = 0, status = 0, hl = 0;
    double rn = 0.;
    bool conv = false;
    rc = bicgstab2_run<double, double>(ctx, op, b, x, tol_abs, p->max_it, p->history, p->history_cap, &hl, &iters,
                                       &rn, &status, &conv, work, 2, false, st, ws);
    if (rc) return rc;
    res->converged = conv ? 1 : 0;
    res->iterations = iters;
    res->status = status ? DARE_ERR_BREAKDOWN : 0;
    res->refinements = 0;
    res->residual = rn;
    res->bnorm = bnorm;
    res->history_len = hl;
    return DARE_OK;
}

int dare_refine_solve_mixed(dare_ctx* ctx, const dare_csr* Ac, const double* b, double* x,
                            const dare_solver_params* p, dare_solver_result* res) {
    DARE_CHECK_ARG(ctx && Ac && b && x && p && res, "NULL argument");
    dare_csr* A = const_cast<dare_csr*>(Ac);
    const long long n = A->nrows;
    DARE_CHECK_ARG(n > 0, "empty matrix");
    // FP32 copy of the operator (values change every step: converted on every call)
    SolveOp<float> op32;
    op32.n = n;
    DARE_CHECK_ARG(p->precond >= 0 && p->precond <= 2, "precond must be 0 (none), 1 (Jacobi) or 2 (multigrid)");
    const bool pre = p->precond == 1;
    const bool use_mg = p->precond == DARE_PRECOND_MG;
    if (A->has_dia && kXcd * make_dia<double>(A->dia_view, A->dia).per_xcd <= kReduceBlocks) {
        if (pre || use_mg) {    // the scaled FP32 view and the multigrid levels read all seven slots
            const int erc = dia_ensure_full(A, ctx->stream);
            if (erc) return erc;
        }
        const int slots = A->dia_view.cross ? 8 : 7;
        const int64_t need = static_cast<int64_t>(slots) * n;
        if (A->cap_dia32 < need) {
            if (A->dia32) hipFree(A->dia32);
            A->dia32 = nullptr;
            A->cap_dia32 = 0;
            A->dia32_fresh = 0;
            if (hipMalloc(&A->dia32, sizeof(float) * need) != hipSuccess) {
                hipGetLastError();
                set_error("FP32 diagonal allocation failed");
                return DARE_ERR_OOM;
            }
            A->cap_dia32 = need;
        }
        if (pre) {       // FP32 copy of A M^-1 (the inner solves run right-preconditioned)
            int prc = jacobi_setup(ctx, A);
            if (prc) return prc;
            hipLaunchKernelGGL((k_dia_colscale<float>), dim3(grid_for(n)), dim3(kThreads), 0, ctx->stream,
                               A->dia_view, A->dia, A->dinv, A->dia32);
            A->dia32_fresh = 0;
        } else {
            // written by the last fill when it already ran in FP32 mode, converted here otherwise;
            // from now on the fills of this matrix write it themselves
            if (!A->dia32_fresh)
                hipLaunchKernelGGL(k_to_f32, dim3(grid_for(need)), dim3(kThreads), 0, ctx->stream, (long long)need,
                                   A->dia, A->dia32);
            A->dia32_fresh = 1;
            A->dia32_mode = 1;
        }
        op32.d = make_dia<float>(A->dia_view, A->dia32, true);
        op32.dia = kXcd * op32.d.per_xcd <= kReduceBlocks;
        if (op32.dia && op32.d.pair == 4 && !pre && A->sym_traversal) {
            // the FP32 copy of an exactly symmetric FP64 view is exactly symmetric (Jacobi
            // column scaling off): the quad traversal then reads E, N, T from the neighbours'
            // W, S, B slots. The FP64 view's symmetry is known on the host (sym64_prepare).
            const int src = sym64_prepare(ctx, A);
            if (src) return src;
            op32.d.sym = A->sym64 == 1;
        }
        if (!op32.d.sym) {      // the seven-slot FP32 traversal
            const int erc = dia_ensure_full(A, ctx->stream);
            if (erc) return erc;
        }
    } else {
        int erc = csr_ensure_values(A, ctx->stream);
        if (erc) return erc;
        if (A->cap_val32 < A->nnz) {
            if (A->val32) hipFree(A->val32);
            A->val32 = nullptr;
            if (hipMalloc(&A->val32, sizeof(float) * (A->nnz ? A->nnz : 1)) != hipSuccess) {
                hipGetLastError();
                set_error("FP32 value allocation failed");
                return DARE_ERR_OOM;
            }
            A->cap_val32 = A->nnz;
        }
        if (pre) {       // FP32 CSR values of A M^-1
            erc = jacobi_setup(ctx, A);
            if (erc) return erc;
            erc = jacobi_csr_values<float>(ctx, A, A->val32);
            if (erc) return erc;
        } else {
            hipLaunchKernelGGL(k_to_f32, dim3(grid_for(A->nnz)), dim3(kThreads), 0, ctx->stream, (long long)A->nnz,
                               A->val, A->val32);
        }
        op32.dia = false;
    }
    op32.csr = CsrOp<float>{A->row_ptr, A->col, A->val32, n};
    DARE_LAUNCH_CHECK();
    MgHierarchy* mg = nullptr;
    if (use_mg) {
        int mrc = mg_prepare<float, float>(ctx, A, op32, static_cast<const float*>(A->dia32), &mg);
        if (mrc) return mrc;
    }
    Workspace W;
    // 7 FP32 work vectors + FP32 residual + FP32 correction
    int rc = ws_reserve(ctx, sizeof(float) * 9 * n, &W);
    if (rc) return rc;
    float* work = static_cast<float*>(W.buf);
    float* r32 = work + 7 * n;
    float* d32 = r32 + n;
    // ||b|| comes out of the first residual sweep (EpiResidB), which reads b anyway
    double bnorm = 0., tol_abs = 0.;
    SolveOp<double> op;
    rc = op64(ctx, A, true, &op, {b, x});
    if (rc) return rc;
    if (!op.dia) {
        rc = csr_ensure_values(A, ctx->stream);
        if (rc) return rc;
    }
    const unsigned gr = grid_for(n);
    int total_it = 0, refine = 0, hl = 0;
    double rn = 0.;
    bool conv = false;
    int status = 0;
    const int max_refine = p->max_refine > 0 ? p->max_refine : 20;
    const int inner_max = p->inner_max_it > 0 ? p->inner_max_it : p->max_it;
    const double inner_tol = p->inner_tol > 0. ? p->inner_tol : 1e-4;
    while (true) {
        const bool first = refine == 0;
        if (first && n >= kStreamRows)
            op.launch(ctx->stream, x, EpiResidB<true>{b, r32}, ctx->ws, W.st, TAIL_STORE, 0);
        else if (first)
            op.launch(ctx->stream, x, EpiResidB<false>{b, r32}, ctx->ws, W.st, TAIL_STORE, 0);
        else if (n >= kStreamRows)
            op.launch(ctx->stream, x, EpiResid<true>{b, r32}, ctx->ws, W.st, TAIL_STORE, 0);
        else
            op.launch(ctx->stream, x, EpiResid<false>{b, r32}, ctx->ws, W.st, TAIL_STORE, 0);   // r = b - A x
        DARE_LAUNCH_CHECK();
        KState h;
        rc = read_state(ctx, W.st, &h);
        if (rc) return rc;
        if (first) {
            bnorm = std::sqrt(h.dots[1]);
            tol_abs = p->tol_kind == 1 ? p->tol * bnorm : p->tol;
        }
        rn = std::sqrt(h.dots[0]);
        if (p->history && hl < p->history_cap) p->history[hl++] = rn;
        if (rn <= tol_abs) {
            conv = true;
            break;
        }
        if (refine >= max_refine || total_it >= p->max_it) break;
        int it = 0, ist = 0, dummy_len = 0;
        double irn = 0.;
        bool iconv = false;
        // an inner solve may stop after one pass once it meets its threshold when a pass costs
        // four V-cycles (multigrid); otherwise the second pass is cheaper than the extra
        // refinement (FP64 residual, correction, new inner start) that stopping early brings
        rc = bicgstab2_run<float, float>(ctx, op32, r32, d32, inner_threshold(inner_tol, rn, tol_abs), inner_max,
                                         nullptr, 0, &dummy_len, &it, &irn, &ist, &iconv, work, use_mg ? 1 : 2,
                                         true, W.st, ctx->ws, mg);   // d32 = 0 to start
        if (rc) return rc;
        total_it += it;
        if (pre)    // the inner solution is y = M d
            hipLaunchKernelGGL(k_add_correction_pre, dim3(gr), dim3(kThreads), 0, ctx->stream, n, x, d32, A->dinv);
        else
            hipLaunchKernelGGL(k_add_correction, dim3(gr), dim3(kThreads), 0, ctx->stream, n, x, d32);
        DARE_LAUNCH_CHECK();
        refine++;
        if (ist && it <= 1) {
            status = ist;
            break;
        }
    }
    res->converged = conv ? 1 : 0;
    res->iterations = total_it;
    res->status = status ? DARE_ERR_BREAKDOWN : 0;
    res->refinements = refine;
    res->residual = rn;
    res->bnorm = bnorm;
    res->history_len = hl;
    return DARE_OK;
}

}  // extern "C"
