// Krylov kernels and the BiCGStab(2) driver for gfx950.
//
// BiCGStab2::Solve (MatrixSystem/BiCGStab2.inl:167-582) is restated as nine fused launches per
// pass. All method scalars (alpha, beta, gamma, rho, omega, mu, nu, tau) live in device memory
// and are updated by the last work-group of each one-pass grid reduction, so the host reads
// back only one small status block per pass (the convergence test). Reductions are
// deterministic: fixed grid, fixed per-thread order, block partials summed in block order.
//
// Vector traffic per pass (FP64, 7-point CSR with int32 columns): four SpMV+dot launches
// (~112 B/row each incl. the fused dot operand) and five fused vector updates (~200 B/row).
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"

namespace dare_hip {
namespace {

constexpr int kThreads = 256;
constexpr int kGridMax = 2048;      // work-groups of every reduction launch (<= partial slots)

// Device-resident BiCGStab2 state (BiCGStab2.inl:180-196 scalars).
struct KState {
    double alpha, beta, gamma, rho0, rho1;
    double om1, om2, mu, nu, tau;
    double rnorm;      // ||r|| after the last update
    double dots[kMaxDots];
    int status;        // breakdown inside the pass (step number), 0 = none
    int next_status;   // breakdown in the prelude of the next pass (steps 3-4)
};

enum TailOp : int {
    TAIL_NONE = 0,
    TAIL_STORE = 1,     // store the dots only
    TAIL_INIT = 2,      // ||r0||, prelude of pass 1 (steps 1, 3, 4)
    TAIL_GAMMA = 3,     // gamma = <v, r0>, alpha = rho0 / gamma (steps 7 / 14)
    TAIL_RHO = 4,       // rho1 = <r0, s>, beta, rho0 = rho1 (step 11)
    TAIL_GCR = 5,       // omega_1,2 (steps 19-23)
    TAIL_END = 6,       // ||r||, prelude of the next pass (steps 3-4)
};

__device__ void apply_tail(int op, KState* st, const double* d) {
    switch (op) {
    case TAIL_STORE:
        for (int j = 0; j < kMaxDots; j++) st->dots[j] = d[j];
        break;
    case TAIL_INIT: {
        st->rnorm = sqrt(d[0]);
        st->alpha = 0.;
        st->rho0 = 1.;
        st->om1 = st->om2 = 1.;
        st->status = 0;
        st->next_status = 0;
        st->rho0 = -st->om2 * st->rho0;                     // (3)
        if (st->rho0 == 0.0) {
            st->next_status = 3;
            break;
        }
        st->rho1 = d[1];                                     // (4)
        st->beta = st->alpha * st->rho1 / st->rho0;
        st->rho0 = st->rho1;
        if (st->rho0 == 0.0) st->next_status = 4;
        break;
    }
    case TAIL_GAMMA:
        st->gamma = d[0];
        if (st->gamma == 0.0) {
            st->status = 7;
            break;
        }
        st->alpha = st->rho0 / st->gamma;
        break;
    case TAIL_RHO:
        st->rho1 = d[0];
        st->beta = st->alpha * st->rho1 / st->rho0;
        st->rho0 = st->rho1;
        break;
    case TAIL_GCR: {
        // omega/mu/nu/tau are long double in the reference; FP64 here (results are cast to
        // double before use there as well, :545-568)
        double om1 = d[0], mu = d[1], nu = d[2], tau = d[3], om2 = d[4];
        st->mu = mu;
        st->nu = nu;
        if (mu == 0.) {
            st->status = 20;
            break;
        }
        tau -= nu * nu / mu;
        st->tau = tau;
        if (tau == 0.0) {
            st->status = 21;
            break;
        }
        om2 = (om2 - (nu * om1) / mu) / tau;
        om1 = (om1 - nu * om2) / mu;
        st->om1 = om1;
        st->om2 = om2;
        break;
    }
    case TAIL_END: {
        st->rnorm = sqrt(d[0]);
        st->next_status = 0;
        st->rho0 = -st->om2 * st->rho0;
        if (st->rho0 == 0.0) {
            st->next_status = 3;
            break;
        }
        st->rho1 = d[1];
        st->beta = st->alpha * st->rho1 / st->rho0;
        st->rho0 = st->rho1;
        if (st->rho0 == 0.0) st->next_status = 4;
        break;
    }
    default: break;
    }
}

// Reduce NDOT per-thread sums over the grid; the last work-group applies the tail.
template <int NDOT>
__device__ inline void grid_reduce(double (&acc)[NDOT], ReduceWs ws, KState* st, int tail) {
    __shared__ double red[NDOT][kThreads / 64];
    __shared__ int is_last;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NDOT; j++) {
        const double v = wave_sum(acc[j]);
        if (lane == 0) red[j][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x < NDOT) {
        double s = 0.;
        for (int w = 0; w < kThreads / 64; w++) s += red[threadIdx.x][w];
        ws.partial[blockIdx.x * kMaxDots + threadIdx.x] = s;
    }
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned prev = atomicAdd(ws.count, 1u);
        is_last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    double tot[NDOT];
#pragma unroll
    for (int j = 0; j < NDOT; j++) {
        double s = 0.;
        for (unsigned b = threadIdx.x; b < gridDim.x; b += kThreads) {
            s += __hip_atomic_load(&ws.partial[b * kMaxDots + j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        s = wave_sum(s);
        __syncthreads();
        if (lane == 0) red[j][wid] = s;
        __syncthreads();
        double t = 0.;
        for (int w = 0; w < kThreads / 64; w++) t += red[j][w];
        tot[j] = t;
    }
    if (threadIdx.x == 0) {
        double d[kMaxDots];
        for (int j = 0; j < kMaxDots; j++) d[j] = j < NDOT ? tot[j] : 0.;
        apply_tail(tail, st, d);
        *ws.count = 0u;
    }
}

template <typename VT, typename XT>
__device__ inline double csr_row(const int* __restrict__ rp, const int* __restrict__ col,
                                 const VT* __restrict__ val, const XT* __restrict__ x, long long r) {
    double s = 0.;
    const int e = rp[r + 1];
    for (int k = rp[r]; k < e; k++) s += static_cast<double>(val[k]) * static_cast<double>(x[col[k]]);
    return s;
}

template <typename VT>
struct Csr {
    const int* rp;
    const int* col;
    const VT* val;
    long long n;
};

// ---- per-pass kernels (status != 0 => the pass has broken down: every launch is a no-op)

// r = b - A x; rh = r; u = v = w = 0; dots <r,r>, <rh,r>     (BiCGStab2.inl:265-273)
template <typename VT, typename XT>
__global__ __launch_bounds__(kThreads) void k_init(Csr<VT> A, const XT* __restrict__ b, const XT* __restrict__ x,
                                                   XT* __restrict__ r, XT* __restrict__ rh, XT* __restrict__ u,
                                                   XT* __restrict__ v, XT* __restrict__ w, ReduceWs ws, KState* st) {
    double acc[2] = {0., 0.};
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < A.n; i += (long long)gridDim.x * kThreads) {
        const double ax = csr_row(A.rp, A.col, A.val, x, i);
        const XT ri = static_cast<XT>(static_cast<double>(b[i]) - ax);
        r[i] = ri;
        rh[i] = ri;
        u[i] = XT(0);
        v[i] = XT(0);
        w[i] = XT(0);
        const double rd = static_cast<double>(ri);
        acc[0] += rd * rd;
        acc[1] += rd * rd;
    }
    grid_reduce<2>(acc, ws, st, TAIL_INIT);
}

// u = r - beta u                                              (5) / (15 part)
template <typename XT>
__global__ __launch_bounds__(kThreads) void k_upd_u(long long n, const XT* __restrict__ r, XT* __restrict__ u,
                                                    const KState* __restrict__ st) {
    if (st->status) return;
    const double beta = st->beta;
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads)
        u[i] = static_cast<XT>(static_cast<double>(r[i]) - beta * static_cast<double>(u[i]));
}

// y = A p; dot <y, rh>; tail (gamma or rho)                   (6)-(7), (9)+(11), (13)-(14)
template <typename VT, typename XT>
__global__ __launch_bounds__(kThreads) void k_spmv_dot(Csr<VT> A, const XT* __restrict__ p, XT* __restrict__ y,
                                                       const XT* __restrict__ rh, ReduceWs ws, KState* st, int tail) {
    double acc[1] = {0.};
    if (st->status == 0) {
        for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < A.n; i += (long long)gridDim.x * kThreads) {
            const XT yi = static_cast<XT>(csr_row(A.rp, A.col, A.val, p, i));
            y[i] = yi;
            acc[0] += static_cast<double>(yi) * static_cast<double>(rh[i]);
        }
    }
    if (st->status) {
        // keep the arrival counter consistent: every block still passes the reduction
    }
    grid_reduce<1>(acc, ws, st, st->status ? TAIL_NONE : tail);
}

// r = r - alpha v; x = x + alpha u                           (8), (10)
template <typename XT, typename XO>
__global__ __launch_bounds__(kThreads) void k_upd_rx(long long n, XT* __restrict__ r, const XT* __restrict__ v,
                                                     XO* __restrict__ x, const XT* __restrict__ u,
                                                     const KState* __restrict__ st) {
    if (st->status) return;
    const double a = st->alpha;
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
        r[i] = static_cast<XT>(static_cast<double>(r[i]) - a * static_cast<double>(v[i]));
        x[i] = static_cast<XO>(1. * static_cast<double>(x[i]) + a * static_cast<double>(u[i]));
    }
}

// v = s - beta v                                              (12)
template <typename XT>
__global__ __launch_bounds__(kThreads) void k_upd_v(long long n, const XT* __restrict__ s, XT* __restrict__ v,
                                                    const KState* __restrict__ st) {
    if (st->status) return;
    const double beta = st->beta;
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads)
        v[i] = static_cast<XT>(static_cast<double>(s[i]) - beta * static_cast<double>(v[i]));
}

// u = r - beta u; r = r - alpha v; s = s - alpha w            (15)-(17)
template <typename XT>
__global__ __launch_bounds__(kThreads) void k_upd_urs(long long n, XT* __restrict__ u, XT* __restrict__ r,
                                                      const XT* __restrict__ v, XT* __restrict__ s,
                                                      const XT* __restrict__ w, const KState* __restrict__ st) {
    if (st->status) return;
    const double a = st->alpha, beta = st->beta;
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
        const double ri = static_cast<double>(r[i]);
        u[i] = static_cast<XT>(ri - beta * static_cast<double>(u[i]));
        r[i] = static_cast<XT>(ri - a * static_cast<double>(v[i]));
        s[i] = static_cast<XT>(static_cast<double>(s[i]) - a * static_cast<double>(w[i]));
    }
}

// t = A s; dots <r,s>, <s,s>, <s,t>, <t,t>, <r,t>             (18)-(23)
template <typename VT, typename XT>
__global__ __launch_bounds__(kThreads) void k_spmv_gcr(Csr<VT> A, const XT* __restrict__ s, XT* __restrict__ t,
                                                       const XT* __restrict__ r, ReduceWs ws, KState* st) {
    double acc[5] = {0., 0., 0., 0., 0.};
    if (st->status == 0) {
        for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < A.n; i += (long long)gridDim.x * kThreads) {
            const XT ti = static_cast<XT>(csr_row(A.rp, A.col, A.val, s, i));
            t[i] = ti;
            const double sd = static_cast<double>(s[i]), td = static_cast<double>(ti), rd = static_cast<double>(r[i]);
            acc[0] += rd * sd;
            acc[1] += sd * sd;
            acc[2] += sd * td;
            acc[3] += td * td;
            acc[4] += rd * td;
        }
    }
    grid_reduce<5>(acc, ws, st, st->status ? TAIL_NONE : TAIL_GCR);
}

// x += om1 r + om2 s; x += alpha u; r -= om1 s + om2 t; u -= om1 v + om2 w; dots <r,r>, <rh,r>
template <typename XT, typename XO>
__global__ __launch_bounds__(kThreads) void k_end(long long n, XO* __restrict__ x, XT* __restrict__ r,
                                                  const XT* __restrict__ s, const XT* __restrict__ t,
                                                  XT* __restrict__ u, const XT* __restrict__ v,
                                                  const XT* __restrict__ w, const XT* __restrict__ rh, ReduceWs ws,
                                                  KState* st) {
    double acc[2] = {0., 0.};
    if (st->status == 0) {
        const double o1 = st->om1, o2 = st->om2, a = st->alpha;
        for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
            const double ri = static_cast<double>(r[i]), si = static_cast<double>(s[i]);
            double xi = static_cast<double>(x[i]);
            xi = 1. * xi + o1 * ri + o2 * si;                                  // (24)
            xi = 1. * xi + a * static_cast<double>(u[i]);
            x[i] = static_cast<XO>(xi);
            const XT rn = static_cast<XT>(1. * ri + (-o1) * si + (-o2) * static_cast<double>(t[i]));   // (25)
            r[i] = rn;
            u[i] = static_cast<XT>(1. * static_cast<double>(u[i]) + (-o1) * static_cast<double>(v[i]) +
                                   (-o2) * static_cast<double>(w[i]));
            const double rd = static_cast<double>(rn);
            acc[0] += rd * rd;
            acc[1] += static_cast<double>(rh[i]) * rd;
        }
    }
    grid_reduce<2>(acc, ws, st, st->status ? TAIL_NONE : TAIL_END);
}

// ---- generic primitives

template <typename VT, typename XT>
__global__ __launch_bounds__(kThreads) void k_spmv(Csr<VT> A, const XT* __restrict__ x, XT* __restrict__ y) {
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < A.n; i += (long long)gridDim.x * kThreads)
        y[i] = static_cast<XT>(csr_row(A.rp, A.col, A.val, x, i));
}

template <typename XT>
__global__ __launch_bounds__(kThreads) void k_dot(long long n, const XT* __restrict__ a, const XT* __restrict__ b,
                                                  ReduceWs ws, KState* st) {
    double acc[1] = {0.};
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads)
        acc[0] += static_cast<double>(a[i]) * static_cast<double>(b[i]);
    grid_reduce<1>(acc, ws, st, TAIL_STORE);
}

__global__ __launch_bounds__(kThreads) void k_axpby(long long n, double a, const double* __restrict__ x, double b,
                                                    double* __restrict__ y) {
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads)
        y[i] = a * x[i] + b * y[i];
}

// refinement: r = b - A x (FP64); r32 = float(r); dot <r,r>
__global__ __launch_bounds__(kThreads) void k_resid_f64(Csr<double> A, const double* __restrict__ b,
                                                        const double* __restrict__ x, float* __restrict__ r32,
                                                        ReduceWs ws, KState* st) {
    double acc[1] = {0.};
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < A.n; i += (long long)gridDim.x * kThreads) {
        const double ri = b[i] - csr_row(A.rp, A.col, A.val, x, i);
        r32[i] = static_cast<float>(ri);
        acc[0] += ri * ri;
    }
    grid_reduce<1>(acc, ws, st, TAIL_STORE);
}

// x += double(d)
__global__ __launch_bounds__(kThreads) void k_add_correction(long long n, double* __restrict__ x,
                                                             const float* __restrict__ d) {
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads)
        x[i] = x[i] + static_cast<double>(d[i]);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_zero(long long n, T* __restrict__ p) {
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads)
        p[i] = T(0);
}

__global__ __launch_bounds__(kThreads) void k_to_f32(long long n, const double* __restrict__ a, float* __restrict__ b) {
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads)
        b[i] = static_cast<float>(a[i]);
}

inline unsigned grid_for(long long n) {
    long long g = (n + kThreads - 1) / kThreads;
    if (g > kGridMax) g = kGridMax;
    if (g < 1) g = 1;
    return static_cast<unsigned>(g);
}

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

// Host-driven BiCGStab2 over device-resident scalars. XO is the type of x (the caller's
// solution vector), XT the type of the work vectors, VT the matrix value type.
template <typename VT, typename XT, typename XO>
int bicgstab2_run(dare_ctx* ctx, const Csr<VT>& A, const XO* b, XO* x, double tol_abs, int max_it,
                  double* hist, int hist_cap, int* hist_len, int* iters, double* rnorm, int* status,
                  bool* converged, XT* work) {
    const long long n = A.n;
    XT* r = work;
    XT* rh = r + n;
    XT* u = rh + n;
    XT* v = u + n;
    XT* s = v + n;
    XT* w = s + n;
    XT* t = w + n;
    KState* st = static_cast<KState*>(ctx->solve_state);
    const unsigned gr = grid_for(n);
    hipStream_t sm = ctx->stream;
    *hist_len = 0;
    *status = 0;
    // (0)-(2): the init kernel consumes b and x of type XO via XT views when XO == XT
    hipLaunchKernelGGL((k_init<VT, XT>), dim3(gr), dim3(kThreads), 0, sm, A, reinterpret_cast<const XT*>(b),
                       reinterpret_cast<const XT*>(x), r, rh, u, v, w, ctx->ws, st);
    DARE_LAUNCH_CHECK();
    KState h;
    int rc = read_state(ctx, st, &h);
    if (rc) return rc;
    double conv = h.rnorm;
    if (conv < tol_abs) {
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
        hipLaunchKernelGGL((k_upd_u<XT>), dim3(gr), dim3(kThreads), 0, sm, n, r, u, st);
        hipLaunchKernelGGL((k_spmv_dot<VT, XT>), dim3(gr), dim3(kThreads), 0, sm, A, u, v, rh, ctx->ws, st,
                           (int)TAIL_GAMMA);
        hipLaunchKernelGGL((k_upd_rx<XT, XO>), dim3(gr), dim3(kThreads), 0, sm, n, r, v, x, u, st);
        hipLaunchKernelGGL((k_spmv_dot<VT, XT>), dim3(gr), dim3(kThreads), 0, sm, A, r, s, rh, ctx->ws, st,
                           (int)TAIL_RHO);
        hipLaunchKernelGGL((k_upd_v<XT>), dim3(gr), dim3(kThreads), 0, sm, n, s, v, st);
        hipLaunchKernelGGL((k_spmv_dot<VT, XT>), dim3(gr), dim3(kThreads), 0, sm, A, v, w, rh, ctx->ws, st,
                           (int)TAIL_GAMMA);
        hipLaunchKernelGGL((k_upd_urs<XT>), dim3(gr), dim3(kThreads), 0, sm, n, u, r, v, s, w, st);
        hipLaunchKernelGGL((k_spmv_gcr<VT, XT>), dim3(gr), dim3(kThreads), 0, sm, A, s, t, r, ctx->ws, st);
        hipLaunchKernelGGL((k_end<XT, XO>), dim3(gr), dim3(kThreads), 0, sm, n, x, r, s, t, u, v, w, rh, ctx->ws, st);
        DARE_LAUNCH_CHECK();
        rc = read_state(ctx, st, &h);
        if (rc) return rc;
        if (h.status) {
            *status = h.status;
            if (h.status == 21) {
                // tau == 0: true residual check (BiCGStab2.inl:514-536)
                hipLaunchKernelGGL((k_init<VT, XT>), dim3(gr), dim3(kThreads), 0, sm, A,
                                   reinterpret_cast<const XT*>(b), reinterpret_cast<const XT*>(x), r, rh, u, v, w,
                                   ctx->ws, st);
                rc = read_state(ctx, st, &h);
                if (rc) return rc;
                conv = h.rnorm;
                if (conv <= tol_abs) {
                    done_ok = true;
                    *status = 0;
                }
            }
            break;
        }
        conv = h.rnorm;
        if (hist && *hist_len < hist_cap) hist[(*hist_len)++] = conv;
        if (conv <= tol_abs && k > 1) break;
        ++k;
    }
    *iters = k;
    *rnorm = conv;
    *converged = done_ok || conv <= tol_abs;
    return DARE_OK;
}

static int dot_dev(dare_ctx* ctx, long long n, const double* a, const double* b, double* out) {
    Workspace W;
    int rc = ws_reserve(ctx, 0, &W);
    if (rc) return rc;
    hipLaunchKernelGGL((k_dot<double>), dim3(grid_for(n)), dim3(kThreads), 0, ctx->stream, n, a, b, ctx->ws, W.st);
    DARE_LAUNCH_CHECK();
    KState h;
    rc = read_state(ctx, W.st, &h);
    if (rc) return rc;
    *out = h.dots[0];
    return DARE_OK;
}

}  // namespace
}  // namespace dare_hip

using namespace dare_hip;

extern "C" {

int dare_spmv(dare_ctx* ctx, const dare_csr* A, const double* x, double* y) {
    DARE_CHECK_ARG(ctx && A && x && y, "NULL argument");
    Csr<double> C{A->row_ptr, A->col, A->val, A->nrows};
    hipLaunchKernelGGL((k_spmv<double, double>), dim3(grid_for(A->nrows)), dim3(kThreads), 0, ctx->stream, C, x, y);
    DARE_LAUNCH_CHECK();
    return DARE_OK;
}

int dare_spmv_dot(dare_ctx* ctx, const dare_csr* A, const double* x, double* y, const double* w,
                  double* out_host) {
    DARE_CHECK_ARG(ctx && A && x && y && w, "NULL argument");
    Workspace W;
    int rc = ws_reserve(ctx, 0, &W);
    if (rc) return rc;
    // the state must not carry a breakdown flag into the fused kernel
    DARE_HIP_TRY(hipMemsetAsync(W.st, 0, sizeof(KState), ctx->stream));
    Csr<double> C{A->row_ptr, A->col, A->val, A->nrows};
    hipLaunchKernelGGL((k_spmv_dot<double, double>), dim3(grid_for(A->nrows)), dim3(kThreads), 0, ctx->stream, C, x,
                       y, w, ctx->ws, W.st, (int)TAIL_STORE);
    DARE_LAUNCH_CHECK();
    if (out_host) {
        KState h;
        rc = read_state(ctx, W.st, &h);
        if (rc) return rc;
        *out_host = h.dots[0];
    }
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
    Workspace W;
    int rc = ws_reserve(ctx, sizeof(double) * 7 * n, &W);
    if (rc) return rc;
    double bnorm2 = 0.;
    rc = dot_dev(ctx, n, b, b, &bnorm2);
    if (rc) return rc;
    const double bnorm = std::sqrt(bnorm2);
    const double tol_abs = p->tol_kind == 1 ? p->tol * bnorm : p->tol;
    Csr<double> C{A->row_ptr, A->col, A->val, n};
    int iters = 0, status = 0, hl = 0;
    double rn = 0.;
    bool conv = false;
    rc = bicgstab2_run<double, double, double>(ctx, C, b, x, tol_abs, p->max_it, p->history, p->history_cap, &hl,
                                               &iters, &rn, &status, &conv, static_cast<double*>(W.buf));
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
    // FP32 copy of the operator values (rebuilt on every call: values change every step)
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
    hipLaunchKernelGGL(k_to_f32, dim3(grid_for(A->nnz)), dim3(kThreads), 0, ctx->stream, (long long)A->nnz, A->val,
                       A->val32);
    DARE_LAUNCH_CHECK();
    Workspace W;
    // 7 FP32 work vectors + FP32 residual + FP32 correction
    int rc = ws_reserve(ctx, sizeof(float) * 9 * n, &W);
    if (rc) return rc;
    float* work = static_cast<float*>(W.buf);
    float* r32 = work + 7 * n;
    float* d32 = r32 + n;
    double bnorm2 = 0.;
    rc = dot_dev(ctx, n, b, b, &bnorm2);
    if (rc) return rc;
    const double bnorm = std::sqrt(bnorm2);
    const double tol_abs = p->tol_kind == 1 ? p->tol * bnorm : p->tol;
    Csr<double> A64{A->row_ptr, A->col, A->val, n};
    Csr<float> A32{A->row_ptr, A->col, A->val32, n};
    const unsigned gr = grid_for(n);
    int total_it = 0, refine = 0, hl = 0;
    double rn = 0.;
    bool conv = false;
    int status = 0;
    const int max_refine = p->max_refine > 0 ? p->max_refine : 20;
    const int inner_max = p->inner_max_it > 0 ? p->inner_max_it : p->max_it;
    const double inner_tol = p->inner_tol > 0. ? p->inner_tol : 1e-4;
    while (true) {
        hipLaunchKernelGGL(k_resid_f64, dim3(gr), dim3(kThreads), 0, ctx->stream, A64, b, x, r32, ctx->ws, W.st);
        DARE_LAUNCH_CHECK();
        KState h;
        rc = read_state(ctx, W.st, &h);
        if (rc) return rc;
        rn = std::sqrt(h.dots[0]);
        if (p->history && hl < p->history_cap) p->history[hl++] = rn;
        if (rn <= tol_abs) {
            conv = true;
            break;
        }
        if (refine >= max_refine || total_it >= p->max_it) break;
        hipLaunchKernelGGL((k_zero<float>), dim3(gr), dim3(kThreads), 0, ctx->stream, n, d32);
        int it = 0, ist = 0, dummy_len = 0;
        double irn = 0.;
        bool iconv = false;
        rc = bicgstab2_run<float, float, float>(ctx, A32, r32, d32, inner_tol * rn, inner_max, nullptr, 0, &dummy_len,
                                                &it, &irn, &ist, &iconv, work);
        if (rc) return rc;
        total_it += it;
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
