// Krylov kernels and the BiCGStab(2) driver for gfx950.
//
// BiCGStab2::Solve (MatrixSystem/BiCGStab2.inl:167-582) is restated as nine fused launches per
// pass. All method scalars (alpha, beta, gamma, rho, omega, mu, nu, tau) live in device memory
// and are updated by the last work-group of each one-pass grid reduction, so the host reads
// back only one small status block per pass (the convergence test). Reductions are
// deterministic: fixed grid, fixed per-thread order, block partials summed in block order.
//
// Operators: a generic CSR traversal (grid-stride, one row per lane) and, for matrices
// assembled from the 7-point structure, a tiled traversal of the slot-major diagonal view:
// 64 (k) x 4 (j) x 16 (i) cells per work-group, the i-1/i/i+1 planes of x rolled through
// registers, tiles ordered so that each XCD (blocks b = x mod 8) owns one contiguous slab.
// Matrix values then stream fully coalesced and x is read from HBM about once per SpMV.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"

namespace dare_hip {
namespace {

constexpr int kThreads = 256;
constexpr int kGridMax = 2048;      // work-groups of the grid-stride launches
constexpr int kTileK = 64, kTileJ = 4, kTileI = 16;
constexpr int kXcd = 8;             // MI355X: 8 accelerator complex dies, blocks dealt round robin

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
        // omega/mu/nu/tau are long double in the reference; FP64 here (they are cast to
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

// Block-level part of a grid reduction: each work-group writes its NDOT partial sums with
// plain stores. No fence or arrival counter here: on MI355X the eight XCD L2 caches are not
// coherent, and an agent-scope release fence in every work-group writes the L2 back each time.
// The kernel boundary publishes the partials; k_finish (one work-group) sums them in block
// order — deterministic — and applies the BiCGStab2 scalar update on the device.
template <int NDOT>
__device__ inline void block_partials(double (&acc)[NDOT], ReduceWs ws) {
    __shared__ double red[NDOT][kThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NDOT; j++) {
        const double v = wave_sum(acc[j]);
        if (lane == 0) red[j][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x < NDOT) {     // dot-major: k_finish reads each dot's partials contiguously
        double s = 0.;
        for (int w = 0; w < kThreads / 64; w++) s += red[threadIdx.x][w];
        ws.partial[threadIdx.x * kReduceBlocks + blockIdx.x] = s;
    }
}

constexpr int kFinishThreads = 1024;

template <int NDOT>
__global__ __launch_bounds__(kFinishThreads) void k_finish(ReduceWs ws, unsigned nblocks, KState* st, int tail,
                                                           int check_status) {
    if (check_status && st->status) return;       // the pass has broken down: keep its state
    __shared__ double red[kMaxDots][kFinishThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double tot[NDOT];
#pragma unroll
    for (int j = 0; j < NDOT; j++) {
        // fixed order: thread t owns blocks t, t + 1024, ...; loads issued before the adds
        const double* p = ws.partial + j * kReduceBlocks;
        double v[kReduceBlocks / kFinishThreads];
#pragma unroll
        for (int q = 0; q < kReduceBlocks / kFinishThreads; q++) {
            const unsigned b = threadIdx.x + q * kFinishThreads;
            v[q] = b < nblocks ? p[b] : 0.;
        }
        double s = 0.;
#pragma unroll
        for (int q = 0; q < kReduceBlocks / kFinishThreads; q++) s += v[q];
        s = wave_sum(s);
        if (lane == 0) red[j][wid] = s;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NDOT; j++) {
        double t = 0.;
        for (int w = 0; w < kFinishThreads / 64; w++) t += red[j][w];
        tot[j] = t;
    }
    if (threadIdx.x == 0) {
        double d[kMaxDots];
        for (int j = 0; j < kMaxDots; j++) d[j] = j < NDOT ? tot[j] : 0.;
        apply_tail(tail, st, d);
    }
}

// grid_reduce keeps the call sites of the kernels: the block part only.
template <int NDOT>
__device__ inline void grid_reduce(double (&acc)[NDOT], ReduceWs ws, KState*, int) {
    block_partials<NDOT>(acc, ws);
}

template <int NDOT>
inline void launch_finish(hipStream_t sm, ReduceWs ws, unsigned nblocks, KState* st, int tail, int check) {
    hipLaunchKernelGGL((k_finish<NDOT>), dim3(1), dim3(kFinishThreads), 0, sm, ws, nblocks, st, tail, check);
}

// ------------------------------------------------------------------------------ operators
template <typename VT>
struct CsrOp {
    const int* rp;
    const int* col;
    const VT* val;
    long long n;
};

template <typename VT>
struct DiaOp {
    DiaView V;
    const VT* D;
    int tiles_k, tiles_j, tiles_i;
    int tile_i;         // planes per tile (grown until the grid fits the reduction slots)
    int per_xcd;        // tiles owned by each XCD (grid = kXcd * per_xcd)
};

template <typename VT>
DiaOp<VT> make_dia(const DiaView& V, const VT* D) {
    DiaOp<VT> A;
    A.V = V;
    A.D = D;
    A.tiles_k = (V.n[2] + kTileK - 1) / kTileK;
    A.tiles_j = (V.n[1] + kTileJ - 1) / kTileJ;
    A.tile_i = kTileI;
    long long T = 0;
    for (;;) {
        A.tiles_i = (V.n[0] + A.tile_i - 1) / A.tile_i;
        T = static_cast<long long>(A.tiles_k) * A.tiles_j * A.tiles_i;
        if (kXcd * ((T + kXcd - 1) / kXcd) <= kReduceBlocks || A.tile_i >= V.n[0]) break;
        A.tile_i *= 2;
    }
    A.per_xcd = static_cast<int>((T + kXcd - 1) / kXcd);
    return A;
}

// ------------------------------------------------------------------------------ epilogues
// load(i) fetches the row's vector operands (issued together with the matrix loads);
// operator() gets y = (A x)_i, xc = x_i and the loaded operands, stores results and adds to
// the per-thread dot accumulators.
template <typename XT>
struct EpiInit {        // r = b - A x; rh = r; u = v = w = 0; <r,r>, <rh,r>   (BiCGStab2.inl:265-273)
    static constexpr int NDOT = 2;
    const XT* b;
    XT *r, *rh, *u, *v, *w;
    __device__ double load(long long i) const { return static_cast<double>(b[i]); }
    __device__ void operator()(long long i, double y, double, double bi, double* acc) const {
        const XT ri = static_cast<XT>(bi - y);
        r[i] = ri;
        rh[i] = ri;
        u[i] = XT(0);
        v[i] = XT(0);
        w[i] = XT(0);
        const double rd = static_cast<double>(ri);
        acc[0] += rd * rd;
        acc[1] += rd * rd;
    }
};

template <typename XT>
struct EpiDot {         // out = A p; <out, w>                   (6)-(7), (9)+(11), (13)-(14)
    static constexpr int NDOT = 1;
    XT* out;
    const XT* w;
    __device__ double load(long long i) const { return static_cast<double>(w[i]); }
    __device__ void operator()(long long i, double y, double, double wi, double* acc) const {
        const XT yi = static_cast<XT>(y);
        out[i] = yi;
        acc[0] += static_cast<double>(yi) * wi;
    }
};

template <typename XT>
struct EpiGcr {         // t = A s; <r,s>, <s,s>, <s,t>, <t,t>, <r,t>   (18)-(20)
    static constexpr int NDOT = 5;
    XT* t;
    const XT* r;
    __device__ double load(long long i) const { return static_cast<double>(r[i]); }
    __device__ void operator()(long long i, double y, double sc, double rd, double* acc) const {
        const XT ti = static_cast<XT>(y);
        t[i] = ti;
        const double td = static_cast<double>(ti);
        acc[0] += rd * sc;
        acc[1] += sc * sc;
        acc[2] += sc * td;
        acc[3] += td * td;
        acc[4] += rd * td;
    }
};

struct EpiResid {       // refinement: r = b - A x in FP64, r32 = float(r), <r,r>
    static constexpr int NDOT = 1;
    const double* b;
    float* r32;
    __device__ double load(long long i) const { return b[i]; }
    __device__ void operator()(long long i, double y, double, double bi, double* acc) const {
        const double ri = bi - y;
        r32[i] = static_cast<float>(ri);
        acc[0] += ri * ri;
    }
};

template <typename XT>
struct EpiStore {
    static constexpr int NDOT = 1;
    XT* out;
    __device__ double load(long long) const { return 0.; }
    __device__ void operator()(long long i, double y, double, double, double*) const { out[i] = static_cast<XT>(y); }
};

// ------------------------------------------------------------------------------ traversals
// Generic CSR: one row per lane; the row's entries are fetched four at a time with clamped
// indices (all loads in flight before the multiply-adds), summed in column order.
template <typename VT, typename XT, class Epi>
__global__ __launch_bounds__(kThreads) void k_apply_csr(CsrOp<VT> A, const XT* __restrict__ x, Epi epi, ReduceWs ws,
                                                       KState* st, int tail, int check_status) {
    double acc[Epi::NDOT];
#pragma unroll
    for (int j = 0; j < Epi::NDOT; j++) acc[j] = 0.;
    const bool run = !(check_status && st->status);
    if (run) {
        for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < A.n; i += (long long)gridDim.x * kThreads) {
            const int b = A.rp[i], e = A.rp[i + 1];
            const double pre = epi.load(i);
            const double xi = static_cast<double>(x[i]);
            double s = 0.;
            for (int k = b; k < e; k += 4) {
                double vv[4], xv[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int kk = (k + u < e) ? k + u : b;
                    vv[u] = static_cast<double>(A.val[kk]);
                    xv[u] = static_cast<double>(x[A.col[kk]]);
                }
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (k + u < e) s += vv[u] * xv[u];
            }
            epi(i, s, xi, pre, acc);
        }
    }
    if (tail >= 0) grid_reduce<Epi::NDOT>(acc, ws, st, run ? tail : TAIL_NONE);
}

// y_row = sum over the row's slots in the order W,S,B,[X],C,[X],T,N,E (the sorted column
// order of a non-wrapped row; X is the cross entry, before C for component 1, after it for 0).
// Every load of a plane is issued before the multiply-adds; entries removed by a boundary
// edit read a valid in-range address and contribute +0.0 (the partial sum starts at +0.0 and
// so is never -0.0: adding +0.0 leaves it bit-identical to skipping the term).
template <int NC, typename VT, typename XT, class Epi>
__global__ __launch_bounds__(kThreads) void k_apply_dia(DiaOp<VT> A, const XT* __restrict__ x, Epi epi, ReduceWs ws,
                                                       KState* st, int tail, int check_status) {
    double acc[Epi::NDOT];
#pragma unroll
    for (int j = 0; j < Epi::NDOT; j++) acc[j] = 0.;
    const bool run = !(check_status && st->status);
    const int tile = static_cast<int>(blockIdx.x % kXcd) * A.per_xcd + static_cast<int>(blockIdx.x / kXcd);
    const int ntiles = A.tiles_k * A.tiles_j * A.tiles_i;
    if (run && tile < ntiles) {
        const DiaView& V = A.V;
        const int tk = tile % A.tiles_k;
        const int t2 = tile / A.tiles_k;
        const int tj = t2 % A.tiles_j;
        const int ti = t2 / A.tiles_j;
        const int k = tk * kTileK + (threadIdx.x & 63);
        const int j = tj * kTileJ + (threadIdx.x >> 6);
        if (k < V.n[2] && j < V.n[1]) {
            const int n0 = V.n[0], n1 = V.n[1], n2 = V.n[2];
            const long long sI = static_cast<long long>(n1) * n2, sJ = n2;
            const long long nr = V.nrows;
            // j and k neighbours: presence (boundary edits) and wrap
            const bool hasS = !(j == 0 && V.rm_lo[1]);
            const bool hasN = !(j == n1 - 1 && V.rm_hi[1] && !(j == 0 && V.rm_lo[1]));
            const bool hasB = !(k == 0 && V.rm_lo[2]);
            const bool hasT = !(k == n2 - 1 && V.rm_hi[2] && !(k == 0 && V.rm_lo[2]));
            const long long dS = !hasS ? 0 : (j == 0) ? (long long)(n1 - 1) * sJ : -sJ;
            const long long dN = !hasN ? 0 : (j == n1 - 1) ? -(long long)(n1 - 1) * sJ : sJ;
            const long long dB = !hasB ? 0 : (k == 0) ? (long long)(n2 - 1) : -1;
            const long long dT = !hasT ? 0 : (k == n2 - 1) ? -(long long)(n2 - 1) : 1;
            const int i0 = ti * A.tile_i;
            const int i1 = (i0 + A.tile_i < n0) ? i0 + A.tile_i : n0;
            long long node = (static_cast<long long>(i0) * n1 + j) * n2 + k;
            double xp[NC], xc[NC];
            {
                const bool need = !(i0 == 0 && V.rm_lo[0]);
                const long long nprev = !need ? node : (i0 == 0) ? node + (long long)(n0 - 1) * sI : node - sI;
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    xc[c] = static_cast<double>(x[node * NC + c]);
                    xp[c] = static_cast<double>(x[nprev * NC + c]);
                }
            }
            for (int i = i0; i < i1; i++) {
                const bool hasW = !(i == 0 && V.rm_lo[0]);
                const bool hasE = !(i == n0 - 1 && V.rm_hi[0] && !(i == 0 && V.rm_lo[0]));
                const long long nnext = !hasE ? node : (i == n0 - 1) ? node - (long long)(n0 - 1) * sI : node + sI;
                double xn[NC], xs[NC], xnn[NC], xb[NC], xt[NC], dv[NC][8], pre[NC];
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    const long long row = node * NC + c;
                    xn[c] = static_cast<double>(x[nnext * NC + c]);
                    xs[c] = static_cast<double>(x[(node + dS) * NC + c]);
                    xnn[c] = static_cast<double>(x[(node + dN) * NC + c]);
                    xb[c] = static_cast<double>(x[(node + dB) * NC + c]);
                    xt[c] = static_cast<double>(x[(node + dT) * NC + c]);
#pragma unroll
                    for (int sl = 0; sl < 7; sl++) dv[c][sl] = static_cast<double>(A.D[sl * nr + row]);
                    dv[c][7] = (NC == 2 && V.cross) ? static_cast<double>(A.D[7 * nr + row]) : 0.;
                    pre[c] = epi.load(row);
                }
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    double s = 0.;
                    s += hasW ? dv[c][0] * xp[c] : 0.;
                    s += hasS ? dv[c][1] * xs[c] : 0.;
                    s += hasB ? dv[c][2] * xb[c] : 0.;
                    if (NC == 2 && V.cross && c == 1) s += dv[c][7] * xc[0];
                    s += dv[c][3] * xc[c];
                    if (NC == 2 && V.cross && c == 0) s += dv[c][7] * xc[NC - 1];
                    s += hasT ? dv[c][4] * xt[c] : 0.;
                    s += hasN ? dv[c][5] * xnn[c] : 0.;
                    s += hasE ? dv[c][6] * xn[c] : 0.;
                    epi(node * NC + c, s, xc[c], pre[c], acc);
                }
#pragma unroll
                for (int c = 0; c < NC; c++) {
                    xp[c] = xc[c];
                    xc[c] = xn[c];
                }
                node += sI;   // a tile never runs past plane n0 - 1, so xc is always a real plane
            }
        }
    }
    if (tail >= 0) grid_reduce<Epi::NDOT>(acc, ws, st, run ? tail : TAIL_NONE);
}

// ------------------------------------------------------------------------------ vector updates
// Grid-stride over chunks of kU * kThreads elements; all loads of a chunk are issued before
// its arithmetic so that every wave keeps several independent HBM requests in flight.
// (status != 0 => the pass has broken down: every launch is a no-op)
constexpr int kU = 4;

#define DARE_CHUNK_LOOP(n)                                                                        \
    for (long long base_ = blockIdx.x * (long long)(kThreads * kU); base_ < (n);                  \
         base_ += (long long)gridDim.x * (kThreads * kU))
#define DARE_IDX(u) (base_ + (long long)(u) * kThreads + threadIdx.x)

template <typename XT>
__global__ __launch_bounds__(kThreads) void k_upd_u(long long n, const XT* __restrict__ r, XT* __restrict__ u,
                                                    const KState* __restrict__ st) {
    if (st->status) return;
    const double beta = st->beta;
    DARE_CHUNK_LOOP(n) {
        double rv[kU], uv[kU];
#pragma unroll
        for (int q = 0; q < kU; q++) {
            const long long i = DARE_IDX(q) < n ? DARE_IDX(q) : 0;
            rv[q] = static_cast<double>(r[i]);
            uv[q] = static_cast<double>(u[i]);
        }
#pragma unroll
        for (int q = 0; q < kU; q++)
            if (DARE_IDX(q) < n) u[DARE_IDX(q)] = static_cast<XT>(rv[q] - beta * uv[q]);          // (5)
    }
}

template <typename XT>
__global__ __launch_bounds__(kThreads) void k_upd_rx(long long n, XT* __restrict__ r, const XT* __restrict__ v,
                                                     XT* __restrict__ x, const XT* __restrict__ u,
                                                     const KState* __restrict__ st) {
    if (st->status) return;
    const double a = st->alpha;
    DARE_CHUNK_LOOP(n) {
        double rv[kU], vv[kU], xv[kU], uv[kU];
#pragma unroll
        for (int q = 0; q < kU; q++) {
            const long long i = DARE_IDX(q) < n ? DARE_IDX(q) : 0;
            rv[q] = static_cast<double>(r[i]);
            vv[q] = static_cast<double>(v[i]);
            xv[q] = static_cast<double>(x[i]);
            uv[q] = static_cast<double>(u[i]);
        }
#pragma unroll
        for (int q = 0; q < kU; q++) {
            const long long i = DARE_IDX(q);
            if (i < n) {
                r[i] = static_cast<XT>(rv[q] - a * vv[q]);                                         // (8)
                x[i] = static_cast<XT>(1. * xv[q] + a * uv[q]);                                    // (10)
            }
        }
    }
}

template <typename XT>
__global__ __launch_bounds__(kThreads) void k_upd_v(long long n, const XT* __restrict__ s, XT* __restrict__ v,
                                                    const KState* __restrict__ st) {
    if (st->status) return;
    const double beta = st->beta;
    DARE_CHUNK_LOOP(n) {
        double sv[kU], vv[kU];
#pragma unroll
        for (int q = 0; q < kU; q++) {
            const long long i = DARE_IDX(q) < n ? DARE_IDX(q) : 0;
            sv[q] = static_cast<double>(s[i]);
            vv[q] = static_cast<double>(v[i]);
        }
#pragma unroll
        for (int q = 0; q < kU; q++)
            if (DARE_IDX(q) < n) v[DARE_IDX(q)] = static_cast<XT>(sv[q] - beta * vv[q]);          // (12)
    }
}

template <typename XT>
__global__ __launch_bounds__(kThreads) void k_upd_urs(long long n, XT* __restrict__ u, XT* __restrict__ r,
                                                      const XT* __restrict__ v, XT* __restrict__ s,
                                                      const XT* __restrict__ w, const KState* __restrict__ st) {
    if (st->status) return;
    const double a = st->alpha, beta = st->beta;
    DARE_CHUNK_LOOP(n) {
        double uv[kU], rv[kU], vv[kU], sv[kU], wv[kU];
#pragma unroll
        for (int q = 0; q < kU; q++) {
            const long long i = DARE_IDX(q) < n ? DARE_IDX(q) : 0;
            uv[q] = static_cast<double>(u[i]);
            rv[q] = static_cast<double>(r[i]);
            vv[q] = static_cast<double>(v[i]);
            sv[q] = static_cast<double>(s[i]);
            wv[q] = static_cast<double>(w[i]);
        }
#pragma unroll
        for (int q = 0; q < kU; q++) {
            const long long i = DARE_IDX(q);
            if (i < n) {
                u[i] = static_cast<XT>(rv[q] - beta * uv[q]);                                      // (15)
                r[i] = static_cast<XT>(rv[q] - a * vv[q]);                                         // (16)
                s[i] = static_cast<XT>(sv[q] - a * wv[q]);                                         // (17)
            }
        }
    }
}

// x += om1 r + om2 s; x += alpha u; r -= om1 s + om2 t; u -= om1 v + om2 w; <r,r>, <rh,r>
template <typename XT>
__global__ __launch_bounds__(kThreads) void k_end(long long n, XT* __restrict__ x, XT* __restrict__ r,
                                                  const XT* __restrict__ s, const XT* __restrict__ t,
                                                  XT* __restrict__ u, const XT* __restrict__ v,
                                                  const XT* __restrict__ w, const XT* __restrict__ rh, ReduceWs ws,
                                                  KState* st) {
    double acc[2] = {0., 0.};
    const bool run = st->status == 0;
    if (run) {
        const double o1 = st->om1, o2 = st->om2, a = st->alpha;
        DARE_CHUNK_LOOP(n) {
            double xv[kU], rv[kU], sv[kU], tv[kU], uv[kU], vv[kU], wv[kU], hv[kU];
#pragma unroll
            for (int q = 0; q < kU; q++) {
                const long long i = DARE_IDX(q) < n ? DARE_IDX(q) : 0;
                xv[q] = static_cast<double>(x[i]);
                rv[q] = static_cast<double>(r[i]);
                sv[q] = static_cast<double>(s[i]);
                tv[q] = static_cast<double>(t[i]);
                uv[q] = static_cast<double>(u[i]);
                vv[q] = static_cast<double>(v[i]);
                wv[q] = static_cast<double>(w[i]);
                hv[q] = static_cast<double>(rh[i]);
            }
#pragma unroll
            for (int q = 0; q < kU; q++) {
                const long long i = DARE_IDX(q);
                if (i < n) {
                    double xi = 1. * xv[q] + o1 * rv[q] + o2 * sv[q];                           // (24)
                    xi = 1. * xi + a * uv[q];
                    x[i] = static_cast<XT>(xi);
                    const XT rn = static_cast<XT>(1. * rv[q] + (-o1) * sv[q] + (-o2) * tv[q]);  // (25)
                    r[i] = rn;
                    u[i] = static_cast<XT>(1. * uv[q] + (-o1) * vv[q] + (-o2) * wv[q]);
                    const double rd = static_cast<double>(rn);
                    acc[0] += rd * rd;
                    acc[1] += hv[q] * rd;
                }
            }
        }
    }
    grid_reduce<2>(acc, ws, st, run ? TAIL_END : TAIL_NONE);
}

template <typename XT>
__global__ __launch_bounds__(kThreads) void k_dot(long long n, const XT* __restrict__ a, const XT* __restrict__ b,
                                                  ReduceWs ws, KState* st) {
    double acc[1] = {0.};
    DARE_CHUNK_LOOP(n) {
        double av[kU], bv[kU];
#pragma unroll
        for (int q = 0; q < kU; q++) {
            const long long i = DARE_IDX(q) < n ? DARE_IDX(q) : 0;
            av[q] = static_cast<double>(a[i]);
            bv[q] = static_cast<double>(b[i]);
        }
#pragma unroll
        for (int q = 0; q < kU; q++)
            if (DARE_IDX(q) < n) acc[0] += av[q] * bv[q];
    }
    grid_reduce<1>(acc, ws, st, TAIL_STORE);
}

__global__ __launch_bounds__(kThreads) void k_axpby(long long n, double a, const double* __restrict__ x, double b,
                                                    double* __restrict__ y) {
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads)
        y[i] = a * x[i] + b * y[i];
}

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

__global__ void k_clear_state(KState* st) {
    if (threadIdx.x == 0) {
        st->status = 0;
        st->next_status = 0;
    }
}

inline unsigned grid_rows(long long n) {      // one row per lane per sweep (CSR traversal)
    long long g = (n + kThreads - 1) / kThreads;
    if (g > kGridMax) g = kGridMax;
    if (g < 1) g = 1;
    return static_cast<unsigned>(g);
}

inline unsigned grid_for(long long n) {
    long long g = (n + kThreads * kU - 1) / (kThreads * kU);
    if (g > kGridMax) g = kGridMax;
    if (g < 1) g = 1;
    return static_cast<unsigned>(g);
}

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

// Host-driven BiCGStab2 over device-resident scalars; XT is the type of every vector.
template <typename VT, typename XT>
int bicgstab2_run(dare_ctx* ctx, const SolveOp<VT>& A, const XT* b, XT* x, double tol_abs, int max_it,
                  double* hist, int hist_cap, int* hist_len, int* iters, double* rnorm, int* status,
                  bool* converged, XT* work, int min_passes) {
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
    ReduceWs ws = ctx->ws;
    *hist_len = 0;
    *status = 0;
    A.launch(sm, x, EpiInit<XT>{b, r, rh, u, v, w}, ws, st, TAIL_INIT, 0);     // (0)-(2)
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
        A.launch(sm, u, EpiDot<XT>{v, rh}, ws, st, TAIL_GAMMA, 1);
        hipLaunchKernelGGL((k_upd_rx<XT>), dim3(gr), dim3(kThreads), 0, sm, n, r, v, x, u, st);
        A.launch(sm, r, EpiDot<XT>{s, rh}, ws, st, TAIL_RHO, 1);
        hipLaunchKernelGGL((k_upd_v<XT>), dim3(gr), dim3(kThreads), 0, sm, n, s, v, st);
        A.launch(sm, v, EpiDot<XT>{w, rh}, ws, st, TAIL_GAMMA, 1);
        hipLaunchKernelGGL((k_upd_urs<XT>), dim3(gr), dim3(kThreads), 0, sm, n, u, r, v, s, w, st);
        A.launch(sm, s, EpiGcr<XT>{t, r}, ws, st, TAIL_GCR, 1);
        hipLaunchKernelGGL((k_end<XT>), dim3(gr), dim3(kThreads), 0, sm, n, x, r, s, t, u, v, w, rh, ws, st);
        launch_finish<2>(sm, ws, gr, st, TAIL_END, 1);
        DARE_LAUNCH_CHECK();
        rc = read_state(ctx, st, &h);
        if (rc) return rc;
        if (h.status) {
            *status = h.status;
            if (h.status == 21) {
                // tau == 0: true residual check (BiCGStab2.inl:514-536)
                A.launch(sm, x, EpiInit<XT>{b, r, rh, u, v, w}, ws, st, TAIL_INIT, 0);
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
        if (conv <= tol_abs && k >= min_passes) break;   // reference: k > 1 (min_passes 2)
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
    launch_finish<1>(ctx->stream, ctx->ws, grid_for(n), W.st, TAIL_STORE, 0);
    DARE_LAUNCH_CHECK();
    KState h;
    rc = read_state(ctx, W.st, &h);
    if (rc) return rc;
    *out = h.dots[0];
    return DARE_OK;
}

static SolveOp<double> op64(const dare_csr* A, bool allow_dia) {
    SolveOp<double> op;
    op.n = A->nrows;
    op.csr = CsrOp<double>{A->row_ptr, A->col, A->val, A->nrows};
    op.dia = allow_dia && A->has_dia;
    if (op.dia) {
        op.d = make_dia<double>(A->dia_view, A->dia);
        op.dia = kXcd * op.d.per_xcd <= kReduceBlocks;   // else the CSR traversal
    }
    return op;
}

}  // namespace
}  // namespace dare_hip

using namespace dare_hip;

extern "C" {

int dare_spmv(dare_ctx* ctx, const dare_csr* A, const double* x, double* y) {
    DARE_CHECK_ARG(ctx && A && x && y, "NULL argument");
    int rc = csr_ensure_values(const_cast<dare_csr*>(A), ctx->stream);
    if (rc) return rc;
    // the Tpetra apply replacement: CSR, per-row sum in sorted column order
    SolveOp<double> op = op64(A, false);
    op.launch(ctx->stream, x, EpiStore<double>{y}, ctx->ws, nullptr, -1, 0);
    DARE_LAUNCH_CHECK();
    return DARE_OK;
}

int dare_spmv_dot(dare_ctx* ctx, const dare_csr* A, const double* x, double* y, const double* w,
                  double* out_host) {
    DARE_CHECK_ARG(ctx && A && x && y && w, "NULL argument");
    Workspace W;
    int rc = ws_reserve(ctx, 0, &W);
    if (rc) return rc;
    hipLaunchKernelGGL(k_clear_state, dim3(1), dim3(64), 0, ctx->stream, W.st);
    // the operator the solver uses (diagonal view when the matrix has one)
    SolveOp<double> op = op64(A, true);
    op.launch(ctx->stream, x, EpiDot<double>{y, w}, ctx->ws, W.st, TAIL_STORE, 1);
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
    const SolveOp<double> op = op64(A, true);
    if (!op.dia) {
        rc = csr_ensure_values(const_cast<dare_csr*>(A), ctx->stream);
        if (rc) return rc;
    }
    int iters = 0, status = 0, hl = 0;
    double rn = 0.;
    bool conv = false;
    rc = bicgstab2_run<double, double>(ctx, op, b, x, tol_abs, p->max_it, p->history, p->history_cap, &hl, &iters,
                                       &rn, &status, &conv, static_cast<double*>(W.buf), 2);
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
    if (A->has_dia && kXcd * make_dia<double>(A->dia_view, A->dia).per_xcd <= kReduceBlocks) {
        const int slots = A->dia_view.cross ? 8 : 7;
        const int64_t need = static_cast<int64_t>(slots) * n;
        if (A->cap_dia32 < need) {
            if (A->dia32) hipFree(A->dia32);
            A->dia32 = nullptr;
            A->cap_dia32 = 0;
            if (hipMalloc(&A->dia32, sizeof(float) * need) != hipSuccess) {
                hipGetLastError();
                set_error("FP32 diagonal allocation failed");
                return DARE_ERR_OOM;
            }
            A->cap_dia32 = need;
        }
        hipLaunchKernelGGL(k_to_f32, dim3(grid_for(need)), dim3(kThreads), 0, ctx->stream, (long long)need, A->dia,
                           A->dia32);
        op32.d = make_dia<float>(A->dia_view, A->dia32);
        op32.dia = kXcd * op32.d.per_xcd <= kReduceBlocks;
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
        hipLaunchKernelGGL(k_to_f32, dim3(grid_for(A->nnz)), dim3(kThreads), 0, ctx->stream, (long long)A->nnz, A->val,
                           A->val32);
        op32.dia = false;
    }
    op32.csr = CsrOp<float>{A->row_ptr, A->col, A->val32, n};
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
    const SolveOp<double> op = op64(A, true);
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
        op.launch(ctx->stream, x, EpiResid{b, r32}, ctx->ws, W.st, TAIL_STORE, 0);      // r = b - A x (FP64)
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
        rc = bicgstab2_run<float, float>(ctx, op32, r32, d32, inner_tol * rn, inner_max, nullptr, 0, &dummy_len, &it,
                                         &irn, &ist, &iconv, work, 2);
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
