// Krylov building blocks shared by the single-system solver (krylov.hip) and the batched
// solver (batch.hip): device-resident BiCGStab(2) scalars and their update rules, the
// block part of the two-kernel deterministic reductions, the operator traversals (CSR and the
// tiled diagonal view) with their fused epilogues, and the vector-update kernels.
#pragma once

#include <cmath>
#include <cstring>

#include <type_traits>

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
// block_partials_at writes dot j of this work-group to partial[j * stride + slot]; a work-group
// that calls it more than once (persistent loops) separates the calls with a barrier.
template <int NDOT>
__device__ inline void block_partials_at(double (&acc)[NDOT], double* partial, long long stride, long long slot) {
    __shared__ double red[NDOT][kThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NDOT; j++) {
        const double v = wave_sum(acc[j]);
        if (lane == 0) red[j][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x < NDOT) {     // dot-major: the finish kernel reads each dot's partials contiguously
        double s = 0.;
        for (int w = 0; w < kThreads / 64; w++) s += red[threadIdx.x][w];
        partial[threadIdx.x * stride + slot] = s;
    }
}

template <int NDOT>
__device__ inline void block_partials(double (&acc)[NDOT], ReduceWs ws) {
    block_partials_at<NDOT>(acc, ws.partial, kReduceBlocks, blockIdx.x);
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
    int pair;           // cells per lane along k (2: tiles of 128 (k) x 4 (j), one component, n2 even)
    bool sym = false;            // the view is exactly symmetric: the four-slot pair / quad traversals
};

template <typename VT>
DiaOp<VT> make_dia(const DiaView& V, const VT* D, bool allow_pair = false) {
    DiaOp<VT> A;
    A.V = V;
    A.D = D;
    A.pair = (allow_pair && V.N == 1 && V.n[2] % 2 == 0) ? 2 : 1;
    if (allow_pair && std::is_same_v<VT, float> && V.N == 1 && V.n[2] % 4 == 0) A.pair = 4;
    A.tiles_k = (V.n[2] + kTileK * A.pair - 1) / (kTileK * A.pair);
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
template <typename XT, bool NT = false>
// r = b - A x; rh = r; u = 0; <r,r>, <rh,r>   (BiCGStab2.inl:265-273). The reference also sets
// v = w = 0; here both are written (by the SpMVs of steps 6 and 13) before the first pass reads
// them, so those two zero sweeps are left out. v and w are kept in the member list for the
// batched solver's call sites.
struct EpiInit {
    static constexpr int NDOT = 2;
    static constexpr bool kNT = NT;
    const XT* b;
    XT *r, *rh, *u, *v, *w;
    __device__ double load(long long i) const { return static_cast<double>(b[i]); }
    __device__ void operator()(long long i, double y, double, double bi, double* acc) const {
        const XT ri = static_cast<XT>(bi - y);
        put<NT>(r + i, ri);
        put<NT>(rh + i, ri);
        put<NT>(u + i, XT(0));
        const double rd = static_cast<double>(ri);
        acc[0] += rd * rd;
        acc[1] += rd * rd;
    }
};

template <typename XT, bool NT = false>
struct EpiDot {         // out = A p; <out, w>                   (6)-(7), (9)+(11), (13)-(14)
    static constexpr int NDOT = 1;
    static constexpr bool kNT = NT;
    XT* out;
    const XT* w;
    __device__ double load(long long i) const { return static_cast<double>(w[i]); }
    __device__ void operator()(long long i, double y, double, double wi, double* acc) const {
        const XT yi = static_cast<XT>(y);
        put<NT>(out + i, yi);
        acc[0] += static_cast<double>(yi) * wi;
    }
};

template <typename XT, bool NT = false>
struct EpiGcr {         // t = A s; <r,s>, <s,s>, <s,t>, <t,t>, <r,t>   (18)-(20)
    static constexpr int NDOT = 5;
    static constexpr bool kNT = NT;
    XT* t;
    const XT* r;
    __device__ double load(long long i) const { return static_cast<double>(r[i]); }
    __device__ void operator()(long long i, double y, double sc, double rd, double* acc) const {
        const XT ti = static_cast<XT>(y);
        put<NT>(t + i, ti);
        const double td = static_cast<double>(ti);
        acc[0] += rd * sc;
        acc[1] += sc * sc;
        acc[2] += sc * td;
        acc[3] += td * td;
        acc[4] += rd * td;
    }
};

// the same step when a right preconditioner sits between s and the operator: t = A (M^-1 s),
// with s read from its own vector (the operand of the traversal is M^-1 s)
template <typename XT, bool NT = false>
struct EpiGcrP {
    static constexpr int NDOT = 5;
    static constexpr bool kNT = NT;
    XT* t;
    const XT* r;
    const XT* s;
    __device__ double load(long long i) const { return static_cast<double>(r[i]); }
    __device__ void operator()(long long i, double y, double, double rd, double* acc) const {
        const XT ti = static_cast<XT>(y);
        put<NT>(t + i, ti);
        const double td = static_cast<double>(ti);
        const double sc = static_cast<double>(s[i]);
        acc[0] += rd * sc;
        acc[1] += sc * sc;
        acc[2] += sc * td;
        acc[3] += td * td;
        acc[4] += rd * td;
    }
};

template <bool NT = false>
struct EpiResid {       // refinement: r = b - A x in FP64, r32 = float(r), <r,r>
    static constexpr int NDOT = 1;
    static constexpr bool kNT = NT;
    const double* b;
    float* r32;
    __device__ double load(long long i) const { return b[i]; }
    __device__ void operator()(long long i, double y, double, double bi, double* acc) const {
        const double ri = bi - y;
        put<NT>(r32 + i, static_cast<float>(ri));
        acc[0] += ri * ri;
    }
};

template <bool NT = false>
struct EpiResidB {      // the first refinement residual: as EpiResid, and <b,b> from the same read of b
    static constexpr int NDOT = 2;
    static constexpr bool kNT = NT;
    const double* b;
    float* r32;
    __device__ double load(long long i) const { return b[i]; }
    __device__ void operator()(long long i, double y, double, double bi, double* acc) const {
        const double ri = bi - y;
        put<NT>(r32 + i, static_cast<float>(ri));
        acc[0] += ri * ri;
        acc[1] += bi * bi;
    }
};

template <bool NT = false>
struct EpiResid64 {     // r = b - A x in FP64, <r,r>
    static constexpr int NDOT = 1;
    static constexpr bool kNT = NT;
    const double* b;
    double* r;
    __device__ double load(long long i) const { return b[i]; }
    __device__ void operator()(long long i, double y, double, double bi, double* acc) const {
        const double ri = bi - y;
        put<NT>(r + i, ri);
        acc[0] += ri * ri;
    }
};

template <typename XT, bool NT = false>
struct EpiStore {
    static constexpr int NDOT = 1;
    static constexpr bool kNT = NT;
    XT* out;
    __device__ double load(long long) const { return 0.; }
    __device__ void operator()(long long i, double y, double, double, double*) const {
        put<NT>(out + i, static_cast<XT>(y));
    }
};

// The epilogue of a lane's two adjacent rows i, i+1 (paired traversal): by default two calls,
// whose single stores land next to each other and merge into one wide store; EpiInit writes
// three vectors per row, so its pair form orders the stores vector by vector for the same
// effect. The dot accumulators see row i, then row i+1, as with two calls.
template <class Epi>
__device__ __forceinline__ void epi_pair(const Epi& e, long long i, double y0, double y1, double x0, double x1,
                                         double p0, double p1, double* acc) {
    e(i, y0, x0, p0, acc);
    e(i + 1, y1, x1, p1, acc);
}
template <typename XT, bool NT>
__device__ __forceinline__ void epi_pair(const EpiInit<XT, NT>& e, long long i, double y0, double y1, double,
                                         double, double b0, double b1, double* acc) {
    const XT r0 = static_cast<XT>(b0 - y0), r1 = static_cast<XT>(b1 - y1);
    put<NT>(e.r + i, r0);
    put<NT>(e.r + i + 1, r1);
    put<NT>(e.rh + i, r0);
    put<NT>(e.rh + i + 1, r1);
    put<NT>(e.u + i, XT(0));
    put<NT>(e.u + i + 1, XT(0));
    const double d0 = static_cast<double>(r0), d1 = static_cast<double>(r1);
    acc[0] += d0 * d0;
    acc[1] += d0 * d0;
    acc[0] += d1 * d1;
    acc[1] += d1 * d1;
}

// The epilogue of four adjacent rows (the FP32 quad traversal): four calls, or for EpiInit the
// stores ordered vector by vector, as epi_pair.
template <class Epi>
__device__ __forceinline__ void epi_quad(const Epi& e, long long i, const double* y, const double* xc,
                                         const double* pre, double* acc) {
#pragma unroll
    for (int m = 0; m < 4; m++) e(i + m, y[m], xc[m], pre[m], acc);
}
template <typename XT, bool NT>
__device__ __forceinline__ void epi_quad(const EpiInit<XT, NT>& e, long long i, const double* y, const double*,
                                         const double* b, double* acc) {
    XT r[4];
#pragma unroll
    for (int m = 0; m < 4; m++) r[m] = static_cast<XT>(b[m] - y[m]);
#pragma unroll
    for (int m = 0; m < 4; m++) put<NT>(e.r + i + m, r[m]);
#pragma unroll
    for (int m = 0; m < 4; m++) put<NT>(e.rh + i + m, r[m]);
#pragma unroll
    for (int m = 0; m < 4; m++) put<NT>(e.u + i + m, XT(0));
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const double dd = static_cast<double>(r[m]);
        acc[0] += dd * dd;
        acc[1] += dd * dd;
    }
}

// ------------------------------------------------------------------------------ traversals
// Generic CSR, one wave per 64 consecutive rows. The wave walks its rows' run of the CSR
// arrays -- one contiguous range -- with coalesced loads (lane l takes entries l, l + 64, ...),
// forms each product val * x[col] (rounded to double, as the row loop rounds it) and stages it
// in its own slice of LDS; then lane l sums row r0 + l from LDS in column order, starting from
// +0.0: the sequential row sum of the reference's row loop, so y is bit-identical to it. A wave
// whose rows hold more than kCsrWaveNnz entries sums them from global memory instead. Wave
// tiles are dealt so that each XCD (blocks b = x mod 8) owns one contiguous eighth of the rows.
// Only the wave that wrote a slice reads it, so the phases need no work-group barrier.
constexpr int kCsrWaveRows = 64;
constexpr int kCsrWaveNnz = 8 * 64;   // eight entries per lane
constexpr int kCsrWaves = kThreads / kCsrWaveRows;

__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// row bounds of wave tile t: the run [B, E) of the CSR arrays and this lane's row [b, e)
struct CsrTileRp {
    int B, E, b, e;
};
template <typename VT>
__device__ inline CsrTileRp csr_tile_rp(const CsrOp<VT>& A, long long t, int lane) {
    const long long r0 = t * kCsrWaveRows;
    const long long r1 = (r0 + kCsrWaveRows < A.n) ? r0 + kCsrWaveRows : A.n;
    const long long i = r0 + lane;
    CsrTileRp R;
    R.B = A.rp[r0];
    R.E = A.rp[r1];
    R.b = i < r1 ? A.rp[i] : R.E;
    R.e = i < r1 ? A.rp[i + 1] : R.E;
    return R;
}

// columns and values of a staged wave tile: eight entries per lane, entries past the tile's run
// repeating its last one; streamed with non-temporal loads so that x keeps the caches
template <typename VT>
__device__ inline void csr_tile_cv(const CsrOp<VT>& A, const CsrTileRp& R, int lane, int* c, double* v) {
    const int nnz = R.E - R.B;
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const int k = u * kCsrWaveRows + lane;
        const int kk = R.B + (k < nnz ? k : nnz - 1);
        c[u] = __builtin_nontemporal_load(A.col + kk);
        v[u] = static_cast<double>(__builtin_nontemporal_load(A.val + kk));
    }
}

template <typename VT, typename XT, class Epi>
__global__ __launch_bounds__(kThreads) void k_apply_csr(CsrOp<VT> A, const XT* __restrict__ x, Epi epi, ReduceWs ws,
                                                       KState* st, int tail, int check_status) {
    __shared__ double s_prod[kCsrWaves][kCsrWaveNnz];
    double acc[Epi::NDOT];
#pragma unroll
    for (int j = 0; j < Epi::NDOT; j++) acc[j] = 0.;
    const bool run = !(check_status && st->status);
    if (run) {
        const int lane = threadIdx.x % kCsrWaveRows, wv = threadIdx.x / kCsrWaveRows;
        double* prod = s_prod[wv];
        const long long ntiles = (A.n + kCsrWaveRows - 1) / kCsrWaveRows;
        const int xcd = static_cast<int>(blockIdx.x % kXcd);
        const long long t0 = ntiles * xcd / kXcd, t1 = ntiles * (xcd + 1) / kXcd;
        const long long stride = static_cast<long long>(gridDim.x / kXcd) * kCsrWaves;
        long long tile = t0 + static_cast<long long>(blockIdx.x / kXcd) * kCsrWaves + wv;
        // software pipeline: the next tile's columns and values, and the row bounds of the one
        // after it, are loaded while this tile's x entries are gathered and summed
        CsrTileRp R0{}, R1{};
        int c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        double v[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
        if (tile < t1) {
            R0 = csr_tile_rp(A, tile, lane);
            if (R0.E - R0.B > 0 && R0.E - R0.B <= kCsrWaveNnz) csr_tile_cv(A, R0, lane, c, v);
        }
        if (tile + stride < t1) R1 = csr_tile_rp(A, tile + stride, lane);
        for (; tile < t1; tile += stride) {
            const long long i = tile * kCsrWaveRows + lane;
            const bool own = i < A.n;
            const long long ic = own ? i : A.n - 1;      // a valid address for every lane
            const double xi = static_cast<double>(x[ic]);
            const double pre = epi.load(ic);
            const int nnz = R0.E - R0.B, b = R0.b - R0.B, e = R0.e - R0.B;
            const bool staged = nnz <= kCsrWaveNnz;
            int cn[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            double vn[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
            CsrTileRp R2{};
            if (tile + stride < t1 && R1.E - R1.B > 0 && R1.E - R1.B <= kCsrWaveNnz) csr_tile_cv(A, R1, lane, cn, vn);
            if (tile + 2 * stride < t1) R2 = csr_tile_rp(A, tile + 2 * stride, lane);
            double s = 0.;
            if (staged) {
                // every slot of the slice is written (entries past nnz repeat the last one and
                // are never read), so no lane branches
                if (nnz > 0) {
                    double xv[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) xv[u] = static_cast<double>(x[c[u]]);
#pragma unroll
                    for (int u = 0; u < 8; u++) prod[u * kCsrWaveRows + lane] = v[u] * xv[u];
                }
                wave_lds_sync();
                for (int k = b; k < e; k += 8) {
                    double pv[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) pv[u] = prod[(k + u < e) ? k + u : b];
#pragma unroll
                    for (int u = 0; u < 8; u++)
                        if (k + u < e) s += pv[u];
                }
                wave_lds_sync();   // the slice is rewritten by the wave's next tile
            } else {
                for (int k = R0.B + b; k < R0.B + e; k++)
                    s += static_cast<double>(A.val[k]) * static_cast<double>(x[A.col[k]]);
            }
            if (own) epi(i, s, xi, pre, acc);
            R0 = R1;
            R1 = R2;
#pragma unroll
            for (int u = 0; u < 8; u++) {
                c[u] = cn[u];
                v[u] = vn[u];
            }
        }
    }
    if (tail >= 0) grid_reduce<Epi::NDOT>(acc, ws, st, run ? tail : TAIL_NONE);
}

// y_row = sum over the row's slots in the order W,S,B,[X],C,[X],T,N,E (the sorted column
// order of a non-wrapped row; X is the cross entry, before C for component 1, after it for 0).
// Every load of a plane is issued before the multiply-adds; entries removed by a boundary
// edit read a valid in-range address and contribute +0.0 (the partial sum starts at +0.0 and
// so is never -0.0: adding +0.0 leaves it bit-identical to skipping the term).
//
// dia_column runs one (j, k) column of a tile over the planes i0 <= i < i1 with the i-1/i/i+1
// values of x rolled through registers. x and D belong to one system; row0 is that system's
// first row in the epilogue's vectors (0 for a single system).
template <int NC, typename VT, typename XT, class Epi>
__device__ inline void dia_column(const DiaView& V, const VT* __restrict__ D, const XT* __restrict__ x,
                                  long long row0, int k, int j, int i0, int i1, const Epi& epi, double* acc) {
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
        if constexpr (NC == 2) {
            // the two component rows of a cell are adjacent (row = 2 node + c, nrows even, the
            // arrays start on 2-element boundaries): one 2-wide load per slot and per x entry
            typedef VT v2 __attribute__((ext_vector_type(2)));
            typedef XT x2 __attribute__((ext_vector_type(2)));
            auto ldx = [&](long long nd, double* o) {
                const x2 v = *reinterpret_cast<const x2*>(x + nd * 2);
                o[0] = static_cast<double>(v.x);
                o[1] = static_cast<double>(v.y);
            };
            ldx(nnext, xn);
            ldx(node + dS, xs);
            ldx(node + dN, xnn);
            ldx(node + dB, xb);
            ldx(node + dT, xt);
            const int nsl = V.cross ? 8 : 7;
#pragma unroll
            for (int sl = 0; sl < 8; sl++) {
                v2 d = v2{VT(0), VT(0)};
                if (sl < nsl) {
                    const v2* p = reinterpret_cast<const v2*>(D + sl * nr + node * 2);
                    if constexpr (Epi::kNT) d = __builtin_nontemporal_load(p);
                    else d = *p;
                }
                dv[0][sl] = static_cast<double>(d.x);
                dv[1][sl] = static_cast<double>(d.y);
            }
#pragma unroll
            for (int c = 0; c < NC; c++) pre[c] = epi.load(row0 + node * NC + c);
        } else {
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const long long row = node * NC + c;
                xn[c] = static_cast<double>(x[nnext * NC + c]);
                xs[c] = static_cast<double>(x[(node + dS) * NC + c]);
                xnn[c] = static_cast<double>(x[(node + dN) * NC + c]);
                xb[c] = static_cast<double>(x[(node + dB) * NC + c]);
                xt[c] = static_cast<double>(x[(node + dT) * NC + c]);
#pragma unroll
                for (int sl = 0; sl < 7; sl++) {
                    if constexpr (Epi::kNT) dv[c][sl] = static_cast<double>(ld_nt(D + sl * nr + row));
                    else dv[c][sl] = static_cast<double>(D[sl * nr + row]);
                }
                dv[c][7] = 0.;
                pre[c] = epi.load(row0 + row);
            }
        }
        // straight-line per component (no loop around the epilogue's stores: LLVM's loop
        // invariant code motion promotes a store out of a one-trip loop and drops its
        // non-temporal hint on the way)
        auto row_out = [&](const int c) {
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
            epi(row0 + node * NC + c, s, xc[c], pre[c], acc);
        };
        row_out(0);
        if constexpr (NC == 2) row_out(1);
#pragma unroll
        for (int c = 0; c < NC; c++) {
            xp[c] = xc[c];
            xc[c] = xn[c];
        }
        node += sI;   // a tile never runs past plane n0 - 1, so xc is always a real plane
    }
}

// dia_column_pair: one component, the two cells k (even) and k+1 of a lane. Their slot values,
// x entries along i and j and epilogue operands are adjacent in memory, so each pair is one
// access twice as wide (a wave moves twice the bytes per memory instruction). Within the pair
// the top neighbour of cell k and the bottom neighbour of cell k+1 are the pair itself; only
// k-1 (cell k) and k+2 (cell k+1) are single loads. Each cell's sum keeps the order of
// dia_column, so y is bit-identical; the per-lane dot accumulation order differs.
template <typename VT, typename XT, class Epi>
__device__ inline void dia_column_pair(const DiaView& V, const VT* __restrict__ D, const XT* __restrict__ x,
                                       int k, int j, int i0, int i1, const Epi& epi, double* acc) {
    const int n0 = V.n[0], n1 = V.n[1], n2 = V.n[2];
    const long long sI = static_cast<long long>(n1) * n2, sJ = n2;
    const long long nr = V.nrows;
    const bool hasS = !(j == 0 && V.rm_lo[1]);
    const bool hasN = !(j == n1 - 1 && V.rm_hi[1] && !(j == 0 && V.rm_lo[1]));
    const bool hasB = !(k == 0 && V.rm_lo[2]);               // cell k   (its top is cell k+1)
    const bool hasT = !(k + 1 == n2 - 1 && V.rm_hi[2]);      // cell k+1 (its bottom is cell k; n2 >= 2)
    const long long dS = !hasS ? 0 : (j == 0) ? (long long)(n1 - 1) * sJ : -sJ;
    const long long dN = !hasN ? 0 : (j == n1 - 1) ? -(long long)(n1 - 1) * sJ : sJ;
    const long long dB = !hasB ? 0 : (k == 0) ? (long long)(n2 - 1) : -1;            // from cell k
    const long long dT = !hasT ? 0 : (k + 1 == n2 - 1) ? -(long long)(n2 - 1) : 1;   // from cell k+1
    long long node = (static_cast<long long>(i0) * n1 + j) * n2 + k;
    double xp0, xp1, xc0, xc1;
    {
        const bool need = !(i0 == 0 && V.rm_lo[0]);
        const long long nprev = !need ? node : (i0 == 0) ? node + (long long)(n0 - 1) * sI : node - sI;
        xc0 = static_cast<double>(x[node]);
        xc1 = static_cast<double>(x[node + 1]);
        xp0 = static_cast<double>(x[nprev]);
        xp1 = static_cast<double>(x[nprev + 1]);
    }
    for (int i = i0; i < i1; i++) {
        const bool hasW = !(i == 0 && V.rm_lo[0]);
        const bool hasE = !(i == n0 - 1 && V.rm_hi[0] && !(i == 0 && V.rm_lo[0]));
        const long long nnext = !hasE ? node : (i == n0 - 1) ? node - (long long)(n0 - 1) * sI : node + sI;
        const double xn0 = static_cast<double>(x[nnext]), xn1 = static_cast<double>(x[nnext + 1]);
        const double xs0 = static_cast<double>(x[node + dS]), xs1 = static_cast<double>(x[node + dS + 1]);
        const double xq0 = static_cast<double>(x[node + dN]), xq1 = static_cast<double>(x[node + dN + 1]);
        const double xb = static_cast<double>(x[node + dB]);
        const double xt = static_cast<double>(x[node + 1 + dT]);
        double d0[7], d1[7];
#pragma unroll
        for (int sl = 0; sl < 7; sl++) {
            if constexpr (Epi::kNT) {
                d0[sl] = static_cast<double>(ld_nt(D + sl * nr + node));
                d1[sl] = static_cast<double>(ld_nt(D + sl * nr + node + 1));
            } else {
                d0[sl] = static_cast<double>(D[sl * nr + node]);
                d1[sl] = static_cast<double>(D[sl * nr + node + 1]);
            }
        }
        const double pre0 = epi.load(node), pre1 = epi.load(node + 1);
        double s0 = 0.;
        s0 += hasW ? d0[0] * xp0 : 0.;
        s0 += hasS ? d0[1] * xs0 : 0.;
        s0 += hasB ? d0[2] * xb : 0.;
        s0 += d0[3] * xc0;
        s0 += d0[4] * xc1;
        s0 += hasN ? d0[5] * xq0 : 0.;
        s0 += hasE ? d0[6] * xn0 : 0.;
        double s1 = 0.;
        s1 += hasW ? d1[0] * xp1 : 0.;
        s1 += hasS ? d1[1] * xs1 : 0.;
        s1 += d1[2] * xc0;
        s1 += d1[3] * xc1;
        s1 += hasT ? d1[4] * xt : 0.;
        s1 += hasN ? d1[5] * xq1 : 0.;
        s1 += hasE ? d1[6] * xn1 : 0.;
        epi_pair(epi, node, s0, s1, xc0, xc1, pre0, pre1, acc);
        xp0 = xc0;
        xp1 = xc1;
        xc0 = xn0;
        xc1 = xn1;
        node += sI;
    }
}

// dia_column_pair for an exactly symmetric view (symmetric by construction, or found so by
// k_dia_asym): the E, N and T coefficients of a row are read as the W, S and B coefficients
// of its east, north and top neighbours, so four slots stream from HBM instead of seven. The
// partner values are the same numbers and each sum keeps dia_column_pair's order, so y is
// bit-identical. The east neighbour's W is the next plane's own W: it is loaded once and kept
// in registers. S and B are read again as partners of the rows south and below (from the
// caches), so they carry no non-temporal hint; C and the east W do.
template <typename VT, typename XT, class Epi>
__device__ inline void dia_column_pair_sym(const DiaView& V, const VT* __restrict__ D, const XT* __restrict__ x,
                                           int k, int j, int i0, int i1, const Epi& epi, double* acc) {
    const int n0 = V.n[0], n1 = V.n[1], n2 = V.n[2];
    const long long sI = static_cast<long long>(n1) * n2, sJ = n2;
    const long long nr = V.nrows;
    const bool hasS = !(j == 0 && V.rm_lo[1]);
    const bool hasN = !(j == n1 - 1 && V.rm_hi[1] && !(j == 0 && V.rm_lo[1]));
    const bool hasB = !(k == 0 && V.rm_lo[2]);
    const bool hasT = !(k + 1 == n2 - 1 && V.rm_hi[2]);
    const long long dS = !hasS ? 0 : (j == 0) ? (long long)(n1 - 1) * sJ : -sJ;
    const long long dN = !hasN ? 0 : (j == n1 - 1) ? -(long long)(n1 - 1) * sJ : sJ;
    const long long dB = !hasB ? 0 : (k == 0) ? (long long)(n2 - 1) : -1;
    const long long dT = !hasT ? 0 : (k + 1 == n2 - 1) ? -(long long)(n2 - 1) : 1;
    const VT* DW = D;
    const VT* DS = D + nr;
    const VT* DB = D + 2 * nr;
    const VT* DC = D + 3 * nr;
    auto ldnt = [](const VT* p) {
        if constexpr (Epi::kNT) return static_cast<double>(ld_nt(p));
        else return static_cast<double>(*p);
    };
    long long node = (static_cast<long long>(i0) * n1 + j) * n2 + k;
    double xp0, xp1, xc0, xc1;
    {
        const bool need = !(i0 == 0 && V.rm_lo[0]);
        const long long nprev = !need ? node : (i0 == 0) ? node + (long long)(n0 - 1) * sI : node - sI;
        xc0 = static_cast<double>(x[node]);
        xc1 = static_cast<double>(x[node + 1]);
        xp0 = static_cast<double>(x[nprev]);
        xp1 = static_cast<double>(x[nprev + 1]);
    }
    double w0 = ldnt(DW + node), w1 = ldnt(DW + node + 1);
    for (int i = i0; i < i1; i++) {
        const bool hasW = !(i == 0 && V.rm_lo[0]);
        const bool hasE = !(i == n0 - 1 && V.rm_hi[0] && !(i == 0 && V.rm_lo[0]));
        // a following plane of this tile is always node + sI (i < n0 - 1 there), so the east
        // W loaded here is that plane's own W
        const long long nnext = !hasE ? node : (i == n0 - 1) ? node - (long long)(n0 - 1) * sI : node + sI;
        const double xn0 = static_cast<double>(x[nnext]), xn1 = static_cast<double>(x[nnext + 1]);
        const double xs0 = static_cast<double>(x[node + dS]), xs1 = static_cast<double>(x[node + dS + 1]);
        const double xq0 = static_cast<double>(x[node + dN]), xq1 = static_cast<double>(x[node + dN + 1]);
        const double xb = static_cast<double>(x[node + dB]);
        const double xt = static_cast<double>(x[node + 1 + dT]);
        const double sv0 = static_cast<double>(DS[node]), sv1 = static_cast<double>(DS[node + 1]);
        const double bv0 = static_cast<double>(DB[node]), bv1 = static_cast<double>(DB[node + 1]);
        const double cv0 = ldnt(DC + node), cv1 = ldnt(DC + node + 1);
        const double ev0 = ldnt(DW + nnext), ev1 = ldnt(DW + nnext + 1);        // W of the east neighbour
        const double nv0 = static_cast<double>(DS[node + dN]), nv1 = static_cast<double>(DS[node + dN + 1]);
        const double tv1 = static_cast<double>(DB[node + 1 + dT]);              // B of cell k+1's top neighbour
        const double pre0 = epi.load(node), pre1 = epi.load(node + 1);
        double s0 = 0.;
        s0 += hasW ? w0 * xp0 : 0.;
        s0 += hasS ? sv0 * xs0 : 0.;
        s0 += hasB ? bv0 * xb : 0.;
        s0 += cv0 * xc0;
        s0 += bv1 * xc1;              // T of cell k = B of cell k+1
        s0 += hasN ? nv0 * xq0 : 0.;
        s0 += hasE ? ev0 * xn0 : 0.;
        double s1 = 0.;
        s1 += hasW ? w1 * xp1 : 0.;
        s1 += hasS ? sv1 * xs1 : 0.;
        s1 += bv1 * xc0;
        s1 += cv1 * xc1;
        s1 += hasT ? tv1 * xt : 0.;
        s1 += hasN ? nv1 * xq1 : 0.;
        s1 += hasE ? ev1 * xn1 : 0.;
        epi_pair(epi, node, s0, s1, xc0, xc1, pre0, pre1, acc);
        xp0 = xc0;
        xp1 = xc1;
        xc0 = xn0;
        xc1 = xn1;
        w0 = ev0;
        w1 = ev1;
        node += sI;
    }
}

// dia_column_quad: the FP32 operator of the inner solves (one component, FP32 slots and
// vectors), four k-adjacent rows per lane, so every slot, x and vector access is one 16-byte
// float4 and a wave moves 1 KB per memory instruction. The products and sums are FP32 (the
// precision of the inner iterations; the slot order of each row is dia_column's), the dot
// accumulators stay FP64. Rows k..k+3 with k a multiple of 4 and n2 % 4 == 0.
template <class Epi>
__device__ inline void dia_column_quad(const DiaView& V, const float* __restrict__ D, const float* __restrict__ x,
                                       int k, int j, int i0, int i1, const Epi& epi, double* acc) {
    const int n0 = V.n[0], n1 = V.n[1], n2 = V.n[2];
    const long long sI = static_cast<long long>(n1) * n2, sJ = n2;
    const long long nr = V.nrows;
    const bool hasS = !(j == 0 && V.rm_lo[1]);
    const bool hasN = !(j == n1 - 1 && V.rm_hi[1] && !(j == 0 && V.rm_lo[1]));
    const bool hasB = !(k == 0 && V.rm_lo[2]);               // cell k   (its top is cell k+1)
    const bool hasT = !(k + 3 == n2 - 1 && V.rm_hi[2]);      // cell k+3 (n2 >= 4)
    const long long dS = !hasS ? 0 : (j == 0) ? (long long)(n1 - 1) * sJ : -sJ;
    const long long dN = !hasN ? 0 : (j == n1 - 1) ? -(long long)(n1 - 1) * sJ : sJ;
    const long long dB = !hasB ? 0 : (k == 0) ? (long long)(n2 - 1) : -1;            // from cell k
    const long long dT = !hasT ? 0 : (k + 3 == n2 - 1) ? -(long long)(n2 - 1) : 1;   // from cell k+3
    typedef float f4 __attribute__((ext_vector_type(4)));
    auto ld4 = [](const float* p) { return *reinterpret_cast<const f4*>(p); };
    auto ld4nt = [](const float* p) {
        if constexpr (Epi::kNT) return __builtin_nontemporal_load(reinterpret_cast<const f4*>(p));
        else return *reinterpret_cast<const f4*>(p);
    };
    long long node = (static_cast<long long>(i0) * n1 + j) * n2 + k;
    f4 xp, xc;
    {
        const bool need = !(i0 == 0 && V.rm_lo[0]);
        const long long nprev = !need ? node : (i0 == 0) ? node + (long long)(n0 - 1) * sI : node - sI;
        xc = ld4(x + node);
        xp = ld4(x + nprev);
    }
    for (int i = i0; i < i1; i++) {
        const bool hasW = !(i == 0 && V.rm_lo[0]);
        const bool hasE = !(i == n0 - 1 && V.rm_hi[0] && !(i == 0 && V.rm_lo[0]));
        const long long nnext = !hasE ? node : (i == n0 - 1) ? node - (long long)(n0 - 1) * sI : node + sI;
        const f4 xn = ld4(x + nnext), xs = ld4(x + node + dS), xq = ld4(x + node + dN);
        const float xb = x[node + dB], xt = x[node + 3 + dT];
        f4 d[7];
#pragma unroll
        for (int sl = 0; sl < 7; sl++) d[sl] = ld4nt(D + sl * nr + node);
        const double p0 = epi.load(node), p1 = epi.load(node + 1), p2 = epi.load(node + 2), p3 = epi.load(node + 3);
        const float xcv[4] = {xc.x, xc.y, xc.z, xc.w}, xpv[4] = {xp.x, xp.y, xp.z, xp.w};
        const float xsv[4] = {xs.x, xs.y, xs.z, xs.w}, xqv[4] = {xq.x, xq.y, xq.z, xq.w};
        const float xnv[4] = {xn.x, xn.y, xn.z, xn.w};
        double y[4];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            auto dv = [&](int sl) { return m == 0 ? d[sl].x : m == 1 ? d[sl].y : m == 2 ? d[sl].z : d[sl].w; };
            float s = 0.f;
            s += hasW ? dv(0) * xpv[m] : 0.f;
            s += hasS ? dv(1) * xsv[m] : 0.f;
            if (m == 0) s += hasB ? dv(2) * xb : 0.f;
            else s += dv(2) * xcv[m - 1];
            s += dv(3) * xcv[m];
            if (m == 3) s += hasT ? dv(4) * xt : 0.f;
            else s += dv(4) * xcv[m + 1];
            s += hasN ? dv(5) * xqv[m] : 0.f;
            s += hasE ? dv(6) * xnv[m] : 0.f;
            y[m] = static_cast<double>(s);
        }
        const double xcd[4] = {xcv[0], xcv[1], xcv[2], xcv[3]}, pre[4] = {p0, p1, p2, p3};
        epi_quad(epi, node, y, xcd, pre, acc);
        xp = xc;
        xc = xn;
        node += sI;
    }
}

// dia_column_quad for an exactly symmetric FP32 view (the FP32 copy of a symmetric FP64 view): the E, N and T
// coefficients of a row are read as the W, S and B coefficients of its east, north and top
// neighbours, so four slots stream from HBM instead of seven. The partner values are the same
// floats and the arithmetic is that of dia_column_quad, so the results are bit-identical. The
// W, S and B slots are read twice (as a row's own and as its neighbour's partner), the second
// time from the caches, so they carry no non-temporal hint.
template <class Epi>
__device__ inline void dia_column_quad_sym(const DiaView& V, const float* __restrict__ D, const float* __restrict__ x,
                                           int k, int j, int i0, int i1, const Epi& epi, double* acc) {
    const int n0 = V.n[0], n1 = V.n[1], n2 = V.n[2];
    const long long sI = static_cast<long long>(n1) * n2, sJ = n2;
    const long long nr = V.nrows;
    const bool hasS = !(j == 0 && V.rm_lo[1]);
    const bool hasN = !(j == n1 - 1 && V.rm_hi[1] && !(j == 0 && V.rm_lo[1]));
    const bool hasB = !(k == 0 && V.rm_lo[2]);
    const bool hasT = !(k + 3 == n2 - 1 && V.rm_hi[2]);
    const long long dS = !hasS ? 0 : (j == 0) ? (long long)(n1 - 1) * sJ : -sJ;
    const long long dN = !hasN ? 0 : (j == n1 - 1) ? -(long long)(n1 - 1) * sJ : sJ;
    const long long dB = !hasB ? 0 : (k == 0) ? (long long)(n2 - 1) : -1;
    const long long dT = !hasT ? 0 : (k + 3 == n2 - 1) ? -(long long)(n2 - 1) : 1;
    typedef float f4 __attribute__((ext_vector_type(4)));
    auto ld4 = [](const float* p) { return *reinterpret_cast<const f4*>(p); };
    auto ld4nt = [](const float* p) {
        if constexpr (Epi::kNT) return __builtin_nontemporal_load(reinterpret_cast<const f4*>(p));
        else return *reinterpret_cast<const f4*>(p);
    };
    const float* DW = D;
    const float* DS = D + nr;
    const float* DB = D + 2 * nr;
    const float* DC = D + 3 * nr;
    // everything a plane reads, loaded one plane ahead (the next plane's loads are in flight
    // while this plane's sums are formed)
    struct Plane {
        f4 xn, xs, xq, cW, cS, cB, cC, cN, cE;
        float xb, xt, cT3;
        double pre[4];
        bool hasW, hasE;
    };
    auto load = [&](int i, long long node, Plane& P) {
        P.hasW = !(i == 0 && V.rm_lo[0]);
        P.hasE = !(i == n0 - 1 && V.rm_hi[0] && !(i == 0 && V.rm_lo[0]));
        const long long nnext = !P.hasE ? node : (i == n0 - 1) ? node - (long long)(n0 - 1) * sI : node + sI;
        P.xn = ld4(x + nnext);
        P.xs = ld4(x + node + dS);
        P.xq = ld4(x + node + dN);
        P.xb = x[node + dB];
        P.xt = x[node + 3 + dT];
        P.cW = ld4(DW + node);
        P.cS = ld4(DS + node);
        P.cB = ld4(DB + node);
        P.cC = ld4nt(DC + node);
        P.cN = ld4(DS + node + dN);          // S of the north neighbour (unused without it)
        P.cE = ld4(DW + nnext);              // W of the east neighbour
        P.cT3 = DB[node + 3 + dT];           // B of the top neighbour of cell k + 3
#pragma unroll
        for (int m = 0; m < 4; m++) P.pre[m] = epi.load(node + m);
    };
    long long node = (static_cast<long long>(i0) * n1 + j) * n2 + k;
    f4 xp, xc;
    {
        const bool need = !(i0 == 0 && V.rm_lo[0]);
        const long long nprev = !need ? node : (i0 == 0) ? node + (long long)(n0 - 1) * sI : node - sI;
        xc = ld4(x + node);
        xp = ld4(x + nprev);
    }
    Plane cur;
    load(i0, node, cur);
    for (int i = i0; i < i1; i++) {
        Plane nxt;
        const int inext = i + 1 < i1 ? i + 1 : i;            // the last plane reloads itself
        load(inext, node + (i + 1 < i1 ? sI : 0), nxt);
        const float xcv[4] = {xc.x, xc.y, xc.z, xc.w}, xpv[4] = {xp.x, xp.y, xp.z, xp.w};
        const float xsv[4] = {cur.xs.x, cur.xs.y, cur.xs.z, cur.xs.w};
        const float xqv[4] = {cur.xq.x, cur.xq.y, cur.xq.z, cur.xq.w};
        const float xnv[4] = {cur.xn.x, cur.xn.y, cur.xn.z, cur.xn.w};
        const float wv[4] = {cur.cW.x, cur.cW.y, cur.cW.z, cur.cW.w}, sv[4] = {cur.cS.x, cur.cS.y, cur.cS.z, cur.cS.w};
        const float bv[4] = {cur.cB.x, cur.cB.y, cur.cB.z, cur.cB.w}, cv[4] = {cur.cC.x, cur.cC.y, cur.cC.z, cur.cC.w};
        const float tv[4] = {cur.cB.y, cur.cB.z, cur.cB.w, cur.cT3};
        const float nv[4] = {cur.cN.x, cur.cN.y, cur.cN.z, cur.cN.w}, ev[4] = {cur.cE.x, cur.cE.y, cur.cE.z, cur.cE.w};
        double y[4];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            float s = 0.f;
            s += cur.hasW ? wv[m] * xpv[m] : 0.f;
            s += hasS ? sv[m] * xsv[m] : 0.f;
            if (m == 0) s += hasB ? bv[m] * cur.xb : 0.f;
            else s += bv[m] * xcv[m - 1];
            s += cv[m] * xcv[m];
            if (m == 3) s += hasT ? tv[m] * cur.xt : 0.f;
            else s += tv[m] * xcv[m + 1];
            s += hasN ? nv[m] * xqv[m] : 0.f;
            s += cur.hasE ? ev[m] * xnv[m] : 0.f;
            y[m] = static_cast<double>(s);
        }
        const double xcd[4] = {xcv[0], xcv[1], xcv[2], xcv[3]};
        epi_quad(epi, node, y, xcd, cur.pre, acc);
        xp = xc;
        xc = cur.xn;
        cur = nxt;
        node += sI;
    }
}

// Symmetry check of a one-component view (FP32 or FP64 slots): flag = 1 if the E, N or T
// coefficient of any row differs (bit for bit) from the W, S or B coefficient of the neighbour
// it couples to, with the neighbour and presence rules of dia_column. flag must be 0 before the
// launch.
template <typename VT>
__global__ __launch_bounds__(kThreads) void k_dia_asym(DiaView V, const VT* __restrict__ D, int* __restrict__ flag) {
    const long long nr = V.nrows;
    const int n0 = V.n[0], n1 = V.n[1], n2 = V.n[2];
    const long long sI = static_cast<long long>(n1) * n2, sJ = n2;
    auto bits = [](VT v) {
        if constexpr (std::is_same_v<VT, float>) return __float_as_uint(v);
        else return static_cast<unsigned long long>(__double_as_longlong(v));
    };
    bool bad = false;
    for (long long row = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x; row < nr;
         row += static_cast<long long>(gridDim.x) * kThreads) {
        const int k = static_cast<int>(row % n2), j = static_cast<int>((row / n2) % n1), i = static_cast<int>(row / sI);
        if (!(i == n0 - 1 && V.rm_hi[0] && !(i == 0 && V.rm_lo[0]))) {
            const long long nb = (i == n0 - 1) ? row - (long long)(n0 - 1) * sI : row + sI;
            bad |= bits(D[6 * nr + row]) != bits(D[nb]);
        }
        if (!(j == n1 - 1 && V.rm_hi[1] && !(j == 0 && V.rm_lo[1]))) {
            const long long nb = (j == n1 - 1) ? row - (long long)(n1 - 1) * sJ : row + sJ;
            bad |= bits(D[5 * nr + row]) != bits(D[nr + nb]);
        }
        if (!(k == n2 - 1 && V.rm_hi[2])) {
            const long long nb = (k == n2 - 1) ? row - (n2 - 1) : row + 1;
            bad |= bits(D[4 * nr + row]) != bits(D[2 * nr + nb]);
        }
    }
    if (bad) flag[0] = 1;
}

__global__ void k_set_int(int* __restrict__ p, int v) { p[0] = v; }

// PAIR: cells per lane along k (1, 2, or 4 for FP32); SYM (pairs and quads): the symmetric
// traversal, a separate instantiation so that the seven-slot one keeps its own register count
template <int NC, typename VT, typename XT, class Epi, int PAIR = 1, bool SYM = false>
__global__ __launch_bounds__(kThreads) void k_apply_dia(DiaOp<VT> A, const XT* __restrict__ x, Epi epi, ReduceWs ws,
                                                       KState* st, int tail, int check_status) {
    double acc[Epi::NDOT];
#pragma unroll
    for (int j = 0; j < Epi::NDOT; j++) acc[j] = 0.;
    const bool run = !(check_status && st->status);
    const int tile = static_cast<int>(blockIdx.x % kXcd) * A.per_xcd + static_cast<int>(blockIdx.x / kXcd);
    const int ntiles = A.tiles_k * A.tiles_j * A.tiles_i;
    if (run && tile < ntiles) {
        const int tk = tile % A.tiles_k;
        const int t2 = tile / A.tiles_k;
        const int tj = t2 % A.tiles_j;
        const int ti = t2 / A.tiles_j;
        const int k = (tk * kTileK + (threadIdx.x & 63)) * PAIR;
        const int j = tj * kTileJ + (threadIdx.x >> 6);
        if (k < A.V.n[2] && j < A.V.n[1]) {
            const int i0 = ti * A.tile_i;
            const int i1 = (i0 + A.tile_i < A.V.n[0]) ? i0 + A.tile_i : A.V.n[0];
            if constexpr (PAIR == 4) {
                static_assert(std::is_same_v<VT, float> && std::is_same_v<XT, float>, "quad traversal is FP32");
                if constexpr (SYM) dia_column_quad_sym<Epi>(A.V, A.D, x, k, j, i0, i1, epi, acc);
                else dia_column_quad<Epi>(A.V, A.D, x, k, j, i0, i1, epi, acc);
            } else if constexpr (PAIR == 2 && SYM) {
                dia_column_pair_sym<VT, XT, Epi>(A.V, A.D, x, k, j, i0, i1, epi, acc);
            } else if constexpr (PAIR == 2) {
                dia_column_pair<VT, XT, Epi>(A.V, A.D, x, k, j, i0, i1, epi, acc);
            } else {
                dia_column<NC, VT, XT, Epi>(A.V, A.D, x, 0, k, j, i0, i1, epi, acc);
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

// FIRST: the first pass from a zero guess (k_init_zero), where r = b and u = 0 were never
// written: r is b itself, u reads as 0 (the same arithmetic, r - beta * 0).
template <typename XT, bool NT, bool FIRST = false>
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
            uv[q] = FIRST ? 0. : static_cast<double>(u[i]);
        }
#pragma unroll
        for (int q = 0; q < kU; q++)
            if (DARE_IDX(q) < n) put<NT>(u + DARE_IDX(q), static_cast<XT>(rv[q] - beta * uv[q]));   // (5)
    }
}

// FIRST: the first pass from a zero guess: the old r is b (read from b0) and the old x is 0,
// both written here for the first time -- also when the pass has already broken down, so the
// solve then returns x = 0, r = b as the zero-guess initialisation leaves them.
template <typename XT, bool NT, bool FIRST = false>
__global__ __launch_bounds__(kThreads) void k_upd_rx(long long n, XT* __restrict__ r, const XT* __restrict__ v,
                                                     XT* __restrict__ x, const XT* __restrict__ u,
                                                     const KState* __restrict__ st, const XT* __restrict__ b0 = nullptr) {
    if (st->status) {
        if constexpr (FIRST) {
            for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
                put<NT>(r + i, b0[i]);
                put<NT>(x + i, XT(0));
            }
        }
        return;
    }
    const double a = st->alpha;
    DARE_CHUNK_LOOP(n) {
        double rv[kU], vv[kU], xv[kU], uv[kU];
#pragma unroll
        for (int q = 0; q < kU; q++) {
            const long long i = DARE_IDX(q) < n ? DARE_IDX(q) : 0;
            rv[q] = static_cast<double>(FIRST ? b0[i] : r[i]);
            vv[q] = static_cast<double>(v[i]);
            xv[q] = FIRST ? 0. : static_cast<double>(x[i]);
            uv[q] = static_cast<double>(u[i]);
        }
#pragma unroll
        for (int q = 0; q < kU; q++) {
            const long long i = DARE_IDX(q);
            if (i < n) {
                put<NT>(r + i, static_cast<XT>(rv[q] - a * vv[q]));                                // (8)
                put<NT>(x + i, static_cast<XT>(1. * xv[q] + a * uv[q]));                           // (10)
            }
        }
    }
}

template <typename XT, bool NT>
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
            if (DARE_IDX(q) < n) put<NT>(v + DARE_IDX(q), static_cast<XT>(sv[q] - beta * vv[q]));   // (12)
    }
}

template <typename XT, bool NT>
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
                put<NT>(u + i, static_cast<XT>(rv[q] - beta * uv[q]));                             // (15)
                put<NT>(r + i, static_cast<XT>(rv[q] - a * vv[q]));                                // (16)
                put<NT>(s + i, static_cast<XT>(sv[q] - a * wv[q]));                                // (17)
            }
        }
    }
}

// x += om1 r + om2 s; x += alpha u; r -= om1 s + om2 t; u -= om1 v + om2 w; <r,r>, <rh,r>
template <typename XT, bool NT>
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
                    put<NT>(x + i, static_cast<XT>(xi));
                    const XT rn = static_cast<XT>(1. * rv[q] + (-o1) * sv[q] + (-o2) * tv[q]);  // (25)
                    put<NT>(r + i, rn);
                    put<NT>(u + i, static_cast<XT>(1. * uv[q] + (-o1) * vv[q] + (-o2) * wv[q]));
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

// Initialisation (0)-(2) for a zero initial guess: A x0 is +0 in every row, so r = b - A x0 = b
// bit for bit, and <r,r>, <rh,r> go to the reduction as EpiInit's do. Nothing is written: the
// first pass reads b where it would read r and 0 where it would read x or u
// (k_upd_u / k_upd_rx <..., FIRST>), and writes them. rh = r0 is never written again, so the
// caller reads b itself as rh (b stays untouched during the solve); v and w are written before
// they are read (EpiInit).
template <typename XT>
__global__ __launch_bounds__(kThreads) void k_init_zero(long long n, const XT* __restrict__ b, ReduceWs ws, KState* st) {
    double acc[2] = {0., 0.};
    DARE_CHUNK_LOOP(n) {
        XT bv[kU];
#pragma unroll
        for (int q = 0; q < kU; q++) bv[q] = b[DARE_IDX(q) < n ? DARE_IDX(q) : 0];
#pragma unroll
        for (int q = 0; q < kU; q++) {
            if (DARE_IDX(q) < n) {
                const double rd = static_cast<double>(bv[q]);
                acc[0] += rd * rd;
                acc[1] += rd * rd;
            }
        }
    }
    grid_reduce<2>(acc, ws, st, TAIL_INIT);
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

inline unsigned grid_rows(long long n) {      // CSR traversal: kThreads rows (one per wave and 64 rows) per block and sweep
    long long g = (n + kThreads - 1) / kThreads;
    if (g > kReduceBlocks) g = kReduceBlocks;
    g = (g + kXcd - 1) / kXcd * kXcd;             // every XCD gets the same number of blocks
    return static_cast<unsigned>(g);
}

inline unsigned grid_for(long long n) {
    long long g = (n + kThreads * kU - 1) / (kThreads * kU);
    if (g > kGridMax) g = kGridMax;
    if (g < 1) g = 1;
    return static_cast<unsigned>(g);
}

}  // namespace
}  // namespace dare_hip
