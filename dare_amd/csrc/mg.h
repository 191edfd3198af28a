// Aggregation multigrid on the Cartesian diagonal view: the right preconditioner of
// BiCGStab2 that stands in for MueLu's AMG (MatrixSystem/TrilinosSolver.inl:180-246, used as
// PreCondPackage::MueLu "AMG" in test_TrilinosSolver.cpp:214-231). Included by krylov.hip.
//
// Hierarchy: each level aggregates 2x2x2 cells of the finer one (an odd extent leaves a last
// layer of one cell; a periodic axis keeps its wrap). With P the piecewise-constant
// prolongation, the coarse operator is P^T E P + beta P^T L P, where A = E + L splits each row
// into its excess E_ii = sum of the row (mass, boundary and convection terms) and the
// zero-row-sum coupling part L. beta = 1/2 corrects the known factor two by which piecewise
// constant aggregation overweights diffusive couplings (the sum of four fine faces across a
// coarse face against one coarse face of twice the width). The coarse operator is again a
// seven-point diagonal view, held in the vector type CT of the solve: FP64 for the FP64
// solve, so that M^-1 is linear to FP64 rounding and the recurrence residual tracks the true
// one as closely as without preconditioning; FP32 for the inner solve of mixed refinement,
// whose outer loop recomputes the true residual in FP64.
//
// V-cycle (a fixed linear operator, as BiCGStab needs), with W = omega / C the damped-Jacobi
// weight of each level:
//   z0 = W r (pre-smoothing from zero)                              t = r - A z0
//   rc = P^T t (the sum over each aggregate, in a fixed order)      zc = coarse cycle of rc
//   z  = z0 + P zc                                                  out = z + W (r - A z)
// Each level runs as two kernels: k_mg_down forms rc straight from r (z0 and t are formed per
// cell and never stored; the eight cells of an aggregate sit in one wave and are summed across
// lanes) and k_mg_up forms out with z evaluated at each neighbour on the fly. The coarsest
// level (at most 1024 rows) runs its sweeps inside one work-group. Work is dealt so that each
// XCD (blocks b = x mod 8) owns one contiguous eighth of a level, keeping the neighbour planes
// of a row in that XCD's L2.
#pragma once

namespace dare_hip {
namespace {

constexpr int kMgMaxLevels = 12;
constexpr long long kMgCoarsestRows = 1024;

// coarse-level arrays are of the hierarchy's coarse type CT (double or float), set up by
// mg_setup<VT, XT> with CT = XT; they are typed at each use with mg_ptr<CT>
struct MgLevel {
    DiaView V{};
    long long rows = 0;
    void* D = nullptr;         // seven slots (coarse levels)
    void* W = nullptr;         // omega / C (coarse levels)
    void *r = nullptr, *z = nullptr;   // right-hand side and cycle result (coarse levels)
};
template <typename CT>
inline CT* mg_ptr(void* p) { return static_cast<CT*>(p); }

struct MgHierarchy {
    int nlev = 0;              // levels including the fine one
    MgLevel lev[kMgMaxLevels];
    void* coarse = nullptr;    // one allocation for every coarse level's arrays
    size_t coarse_bytes = 0;
    void* fine = nullptr;      // fine level: W (VT) and the cycle result (XT)
    size_t fine_bytes = 0;
    const void* D0 = nullptr;  // fine-level slots (VT) of the last set-up
    double omega = 0.8;
    int coarse_sweeps = 12;
};

// Neighbour of cell p[3] through slot s (0 W, 1 S, 2 B, 4 T, 5 N, 6 E) with the diagonal
// view's presence and wrap rules (dia_column): false when the boundary edit removed it, else
// its coordinates in q.
__device__ inline bool mg_nb(const DiaView& V, const int p[3], int s, int q[3]) {
    const int d = s < 3 ? s : 6 - s;
    const int n = V.n[d];
    int c = p[d];
    if (s < 3) {
        if (c == 0 && V.rm_lo[d]) return false;
        c = (c == 0) ? n - 1 : c - 1;
    } else {
        if (c == n - 1 && V.rm_hi[d] && !(c == 0 && V.rm_lo[d])) return false;
        c = (c == n - 1) ? 0 : c + 1;
    }
    q[0] = p[0];
    q[1] = p[1];
    q[2] = p[2];
    q[d] = c;
    return true;
}
__device__ inline long long mg_row(const DiaView& V, const int q[3]) {
    return (static_cast<long long>(q[0]) * V.n[1] + q[1]) * V.n[2] + q[2];
}
__device__ inline void mg_cell(const DiaView& V, long long row, int p[3]) {
    p[2] = static_cast<int>(row % V.n[2]);
    p[1] = static_cast<int>((row / V.n[2]) % V.n[1]);
    p[0] = static_cast<int>(row / (static_cast<long long>(V.n[1]) * V.n[2]));
}
// aggregate (coarse row) of fine cell q
__device__ inline long long mg_agg(const DiaView& Vc, const int q[3]) {
    return (static_cast<long long>(q[0] / 2) * Vc.n[1] + q[1] / 2) * Vc.n[2] + q[2] / 2;
}

// this lane's rows of an n-row level: XCD x (blocks b = x mod 8) owns rows [n x / 8, n (x+1) / 8)
struct MgSlab {
    long long first, end, step;
};
__device__ inline MgSlab mg_slab(long long n) {
    const int xcd = static_cast<int>(blockIdx.x % kXcd);
    const long long per = gridDim.x / kXcd;
    MgSlab S;
    S.first = n * xcd / kXcd + static_cast<long long>(blockIdx.x / kXcd) * kThreads + threadIdx.x;
    S.end = n * (xcd + 1) / kXcd;
    S.step = per * kThreads;
    return S;
}

// coarse slots and W of level l+1 from the slots of level l (one lane per coarse row)
template <typename VT, typename CT>
__global__ __launch_bounds__(kThreads) void k_mg_coarsen(DiaView Vf, const VT* __restrict__ Df, DiaView Vc,
                                                        CT* __restrict__ Dc, CT* __restrict__ Wc, double beta,
                                                        double omega) {
    const long long nrf = Vf.nrows, nrc = Vc.nrows;
    const MgSlab S = mg_slab(nrc);
    for (long long row = S.first; row < S.end; row += S.step) {
        int agg[3];
        mg_cell(Vc, row, agg);
        double excess = 0., cpl[7] = {0., 0., 0., 0., 0., 0., 0.};
        for (int di = 0; di < 2; di++)
            for (int dj = 0; dj < 2; dj++)
                for (int dk = 0; dk < 2; dk++) {
                    const int p[3] = {2 * agg[0] + di, 2 * agg[1] + dj, 2 * agg[2] + dk};
                    if (p[0] >= Vf.n[0] || p[1] >= Vf.n[1] || p[2] >= Vf.n[2]) continue;
                    const long long frow = mg_row(Vf, p);
                    excess += static_cast<double>(Df[3 * nrf + frow]);
                    for (int s = 0; s < 7; s++) {
                        if (s == 3) continue;
                        int q[3];
                        if (!mg_nb(Vf, p, s, q)) continue;                 // removed boundary entry (0)
                        const double a = static_cast<double>(Df[s * nrf + frow]);
                        excess += a;
                        const int d = s < 3 ? s : 6 - s;
                        if (q[d] / 2 != agg[d]) cpl[s] += beta * a;        // coupling to a neighbour aggregate
                    }
                }
        double c = excess;
        for (int s = 0; s < 7; s++)
            if (s != 3) {
                c -= cpl[s];
                Dc[s * nrc + row] = static_cast<CT>(cpl[s]);
            }
        Dc[3 * nrc + row] = static_cast<CT>(c);
        Wc[row] = static_cast<CT>(omega / c);
    }
}

// W = omega / C of the fine level
template <typename VT>
__global__ __launch_bounds__(kThreads) void k_mg_weight(long long n, const VT* __restrict__ C, VT* __restrict__ W,
                                                       double omega) {
    const MgSlab S = mg_slab(n);
    for (long long i = S.first; i < S.end; i += S.step) W[i] = static_cast<VT>(omega / static_cast<double>(C[i]));
}

// Fine cells of a level in wave-sized bricks of 16 (k) x 2 (j) x 2 (i) cells: lane l of brick
// (I, J, G) holds cell (2I + (l >> 5 & 1), 2J + (l >> 4 & 1), 16 G + (l & 15)), so the 2x2x2
// cells of one aggregate sit in lanes that differ in bits 0, 4 and 5. Bricks are dealt so that
// XCD x (blocks b = x mod 8) owns one contiguous eighth of them.
struct MgBricks {
    int ng;                    // bricks along k
    long long total;
};
inline MgBricks mg_bricks(const DiaView& V) {
    MgBricks B;
    B.ng = (V.n[2] + 15) / 16;
    B.total = static_cast<long long>((V.n[0] + 1) / 2) * ((V.n[1] + 1) / 2) * B.ng;
    return B;
}
inline unsigned mg_brick_grid(const MgBricks& B) {   // four bricks (waves) per work-group, a multiple of kXcd
    long long g = (B.total + 3) / 4;
    if (g > 65536) g = 65536;                         // one brick per wave up to 256^3
    g = (g + kXcd - 1) / kXcd * kXcd;
    return static_cast<unsigned>(g);
}
struct MgBrickIter {
    long long first, end, step;
};
__device__ inline MgBrickIter mg_brick_iter(const MgBricks& B) {
    const int xcd = static_cast<int>(blockIdx.x % kXcd);
    MgBrickIter it;
    it.first = B.total * xcd / kXcd + static_cast<long long>(blockIdx.x / kXcd) * 4 + threadIdx.x / 64;
    it.end = B.total * (xcd + 1) / kXcd;
    it.step = static_cast<long long>(gridDim.x / kXcd) * 4;
    return it;
}
// this lane's cell of brick w; false when it lies outside the level
__device__ inline bool mg_brick_cell(const DiaView& V, const MgBricks& B, long long w, int p[3]) {
    const int lane = threadIdx.x % 64;
    const unsigned nj = static_cast<unsigned>((V.n[1] + 1) / 2);
    const unsigned wu = static_cast<unsigned>(w);     // bricks < 2^32 (mg_setup)
    const unsigned ij = wu / static_cast<unsigned>(B.ng);
    p[0] = 2 * static_cast<int>(ij / nj) + ((lane >> 5) & 1);
    p[1] = 2 * static_cast<int>(ij % nj) + ((lane >> 4) & 1);
    p[2] = 16 * static_cast<int>(wu - ij * static_cast<unsigned>(B.ng)) + (lane & 15);
    return p[0] < V.n[0] && p[1] < V.n[1] && p[2] < V.n[2];
}

// Row of the neighbour of cell p through slot s, or the cell's own row when the boundary edit
// removed that neighbour: its slot holds 0, so the term it feeds is +0.0 and every load can
// be issued unconditionally (all of a cell's loads in flight together).
__device__ inline long long mg_nb_row(const DiaView& V, const int p[3], int s, long long self, int q[3]) {
    if (!mg_nb(V, p, s, q)) {
        q[0] = p[0];
        q[1] = p[1];
        q[2] = p[2];
        return self;
    }
    return mg_row(V, q);
}

// rc = P^T (r - A W r) (PRE) or P^T r: one lane per fine cell, the aggregate's eight values
// summed across lanes in a fixed order
template <bool PRE, typename DT, typename XT, typename CT>
__global__ __launch_bounds__(kThreads) void k_mg_down(DiaView Vf, MgBricks Bf, const DT* __restrict__ Df,
                                                     const DT* __restrict__ Wf, const XT* __restrict__ r, DiaView Vc,
                                                     CT* __restrict__ rc) {
    const long long nrf = Vf.nrows;
    const MgBrickIter it = mg_brick_iter(Bf);
    for (long long w = it.first; w < it.end; w += it.step) {
        int p[3];
        const bool in = mg_brick_cell(Vf, Bf, w, p);
        if (!in) {                    // a valid cell for the loads; the lane's value is dropped
            p[0] = p[0] < Vf.n[0] ? p[0] : Vf.n[0] - 1;
            p[1] = p[1] < Vf.n[1] ? p[1] : Vf.n[1] - 1;
            p[2] = p[2] < Vf.n[2] ? p[2] : Vf.n[2] - 1;
        }
        const long long frow = mg_row(Vf, p);
        double t = static_cast<double>(r[frow]);
        if (PRE) {
            long long nb[7];
            double a[7], wn[7], rn[7];
#pragma unroll
            for (int s = 0; s < 7; s++) {
                int q[3];
                nb[s] = s == 3 ? frow : mg_nb_row(Vf, p, s, frow, q);
            }
#pragma unroll
            for (int s = 0; s < 7; s++) {
                a[s] = static_cast<double>(Df[s * nrf + frow]);
                wn[s] = static_cast<double>(Wf[nb[s]]);
                rn[s] = static_cast<double>(r[nb[s]]);
            }
            double y = a[3] * (wn[3] * rn[3]);
#pragma unroll
            for (int s = 0; s < 7; s++)
                if (s != 3) y += a[s] * (wn[s] * rn[s]);
            t -= y;
        }
        if (!in) t = 0.;
        t += __shfl_xor(t, 1);
        t += __shfl_xor(t, 16);
        t += __shfl_xor(t, 32);
        const int lane = threadIdx.x % 64;
        if (in && (lane & 0x31) == 0) {
            const int ag[3] = {p[0] / 2, p[1] / 2, p[2] / 2};
            rc[mg_row(Vc, ag)] = static_cast<CT>(t);
        }
    }
}

// out = z + W (r - A z) with z = [W r] + P zc evaluated at each neighbour: one lane per row
template <bool PRE, typename DT, typename XT, typename CT>
__global__ __launch_bounds__(kThreads) void k_mg_up(DiaView Vf, MgBricks Bf, const DT* __restrict__ Df,
                                                   const DT* __restrict__ Wf, const XT* __restrict__ r, DiaView Vc,
                                                   const CT* __restrict__ zc, XT* __restrict__ out) {
    const long long nrf = Vf.nrows;
    const MgBrickIter it = mg_brick_iter(Bf);
    for (long long w = it.first; w < it.end; w += it.step) {
        int p[3];
        if (!mg_brick_cell(Vf, Bf, w, p)) continue;
        const long long row = mg_row(Vf, p);
        long long nb[7], ag[7];
#pragma unroll
        for (int s = 0; s < 7; s++) {
            int q[3];
            nb[s] = s == 3 ? row : mg_nb_row(Vf, p, s, row, q);
            ag[s] = s == 3 ? mg_agg(Vc, p) : mg_agg(Vc, q);
        }
        double a[7], wn[7], rn[7], zn[7];
#pragma unroll
        for (int s = 0; s < 7; s++) {
            a[s] = static_cast<double>(Df[s * nrf + row]);
            wn[s] = PRE ? static_cast<double>(Wf[nb[s]]) : 0.;
            rn[s] = PRE ? static_cast<double>(r[nb[s]]) : 0.;
            zn[s] = static_cast<double>(zc[ag[s]]);
        }
        const double wr = static_cast<double>(Wf[row]), rr = static_cast<double>(r[row]);
        const double zi = (PRE ? wn[3] * rn[3] : 0.) + zn[3];
        double y = a[3] * zi;
#pragma unroll
        for (int s = 0; s < 7; s++)
            if (s != 3) y += a[s] * ((PRE ? wn[s] * rn[s] : 0.) + zn[s]);
        out[row] = static_cast<XT>(zi + wr * (rr - y));
    }
}

// the coarsest level: z = W r, then sweeps - 1 damped-Jacobi sweeps, all in one work-group
// (rows <= 1024) with z held in LDS
template <typename CT>
__global__ __launch_bounds__(1024) void k_mg_coarsest(DiaView V, const CT* __restrict__ D, const CT* __restrict__ W,
                                                     const CT* __restrict__ r, CT* __restrict__ z, int sweeps) {
    __shared__ CT za[kMgCoarsestRows], zb[kMgCoarsestRows];
    const long long nr = V.nrows;
    const int row = threadIdx.x;
    CT a[7] = {0, 0, 0, 0, 0, 0, 0};
    int nb[7] = {-1, -1, -1, -1, -1, -1, -1};
    CT rv = 0, wv = 0;
    if (row < nr) {
        int p[3];
        mg_cell(V, row, p);
        for (int s = 0; s < 7; s++) {
            a[s] = D[s * nr + row];
            int q[3];
            nb[s] = s == 3 ? row : mg_nb(V, p, s, q) ? static_cast<int>(mg_row(V, q)) : -1;
        }
        rv = r[row];
        wv = W[row];
        za[row] = wv * rv;
    }
    __syncthreads();
    CT* cur = za;
    CT* nxt = zb;
    for (int it = 1; it < sweeps; it++) {
        if (row < nr) {
            CT y = 0;
            for (int s = 0; s < 7; s++)
                if (nb[s] >= 0) y += a[s] * cur[nb[s]];
            nxt[row] = cur[row] + wv * (rv - y);
        }
        __syncthreads();
        CT* tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    if (row < nr) z[row] = cur[row];
}

// out = x + z (x may be NULL: out = z); out may alias x
template <typename XT>
__global__ __launch_bounds__(kThreads) void k_mg_combine(long long n, const XT* x, const XT* __restrict__ z, XT* out) {
    for (long long i = blockIdx.x * (long long)kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads)
        out[i] = x ? static_cast<XT>(x[i] + z[i]) : z[i];
}

inline unsigned mg_grid(long long n) {   // a multiple of kXcd (mg_slab)
    long long g = (n + kThreads - 1) / kThreads;
    if (g > 4096) g = 4096;
    g = (g + kXcd - 1) / kXcd * kXcd;
    return static_cast<unsigned>(g);
}

// Builds (or refreshes) the hierarchy of a one-component 3-D diagonal view whose level-0 slots
// D0 are of type VT; XT is the vector type of the solve and of the coarse levels.
template <typename VT, typename XT>
int mg_setup(dare_ctx* ctx, MgHierarchy* H, const DiaView& V0, const VT* D0) {
    using CT = XT;
    if (V0.N != 1 || V0.cross) {
        set_error("the multigrid preconditioner needs a one-component operator without cross entries");
        return DARE_ERR_UNSUPPORTED;
    }
    MgLevel lv[kMgMaxLevels];
    lv[0].V = V0;
    lv[0].rows = V0.nrows;
    int nlev = 1;
    while (nlev < kMgMaxLevels && (nlev == 1 || lv[nlev - 1].rows > kMgCoarsestRows)) {
        const DiaView& F = lv[nlev - 1].V;
        DiaView C = F;
        for (int d = 0; d < 3; d++) C.n[d] = (F.n[d] + 1) / 2;
        C.nrows = static_cast<long long>(C.n[0]) * C.n[1] * C.n[2];
        if (C.nrows == F.nrows) break;
        lv[nlev].V = C;
        lv[nlev].rows = C.nrows;
        nlev++;
    }
    if (nlev < 2 || lv[nlev - 1].rows > kMgCoarsestRows) {
        set_error("multigrid needs a grid that coarsens (at least two cells along an axis) to at most 1024 rows");
        return DARE_ERR_UNSUPPORTED;
    }
    if (mg_bricks(V0).total >= (1LL << 32)) {
        set_error("grid too large for the multigrid preconditioner");
        return DARE_ERR_UNSUPPORTED;
    }
    long long need = 0;
    for (int l = 1; l < nlev; l++) need += 10 * lv[l].rows;            // 7 slots, W, r, z
    const size_t coarse_need = sizeof(CT) * static_cast<size_t>(need);
    const size_t fine_need = (sizeof(VT) + sizeof(XT)) * static_cast<size_t>(V0.nrows);
    if (H->coarse_bytes < coarse_need) {
        if (H->coarse) (void)hipFree(H->coarse);
        H->coarse = nullptr;
        H->coarse_bytes = 0;
        if (hipMalloc(&H->coarse, coarse_need) != hipSuccess) {
            (void)hipGetLastError();
            set_error("multigrid hierarchy allocation failed");
            return DARE_ERR_OOM;
        }
        H->coarse_bytes = coarse_need;
    }
    if (H->fine_bytes < fine_need) {
        if (H->fine) (void)hipFree(H->fine);
        H->fine = nullptr;
        H->fine_bytes = 0;
        if (hipMalloc(&H->fine, fine_need) != hipSuccess) {
            (void)hipGetLastError();
            set_error("multigrid work vector allocation failed");
            return DARE_ERR_OOM;
        }
        H->fine_bytes = fine_need;
    }
    CT* p = static_cast<CT*>(H->coarse);
    for (int l = 1; l < nlev; l++) {
        const long long n = lv[l].rows;
        lv[l].D = p;
        p += 7 * n;
        lv[l].W = p;
        p += n;
        lv[l].r = p;
        p += n;
        lv[l].z = p;
        p += n;
    }
    lv[0].W = H->fine;
    lv[0].z = static_cast<char*>(H->fine) + sizeof(VT) * V0.nrows;
    H->nlev = nlev;
    H->D0 = D0;
    for (int l = 0; l < nlev; l++) H->lev[l] = lv[l];
    hipLaunchKernelGGL((k_mg_weight<VT>), dim3(mg_grid(V0.nrows)), dim3(kThreads), 0, ctx->stream, V0.nrows,
                       D0 + 3 * V0.nrows, mg_ptr<VT>(lv[0].W), H->omega);
    hipLaunchKernelGGL((k_mg_coarsen<VT, CT>), dim3(mg_grid(lv[1].rows)), dim3(kThreads), 0, ctx->stream, V0, D0,
                       lv[1].V, mg_ptr<CT>(lv[1].D), mg_ptr<CT>(lv[1].W), 0.5, H->omega);
    for (int l = 2; l < nlev; l++)
        hipLaunchKernelGGL((k_mg_coarsen<CT, CT>), dim3(mg_grid(lv[l].rows)), dim3(kThreads), 0, ctx->stream,
                           lv[l - 1].V, static_cast<const CT*>(lv[l - 1].D), lv[l].V, mg_ptr<CT>(lv[l].D),
                           mg_ptr<CT>(lv[l].W), 0.5, H->omega);
    DARE_LAUNCH_CHECK();
    return DARE_OK;
}

// z = M^-1 r on the fine level (slots of type VT, vectors XT, coarse levels of type XT, as set
// up by mg_setup<VT, XT>); returns the device vector that holds z (the hierarchy's fine result
// vector, rewritten by the next call)
template <typename VT, typename XT>
XT* mg_apply(hipStream_t sm, const MgHierarchy& H, const XT* r) {
    using CT = XT;
    const VT* D0 = static_cast<const VT*>(H.D0);
    const int L = H.nlev - 1;
    const MgLevel& F = H.lev[0];
    const VT* W0 = static_cast<const VT*>(F.W);
    const MgBricks B0 = mg_bricks(F.V);
    hipLaunchKernelGGL((k_mg_down<true, VT, XT, CT>), dim3(mg_brick_grid(B0)), dim3(kThreads), 0, sm, F.V, B0, D0,
                       W0, r, H.lev[1].V, mg_ptr<CT>(H.lev[1].r));
    for (int l = 1; l < L; l++) {
        const MgLevel& A = H.lev[l];
        const MgLevel& C = H.lev[l + 1];
        const MgBricks B = mg_bricks(A.V);
        hipLaunchKernelGGL((k_mg_down<true, CT, CT, CT>), dim3(mg_brick_grid(B)), dim3(kThreads), 0, sm, A.V, B,
                           static_cast<const CT*>(A.D), static_cast<const CT*>(A.W), static_cast<const CT*>(A.r), C.V,
                           mg_ptr<CT>(C.r));
    }
    const MgLevel& Z = H.lev[L];
    hipLaunchKernelGGL((k_mg_coarsest<CT>), dim3(1), dim3(1024), 0, sm, Z.V, static_cast<const CT*>(Z.D),
                       static_cast<const CT*>(Z.W), static_cast<const CT*>(Z.r), mg_ptr<CT>(Z.z), H.coarse_sweeps);
    for (int l = L - 1; l >= 1; l--) {
        const MgLevel& A = H.lev[l];
        const MgLevel& C = H.lev[l + 1];
        const MgBricks B = mg_bricks(A.V);
        hipLaunchKernelGGL((k_mg_up<true, CT, CT, CT>), dim3(mg_brick_grid(B)), dim3(kThreads), 0, sm, A.V, B,
                           static_cast<const CT*>(A.D), static_cast<const CT*>(A.W), static_cast<const CT*>(A.r), C.V,
                           static_cast<const CT*>(C.z), mg_ptr<CT>(A.z));
    }
    XT* z = mg_ptr<XT>(F.z);
    hipLaunchKernelGGL((k_mg_up<true, VT, XT, CT>), dim3(mg_brick_grid(B0)), dim3(kThreads), 0, sm, F.V, B0, D0, W0,
                       r, H.lev[1].V, static_cast<const CT*>(H.lev[1].z), z);
    return z;
}

}  // namespace
}  // namespace dare_hip
