// Device assembly of the per-cell operator into a sorted CSR matrix.
//
// Each row (cell q, component c) re-evaluates on the device the functor a reference user
// hands to Trilinos<SC>::Build (MatrixSystem/Trilinos.inl:308-328):
//     mb = ddt(phi); mb -= div(D_f * grad(mb)); mb += div(tvd * phi); rhs += V s;
//     reaction; boundary edits (Get(C) -= Get(NB); rhs -= 2 v Get(NB); Remove(NB))
// with the reference's floating-point operation order (DDT.inl:41-81, Gradient_Cartesian.inl:
// 39-58, Divergence_Cartesian.inl:141-207, TVD_Cartesian.inl:185-261, MatrixBlock_Cartesian.inl:
// 487-593) and MatrixBlock::Finalize's column rule (:128-240) followed by the column sort of
// Tpetra's fillComplete. Structure: count -> exclusive scan -> columns. Values are re-run
// every step without touching the structure.
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "common.h"

namespace dare_hip {

namespace {

constexpr int kMaxEntries = 8;     // 7-point stencil + one cross-component entry

struct Terms {
    double tt;                 // (1 V) / dt, or 0 if no transient term
    int has_ddt;
    double diff_const;
    const double* diff_cell;
    double diff_comp[kMaxComp];
    int use_diff_comp;
    int conv_limiter;
    const double* vel[3];
    const double* source;
    const double* react;
    int cross;                 // ncomp == 2 with a reaction block: one cross-component entry
    int bc_type[6];
    double bc_value[6];
    long long fs[3][3];        // face array strides
    const double* phi;         // phi_old, padded
    double div_w[3];           // the divergences' time-scheme weights (div_w[0] = 1: EULER_BACKWARD)
    int div_nt;                // their time levels (1: EULER_BACKWARD)
};

__device__ inline double lim_psi(int id, double r) {
    switch (id) {
    case DARE_LIMITER_CDS: return 1.;
    case DARE_LIMITER_UPWIND: return 0.;
    case DARE_LIMITER_VANALBADA: {
        r = isnan(r) ? 0. : r;
        const double r2 = r * r;
        return (r + r2) / (1. + r2);
    }
    case DARE_LIMITER_MINMOD: {
        r = isnan(r) ? 0. : r;
        const double m = (1. < r) ? 1. : r;
        return (0. < m) ? m : 0.;
    }
    case DARE_LIMITER_VANLEER: {
        if (isnan(r)) return 0.;
        if (isinf(r)) return r > 0. ? 2. : 0.;
        const double a = fabs(r);
        return (r + a) / (1. + a);
    }
    default: return 0.;
    }
}

// Column of neighbor position pos (0 = centre, 1..6 = W,E,S,N,B,T) with the periodic wrap of
// MapIndexToOrdinalGlobalInternal(ind, true) (CartesianRepresentation.inl:418-450).
__device__ inline long long nb_col(const Geo& G, const int q[3], int pos) {
    int p[3] = {q[0], q[1], q[2]};
    if (pos > 0) {
        const int d = (pos - 1) / 2;
        p[d] += (pos & 1) ? -1 : 1;
    }
    long long o = 0;
    for (int d = 0; d < G.dim; d++) {
        int v = p[d];
        if (G.any_periodic) {
            if (v < 0) v = G.n[d] + v;
            else if (v >= G.n[d]) v = v - G.n[d];
        }
        o += G.si[d] * v;
    }
    return o;
}

// Builds the (unsorted) entries of one row in Finalize order: explicit cross entry first, then
// W,S,B,C,T,N,E for set positions. Then stable-sorts by column and merges duplicates.
// MODE 0: structure only; 1: evaluate the functor; 2: values taken from the diagonal-view
// slots (slot_in) written earlier by mode 1. SORT = false stops after the values (the
// diagonal-view fill needs no column order).
enum { kRowStructure = 0, kRowValues = 1, kRowFromSlots = 2 };
constexpr int kPosOfSlot[7] = {1, 3, 5, 0, 6, 4, 2};

// DIM = 0 takes the dimension from G at run time; DIM = 3 (the diagonal-view fill) lets every
// per-axis loop unroll, so the per-position arrays live in registers.
// FEAT: the terms the instantiation may meet (kFeat* bits); a cleared bit drops that term's
// code (the host dispatches the diagonal-view fill of operators without them to FEAT = 0).
enum { kFeatConv = 1, kFeatDiffCell = 2, kFeatSource = 4, kFeatCross = 8, kFeatFixed = 16, kFeatAll = 31 };
template <int MODE, bool SORT, int DIM = 0, int FEAT = kFeatAll>
__device__ int row_entries(const Geo& G, const Terms& T, const int q[3], int c, long long node,
                           int* cols, double* vals, double* rhs_out, double* x0_out,
                           double* slotv = nullptr, const double* slot_in = nullptr) {
    constexpr bool VALUES = MODE == kRowValues;
    const int dim = DIM ? DIM : G.dim;
    double nbv[7] = {0., 0., 0., 0., 0., 0., 0.};
    double rhs = 0.;
    bool removed[7] = {false, false, false, false, false, false, false};
    double cross = 0.;
    const bool has_cross = (FEAT & kFeatCross) && T.cross != 0;
    long long lp = 0;
    for (int d = 0; d < dim; d++) lp += G.sl[d] * (q[d] + G.g);
    const int N = G.N;

    // a cell on a DARE_BC_FIXED face (the first axis that has one, low face first) is an
    // identity row: every neighbour and the cross-component entry removed
    int fixed_face = -1;
    for (int d = 0; (FEAT & kFeatFixed) && d < dim && fixed_face < 0; d++) {
        if (q[d] == 0 && T.bc_type[2 * d] == DARE_BC_FIXED) fixed_face = 2 * d;
        else if (q[d] == G.n[d] - 1 && T.bc_type[2 * d + 1] == DARE_BC_FIXED) fixed_face = 2 * d + 1;
    }
    const bool fixed = fixed_face >= 0;
    // boundary faces hit by an edit (they only depend on the structure)
    int bc_face[3] = {-1, -1, -1};
    for (int d = 0; d < dim; d++) {
        if (q[d] == 0 && T.bc_type[2 * d] != 0) bc_face[d] = 2 * d;
        else if (q[d] == G.n[d] - 1 && T.bc_type[2 * d + 1] != 0) bc_face[d] = 2 * d + 1;
        // position bc_face + 1, written with constant indices
        if (bc_face[d] == 2 * d) removed[2 * d + 1] = true;
        else if (bc_face[d] == 2 * d + 1) removed[2 * d + 2] = true;
    }
    if (fixed)
        for (int pos = 1; pos < 7; pos++) removed[pos] = true;

    if (VALUES) {
        const double* phi = T.phi;
        if (T.has_ddt) {                                   // Set(ddt(field))
            nbv[0] = T.tt;
            rhs = 0.;
            rhs += T.tt * phi[lp * N + c];
        }
        {   // mb -= div(D_f * grad(mb))
            double o[7] = {0., 0., 0., 0., 0., 0., 0.};
            for (int d = 0; d < dim; d++) {
                double Df[2];
                for (int h = 0; h < 2; h++) {
                    if ((FEAT & kFeatDiffCell) && T.diff_cell) {
                        const long long lq = lp + (h ? G.sl[d] : -G.sl[d]);
                        const double Dp = T.diff_cell[lp * N + c], Dq = T.diff_cell[lq * N + c];
                        const double Dlo = h ? Dp : Dq, Dhi = h ? Dq : Dp;
                        Df[h] = ((2. * Dlo) * Dhi) / (Dlo + Dhi);
                    } else if (T.use_diff_comp) {
                        Df[h] = T.diff_comp[c];
                    } else {
                        Df[h] = T.diff_const;
                    }
                }
                const double nbW = -G.dnr[d] * Df[0], cW = G.dnr[d] * Df[0];
                const double nbE = G.dnr[d] * Df[1], cE = -G.dnr[d] * Df[1];
                o[2 * d + 2] = G.A[d] * nbE;
                if (d == 0) o[0] = G.A[0] * cE;
                else o[0] += G.A[d] * cE;
                o[2 * d + 1] = -G.A[d] * nbW;
                o[0] -= G.A[d] * cW;
            }
            // the scheme weight of the implicit level (s_c *= coef_td); the gradient's face
            // right-hand sides are 0, so the explicit levels add nothing here
            for (int pos = 0; pos < 2 * dim + 1; pos++) nbv[pos] += (o[pos] * T.div_w[0]) * -1.;
            rhs += 0. * -1.;
        }
        if ((FEAT & kFeatConv) && T.conv_limiter >= 0) {   // mb += div(tvd * phi)
            double o[7] = {0., 0., 0., 0., 0., 0., 0.};
            double orhs = 0., frd[3] = {0., 0., 0.};
            for (int d = 0; d < dim; d++) {
                long long fi = 0;
                for (int e = 0; e < dim; e++) fi += T.fs[d][e] * q[e];
                const double u[2] = {T.vel[d][fi], T.vel[d][fi + T.fs[d][d]]};
                double nb[2], cc[2], fr[2];
                const long long s = G.sl[d] * N;
                for (int h = 0; h < 2; h++) {
                    const bool up = u[h] >= 0.;
                    const double vp = up ? 1. : 0., vn = up ? 0. : 1.;
                    const double* base = T.phi + (lp + (h ? 0 : -G.sl[d])) * N + c;   // cell L
                    double pUU, pU, pD;
                    if (up) {
                        pUU = base[-s];
                        pU = base[0];
                        pD = base[s];
                    } else {
                        pUU = base[2 * s];
                        pU = base[s];
                        pD = base[0];
                    }
                    const double r = (pU - pUU) / (pD - pU);
                    const double psi = lim_psi(T.conv_limiter, r);
                    if (h == 0) {
                        nb[0] = vp * u[0];
                        cc[0] = vn * u[0];
                    } else {
                        nb[1] = vn * u[1];
                        cc[1] = vp * u[1];
                    }
                    fr[h] = ((psi * -0.5) * (pD - pU)) * u[h];
                }
                o[2 * d + 2] = G.A[d] * nb[1];
                if (d == 0) o[0] = G.A[0] * cc[1];
                else o[0] += G.A[d] * cc[1];
                o[2 * d + 1] = -G.A[d] * nb[0];
                o[0] -= G.A[d] * cc[0];
                frd[d] = fr[1] - fr[0];
                if (d == 0) orhs = G.A[0] * frd[0];
                else orhs += G.A[d] * frd[d];
            }
            // time scheme (Divergence_Cartesian.inl:183-204): coefficients times w0, the face
            // right-hand sides once more per explicit level t, weighted by w_t
            for (int t = 1; t < T.div_nt; t++)
                for (int d = 0; d < dim; d++) orhs += T.div_w[t] * G.A[d] * frd[d];
            for (int pos = 0; pos < 2 * dim + 1; pos++) nbv[pos] += o[pos] * T.div_w[0];
            rhs += orhs;
        }
        if ((FEAT & kFeatSource) && T.source) rhs += G.V * T.source[lp * N + c];
        if (has_cross) {
            const double* R = T.react + node * 4;
            nbv[0] -= G.V * R[c * 2 + c];
            cross = -(G.V * R[c * 2 + (1 - c)]);
        }
        for (int d = 0; d < dim; d++) {                  // boundary edits
            const int f = bc_face[d];
            if (f < 0) continue;
            const double nbf = (f == 2 * d) ? nbv[2 * d + 1] : nbv[2 * d + 2];   // nbv[f + 1]
            if (T.bc_type[f] == DARE_BC_DIRICHLET) {
                nbv[0] -= nbf;
                rhs -= (2. * T.bc_value[f]) * nbf;
            } else {
                nbv[0] += nbf;
            }
        }
        *rhs_out = rhs;
        *x0_out = phi[lp * N + c];
        if (fixed) {                                     // Get<CENTER> = 1, rhs = x0 = value
            nbv[0] = 1.;
            cross = 0.;
            *rhs_out = T.bc_value[fixed_face];
            *x0_out = T.bc_value[fixed_face];
        }
        if (slotv) {      // diagonal view: W,S,B,C,T,N,E,cross; removed entries hold 0
            for (int sl = 0; sl < 7; sl++) {
                const int pos = kPosOfSlot[sl];
                slotv[sl] = (pos <= 2 * dim && !removed[pos]) ? nbv[pos] : 0.;
            }
            slotv[7] = cross;
        }
    }
    if (MODE == kRowFromSlots) {
        for (int sl = 0; sl < 7; sl++) nbv[kPosOfSlot[sl]] = slot_in[sl];
        cross = has_cross ? slot_in[7] : 0.;
    }
    if (!SORT) return 0;

    // Finalize order, then stable sort by column + merge of equal columns
    long long kc[kMaxEntries];
    double kv[kMaxEntries];
    int k = 0;
    if (has_cross && !fixed) {
        kc[k] = node * N + (1 - c);
        kv[k] = cross;
        k++;
    }
    const int order[7] = {1, 3, 5, 0, 6, 4, 2};
    for (int j = 0; j < 7; j++) {
        const int pos = order[j];
        if (pos > 2 * dim || removed[pos]) continue;
        kc[k] = nb_col(G, q, pos) * N + c;
        kv[k] = nbv[pos];
        k++;
    }
    for (int i = 1; i < k; i++) {                          // insertion sort (stable)
        const long long ck = kc[i];
        const double cv = kv[i];
        int j = i - 1;
        while (j >= 0 && kc[j] > ck) {
            kc[j + 1] = kc[j];
            kv[j + 1] = kv[j];
            j--;
        }
        kc[j + 1] = ck;
        kv[j + 1] = cv;
    }
    int w = 0;
    for (int i = 0; i < k; i++) {
        if (w > 0 && kc[w - 1] == kc[i]) {
            kv[w - 1] += kv[i];
        } else {
            kc[w] = kc[i];
            kv[w] = kv[i];
            w++;
        }
    }
    for (int i = 0; i < w; i++) {
        if (cols) cols[i] = static_cast<int>(kc[i]);
        if (MODE != kRowStructure) vals[i] = kv[i];
    }
    return w;
}

__device__ inline void row_to_cell(const Geo& G, long long row, int q[3], int* c, long long* node) {
    *c = static_cast<int>(row % G.N);
    long long nd = row / G.N;
    *node = nd;
    q[0] = q[1] = q[2] = 0;
    for (int d = 0; d < G.dim; d++) {
        q[d] = static_cast<int>(nd / G.si[d]);
        nd -= static_cast<long long>(q[d]) * G.si[d];
    }
}

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ inline int row_count(const Geo& G, const Terms& T, long long row) {
    int q[3], c;
    long long node;
    row_to_cell(G, row, q, &c, &node);
    int cols[kMaxEntries];
    return row_entries<kRowStructure, true>(G, T, q, c, node, cols, nullptr, nullptr, nullptr);
}

// block-wide inclusive scan of one int per thread (wave64 shuffles + LDS)
__device__ inline int block_scan_incl(int v, int* lds) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    if (lane == 63) lds[wid] = v;
    __syncthreads();
    int add = 0;
    for (int w = 0; w < wid; w++) add += lds[w];
    __syncthreads();
    return v + add;
}

__global__ __launch_bounds__(kScanThreads) void k_count_tiles(Geo G, Terms T, long long nrows,
                                                              long long* __restrict__ tile_sums) {
    __shared__ int lds[kScanThreads / 64];
    const long long base = static_cast<long long>(blockIdx.x) * kScanTile;
    int s = 0;
    for (int i = 0; i < kScanItems; i++) {
        const long long row = base + static_cast<long long>(i) * kScanThreads + threadIdx.x;
        if (row < nrows) s += row_count(G, T, row);
    }
    const int incl = block_scan_incl(s, lds);
    if (threadIdx.x == kScanThreads - 1) tile_sums[blockIdx.x] = incl;
}

// single work-group exclusive scan of the tile sums; tile_sums[ntiles] receives the total
__global__ __launch_bounds__(1024) void k_scan_tiles(long long* __restrict__ tile_sums, long long ntiles) {
    __shared__ long long lds[1024];
    long long carry = 0;
    for (long long base = 0; base < ntiles; base += 1024) {
        const long long i = base + threadIdx.x;
        const long long v = i < ntiles ? tile_sums[i] : 0;
        lds[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {          // Hillis-Steele on LDS
            const long long t = threadIdx.x >= static_cast<unsigned>(off) ? lds[threadIdx.x - off] : 0;
            __syncthreads();
            lds[threadIdx.x] += t;
            __syncthreads();
        }
        const long long incl = lds[threadIdx.x];
        if (i < ntiles) tile_sums[i] = carry + incl - v;
        __syncthreads();
        carry += lds[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) tile_sums[ntiles] = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_write_structure(Geo G, Terms T, long long nrows,
                                                                  const long long* __restrict__ tile_off,
                                                                  int* __restrict__ row_ptr,
                                                                  int* __restrict__ col) {
    __shared__ int lds[kScanThreads / 64];
    const long long base = static_cast<long long>(blockIdx.x) * kScanTile;
    // each thread owns kScanItems consecutive rows so that a block scan gives row offsets
    int cnt[kScanItems];
    int s = 0;
    for (int i = 0; i < kScanItems; i++) {
        const long long row = base + static_cast<long long>(threadIdx.x) * kScanItems + i;
        cnt[i] = row < nrows ? row_count(G, T, row) : 0;
        s += cnt[i];
    }
    const int incl = block_scan_incl(s, lds);
    long long off = tile_off[blockIdx.x] + incl - s;
    for (int i = 0; i < kScanItems; i++) {
        const long long row = base + static_cast<long long>(threadIdx.x) * kScanItems + i;
        if (row >= nrows) break;
        row_ptr[row] = static_cast<int>(off);
        int q[3], c;
        long long node;
        row_to_cell(G, row, q, &c, &node);
        int cols[kMaxEntries];
        const int k = row_entries<kRowStructure, true>(G, T, q, c, node, cols, nullptr, nullptr, nullptr);
        for (int j = 0; j < k; j++) col[off + j] = cols[j];
        off += k;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) row_ptr[nrows] = static_cast<int>(tile_off[gridDim.x]);
}

__global__ __launch_bounds__(256) void k_fill_values(Geo G, Terms T, long long nrows,
                                                     const int* __restrict__ row_ptr,
                                                     double* __restrict__ val, double* __restrict__ rhs,
                                                     double* __restrict__ x0) {
    const long long row = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (row >= nrows) return;
    int q[3], c;
    long long node;
    row_to_cell(G, row, q, &c, &node);
    int cols[kMaxEntries];
    double vals[kMaxEntries];
    double r = 0., x = 0.;
    const int k = row_entries<kRowValues, true>(G, T, q, c, node, cols, vals, &r, &x);
    const int o = row_ptr[row];
    for (int j = 0; j < k; j++) val[o + j] = vals[j];
    rhs[row] = r;
    if (x0) x0[row] = x;
}

// Diagonal-view fill (3-D): one lane per cell, lanes along k, no column sort. Writes the slot
// values (coalesced, slot-major), rhs and x0. The CSR values are materialized from the slots
// only when a CSR consumer asks for them (k_materialize_csr, same sort and merge). HALF (an
// operator symmetric by construction): only W, S, B, C are written; E, N, T equal the W, S, B
// of the east, north and top neighbours, bit for bit (k_dia_mirror writes them on demand).
template <bool NT, bool F32, int FEAT, bool HALF>
__global__ __launch_bounds__(256) void k_fill_dia(Geo G, Terms T, double* __restrict__ dia, float* __restrict__ dia32,
                                                  int slots, double* __restrict__ rhs, double* __restrict__ x0) {
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int i = blockIdx.z;
    if (k >= G.n[2] || j >= G.n[1]) return;
    const int q[3] = {i, j, k};
    const long long node = (static_cast<long long>(i) * G.n[1] + j) * G.n[2] + k;
    const long long nrows = G.ncell_int * G.N;
    for (int c = 0; c < G.N; c++) {
        const long long row = node * G.N + c;
        double slotv[kDiaSlots];
        double r = 0., x = 0.;
        row_entries<kRowValues, false, 3, FEAT>(G, T, q, c, node, nullptr, nullptr, &r, &x, slotv);
        if constexpr (HALF) {                           // W, S, B, C
#pragma unroll
            for (int sl = 0; sl < 4; sl++) put<NT>(dia + sl * nrows + row, slotv[sl]);
            if constexpr (F32)
#pragma unroll
                for (int sl = 0; sl < 4; sl++) put<NT>(dia32 + sl * nrows + row, static_cast<float>(slotv[sl]));
        } else if constexpr ((FEAT & kFeatCross) == 0) {      // seven slots
#pragma unroll
            for (int sl = 0; sl < 7; sl++) put<NT>(dia + sl * nrows + row, slotv[sl]);
            if constexpr (F32)
#pragma unroll
                for (int sl = 0; sl < 7; sl++) put<NT>(dia32 + sl * nrows + row, static_cast<float>(slotv[sl]));
        } else {
            for (int sl = 0; sl < slots; sl++) put<NT>(dia + sl * nrows + row, slotv[sl]);
            if constexpr (F32)
                for (int sl = 0; sl < slots; sl++) put<NT>(dia32 + sl * nrows + row, static_cast<float>(slotv[sl]));
        }
        put<NT>(rhs + row, r);
        if (x0) put<NT>(x0 + row, x);
    }
}

// E, N, T of every row of a one-component view from the W, S, B slots of its east, north and
// top neighbours (periodic wraps as in the traversals); entries a boundary edit removed get 0
// (never read). Exact for a view whose fill left them out (k_fill_dia<..., HALF>).
template <typename VT>
__global__ __launch_bounds__(256) void k_dia_mirror(DiaView V, VT* __restrict__ D) {
    const long long r = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    if (r >= V.nrows) return;
    const int n0 = V.n[0], n1 = V.n[1], n2 = V.n[2];
    const long long sI = static_cast<long long>(n1) * n2, sJ = n2, nr = V.nrows;
    const int k = static_cast<int>(r % n2);
    const long long t = r / n2;
    const int j = static_cast<int>(t % n1), i = static_cast<int>(t / n1);
    const bool hasE = !(i == n0 - 1 && V.rm_hi[0]);
    const bool hasN = !(j == n1 - 1 && V.rm_hi[1]);
    const bool hasT = !(k == n2 - 1 && V.rm_hi[2]);
    const long long east = (i == n0 - 1) ? r - static_cast<long long>(n0 - 1) * sI : r + sI;
    const long long north = (j == n1 - 1) ? r - static_cast<long long>(n1 - 1) * sJ : r + sJ;
    const long long top = (k == n2 - 1) ? r - (n2 - 1) : r + 1;
    D[4 * nr + r] = hasT ? D[2 * nr + top] : VT(0);
    D[5 * nr + r] = hasN ? D[1 * nr + north] : VT(0);
    D[6 * nr + r] = hasE ? D[0 * nr + east] : VT(0);
}

__global__ __launch_bounds__(256) void k_materialize_csr(Geo G, Terms T, long long nrows,
                                                         const int* __restrict__ row_ptr,
                                                         const double* __restrict__ dia, int slots,
                                                         double* __restrict__ val) {
    const long long row = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (row >= nrows) return;
    int q[3], c;
    long long node;
    row_to_cell(G, row, q, &c, &node);
    double sl_in[kDiaSlots];
    for (int sl = 0; sl < kDiaSlots; sl++) sl_in[sl] = sl < slots ? dia[sl * nrows + row] : 0.;
    double vals[kMaxEntries];
    const int k = row_entries<kRowFromSlots, true>(G, T, q, c, node, nullptr, vals, nullptr, nullptr, nullptr,
                                                   sl_in);
    const int o = row_ptr[row];
    for (int jj = 0; jj < k; jj++) val[o + jj] = vals[jj];
}

int make_terms(const Geo& G, const dare_terms* t, const double* phi, Terms* T) {
    std::memset(T, 0, sizeof(*T));
    T->has_ddt = t->dt > 0.;
    T->tt = T->has_ddt ? (1. * G.V) / t->dt : 0.;
    T->diff_const = t->diff_const;
    T->diff_cell = t->diff_cell;
    for (int c = 0; c < kMaxComp; c++) T->diff_comp[c] = t->diff_comp[c];
    T->use_diff_comp = t->use_diff_comp;
    T->conv_limiter = t->conv_limiter;
    for (int d = 0; d < 3; d++) T->vel[d] = t->vel[d];
    T->source = t->source;
    T->react = t->react;
    T->cross = (t->react != nullptr && G.N == 2) ? 1 : 0;
    if (t->div_levels < 0 || t->div_levels > 3) {
        set_error("div_levels must be 0..3");
        return DARE_ERR_INVALID;
    }
    T->div_nt = t->div_levels > 0 ? t->div_levels : 1;
    for (int l = 0; l < 3; l++) T->div_w[l] = (t->div_levels > 0 && l < t->div_levels) ? t->div_weights[l] : (l == 0 ? 1. : 0.);
    for (int f = 0; f < 6; f++) {
        T->bc_type[f] = f < 2 * G.dim ? t->bc_type[f] : 0;
        T->bc_value[f] = t->bc_value[f];
        if (T->bc_type[f] < 0 || T->bc_type[f] > DARE_BC_FIXED) {
            set_error("invalid boundary type");
            return DARE_ERR_INVALID;
        }
    }
    if (t->conv_limiter > DARE_LIMITER_VANLEER) {
        set_error("unknown limiter");
        return DARE_ERR_INVALID;
    }
    if (t->conv_limiter >= 0) {
        if (G.g < 2) {
            set_error("implicit TVD convection needs >= 2 ghost layers");
            return DARE_ERR_INVALID;
        }
        for (int d = 0; d < G.dim; d++)
            if (!t->vel[d]) {
                set_error("convection needs face velocities for every axis");
                return DARE_ERR_INVALID;
            }
    }
    if (G.g < 1) {
        set_error("assembly needs >= 1 ghost layer");
        return DARE_ERR_INVALID;
    }
    for (int ax = 0; ax < 3; ax++) {
        long long nf[3];
        for (int e = 0; e < 3; e++) nf[e] = G.n[e] + (e == ax ? 1 : 0);
        for (int d = 0; d < 3; d++) {
            long long s = 1;
            for (int e = d + 1; e < G.dim; e++) s *= nf[e];
            T->fs[ax][d] = d < G.dim ? s : 0;
        }
    }
    T->phi = phi;
    return DARE_OK;
}

}  // namespace

int dia_ensure_full(dare_csr* A, hipStream_t stream) {
    if (!A->ent_lazy) return DARE_OK;
    const unsigned grid = static_cast<unsigned>(ceil_div(A->nrows, 256));
    hipLaunchKernelGGL(k_dia_mirror<double>, dim3(grid), dim3(256), 0, stream, A->dia_view, A->dia);
    if (A->dia32 && A->dia32_fresh)
        hipLaunchKernelGGL(k_dia_mirror<float>, dim3(grid), dim3(256), 0, stream, A->dia_view, A->dia32);
    DARE_LAUNCH_CHECK();
    A->ent_lazy = 0;
    A->ent_full = 1;
    return DARE_OK;
}

int csr_ensure_values(dare_csr* A, hipStream_t stream) {
    if (!A->vals_stale) return DARE_OK;
    int rc = dia_ensure_full(A, stream);
    if (rc) return rc;
    Geo G;
    rc = make_geo(&A->grid_desc, &G);
    if (rc) return rc;
    Terms T;
    std::memset(&T, 0, sizeof(T));
    T.cross = A->cross;
    for (int f = 0; f < 6; f++) T.bc_type[f] = A->bc_type[f];
    hipLaunchKernelGGL(k_materialize_csr, dim3(ceil_div(A->nrows, 256)), dim3(256), 0, stream, G, T, A->nrows,
                       A->row_ptr, A->dia, A->cross ? 8 : 7, A->val);
    DARE_LAUNCH_CHECK();
    A->vals_stale = 0;
    return DARE_OK;
}

}  // namespace dare_hip

using namespace dare_hip;

namespace {
struct DeviceFree {
    void operator()(void* p) const {
        if (p) (void)hipFree(p);
    }
};
}  // namespace

static int ensure_rows(dare_csr* A, int64_t nrows) {
    if (A->cap_rows >= nrows + 1 && A->row_ptr) return DARE_OK;
    if (A->row_ptr) hipFree(A->row_ptr);
    A->row_ptr = nullptr;
    if (hipMalloc(&A->row_ptr, sizeof(int32_t) * (nrows + 1)) != hipSuccess) {
        hipGetLastError();
        set_error("CSR row pointer allocation failed");
        return DARE_ERR_OOM;
    }
    A->cap_rows = nrows + 1;
    return DARE_OK;
}

static int ensure_nnz(dare_csr* A, int64_t nnz) {
    if (A->cap_nnz >= nnz && A->col && A->val) return DARE_OK;
    if (A->col) hipFree(A->col);
    if (A->val) hipFree(A->val);
    A->col = nullptr;
    A->val = nullptr;
    if (hipMalloc(&A->col, sizeof(int32_t) * (nnz ? nnz : 1)) != hipSuccess ||
        hipMalloc(&A->val, sizeof(double) * (nnz ? nnz : 1)) != hipSuccess) {
        hipGetLastError();
        set_error("CSR value allocation failed");
        return DARE_ERR_OOM;
    }
    A->cap_nnz = nnz;
    return DARE_OK;
}

extern "C" {

int dare_csr_create(dare_ctx* ctx, dare_csr** out) {
    DARE_CHECK_ARG(ctx && out, "NULL argument");
    dare_csr* A = new dare_csr();
    A->device = ctx->device;
    *out = A;
    return DARE_OK;
}

int dare_csr_destroy(dare_csr* A) {
    if (!A) return DARE_OK;
    // hipFree synchronizes with outstanding work on the device before releasing memory
    hipSetDevice(A->device);
    if (A->row_ptr) hipFree(A->row_ptr);
    if (A->col) hipFree(A->col);
    if (A->val) hipFree(A->val);
    if (A->val32) hipFree(A->val32);
    if (A->dia) hipFree(A->dia);
    if (A->dia32) hipFree(A->dia32);
    if (A->asym64) hipFree(A->asym64);
    if (A->dinv) hipFree(A->dinv);
    if (A->dia_pre) hipFree(A->dia_pre);
    mg_free(A->mg);
    delete A;
    return DARE_OK;
}

int dare_csr_build_structure(dare_ctx* ctx, dare_csr* A, const dare_grid* g, const dare_terms* t) {
    DARE_CHECK_ARG(ctx && A && g && t, "NULL argument");
    Geo G;
    int rc = make_geo(g, &G);
    if (rc) return rc;
    Terms T;
    rc = make_terms(G, t, nullptr, &T);
    if (rc) return rc;
    DARE_CHECK_ARG(!(t->react && G.N != 2), "reaction blocks need ncomp == 2");
    // Without a periodic axis, MatrixBlock::Finalize maps the outside neighbor of a boundary
    // cell to an out-of-range column unless the functor removed it: refuse such operators.
    if (!G.any_periodic) {
        for (int f = 0; f < 2 * G.dim; f++)
            DARE_CHECK_ARG(t->bc_type[f] != DARE_BC_NONE,
                           "non-periodic boundary face without a boundary edit (out-of-range columns)");
    }
    // A boundary edit touches one face per axis (the reference functor's if / else-if on the low
    // and the high index, counter_diffusion2D.cpp:125-161): on a single-cell non-periodic axis the
    // high neighbour stays and maps past the last row. Refuse such grids.
    for (int d = 0; d < G.dim; d++)
        DARE_CHECK_ARG(G.n[d] > 1 || G.per[d],
                       "single-cell non-periodic axis (its high neighbour column would be out of range)");
    const long long nrows = G.ncell_int * G.N;
    rc = ensure_rows(A, nrows);
    if (rc) return rc;
    const long long ntiles = (nrows + kScanTile - 1) / kScanTile;
    long long* tiles_raw = nullptr;
    DARE_HIP_TRY(hipMalloc(reinterpret_cast<void**>(&tiles_raw), sizeof(long long) * (ntiles + 1)));
    // freed on every return path; hipFree waits for the kernels that still use it
    std::unique_ptr<long long, DeviceFree> tiles(tiles_raw);
    hipLaunchKernelGGL(k_count_tiles, dim3(ntiles), dim3(kScanThreads), 0, ctx->stream, G, T, nrows, tiles.get());
    DARE_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, ctx->stream, tiles.get(), ntiles);
    DARE_LAUNCH_CHECK();
    long long total = 0;
    DARE_HIP_TRY(hipMemcpyAsync(&total, tiles.get() + ntiles, sizeof(long long), hipMemcpyDeviceToHost, ctx->stream));
    DARE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    DARE_CHECK_ARG(total < (1LL << 31), "nnz exceeds 32-bit row pointers");
    rc = ensure_nnz(A, total);
    if (rc) return rc;
    hipLaunchKernelGGL(k_write_structure, dim3(ntiles), dim3(kScanThreads), 0, ctx->stream, G, T, nrows, tiles.get(),
                       A->row_ptr, A->col);
    DARE_LAUNCH_CHECK();
    DARE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    DARE_HIP_TRY(hipFree(tiles.release()));
    A->nrows = nrows;
    A->nnz = total;
    A->stencil_structured = 1;
    A->grid_desc = *g;
    for (int f = 0; f < 6; f++) A->bc_type[f] = T.bc_type[f];
    A->cross = T.cross;
    A->vals_stale = 0;
    A->dia32_fresh = 0;
    A->sym_fill = 0;
    // diagonal view for the solver (3-D, 1 or 2 components)
    A->has_dia = 0;
    if (G.dim == 3 && (G.N == 1 || G.N == 2)) {
        DiaView& V = A->dia_view;
        for (int d = 0; d < 3; d++) {
            V.n[d] = G.n[d];
            V.per[d] = G.per[d];
            V.rm_lo[d] = T.bc_type[2 * d] != 0;
            V.rm_hi[d] = T.bc_type[2 * d + 1] != 0;
        }
        V.N = G.N;
        V.cross = (t->react != nullptr && G.N == 2) ? 1 : 0;
        V.nrows = nrows;
        const int slots = V.cross ? 8 : 7;
        const int64_t need = static_cast<int64_t>(slots) * nrows;
        if (A->cap_dia < need) {
            if (A->dia) hipFree(A->dia);
            A->dia = nullptr;
            A->cap_dia = 0;
            if (hipMalloc(&A->dia, sizeof(double) * need) != hipSuccess) {
                hipGetLastError();
                set_error("diagonal view allocation failed");
                return DARE_ERR_OOM;
            }
            A->cap_dia = need;
        }
        if (!A->asym64 && hipMalloc(&A->asym64, sizeof(int)) != hipSuccess) {
            hipGetLastError();
            set_error("symmetry flag allocation failed");
            return DARE_ERR_OOM;
        }
        A->sym64 = -1;
        A->ent_lazy = 0;
        A->has_dia = 1;
    }
    return DARE_OK;
}

int dare_csr_fill_values(dare_ctx* ctx, dare_csr* A, const dare_grid* g, const dare_terms* t,
                         const double* phi_old, double* rhs, double* x0) {
    DARE_CHECK_ARG(ctx && A && g && t && phi_old && rhs, "NULL argument");
    Geo G;
    int rc = make_geo(g, &G);
    if (rc) return rc;
    Terms T;
    rc = make_terms(G, t, phi_old, &T);
    if (rc) return rc;
    const long long nrows = G.ncell_int * G.N;
    DARE_CHECK_ARG(A->stencil_structured && A->nrows == nrows, "CSR structure does not match the grid");
    // the diagonal view and the lazy CSR values walk the build-time grid: the fill must use the same one
    {
        const dare_grid& B = A->grid_desc;
        bool same = B.dim == g->dim && B.ncomp == g->ncomp;
        for (int d = 0; d < g->dim && same; d++)
            same = B.n[d] == g->n[d] && (B.periodic[d] != 0) == (g->periodic[d] != 0);
        DARE_CHECK_ARG(same, "grid differs from the one the CSR structure was built on (dim, ncomp, n, periodic)");
    }
    DARE_CHECK_ARG(T.cross == A->cross, "reaction terms differ from the ones the structure was built with");
    for (int f = 0; f < 6; f++)
        DARE_CHECK_ARG(T.bc_type[f] == A->bc_type[f], "boundary edits differ from the structure's");
    if (A->has_dia) {
        const dim3 grid(ceil_div(G.n[2], 64), ceil_div(G.n[1], 4), G.n[0]);
        const int slots = A->cross ? 8 : 7;
        // the FP32 copy rides along once a mixed solve has used this matrix
        const bool f32 = A->dia32_mode && A->dia32 && A->cap_dia32 >= static_cast<int64_t>(slots) * nrows;
        const bool nt = A->nrows >= kStreamRows;
        // operators with only DDT, constant / per-component diffusion and Dirichlet / Neumann
        // edits take the instantiation without the other terms' code
        bool plain = T.conv_limiter < 0 && !T.diff_cell && !T.source && !T.cross;
        for (int f = 0; f < 6; f++) plain = plain && T.bc_type[f] != DARE_BC_FIXED;
        // Symmetric by construction: a face coefficient is A_d * (dnr_d * D_f) in both rows it
        // couples (with the signs exchanged twice), and the per-cell harmonic mean takes its two
        // diffusivities in face order (low cell first) from either side. Not claimed for fixed
        // faces (identity rows), for per-cell diffusivity across a periodic wrap (the wrapped
        // face reads the caller's ghost values, which need not be the periodic images) or for a
        // single-cell axis (a cell coupled to itself through both faces; the structure refuses
        // it unless periodic).
        bool sym = G.N == 1 && t->conv_limiter < 0 && t->react == nullptr;
        for (int f = 0; f < 6; f++) sym = sym && t->bc_type[f] != DARE_BC_FIXED;
        for (int d = 0; d < G.dim; d++) sym = sym && G.n[d] > 1 && !(t->diff_cell && G.per[d]);
        // E, N, T are left out when the solvers will take the symmetric traversals (an even k
        // extent for the FP64 pairs, a multiple of four for the FP32 quads) and no consumer of
        // this matrix has needed the seven slots yet
        const bool half = sym && !A->cross && A->sym_traversal && !A->ent_full && G.n[2] % 2 == 0 &&
                          (!f32 || G.n[2] % 4 == 0);
        auto launch = [&](auto ntc, auto f32c, auto halfc) {
            constexpr bool NTV = decltype(ntc)::value, F32V = decltype(f32c)::value, HV = decltype(halfc)::value;
            if (plain)
                hipLaunchKernelGGL((k_fill_dia<NTV, F32V, 0, HV>), grid, dim3(256), 0, ctx->stream, G, T, A->dia,
                                   F32V ? A->dia32 : nullptr, slots, rhs, x0);
            else
                hipLaunchKernelGGL((k_fill_dia<NTV, F32V, kFeatAll, HV>), grid, dim3(256), 0, ctx->stream, G, T,
                                   A->dia, F32V ? A->dia32 : nullptr, slots, rhs, x0);
        };
        using Yes = std::true_type;
        using No = std::false_type;
        if (half) {
            if (nt && f32) launch(Yes{}, Yes{}, Yes{});
            else if (nt) launch(Yes{}, No{}, Yes{});
            else if (f32) launch(No{}, Yes{}, Yes{});
            else launch(No{}, No{}, Yes{});
        } else {
            if (nt && f32) launch(Yes{}, Yes{}, No{});
            else if (nt) launch(Yes{}, No{}, No{});
            else if (f32) launch(No{}, Yes{}, No{});
            else launch(No{}, No{}, No{});
        }
        DARE_LAUNCH_CHECK();
        A->dia32_fresh = f32 ? 1 : 0;
        A->vals_stale = 1;
        A->ent_lazy = half ? 1 : 0;
        A->sym_fill = sym ? 1 : 0;
        // the cases just excluded may still give an exactly symmetric view: the solvers check
        // those on the device
        A->sym_cand = (G.N == 1 && t->conv_limiter < 0 && t->react == nullptr) ? 1 : 0;
        A->sym64 = -1;
        return DARE_OK;
    }
    hipLaunchKernelGGL(k_fill_values, dim3(ceil_div(nrows, 256)), dim3(256), 0, ctx->stream, G, T, nrows,
                       A->row_ptr, A->val, rhs, x0);
    DARE_LAUNCH_CHECK();
    A->vals_stale = 0;
    return DARE_OK;
}

int dare_csr_upload_host(dare_ctx* ctx, dare_csr* A, int64_t nrows, int64_t nnz, const int32_t* row_ptr,
                         const int32_t* col, const double* val) {
    DARE_CHECK_ARG(ctx && A && row_ptr && (nnz == 0 || (col && val)) && nrows >= 0 && nnz >= 0,
                   "invalid CSR upload arguments");
    DARE_CHECK_ARG(row_ptr[0] == 0 && row_ptr[nrows] == nnz, "row_ptr inconsistent with nnz");
    for (int64_t r = 0; r < nrows; r++) DARE_CHECK_ARG(row_ptr[r + 1] >= row_ptr[r], "row_ptr not monotone");
    for (int64_t k = 0; k < nnz; k++) DARE_CHECK_ARG(col[k] >= 0 && col[k] < nrows, "column index out of range");
    int rc = ensure_rows(A, nrows);
    if (rc) return rc;
    rc = ensure_nnz(A, nnz);
    if (rc) return rc;
    DARE_HIP_TRY(hipMemcpyAsync(A->row_ptr, row_ptr, sizeof(int32_t) * (nrows + 1), hipMemcpyHostToDevice, ctx->stream));
    if (nnz) {
        DARE_HIP_TRY(hipMemcpyAsync(A->col, col, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, ctx->stream));
        DARE_HIP_TRY(hipMemcpyAsync(A->val, val, sizeof(double) * nnz, hipMemcpyHostToDevice, ctx->stream));
    }
    DARE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    A->nrows = nrows;
    A->nnz = nnz;
    A->stencil_structured = 0;
    A->has_dia = 0;
    A->vals_stale = 0;
    A->dia32_fresh = 0;
    A->sym_fill = 0;
    A->sym_cand = 0;
    A->sym64 = -1;
    A->ent_lazy = 0;
    return DARE_OK;
}

int dare_csr_info(const dare_csr* A, int64_t* nrows, int64_t* nnz, const int32_t** row_ptr,
                  const int32_t** col, const double** val) {
    DARE_CHECK_ARG(A != nullptr, "NULL csr");
    if (A->vals_stale) {    // hand out valid CSR values: materialize on the null stream
        // the fill ran on a context stream that may not synchronize with the null stream
        DARE_HIP_TRY(hipDeviceSynchronize());
        int rc = csr_ensure_values(const_cast<dare_csr*>(A), nullptr);
        if (rc) return rc;
        DARE_HIP_TRY(hipDeviceSynchronize());
    }
    if (nrows) *nrows = A->nrows;
    if (nnz) *nnz = A->nnz;
    if (row_ptr) *row_ptr = A->row_ptr;
    if (col) *col = A->col;
    if (val) *val = A->val;
    return DARE_OK;
}

int dare_csr_download(dare_ctx* ctx, const dare_csr* A, int32_t* row_ptr, int32_t* col, double* val) {
    DARE_CHECK_ARG(ctx && A, "NULL argument");
    int rc = csr_ensure_values(const_cast<dare_csr*>(A), ctx->stream);
    if (rc) return rc;
    if (row_ptr)
        DARE_HIP_TRY(hipMemcpyAsync(row_ptr, A->row_ptr, sizeof(int32_t) * (A->nrows + 1), hipMemcpyDeviceToHost, ctx->stream));
    if (col && A->nnz)
        DARE_HIP_TRY(hipMemcpyAsync(col, A->col, sizeof(int32_t) * A->nnz, hipMemcpyDeviceToHost, ctx->stream));
    if (val && A->nnz)
        DARE_HIP_TRY(hipMemcpyAsync(val, A->val, sizeof(double) * A->nnz, hipMemcpyDeviceToHost, ctx->stream));
    DARE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return DARE_OK;
}

}  // extern "C"
