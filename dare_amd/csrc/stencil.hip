// Field-wide FV stencils for gfx950: ghost fill, face gradient, value divergence, TVD face
// flux and the fused explicit TVD step. One work-item per internal cell; wave64 lanes run
// along the fastest (last) axis so every plane-neighbour load is a coalesced row.
// Floating-point operation order follows the reference per-cell code exactly (the file is
// built with -ffp-contract=off), so results are bit-identical to the per-cell evaluation.
#include "common.h"

namespace dare_hip {

struct FaceGeo {            // strides of the three face arrays (axis d has n[d]+1 faces)
    long long fs[3][3];
};

static FaceGeo make_face_geo(const Geo& G) {
    FaceGeo F;
    for (int ax = 0; ax < 3; ax++) {
        long long nf[3];
        for (int e = 0; e < 3; e++) nf[e] = G.n[e] + (e == ax ? 1 : 0);
        for (int d = 0; d < 3; d++) {
            long long s = 1;
            for (int e = d + 1; e < G.dim; e++) s *= nf[e];
            F.fs[ax][d] = d < G.dim ? s : 0;
        }
    }
    return F;
}

// Launch shape: x along axis dim-1 (64 lanes), y along axis dim-2, z along axis dim-3.
struct Launch {
    dim3 grid, block;
};

static Launch cell_launch(const Geo& G) {
    Launch L;
    L.block = dim3(64, G.dim >= 2 ? 4 : 1, 1);
    const int ax_x = G.dim - 1, ax_y = G.dim - 2, ax_z = G.dim - 3;
    L.grid = dim3(ceil_div(G.n[ax_x], 64), ax_y >= 0 ? ceil_div(G.n[ax_y], 4) : 1, ax_z >= 0 ? G.n[ax_z] : 1);
    return L;
}

template <int DIM>
__device__ inline bool cell_coord(const Geo& G, int q[3]) {
    q[0] = q[1] = q[2] = 0;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    q[DIM - 1] = a;
    if (a >= G.n[DIM - 1]) return false;
    if (DIM >= 2) {
        const int b = blockIdx.y * blockDim.y + threadIdx.y;
        if (b >= G.n[DIM - 2]) return false;
        q[DIM - 2] = b;
    }
    if (DIM == 3) q[0] = blockIdx.z;
    return true;
}

template <int DIM>
__device__ inline long long pad_ord(const Geo& G, const int q[3]) {   // internal q -> padded ordinal
    long long o = 0;
#pragma unroll
    for (int d = 0; d < DIM; d++) o += G.sl[d] * (q[d] + G.g);
    return o;
}

template <int DIM>
__device__ inline long long face_ord(const FaceGeo& F, int ax, const int q[3]) {
    long long o = 0;
#pragma unroll
    for (int d = 0; d < DIM; d++) o += F.fs[ax][d] * q[d];
    return o;
}

// Limiter Psi(r): FluxLimiter.h:41-173 with std::min/std::max as comparisons; van Leer from
// its closed form with the same NaN -> 0 convention and limits 2 / 0 at +inf / -inf.
__device__ inline double limiter(int id, double r) {
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
        const double m = (1. < r) ? 1. : r;       // std::min(r, 1)
        return (0. < m) ? m : 0.;                 // std::max(0, m)
    }
    case DARE_LIMITER_VANLEER: {
        // branch-free: the quotient is formed for every r (inf/inf for infinite r) and the
        // special cases are selected over it, so a wave never splits and the loads of the
        // stencil can all be in flight before the first division
        const double a = fabs(r);
        double q = (r + a) / (1. + a);
        asm volatile("" : "+v"(q));   // keeps the division unconditional (not sunk into a branch)
        const double psi = isinf(r) ? (r > 0. ? 2. : 0.) : q;
        return isnan(r) ? 0. : psi;
    }
    default: return 0.;
    }
}

// TVD face value between cells L and L+1 along a stride s (TVD_Cartesian.inl:133-157).
// base points at component c of cell L.
__device__ inline double tvd_face(const double* __restrict__ base, long long s, int lim, double u) {
    double pUU, pU, pD;
    if (u >= 0.) {
        pUU = base[-s];
        pU = base[0];
        pD = base[s];
    } else {
        pUU = base[2 * s];
        pU = base[s];
        pD = base[0];
    }
    const double r = (pU - pUU) / (pD - pU);
    const double psi = limiter(lim, r);
    return pU + (psi * 0.5) * (pD - pU);
}

// ------------------------------------------------------------------------------ ghost fill
struct GhostSlabs {
    long long start[7];      // prefix offsets of the 6 slabs (W,E,S,N,B,T)
    int other[6][2];         // the two non-normal axes per slab (or -1)
    int mode[6];
    double value[6];
};

template <int DIM>
__global__ void k_ghost_fill(Geo G, GhostSlabs S, double* __restrict__ phi) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= S.start[2 * DIM]) return;
    int f = 0;
    while (f < 2 * DIM - 1 && t >= S.start[f + 1]) f++;
    if (S.mode[f] == DARE_GHOST_LEAVE) return;
    long long r = t - S.start[f];
    const int d = f / 2;
    const bool low = (f % 2) == 0;
    int p[3] = {0, 0, 0};       // padded index
    // decompose r over (layer, other axes) with the last other axis fastest
    int o1 = S.other[f][0], o2 = S.other[f][1];
    int i2 = 0, i1 = 0;
    if (o2 >= 0) {
        i2 = static_cast<int>(r % G.n[o2]);
        r /= G.n[o2];
    }
    if (o1 >= 0) {
        i1 = static_cast<int>(r % G.n[o1]);
        r /= G.n[o1];
    }
    const int l = static_cast<int>(r) + 1;   // ghost layer 1..g
    if (o1 >= 0) p[o1] = i1 + G.g;
    if (o2 >= 0) p[o2] = i2 + G.g;
    p[d] = low ? G.g - l : G.g + G.n[d] - 1 + l;
    int src[3] = {p[0], p[1], p[2]};
    if (S.mode[f] == DARE_GHOST_PERIODIC) src[d] = G.g + (low ? G.n[d] - l : l - 1);
    else src[d] = G.g + (low ? l - 1 : G.n[d] - l);
    long long od = 0, os = 0;
    for (int e = 0; e < DIM; e++) {
        od += G.sl[e] * p[e];
        os += G.sl[e] * src[e];
    }
    for (int c = 0; c < G.N; c++) {
        const double s = phi[os * G.N + c];
        double v;
        switch (S.mode[f]) {
        case DARE_GHOST_DIRICHLET: v = 2. * S.value[f] - s; break;
        case DARE_GHOST_CONSTANT: v = S.value[f]; break;
        default: v = s; break;
        }
        phi[od * G.N + c] = v;
    }
}

// --------------------------------------------------------------------------- face gradient
// Gradient_Cartesian.inl:62-96: W = (phiC - phiW) dn_r, E = (phiE - phiC) dn_r, ...
template <int DIM>
__global__ __launch_bounds__(256) void k_grad_faces(Geo G, FaceGeo F, const double* __restrict__ phi,
                                                    double* __restrict__ f0, double* __restrict__ f1,
                                                    double* __restrict__ f2) {
    int q[3];
    if (!cell_coord<DIM>(G, q)) return;
    double* out[3] = {f0, f1, f2};
    const long long lp = pad_ord<DIM>(G, q);
    const int N = G.N;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const long long s = G.sl[d] * N;
        const long long fo = face_ord<DIM>(F, d, q) * N;
        const bool last = q[d] == G.n[d] - 1;
        for (int c = 0; c < N; c++) {
            const double* b = phi + lp * N + c;
            const double pc = b[0];
            out[d][fo + c] = (pc - b[-s]) * G.dnr[d];
            if (last) out[d][fo + F.fs[d][d] * N + c] = (b[s] - pc) * G.dnr[d];
        }
    }
}

// ------------------------------------------------------------ cell-centred velocity -> faces
// InterpolateToFace of a cell-centred one-component field (Interpolation_Cartesian.h:203-219,
// 298-354), as the TVD constructor evaluates it for every face (TVD_Cartesian.inl:27-38): the
// neighbour's value plus the cell's, times Divisor<2> = 0.5. Face m along axis d is shared by
// cells m-1 and m; both orders of the sum give the same bits, so each face is written once.
template <int DIM>
__global__ __launch_bounds__(256) void k_interp_faces(Geo G, FaceGeo F, const double* __restrict__ u0,
                                                      const double* __restrict__ u1, const double* __restrict__ u2,
                                                      double* __restrict__ f0, double* __restrict__ f1,
                                                      double* __restrict__ f2) {
    int q[3];
    if (!cell_coord<DIM>(G, q)) return;
    const double* in[3] = {u0, u1, u2};
    double* out[3] = {f0, f1, f2};
    const long long lp = pad_ord<DIM>(G, q);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const long long s = G.sl[d];
        const long long fo = face_ord<DIM>(F, d, q);
        const double uc = in[d][lp];
        double v = in[d][lp - s];
        v += uc;
        out[d][fo] = v * 0.5;
        if (q[d] == G.n[d] - 1) {
            double w = uc;
            w += in[d][lp + s];
            out[d][fo + F.fs[d][d]] = w * 0.5;
        }
    }
}

// ---------------------------------------------------------------------- value divergence
// Divergence_Cartesian.inl:211-235: A0 E - A0 W + A1 N - A1 S + A2 T - A2 B, in that order.
template <int DIM>
__global__ __launch_bounds__(256) void k_div_faces(Geo G, FaceGeo F, const double* __restrict__ f0,
                                                   const double* __restrict__ f1,
                                                   const double* __restrict__ f2,
                                                   double* __restrict__ out) {
    int q[3];
    if (!cell_coord<DIM>(G, q)) return;
    const double* in[3] = {f0, f1, f2};
    const long long lp = pad_ord<DIM>(G, q);
    const int N = G.N;
    for (int c = 0; c < N; c++) {
        double v = 0.;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const long long fo = face_ord<DIM>(F, d, q) * N + c;
            const double lo = in[d][fo], hi = in[d][fo + F.fs[d][d] * N];
            if (d == 0) v = G.A[0] * hi;
            else v += G.A[d] * hi;
            v -= G.A[d] * lo;
        }
        out[lp * N + c] = v;
    }
}

// -------------------------------------------------------------------------- TVD face flux
template <int DIM>
__global__ __launch_bounds__(256) void k_tvd_flux(Geo G, FaceGeo F, int lim, const double* __restrict__ phi,
                                                  const double* __restrict__ u0,
                                                  const double* __restrict__ u1,
                                                  const double* __restrict__ u2, double* __restrict__ f0,
                                                  double* __restrict__ f1, double* __restrict__ f2) {
    int q[3];
    if (!cell_coord<DIM>(G, q)) return;
    const double* vel[3] = {u0, u1, u2};
    double* out[3] = {f0, f1, f2};
    const long long lp = pad_ord<DIM>(G, q);
    const int N = G.N;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const long long s = G.sl[d] * N;
        const long long fi = face_ord<DIM>(F, d, q);
        const double ulo = vel[d][fi];
        const bool last = q[d] == G.n[d] - 1;
        const double uhi = last ? vel[d][fi + F.fs[d][d]] : 0.;
        for (int c = 0; c < N; c++) {
            const double* b = phi + lp * N + c;
            out[d][fi * N + c] = tvd_face(b - s, s, lim, ulo) * ulo;
            if (last) out[d][(fi + F.fs[d][d]) * N + c] = tvd_face(b, s, lim, uhi) * uhi;
        }
    }
}

// ------------------------------------------------------------------ fused explicit TVD step
// phi_out = phi - (dt/V) * div(phi_f * u_f); reads phi (13-point star) + 2 face velocities per
// axis, writes one value per cell.
template <int DIM>
__global__ __launch_bounds__(256) void k_tvd_step(Geo G, FaceGeo F, int lim, const double* __restrict__ phi,
                                                  const double* __restrict__ u0,
                                                  const double* __restrict__ u1,
                                                  const double* __restrict__ u2, double fac,
                                                  double* __restrict__ phi_out) {
    int q[3];
    if (!cell_coord<DIM>(G, q)) return;
    const double* vel[3] = {u0, u1, u2};
    const long long lp = pad_ord<DIM>(G, q);
    const int N = G.N;
    double ulo[3], uhi[3];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const long long fi = face_ord<DIM>(F, d, q);
        ulo[d] = vel[d][fi];
        uhi[d] = vel[d][fi + F.fs[d][d]];
    }
    for (int c = 0; c < N; c++) {
        const double* b = phi + lp * N + c;
        double v = 0.;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const long long s = G.sl[d] * N;
            const double flo = tvd_face(b - s, s, lim, ulo[d]) * ulo[d];
            const double fhi = tvd_face(b, s, lim, uhi[d]) * uhi[d];
            if (d == 0) v = G.A[0] * fhi;
            else v += G.A[d] * fhi;
            v -= G.A[d] * flo;
        }
        phi_out[lp * N + c] = b[0] - fac * v;
    }
}

// Face value between L and L+1 given the four cells a = L-1, b = L, c = L+1, d = L+2
// (same selection and arithmetic as tvd_face), times the face velocity.
__device__ inline double tvd_flux4(double a, double b, double c, double d, int lim, double u) {
    double pUU, pU, pD;
    if (u >= 0.) {
        pUU = a;
        pU = b;
        pD = c;
    } else {
        pUU = d;
        pU = c;
        pD = b;
    }
    const double r = (pU - pUU) / (pD - pU);
    const double psi = limiter(lim, r);
    return (pU + (psi * 0.5) * (pD - pU)) * u;
}

constexpr int kStepTileK = 64, kStepTileJ = 4, kStepXcd = 8, kStepMinPlanes = 8;

// One x-march of a 64 (k) x 4 (j) column over the planes [i0, i1): the planes i-2..i+2 of phi
// stay in registers and the x-face flux at (i, i+1) is carried to the next plane as its low
// face. NT streams the x-face velocities in and the result out with the non-temporal hint.
template <bool NT>
__device__ __forceinline__ void tvd_march3(const Geo& G, int lim, const double* __restrict__ phi,
                                           const double* __restrict__ ux, const double* __restrict__ uy,
                                           const double* __restrict__ uz, double fac, double* __restrict__ out,
                                           int k, int j, int i0, int i1) {
    const int n1 = G.n[1], n2 = G.n[2];
    const long long sI = G.sl[0], sJ = G.sl[1];
    long long lp = (static_cast<long long>(i0 + G.g) * G.np[1] + (j + G.g)) * G.np[2] + (k + G.g);
    // face arrays: x (n0+1, n1, n2), y (n0, n1+1, n2), z (n0, n1, n2+1)
    const long long fxs = static_cast<long long>(n1) * n2;
    long long fx = static_cast<long long>(i0) * fxs + static_cast<long long>(j) * n2 + k;
    long long fy = (static_cast<long long>(i0) * (n1 + 1) + j) * n2 + k;
    const long long fys = static_cast<long long>(n1 + 1) * n2;
    long long fz = (static_cast<long long>(i0) * n1 + j) * (n2 + 1) + k;
    const long long fzs = static_cast<long long>(n1) * (n2 + 1);
    double pm2 = phi[lp - 2 * sI], pm1 = phi[lp - sI], p0 = phi[lp], pp1 = phi[lp + sI];
    double f_lo = tvd_flux4(pm2, pm1, p0, pp1, lim, ux[fx]);
    const double A0 = G.A[0], A1 = G.A[1], A2 = G.A[2];
    for (int i = i0; i < i1; i++) {
        // every load of the plane first: x (i+2 and the face velocity), y (cells j-2..j+2),
        // z (cells k-2..k+2); the scheduling barrier keeps them ahead of the divisions
        const double pp2 = phi[lp + 2 * sI];
        double u_hi;
        if constexpr (NT) u_hi = ld_nt(ux + fx + fxs);
        else u_hi = ux[fx + fxs];
        const double ym2 = phi[lp - 2 * sJ], ym1 = phi[lp - sJ], yp1 = phi[lp + sJ], yp2 = phi[lp + 2 * sJ];
        const double vy_lo = uy[fy], vy_hi = uy[fy + n2];
        const double zm2 = phi[lp - 2], zm1 = phi[lp - 1], zp1 = phi[lp + 1], zp2 = phi[lp + 2];
        const double wz_lo = uz[fz], wz_hi = uz[fz + 1];
        __builtin_amdgcn_sched_barrier(0);
        const double f_hi = tvd_flux4(pm1, p0, pp1, pp2, lim, u_hi);
        const double gy_lo = tvd_flux4(ym2, ym1, p0, yp1, lim, vy_lo);
        const double gy_hi = tvd_flux4(ym1, p0, yp1, yp2, lim, vy_hi);
        const double hz_lo = tvd_flux4(zm2, zm1, p0, zp1, lim, wz_lo);
        const double hz_hi = tvd_flux4(zm1, p0, zp1, zp2, lim, wz_hi);
        double v = A0 * f_hi;
        v -= A0 * f_lo;
        v += A1 * gy_hi;
        v -= A1 * gy_lo;
        v += A2 * hz_hi;
        v -= A2 * hz_lo;
        put<NT>(out + lp, p0 - fac * v);
        // roll
        pm2 = pm1;
        pm1 = p0;
        p0 = pp1;
        pp1 = pp2;
        f_lo = f_hi;
        lp += sI;
        fx += fxs;
        fy += fys;
        fz += fzs;
    }
}

// Walks logical block L's column-plane range [L*chunk, (L+1)*chunk) (columns of 64 PW (k) x 4
// (j) cells ordered k-tile fastest, planes along i; lane l takes cells PW l .. PW l + PW - 1 of
// its column) and calls march(k, j, i0, i1) for each column piece with k its first cell; blocks
// are dealt XCD by XCD (blocks b = x mod 8 own a contiguous run).
template <int PW = 1, class F>
__device__ __forceinline__ void march_columns(const Geo& G, int tiles_k, long long total, long long chunk,
                                              int per_xcd, F&& march) {
    const long long L = static_cast<long long>(blockIdx.x % kStepXcd) * per_xcd + blockIdx.x / kStepXcd;
    const int n0 = G.n[0];
    long long c = L * chunk;
    const long long cend = (c + chunk < total) ? c + chunk : total;
    while (c < cend) {
        const long long col = c / n0;
        const int i0 = static_cast<int>(c - col * n0);
        const int i1 = (cend - c < n0 - i0) ? i0 + static_cast<int>(cend - c) : n0;
        const int k = (static_cast<int>(col % tiles_k) * kStepTileK + (threadIdx.x & 63)) * PW;   // PW cells per lane
        const int j = static_cast<int>(col / tiles_k) * kStepTileJ + (threadIdx.x >> 6);
        if (k < G.n[2] && j < G.n[1]) march(k, j, i0, i1);
        c += i1 - i0;
    }
}

// LDS-staged march of one full 64 (k) x TJ (j) tile (all 64 TJ threads active, so the block can
// synchronise). Plane i of phi is staged with two ghost rows and columns (rows j0-2..j0+TJ+1,
// columns k0-2..k0+65), each y and z face flux of the plane is computed once, by one lane, and
// shared through LDS (the face between rows j-1 and j by the lane of cell j; the top face row
// by the last wave; the right face column by lanes 0..TJ-1 of wave 0), and every cell combines
// its six fluxes in the order of tvd_march3. Each flux has the same four cells and velocity as
// there, so the result is bit-identical with about (3 + 2/TJ) / 5 of the flux evaluations and
// five loads per cell and plane instead of fourteen. The x faces stay in registers as in
// tvd_march3. Waves 0, 1 stage the two ghost rows below the tile, waves 2, 3 the two above.
constexpr int kTvdRow = 68;   // staged row: 2 + 64 + 2 doubles
template <bool NT, int TJ>
__device__ __forceinline__ void tvd_march3_lds(const Geo& G, int lim, const double* __restrict__ phi,
                                               const double* __restrict__ ux, const double* __restrict__ uy,
                                               const double* __restrict__ uz, double fac, double* __restrict__ out,
                                               int k0, int j0, int i0, int i1, double* T, double* Fy, double* Fz) {
    static_assert(TJ >= 4 && TJ <= 16, "ghost rows are staged by waves 0..3, ghost columns by wave 1");
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int k = k0 + lane, j = j0 + w;
    const int n1 = G.n[1], n2 = G.n[2];
    const long long sI = G.sl[0], sJ = G.sl[1];
    long long lp = (static_cast<long long>(i0 + G.g) * G.np[1] + (j + G.g)) * G.np[2] + (k + G.g);
    const long long fxs = static_cast<long long>(n1) * n2;
    long long fx = static_cast<long long>(i0) * fxs + static_cast<long long>(j) * n2 + k;
    long long fy = (static_cast<long long>(i0) * (n1 + 1) + j) * n2 + k;
    const long long fys = static_cast<long long>(n1 + 1) * n2;
    long long fz = (static_cast<long long>(i0) * n1 + j) * (n2 + 1) + k;
    const long long fzs = static_cast<long long>(n1) * (n2 + 1);
    // ghost row staged by waves 0..3: j0-2, j0-1, j0+TJ, j0+TJ+1 (staged rows 0, 1, TJ+2, TJ+3)
    const bool hload = w < 4;
    const int hj = w < 2 ? w - 2 : TJ + w - 2;                  // tile-relative row
    const long long hoff = static_cast<long long>(hj - w) * sJ;
    const int hrow = hj + 2;
    // ghost columns (wave 1, lanes 0..4 TJ-1): rows j0..j0+TJ-1 x columns k0-2, k0-1, k0+64, k0+65
    const int gr = lane >> 2, gq = lane & 3;
    const int gdk = gq < 2 ? gq - 2 : 62 + gq;                  // -2, -1, 64, 65
    const long long goff = static_cast<long long>(gr - w) * sJ + (gdk - lane);   // from this lane's cell
    const int gslot = (2 + gr) * kTvdRow + (gdk + 2);
    const bool gload = w == 1 && lane < 4 * TJ;
    // right face column (wave 0, lanes 0..TJ-1): row j0+lane, face k0+64
    const long long zx_off = (static_cast<long long>(lane) - w) * (n2 + 1) + (64 - lane);
    double pm2 = phi[lp - 2 * sI], pm1 = phi[lp - sI], p0 = phi[lp], pp1 = phi[lp + sI];
    double f_lo = tvd_flux4(pm2, pm1, p0, pp1, lim, ux[fx]);
    const double A0 = G.A[0], A1 = G.A[1], A2 = G.A[2];
    double* Trow = T + (2 + w) * kTvdRow + 2 + lane;
    for (int i = i0; i < i1; i++) {
        const double pp2 = phi[lp + 2 * sI];
        double u_hi;
        if constexpr (NT) u_hi = ld_nt(ux + fx + fxs);
        else u_hi = ux[fx + fxs];
        const double hal = hload ? phi[lp + hoff] : 0.;
        const double vy = uy[fy], wz = uz[fz];
        const double vyt = w == TJ - 1 ? uy[fy + n2] : 0.;
        const double wzr = (w == 0 && lane < TJ) ? uz[fz + zx_off] : 0.;
        const double gv = gload ? phi[lp + goff] : 0.;
        Trow[0] = p0;
        if (hload) T[hrow * kTvdRow + 2 + lane] = hal;
        if (gload) T[gslot] = gv;
        __syncthreads();
        const double f_hi = tvd_flux4(pm1, p0, pp1, pp2, lim, u_hi);
        const double* Tc = T + w * kTvdRow + 2 + lane;             // column k, rows j-2 ..
        const double gy_lo = tvd_flux4(Tc[0], Tc[kTvdRow], Tc[2 * kTvdRow], Tc[3 * kTvdRow], lim, vy);
        const double* Tr = T + (2 + w) * kTvdRow + lane;           // row j, columns k-2 ..
        const double hz_lo = tvd_flux4(Tr[0], Tr[1], Tr[2], Tr[3], lim, wz);
        Fy[w * 64 + lane] = gy_lo;
        Fz[w * 65 + lane] = hz_lo;
        if (w == TJ - 1) {                                        // top face row j0+TJ
            const double* Tt = T + TJ * kTvdRow + 2 + lane;
            Fy[TJ * 64 + lane] = tvd_flux4(Tt[0], Tt[kTvdRow], Tt[2 * kTvdRow], Tt[3 * kTvdRow], lim, vyt);
        }
        if (w == 0 && lane < TJ) {                                // right face column k0+64
            const double* Te = T + (2 + lane) * kTvdRow + 64;
            Fz[lane * 65 + 64] = tvd_flux4(Te[0], Te[1], Te[2], Te[3], lim, wzr);
        }
        __syncthreads();
        const double gy_hi = Fy[(w + 1) * 64 + lane];
        const double hz_hi = Fz[w * 65 + lane + 1];
        double v = A0 * f_hi;
        v -= A0 * f_lo;
        v += A1 * gy_hi;
        v -= A1 * gy_lo;
        v += A2 * hz_hi;
        v -= A2 * hz_lo;
        put<NT>(out + lp, p0 - fac * v);
        pm2 = pm1;
        pm1 = p0;
        p0 = pp1;
        pp1 = pp2;
        f_lo = f_hi;
        lp += sI;
        fx += fxs;
        fy += fys;
        fz += fzs;
    }
}

// 3-D, one component, persistent and balanced TVD step over 64 (k) x kStepTvdJ (j) tiles.
// Logical block L takes the column-plane range [L*chunk, (L+1)*chunk) (the walk of
// march_columns); a piece of a full tile marches with the LDS-staged form, a piece of an edge
// tile (k or j extent cut by the grid) with the register form, each lane on its own. The
// limiter is a template parameter so that each instantiation carries only its own branch.
constexpr int kStepTvdJ = 8;
template <bool NT, int LIM>
__global__ __launch_bounds__(64 * kStepTvdJ) void k_tvd_step3(Geo G, const double* __restrict__ phi,
                                                              const double* __restrict__ ux,
                                                              const double* __restrict__ uy,
                                                              const double* __restrict__ uz, double fac,
                                                              double* __restrict__ out, int tiles_k, long long total,
                                                              long long chunk, int per_xcd) {
    __shared__ double T[(kStepTvdJ + 4) * kTvdRow];
    __shared__ double Fy[(kStepTvdJ + 1) * 64];
    __shared__ double Fz[kStepTvdJ * 65];
    const long long L = static_cast<long long>(blockIdx.x % kStepXcd) * per_xcd + blockIdx.x / kStepXcd;
    const int n0 = G.n[0];
    long long c = L * chunk;
    const long long cend = (c + chunk < total) ? c + chunk : total;
    while (c < cend) {
        const long long col = c / n0;
        const int i0 = static_cast<int>(c - col * n0);
        const int i1 = (cend - c < n0 - i0) ? i0 + static_cast<int>(cend - c) : n0;
        const int k0 = static_cast<int>(col % tiles_k) * kStepTileK;
        const int j0 = static_cast<int>(col / tiles_k) * kStepTvdJ;
        if (k0 + kStepTileK <= G.n[2] && j0 + kStepTvdJ <= G.n[1]) {      // block-uniform
            tvd_march3_lds<NT, kStepTvdJ>(G, LIM, phi, ux, uy, uz, fac, out, k0, j0, i0, i1, T, Fy, Fz);
        } else {
            const int k = k0 + (threadIdx.x & 63), j = j0 + (threadIdx.x >> 6);
            if (k < G.n[2] && j < G.n[1]) tvd_march3<NT>(G, LIM, phi, ux, uy, uz, fac, out, k, j, i0, i1);
        }
        c += i1 - i0;
    }
}

// 3-D, one component: face gradient (same expressions as k_grad_faces) marching along i with
// phi(i-1) in a register; face i of an axis is (phi(i) - phi(i-1)) dn_r, the high boundary face
// (phi(n) - phi(n-1)) dn_r is written by the last cell.
template <bool NT>
__global__ __launch_bounds__(256) void k_grad3(Geo G, const double* __restrict__ phi, double* __restrict__ fx,
                                               double* __restrict__ fy, double* __restrict__ fz, int tiles_k,
                                               long long total, long long chunk, int per_xcd) {
    march_columns(G, tiles_k, total, chunk, per_xcd, [&](int k, int j, int i0, int i1) {
        const int n0 = G.n[0], n1 = G.n[1], n2 = G.n[2];
        const long long sI = G.sl[0], sJ = G.sl[1];
        const double d0 = G.dnr[0], d1 = G.dnr[1], d2 = G.dnr[2];
        const bool lastj = j == n1 - 1, lastk = k == n2 - 1;
        long long lp = (static_cast<long long>(i0 + G.g) * G.np[1] + (j + G.g)) * G.np[2] + (k + G.g);
        const long long fxs = static_cast<long long>(n1) * n2;
        long long ox = static_cast<long long>(i0) * fxs + static_cast<long long>(j) * n2 + k;
        long long oy = (static_cast<long long>(i0) * (n1 + 1) + j) * n2 + k;
        const long long fys = static_cast<long long>(n1 + 1) * n2;
        long long oz = (static_cast<long long>(i0) * n1 + j) * (n2 + 1) + k;
        const long long fzs = static_cast<long long>(n1) * (n2 + 1);
        double pm = phi[lp - sI];
        for (int i = i0; i < i1; i++) {
            const double p0 = phi[lp], pS = phi[lp - sJ], pB = phi[lp - 1];
            const double pN = phi[lastj ? lp + sJ : lp], pT = phi[lastk ? lp + 1 : lp];
            put<NT>(fx + ox, (p0 - pm) * d0);
            put<NT>(fy + oy, (p0 - pS) * d1);
            if (lastj) put<NT>(fy + oy + n2, (pN - p0) * d1);
            put<NT>(fz + oz, (p0 - pB) * d2);
            if (lastk) put<NT>(fz + oz + 1, (pT - p0) * d2);
            pm = p0;
            lp += sI;
            ox += fxs;
            oy += fys;
            oz += fzs;
        }
        if (i1 == n0) put<NT>(fx + ox, (phi[lp] - pm) * d0);   // lp is now the ghost plane n0
    });
}

// k_grad3 with two k-adjacent cells per lane (n2 even, 16-byte aligned arrays): phi and the x,
// y and z faces move as 16-byte vectors (z-face rows hold n2 + 1 faces, so a row starting on an
// odd face pairs each lane's second face with the next lane's first). Each face is the same
// difference times the same factor as in k_grad3, so the results are bit-identical.
template <bool NT>
__global__ __launch_bounds__(256) void k_grad3_pair(Geo G, const double* __restrict__ phi, double* __restrict__ fx,
                                                    double* __restrict__ fy, double* __restrict__ fz, int tiles_k,
                                                    long long total, long long chunk, int per_xcd) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    auto ld2 = [](const double* p) { return *reinterpret_cast<const d2*>(p); };
    auto st2 = [](double* p, d2 v) {
        if constexpr (NT) __builtin_nontemporal_store(v, reinterpret_cast<d2*>(p));
        else *reinterpret_cast<d2*>(p) = v;
    };
    march_columns<2>(G, tiles_k, total, chunk, per_xcd, [&](int k, int j, int i0, int i1) {
        const int n0 = G.n[0], n1 = G.n[1], n2 = G.n[2];
        const long long sI = G.sl[0], sJ = G.sl[1];
        const double d0 = G.dnr[0], d1 = G.dnr[1], dz2 = G.dnr[2];
        const bool lastj = j == n1 - 1, lastk = k == n2 - 2;     // cell k + 1 is the last one
        long long lp = (static_cast<long long>(i0 + G.g) * G.np[1] + (j + G.g)) * G.np[2] + (k + G.g);
        const long long fxs = static_cast<long long>(n1) * n2;
        long long ox = static_cast<long long>(i0) * fxs + static_cast<long long>(j) * n2 + k;
        long long oy = (static_cast<long long>(i0) * (n1 + 1) + j) * n2 + k;
        const long long fys = static_cast<long long>(n1 + 1) * n2;
        long long oz = (static_cast<long long>(i0) * n1 + j) * (n2 + 1) + k;
        const long long fzs = static_cast<long long>(n1) * (n2 + 1);
        d2 pm = ld2(phi + lp - sI);
        for (int i = i0; i < i1; i++) {
            const d2 p0 = ld2(phi + lp), pS = ld2(phi + lp - sJ);
            const double pB = phi[lp - 1];
            const d2 pN = ld2(phi + (lastj ? lp + sJ : lp));
            const double pT = phi[lastk ? lp + 2 : lp + 1];
            st2(fx + ox, (p0 - pm) * d0);
            st2(fy + oy, (p0 - pS) * d1);
            if (lastj) st2(fy + oy + n2, (pN - p0) * d1);
            const double z0 = (p0.x - pB) * dz2, z1 = (p0.y - p0.x) * dz2;
            if ((oz & 1) == 0) {              // wave-uniform: k is even in every lane
                st2(fz + oz, d2{z0, z1});
                if (lastk) put<NT>(fz + oz + 2, (pT - p0.y) * dz2);
            } else {
                // a row starting on an odd face: the aligned pairs are (k + 1, k + 2), face k + 2
                // coming from the next lane (active whenever this one is not the row's last
                // pair); lane 0 writes its face k alone, lane 63 its face k + 1 alone (the next
                // tile's lane 0 writes k + 2)
                const int lane = threadIdx.x & 63;
                const double zn = __shfl_down(z0, 1);
                if (lane == 0) put<NT>(fz + oz, z0);
                if (lastk) st2(fz + oz + 1, d2{z1, (pT - p0.y) * dz2});
                else if (lane < 63) st2(fz + oz + 1, d2{z1, zn});
                else put<NT>(fz + oz + 1, z1);
            }
            pm = p0;
            lp += sI;
            ox += fxs;
            oy += fys;
            oz += fzs;
        }
        if (i1 == n0) st2(fx + ox, (ld2(phi + lp) - pm) * d0);   // lp is now the ghost plane n0
    });
}

// 3-D, one component: value divergence A0 E - A0 W + A1 N - A1 S + A2 T - A2 B (the order of
// k_div_faces) marching along i with the x face shared by consecutive planes in a register.
template <bool NT>
__global__ __launch_bounds__(256) void k_div3(Geo G, const double* __restrict__ fx, const double* __restrict__ fy,
                                              const double* __restrict__ fz, double* __restrict__ out, int tiles_k,
                                              long long total, long long chunk, int per_xcd) {
    march_columns(G, tiles_k, total, chunk, per_xcd, [&](int k, int j, int i0, int i1) {
        const int n1 = G.n[1], n2 = G.n[2];
        const double A0 = G.A[0], A1 = G.A[1], A2 = G.A[2];
        long long lp = (static_cast<long long>(i0 + G.g) * G.np[1] + (j + G.g)) * G.np[2] + (k + G.g);
        const long long sI = G.sl[0];
        const long long fxs = static_cast<long long>(n1) * n2;
        long long ox = static_cast<long long>(i0) * fxs + static_cast<long long>(j) * n2 + k;
        long long oy = (static_cast<long long>(i0) * (n1 + 1) + j) * n2 + k;
        const long long fys = static_cast<long long>(n1 + 1) * n2;
        long long oz = (static_cast<long long>(i0) * n1 + j) * (n2 + 1) + k;
        const long long fzs = static_cast<long long>(n1) * (n2 + 1);
        double xlo = fx[ox];
        for (int i = i0; i < i1; i++) {
            double xhi;
            if constexpr (NT) xhi = ld_nt(fx + ox + fxs);
            else xhi = fx[ox + fxs];
            const double ylo = fy[oy], yhi = fy[oy + n2], zlo = fz[oz], zhi = fz[oz + 1];
            double v = A0 * xhi;
            v -= A0 * xlo;
            v += A1 * yhi;
            v -= A1 * ylo;
            v += A2 * zhi;
            v -= A2 * zlo;
            put<NT>(out + lp, v);
            xlo = xhi;
            lp += sI;
            ox += fxs;
            oy += fys;
            oz += fzs;
        }
    });
}

__global__ void k_fill(double* __restrict__ p, long long n, double v) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Read + write stream, the attainable-bandwidth probe of the measurements: each lane reads R
// 16-byte chunks (non-temporal, all in flight before the store) and writes their sum, one
// 16-byte chunk; a work-group covers 256 consecutive chunks per read, so every access is
// coalesced. R = 1 is a copy; larger R gives the read-heavy mixes of the solver kernels
// (the FP64 SpMV moves 9 bytes read per byte written).
typedef double dvec2 __attribute__((ext_vector_type(2)));
// LAYOUT 0: XCD x (blocks b = x mod 8) owns one contiguous eighth of the chunks and each chunk
// reads R consecutive 4 KiB pieces of src; 1: the same pieces with the grid striding over all
// chunks; 2: XCD slabs, piece u of a chunk read from the u-th of R separate arrays (src + u *
// n_dst: the per-slot arrays of the diagonal view); 3: layout 2 read-only, each lane adding its
// chunks' sums and writing the total once, to dst[block * 256 + lane].
template <int R, int LAYOUT>
__global__ __launch_bounds__(256) void k_stream_probe(const dvec2* __restrict__ src, dvec2* __restrict__ dst,
                                                      long long n2_dst) {
    const long long nblk = n2_dst / 256;
    long long b0 = blockIdx.x, b1 = nblk, step = gridDim.x;
    if (LAYOUT != 1) {
        const int xcd = static_cast<int>(blockIdx.x % 8);
        b0 = nblk * xcd / 8 + blockIdx.x / 8;
        b1 = nblk * (xcd + 1) / 8;
        step = gridDim.x / 8;
    }
    dvec2 total = {0., 0.};
    for (long long blk = b0; blk < b1; blk += step) {
        const long long j = blk * 256 + threadIdx.x;
        dvec2 v[R];
#pragma unroll
        for (int u = 0; u < R; u++)
            v[u] = __builtin_nontemporal_load(LAYOUT >= 2 ? src + u * n2_dst + j : src + (blk * R + u) * 256 + threadIdx.x);
        dvec2 sum = v[0];
#pragma unroll
        for (int u = 1; u < R; u++) sum += v[u];
        if (LAYOUT == 3) total += sum;
        else __builtin_nontemporal_store(sum, dst + j);
    }
    if (LAYOUT == 3) dst[static_cast<long long>(blockIdx.x) * 256 + threadIdx.x] = total;
}

template <int DIM>
__global__ __launch_bounds__(256) void k_copy_to_field(Geo G, const double* __restrict__ x, double* __restrict__ phi) {
    int q[3];
    if (!cell_coord<DIM>(G, q)) return;
    long long nd = 0;
#pragma unroll
    for (int d = 0; d < DIM; d++) nd += G.si[d] * q[d];
    const long long lp = pad_ord<DIM>(G, q);
    for (int c = 0; c < G.N; c++) phi[lp * G.N + c] = x[nd * G.N + c];
}

// Balanced persistent schedule of the 3-D marching kernels: as many 256-thread blocks as stay
// resident at once (per_cu, queried once per kernel), each with at least kStepMinPlanes planes.
struct MarchPlan {
    int tiles_k, per_xcd;
    long long total, chunk;
    unsigned grid;
};

inline MarchPlan march_plan(const dare_ctx* ctx, const Geo& G, int per_cu, int tile_j = kStepTileJ,
                            int tile_k = kStepTileK) {
    MarchPlan P;
    P.tiles_k = static_cast<int>(ceil_div(G.n[2], tile_k));
    P.total = static_cast<long long>(P.tiles_k) * ceil_div(G.n[1], tile_j) * G.n[0];
    long long nb = static_cast<long long>(ctx->num_cu) * per_cu;
    const long long by_work = ceil_div(P.total, kStepMinPlanes);
    if (nb > by_work) nb = by_work;
    P.per_xcd = static_cast<int>(ceil_div(nb, kStepXcd));
    P.chunk = ceil_div(P.total, static_cast<long long>(P.per_xcd) * kStepXcd);
    P.grid = static_cast<unsigned>(kStepXcd * P.per_xcd);
    return P;
}

template <class K>
int resident_per_cu(K kernel, int threads = 256) {
    int n = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, 0);
    return (e == hipSuccess && n > 0) ? n : 2;
}

template <bool NT, int LIM>
void launch_step3(dare_ctx* ctx, const Geo& G, const double* phi, const double* ux, const double* uy,
                  const double* uz, double fac, double* out) {
    static const int per_cu = resident_per_cu(k_tvd_step3<NT, LIM>, 64 * kStepTvdJ);
    const MarchPlan P = march_plan(ctx, G, per_cu, kStepTvdJ);
    hipLaunchKernelGGL((k_tvd_step3<NT, LIM>), dim3(P.grid), dim3(64 * kStepTvdJ), 0, ctx->stream, G, phi, ux, uy, uz,
                       fac, out, P.tiles_k, P.total, P.chunk, P.per_xcd);
}

template <bool NT>
void launch_grad3(dare_ctx* ctx, const Geo& G, const double* phi, double* fx, double* fy, double* fz) {
    const auto aligned = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    if (G.n[2] % 2 == 0 && G.np[2] % 2 == 0 && G.g % 2 == 0 && aligned(phi) && aligned(fx) && aligned(fy) &&
        aligned(fz)) {
        static const int per_cu2 = resident_per_cu(k_grad3_pair<NT>);
        const MarchPlan P = march_plan(ctx, G, per_cu2, kStepTileJ, 2 * kStepTileK);
        hipLaunchKernelGGL((k_grad3_pair<NT>), dim3(P.grid), dim3(256), 0, ctx->stream, G, phi, fx, fy, fz, P.tiles_k,
                           P.total, P.chunk, P.per_xcd);
        return;
    }
    static const int per_cu = resident_per_cu(k_grad3<NT>);
    const MarchPlan P = march_plan(ctx, G, per_cu);
    hipLaunchKernelGGL((k_grad3<NT>), dim3(P.grid), dim3(256), 0, ctx->stream, G, phi, fx, fy, fz, P.tiles_k, P.total,
                       P.chunk, P.per_xcd);
}

template <bool NT>
void launch_div3(dare_ctx* ctx, const Geo& G, const double* fx, const double* fy, const double* fz, double* out) {
    static const int per_cu = resident_per_cu(k_div3<NT>);
    const MarchPlan P = march_plan(ctx, G, per_cu);
    hipLaunchKernelGGL((k_div3<NT>), dim3(P.grid), dim3(256), 0, ctx->stream, G, fx, fy, fz, out, P.tiles_k, P.total,
                       P.chunk, P.per_xcd);
}

template <bool NT>
void launch_step3_lim(dare_ctx* ctx, const Geo& G, int lim, const double* phi, const double* ux,
                      const double* uy, const double* uz, double fac, double* out) {
    switch (lim) {
    case DARE_LIMITER_CDS: launch_step3<NT, DARE_LIMITER_CDS>(ctx, G, phi, ux, uy, uz, fac, out); break;
    case DARE_LIMITER_UPWIND: launch_step3<NT, DARE_LIMITER_UPWIND>(ctx, G, phi, ux, uy, uz, fac, out); break;
    case DARE_LIMITER_VANALBADA: launch_step3<NT, DARE_LIMITER_VANALBADA>(ctx, G, phi, ux, uy, uz, fac, out); break;
    case DARE_LIMITER_MINMOD: launch_step3<NT, DARE_LIMITER_MINMOD>(ctx, G, phi, ux, uy, uz, fac, out); break;
    default: launch_step3<NT, DARE_LIMITER_VANLEER>(ctx, G, phi, ux, uy, uz, fac, out); break;
    }
}

#define DARE_DISPATCH_DIM(G, KERNEL, L, ...)                                                       \
    do {                                                                                         \
        switch ((G).dim) {                                                                       \
        case 1: hipLaunchKernelGGL(KERNEL<1>, (L).grid, (L).block, 0, ctx->stream, __VA_ARGS__); break; \
        case 2: hipLaunchKernelGGL(KERNEL<2>, (L).grid, (L).block, 0, ctx->stream, __VA_ARGS__); break; \
        default: hipLaunchKernelGGL(KERNEL<3>, (L).grid, (L).block, 0, ctx->stream, __VA_ARGS__); break; \
        }                                                                                        \
    } while (0)

}  // namespace dare_hip

using namespace dare_hip;

extern "C" {

int dare_fill_f64(dare_ctx* ctx, double* p, int64_t n, double v) {
    DARE_CHECK_ARG(ctx && (p || n == 0) && n >= 0, "invalid fill arguments");
    if (n == 0) return DARE_OK;
    hipLaunchKernelGGL(k_fill, dim3(ceil_div(n, 256)), dim3(256), 0, ctx->stream, p, static_cast<long long>(n), v);
    DARE_LAUNCH_CHECK();
    return DARE_OK;
}

int dare_stream_probe(dare_ctx* ctx, int64_t n_dst, const double* src, double* dst, int32_t ratio, int32_t grid,
                      int32_t layout) {
    DARE_CHECK_ARG(ctx && src && dst && n_dst >= 0 && n_dst % 512 == 0,
                   "invalid stream-probe arguments (n_dst a multiple of 512)");
    DARE_CHECK_ARG(ratio == 1 || ratio == 2 || ratio == 4 || ratio == 8, "ratio must be 1, 2, 4 or 8");
    DARE_CHECK_ARG(layout >= 0 && layout <= 3,
                   "layout must be 0 (XCD slabs), 1 (grid stride), 2 (separate arrays) or 3 (read-only)");
    DARE_CHECK_ARG((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) % 16 == 0,
                   "stream probe needs 16-byte aligned vectors");
    if (n_dst == 0) return DARE_OK;
    const long long n2 = n_dst / 2;
    long long gr = grid > 0 ? grid : 8 * ctx->num_cu;
    if (gr > n2 / 256) gr = n2 / 256;
    if (layout != 1) gr = (gr + 7) / 8 * 8;
    DARE_CHECK_ARG(layout != 3 || 512 * gr <= n_dst, "read-only probe: dst holds 512 doubles per work-group");
    const dim3 g(static_cast<unsigned>(gr)), b(256);
    const dvec2* s2 = reinterpret_cast<const dvec2*>(src);
    dvec2* d2 = reinterpret_cast<dvec2*>(dst);
#define DARE_PROBE(R)                                                                                   \
    if (layout == 0) hipLaunchKernelGGL((k_stream_probe<R, 0>), g, b, 0, ctx->stream, s2, d2, n2);     \
    else if (layout == 1) hipLaunchKernelGGL((k_stream_probe<R, 1>), g, b, 0, ctx->stream, s2, d2, n2); \
    else if (layout == 2) hipLaunchKernelGGL((k_stream_probe<R, 2>), g, b, 0, ctx->stream, s2, d2, n2); \
    else hipLaunchKernelGGL((k_stream_probe<R, 3>), g, b, 0, ctx->stream, s2, d2, n2);
    switch (ratio) {
    case 1: DARE_PROBE(1) break;
    case 2: DARE_PROBE(2) break;
    case 4: DARE_PROBE(4) break;
    default: DARE_PROBE(8) break;
    }
#undef DARE_PROBE
    DARE_LAUNCH_CHECK();
    return DARE_OK;
}

int dare_ghost_fill(dare_ctx* ctx, const dare_grid* g, const int32_t* mode6, const double* value6,
                    double* phi) {
    DARE_CHECK_ARG(ctx && g && mode6 && phi, "NULL argument");
    Geo G;
    int rc = make_geo(g, &G);
    if (rc) return rc;
    GhostSlabs S;
    S.start[0] = 0;
    for (int f = 0; f < 6; f++) {
        const int d = f / 2;
        S.mode[f] = f < 2 * G.dim ? mode6[f] : DARE_GHOST_LEAVE;
        S.value[f] = value6 ? value6[f] : 0.;
        DARE_CHECK_ARG(S.mode[f] >= 0 && S.mode[f] <= DARE_GHOST_CONSTANT, "invalid ghost mode");
        DARE_CHECK_ARG(!(S.mode[f] == DARE_GHOST_PERIODIC && !G.per[d]), "periodic ghost fill on a non-periodic axis");
        int k = 0;
        S.other[f][0] = S.other[f][1] = -1;
        long long cnt = 0;
        if (f < 2 * G.dim) {
            cnt = G.g;
            for (int e = 0; e < G.dim; e++)
                if (e != d) {
                    S.other[f][k++] = e;
                    cnt *= G.n[e];
                }
        }
        S.start[f + 1] = S.start[f] + cnt;
    }
    const long long total = S.start[2 * G.dim];
    if (total == 0) return DARE_OK;
    const dim3 grid(ceil_div(total, 256)), block(256);
    switch (G.dim) {
    case 1: hipLaunchKernelGGL(k_ghost_fill<1>, grid, block, 0, ctx->stream, G, S, phi); break;
    case 2: hipLaunchKernelGGL(k_ghost_fill<2>, grid, block, 0, ctx->stream, G, S, phi); break;
    default: hipLaunchKernelGGL(k_ghost_fill<3>, grid, block, 0, ctx->stream, G, S, phi); break;
    }
    DARE_LAUNCH_CHECK();
    return DARE_OK;
}

int dare_grad_faces(dare_ctx* ctx, const dare_grid* g, const double* phi, double* fx, double* fy,
                    double* fz) {
    DARE_CHECK_ARG(ctx && g && phi && fx, "NULL argument");
    Geo G;
    int rc = make_geo(g, &G);
    if (rc) return rc;
    DARE_CHECK_ARG(G.g >= 1, "face gradient needs >= 1 ghost layer");
    DARE_CHECK_ARG(G.dim < 2 || fy, "fy is NULL");
    DARE_CHECK_ARG(G.dim < 3 || fz, "fz is NULL");
    if (G.dim == 3 && G.N == 1) {
        if (G.ncell_int >= kStreamRows) launch_grad3<true>(ctx, G, phi, fx, fy, fz);
        else launch_grad3<false>(ctx, G, phi, fx, fy, fz);
        DARE_LAUNCH_CHECK();
        return DARE_OK;
    }
    const FaceGeo F = make_face_geo(G);
    const Launch L = cell_launch(G);
    DARE_DISPATCH_DIM(G, k_grad_faces, L, G, F, phi, fx, fy, fz);
    DARE_LAUNCH_CHECK();
    return DARE_OK;
}

int dare_interpolate_to_faces(dare_ctx* ctx, const dare_grid* g, const double* ux, const double* uy,
                              const double* uz, double* fx, double* fy, double* fz) {
    DARE_CHECK_ARG(ctx && g && ux && fx, "NULL argument");
    dare_grid g1 = *g;
    g1.ncomp = 1;              // velocity fields carry one component
    Geo G;
    int rc = make_geo(&g1, &G);
    if (rc) return rc;
    DARE_CHECK_ARG(G.g >= 1, "face interpolation needs >= 1 ghost layer");
    DARE_CHECK_ARG(G.dim < 2 || (uy && fy), "y arrays are NULL");
    DARE_CHECK_ARG(G.dim < 3 || (uz && fz), "z arrays are NULL");
    const FaceGeo F = make_face_geo(G);
    const Launch L = cell_launch(G);
    DARE_DISPATCH_DIM(G, k_interp_faces, L, G, F, ux, uy, uz, fx, fy, fz);
    DARE_LAUNCH_CHECK();
    return DARE_OK;
}

int dare_div_faces(dare_ctx* ctx, const dare_grid* g, const double* fx, const double* fy,
                   const double* fz, double* out) {
    DARE_CHECK_ARG(ctx && g && fx && out, "NULL argument");
    Geo G;
    int rc = make_geo(g, &G);
    if (rc) return rc;
    DARE_CHECK_ARG(G.dim < 2 || fy, "fy is NULL");
    DARE_CHECK_ARG(G.dim < 3 || fz, "fz is NULL");
    if (G.dim == 3 && G.N == 1) {
        if (G.ncell_int >= kStreamRows) launch_div3<true>(ctx, G, fx, fy, fz, out);
        else launch_div3<false>(ctx, G, fx, fy, fz, out);
        DARE_LAUNCH_CHECK();
        return DARE_OK;
    }
    const FaceGeo F = make_face_geo(G);
    const Launch L = cell_launch(G);
    DARE_DISPATCH_DIM(G, k_div_faces, L, G, F, fx, fy, fz, out);
    DARE_LAUNCH_CHECK();
    return DARE_OK;
}

int dare_tvd_face_flux(dare_ctx* ctx, const dare_grid* g, int32_t limiter, const double* phi,
                       const double* ux, const double* uy, const double* uz, double* fx,
                       double* fy, double* fz) {
    DARE_CHECK_ARG(ctx && g && phi && ux && fx, "NULL argument");
    Geo G;
    int rc = make_geo(g, &G);
    if (rc) return rc;
    DARE_CHECK_ARG(G.g >= 2, "TVD needs >= 2 ghost layers");
    DARE_CHECK_ARG(limiter >= 0 && limiter <= DARE_LIMITER_VANLEER, "unknown limiter");
    DARE_CHECK_ARG(G.dim < 2 || (uy && fy), "y arrays are NULL");
    DARE_CHECK_ARG(G.dim < 3 || (uz && fz), "z arrays are NULL");
    const FaceGeo F = make_face_geo(G);
    const Launch L = cell_launch(G);
    DARE_DISPATCH_DIM(G, k_tvd_flux, L, G, F, limiter, phi, ux, uy, uz, fx, fy, fz);
    DARE_LAUNCH_CHECK();
    return DARE_OK;
}

int dare_tvd_explicit_step(dare_ctx* ctx, const dare_grid* g, int32_t limiter, const double* phi,
                           const double* ux, const double* uy, const double* uz, double dt,
                           double* phi_out) {
    DARE_CHECK_ARG(ctx && g && phi && ux && phi_out && phi != phi_out, "invalid arguments");
    Geo G;
    int rc = make_geo(g, &G);
    if (rc) return rc;
    DARE_CHECK_ARG(G.g >= 2, "TVD needs >= 2 ghost layers");
    DARE_CHECK_ARG(limiter >= 0 && limiter <= DARE_LIMITER_VANLEER, "unknown limiter");
    DARE_CHECK_ARG(G.dim < 2 || uy, "uy is NULL");
    DARE_CHECK_ARG(G.dim < 3 || uz, "uz is NULL");
    const FaceGeo F = make_face_geo(G);
    const Launch L = cell_launch(G);
    const double fac = dt / G.V;
    if (G.dim == 3 && G.N == 1) {
        if (G.ncell_int >= kStreamRows) launch_step3_lim<true>(ctx, G, limiter, phi, ux, uy, uz, fac, phi_out);
        else launch_step3_lim<false>(ctx, G, limiter, phi, ux, uy, uz, fac, phi_out);
        DARE_LAUNCH_CHECK();
        return DARE_OK;
    }
    DARE_DISPATCH_DIM(G, k_tvd_step, L, G, F, limiter, phi, ux, uy, uz, fac, phi_out);
    DARE_LAUNCH_CHECK();
    return DARE_OK;
}

int dare_copy_to_field(dare_ctx* ctx, const dare_grid* g, const double* x, double* phi) {
    DARE_CHECK_ARG(ctx && g && x && phi, "NULL argument");
    Geo G;
    int rc = make_geo(g, &G);
    if (rc) return rc;
    const Launch L = cell_launch(G);
    DARE_DISPATCH_DIM(G, k_copy_to_field, L, G, x, phi);
    DARE_LAUNCH_CHECK();
    return DARE_OK;
}

}  // extern "C"
