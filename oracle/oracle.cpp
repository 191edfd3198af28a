/*
 * CPU oracle for the DaRe finite-volume hot path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++ restatement of hpc-dave/dare v0.3.1 (the reference), written so that the
 * floating-point operation order of every stencil and assembly term is the one the
 * reference executes; citations are reference file:line. Only tests/ and smoke() (to check
 * the HIP results against it) and the CPU-baseline leg of bench.py (to time it) use it. The
 * shipped library never does.
 *
 * Parity pins (which check fixes each part; see tests/test_oracle.py):
 *   limiters, Diffusion_N0D1 ........ reference headers compiled into oracle/_ref (exact)
 *   index maps, CSR structure ........ closed forms + test_Trilinos.cpp placement (exact)
 *   gradient/divergence/matrix ....... test_Cartesian_Operators.cpp KATs (exact)
 *   TVD value/matrix (1-D) ........... test_Cartesian_Operators.cpp:967-1501 KATs
 *   TVD 2-D/3-D, van Leer ............ brute-force restatement on tiny grids, closed form
 *   DDT .............................. test_DDT.cpp closed form
 *   ghost fill (periodic) ............ test_HaloBuffer.cpp value rule
 *   BiCGStab2 ........................ closed-form solves (test_TrilinosSolver.cpp,
 *                                      test_integration_Cartesian_Trilinos.cpp), dense LU
 *   harmonic-mean face diffusivity ... series-conductance steady profile (closed form)
 *   2x2 reaction block, source ....... exact per-cell implicit-Euler step, conservation
 *   explicit TVD step, multi-D TVD ... conservation, quadratic-field closed form of r,
 *                                      UPWIND step in numpy, matrix form = flux divergence
 *   convection sign .................. steady upwind advection-diffusion geometric series
 *   ghost modes 2-4 .................. brute force + face mean / zero-jump properties
 *   (tests/test_oracle_pins.py also plants each targeted mistake in a mutant build and
 *    requires its pin to fail)
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int MAXN = 8;   // max components handled by the oracle
static int g_threads = 1;

// ---------------------------------------------------------------- geometry
// CartesianMesh.inl:75-83 (cell width, volume, face area),
// CartesianRepresentation.inl:82-112 (ghost padding, hierarchic strides, k fastest).
struct Geo {
    int dim = 3, N = 1, g = 2;
    int n[3] = {1, 1, 1}, np[3] = {1, 1, 1};
    bool per[3] = {false, false, false};
    bool any_periodic = false;
    double dx[3] = {1, 1, 1}, dnr[3] = {1, 1, 1}, A[3] = {1, 1, 1}, V = 1.;
    int64_t sl[3] = {0, 0, 0}, si[3] = {0, 0, 0};
    int64_t ncell_pad = 1, ncell_int = 1;
};

bool make_geo(const orc_grid* gr, Geo* G) {
    if (!gr || gr->dim < 1 || gr->dim > 3 || gr->ncomp < 1 || gr->ncomp > MAXN || gr->ghost < 0)
        return false;
    G->dim = gr->dim;
    G->N = gr->ncomp;
    G->g = gr->ghost;
    for (int d = 0; d < 3; d++) {
        const bool act = d < G->dim;
        G->n[d] = act ? gr->n[d] : 1;
        if (G->n[d] < 1) return false;
        G->np[d] = act ? G->n[d] + 2 * G->g : 1;
        G->per[d] = act && gr->periodic[d] != 0;
        G->any_periodic |= G->per[d];
    }
    G->V = 1.;
    for (int d = 0; d < G->dim; d++) G->dx[d] = gr->size[d] / static_cast<double>(G->n[d]);
    for (int d = 0; d < G->dim; d++) G->V *= G->dx[d];
    for (int d = 0; d < G->dim; d++) G->A[d] = G->V / G->dx[d];
    for (int d = 0; d < G->dim; d++) {   // Gradient_Cartesian.inl:31-33: 1 / distance
        double e = 1.;
        e /= G->dx[d];
        G->dnr[d] = e;
    }
    G->sl[G->dim - 1] = 1;
    G->si[G->dim - 1] = 1;
    for (int d = G->dim - 2; d >= 0; d--) {
        G->sl[d] = G->sl[d + 1] * G->np[d + 1];
        G->si[d] = G->si[d + 1] * G->n[d + 1];
    }
    G->ncell_pad = 1;
    G->ncell_int = 1;
    for (int d = 0; d < G->dim; d++) {
        G->ncell_pad *= G->np[d];
        G->ncell_int *= G->n[d];
    }
    return true;
}

struct Ind { int64_t v[3] = {0, 0, 0}; };

inline int64_t local_ord(const Geo& G, const Ind& p) {   // MapIndexToOrdinalLocal :373-384
    int64_t o = 0;
    for (int d = 0; d < G.dim; d++) o += G.sl[d] * p.v[d];
    return o;
}

inline Ind internal_index(const Geo& G, int64_t n_int) {  // MapOrdinalToIndexLocalInternal
    Ind p;
    for (int d = 0; d < G.dim; d++) {
        p.v[d] = n_int / G.si[d];
        n_int -= p.v[d] * G.si[d];
    }
    return p;
}

inline Ind internal_to_padded(const Geo& G, Ind p) {      // MapInternalToLocal(Index)
    for (int d = 0; d < G.dim; d++) p.v[d] += G.g;
    return p;
}

// MapIndexToOrdinalGlobalInternal with periodic wrap (:418-450). On one rank global == local.
inline int64_t global_internal_ord(const Geo& G, Ind p, bool map_periodic) {
    if (map_periodic && G.any_periodic) {
        for (int d = 0; d < G.dim; d++) {
            if (p.v[d] < 0) p.v[d] = G.n[d] + p.v[d];
            else if (p.v[d] >= G.n[d]) p.v[d] = p.v[d] - G.n[d];
        }
    }
    int64_t o = 0;
    for (int d = 0; d < G.dim; d++) o += G.si[d] * p.v[d];
    return o;
}

inline double at(const Geo& G, const double* f, const Ind& p, int c) {
    return f[local_ord(G, p) * G.N + c];
}

// face array index of the low face of internal cell q (internal index) along axis ax
inline int64_t face_index(const Geo& G, int ax, const Ind& q) {
    int64_t nf[3];
    for (int e = 0; e < 3; e++) nf[e] = G.n[e] + (e == ax ? 1 : 0);
    int64_t o = 0;
    for (int d = 0; d < G.dim; d++) {
        int64_t s = 1;
        for (int e = d + 1; e < G.dim; e++) s *= nf[e];
        o += s * q.v[d];
    }
    return o;
}

// ---------------------------------------------------------------- limiters
// FluxLimiter.h:41-173. NaN -> 0 for VANALBADA/MINMOD (:129,170); std::min/std::max
// semantics are kept as comparisons. VANLEER is not in the reference: textbook
// (r+|r|)/(1+|r|) with the same NaN -> 0 convention and its limits 2 (r=+inf), 0 (r=-inf).
inline double std_min(double a, double b) { return (b < a) ? b : a; }
inline double std_max(double a, double b) { return (a < b) ? b : a; }

double limiter(int id, double r) {
    switch (id) {
    case ORC_CDS: return 1.;
    case ORC_UPWIND: return 0.;
    case ORC_VANALBADA: {
        r = std::isnan(r) ? 0. : r;
        const double r2 = r * r;
        return (r + r2) / (1. + r2);
    }
    case ORC_MINMOD: {
        r = std::isnan(r) ? 0. : r;
        return std_max(0., std_min(r, 1.));
    }
    case ORC_VANLEER: {
        if (std::isnan(r)) return 0.;
        if (std::isinf(r)) return r > 0 ? 2. : 0.;
        const double a = std::fabs(r);
        return (r + a) / (1. + a);
    }
    default: return 0.;
    }
}

// ---------------------------------------------------------------- stencils
// Faces are numbered pos-1 with CartesianNeighbor W=1,E=2,S=3,N=4,B=5,T=6 (CartesianMesh.h:67-77).
struct FaceVal { double v[MAXN][6]; };
struct FaceMat { double nb[MAXN][6], c[MAXN][6], rhs[MAXN][6]; };
struct CenterMat { double a[MAXN][7]; double rhs[MAXN]; };   // a[n][pos], pos 0 = centre

void zero(FaceVal* s) { std::memset(s, 0, sizeof(*s)); }
void zero(FaceMat* s) { std::memset(s, 0, sizeof(*s)); }
void zero(CenterMat* s) { std::memset(s, 0, sizeof(*s)); }

// Gradient::operator()(GridVector) (Gradient_Cartesian.inl:62-96)
FaceVal grad_values(const Geo& G, const double* phi, const Ind& p) {
    FaceVal s;
    zero(&s);
    for (int c = 0; c < G.N; c++) {
        const double pc = at(G, phi, p, c);
        for (int d = 0; d < G.dim; d++) {
            Ind lo = p, hi = p;
            lo.v[d]--;
            hi.v[d]++;
            s.v[c][2 * d] = (pc - at(G, phi, lo, c)) * G.dnr[d];
            s.v[c][2 * d + 1] = (at(G, phi, hi, c) - pc) * G.dnr[d];
        }
    }
    return s;
}

// Gradient::operator()(MatrixBlock) (Gradient_Cartesian.inl:39-58)
FaceMat grad_matrix(const Geo& G) {
    FaceMat s;
    zero(&s);
    for (int d = 0; d < G.dim; d++) {
        s.nb[0][2 * d] = -G.dnr[d];
        s.c[0][2 * d] = G.dnr[d];
        s.nb[0][2 * d + 1] = G.dnr[d];
        s.c[0][2 * d + 1] = -G.dnr[d];
    }
    for (int c = 1; c < G.N; c++)
        for (int f = 0; f < 6; f++) {
            s.nb[c][f] = s.nb[c - 1][f];
            s.c[c][f] = s.c[c - 1][f];
        }
    return s;
}

// Divergence::ApplyDivergence(FaceValueStencil) (Divergence_Cartesian.inl:211-235)
double div_value(const Geo& G, const FaceVal& s, int c) {
    double v = G.A[0] * s.v[c][1];
    v -= G.A[0] * s.v[c][0];
    if (G.dim > 1) {
        v += G.A[1] * s.v[c][3];
        v -= G.A[1] * s.v[c][2];
    }
    if (G.dim > 2) {
        v += G.A[2] * s.v[c][5];
        v -= G.A[2] * s.v[c][4];
    }
    return v;
}

// Divergence::ApplyDivergence(TFaceMatrixStencil) (Divergence_Cartesian.inl:141-207) for a
// one-level scheme (EULER_BACKWARD, weight 1), the scheme the reference drivers pass.
// The matrix-form divergence of a face stencil under a time scheme with nt levels and weights
// w (Divergence_Cartesian.inl:141-207): the explicit levels carry copies of the implicit
// level's face stencil (:58-66), so the coefficients are scaled by w[0] and the face
// right-hand sides enter once unweighted and once per explicit level t with weight w[t].
CenterMat div_matrix(const Geo& G, const FaceMat& s, const double* w, int nt) {
    const double w0 = w[0];
    CenterMat o;
    zero(&o);
    for (int c = 0; c < G.N; c++) {
        o.a[c][2] = G.A[0] * s.nb[c][1];
        o.a[c][0] = G.A[0] * s.c[c][1];
        o.a[c][1] = -G.A[0] * s.nb[c][0];
        o.a[c][0] -= G.A[0] * s.c[c][0];
        if (G.dim > 1) {
            o.a[c][4] = G.A[1] * s.nb[c][3];
            o.a[c][0] += G.A[1] * s.c[c][3];
            o.a[c][3] = -G.A[1] * s.nb[c][2];
            o.a[c][0] -= G.A[1] * s.c[c][2];
        }
        if (G.dim > 2) {
            o.a[c][6] = G.A[2] * s.nb[c][5];
            o.a[c][0] += G.A[2] * s.c[c][5];
            o.a[c][5] = -G.A[2] * s.nb[c][4];
            o.a[c][0] -= G.A[2] * s.c[c][4];
        }
    }
    for (int c = 0; c < G.N; c++) {   // s_c *= coef_td
        for (int q = 0; q < 7; q++) o.a[c][q] *= w0;
        o.rhs[c] *= w0;
    }
    for (int c = 0; c < G.N; c++) {
        o.rhs[c] = G.A[0] * (s.rhs[c][1] - s.rhs[c][0]);
        if (G.dim > 1) o.rhs[c] += G.A[1] * (s.rhs[c][3] - s.rhs[c][2]);
        if (G.dim > 2) o.rhs[c] += G.A[2] * (s.rhs[c][5] - s.rhs[c][4]);
    }
    for (int t = 1; t < nt; t++) {
        for (int c = 0; c < G.N; c++) {
            o.rhs[c] += w[t] * G.A[0] * (s.rhs[c][1] - s.rhs[c][0]);
            if (G.dim > 1) o.rhs[c] += w[t] * G.A[1] * (s.rhs[c][3] - s.rhs[c][2]);
            if (G.dim > 2) o.rhs[c] += w[t] * G.A[2] * (s.rhs[c][5] - s.rhs[c][4]);
        }
    }
    return o;
}

// Face velocities of a cell from face arrays (staggered source: InterpolateToFace picks the
// face value directly, Interpolation_Cartesian.h:203-219,298-354) and the upwind flags
// (TVD_Cartesian.inl:27-38: upwind = u_f >= 0).
void face_velocities(const Geo& G, const double* const* vel, const Ind& p, double* uf) {
    Ind q;
    for (int d = 0; d < G.dim; d++) q.v[d] = p.v[d] - G.g;
    for (int d = 0; d < G.dim; d++) {
        Ind qh = q;
        qh.v[d]++;
        uf[2 * d] = vel[d][face_index(G, d, q)];
        uf[2 * d + 1] = vel[d][face_index(G, d, qh)];
    }
}

// Upwind / central / downstream cells of one face (TVD_Cartesian.inl:137-143, 162-167).
void tvd_cells(const Ind& p, int d, bool high, bool up, Ind* UU, Ind* U, Ind* D) {
    *UU = *U = *D = p;
    const int64_t vp = up ? 1 : 0, vn = up ? 0 : 1;
    if (!high) {
        UU->v[d] += vn - 2 * vp;
        U->v[d] -= vp;
        D->v[d] -= vn;
    } else {
        UU->v[d] += 2 * vn - vp;
        U->v[d] += vn;
        D->v[d] += vp;
    }
}

// TVD::Interpolate(GridVector) (TVD_Cartesian.inl:122-183)
FaceVal tvd_interpolate(const Geo& G, int lim, const double* phi, const Ind& p, const double* uf) {
    FaceVal s;
    zero(&s);
    for (int d = 0; d < G.dim; d++) {
        for (int h = 0; h < 2; h++) {
            const int f = 2 * d + h;
            Ind UU, U, D;
            tvd_cells(p, d, h == 1, uf[f] >= 0., &UU, &U, &D);
            for (int c = 0; c < G.N; c++) {
                const double pUU = at(G, phi, UU, c), pU = at(G, phi, U, c), pD = at(G, phi, D, c);
                const double r = (pU - pUU) / (pD - pU);
                const double psi = limiter(lim, r);
                s.v[c][f] = pU + (psi * 0.5) * (pD - pU);
            }
        }
    }
    return s;
}

// TVD::operator*(GridVector) (TVD_Cartesian.inl:185-261): implicit upwind part plus the
// deferred-correction face RHS (-0.5 Psi)(phiD - phiU) u.
FaceMat tvd_matrix(const Geo& G, int lim, const double* phi, const Ind& p, const double* uf) {
    FaceMat s;
    zero(&s);
    for (int d = 0; d < G.dim; d++) {
        for (int h = 0; h < 2; h++) {
            const int f = 2 * d + h;
            const bool up = uf[f] >= 0.;
            const double vp = up ? 1. : 0., vn = up ? 0. : 1.;
            const double vel = uf[f];
            Ind UU, U, D;
            tvd_cells(p, d, h == 1, up, &UU, &U, &D);
            for (int c = 0; c < G.N; c++) {
                const double pUU = at(G, phi, UU, c), pU = at(G, phi, U, c), pD = at(G, phi, D, c);
                const double r = (pU - pUU) / (pD - pU);
                const double psi = limiter(lim, r);
                if (h == 0) {
                    s.nb[c][f] = vp * vel;
                    s.c[c][f] = vn * vel;
                } else {
                    s.nb[c][f] = vn * vel;
                    s.c[c][f] = vp * vel;
                }
                s.rhs[c][f] = ((psi * -0.5) * (pD - pU)) * vel;
            }
        }
    }
    return s;
}

// ---------------------------------------------------------------- MatrixBlock
// MatrixBlock_Cartesian.inl:39-593 and MatrixBlockBase.inl:72-295 (one rank, GO ordinals).
struct MB {
    const Geo* G = nullptr;
    Ind ind_int;                       // internal index
    int64_t node = 0;
    double nbv[MAXN][7];               // neighbor coefficients by position (pos 0 = centre)
    uint8_t bits[MAXN];
    double rhs[MAXN], x0[MAXN];
    std::vector<int64_t> ord[MAXN];    // finalized (or explicitly set) column ordinals
    std::vector<double> coef[MAXN];

    void init(const Geo* g, int64_t nd) {
        G = g;
        node = nd;
        ind_int = internal_index(*G, nd);
        std::memset(nbv, 0, sizeof(nbv));
        std::memset(bits, 0, sizeof(bits));
        std::memset(rhs, 0, sizeof(rhs));
        std::memset(x0, 0, sizeof(x0));
        for (int c = 0; c < MAXN; c++) {
            ord[c].clear();
            coef[c].clear();
        }
    }
    int64_t row(int c) const { return node * G->N + c; }      // GetRow :110-112
    double& get(int c, int pos) {                               // Get<CNB> marks the bit
        bits[c] |= static_cast<uint8_t>(1u << pos);
        return nbv[c][pos];
    }
    void remove(int c, int pos) { bits[c] &= static_cast<uint8_t>(~(1u << pos)); }
    bool used_pos(int pos) const { return pos == 0 || (pos + 1) / 2 <= G->dim; }
    void set(const CenterMat& s) {                              // Set(CenterMatrixStencil)
        for (int c = 0; c < G->N; c++) {
            for (int pos = 0; pos < 7; pos++)
                if (used_pos(pos)) get(c, pos) = s.a[c][pos];
            rhs[c] = s.rhs[c];
        }
    }
    void add(const CenterMat& s) {                              // Add(CenterMatrixStencil)
        for (int c = 0; c < G->N; c++) {
            for (int pos = 0; pos < 7; pos++)
                if (used_pos(pos)) get(c, pos) += s.a[c][pos];
            rhs[c] += s.rhs[c];
        }
    }
    void subtract(CenterMat s) {                                // operator-=: Add(-1. * s)
        for (int c = 0; c < G->N; c++) {
            for (int pos = 0; pos < 7; pos++) s.a[c][pos] *= -1.;
            s.rhs[c] *= -1.;
        }
        add(s);
    }
    int64_t num_entries(int c) const {
        int64_t k = 0;
        for (int64_t o : ord[c]) k += o > -1;
        return k;
    }
    void resize(int c, size_t sz) {                             // MatrixBlockBase::Resize
        const size_t old = ord[c].size();
        ord[c].resize(sz);
        coef[c].resize(sz);
        for (size_t i = old; i < sz; i++) {
            ord[c][i] = -1;
            coef[c][i] = 0.;
        }
    }
    void set_coefficient(int c, int64_t col, double v) {        // SetCoefficient :229-254
        for (size_t i = 0; i < ord[c].size(); i++) {
            if (ord[c][i] == col || ord[c][i] == -1) {
                ord[c][i] = col;
                coef[c][i] = v;
                break;
            }
        }
    }
    void set_coefficients(int c, const int64_t* cols, const double* vals, size_t sz) {  // :256-295
        resize(c, static_cast<size_t>(num_entries(c)) + sz);
        const bool empty = ord[c].empty() || ord[c][0] < 0;
        if (empty) {
            for (size_t i = 0; i < sz; i++) {
                ord[c][i] = cols[i];
                coef[c][i] = vals[i];
            }
        } else {
            for (size_t i = 0; i < sz; i++) set_coefficient(c, cols[i], vals[i]);
        }
        resize(c, static_cast<size_t>(num_entries(c)));
    }
    void finalize() {                                           // Finalize :128-240, W,S,B,C,T,N,E
        static const int order[7] = {1, 3, 5, 0, 6, 4, 2};
        for (int c = 0; c < G->N; c++) {
            int64_t cols[7];
            double vals[7];
            size_t k = 0;
            for (int q = 0; q < 7; q++) {
                const int pos = order[q];
                if (!used_pos(pos) || !(bits[c] & (1u << pos))) continue;
                Ind nb = ind_int;
                if (pos > 0) {
                    const int d = (pos - 1) / 2;
                    nb.v[d] += (pos % 2 == 1) ? -1 : 1;
                }
                cols[k] = global_internal_ord(*G, nb, true) * G->N + c;
                vals[k] = nbv[c][pos];
                k++;
            }
            set_coefficients(c, cols, vals, k);
        }
    }
};

// The assembled operator: the per-cell functor a reference user writes for these terms
// (pattern of counter_diffusion2D.cpp:106-162 and advection_periodic2D.cpp:108-125):
//   mb = ddt(field); mb -= div(D_f * grad(mb)); mb += div(tvd * field);
//   rhs += V s;  reaction; boundary edits (Dirichlet mirror / Neumann) + Remove; fixed
//   boundary cells (identity rows) instead of all of it.
void functor(const Geo& G, const orc_terms* t, const double* phi, MB* mb) {
    const Ind p = internal_to_padded(G, mb->ind_int);
    const int64_t lp = local_ord(G, p);
    // a cell on a fixed face (bc type 3; the first axis that has one, low face first) carries
    // the boundary value itself: the functor of the reference's staggered tests
    // (test_integration_Cartesian_Trilinos.cpp:155-220): Get<CENTER> = 1, GetRhs = value,
    // SetInitialGuess(value), nothing else
    for (int d = 0; d < G.dim; d++) {
        int face = -1;
        if (mb->ind_int.v[d] == 0 && t->bc_type[2 * d] == 3) face = 2 * d;
        else if (mb->ind_int.v[d] == G.n[d] - 1 && t->bc_type[2 * d + 1] == 3) face = 2 * d + 1;
        if (face < 0) continue;
        for (int c = 0; c < G.N; c++) {
            mb->get(c, 0) = 1.;
            mb->rhs[c] = t->bc_value[face];
            mb->x0[c] = t->bc_value[face];
        }
        return;
    }
    if (t->react && G.N == 2) {                    // explicit cross-component entries first
        for (int c = 0; c < 2; c++) mb->resize(c, 1);
    }
    // the divergences' time scheme (0 levels: EULER_BACKWARD, weights {1})
    const double euler[1] = {1.};
    const int dn = t->div_levels > 0 ? t->div_levels : 1;
    const double* dw = t->div_levels > 0 ? t->div_weights : euler;
    if (t->dt > 0.) {                              // DDT (DDT.inl:41-81), IMPLICIT weights (1,1)
        CenterMat s;
        zero(&s);
        const double tt0 = (1. * G.V) / t->dt;
        const double tt1 = (1. * G.V) / t->dt;
        for (int c = 0; c < G.N; c++) {
            s.a[c][0] = tt0;
            s.rhs[c] += tt1 * phi[lp * G.N + c];
        }
        mb->set(s);
    }
    {                                              // diffusion: mb -= div(D_f * grad(mb))
        FaceMat gm = grad_matrix(G);
        for (int c = 0; c < G.N; c++) {
            for (int d = 0; d < G.dim; d++) {
                for (int h = 0; h < 2; h++) {
                    const int f = 2 * d + h;
                    double Df;
                    if (t->diff_cell) {                // harmonic mean ((2 Dlo) Dhi) / (Dlo + Dhi)
                        Ind q = p;
                        q.v[d] += h ? 1 : -1;
                        const double Dp = t->diff_cell[lp * G.N + c];
                        const double Dq = t->diff_cell[local_ord(G, q) * G.N + c];
                        const double Dlo = h ? Dp : Dq, Dhi = h ? Dq : Dp;
                        Df = ((2. * Dlo) * Dhi) / (Dlo + Dhi);
                    } else if (t->use_diff_comp) {
                        Df = t->diff_comp[c];
                    } else {
                        Df = t->diff_const;
                    }
                    gm.nb[c][f] *= Df;
                    gm.c[c][f] *= Df;
                }
            }
        }
        mb->subtract(div_matrix(G, gm, dw, dn));
    }
    if (t->conv_limiter >= 0) {                    // implicit TVD convection, deferred correction
        double uf[6] = {0, 0, 0, 0, 0, 0};
        face_velocities(G, t->vel, p, uf);
        mb->add(div_matrix(G, tvd_matrix(G, t->conv_limiter, phi, p, uf), dw, dn));
    }
    if (t->source) {
        for (int c = 0; c < G.N; c++) mb->rhs[c] += G.V * t->source[lp * G.N + c];
    }
    if (t->react && G.N == 2) {                    // cell reaction block R = [[a,b],[c,d]]
        const double* R = t->react + mb->node * 4;
        for (int c = 0; c < 2; c++) {
            mb->get(c, 0) -= G.V * R[c * 2 + c];
            const int o = 1 - c;
            mb->set_coefficient(c, mb->row(o), -(G.V * R[c * 2 + o]));
        }
    }
    for (int d = 0; d < G.dim; d++) {              // boundary edits, counter_diffusion2D.cpp:125-161
        int face = -1;
        if (mb->ind_int.v[d] == 0 && t->bc_type[2 * d] != 0) face = 2 * d;
        else if (mb->ind_int.v[d] == G.n[d] - 1 && t->bc_type[2 * d + 1] != 0) face = 2 * d + 1;
        if (face < 0) continue;
        const int pos = face + 1;
        for (int c = 0; c < G.N; c++) {
            if (t->bc_type[face] == 1) {
                mb->get(c, 0) -= mb->get(c, pos);
                mb->rhs[c] -= (2. * t->bc_value[face]) * mb->get(c, pos);
            } else {
                mb->get(c, 0) += mb->get(c, pos);
            }
            mb->remove(c, pos);
        }
    }
}

// Tpetra fillComplete leaves every local row sorted by column with duplicates merged.
void sorted_row(const MB& mb, int c, std::vector<std::pair<int64_t, double>>* out) {
    out->clear();
    for (size_t i = 0; i < mb.ord[c].size(); i++)
        if (mb.ord[c][i] > -1) out->push_back({mb.ord[c][i], mb.coef[c][i]});
    std::stable_sort(out->begin(), out->end(),
                     [](const auto& a, const auto& b) { return a.first < b.first; });
    size_t w = 0;
    for (size_t i = 0; i < out->size(); i++) {
        if (w > 0 && (*out)[w - 1].first == (*out)[i].first) (*out)[w - 1].second += (*out)[i].second;
        else (*out)[w++] = (*out)[i];
    }
    out->resize(w);
}

// deterministic blocked sum (fixed block order, independent of the thread count)
double blocked_dot(int64_t n, const double* a, const double* b) {
    const int64_t B = 4096;
    const int64_t nb = (n + B - 1) / B;
    std::vector<double> part(static_cast<size_t>(nb > 0 ? nb : 1), 0.);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t k = 0; k < nb; k++) {
        double s = 0.;
        const int64_t e = std::min(n, (k + 1) * B);
        for (int64_t i = k * B; i < e; i++) s += a[i] * b[i];
        part[k] = s;
    }
    double s = 0.;
    for (int64_t k = 0; k < nb; k++) s += part[k];
    return s;
}

void spmv(int64_t nrows, const int32_t* rp, const int32_t* col, const double* val,
          const double* x, double* y) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t r = 0; r < nrows; r++) {
        double s = 0.;
        for (int32_t k = rp[r]; k < rp[r + 1]; k++) s += val[k] * x[col[k]];
        y[r] = s;
    }
}

}  // namespace

extern "C" {

double orc_limiter(int32_t id, double r) { return limiter(id, r); }

// Diffusion_N0D1 (AnalyticalSolutions/Diffusion.h:46-58): phi = 1 - sum_n 2 cos(n pi)/l_n
// exp(-l_n^2 tau) cos(l_n zeta), l_n = (n + 1/2) pi.
double orc_diffusion_n0d1(double tau, double zeta, int32_t n_modes) {
    double tr = 0.;
    for (int n = 0; n < n_modes; ++n) {
        const double l = (n + 0.5) * M_PI;
        tr += 2 * std::cos(n * M_PI) / l * std::exp(-l * l * tau) * std::cos(l * zeta);
    }
    return 1. - tr;
}

void orc_set_threads(int32_t n) { g_threads = n < 1 ? 1 : n; }
int32_t orc_get_threads(void) { return g_threads; }

int64_t orc_num_cells_padded(const orc_grid* gr) {
    Geo G;
    return make_geo(gr, &G) ? G.ncell_pad : -1;
}
int64_t orc_num_cells_internal(const orc_grid* gr) {
    Geo G;
    return make_geo(gr, &G) ? G.ncell_int : -1;
}
int64_t orc_map_internal_to_local(const orc_grid* gr, int64_t n_int) {
    Geo G;
    if (!make_geo(gr, &G)) return -1;
    return local_ord(G, internal_to_padded(G, internal_index(G, n_int)));
}
int64_t orc_num_faces(const orc_grid* gr, int32_t axis) {
    Geo G;
    if (!make_geo(gr, &G) || axis < 0 || axis >= G.dim) return -1;
    int64_t k = 1;
    for (int d = 0; d < G.dim; d++) k *= G.n[d] + (d == axis ? 1 : 0);
    return k;
}
void orc_cell_center(const orc_grid* gr, const int32_t* ind, double* pos) {
    Geo G;
    if (!make_geo(gr, &G)) return;
    for (int d = 0; d < G.dim; d++)   // offset 0 on one rank
        pos[d] = 0. + (ind[d] - G.g + 0.5) * G.dx[d];
}

int orc_grad_faces(const orc_grid* gr, const double* phi, double* fx, double* fy, double* fz) {
    Geo G;
    if (!make_geo(gr, &G) || G.g < 1) return -1;
    double* out[3] = {fx, fy, fz};
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t nd = 0; nd < G.ncell_int; nd++) {
        const Ind q = internal_index(G, nd);
        const Ind p = internal_to_padded(G, q);
        const FaceVal s = grad_values(G, phi, p);
        for (int d = 0; d < G.dim; d++) {
            Ind qh = q;
            qh.v[d]++;
            for (int c = 0; c < G.N; c++) {
                out[d][face_index(G, d, q) * G.N + c] = s.v[c][2 * d];
                if (q.v[d] == G.n[d] - 1) out[d][face_index(G, d, qh) * G.N + c] = s.v[c][2 * d + 1];
            }
        }
    }
    return 0;
}

int orc_div_faces(const orc_grid* gr, const double* fx, const double* fy, const double* fz,
                  double* out) {
    Geo G;
    if (!make_geo(gr, &G)) return -1;
    const double* in[3] = {fx, fy, fz};
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t nd = 0; nd < G.ncell_int; nd++) {
        const Ind q = internal_index(G, nd);
        FaceVal s;
        zero(&s);
        for (int d = 0; d < G.dim; d++) {
            Ind qh = q;
            qh.v[d]++;
            for (int c = 0; c < G.N; c++) {
                s.v[c][2 * d] = in[d][face_index(G, d, q) * G.N + c];
                s.v[c][2 * d + 1] = in[d][face_index(G, d, qh) * G.N + c];
            }
        }
        const int64_t lp = local_ord(G, internal_to_padded(G, q));
        for (int c = 0; c < G.N; c++) out[lp * G.N + c] = div_value(G, s, c);
    }
    return 0;
}

// Face velocities from cell-centred velocity fields (padded layout, one component each):
// math::InterpolateToFace for a cell-centred source and target (Interpolation_Cartesian.h:
// 203-219,298-354) adds the neighbour's value, then the cell's, and multiplies by
// Divisor<2> = 0.5 (Math/Divisors.h:59-79); the TVD constructor of TVD_Cartesian.inl:27-38
// evaluates it for every face. Face m along axis d lies between cells m-1 and m, and both
// cells see (u[m-1] + u[m]) * 0.5 (the sum is commutative), so one value per face.
int orc_interpolate_to_faces(const orc_grid* gr, const double* ux, const double* uy, const double* uz,
                             double* fx, double* fy, double* fz) {
    Geo G;
    if (!make_geo(gr, &G) || G.g < 1) return -1;
    const double* in[3] = {ux, uy, uz};
    double* out[3] = {fx, fy, fz};
    for (int d = 0; d < G.dim; d++) {
        if (!in[d] || !out[d]) return -1;
        Ind q;
        int64_t nf[3] = {1, 1, 1};
        for (int e = 0; e < G.dim; e++) nf[e] = G.n[e] + (e == d ? 1 : 0);
        const int64_t total = nf[0] * nf[1] * nf[2];
        for (int64_t f = 0; f < total; f++) {
            int64_t r = f;
            for (int e = G.dim - 1; e >= 0; e--) {
                q.v[e] = r % nf[e];
                r /= nf[e];
            }
            Ind hi = internal_to_padded(G, q);     // cell m (padded)
            Ind lo = hi;
            lo.v[d]--;                             // cell m - 1
            double v = in[d][local_ord(G, lo)];
            v += in[d][local_ord(G, hi)];
            out[d][face_index(G, d, q)] = v * 0.5;
        }
    }
    return 0;
}

int orc_tvd_face_flux(const orc_grid* gr, int32_t lim, const double* phi,
                      const double* ux, const double* uy, const double* uz,
                      double* fx, double* fy, double* fz) {
    Geo G;
    if (!make_geo(gr, &G) || G.g < 2) return -1;
    const double* vel[3] = {ux, uy, uz};
    double* out[3] = {fx, fy, fz};
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t nd = 0; nd < G.ncell_int; nd++) {
        const Ind q = internal_index(G, nd);
        const Ind p = internal_to_padded(G, q);
        double uf[6] = {0, 0, 0, 0, 0, 0};
        face_velocities(G, vel, p, uf);
        FaceVal s = tvd_interpolate(G, lim, phi, p, uf);
        for (int c = 0; c < G.N; c++)      // face_values *= velocity (Stencils_Cartesian.inl:652)
            for (int f = 0; f < 2 * G.dim; f++) s.v[c][f] *= uf[f];
        for (int d = 0; d < G.dim; d++) {
            Ind qh = q;
            qh.v[d]++;
            for (int c = 0; c < G.N; c++) {
                out[d][face_index(G, d, q) * G.N + c] = s.v[c][2 * d];
                if (q.v[d] == G.n[d] - 1) out[d][face_index(G, d, qh) * G.N + c] = s.v[c][2 * d + 1];
            }
        }
    }
    return 0;
}

int orc_tvd_matrix_cell(const orc_grid* gr, int32_t lim, const double* phi,
                        const double* ux, const double* uy, const double* uz,
                        int64_t cell_int, int32_t comp, double* nb6, double* c6, double* rhs6) {
    Geo G;
    if (!make_geo(gr, &G) || G.g < 2 || comp < 0 || comp >= G.N) return -1;
    const double* vel[3] = {ux, uy, uz};
    const Ind p = internal_to_padded(G, internal_index(G, cell_int));
    double uf[6] = {0, 0, 0, 0, 0, 0};
    face_velocities(G, vel, p, uf);
    const FaceMat s = tvd_matrix(G, lim, phi, p, uf);
    for (int f = 0; f < 6; f++) {
        nb6[f] = s.nb[comp][f];
        c6[f] = s.c[comp][f];
        rhs6[f] = s.rhs[comp][f];
    }
    return 0;
}

int orc_tvd_explicit_step(const orc_grid* gr, int32_t lim, const double* phi,
                          const double* ux, const double* uy, const double* uz,
                          double dt, double* phi_out) {
    Geo G;
    if (!make_geo(gr, &G) || G.g < 2) return -1;
    const double* vel[3] = {ux, uy, uz};
    const double fac = dt / G.V;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t nd = 0; nd < G.ncell_int; nd++) {
        const Ind p = internal_to_padded(G, internal_index(G, nd));
        double uf[6] = {0, 0, 0, 0, 0, 0};
        face_velocities(G, vel, p, uf);
        FaceVal s = tvd_interpolate(G, lim, phi, p, uf);
        for (int c = 0; c < G.N; c++)
            for (int f = 0; f < 2 * G.dim; f++) s.v[c][f] *= uf[f];
        const int64_t lp = local_ord(G, p);
        for (int c = 0; c < G.N; c++) phi_out[lp * G.N + c] = phi[lp * G.N + c] - fac * div_value(G, s, c);
    }
    return 0;
}

// Single-rank HaloBuffer (HaloBuffer.inl:32-208, CartesianRepresentation.inl:117-226): only
// ghost cells outside the internal range along exactly one axis are filled; along a periodic
// axis they take the value of the wrapped internal cell. Modes 2-4 are boundary fills that the
// reference leaves to the user (ghost values enter only via the stencils that read them).
int orc_ghost_fill(const orc_grid* gr, const int32_t* mode, const double* value, double* phi) {
    Geo G;
    if (!make_geo(gr, &G)) return -1;
    for (int f = 0; f < 2 * G.dim; f++)
        if (mode[f] == 1 && !G.per[f / 2]) return -2;
    for (int64_t lp = 0; lp < G.ncell_pad; lp++) {
        Ind p;
        int64_t r = lp;
        for (int d = 0; d < G.dim; d++) {
            p.v[d] = r / G.sl[d];
            r -= p.v[d] * G.sl[d];
        }
        int out_axis = -1, cnt = 0;
        for (int d = 0; d < G.dim; d++) {
            const int64_t i = p.v[d] - G.g;
            if (i < 0 || i >= G.n[d]) {
                cnt++;
                out_axis = d;
            }
        }
        if (cnt != 1) continue;
        const int d = out_axis;
        const int64_t i = p.v[d] - G.g;
        const bool low = i < 0;
        const int f = 2 * d + (low ? 0 : 1);
        const int64_t l = low ? -i : i - (G.n[d] - 1);    // layer 1..g
        Ind src = p;
        if (mode[f] == 1) src.v[d] = G.g + (low ? G.n[d] - l : l - 1);
        else src.v[d] = G.g + (low ? l - 1 : G.n[d] - l);
        for (int c = 0; c < G.N; c++) {
            double* dst = &phi[lp * G.N + c];
            const double s = phi[local_ord(G, src) * G.N + c];
            switch (mode[f]) {
            case 1: *dst = s; break;
            case 2: *dst = 2. * value[f] - s; break;
            case 3: *dst = s; break;
            case 4: *dst = value[f]; break;
            default: break;
            }
        }
    }
    return 0;
}

// Trilinos<SC>::BuildNew (Trilinos.inl:287-385): initial guess from the field, functor per
// cell, Finalize, rows sorted by column (fillComplete). Row = internal ordinal * N + n.
int64_t orc_assemble(const orc_grid* gr, const orc_terms* t, const double* phi_old,
                     int32_t* row_ptr, int32_t* col, double* val, int64_t cap,
                     double* rhs, double* x0) {
    Geo G;
    if (!make_geo(gr, &G) || !t || t->div_levels < 0 || t->div_levels > 3) return -1;
    const int64_t ncell = G.ncell_int;
    const int64_t nrows = ncell * G.N;
    std::vector<int32_t> cnt(static_cast<size_t>(nrows), 0);
    // pass 1: row lengths
#pragma omp parallel num_threads(g_threads)
    {
        MB mb;
        std::vector<std::pair<int64_t, double>> row;
#pragma omp for schedule(static)
        for (int64_t nd = 0; nd < ncell; nd++) {
            mb.init(&G, nd);
            functor(G, t, phi_old, &mb);
            mb.finalize();
            for (int c = 0; c < G.N; c++) {
                sorted_row(mb, c, &row);
                cnt[nd * G.N + c] = static_cast<int32_t>(row.size());
            }
        }
    }
    int64_t nnz = 0;
    row_ptr[0] = 0;
    for (int64_t r = 0; r < nrows; r++) {
        nnz += cnt[r];
        if (nnz > INT32_MAX) return -3;
        row_ptr[r + 1] = static_cast<int32_t>(nnz);
    }
    if (nnz > cap) return -2;
    // pass 2: values
#pragma omp parallel num_threads(g_threads)
    {
        MB mb;
        std::vector<std::pair<int64_t, double>> row;
#pragma omp for schedule(static)
        for (int64_t nd = 0; nd < ncell; nd++) {
            mb.init(&G, nd);
            const int64_t lp = local_ord(G, internal_to_padded(G, mb.ind_int));
            for (int c = 0; c < G.N; c++) mb.x0[c] = phi_old[lp * G.N + c];
            functor(G, t, phi_old, &mb);
            mb.finalize();
            for (int c = 0; c < G.N; c++) {
                const int64_t r = nd * G.N + c;
                sorted_row(mb, c, &row);
                int64_t k = row_ptr[r];
                for (const auto& e : row) {
                    col[k] = static_cast<int32_t>(e.first);
                    val[k] = e.second;
                    k++;
                }
                rhs[r] = mb.rhs[c];
                x0[r] = mb.x0[c];
            }
        }
    }
    return nnz;
}

void orc_spmv(int64_t nrows, const int32_t* rp, const int32_t* col, const double* val,
              const double* x, double* y) {
    spmv(nrows, rp, col, val, x, y);
}

double orc_dot(int64_t n, const double* a, const double* b) { return blocked_dot(n, a, b); }

// BiCGStab2::Solve (BiCGStab2.inl:167-582), unpreconditioned, stop criterion RNORM
// (normalizer 1, :34,288-306). GCR(2) scalars in long double (:186-190).
int orc_bicgstab2(int64_t n, const int32_t* rp, const int32_t* col, const double* val,
                  const double* b, double* x, double tol, int32_t max_it,
                  int32_t* iters, double* resid, double* hist, int32_t hist_cap) {
    std::vector<double> r(n), rh(n), u(n, 0.), v(n, 0.), s(n, 0.), w(n, 0.), t(n, 0.);
    int k = 0;
    double alpha = 0., rho[2] = {0., 0.}, gamma = 0., beta = 0.;
    long double om1 = 0.0L, om2 = 0.0L, mu = 0.0L, nu = 0.0L, tau = 0.0L;
    double conv = 0.;
    const double normalizer = 1.;
    auto axpby = [&](std::vector<double>& y, double a, const std::vector<double>& xx, double bb) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int64_t i = 0; i < n; i++) y[i] = bb * y[i] + a * xx[i];
    };
    // (0) r = b - A x0, r_hat_0 = r
    spmv(n, rp, col, val, x, v.data());
    for (int64_t i = 0; i < n; i++) r[i] = b[i] - v[i];
    rh = r;
    // (1)
    std::fill(u.begin(), u.end(), 0.);
    std::fill(v.begin(), v.end(), 0.);
    std::fill(w.begin(), w.end(), 0.);
    rho[0] = 1.;
    om1 = om2 = 1.0L;
    const double r0 = std::sqrt(blocked_dot(n, r.data(), r.data()));
    conv = r0 / normalizer;
    if (conv < tol) {
        *iters = 0;
        *resid = conv;
        return 1;
    }
    ++k;
    int hk = 0;
    bool ok_exit = false;
    while (true) {
        if (k > max_it) break;
        rho[0] = static_cast<double>(-om2 * rho[0]);                  // (3)
        if (rho[0] == 0.0) break;
        rho[1] = blocked_dot(n, rh.data(), r.data());                 // (4)
        beta = alpha * rho[1] / rho[0];
        rho[0] = rho[1];
        if (rho[0] == 0.0) break;
        axpby(u, 1., r, -beta);                                       // (5) u = r - beta u
        spmv(n, rp, col, val, u.data(), v.data());                    // (6) v = A u
        gamma = blocked_dot(n, v.data(), rh.data());                  // (7)
        if (gamma == 0.0) break;
        alpha = rho[0] / gamma;
        axpby(r, -alpha, v, 1.);                                      // (8)
        spmv(n, rp, col, val, r.data(), s.data());                    // (9) s = A r
        for (int64_t i = 0; i < n; i++) x[i] = 1. * x[i] + alpha * u[i];   // (10)
        rho[1] = blocked_dot(n, rh.data(), s.data());                 // (11)
        beta = alpha * rho[1] / rho[0];
        rho[0] = rho[1];
        axpby(v, 1., s, -beta);                                       // (12) v = s - beta v
        spmv(n, rp, col, val, v.data(), w.data());                    // (13) w = A v
        gamma = blocked_dot(n, w.data(), rh.data());                  // (14)
        if (gamma == 0.0) break;
        alpha = rho[0] / gamma;
        axpby(u, 1., r, -beta);                                       // (15)
        axpby(r, -alpha, v, 1.);                                      // (16)
        axpby(s, -alpha, w, 1.);                                      // (17)
        spmv(n, rp, col, val, s.data(), t.data());                    // (18) t = A s
        om1 = blocked_dot(n, r.data(), s.data());                     // (19)-(20)
        mu = blocked_dot(n, s.data(), s.data());
        nu = blocked_dot(n, s.data(), t.data());
        tau = blocked_dot(n, t.data(), t.data());
        om2 = blocked_dot(n, r.data(), t.data());
        if (mu == 0.) break;
        tau -= nu * nu / mu;                                          // (21)
        if (tau == 0.0) {
            // true residual check (BiCGStab2.inl:514-536)
            std::vector<double> av(n);
            spmv(n, rp, col, val, x, av.data());
            for (int64_t i = 0; i < n; i++) r[i] = b[i] - av[i];
            conv = std::sqrt(blocked_dot(n, r.data(), r.data())) / normalizer;
            if (conv <= tol) {
                ok_exit = true;
                *iters = k;   // the reference returns before storing num_iteration
            }
            break;
        }
        om2 = (om2 - (nu * om1) / mu) / tau;                          // (22)
        om1 = (om1 - nu * om2) / mu;                                  // (23)
        const double o1 = static_cast<double>(om1), o2 = static_cast<double>(om2);
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int64_t i = 0; i < n; i++) {                             // (24)
            x[i] = 1. * x[i] + o1 * r[i] + o2 * s[i];
            x[i] = 1. * x[i] + alpha * u[i];
        }
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int64_t i = 0; i < n; i++) r[i] = 1. * r[i] + (-o1) * s[i] + (-o2) * t[i];   // (25)
        conv = std::sqrt(blocked_dot(n, r.data(), r.data())) / normalizer;
        if (hist && hk < hist_cap) hist[hk++] = conv;
        if (conv <= tol && k > 1) break;
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int64_t i = 0; i < n; i++) u[i] = 1. * u[i] + (-o1) * v[i] + (-o2) * w[i];
        ++k;
    }
    if (ok_exit) {
        *resid = conv;
        return 1;
    }
    *iters = k;
    *resid = conv;
    return conv <= tol ? 1 : 0;
}

void orc_copy_to_field(const orc_grid* gr, const double* x, double* phi) {
    Geo G;
    if (!make_geo(gr, &G)) return;
    for (int64_t nd = 0; nd < G.ncell_int; nd++) {
        const int64_t lp = local_ord(G, internal_to_padded(G, internal_index(G, nd)));
        for (int c = 0; c < G.N; c++) phi[lp * G.N + c] = x[nd * G.N + c];
    }
}

}  // extern "C"
