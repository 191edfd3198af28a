// Reference-style functors on the host mirror's operator classes (dare/Operators.h), assembled
// by the generic MatrixBlock path, against the CPU oracle (mode "cpu") and against the device
// assembly of the same terms (mode "gpu"), bit for bit: row pointers, columns, values,
// right-hand side and initial guess.
//
//   CounterDiffusion2D  the functor pattern of test/Verification/counter_diffusion2D.cpp:106-162:
//                       (*mb) = ddt(field); (*mb) -= div(grad(*mb)); per-component Dirichlet /
//                       Neumann edits through Get / GetRhs / Remove, boundary found through
//                       MapLocalToGlobal(MapOrdinalToIndexLocalInternal(ordinal))
//   AdvectionPeriodic2D the pattern of test/Verification/advection_periodic2D.cpp:108-125: one
//                       constant-velocity TVD per component, mb->Add(n, div(tvd * phi).GetSlice(n))
//   General3D           ddt, div(D_f * grad) with the harmonic-mean face diffusivity, TVD from
//                       cell-centred velocity fields (InterpolateToFace), a source, Dirichlet and
//                       Neumann edits
//   Reaction3D          two components with the 2x2 reaction block entered as an explicit
//                       cross-component coefficient (test_TrilinosSolver.cpp:157-164 pattern)
// The components of the two 2-D cases do not couple, so each component's rows are compared
// with the one-component oracle / device operator of that component's terms.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "dare/dare.h"
#include "oracle.h"

using namespace dare;
using CN = dare::Grid::CartesianNeighbor;

static int g_fail = 0;
#define EXPECT(cond, ...)                                      \
    do {                                                       \
        if (!(cond)) {                                         \
            std::printf("  FAIL %s:%d: ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                          \
            std::printf("\n");                                 \
            g_fail++;                                          \
        }                                                      \
    } while (0)

struct Csr {
    std::vector<int32_t> rp, col;
    std::vector<double> val, rhs, x0;
};

static bool same_bits(const std::vector<double>& a, const std::vector<double>& b) {
    return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(double)) == 0);
}

static void expect_equal(const Csr& a, const Csr& b, const char* what) {
    EXPECT(a.rp == b.rp, "%s: row pointers differ", what);
    EXPECT(a.col == b.col, "%s: columns differ", what);
    EXPECT(same_bits(a.val, b.val), "%s: values differ", what);
    EXPECT(same_bits(a.rhs, b.rhs), "%s: right-hand sides differ", what);
    EXPECT(same_bits(a.x0, b.x0), "%s: initial guesses differ", what);
}

// rows node * N + c of an N-component matrix (its columns are nb * N + c) as a 1-component one
static Csr component(const Csr& A, int N, int c) {
    Csr o;
    const int64_t nrows = static_cast<int64_t>(A.rhs.size()) / N;
    o.rp.push_back(0);
    for (int64_t node = 0; node < nrows; node++) {
        const int64_t r = node * N + c;
        for (int32_t k = A.rp[r]; k < A.rp[r + 1]; k++) {
            if (A.col[k] % N != c) std::printf("  component %d couples to another component\n", c), g_fail++;
            o.col.push_back(A.col[k] / N);
            o.val.push_back(A.val[k]);
        }
        o.rp.push_back(static_cast<int32_t>(o.col.size()));
        o.rhs.push_back(A.rhs[r]);
        o.x0.push_back(A.x0[r]);
    }
    return o;
}

template <typename H>
static Csr from_host(const H& h) {
    return Csr{h.row_ptr, h.col, h.val, h.rhs, h.x0};
}

static Csr oracle(const orc_grid& g, const orc_terms& t, const std::vector<double>& phi) {
    const int64_t ncell = orc_num_cells_internal(&g), nrows = ncell * g.ncomp;
    const int64_t cap = nrows * (2 * g.dim + 2);
    Csr o;
    o.rp.resize(nrows + 1);
    o.col.resize(cap);
    o.val.resize(cap);
    o.rhs.resize(nrows);
    o.x0.resize(nrows);
    const int64_t nnz = orc_assemble(&g, &t, phi.data(), o.rp.data(), o.col.data(), o.val.data(), cap, o.rhs.data(),
                                     o.x0.data());
    if (nnz < 0) {
        std::printf("  oracle assembly failed (%lld)\n", static_cast<long long>(nnz));
        g_fail++;
        return Csr{};
    }
    o.col.resize(nnz);
    o.val.resize(nnz);
    return o;
}

template <std::size_t Dim>
static orc_grid ogrid(const std::array<int, Dim>& n, const std::array<double, Dim>& size, int ghost, int ncomp,
                      const std::array<int, Dim>& periodic = {}) {
    orc_grid g{};
    g.dim = static_cast<int32_t>(Dim);
    g.ncomp = ncomp;
    g.ghost = ghost;
    for (int d = 0; d < 3; d++) {
        const bool act = d < static_cast<int>(Dim);
        g.n[d] = act ? n[d] : 1;
        g.periodic[d] = act ? periodic[d] : 0;
        g.size[d] = act ? size[d] : 1.;
    }
    return g;
}

template <typename GV>
static std::vector<double> host_values(const GV& v) {
    std::vector<double> h(v.GetSize());
    for (std::size_t i = 0; i < h.size(); i++) h[i] = v[i];
    return h;
}

template <typename GV1, typename GVN>
static void copy_component(const GVN& src, int N, int c, GV1* dst) {
    for (std::size_t i = 0; i < dst->GetSize(); i++) (*dst)[i] = src[i * N + c];
}

// the device assembly of `terms` on a one-component (or N-component) field, read back
static bool g_gpu = false;
template <typename Grid, std::size_t N>
static Csr device(const typename Grid::Representation& grep, Data::GridVector<Grid, double, N>& field,
                  const Matrix::Terms& terms) {
    Matrix::Trilinos<double> sys;
    sys.Build(grep, field, terms);
    Csr o;
    sys.GetA()->Download(&o.rp, &o.col, &o.val);
    o.rhs = sys.GetB()->getData();
    o.x0 = sys.GetX()->getData();
    return o;
}
template <typename Grid, std::size_t N, typename F>
static Csr generic_device(const typename Grid::Representation& grep, Data::GridVector<Grid, double, N>& field, F f) {
    Matrix::Trilinos<double> sys;
    sys.Build(grep, field, f);
    Csr o;
    sys.GetA()->Download(&o.rp, &o.col, &o.val);
    o.rhs = sys.GetB()->getData();
    o.x0 = sys.GetX()->getData();
    return o;
}

// ---------------------------------------------------------------------------------------------
static void CounterDiffusion2D() {
    using Grid = dare::Grid::Cartesian<2>;
    using SC = double;
    using LO = Grid::LocalOrdinalType;
    using IndexLocal = Grid::Index;
    using IndexGlobal = Grid::IndexGlobal;
    using Field = Data::Field<Grid, SC, 2>;
    const int nx = 17, ny = 13;
    const SC dt = 0.05;
    Grid grid({nx, ny}, {1.0, 0.8}, 2);
    auto grep = grid.GetRepresentation();
    Field field("conc", grep, 2);
    std::mt19937_64 rng(11);
    std::uniform_real_distribution<double> U(0., 1.);
    for (std::size_t i = 0; i < field.GetDataVector(0).GetSize(); i++) {
        const double v = U(rng);
        field.GetDataVector(0)[i] = v;
        field.GetDataVector(1)[i] = v;     // the state after CopyDataVectorsToOldTimeStep
    }
    auto build_coef = [&](auto mblock) {
        using Gradient = dare::Matrix::Gradient<Grid>;
        using TimeScheme = dare::Matrix::EULER_BACKWARD;
        using Divergence = dare::Matrix::Divergence<Grid, TimeScheme>;
        using DDT = dare::Matrix::DDT<Grid>;
        auto g_r = mblock->GetRepresentation();
        LO loc_o{mblock->GetLocalOrdinal()};
        Gradient grad(*g_r, loc_o);
        Divergence div(*g_r, loc_o);
        DDT ddt(*g_r, loc_o, dt);
        (*mblock) = ddt(field);
        (*mblock) -= div(grad(*mblock));
        IndexLocal ind_l = g_r->MapOrdinalToIndexLocalInternal(loc_o);
        IndexGlobal ind_g = g_r->MapLocalToGlobal(ind_l);
        if (ind_g.i() == 0) {                 // W: Dirichlet 1 for component 0, Neumann for 1
            SC bc_0 = 1.;
            mblock->Get(0, CN::CENTER) -= mblock->Get(0, CN::WEST);
            mblock->GetRhs(0) -= 2. * bc_0 * mblock->Get(0, CN::WEST);
            mblock->Remove(0, CN::WEST);
            mblock->Get(1, CN::CENTER) += mblock->Get(1, CN::WEST);
            mblock->Remove(1, CN::WEST);
        } else if (ind_g.i() == (nx - 1)) {  // E: Dirichlet 1 for component 1, Neumann for 0
            SC bc_1 = 1.;
            mblock->Get(1, CN::CENTER) -= mblock->Get(1, CN::EAST);
            mblock->GetRhs(1) -= 2. * bc_1 * mblock->Get(1, CN::EAST);
            mblock->Remove(1, CN::EAST);
            mblock->Get(0, CN::CENTER) += mblock->Get(0, CN::EAST);
            mblock->Remove(0, CN::EAST);
        }
        if (ind_g.j() == 0) {                 // S and N: Neumann for both
            mblock->Get(0, CN::CENTER) += mblock->Get(0, CN::SOUTH);
            mblock->Remove(0, CN::SOUTH);
            mblock->Get(1, CN::CENTER) += mblock->Get(1, CN::SOUTH);
            mblock->Remove(1, CN::SOUTH);
        } else if (ind_g.j() == (ny - 1)) {
            mblock->Get(0, CN::CENTER) += mblock->Get(0, CN::NORTH);
            mblock->Remove(0, CN::NORTH);
            mblock->Get(1, CN::CENTER) += mblock->Get(1, CN::NORTH);
            mblock->Remove(1, CN::NORTH);
        }
    };
    const Csr host = from_host(Matrix::Trilinos<SC>::AssembleHost(grep, field.GetDataVector(), build_coef));
    const int bc[2][4] = {{DARE_BC_DIRICHLET, DARE_BC_NEUMANN, DARE_BC_NEUMANN, DARE_BC_NEUMANN},
                          {DARE_BC_NEUMANN, DARE_BC_DIRICHLET, DARE_BC_NEUMANN, DARE_BC_NEUMANN}};
    const double bv[2][4] = {{1., 0., 0., 0.}, {0., 1., 0., 0.}};
    for (int c = 0; c < 2; c++) {
        const Csr mine = component(host, 2, c);
        Data::GridVector<Grid, SC, 1> phi_c("phi_c", grep);
        copy_component(field.GetDataVector(0), 2, c, &phi_c);
        if (!g_gpu) {
            orc_terms t{};
            t.dt = dt;
            t.diff_const = 1.;
            t.conv_limiter = -1;
            for (int f = 0; f < 4; f++) {
                t.bc_type[f] = bc[c][f];
                t.bc_value[f] = bv[c][f];
            }
            expect_equal(mine, oracle(ogrid<2>({nx, ny}, {1.0, 0.8}, 2, 1), t, host_values(phi_c)),
                         c ? "counter-diffusion component 1 vs oracle" : "counter-diffusion component 0 vs oracle");
        } else {
            Matrix::Terms terms;
            terms.DDT(dt).Diffusion(1.);
            for (int f = 0; f < 4; f++) {
                if (bc[c][f] == DARE_BC_DIRICHLET) terms.Dirichlet(f, bv[c][f]);
                else terms.Neumann(f);
            }
            expect_equal(mine, device(grep, phi_c, terms),
                         c ? "counter-diffusion component 1 vs device" : "counter-diffusion component 0 vs device");
        }
    }
    if (g_gpu) expect_equal(generic_device(grep, field.GetDataVector(), build_coef), host, "generic Build upload");
}

// ---------------------------------------------------------------------------------------------
template <typename SCHEME, int LIMITER>
static void AdvectionPeriodic2DWith(const char* name) {
    using Grid = dare::Grid::Cartesian<2>;
    using SC = double;
    using LO = Grid::LocalOrdinalType;
    using VecSC = Grid::VecSC;
    using VecLO = Grid::VecLO;
    const int nx = 16, ny = 12;
    const SC dt = 1e-3;
    VecLO periodic(1, 1);
    VecSC velocity_0(0.7, 0.), velocity_1(0., -1.3);
    Grid grid({nx, ny}, {1.0, 1.5}, 2, periodic);
    auto grep = grid.GetRepresentation();
    Data::Field<Grid, SC, 2> field("dancers", grep, 2);
    std::mt19937_64 rng(12);
    std::uniform_real_distribution<double> U(0., 1.);
    for (std::size_t i = 0; i < field.GetDataVector(0).GetSize(); i++) {
        const double v = (U(rng) < 0.3) ? 1. : U(rng);      // flat patches: 0/0 limiter ratios
        field.GetDataVector(0)[i] = v;
        field.GetDataVector(1)[i] = v;
    }
    auto build_coef = [&](auto mblock) {
        using TimeScheme = dare::Matrix::EULER_BACKWARD;
        using Divergence = dare::Matrix::Divergence<Grid, TimeScheme>;
        using TVD = dare::Matrix::TVD<Grid, SC, SCHEME>;
        using DDT = dare::Matrix::DDT<Grid>;
        auto g_r = mblock->GetRepresentation();
        LO loc_o{mblock->GetLocalOrdinal()};
        TVD u_0(*g_r, loc_o, velocity_0);
        TVD u_1(*g_r, loc_o, velocity_1);
        Divergence div(*g_r, loc_o);
        DDT ddt(*g_r, loc_o, dt);
        (*mblock) = ddt(field);
        mblock->Add(0, div(u_0 * field.GetDataVector()).GetSlice(0));
        mblock->Add(1, div(u_1 * field.GetDataVector()).GetSlice(1));
    };
    const Csr host = from_host(Matrix::Trilinos<SC>::AssembleHost(grep, field.GetDataVector(), build_coef));
    const VecSC vel[2] = {velocity_0, velocity_1};
    for (int c = 0; c < 2; c++) {
        const Csr mine = component(host, 2, c);
        Data::GridVector<Grid, SC, 1> phi_c("phi_c", grep);
        copy_component(field.GetDataVector(0), 2, c, &phi_c);
        const std::size_t nfx = static_cast<std::size_t>(nx + 1) * ny, nfy = static_cast<std::size_t>(nx) * (ny + 1);
        std::vector<double> fx(nfx, vel[c][0]), fy(nfy, vel[c][1]);
        const std::string what = std::string(name) + (c ? " component 1" : " component 0");
        if (!g_gpu) {
            orc_terms t{};
            t.dt = dt;
            t.conv_limiter = LIMITER;
            t.vel[0] = fx.data();
            t.vel[1] = fy.data();
            expect_equal(mine, oracle(ogrid<2>({nx, ny}, {1.0, 1.5}, 2, 1, {1, 1}), t, host_values(phi_c)),
                         (what + " vs oracle").c_str());
        } else {
            auto dev = Device::Default();
            DeviceBuffer<double> dx(dev, nfx), dy(dev, nfy);
            dx.Upload(fx.data());
            dy.Upload(fy.data());
            Matrix::Terms terms;
            terms.DDT(dt).TVD(LIMITER, dx.data(), dy.data(), nullptr);
            expect_equal(mine, device(grep, phi_c, terms), (what + " vs device").c_str());
        }
    }
}
static void AdvectionPeriodic2D() {
    AdvectionPeriodic2DWith<Matrix::MINMOD, DARE_LIMITER_MINMOD>("periodic advection (MINMOD)");
    AdvectionPeriodic2DWith<Matrix::VANALBADA, DARE_LIMITER_VANALBADA>("periodic advection (VANALBADA)");
    AdvectionPeriodic2DWith<Matrix::VANLEER, DARE_LIMITER_VANLEER>("periodic advection (VANLEER)");
}

// ---------------------------------------------------------------------------------------------
// DDT + harmonic-mean diffusion + VANLEER convection + source under the divergence time scheme
// SCHEME (Divergence<Grid, SCHEME> for both divergences)
template <typename SCHEME>
static void General3DWith(const char* name) {
    using Grid = dare::Grid::Cartesian<3>;
    using SC = double;
    using LO = Grid::LocalOrdinalType;
    using Index = Grid::Index;
    using GV1 = Data::GridVector<Grid, SC, 1>;
    const int n0 = 7, n1 = 6, n2 = 5;
    const SC dt = 5e-3;
    Grid grid({n0, n1, n2}, {1.0, 0.8, 1.2}, 2);
    auto grep = grid.GetRepresentation();
    Data::Field<Grid, SC, 1> phi("phi", grep, 2);
    GV1 D("D", grep), src("s", grep), ux("ux", grep), uy("uy", grep), uz("uz", grep);
    std::mt19937_64 rng(13);
    std::uniform_real_distribution<double> U(0., 1.), V(-1., 1.);
    std::normal_distribution<double> Z(std::log(1e-3), 1.0);
    for (std::size_t i = 0; i < D.GetSize(); i++) {
        const double p = (U(rng) < 0.2) ? 1. : U(rng);
        phi.GetDataVector(0)[i] = p;
        phi.GetDataVector(1)[i] = p;
        D[i] = std::exp(Z(rng));
        src[i] = U(rng);
        ux[i] = V(rng);
        uy[i] = V(rng);
        uz[i] = V(rng);
    }
    const int bc[6] = {DARE_BC_DIRICHLET, DARE_BC_NEUMANN, DARE_BC_DIRICHLET, DARE_BC_DIRICHLET, DARE_BC_NEUMANN,
                       DARE_BC_DIRICHLET};
    const double bv[6] = {0.3, 0., -1., 0.1, 0., 0.25};
    const SC volume = grep.GetCellVolume();
    auto functor = [&](auto mb) {
        using Gradient = Matrix::Gradient<Grid>;
        using Divergence = Matrix::Divergence<Grid, SCHEME>;
        using TVD = Matrix::TVD<Grid, SC, Matrix::VANLEER>;
        using DDT = Matrix::DDT<Grid>;
        auto g_r = mb->GetRepresentation();
        const LO loc = mb->GetLocalOrdinal();
        Gradient grad(*g_r, loc);
        Divergence div(*g_r, loc);
        DDT ddt(*g_r, loc, dt);
        TVD tvd(*g_r, loc, utils::Vector<3, const GV1*>(&ux, &uy, &uz));
        const Index ind_i = g_r->MapOrdinalToIndexLocalInternal(loc);
        const Index ind = g_r->MapInternalToLocal(ind_i);
        // harmonic-mean face diffusivity ((2 D_lo) D_hi) / (D_lo + D_hi)
        Data::FaceValueStencil<Grid, SC, 1> Df;
        for (std::size_t d = 0; d < 3; d++)
            for (int h = 0; h < 2; h++) {
                Index q = ind;
                q[d] += h ? 1 : -1;
                const SC Dp = D.At(ind, 0), Dq = D.At(q, 0);
                const SC Dlo = h ? Dp : Dq, Dhi = h ? Dq : Dp;
                Df.SetValue(dare::Grid::ToCartesianNeighbor(2 * d + 1 + h), 0, ((2. * Dlo) * Dhi) / (Dlo + Dhi));
            }
        (*mb) = ddt(phi);
        (*mb) -= div(Df * grad(*mb));
        (*mb) += div(tvd * phi.GetDataVector());
        mb->GetRhs(0) += volume * src.At(ind, 0);
        const int nn[3] = {n0, n1, n2};
        for (int d = 0; d < 3; d++) {          // one edited face per axis, low face first
            int f = -1;
            if (ind_i[d] == 0) f = 2 * d;
            else if (ind_i[d] == nn[d] - 1) f = 2 * d + 1;
            if (f < 0) continue;
            const CN nb = dare::Grid::ToCartesianNeighbor(f + 1);
            if (bc[f] == DARE_BC_DIRICHLET) {
                mb->Get(0, CN::CENTER) -= mb->Get(0, nb);
                mb->GetRhs(0) -= 2. * bv[f] * mb->Get(0, nb);
            } else {
                mb->Get(0, CN::CENTER) += mb->Get(0, nb);
            }
            mb->Remove(0, nb);
        }
    };
    const Csr host = from_host(Matrix::Trilinos<SC>::AssembleHost(grep, phi.GetDataVector(), functor));
    if (!g_gpu) {
        const orc_grid g = ogrid<3>({n0, n1, n2}, {1.0, 0.8, 1.2}, 2, 1);
        std::vector<double> f[3];
        for (int d = 0; d < 3; d++) f[d].resize(orc_num_faces(&g, d));
        const std::vector<double> hx = host_values(ux), hy = host_values(uy), hz = host_values(uz);
        orc_interpolate_to_faces(&g, hx.data(), hy.data(), hz.data(), f[0].data(), f[1].data(), f[2].data());
        const std::vector<double> hD = host_values(D), hs = host_values(src);
        orc_terms t{};
        t.dt = dt;
        t.diff_cell = hD.data();
        t.conv_limiter = ORC_VANLEER;
        for (int d = 0; d < 3; d++) t.vel[d] = f[d].data();
        t.source = hs.data();
        for (int k = 0; k < 6; k++) {
            t.bc_type[k] = bc[k];
            t.bc_value[k] = bv[k];
        }
        const auto w = SCHEME::template GetWeights<double>();
        t.div_levels = static_cast<int32_t>(w.size());
        for (std::size_t l = 0; l < w.size(); l++) t.div_weights[l] = w[l];
        expect_equal(host, oracle(g, t, host_values(phi.GetDataVector())), (std::string(name) + " vs oracle").c_str());
    } else {
        auto dev = Device::Default();
        const dare_grid g = ux.Descriptor();
        DeviceBuffer<double> f0(dev, (n0 + 1) * n1 * n2), f1(dev, n0 * (n1 + 1) * n2), f2(dev, n0 * n1 * (n2 + 1));
        Check(dare_interpolate_to_faces(dev->Get(), &g, ux.GetDeviceView(), uy.GetDeviceView(), uz.GetDeviceView(),
                                        f0.data(), f1.data(), f2.data()),
              "dare_interpolate_to_faces");
        Matrix::Terms terms;
        terms.DDT(dt).DiffusionField(D.GetDeviceView()).TVD(DARE_LIMITER_VANLEER, f0.data(), f1.data(), f2.data());
        terms.Source(src.GetDeviceView());
        terms.template DivergenceScheme<SCHEME>();
        for (int k = 0; k < 6; k++) {
            if (bc[k] == DARE_BC_DIRICHLET) terms.Dirichlet(k, bv[k]);
            else terms.Neumann(k);
        }
        expect_equal(host, device(grep, phi.GetDataVector(), terms), (std::string(name) + " vs device").c_str());
    }
}
static void General3D() {
    General3DWith<Matrix::EULER_BACKWARD>("general 3-D");
    General3DWith<Matrix::CRANK_NICHOLSON>("general 3-D, CRANK_NICHOLSON divergences");
    General3DWith<Matrix::EULER_FORWARD>("general 3-D, EULER_FORWARD divergences");
    General3DWith<Matrix::ADAMS_BASHFORT>("general 3-D, ADAMS_BASHFORT divergences");
    General3DWith<Matrix::ADAMS_MOULTON>("general 3-D, ADAMS_MOULTON divergences");
}

// ---------------------------------------------------------------------------------------------
static void Reaction3D() {
    using Grid = dare::Grid::Cartesian<3>;
    using SC = double;
    using LO = Grid::LocalOrdinalType;
    using Index = Grid::Index;
    const int n0 = 5, n1 = 4, n2 = 6;
    const SC dt = 0.1;
    const SC Dc[2] = {1e-3, 5e-3};
    Grid grid({n0, n1, n2}, {1.0, 1.0, 1.0}, 1);
    auto grep = grid.GetRepresentation();
    Data::Field<Grid, SC, 2> phi("phi", grep, 2);
    std::mt19937_64 rng(14);
    std::uniform_real_distribution<double> U(0., 1.), K(0.1, 10.);
    for (std::size_t i = 0; i < phi.GetDataVector(0).GetSize(); i++) {
        const double p = U(rng);
        phi.GetDataVector(0)[i] = p;
        phi.GetDataVector(1)[i] = p;
    }
    const int ncell = n0 * n1 * n2;
    std::vector<double> R(4 * ncell);
    for (int c = 0; c < ncell; c++) {
        const double k1 = K(rng), k2 = K(rng);
        R[4 * c + 0] = -k1;
        R[4 * c + 1] = k2;
        R[4 * c + 2] = k1;
        R[4 * c + 3] = -k2;
    }
    const double bv[6] = {0.1, -0.2, 0.3, 0., 0., 0.5};
    const SC volume = grep.GetCellVolume();
    auto functor = [&](auto mb) {
        auto g_r = mb->GetRepresentation();
        const LO loc = mb->GetLocalOrdinal();
        Matrix::Gradient<Grid> grad(*g_r, loc);
        Matrix::Divergence<Grid, Matrix::EULER_BACKWARD> div(*g_r, loc);
        Matrix::DDT<Grid> ddt(*g_r, loc, dt);
        for (std::size_t n = 0; n < 2; n++) mb->Resize(n, 1);     // room for the cross entry
        Data::FaceValueStencil<Grid, SC, 2> Df;
        for (std::size_t n = 0; n < 2; n++) Df.GetData()[n].SetAllValues(Dc[n]);
        (*mb) = ddt(phi);
        (*mb) -= div(Df * grad(*mb));
        const double* r = R.data() + 4 * mb->GetNode();
        for (std::size_t n = 0; n < 2; n++) {
            mb->Get(n, CN::CENTER) -= volume * r[n * 2 + n];
            mb->SetCoefficient(n, mb->GetRow(1 - n), -(volume * r[n * 2 + (1 - n)]));
        }
        const Index ind_i = g_r->MapOrdinalToIndexLocalInternal(loc);
        const int nn[3] = {n0, n1, n2};
        for (int d = 0; d < 3; d++) {
            int f = -1;
            if (ind_i[d] == 0) f = 2 * d;
            else if (ind_i[d] == nn[d] - 1) f = 2 * d + 1;
            if (f < 0) continue;
            const CN nb = dare::Grid::ToCartesianNeighbor(f + 1);
            for (std::size_t n = 0; n < 2; n++) {
                mb->Get(n, CN::CENTER) -= mb->Get(n, nb);
                mb->GetRhs(n) -= 2. * bv[f] * mb->Get(n, nb);
                mb->Remove(n, nb);
            }
        }
    };
    const Csr host = from_host(Matrix::Trilinos<SC>::AssembleHost(grep, phi.GetDataVector(), functor));
    if (!g_gpu) {
        orc_terms t{};
        t.dt = dt;
        t.use_diff_comp = 1;
        t.diff_comp[0] = Dc[0];
        t.diff_comp[1] = Dc[1];
        t.conv_limiter = -1;
        t.react = R.data();
        for (int k = 0; k < 6; k++) {
            t.bc_type[k] = DARE_BC_DIRICHLET;
            t.bc_value[k] = bv[k];
        }
        expect_equal(host, oracle(ogrid<3>({n0, n1, n2}, {1., 1., 1.}, 1, 2), t, host_values(phi.GetDataVector())),
                     "reaction 3-D vs oracle");
    } else {
        auto dev = Device::Default();
        DeviceBuffer<double> Rd(dev, R.size());
        Rd.Upload(R.data());
        Matrix::Terms terms;
        terms.DDT(dt).DiffusionPerComponent({Dc[0], Dc[1]}).Reaction(Rd.data());
        for (int k = 0; k < 6; k++) terms.Dirichlet(k, bv[k]);
        expect_equal(host, device(grep, phi.GetDataVector(), terms), "reaction 3-D vs device");
    }
}

static int run(const char* name, void (*fn)()) {
    const int before = g_fail;
    try {
        fn();
    } catch (const std::exception& e) {
        std::printf("  FAIL exception: %s\n", e.what());
        g_fail++;
    }
    std::printf("[%s] %s\n", g_fail == before ? " OK " : "FAIL", name);
    return g_fail == before;
}

int main(int argc, char** argv) {
    g_gpu = argc > 1 && std::string(argv[1]) == "gpu";
    std::printf("operator classes vs %s\n", g_gpu ? "the device assembly" : "the oracle");
    run("CounterDiffusion2D", CounterDiffusion2D);
    run("AdvectionPeriodic2D", AdvectionPeriodic2D);
    run("General3D", General3D);
    run("Reaction3D", Reaction3D);
    if (g_gpu) Device::ReleaseDefault();
    return g_fail ? 1 : 0;
}
