// Reference-style drivers on the MI355X host mirror (dare_amd/include/dare). Each case restates
// a test of the reference with the same functor and the same closed-form expectation:
//   TrilinosBuild     — MatrixSystem/test/test_Trilinos.cpp:60-191 (placement of A, x, B)
//   SolveLaplace      — MatrixSystem/test/test_TrilinosSolver.cpp:98-205 (BumbleBee/BICGSTAB2)
//   SolveScalar1D     — test/test_integration_Cartesian_Trilinos.cpp:82-160
//   HaloExchange      — MPI/test/test_HaloBuffer.cpp:33-115 (one rank, periodic)
//   DeviceDiffusion3D — counter_diffusion pattern on the device fast path vs the heat kernel
//   VTKWrite          — IO/test/test_VTKWriter.cpp pattern: a field changed on the device
//                       (periodic ghost exchange) written through dare::io::VTKWriter into
//                       argv[1] (checked by tests/test_host_mirror.py)
//   StaggeredStorage  — test/test_integration_Cartesian_GridVector.cpp:104-158 and
//                       test_integration_Cartesian_Field.cpp:84-142 (sizes and names of
//                       staggered representations), plus their cell positions
//   SolveStaggered1D  — test/test_integration_Cartesian_Trilinos.cpp:155-220
//   SolveStaggered2D  — test/test_integration_Cartesian_Trilinos.cpp:410-585 (X and Y)
//   DeviceStaggered   — the face gradient of a y-staggered linear field on the device (the
//                       descriptor of a staggered representation keeps the grid's dx), and a
//                       device-path solve with fixed boundary cells (SolveStaggeredY's profile)
//   SolveJacobi       — SolveLaplace with Ifpack2 RELAXATION/Jacobi as M
//                       (TrilinosSolver.inl:97-175), and the device fast path with none, Jacobi
//                       and MueLu AMG (the aggregation multigrid, TrilinosSolver.inl:180-246)
// Exit code 0 when every case passes.
#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "dare/dare.h"

using namespace dare;
using CN = dare::Grid::CartesianNeighbor;

static int g_fail = 0;
#define EXPECT(cond, ...)                                  \
    do {                                                   \
        if (!(cond)) {                                     \
            std::printf("  FAIL %s:%d: ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                      \
            std::printf("\n");                             \
            g_fail++;                                      \
        }                                                  \
    } while (0)

static void TrilinosBuild() {
    constexpr std::size_t N = 4;
    using G = dare::Grid::Cartesian<1>;
    G grid({23}, {1.}, 1);
    auto g_rep = grid.GetRepresentation();
    Data::GridVector<G, double, N> field("test", g_rep);
    const double offset_rhs = 1.5, offset_init = 0.75, offset_coef = 0.3;
    for (int node = 0; node < 23; node++)
        for (std::size_t i = 0; i < N; i++) field.At(g_rep.MapInternalToLocal(node), i) = node * N + i;
    const int64_t num_rows = 23 * N;
    auto functor = [&](auto mblock) {
        for (std::size_t n = 0; n < N; n++) {
            const int64_t row = mblock->GetRow(n);
            mblock->Resize(n, (row == 0 || row == num_rows - 1) ? 2 : 3);
            if (row > 0) mblock->SetCoefficient(n, row - 1, mblock->GetInitialGuess(n));
            mblock->SetCoefficient(n, row, mblock->GetInitialGuess(n) + offset_coef);
            if (row < num_rows - 1) mblock->SetCoefficient(n, row + 1, mblock->GetInitialGuess(n) + 2. * offset_coef);
            mblock->SetRhs(n, row + offset_rhs);
            mblock->SetInitialGuess(n, row + offset_init);
        }
    };
    Matrix::Trilinos<double> trilinos;
    trilinos.Build(g_rep, field, functor, false);
    std::vector<int32_t> rp, col;
    std::vector<double> val;
    trilinos.GetA()->Download(&rp, &col, &val);
    auto x = trilinos.GetX()->getData();
    auto B = trilinos.GetB()->getData();
    for (int64_t row = 0; row < num_rows; row++) {
        EXPECT(x[row] == row + offset_init, "x[%ld]", static_cast<long>(row));
        EXPECT(B[row] == row + offset_rhs, "B[%ld]", static_cast<long>(row));
        const double f = static_cast<double>(row);   // field value of that row
        std::vector<std::pair<int, double>> exp;
        if (row > 0) exp.push_back({static_cast<int>(row - 1), f});
        exp.push_back({static_cast<int>(row), f + offset_coef});
        if (row < num_rows - 1) exp.push_back({static_cast<int>(row + 1), f + 2. * offset_coef});
        EXPECT(rp[row + 1] - rp[row] == static_cast<int>(exp.size()), "entries of row %ld", static_cast<long>(row));
        for (std::size_t k = 0; k < exp.size() && rp[row] + k < col.size(); k++) {
            EXPECT(col[rp[row] + k] == exp[k].first, "col of row %ld", static_cast<long>(row));
            EXPECT(val[rp[row] + k] == exp[k].second, "value of row %ld", static_cast<long>(row));
        }
    }
}

static void SolveLaplaceWith(Matrix::PreCondPackage pack, const std::string& ptype) {
    constexpr std::size_t N = 4;
    using G = dare::Grid::Cartesian<1>;
    G grid({23}, {1.}, 1);
    auto g_rep = grid.GetRepresentation();
    Data::GridVector<G, double, N> field("test field", g_rep);
    field.SetValues(0.5);
    const int64_t last = 22;
    auto functor = [&](auto mblock) {
        const int64_t node_g = mblock->GetNode();
        const bool left = node_g == 0, right = node_g == last;
        for (std::size_t n = 0; n < N; n++) {
            std::size_t size_stencil = (left || right) ? 4 : 5;
            if (n == 0) size_stencil--;
            if (n == N - 1) size_stencil--;
            mblock->Resize(n, size_stencil);
            const int64_t row = mblock->GetRow(n);
            if (left) {
                mblock->SetCoefficient(n, row, 3.);
                mblock->SetCoefficient(n, row + N, -1.);
                mblock->SetRhs(n, 0.);
                mblock->SetInitialGuess(n, 0.);
            } else if (right) {
                mblock->SetCoefficient(n, row - N, -1.);
                mblock->SetCoefficient(n, row, 3.);
                mblock->SetRhs(n, 2.);
                mblock->SetInitialGuess(n, 1.);
            } else {
                mblock->SetCoefficient(n, row - N, -1.);
                mblock->SetCoefficient(n, row, 2.);
                mblock->SetCoefficient(n, row + N, -1.);
                mblock->SetRhs(n, 0.);
            }
            if (n != 0) {
                mblock->SetCoefficient(n, row - 1, -1.);
                mblock->GetCoefficientByOrdinal(n, row) += 1;
            }
            if (n != N - 1) {
                mblock->SetCoefficient(n, row + 1, -1.);
                mblock->GetCoefficientByOrdinal(n, row) += 1;
            }
        }
    };
    Matrix::Trilinos<double> trilinos;
    trilinos.Build(g_rep, field, functor, false);
    auto params = std::make_shared<ParameterList>();
    params->set("Convergence Tolerance", 1e-12);
    params->set("Maximum Iterations", 5000);
    Matrix::TrilinosSolver<double> solver;
    auto pparams = std::make_shared<ParameterList>();
    pparams->set("relaxation: type", std::string("Jacobi"));
    auto M = solver.BuildPreconditioner(pack, ptype, pparams, trilinos.GetA());
    EXPECT((pack == Matrix::PreCondPackage::None) == !M, "preconditioner %s", ptype.c_str());
    auto ret = solver.Solve(Matrix::SolverPackage::BumbleBee, "BICGSTAB2", M, trilinos.GetA(), trilinos.GetX(),
                            trilinos.GetB(), params);
    EXPECT(ret == Matrix::ReturnType::Converged, "SolveLaplace did not converge");
    auto x = trilinos.GetX()->getData();
    for (int node = 0; node < 23; node++)
        for (std::size_t n = 0; n < N; n++) {
            const double phi = (node + 0.5) / 23.;
            EXPECT(std::fabs(phi - x[node * N + n]) < 1e-8, "node %d comp %zu: %g vs %g", node, n, x[node * N + n], phi);
        }
    std::printf("  SolveLaplace (%s): %d passes\n", ptype.c_str(), solver.GetNumIterations());
}

static void SolveLaplace() { SolveLaplaceWith(Matrix::PreCondPackage::None, "none"); }

// Ifpack2 RELAXATION / Jacobi as the right preconditioner M of TrilinosSolver::Solve
// (TrilinosSolver.inl:71-127): on the generic host-functor matrix (scaled CSR values) and on
// the device fast path (scaled diagonal view); same solutions as without it
static void SolveJacobi() {
    SolveLaplaceWith(Matrix::PreCondPackage::Ifpack2, "RELAXATION");
    Matrix::TrilinosSolver<double> solver;
    bool threw = false;
    try {
        solver.BuildPreconditioner(Matrix::PreCondPackage::Ifpack2, "ILUT", nullptr, nullptr);
    } catch (const std::exception&) {
        threw = true;
    }
    EXPECT(threw, "an unknown preconditioner type must be refused");
    using G = dare::Grid::Cartesian<3>;
    const int n = 20;
    G grid({n, n + 2, n + 4}, {1., 1.1, 1.2}, 2);
    auto grep = grid.GetRepresentation();
    Data::GridVector<G, double, 1> phi("phi", grep);
    const auto res = grep.GetLocalResolution();
    for (int i = 0; i < res[0]; i++)
        for (int j = 0; j < res[1]; j++)
            for (int k = 0; k < res[2]; k++) phi.At({i, j, k}, 0) = std::sin(0.3 * i) + std::cos(0.2 * j * k);
    Matrix::Terms terms;
    terms.DDT(0.05).Diffusion(1e-2);
    for (int f = 0; f < 6; f++) terms.Dirichlet(f, 0.25 * f);
    auto params = std::make_shared<ParameterList>();
    params->set("Maximum Iterations", 500);
    // none, Ifpack2 RELAXATION/Jacobi, MueLu AMG (the aggregation multigrid, with the
    // reference test's MueLu parameters)
    auto amg_param = std::make_shared<ParameterList>();
    amg_param->set("problem: type", std::string("MHD"));
    amg_param->set("verbosity", std::string("none"));
    std::vector<double> sol[3];
    int passes[3];
    for (int use = 0; use < 3; use++) {
        Matrix::Trilinos<double> msystem;
        msystem.Build(grep, phi, terms, false);
        params->set("Convergence Tolerance", 1e-13 * msystem.GetB()->norm2());
        std::shared_ptr<Matrix::Preconditioner> M;
        if (use == 1) M = solver.BuildPreconditioner(Matrix::PreCondPackage::Ifpack2, "RELAXATION", nullptr, msystem.GetA());
        if (use == 2) M = solver.BuildPreconditioner(Matrix::PreCondPackage::MueLu, "AMG", amg_param, msystem.GetA());
        auto ret = solver.Solve(Matrix::SolverPackage::BumbleBee, "BICGSTAB2", M, msystem.GetA(), msystem.GetX(),
                                msystem.GetB(), params);
        EXPECT(ret == Matrix::ReturnType::Converged, "device solve %d did not converge", use);
        sol[use] = msystem.GetX()->getData();
        passes[use] = solver.GetNumIterations();
    }
    double err[3] = {0., 0., 0.}, mx = 0.;
    for (std::size_t r = 0; r < sol[0].size(); r++) {
        for (int u = 1; u < 3; u++) err[u] = std::fmax(err[u], std::fabs(sol[0][r] - sol[u][r]));
        mx = std::fmax(mx, std::fabs(sol[0][r]));
    }
    EXPECT(err[1] <= 1e-10 * mx, "Jacobi solution differs by %g (max %g)", err[1], mx);
    EXPECT(err[2] <= 1e-10 * mx, "multigrid solution differs by %g (max %g)", err[2], mx);
    // constant diffusivity: diag(A) is nearly a multiple of I, so Jacobi saves little here
    EXPECT(passes[2] < passes[1] && passes[1] <= passes[0], "passes none %d, Jacobi %d, AMG %d", passes[0],
           passes[1], passes[2]);
    std::printf("  SolveJacobi device path: %d passes plain, %d with Jacobi, %d with AMG, difference %.1e / %.1e\n",
                passes[0], passes[1], passes[2], err[1] / mx, err[2] / mx);
}

static void SolveScalar1D() {
    constexpr std::size_t N = 3;
    using G = dare::Grid::Cartesian<1>;
    G grid({20}, {1.}, 2);
    auto g_rep = grid.GetRepresentation();
    Data::GridVector<G, double, N> data("test", g_rep);
    const double value_west = 0., value_east = 1.;
    auto functor = [&](auto mblock) {
        const int64_t node_g = mblock->GetNode();
        const bool west = node_g == 0, east = node_g == g_rep.GetGlobalResolutionInternal()[0] - 1;
        for (std::size_t n = 0; n < mblock->GetNumComponents(); n++) {
            mblock->GetRhs(n) = 0.;
            mblock->template Get<CN::CENTER>(n) = 2.;
            if (!west) {
                mblock->template Get<CN::WEST>(n) = -1.;
            } else {
                mblock->template Get<CN::CENTER>(n) += 1.;
                mblock->GetRhs(n) = 2. * value_west;
                mblock->SetInitialGuess(n, value_west);
            }
            if (!east) {
                mblock->template Get<CN::EAST>(n) = -1.;
            } else {
                mblock->template Get<CN::CENTER>(n) += 1.;
                mblock->GetRhs(n) = 2. * value_east;
                mblock->SetInitialGuess(n, value_east);
            }
        }
        mblock->Finalize();
    };
    Matrix::Trilinos<double> trilinos;
    Matrix::TrilinosSolver<double> solver;
    auto params = std::make_shared<ParameterList>();
    params->set("Convergence Tolerance", 1e-14);
    params->set("Maximum Iterations", 5000);
    params->set("Verbosity", "high");
    trilinos.Build(g_rep, data, functor, false);
    auto ret = solver.Solve(Matrix::SolverPackage::BumbleBee, "BICGSTAB2", trilinos.GetA(), trilinos.GetX(),
                            trilinos.GetB(), params);
    EXPECT(ret == Matrix::ReturnType::Converged, "SolveScalar1D did not converge");
    trilinos.CopyTo(&data);
    const double d_phi = (value_east - value_west) / 20.;
    for (int i = 0; i < 20; i++)
        for (std::size_t n = 0; n < N; n++) {
            const double phi = value_west + (i + 0.5) * d_phi;
            const double v = data.At(g_rep.MapInternalToLocal(i), n);
            EXPECT(std::fabs(phi - v) < 1e-8, "cell %d: %g vs %g", i, v, phi);
        }
}

static void HaloExchange() {
    constexpr std::size_t N = 5;
    using G = dare::Grid::Cartesian<1>;
    G grid({2}, {1.}, 1, {1});
    auto g_rep = grid.GetRepresentation();
    Data::Field<G, double, N> field("f", g_rep, 1);
    auto& v = field.GetDataVector();
    v.SetValues(-1.);
    for (int n = 1; n < 3; n++)
        for (std::size_t e = 0; e < N; e++) v.At(n, e) = n * N + e;
    field.ExchangeHaloCells();
    v.GetDevice().Synchronize();
    v.MarkDeviceModified();
    v.SyncToHost();
    for (std::size_t e = 0; e < N; e++) {
        EXPECT(std::fabs(v.At(0, e) - (2 * N + e)) < 2.3e-16, "low ghost");
        EXPECT(std::fabs(v.At(3, e) - (1 * N + e)) < 2.3e-16, "high ghost");
    }
}

static void DeviceDiffusion3D() {
    using G = dare::Grid::Cartesian<3>;
    const int n = 24;
    G grid({n, n, n}, {1., 1., 1.}, 2);
    auto grep = grid.GetRepresentation();
    Data::Field<G, double, 1> field("phi", grep, 2);
    auto& phi = field.GetDataVector();
    const double sig = 0.1, D = 1e-3, dt = 1e-2;
    for (int i = 0; i < n + 4; i++)
        for (int j = 0; j < n + 4; j++)
            for (int k = 0; k < n + 4; k++) {
                auto c = grep.GetCoordinatesCenter({i, j, k});
                const double r2 = (c[0] - .5) * (c[0] - .5) + (c[1] - .5) * (c[1] - .5) + (c[2] - .5) * (c[2] - .5);
                phi.At({i, j, k}, 0) = std::exp(-r2 / (2 * sig * sig));
            }
    Matrix::Terms terms;
    terms.DDT(dt).Diffusion(D);
    for (int f = 0; f < 6; f++) terms.Dirichlet(f, 0.);
    Matrix::Trilinos<double> msystem;
    Matrix::TrilinosSolver<double> solver;
    auto params = std::make_shared<ParameterList>();
    params->set("Maximum Iterations", 200);
    for (int step = 0; step < 10; step++) {
        msystem.Build(grep, phi, terms, false);
        params->set("Convergence Tolerance", 1e-10 * msystem.GetB()->norm2());
        auto ret = solver.Solve(Matrix::SolverPackage::BumbleBee, "BICGSTAB2", msystem.GetA(), msystem.GetX(),
                                msystem.GetB(), params);
        EXPECT(ret == Matrix::ReturnType::Converged, "step %d did not converge", step);
        msystem.CopyTo(&phi);
    }
    const double T = 10 * dt, s2 = sig * sig + 2 * D * T, amp = std::pow(sig * sig / s2, 1.5);
    double err = 0.;
    for (int i = 2; i < n + 2; i++)
        for (int j = 2; j < n + 2; j++)
            for (int k = 2; k < n + 2; k++) {
                auto c = grep.GetCoordinatesCenter({i, j, k});
                const double r2 = (c[0] - .5) * (c[0] - .5) + (c[1] - .5) * (c[1] - .5) + (c[2] - .5) * (c[2] - .5);
                err = std::fmax(err, std::fabs(phi.At({i, j, k}, 0) - amp * std::exp(-r2 / (2 * s2))));
            }
    EXPECT(err < 2e-3, "heat kernel error %g", err);
    std::printf("  DeviceDiffusion3D: max error vs heat kernel %.2e\n", err);
}

static void StaggeredStorage() {
    constexpr std::size_t N = 3;
    static const char axis[3] = {'X', 'Y', 'Z'};
    using G = dare::Grid::Cartesian<3>;
    G grid({100, 101, 102}, {1., 2., 3.}, 2);
    for (int d = 0; d < 3; d++) {
        G::Options opt{0, 0, 0};
        opt[d] = 1;
        auto rep = grid.GetRepresentation(opt);
        Data::GridVector<G, double, N> data("test", rep);
        int64_t cells = 1;
        for (int e = 0; e < 3; e++) cells *= 100 + e + (e == d) + 4;
        EXPECT(rep.GetNumberLocalCells() == cells, "axis %d: %d local cells, expected %lld", d,
               rep.GetNumberLocalCells(), (long long)cells);
        EXPECT(data.GetSize() == static_cast<std::size_t>(rep.GetNumberLocalCells()) * N, "axis %d: size", d);
        EXPECT(rep.GetGlobalResolutionInternal()[d] == 100 + d + 1, "axis %d: internal cells", d);
        EXPECT(rep.GetName() == std::string("Cartesian3D_") + axis[d], "axis %d: name %s", d, rep.GetName().c_str());
        // the first internal cell sits on the lower boundary face, the last on the upper one
        G::Index lo{2, 2, 2}, hi{2, 2, 2};
        hi[d] = 2 + 100 + d;
        const double L = 1. + d;
        EXPECT(std::fabs(rep.GetCoordinatesCenter(lo)[d]) < 1e-15, "axis %d: first centre %g", d,
               rep.GetCoordinatesCenter(lo)[d]);
        EXPECT(std::fabs(rep.GetCoordinatesCenter(hi)[d] - L) < 1e-13, "axis %d: last centre %g", d,
               rep.GetCoordinatesCenter(hi)[d]);
    }
    {   // Field of a staggered 1-D representation: level names and sizes
        using G1 = dare::Grid::Cartesian<1>;
        G1 grid1({100}, {1.}, 2);
        auto rep_s = grid1.GetRepresentation({1});
        Data::Field<G1, double, N> field_s("test_1D_staggered", rep_s, 2);
        for (std::size_t t = 0; t < 2; t++) {
            const auto& v = field_s.GetDataVector(t);
            EXPECT(v.GetSize() == static_cast<std::size_t>(rep_s.GetNumberLocalCells()) * N, "level %zu size", t);
            EXPECT(v.GetSize() == (100 + 1 + 4) * N, "level %zu: %zu entries", t, v.GetSize());
            EXPECT(v.GetName() == "test_1D_staggered_" + std::to_string(t), "level %zu name %s", t, v.GetName().c_str());
        }
    }
    {   // a periodic axis has no boundary face to add (CartesianRepresentation.inl:66-68)
        using G2 = dare::Grid::Cartesian<2>;
        G2 grid2({10, 11}, {1., 1.}, 2, {1, 0});
        auto rep = grid2.GetRepresentation({1, 1});
        EXPECT(rep.GetGlobalResolutionInternal()[0] == 10 && rep.GetGlobalResolutionInternal()[1] == 12,
               "periodic staggered: %d x %d", rep.GetGlobalResolutionInternal()[0], rep.GetGlobalResolutionInternal()[1]);
        EXPECT(rep.GetName() == "Cartesian2D_X_Y", "name %s", rep.GetName().c_str());
    }
}

static void SolveStaggered1D() {
    constexpr std::size_t N = 3;
    using G = dare::Grid::Cartesian<1>;
    G grid({20}, {1.}, 2);
    auto g_rep = grid.GetRepresentation({1});
    Data::GridVector<G, double, N> data("test", g_rep);
    const double value_west = 0., value_east = 1.;
    const int64_t last = g_rep.GetGlobalResolutionInternal()[0] - 1;
    auto functor = [&](auto mblock) {
        const int64_t node_g = mblock->GetNode();
        const bool west = node_g == 0, east = node_g == last;
        for (std::size_t n = 0; n < mblock->GetNumComponents(); n++) {
            mblock->Resize(n, (west || east) ? 1 : 3);
            if (west) {
                mblock->template Get<CN::CENTER>(n) = 1.;
                mblock->GetRhs(n) = value_west;
                mblock->SetInitialGuess(n, value_west);
            } else if (east) {
                mblock->template Get<CN::CENTER>(n) = 1.;
                mblock->GetRhs(n) = value_east;
                mblock->SetInitialGuess(n, value_east);
            } else {
                mblock->template Get<CN::CENTER>(n) = 2.;
                mblock->template Get<CN::WEST>(n) = -1.;
                mblock->template Get<CN::EAST>(n) = -1.;
                mblock->GetRhs(n) = 0.;
            }
        }
        mblock->Finalize();
    };
    Matrix::Trilinos<double> trilinos;
    Matrix::TrilinosSolver<double> solver;
    auto params = std::make_shared<ParameterList>();
    params->set("Convergence Tolerance", 1e-14);
    params->set("Maximum Iterations", 5000);
    trilinos.Build(g_rep, data, functor, false);
    EXPECT(trilinos.GetX()->getLocalLength() == 21 * N, "rows %zu", trilinos.GetX()->getLocalLength());
    auto ret = solver.Solve(Matrix::SolverPackage::BumbleBee, "BICGSTAB2", trilinos.GetM(), trilinos.GetA(),
                            trilinos.GetX(), trilinos.GetB(), params);
    EXPECT(ret == Matrix::ReturnType::Converged, "SolveStaggered1D did not converge");
    const double d_phi = (value_east - value_west) / last;
    auto x = trilinos.GetX()->getData();
    for (int32_t node_l = 0; node_l < g_rep.GetNumberLocalCellsInternal(); node_l++) {
        const auto ind = g_rep.MapOrdinalToIndexGlobalInternal(g_rep.MapLocalToGlobalInternal(node_l));
        for (std::size_t n = 0; n < N; n++) {
            const double phi = value_west + ind[0] * d_phi;
            EXPECT(std::fabs(phi - x[node_l * N + n]) < 1e-8, "node %d: %g vs %g", node_l, x[node_l * N + n], phi);
        }
    }
}

// the 2-D functor of test_integration_Cartesian_Trilinos.cpp:427-470 with the staggered axis
// `sd` carrying the Dirichlet values (identity rows) and the other axis closed (no neighbour)
template <int SD>
static void SolveStaggered2DAxis() {
    constexpr std::size_t N = 3;
    using G = dare::Grid::Cartesian<2>;
    G grid({20, 21}, {1., 2.}, 2);
    G::Options opt{0, 0};
    opt[SD] = 1;
    auto g_rep = grid.GetRepresentation(opt);
    Data::GridVector<G, double, N> data("test", g_rep);
    const double value_lo = 0., value_hi = 1.;
    const auto res = g_rep.GetGlobalResolutionInternal();
    constexpr int OD = 1 - SD;
    auto functor = [&](auto mblock) {
        const auto ind = mblock->GetIndexInternal();
        const bool lo = ind[SD] == 0, hi = ind[SD] == res[SD] - 1;
        const bool olo = ind[OD] == 0, ohi = ind[OD] == res[OD] - 1;
        std::size_t stencil_size = 5;
        if (lo || hi) {
            stencil_size = 1;
        } else {
            stencil_size -= olo;
            stencil_size -= ohi;
        }
        constexpr CN slo = SD == 0 ? CN::WEST : CN::SOUTH, shi = SD == 0 ? CN::EAST : CN::NORTH;
        constexpr CN olo_nb = OD == 0 ? CN::WEST : CN::SOUTH, ohi_nb = OD == 0 ? CN::EAST : CN::NORTH;
        for (std::size_t n = 0; n < mblock->GetNumComponents(); n++) {
            mblock->Resize(n, stencil_size);
            if (lo || hi) {
                mblock->template Get<CN::CENTER>(n) = 1.;
                mblock->GetRhs(n) = lo ? value_lo : value_hi;
                mblock->SetInitialGuess(n, lo ? value_lo : value_hi);
            } else {
                mblock->template Get<CN::CENTER>(n) = (olo || ohi) ? 3. : 4.;
                mblock->template Get<slo>(n) = -1.;
                mblock->template Get<shi>(n) = -1.;
                if (!olo) mblock->template Get<olo_nb>(n) = -1.;
                if (!ohi) mblock->template Get<ohi_nb>(n) = -1.;
                mblock->GetRhs(n) = 0.;
            }
        }
        mblock->Finalize();
    };
    Matrix::Trilinos<double> trilinos;
    Matrix::TrilinosSolver<double> solver;
    auto params = std::make_shared<ParameterList>();
    params->set("Convergence Tolerance", 1e-14);
    params->set("Maximum Iterations", 5000);
    trilinos.Build(g_rep, data, functor, false);
    auto ret = solver.Solve(Matrix::SolverPackage::BumbleBee, "BICGSTAB2", trilinos.GetM(), trilinos.GetA(),
                            trilinos.GetX(), trilinos.GetB(), params);
    EXPECT(ret == Matrix::ReturnType::Converged, "SolveStaggered2D axis %d did not converge", SD);
    const double d_phi = (value_hi - value_lo) / (res[SD] - 1);
    auto x = trilinos.GetX()->getData();
    double err = 0.;
    for (int32_t node_l = 0; node_l < g_rep.GetNumberLocalCellsInternal(); node_l++) {
        const auto ind = g_rep.MapOrdinalToIndexGlobalInternal(g_rep.MapLocalToGlobalInternal(node_l));
        for (std::size_t n = 0; n < N; n++) err = std::fmax(err, std::fabs(value_lo + ind[SD] * d_phi - x[node_l * N + n]));
    }
    EXPECT(err < 1e-8, "SolveStaggered2D axis %d: max deviation %g", SD, err);
    std::printf("  SolveStaggered2D axis %d: %d x %d cells, %d passes, max deviation %.1e\n", SD, res[0], res[1],
                solver.GetNumIterations(), err);
}

static void SolveStaggered2D() {
    SolveStaggered2DAxis<0>();
    SolveStaggered2DAxis<1>();
}

static void DeviceStaggered() {
    using G = dare::Grid::Cartesian<3>;
    G grid({16, 12, 10}, {1., 1.2, 1.4}, 2);
    auto rep = grid.GetRepresentation({0, 1, 0});
    Data::GridVector<G, double, 1> v("v", rep);
    const auto res = rep.GetLocalResolution();
    for (int i = 0; i < res[0]; i++)
        for (int j = 0; j < res[1]; j++)
            for (int k = 0; k < res[2]; k++) v.At({i, j, k}, 0) = 3. * rep.GetCoordinatesCenter({i, j, k})[1];
    const auto ni = rep.GetGlobalResolutionInternal();
    EXPECT(ni[0] == 16 && ni[1] == 13 && ni[2] == 10, "staggered internal cells %d %d %d", ni[0], ni[1], ni[2]);
    const dare_grid g = v.Descriptor();
    auto dev = Device::Default();
    const std::size_t nf[3] = {static_cast<std::size_t>(ni[0] + 1) * ni[1] * ni[2],
                               static_cast<std::size_t>(ni[0]) * (ni[1] + 1) * ni[2],
                               static_cast<std::size_t>(ni[0]) * ni[1] * (ni[2] + 1)};
    DeviceBuffer<double> fx(dev, nf[0]), fy(dev, nf[1]), fz(dev, nf[2]);
    Check(dare_grad_faces(dev->Get(), &g, v.GetDeviceView(), fx.data(), fy.data(), fz.data()), "dare_grad_faces");
    std::vector<double> hx(nf[0]), hy(nf[1]), hz(nf[2]);
    fx.Download(hx.data());
    fy.Download(hy.data());
    fz.Download(hz.data());
    double ex = 0., ey = 0., ez = 0.;
    for (double e : hx) ex = std::fmax(ex, std::fabs(e));
    for (double e : hy) ey = std::fmax(ey, std::fabs(e - 3.));
    for (double e : hz) ez = std::fmax(ez, std::fabs(e));
    EXPECT(ex == 0. && ez == 0., "cross-axis gradients %g %g", ex, ez);
    EXPECT(ey < 1e-12, "y gradient deviates from 3 by %g", ey);
    std::printf("  DeviceStaggered: 16 x 13 x 10 y-staggered cells, |grad_y - 3| <= %.1e\n", ey);
    // the device fast path on the staggered representation: steady diffusion with the boundary
    // cells of the S/N faces fixed at 0 and 1 (identity rows), Neumann elsewhere -> phi linear
    // in the staggered index, j / 12 (the expectation of SolveStaggeredY,
    // test_integration_Cartesian_Trilinos.cpp:497-585)
    Matrix::Terms terms;
    terms.Diffusion(1.).Fixed(2, 0.).Fixed(3, 1.).Neumann(0).Neumann(1).Neumann(4).Neumann(5);
    Matrix::Trilinos<double> msystem;
    Matrix::TrilinosSolver<double> solver;
    auto params = std::make_shared<ParameterList>();
    params->set("Convergence Tolerance", 1e-13);
    params->set("Maximum Iterations", 2000);
    msystem.Build(rep, v, terms, false);
    auto ret = solver.Solve(Matrix::SolverPackage::BumbleBee, "BICGSTAB2", msystem.GetA(), msystem.GetX(),
                            msystem.GetB(), params);
    EXPECT(ret == Matrix::ReturnType::Converged, "staggered device solve did not converge");
    auto x = msystem.GetX()->getData();
    double err = 0.;
    for (int i = 0; i < ni[0]; i++)
        for (int j = 0; j < ni[1]; j++)
            for (int k = 0; k < ni[2]; k++)
                err = std::fmax(err, std::fabs(x[(static_cast<std::size_t>(i) * ni[1] + j) * ni[2] + k] - j / 12.));
    EXPECT(err < 1e-9, "staggered device solve: max deviation from j / 12 is %g", err);
    std::printf("  DeviceStaggered: fixed S/N cells, %d passes, max deviation from the linear profile %.1e\n",
                solver.GetNumIterations(), err);
}

static std::string g_vtk_dir;

static void VTKWrite() {
    if (g_vtk_dir.empty()) return;
    using G = dare::Grid::Cartesian<3>;
    G grid({6, 5, 4}, {1.2, 1., 0.8}, 2, {1, 1, 1});
    grid.SetName("mirror3D");
    auto grep = grid.GetRepresentation();
    Data::Field<G, double, 2> field("conc", grep, 1);
    auto& v = field.GetDataVector();
    v.SetComponentName(0, "a");
    v.SetComponentName(1, "b");
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 5; j++)
            for (int k = 0; k < 4; k++)
                for (int c = 0; c < 2; c++)
                    v.At(grep.MapInternalToLocal(std::array<int32_t, 3>{i, j, k}), c) = 1e6 * i + 1e3 * j + k + 0.125 * c;
    field.ExchangeHaloCells();   // periodic ghosts filled on the device only
    dare::io::VTKWriter<G> writer(0.5, 4);
    EXPECT(writer.Write(g_vtk_dir, &v), "VTK write to %s failed", g_vtk_dir.c_str());
}

int main(int argc, char** argv) {
    if (argc > 1) g_vtk_dir = argv[1];
    const std::vector<std::pair<const char*, std::function<void()>>> cases = {
        {"TrilinosBuild", TrilinosBuild}, {"SolveLaplace", SolveLaplace}, {"SolveScalar1D", SolveScalar1D},
        {"HaloExchange", HaloExchange}, {"DeviceDiffusion3D", DeviceDiffusion3D}, {"VTKWrite", VTKWrite},
        {"StaggeredStorage", StaggeredStorage}, {"SolveStaggered1D", SolveStaggered1D},
        {"SolveStaggered2D", SolveStaggered2D}, {"DeviceStaggered", DeviceStaggered}, {"SolveJacobi", SolveJacobi}};
    for (const auto& c : cases) {
        const int before = g_fail;
        std::printf("[ RUN ] %s\n", c.first);
        std::fflush(stdout);
        try {
            c.second();
        } catch (const std::exception& e) {
            std::printf("  exception: %s\n", e.what());
            g_fail++;
        }
        std::printf("[ %s ] %s\n", g_fail == before ? "OK" : "FAILED", c.first);
        std::fflush(stdout);
    }
    Device::ReleaseDefault();
    return g_fail ? 1 : 0;
}
