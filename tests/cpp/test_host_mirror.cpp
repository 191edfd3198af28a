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

static void SolveLaplace() {
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
    auto M = solver.BuildPreconditioner(Matrix::PreCondPackage::None, "none", nullptr, trilinos.GetA());
    auto ret = solver.Solve(Matrix::SolverPackage::BumbleBee, "BICGSTAB2", M, trilinos.GetA(), trilinos.GetX(),
                            trilinos.GetB(), params);
    EXPECT(ret == Matrix::ReturnType::Converged, "SolveLaplace did not converge");
    auto x = trilinos.GetX()->getData();
    for (int node = 0; node < 23; node++)
        for (std::size_t n = 0; n < N; n++) {
            const double phi = (node + 0.5) / 23.;
            EXPECT(std::fabs(phi - x[node * N + n]) < 1e-8, "node %d comp %zu: %g vs %g", node, n, x[node * N + n], phi);
        }
    std::printf("  SolveLaplace: %d passes\n", solver.GetNumIterations());
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
        {"HaloExchange", HaloExchange}, {"DeviceDiffusion3D", DeviceDiffusion3D}, {"VTKWrite", VTKWrite}};
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
