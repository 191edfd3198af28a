// Host-only driver of dare::io::VTKWriter (no device): writes grid vectors of 1-D, 2-D and 3-D
// Cartesian grids whose cell values encode their local index, into the folder argv[1]. The
// files are checked by tests/test_vtk_writer.py. The stand-in vector has the interface the
// writer reads from a GridVector (name, components, representation, At, SyncToHost).
#include <cstdio>
#include <string>
#include <vector>

#include "dare/Grid.h"
#include "dare/IO.h"

template <typename G, std::size_t N>
struct HostVector {
    std::string name;
    typename G::Representation grep;
    std::vector<double> v;
    std::vector<std::string> cnames = std::vector<std::string>(N);
    HostVector(std::string n, const typename G::Representation& r)
        : name(std::move(n)), grep(r), v(static_cast<std::size_t>(r.GetNumberLocalCells()) * N) {}
    const std::string& GetName() const { return name; }
    std::size_t GetNumEquations() const { return N; }
    const std::string& GetComponentName(std::size_t c) const { return cnames[c]; }
    const typename G::Representation& GetGridRepresentation() const { return grep; }
    double At(int32_t cell, std::size_t c) const { return v[cell * N + c]; }
    void SyncToHost() {}
};

// value of component c of the cell with local index (i, j, k)
static double code(int i, int j, int k, std::size_t c) { return 1e6 * i + 1e3 * j + k + 0.125 * c; }

template <std::size_t Dim, std::size_t N>
static void fill(HostVector<dare::Grid::Cartesian<Dim>, N>& h) {
    for (int32_t cell = 0; cell < h.grep.GetNumberLocalCells(); cell++) {
        const auto ind = h.grep.MapOrdinalToIndexLocal(cell);
        const int i = ind[0], j = Dim > 1 ? ind[Dim > 1 ? 1 : 0] : 0, k = Dim > 2 ? ind[Dim > 2 ? 2 : 0] : 0;
        for (std::size_t c = 0; c < N; c++) h.v[cell * N + c] = code(i, j, k, c);
    }
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string out = argv[1];
    bool ok = true;
    {   // 3-D, two vectors on one grid (one file pair), time stamp
        using G = dare::Grid::Cartesian<3>;
        G grid({5, 6, 7}, {1., 1.2, 1.4}, 2);
        auto rep = grid.GetRepresentation();
        HostVector<G, 1> a("phi", rep);
        HostVector<G, 3> b("vel", rep);
        b.cnames = {"u", "", "w"};
        fill<3, 1>(a);
        fill<3, 3>(b);
        dare::io::VTKWriter<G> writer(0.25, 7);
        ok = writer.Write(out, &a, &b) && ok;
    }
    {   // 2-D, no time stamp, grid renamed
        using G = dare::Grid::Cartesian<2>;
        G grid({4, 3}, {2., 1.5}, 1);
        grid.SetName("plane");
        auto rep = grid.GetRepresentation();
        HostVector<G, 2> a("c", rep);
        fill<2, 2>(a);
        dare::io::VTKWriter<G> writer(-1., 3);
        ok = writer.Write(out, &a) && ok;
    }
    {   // 1-D: pseudo 2-D output
        using G = dare::Grid::Cartesian<1>;
        G grid({9}, {0.9}, 2);
        auto rep = grid.GetRepresentation();
        HostVector<G, 1> a("s", rep);
        fill<1, 1>(a);
        dare::io::VTKWriter<G> writer(1.5, 0);
        ok = writer.Write(out, &a) && ok;
    }
    {   // 3-D, a cell-centred and a y-staggered vector: two grids, "stag" and "stag_Y"
        using G = dare::Grid::Cartesian<3>;
        G grid({5, 6, 7}, {1., 1.2, 1.4}, 2);
        grid.SetName("stag");
        auto rep = grid.GetRepresentation();
        auto rep_y = grid.GetRepresentation({0, 1, 0});
        HostVector<G, 1> a("p", rep);
        HostVector<G, 1> b("v", rep_y);
        fill<3, 1>(a);
        fill<3, 1>(b);
        dare::io::VTKWriter<G> writer(0.5, 2);
        ok = writer.Write(out, &a, &b) && ok;
    }
    {   // 1-D staggered, no time stamp
        using G = dare::Grid::Cartesian<1>;
        G grid({6}, {1.2}, 1);
        grid.SetName("line");
        auto rep = grid.GetRepresentation({1});
        HostVector<G, 1> a("u", rep);
        fill<1, 1>(a);
        dare::io::VTKWriter<G> writer(-1., 1);
        ok = writer.Write(out, &a) && ok;
    }
    std::printf("%s\n", ok ? "written" : "write failed");
    return ok ? 0 : 1;
}
