// dare::io::VTKWriter for Cartesian grid vectors, without the VTK library.
// Mirrors IO/VTKWriter.h:73-157 / VTKWriter.inl:38-95,124-137 (one writer per output step, data
// grouped by representation name, so a staggered field gets its own grid; one cell-data array
// per grid vector named after it, component names, a TimeValue field-data entry when
// time >= 0) and the Cartesian options of Grid/Cartesian/VTKOptions_Cartesian.h:144-204,264-322:
// points at the cell vertices of the ghost-padded local grid, offset by the representation's
// offset (-dx/2 on a staggered axis) minus g*dx; VTK's x-fastest cell order (the Mapper of :78-118);
// 1-D grids written as a pseudo 2-D grid one cell thick with dy = dx; whole extent = the
// internal cells 0..n, piece extent = the local cells including the ghost layers counted from
// the rank's cell offset (0 on one rank), i.e. 0..n+2g as GetPExtentLocal computes it, ghost
// level g.
// Files, as vtkXMLPStructuredGridWriter with SetUseSubdirectory(true) lays them out
// (VTKWriter.cpp:28-46 for the name):
//   <base>/<grid>_<step>.pvts               parallel coordination file (one piece on one rank)
//   <base>/<grid>_<step>/<grid>_<step>_0.vts the piece: points and cell data
// Arrays are Float64 in VTK's raw appended encoding (UInt64 byte-count headers, little endian)
// rather than the library's base64 inline blocks; the grid, extents, names and values are the
// same.
#pragma once

#include <array>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <sys/stat.h>
#include <tuple>
#include <vector>

namespace dare::io {

namespace details {

// <path>/<grid>_<step>.<ext> (VTKWriter.cpp:28-46)
inline std::string VTKGetParallelOutputFileName(const std::string& output_path, const std::string& grid_name,
                                                int step, const std::string& ext) {
    std::string s = output_path;
    if (!s.empty() && s.back() != '/') s += '/';
    return s + grid_name + "_" + std::to_string(step) + "." + ext;
}

inline bool MakeDirectory(const std::string& path) {
    struct stat st;
    if (stat(path.c_str(), &st) == 0) return S_ISDIR(st.st_mode);
    return mkdir(path.c_str(), 0755) == 0;
}

inline std::string ExtentString(const std::array<int, 6>& e) {
    std::string s;
    for (int i = 0; i < 6; i++) s += (i ? " " : "") + std::to_string(e[i]);
    return s;
}

// XML attribute text: the few characters XML reserves
inline std::string Escape(const std::string& in) {
    std::string out;
    for (char c : in) {
        switch (c) {
        case '&': out += "&amp;"; break;
        case '<': out += "&lt;"; break;
        case '>': out += "&gt;"; break;
        case '"': out += "&quot;"; break;
        default: out += c;
        }
    }
    return out;
}

struct CellArray {
    std::string name;
    int ncomp;
    std::vector<std::string> cnames;
    std::vector<double> values;   // VTK cell order, tuples of ncomp
};

}  // namespace details

template <typename Grid>
class VTKWriter {
public:
    using Representation = typename Grid::Representation;
    using LO = typename Grid::LocalOrdinalType;
    static constexpr std::size_t Dim = Grid::DIMENSION;

    // time < 0: no TimeValue entry (VTKWriter.inl:143-152)
    VTKWriter(double time, int step) : time_(time), step_(step) {}

    // Writes every grid vector given (pointers, as the reference's calls pass them) into
    // base_path, one file pair per grid. The base folder is created if missing.
    template <typename... Data>
    bool Write(const std::string& base_path, const Data*... data) {
        static_assert(sizeof...(Data) > 0, "at least one data set has to be provided for writing!");
        if (!details::MakeDirectory(base_path)) return false;
        std::map<std::string, std::vector<details::CellArray>> grouped;
        std::map<std::string, const Representation*> reps;
        bool ok = true;
        auto collect = [&](auto* v) {
            const Representation& grep = v->GetGridRepresentation();
            const std::string gname = grep.GetName();   // staggered representations apart
            reps.emplace(gname, &grep);
            grouped[gname].push_back(Gather(*v));
        };
        (collect(data), ...);
        for (auto& [gname, arrays] : grouped) ok = WriteGrid(base_path, gname, *reps[gname], arrays) && ok;
        return ok;
    }

private:
    // cell data of one vector in VTK order (x fastest), ghost cells included
    template <typename V>
    details::CellArray Gather(const V& v) const {
        const_cast<V&>(v).SyncToHost();
        const Representation& grep = v.GetGridRepresentation();
        details::CellArray a;
        a.name = v.GetName();
        a.ncomp = static_cast<int>(v.GetNumEquations());
        for (int n = 0; n < a.ncomp; n++) {
            std::string c = v.GetComponentName(n);
            a.cnames.push_back(c.empty() ? std::to_string(n) : c);
        }
        const LO ncell = grep.GetNumberLocalCells();
        const auto res = grep.GetLocalResolution();
        std::array<int64_t, 3> hsum{1, 1, 1};
        hsum[1] = res[0];
        if (Dim > 1) hsum[2] = static_cast<int64_t>(res[0]) * res[1];
        a.values.assign(static_cast<std::size_t>(ncell) * a.ncomp, 0.);
        for (LO cell = 0; cell < ncell; cell++) {
            const auto ind = grep.MapOrdinalToIndexLocal(cell);
            int64_t vid = 0;
            for (std::size_t d = 0; d < Dim; d++) vid += hsum[d] * ind[d];
            for (int n = 0; n < a.ncomp; n++) a.values[vid * a.ncomp + n] = v.At(cell, n);
        }
        return a;
    }

    bool WriteGrid(const std::string& base, const std::string& gname, const Representation& grep,
                   const std::vector<details::CellArray>& arrays) const {
        const int g = grep.GetGrid()->GetNumGhost();
        const auto res = grep.GetLocalResolution();
        const auto res_int = grep.GetGlobalResolutionInternal();
        const auto dn = grep.GetDistances();
        std::array<int, 3> npt{1, 1, 1};      // points per axis of the local grid
        std::array<double, 3> dx{0., 0., 0.}, off{0., 0., 0.};
        std::array<int, 6> whole{0, 0, 0, 0, 0, 0}, piece{0, 0, 0, 0, 0, 0};
        for (std::size_t d = 0; d < Dim; d++) {
            npt[d] = res[d] + 1;
            dx[d] = dn[d];
            off[d] = grep.GetOffsetSize()[d] - static_cast<double>(g) * dn[d];
            whole[2 * d + 1] = res_int[d];
            piece[2 * d + 1] = res[d];
        }
        if (Dim == 1) {               // pseudo 2-D, one cell thick (VTKOptions_Cartesian.h:163-167)
            npt[1] = 2;
            dx[1] = dx[0];
            whole[3] = 1;
            piece[3] = 1;
        }
        const std::string pvts = details::VTKGetParallelOutputFileName(base, gname, step_, "pvts");
        const std::string sub_name = gname + "_" + std::to_string(step_);
        std::string dir = base;
        if (!dir.empty() && dir.back() != '/') dir += '/';
        if (!details::MakeDirectory(dir + sub_name)) return false;
        const std::string piece_rel = sub_name + "/" + sub_name + "_0.vts";

        // piece file: header, then the raw appended blocks in declaration order
        std::vector<const std::vector<double>*> blocks;
        std::vector<double> tv{time_};
        std::vector<double> pts(static_cast<std::size_t>(npt[0]) * npt[1] * npt[2] * 3);
        std::size_t q = 0;
        for (int k = 0; k < npt[2]; k++)
            for (int j = 0; j < npt[1]; j++)
                for (int i = 0; i < npt[0]; i++) {
                    pts[q++] = dx[0] * i + off[0];
                    pts[q++] = dx[1] * j + off[1];
                    pts[q++] = dx[2] * k + off[2];
                }
        std::string x;
        x += "<?xml version=\"1.0\"?>\n";
        x += "<VTKFile type=\"StructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
        x += "  <StructuredGrid WholeExtent=\"" + details::ExtentString(piece) + "\">\n";
        uint64_t offset = 0;
        auto decl = [&](const std::vector<double>& b) {
            blocks.push_back(&b);
            const std::string s = std::to_string(offset);
            offset += sizeof(uint64_t) + b.size() * sizeof(double);
            return s;
        };
        if (time_ >= 0.) {
            x += "    <FieldData>\n";
            x += "      <DataArray type=\"Float64\" Name=\"TimeValue\" NumberOfTuples=\"1\" format=\"appended\" offset=\"" +
                 decl(tv) + "\"/>\n";
            x += "    </FieldData>\n";
        }
        x += "    <Piece Extent=\"" + details::ExtentString(piece) + "\">\n";
        x += "      <PointData>\n      </PointData>\n";
        x += "      <CellData>\n";
        for (const auto& a : arrays) {
            const std::string off = decl(a.values);
            x += "        " + ArrayHeader(a, &off) + "\n";
        }
        x += "      </CellData>\n";
        x += "      <Points>\n";
        x += "        <DataArray type=\"Float64\" Name=\"Points\" NumberOfComponents=\"3\" format=\"appended\" offset=\"" +
             decl(pts) + "\"/>\n";
        x += "      </Points>\n";
        x += "    </Piece>\n  </StructuredGrid>\n  <AppendedData encoding=\"raw\">\n   _";
        FILE* f = std::fopen((dir + piece_rel).c_str(), "wb");
        if (!f) return false;
        bool ok = std::fwrite(x.data(), 1, x.size(), f) == x.size();
        for (const auto* b : blocks) {
            const uint64_t nbytes = b->size() * sizeof(double);
            ok = ok && std::fwrite(&nbytes, sizeof(nbytes), 1, f) == 1;
            ok = ok && (b->empty() || std::fwrite(b->data(), sizeof(double), b->size(), f) == b->size());
        }
        const char* tail = "\n  </AppendedData>\n</VTKFile>\n";
        ok = ok && std::fwrite(tail, 1, std::strlen(tail), f) == std::strlen(tail);
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) return false;

        // parallel file: whole extent = internal cells, the piece with its ghost layers
        std::string p;
        p += "<?xml version=\"1.0\"?>\n";
        p += "<VTKFile type=\"PStructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
        p += "  <PStructuredGrid WholeExtent=\"" + details::ExtentString(whole) + "\" GhostLevel=\"" +
             std::to_string(g) + "\">\n";
        if (time_ >= 0.) {
            p += "    <PFieldData>\n      <PDataArray type=\"Float64\" Name=\"TimeValue\"/>\n    </PFieldData>\n";
        }
        p += "    <PPointData>\n    </PPointData>\n";
        p += "    <PCellData>\n";
        for (const auto& a : arrays) p += "      " + ArrayHeader(a, nullptr) + "\n";
        p += "    </PCellData>\n";
        p += "    <PPoints>\n      <PDataArray type=\"Float64\" Name=\"Points\" NumberOfComponents=\"3\"/>\n    </PPoints>\n";
        p += "    <Piece Extent=\"" + details::ExtentString(piece) + "\" Source=\"" + piece_rel + "\"/>\n";
        p += "  </PStructuredGrid>\n</VTKFile>\n";
        FILE* fp = std::fopen(pvts.c_str(), "wb");
        if (!fp) return false;
        ok = std::fwrite(p.data(), 1, p.size(), fp) == p.size();
        return (std::fclose(fp) == 0) && ok;
    }

    // <DataArray ... offset=".."/> in the piece, <PDataArray .../> (offset == nullptr) in the
    // parallel file
    static std::string ArrayHeader(const details::CellArray& a, const std::string* offset) {
        std::string h = std::string(offset ? "<DataArray" : "<PDataArray") + " type=\"Float64\" Name=\"" +
                        details::Escape(a.name) + "\" NumberOfComponents=\"" + std::to_string(a.ncomp) + "\"";
        if (a.ncomp > 1)
            for (int n = 0; n < a.ncomp; n++)
                h += " ComponentName" + std::to_string(n) + "=\"" + details::Escape(a.cnames[n]) + "\"";
        if (offset) h += " format=\"appended\" offset=\"" + *offset + "\"";
        return h + "/>";
    }

    double time_;
    int step_;
};

}  // namespace dare::io
