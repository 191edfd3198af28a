// dare::Grid::Cartesian<Dim> and its representation for one rank (no MPI decomposition).
// Mirrors Grid/Cartesian/CartesianMesh.inl:35-112 and CartesianRepresentation.inl:32-534:
// cell width L/n, volume, face areas; ghost-padded local indices with the last index fastest.
// A representation may be staggered (Options, CartesianMesh.h:209,283): along every axis with
// opt[d] != 0 its cells sit on the faces of the grid's cells, so a non-periodic axis gains one
// cell at its upper end and its centres move by -dx/2 (CartesianRepresentation.inl:50-77).
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "Vector.h"
#include "dare_hip.h"

namespace dare::Grid {

enum class CartesianNeighbor : char {   // CartesianMesh.h:67-77
    CENTER = 0, WEST = 1, EAST = 2, SOUTH = 3, NORTH = 4, BOTTOM = 5, TOP = 6,
};

// position helpers of CartesianMesh.h:79-120: stencil slot, face slot (W = 0), outward normal
inline char ToNum(CartesianNeighbor nb) { return static_cast<char>(nb); }
inline char ToFace(CartesianNeighbor nb) { return static_cast<char>(static_cast<char>(nb) - 1); }
inline char ToNormal(CartesianNeighbor nb) { return (static_cast<char>(nb) % 2 == 1) ? -1 : 1; }
inline CartesianNeighbor ToCartesianNeighbor(std::size_t id) { return static_cast<CartesianNeighbor>(id); }

template <std::size_t Dim>
class CartesianRepresentation;

template <std::size_t Dim>
class Cartesian {
public:
    using LocalOrdinalType = int32_t;
    using GlobalOrdinalType = int64_t;
    using ScalarType = double;
    using Index = utils::Vector<Dim, int32_t>;
    using IndexGlobal = utils::Vector<Dim, int64_t>;
    using VecSC = utils::Vector<Dim, double>;
    using VecLO = utils::Vector<Dim, int32_t>;
    using VecGO = utils::Vector<Dim, int64_t>;
    using Representation = CartesianRepresentation<Dim>;
    using NeighborID = CartesianNeighbor;
    using Options = VecLO;    // per axis: 0 cell centred, != 0 staggered
    static constexpr std::size_t DIMENSION = Dim;
    static constexpr std::size_t STENCIL_SIZE = Dim * 2 + 1;   // CartesianMesh.h:202-203
    static constexpr std::size_t NUM_FACES = Dim * 2;

    Cartesian(const IndexGlobal& res, const VecSC& size, int32_t num_ghost, const VecLO& periodic = VecLO{})
        : res_(res), size_(size), num_ghost_(num_ghost), periodic_(periodic) {
        for (std::size_t d = 0; d < Dim; d++) {
            if (res[d] <= num_ghost)   // CartesianMesh.inl:58-70
                throw std::runtime_error("resolution must be larger than the number of ghost cells");
        }
        volume_ = 1.;
        for (std::size_t d = 0; d < Dim; d++) width_[d] = size_[d] / static_cast<double>(res_[d]);
        for (std::size_t d = 0; d < Dim; d++) volume_ *= width_[d];
        for (std::size_t d = 0; d < Dim; d++) area_[d] = volume_ / width_[d];
    }

    Representation GetRepresentation(const Options& opt = Options{}) const { return Representation(this, opt); }
    // the grid's name (CartesianMesh.h: the first constructor argument); output groups data by it
    const std::string& GetName() const { return name_; }
    void SetName(std::string name) { name_ = std::move(name); }
    const VecSC& GetCellWidth() const { return width_; }
    const VecSC& GetFaceArea() const { return area_; }
    double GetCellVolume() const { return volume_; }
    int32_t GetNumGhost() const { return num_ghost_; }
    const IndexGlobal& GetGlobalResolution() const { return res_; }
    const VecSC& GetGlobalSize() const { return size_; }
    const VecLO& GetPeriodicity() const { return periodic_; }
    bool IsPeriodic() const {
        for (auto p : periodic_)
            if (p) return true;
        return false;
    }

private:
    IndexGlobal res_;
    VecSC size_, width_{}, area_{};
    double volume_ = 1.;
    int32_t num_ghost_;
    VecLO periodic_;
    std::string name_ = "Cartesian" + std::to_string(Dim) + "D";
};

template <std::size_t Dim>
class CartesianRepresentation {
public:
    using GridType = Cartesian<Dim>;
    using LO = int32_t;
    using GO = int64_t;
    using SC = double;
    using Index = typename GridType::Index;
    using IndexGlobal = typename GridType::IndexGlobal;
    using VecSC = typename GridType::VecSC;
    using Options = typename GridType::Options;

    CartesianRepresentation() = default;
    explicit CartesianRepresentation(const GridType* g, const Options& opt = Options{}) : grid_(g), opt_(opt) {
        static const char axis[3] = {'X', 'Y', 'Z'};
        name_ = g->GetName();
        for (std::size_t d = 0; d < Dim; d++)
            if (opt[d] != 0) name_ += std::string("_") + axis[d];
        const int32_t ng = g->GetNumGhost();
        for (std::size_t d = 0; d < Dim; d++) {
            res_int_[d] = static_cast<LO>(g->GetGlobalResolution()[d]);
            offset_[d] = 0.;
            if (opt[d] != 0 && g->GetPeriodicity()[d] == 0) {   // one rank owns the upper boundary
                res_int_[d] += 1;
                offset_[d] = -0.5 * g->GetCellWidth()[d];
            }
            res_[d] = res_int_[d] + 2 * ng;
        }
        stride_[Dim - 1] = stride_int_[Dim - 1] = 1;          // hierarchic sums, :88-112
        for (int d = static_cast<int>(Dim) - 2; d >= 0; d--) {
            stride_[d] = stride_[d + 1] * res_[d + 1];
            stride_int_[d] = stride_int_[d + 1] * res_int_[d + 1];
        }
    }

    const GridType* GetGrid() const { return grid_; }
    // the grid's name with "_X"/"_Y"/"_Z" for each staggered axis; output groups data by it
    const std::string& GetName() const { return name_; }
    const Options& GetOptions() const { return opt_; }
    bool IsStaggered() const {
        for (auto o : opt_)
            if (o) return true;
        return false;
    }
    // position of the first internal cell's lower face (0 on one rank, -dx/2 on a staggered axis)
    const VecSC& GetOffsetSize() const { return offset_; }
    int32_t GetNumberGhostCells() const { return grid_->GetNumGhost(); }
    LO GetNumberLocalCells() const { return product(res_); }
    LO GetNumberLocalCellsInternal() const { return product(res_int_); }
    GO GetNumberGlobalCellsInternal() const { return GetNumberLocalCellsInternal(); }
    const Index& GetLocalResolution() const { return res_; }
    const Index& GetGlobalResolutionInternal() const { return res_int_; }
    const VecSC& GetDistances() const { return grid_->GetCellWidth(); }
    const VecSC& GetFaceArea() const { return grid_->GetFaceArea(); }
    double GetCellVolume(LO = 0) const { return grid_->GetCellVolume(); }

    LO MapIndexToOrdinalLocal(const Index& ind) const {          // :373-384
        LO o = 0;
        for (std::size_t d = 0; d < Dim; d++) o += stride_[d] * ind[d];
        return o;
    }
    Index MapOrdinalToIndexLocal(LO n) const {                    // :493-501
        Index ind;
        for (std::size_t d = 0; d < Dim; d++) {
            ind[d] = n / stride_[d];
            n -= ind[d] * stride_[d];
        }
        return ind;
    }
    Index MapOrdinalToIndexLocalInternal(LO n) const {
        Index ind;
        for (std::size_t d = 0; d < Dim; d++) {
            ind[d] = n / stride_int_[d];
            n -= ind[d] * stride_int_[d];
        }
        return ind;
    }
    LO MapInternalToLocal(LO n_internal) const {                 // :315-333
        Index ind = MapOrdinalToIndexLocalInternal(n_internal);
        for (auto& e : ind) e += grid_->GetNumGhost();
        return MapIndexToOrdinalLocal(ind);
    }
    Index MapInternalToLocal(Index ind) const {
        for (auto& e : ind) e += grid_->GetNumGhost();
        return ind;
    }
    GO MapLocalToGlobalInternal(LO n) const { return n; }          // one rank: global == local
    // CartesianRepresentation.inl:617-623: index + the rank's cell offset, which is 0 on one rank
    IndexGlobal MapLocalToGlobal(const Index& ind) const { return IndexGlobal(ind); }
    Index MapGlobalToLocal(const IndexGlobal& ind) const { return Index(ind); }
    Index GetOffsetCells() const { return Index{}; }
    Index MapOrdinalToIndexGlobalInternal(GO n) const { return MapOrdinalToIndexLocalInternal(static_cast<LO>(n)); }
    LO MapGlobalToLocalInternal(GO n) const { return static_cast<LO>(n); }
    // MapIndexToOrdinalGlobalInternal with the periodic wrap of :418-450
    GO MapIndexToOrdinalGlobalInternal(Index ind, bool map_periodic) const {
        if (map_periodic && grid_->IsPeriodic()) {
            for (std::size_t d = 0; d < Dim; d++) {
                if (ind[d] < 0) ind[d] = res_int_[d] + ind[d];
                else if (ind[d] >= res_int_[d]) ind[d] = ind[d] - res_int_[d];
            }
        }
        GO o = 0;
        for (std::size_t d = 0; d < Dim; d++) o += static_cast<GO>(stride_int_[d]) * ind[d];
        return o;
    }
    GO MapIndexToOrdinalLocalInternal(const Index& ind) const {
        GO o = 0;
        for (std::size_t d = 0; d < Dim; d++) o += static_cast<GO>(stride_int_[d]) * ind[d];
        return o;
    }
    VecSC GetCoordinatesCenter(const Index& ind) const {         // :230-238
        VecSC pos;
        for (std::size_t d = 0; d < Dim; d++)
            pos[d] = offset_[d] + (ind[d] - grid_->GetNumGhost() + 0.5) * grid_->GetCellWidth()[d];
        return pos;
    }

    // the C-ABI grid descriptor of this representation for an N-component field. The device
    // derives the cell width as size / n, so a staggered axis (n + 1 cells) reports the
    // extent (n + 1) dx that keeps dx, volume and face areas those of the grid.
    dare_grid Descriptor(int ncomp) const {
        dare_grid g{};
        g.dim = static_cast<int32_t>(Dim);
        g.ncomp = ncomp;
        g.ghost = grid_->GetNumGhost();
        for (int d = 0; d < 3; d++) {
            const bool act = d < static_cast<int>(Dim);
            g.n[d] = act ? res_int_[d] : 1;
            g.periodic[d] = act ? grid_->GetPeriodicity()[d] : 0;
            g.size[d] = act ? (res_int_[d] == grid_->GetGlobalResolution()[d]
                                   ? grid_->GetGlobalSize()[d]
                                   : grid_->GetCellWidth()[d] * res_int_[d])
                            : 1.;
        }
        return g;
    }

private:
    static LO product(const Index& a) {
        LO p = 1;
        for (auto e : a) p *= e;
        return p;
    }
    const GridType* grid_ = nullptr;
    Options opt_{};
    std::string name_;
    VecSC offset_{};
    Index res_{}, res_int_{};
    std::array<LO, Dim> stride_{}, stride_int_{};
};

}  // namespace dare::Grid
