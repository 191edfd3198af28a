// dare::Grid::Cartesian<Dim> and its representation for one rank (no MPI decomposition).
// Mirrors Grid/Cartesian/CartesianMesh.inl:35-112 and CartesianRepresentation.inl:32-534:
// cell width L/n, volume, face areas; ghost-padded local indices with the last index fastest.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "dare_hip.h"

namespace dare::Grid {

enum class CartesianNeighbor : char {   // CartesianMesh.h:67-77
    CENTER = 0, WEST = 1, EAST = 2, SOUTH = 3, NORTH = 4, BOTTOM = 5, TOP = 6,
};

template <std::size_t Dim>
class CartesianRepresentation;

template <std::size_t Dim>
class Cartesian {
public:
    using LocalOrdinalType = int32_t;
    using GlobalOrdinalType = int64_t;
    using ScalarType = double;
    using Index = std::array<int32_t, Dim>;
    using IndexGlobal = std::array<int64_t, Dim>;
    using VecSC = std::array<double, Dim>;
    using VecLO = std::array<int32_t, Dim>;
    using Representation = CartesianRepresentation<Dim>;
    using NeighborID = CartesianNeighbor;

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

    Representation GetRepresentation() const { return Representation(this); }
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
    using Index = typename GridType::Index;
    using VecSC = typename GridType::VecSC;

    CartesianRepresentation() = default;
    explicit CartesianRepresentation(const GridType* g) : grid_(g) {
        const int32_t ng = g->GetNumGhost();
        for (std::size_t d = 0; d < Dim; d++) {
            res_int_[d] = static_cast<LO>(g->GetGlobalResolution()[d]);
            res_[d] = res_int_[d] + 2 * ng;
        }
        stride_[Dim - 1] = stride_int_[Dim - 1] = 1;          // hierarchic sums, :88-112
        for (int d = static_cast<int>(Dim) - 2; d >= 0; d--) {
            stride_[d] = stride_[d + 1] * res_[d + 1];
            stride_int_[d] = stride_int_[d + 1] * res_int_[d + 1];
        }
    }

    const GridType* GetGrid() const { return grid_; }
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
            pos[d] = 0. + (ind[d] - grid_->GetNumGhost() + 0.5) * grid_->GetCellWidth()[d];
        return pos;
    }

    // the C-ABI grid descriptor of this representation for an N-component field
    dare_grid Descriptor(int ncomp) const {
        dare_grid g{};
        g.dim = static_cast<int32_t>(Dim);
        g.ncomp = ncomp;
        g.ghost = grid_->GetNumGhost();
        for (int d = 0; d < 3; d++) {
            const bool act = d < static_cast<int>(Dim);
            g.n[d] = act ? res_int_[d] : 1;
            g.periodic[d] = act ? grid_->GetPeriodicity()[d] : 0;
            g.size[d] = act ? grid_->GetGlobalSize()[d] : 1.;
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
    Index res_{}, res_int_{};
    std::array<LO, Dim> stride_{}, stride_int_{};
};

}  // namespace dare::Grid
