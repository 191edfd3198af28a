// dare::Data::GridVector and dare::Data::Field on HBM.
// Mirrors Data/GridVector.h:42-251 (ghost-padded, cell-major, component-interleaved storage:
// data[local_ordinal * N + c], GridVector.inl:37-41,69-76) and Data/Field.h:47-138 (time
// levels, CopyDataVectorsToOldTimeStep, ExchangeHaloCells). Like the reference's
// Kokkos::DualView, a host copy and a device copy are kept with explicit synchronisation.
#pragma once

#include <array>
#include <memory>
#include <string>
#include <vector>

#include "Device.h"
#include "Grid.h"

namespace dare::Data {

template <typename Grid, typename SC, std::size_t N>
class GridVector {
public:
    using GridRepresentation = typename Grid::Representation;
    using GridType = Grid;
    using Index = typename Grid::Index;
    using DataType = SC;
    using ScalarType = SC;
    static constexpr std::size_t NUM_COMPONENTS = N;

    GridVector() = default;
    // The device copy is allocated on first device access (dev, or the process's default
    // device), so host-only work (a functor evaluated on the host) needs no GPU.
    GridVector(std::string name, const GridRepresentation& grep, std::shared_ptr<Device> dev = nullptr)
        : name_(std::move(name)), grep_(grep), host_(grep.GetNumberLocalCells() * N, SC(0)), devp_(std::move(dev)),
          host_dirty_(true) {}

    SC& At(int32_t n, std::size_t c) { host_dirty_ = true; return host_[n * N + c]; }
    SC At(int32_t n, std::size_t c) const { return host_[n * N + c]; }
    SC& At(const Index& ind, std::size_t c) { return At(grep_.MapIndexToOrdinalLocal(ind), c); }
    SC At(const Index& ind, std::size_t c) const { return At(grep_.MapIndexToOrdinalLocal(ind), c); }
    // all components of a cell (GridVector.h GetValues)
    utils::Vector<N, SC> GetValues(const Index& ind) const {
        utils::Vector<N, SC> v;
        const int32_t o = grep_.MapIndexToOrdinalLocal(ind);
        for (std::size_t c = 0; c < N; c++) v[c] = host_[o * N + c];
        return v;
    }
    SC& operator[](std::size_t i) { host_dirty_ = true; return host_[i]; }
    SC operator[](std::size_t i) const { return host_[i]; }
    std::size_t GetSize() const { return host_.size(); }
    std::size_t GetNumEquations() const { return N; }
    std::size_t GetNumComponents() const { return N; }
    // component names for output (GridVector.h: SetComponentName / GetComponentName)
    void SetComponentName(std::size_t c, std::string name) { cnames_[c] = std::move(name); }
    const std::string& GetComponentName(std::size_t c) const { return cnames_[c]; }
    const GridRepresentation& GetGridRepresentation() const { return grep_; }
    const std::string& GetName() const { return name_; }
    void SetValues(SC v) {
        for (auto& e : host_) e = v;
        host_dirty_ = true;
    }

    // device side (GridVector::GetDeviceView)
    SC* GetDeviceView() {
        SyncToDevice();
        device_dirty_ = true;   // the caller may write through the pointer
        return Dev().data();
    }
    const SC* GetDeviceView() const {
        const_cast<GridVector*>(this)->SyncToDevice();
        return Dev().data();
    }
    // GridVector::Synchronize (GridVector.inl:124-133): bring the stale side up to date
    void Synchronize() {
        SyncToDevice();
        SyncToHost();
    }
    void SyncToDevice() {
        if (host_dirty_) Dev().Upload(host_.data());
        host_dirty_ = false;
    }
    void SyncToHost() {
        if (device_dirty_) Dev().Download(host_.data());
        device_dirty_ = false;
    }
    void MarkDeviceModified() { device_dirty_ = true; }
    dare_grid Descriptor() const { return grep_.Descriptor(static_cast<int>(N)); }
    Device& GetDevice() const { return Dev().GetDevice(); }

private:
    DeviceBuffer<SC>& Dev() const {
        if (!dev_.data()) {
            if (!devp_) devp_ = Device::Default();
            dev_ = DeviceBuffer<SC>(devp_, host_.size());
        }
        return dev_;
    }

    std::string name_;
    GridRepresentation grep_;
    std::vector<SC> host_;
    mutable std::shared_ptr<Device> devp_;
    mutable DeviceBuffer<SC> dev_;
    bool host_dirty_ = false;
    bool device_dirty_ = false;
    std::array<std::string, N> cnames_{};
};

template <typename Grid, typename SC, std::size_t N>
class Field {
public:
    using GridRepresentation = typename Grid::Representation;
    using GridType = Grid;
    using VectorType = GridVector<Grid, SC, N>;
    using ScalarType = SC;
    static constexpr std::size_t NUM_COMPONENTS = N;

    Field(const std::string& name, const GridRepresentation& grep, std::size_t num_time_levels,
          std::shared_ptr<Device> dev = nullptr) {
        for (std::size_t t = 0; t < num_time_levels; t++)
            levels_.push_back(std::make_unique<VectorType>(name + "_" + std::to_string(t), grep, dev));
    }
    VectorType& GetDataVector(std::size_t t = 0) { return *levels_[t]; }
    const VectorType& GetDataVector(std::size_t t = 0) const { return *levels_[t]; }
    std::size_t GetNumberTimesteps() const { return levels_.size(); }
    void SetValues(SC v) {
        for (auto& l : levels_) l->SetValues(v);
    }
    void SetValues(SC v, std::size_t t) { levels_[t]->SetValues(v); }

    void SetComponentName(std::size_t c, const std::string& name) {
        for (auto& l : levels_) l->SetComponentName(c, name);
    }

    // Field.inl:89-93: level n-1 -> level n, on the device; the host copies follow, so a host
    // functor (DDT reads level 1 per cell) sees the same values
    void CopyDataVectorsToOldTimeStep() {
        for (std::size_t t = levels_.size() - 1; t > 0; t--) {
            SC* dst = levels_[t]->GetDeviceView();
            const SC* src = levels_[t - 1]->GetDeviceView();
            Check(dare_memcpy_d2d(levels_[t]->GetDevice().Get(), dst, src, sizeof(SC) * levels_[t]->GetSize()),
                  "CopyDataVectorsToOldTimeStep");
            levels_[t]->SyncToHost();
        }
    }
    // Field.inl:101-103 -> HaloBuffer::Exchange. One rank: a periodic grid fills its face
    // ghosts from the wrapped internal cells; a non-periodic grid has nothing to exchange.
    void ExchangeHaloCells() {
        VectorType& v = *levels_[0];
        dare_grid g = v.Descriptor();
        int32_t mode[6] = {0, 0, 0, 0, 0, 0};
        bool any = false;
        for (int d = 0; d < g.dim; d++)
            if (g.periodic[d]) {
                mode[2 * d] = mode[2 * d + 1] = DARE_GHOST_PERIODIC;
                any = true;
            }
        if (!any) return;
        Check(dare_ghost_fill(v.GetDevice().Get(), &g, mode, nullptr, v.GetDeviceView()), "ExchangeHaloCells");
        v.SyncToHost();
    }

private:
    std::vector<std::unique_ptr<VectorType>> levels_;
};

}  // namespace dare::Data
