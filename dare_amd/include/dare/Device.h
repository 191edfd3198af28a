// Host mirror of DaRe's call surface over the MI355X C-ABI (include/dare_hip.h).
// Device context and error translation. Replaces the MPI/Tpetra set-up of dare::ScopeGuard
// (ScopeGuard/ScopeGuard.cpp:30-151) for one rank on one GPU.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>

#include "dare_hip.h"

namespace dare {

// Errors of the C-ABI become exceptions, as the reference's API misuse does
// (TEUCHOS_TEST_FOR_EXCEPTION, MatrixSystem/BiCGStab2.inl:139-160).
inline void Check(int rc, const char* what) {
    if (rc != DARE_OK) throw std::runtime_error(std::string(what) + ": " + dare_last_error());
}

class Device {
public:
    explicit Device(int device = 0, void* stream = nullptr) {
        Check(dare_ctx_create(device, &ctx_), "dare_ctx_create");
        Check(dare_ctx_set_stream(ctx_, stream), "dare_ctx_set_stream");
    }
    ~Device() { dare_ctx_destroy(ctx_); }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;

    dare_ctx* Get() const { return ctx_; }
    void Synchronize() const { Check(dare_ctx_synchronize(ctx_), "dare_ctx_synchronize"); }

    // default device of this process (one rank per GPU)
    static std::shared_ptr<Device>& Default() {
        std::shared_ptr<Device>& d = Slot();
        if (!d) d = std::make_shared<Device>(0);
        return d;
    }
    // drop the default context before the HIP runtime is torn down at exit
    static void ReleaseDefault() { Slot().reset(); }

private:
    static std::shared_ptr<Device>& Slot() {
        static std::shared_ptr<Device> d;
        return d;
    }
    dare_ctx* ctx_ = nullptr;
};

// Device buffer owned by a handle (the reference's Kokkos device view).
template <typename T>
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    DeviceBuffer(std::shared_ptr<Device> dev, std::size_t n) : dev_(std::move(dev)), n_(n) {
        void* p = nullptr;
        Check(dare_malloc(dev_->Get(), sizeof(T) * (n ? n : 1), &p), "dare_malloc");
        p_ = static_cast<T*>(p);
    }
    ~DeviceBuffer() {
        if (p_) dare_free(dev_->Get(), p_);
    }
    DeviceBuffer(DeviceBuffer&& o) noexcept : dev_(std::move(o.dev_)), p_(o.p_), n_(o.n_) {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) {
            if (p_) dare_free(dev_->Get(), p_);
            dev_ = std::move(o.dev_);
            p_ = o.p_;
            n_ = o.n_;
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;

    T* data() const { return p_; }
    std::size_t size() const { return n_; }
    void Upload(const T* host) { Check(dare_memcpy_h2d(dev_->Get(), p_, host, sizeof(T) * n_), "upload"); }
    void Download(T* host) const { Check(dare_memcpy_d2h(dev_->Get(), host, p_, sizeof(T) * n_), "download"); }
    void CopyFrom(const DeviceBuffer& o) {
        Check(dare_memcpy_d2d(dev_->Get(), p_, o.p_, sizeof(T) * n_), "device copy");
    }
    Device& GetDevice() const { return *dev_; }
    std::shared_ptr<Device> GetDevicePtr() const { return dev_; }

private:
    std::shared_ptr<Device> dev_;
    T* p_ = nullptr;
    std::size_t n_ = 0;
};

}  // namespace dare
