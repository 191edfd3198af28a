// Context, memory and error plumbing of the C-ABI.
#include <cstring>
#include <string>

#include "common.h"

namespace dare_hip {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int make_geo(const dare_grid* gr, Geo* G) {
    if (!gr || gr->dim < 1 || gr->dim > 3 || gr->ncomp < 1 || gr->ncomp > kMaxComp || gr->ghost < 0) {
        set_error("invalid grid descriptor (dim 1..3, ncomp 1..8, ghost >= 0)");
        return DARE_ERR_INVALID;
    }
    G->dim = gr->dim;
    G->N = gr->ncomp;
    G->g = gr->ghost;
    G->any_periodic = 0;
    for (int d = 0; d < 3; d++) {
        const bool act = d < G->dim;
        G->n[d] = act ? gr->n[d] : 1;
        if (G->n[d] < 1) {
            set_error("grid resolution must be >= 1");
            return DARE_ERR_INVALID;
        }
        G->np[d] = act ? G->n[d] + 2 * G->g : 1;
        G->per[d] = act && gr->periodic[d] != 0;
        G->any_periodic |= G->per[d];
        G->sl[d] = 0;
        G->si[d] = 0;
        G->dnr[d] = 1.;
        G->A[d] = 1.;
    }
    // cell width, volume and face area: CartesianMesh.inl:75-83
    double dx[3] = {1., 1., 1.};
    G->V = 1.;
    for (int d = 0; d < G->dim; d++) dx[d] = gr->size[d] / static_cast<double>(G->n[d]);
    for (int d = 0; d < G->dim; d++) G->V *= dx[d];
    for (int d = 0; d < G->dim; d++) G->A[d] = G->V / dx[d];
    for (int d = 0; d < G->dim; d++) {
        double e = 1.;
        e /= dx[d];
        G->dnr[d] = e;
    }
    G->sl[G->dim - 1] = 1;
    G->si[G->dim - 1] = 1;
    for (int d = G->dim - 2; d >= 0; d--) {
        G->sl[d] = G->sl[d + 1] * G->np[d + 1];
        G->si[d] = G->si[d + 1] * G->n[d + 1];
    }
    G->ncell_int = 1;
    G->ncell_pad = 1;
    for (int d = 0; d < G->dim; d++) {
        G->ncell_int *= G->n[d];
        G->ncell_pad *= G->np[d];
    }
    if (G->ncell_pad * G->N >= (1LL << 31)) {
        set_error("grid too large for 32-bit ordinals");
        return DARE_ERR_INVALID;
    }
    return DARE_OK;
}

}  // namespace dare_hip

using namespace dare_hip;

extern "C" {

const char* dare_last_error(void) { return g_last_error.c_str(); }

int32_t dare_abi_version(void) { return 1; }

int dare_ctx_create(int32_t device, dare_ctx** out) {
    DARE_CHECK_ARG(out != nullptr, "out is NULL");
    int ndev = 0;
    DARE_HIP_TRY(hipGetDeviceCount(&ndev));
    DARE_CHECK_ARG(device >= 0 && device < ndev, "device index out of range");
    DARE_HIP_TRY(hipSetDevice(device));
    const int src = noalloc_scratch_init(device);
    if (src) return src;
    dare_ctx* c = new dare_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cu = prop.multiProcessorCount;
    if (hipMalloc(&c->ws.partial, sizeof(double) * kReduceBlocks * kMaxDots) != hipSuccess ||
        hipMalloc(&c->ws.count, sizeof(unsigned int) * 64) != hipSuccess ||
        hipMalloc(&c->dot_out, sizeof(double) * 64) != hipSuccess ||
        hipHostMalloc(&c->host_pinned, sizeof(double) * 64, hipHostMallocDefault) != hipSuccess) {
        set_error("context allocation failed");
        delete c;
        return DARE_ERR_OOM;
    }
    // counters start at zero; every reduction's last block resets its counter to zero
    DARE_HIP_TRY(hipMemset(c->ws.count, 0, sizeof(unsigned int) * 64));
    DARE_HIP_TRY(hipDeviceSynchronize());
    *out = c;
    return DARE_OK;
}

int dare_ctx_destroy(dare_ctx* c) {
    if (!c) return DARE_OK;
    hipSetDevice(c->device);
    hipDeviceSynchronize();
    hipFree(c->ws.partial);
    hipFree(c->ws.count);
    hipFree(c->dot_out);
    if (c->solve_buf) hipFree(c->solve_buf);
    if (c->solve_state) hipFree(c->solve_state);
    hipHostFree(c->host_pinned);
    delete c;
    return DARE_OK;
}

int dare_ctx_set_stream(dare_ctx* c, void* s) {
    DARE_CHECK_ARG(c != nullptr, "ctx is NULL");
    c->stream = static_cast<hipStream_t>(s);
    return DARE_OK;
}

void* dare_ctx_get_stream(dare_ctx* c) { return c ? static_cast<void*>(c->stream) : nullptr; }

int dare_ctx_synchronize(dare_ctx* c) {
    DARE_CHECK_ARG(c != nullptr, "ctx is NULL");
    DARE_HIP_TRY(hipSetDevice(c->device));
    DARE_HIP_TRY(hipStreamSynchronize(c->stream));
    return DARE_OK;
}

int dare_malloc(dare_ctx* c, size_t bytes, void** out) {
    DARE_CHECK_ARG(c != nullptr && out != nullptr, "NULL argument");
    DARE_HIP_TRY(hipSetDevice(c->device));
    if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess) {
        hipGetLastError();
        set_error("hipMalloc failed");
        return DARE_ERR_OOM;
    }
    return DARE_OK;
}

int dare_free(dare_ctx* c, void* p) {
    DARE_CHECK_ARG(c != nullptr, "ctx is NULL");
    if (p) DARE_HIP_TRY(hipFree(p));
    return DARE_OK;
}

int dare_memcpy_h2d(dare_ctx* c, void* dst, const void* src, size_t bytes) {
    DARE_CHECK_ARG(c != nullptr, "ctx is NULL");
    DARE_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    DARE_HIP_TRY(hipStreamSynchronize(c->stream));
    return DARE_OK;
}

int dare_memcpy_d2h(dare_ctx* c, void* dst, const void* src, size_t bytes) {
    DARE_CHECK_ARG(c != nullptr, "ctx is NULL");
    DARE_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    DARE_HIP_TRY(hipStreamSynchronize(c->stream));
    return DARE_OK;
}

int dare_memcpy_d2d(dare_ctx* c, void* dst, const void* src, size_t bytes) {
    DARE_CHECK_ARG(c != nullptr, "ctx is NULL");
    DARE_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return DARE_OK;
}

}  // extern "C"
