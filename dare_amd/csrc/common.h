// Shared internals of libdare_hip: context, error handling, device geometry, reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "dare_hip.h"

namespace dare_hip {

void set_error(const std::string& msg);

#define DARE_HIP_TRY(expr)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            ::dare_hip::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));       \
            return DARE_ERR_HIP;                                                             \
        }                                                                                    \
    } while (0)

#define DARE_CHECK_ARG(cond, msg)                                                            \
    do {                                                                                     \
        if (!(cond)) {                                                                       \
            ::dare_hip::set_error(msg);                                                      \
            return DARE_ERR_INVALID;                                                         \
        }                                                                                    \
    } while (0)

#define DARE_LAUNCH_CHECK()                                                                  \
    do {                                                                                     \
        hipError_t e_ = hipGetLastError();                                                   \
        if (e_ != hipSuccess) {                                                              \
            ::dare_hip::set_error(std::string("kernel launch: ") + hipGetErrorString(e_));  \
            return DARE_ERR_HIP;                                                             \
        }                                                                                    \
    } while (0)

constexpr int kMaxComp = 8;
constexpr int kReduceBlocks = 8192;     // partial sums per one-pass grid reduction (= max grid)
constexpr int kMaxDots = 8;

// Device-side reduction workspace: block partials + arrival counter (reset by the last block).
struct ReduceWs {
    double* partial;      // kReduceBlocks * kMaxDots
    unsigned int* count;  // one counter
};

}  // namespace dare_hip

struct dare_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    dare_hip::ReduceWs ws{nullptr, nullptr};
    double* dot_out = nullptr;     // device results of dare_dot
    double* host_pinned = nullptr; // 64 doubles of pinned host memory for scalar readback
    int num_cu = 256;
    void* solve_buf = nullptr;     // Krylov work vectors, grown on demand
    size_t solve_bytes = 0;
    void* solve_state = nullptr;   // device-resident solver scalars
};

namespace dare_hip {

// Geometry of a grid, computed on the host with the reference's formulas
// (CartesianMesh.inl:75-83) and passed to kernels by value.
struct Geo {
    int dim, N, g;
    int n[3], np[3];
    int per[3];
    int any_periodic;
    long long sl[3], si[3];
    double dnr[3], A[3], V;
    long long ncell_int, ncell_pad;
};

// Slot-major diagonal view of a stencil-structured CSR (3-D, 1 or 2 components): the value
// of slot s in row r is dia[s * nrows + r]. Slots: 0 W, 1 S, 2 B, 3 C, 4 T, 5 N, 6 E,
// 7 cross-component entry. Entries removed by a boundary edit hold 0 and are never read.
constexpr int kDiaSlots = 8;
struct DiaView {
    int n[3];           // cells along i (slowest), j, k (fastest)
    int N;              // components (1 or 2)
    int per[3];
    int rm_lo[3], rm_hi[3];   // boundary entry removed at the low / high face of each axis
    int cross;          // 1: slot 7 holds the cross-component entry
    long long nrows;
};

}  // namespace dare_hip

struct dare_csr {
    int device = 0;              // owner device; the creating context may be destroyed first
    int64_t nrows = 0, nnz = 0, cap_rows = 0, cap_nnz = 0;
    int32_t* row_ptr = nullptr;
    int32_t* col = nullptr;
    double* val = nullptr;
    float* val32 = nullptr;      // FP32 copy of the values (mixed precision, generic path)
    int64_t cap_val32 = 0;
    int stencil_structured = 0;  // 1 if built by dare_csr_build_structure
    // diagonal view written next to the CSR values by dare_csr_fill_values (structured, 3-D)
    int has_dia = 0;
    dare_hip::DiaView dia_view{};
    double* dia = nullptr;
    float* dia32 = nullptr;
    int64_t cap_dia = 0, cap_dia32 = 0;
    // the FP32 copy for mixed solves: once a mixed solve has asked for it (dia32_mode), every
    // fill writes it next to the FP64 slots; dia32_fresh says it holds float(dia) of the last fill
    int dia32_mode = 0, dia32_fresh = 0;
    // 1: the last device fill's operator is exactly symmetric by construction (one component,
    // no convection, reaction or fixed faces, per-cell diffusivity only without a periodic axis:
    // every face coefficient is the same expression for both rows it couples)
    int sym_fill = 0;
    // 1: the last device fill may be exactly symmetric (one component, no convection or
    // reaction): the FP64 view is then checked on the device unless sym_fill holds (the FP32
    // copy of a symmetric FP64 view is symmetric too)
    int sym_cand = 0;
    // the FP64 view's symmetry for the solvers: -1 not yet known since the last fill, 0 not
    // exactly symmetric, 1 symmetric (by construction, or by the device check whose result
    // lands in the device int asym64, allocated with the view)
    int sym64 = -1;
    int* asym64 = nullptr;
    // 0: every traversal reads the seven slots (dare_csr_set_symmetric_traversal)
    int sym_traversal = 1;
    // ent_lazy 1: the last fill left the E, N, T slots (4-6) of the view -- and of the FP32 copy
    // while dia32_fresh -- unwritten: the operator is symmetric by construction and the symmetric
    // traversals read those coefficients as the W, S, B slots of the east, north and top
    // neighbours. dia_ensure_full writes them for any other consumer (seven-slot traversals,
    // Jacobi scaling, multigrid, CSR values, batches) and sets ent_full, after which the fills
    // of this matrix write every slot.
    int ent_lazy = 0, ent_full = 0;
    // Jacobi right preconditioning: 1 / diag(A) and the column-scaled FP64 operator A diag(A)^-1
    // (diagonal-view slots, or CSR values for a matrix without a diagonal view)
    double* dinv = nullptr;
    double* dia_pre = nullptr;
    int64_t cap_dinv = 0, cap_dia_pre = 0;
    // structural description kept for lazy CSR value materialization from the diagonal view
    dare_grid grid_desc{};
    int bc_type[6] = {0, 0, 0, 0, 0, 0};
    int cross = 0;
    int vals_stale = 0;          // 1: CSR values not yet written for the last fill
    void* mg = nullptr;          // multigrid hierarchy of the right preconditioner (precond = 2)
};

namespace dare_hip {

int make_geo(const dare_grid* gr, Geo* G);

// Looks up, with `device` current, the address of the allocation-free solve's static scratch
// on that device (once per device; dare_ctx_create calls it).
int noalloc_scratch_init(int device);

// Releases the multigrid hierarchy a solve attached to a matrix (dare_csr::mg).
void mg_free(void* mg);

// Writes the CSR values of a matrix whose last fill only wrote the diagonal view.
int csr_ensure_values(dare_csr* A, hipStream_t stream);
// the E, N, T slots of a view the last fill left unwritten (ent_lazy), from the neighbours'
// W, S, B slots, on `stream`; no work when they are present
int dia_ensure_full(dare_csr* A, hipStream_t stream);

inline unsigned int ceil_div(long long a, long long b) { return static_cast<unsigned int>((a + b - 1) / b); }

// Wave64 reduction of a double with DPP-free shuffles.
// Streaming accesses. Outputs of the bandwidth-bound kernels are written once and not read
// again by the same kernel: non-temporal stores keep them from displacing cached operands
// (measured on the fused SpMV + dot: 4.7 -> 5.8 TB/s). Matrix values are read once per sweep.
template <typename T>
__device__ __forceinline__ void st_nt(T* p, T v) {
    __builtin_nontemporal_store(v, p);
}
template <typename T>
__device__ __forceinline__ T ld_nt(const T* p) {
    return __builtin_nontemporal_load(p);
}
// store with the hint only where the vectors exceed the caches (chosen at compile time: a
// run-time branch between the two forms is merged by the compiler into a plain store)
template <bool NT, typename T>
__device__ __forceinline__ void put(T* p, T v) {
    if constexpr (NT) st_nt(p, v);
    else *p = v;
}
// rows from which a solve streams (an FP64 vector of this length is larger than the eight
// XCD L2 caches together); below it, outputs stay cached for the next kernel of the pass
constexpr long long kStreamRows = 1LL << 22;
// Mixed-precision refinement: an inner FP32 solve stops at max(inner_tol * ||r||,
// kLastInnerFrac * outer threshold). Once inner_tol * ||r|| is below the outer threshold the
// last inner solve only has to reach a fraction of that threshold, so a closer initial guess
// ends the last refinement early instead of paying a full inner_tol reduction.
constexpr double kLastInnerFrac = 0.25;
__host__ __device__ inline double inner_threshold(double inner_tol, double rn, double outer_abs) {
    const double t = inner_tol * rn, floor_t = kLastInnerFrac * outer_abs;
    return t > floor_t ? t : floor_t;
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

}  // namespace dare_hip
