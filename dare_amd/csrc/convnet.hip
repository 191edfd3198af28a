// Learned warm start for the implicit step: a 3-level 3-D convolutional encoder-decoder
// (16/32/64 channels, 3x3x3 kernels, no biases) that predicts phi^{n+1} from phi^n,
// phi^{n-1} and the right-hand side, used as BiCGStab2's initial guess. No counterpart in the
// reference (its guess is phi^n, MatrixSystem/Trilinos.inl:316-317); the solver still iterates
// to the same FP64 tolerance, so a poor guess costs passes, never accuracy.
//
// Tensors are channels-last BF16 ([voxel][channel], voxel = (z * H + y) * W + x with x = the
// grid's fastest index k), weights BF16 with FP32 master copies, accumulation in FP32 on the
// matrix cores. Every convolution is an implicit GEMM on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16): a wave computes up to 64 consecutive voxels x all output
// channels, with K running over 27 taps x input channels. Back-propagation of the data gradient is the same kernel with the
// weights flipped and transposed; the weight gradient is a GEMM over voxels (M = output
// channels, N = input channels) written as per-chunk partials and summed in chunk order.
//
// Network (layer: in -> out channels, resolution):
//   L1 16->16 (input channels padded to 16), L2 16->16            e1   full
//   pool, L3 16->32, L4 32->32                                     e2   1/2
//   pool, L5 32->64, L6 64->64                                     e3   1/4
//   up,   L7 64->32, + e2, L8 32->32                               d2   1/2
//   up,   L9 32->16, + e1, L10 16->16                              d1   full
//   L11 16->16 (only output channel 0 is trained and read)        out  full, FP32
// ReLU after every layer but L11. Prediction: x = x_base + S * out, where x_base is a classical
// guess (phi^n, or the extrapolation 2 phi^n - phi^{n-1}), r0 = b - A x_base its residual (an
// input channel) and S = max|r0| dt / V the size of one Richardson step from it.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"

namespace dare_hip {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kNetThreads = 256;
constexpr int kLayers = 11;
constexpr int kInChannels = 8;     // network input: 3 solver channels + up to 5 static fields
constexpr int kOutWidth = 16;       // channels of the output layer's tensor (channel 0 is used)

struct LayerDef {
    int cin, cout, level;    // level 0 full, 1 half, 2 quarter resolution
};
constexpr LayerDef kDefs[kLayers] = {{kInChannels, 16, 0}, {16, 16, 0}, {16, 32, 1}, {32, 32, 1}, {32, 64, 2}, {64, 64, 2},
                                     {64, 32, 1}, {32, 32, 1}, {32, 16, 0}, {16, 16, 0}, {16, 16, 0}};

// FP32 -> BF16, round to nearest even; NaN stays NaN.
__device__ __host__ inline uint16_t f2bf(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

__device__ __host__ inline float bf2f(uint16_t h) {
    return __builtin_bit_cast(float, static_cast<uint32_t>(h) << 16);
}

// ------------------------------------------------------------------------------ convolution
// Y[v][co] = sum_tap sum_ci X[v + off(tap)][ci] * Wt[tap][co][ci], zero outside the grid.
//
// Computed transposed, D[co][v] = sum_k Wt[co][k] X^T[k][v], on v_mfma_f32_16x16x32_bf16
// (K = 32 per instruction). Operand layout: lane l holds A[l % 16][8 (l / 16) + j] and
// B[8 (l / 16) + j][l % 16] (j = 0..7, 16 contiguous bytes each: eight channels of one output
// channel's weights / of one voxel), and D[4 (l / 16) + j][l % 16] (j = 0..3: four consecutive
// output channels of one voxel, stored as one 8-byte write). K runs over two taps x 16 input
// channels. A wave owns NSEG x 16 voxels of one (z, y) row, so each weight fragment feeds NSEG
// instructions; a block is 4 such rows and stages their input in LDS.
// OUT: 0 BF16 all channels, 1 FP32 channel 0 only.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ inline bf16x8 load8(const uint16_t* p, bool ok) {
    if (!ok) return bf16x8{};
    const uint4 u = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, u);
}

// Output-channel blocks of 16 per work-group (a group stages its slice of the weights).
template <int COUT>
constexpr int conv_nbb() { return COUT / 16 < 2 ? COUT / 16 : 2; }

// Fusions of k_conv (MODE bits): kConvUp -- the input is the half-resolution tensor and the
// conv reads its nearest-neighbour upsampling (staging fetches each voxel's parent); kConvAdd --
// the epilogue also writes Y2 = bf16(Y + E) (the decoder's skip connection, as k_add).
constexpr int kConvUp = 1, kConvAdd = 2;

template <int CIN, int COUT, bool RELU, int OUT, int NSEG, int MODE = 0>
__global__ __launch_bounds__(kNetThreads) void k_conv(const uint16_t* __restrict__ X, const uint16_t* __restrict__ Wt,
                                                     uint16_t* __restrict__ Y, float* __restrict__ Yf, int D, int H,
                                                     int W, const uint16_t* __restrict__ E, uint16_t* __restrict__ Y2) {
    // Block tile: 16 NSEG voxels along x, 4 rows along y (one per wave), one plane z, and
    // 16 NBB output channels (blockIdx.y = row block * NGRP + channel group). Per 16-channel
    // slice of the input, the block stages in LDS the input it needs -- (16 NSEG + 2) x 6 x 3
    // voxels, zero outside the box -- as [half][voxel] (16 bytes = 8 channels per voxel), and
    // its weights as [half][tap][co] (16 bytes each). Every MFMA then reads both operands from
    // LDS: a staged input value feeds up to 27 taps x NBB instructions. All global loads of a
    // stage are issued before it is written, in batches of eight per thread.
    // K = 32 is two taps x 16 channels (the 28th tap is zero).
    constexpr int SX = 16 * NSEG + 2, SY = 6, NV = SX * SY * 3;
    // the two channel halves are NVS pieces apart, a multiple of 256 bytes: lanes 0-15 and
    // 16-31 (32-47 and 48-63) of one operand read then start on the same LDS bank, which keeps
    // the 16-byte reads free of bank conflicts (a 396- or 1188-piece stride costs about 40 %)
    constexpr int NVS = (NV + 15) / 16 * 16;
    constexpr int NBB = conv_nbb<COUT>(), NGRP = COUT / (16 * NBB);
    constexpr int NWP = 2 * 27 * 16 * NBB;                     // weight pieces per slice
    constexpr int NPIECE = 2 * NV + NWP;
    __shared__ uint4 stage[2 * NVS];
    __shared__ uint4 wl[NWP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int grp = blockIdx.y % NGRP;
    const int x0 = blockIdx.x * (16 * NSEG), y0 = (blockIdx.y / NGRP) * 4, z = blockIdx.z;
    const int y = y0 + wave;
    const int i = lane & 15, kq = lane >> 4, half = kq & 1;
    floatx4 acc[NSEG][NBB];
#pragma unroll
    for (int s = 0; s < NSEG; s++)
#pragma unroll
        for (int nb = 0; nb < NBB; nb++) acc[s][nb] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int cc = 0; cc < CIN / 16; cc++) {
        for (int b0 = 0; b0 < NPIECE; b0 += 8 * kNetThreads) {
            uint4 u[8];
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const int t = b0 + r * kNetThreads + threadIdx.x;
                u[r] = uint4{0u, 0u, 0u, 0u};
                if (t < 2 * NV) {        // input piece (voxel, half), half fastest
                    const int v = t >> 1, h = t & 1;
                    const int xl = v % SX, yl = (v / SX) % SY, zl = v / (SX * SY);
                    const int gx = x0 - 1 + xl, gy = y0 - 1 + yl, gz = z - 1 + zl;
                    if (gx >= 0 && gx < W && gy >= 0 && gy < H && gz >= 0 && gz < D) {
                        const long long src = (MODE & kConvUp)
                            ? (static_cast<long long>(gz >> 1) * (H >> 1) + (gy >> 1)) * (W >> 1) + (gx >> 1)
                            : (static_cast<long long>(gz) * H + gy) * W + gx;
                        u[r] = *reinterpret_cast<const uint4*>(X + src * CIN + 16 * cc + 8 * h);
                    }
                } else if (t < NPIECE) { // weight piece (half, tap, co), co fastest
                    const int w = t - 2 * NV;
                    const int co = w % (16 * NBB), tap = (w / (16 * NBB)) % 27, h = w / (16 * NBB * 27);
                    u[r] = *reinterpret_cast<const uint4*>(
                        Wt + (static_cast<long long>(tap) * COUT + grp * 16 * NBB + co) * CIN + 16 * cc + 8 * h);
                }
            }
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const int t = b0 + r * kNetThreads + threadIdx.x;
                if (t < 2 * NV) stage[(t & 1) * NVS + (t >> 1)] = u[r];
                else if (t < NPIECE) wl[t - 2 * NV] = u[r];
            }
        }
        __syncthreads();
        if (y < H) {                           // wave-uniform; every wave reaches both barriers
            // operands of step c + 1 are read from LDS before the MFMAs of step c
            bf16x8 wa[NBB], xb[NSEG];
            auto fetch = [&](int c, bf16x8* w, bf16x8* x) {
                const int tap = 2 * c + (kq >> 1);
                const bool tap_ok = tap < 27;
                const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
#pragma unroll
                for (int nb = 0; nb < NBB; nb++)
                    w[nb] = tap_ok ? __builtin_bit_cast(bf16x8, wl[(half * 27 + tap) * 16 * NBB + nb * 16 + i])
                                   : bf16x8{};
                const int base = half * NVS + ((dz + 1) * SY + (wave + dy + 1)) * SX + (i + dx + 1);
#pragma unroll
                for (int s = 0; s < NSEG; s++)
                    x[s] = tap_ok ? __builtin_bit_cast(bf16x8, stage[base + 16 * s]) : bf16x8{};
            };
            fetch(0, wa, xb);
#pragma unroll
            for (int c = 0; c < 14; c++) {
                bf16x8 wn[NBB], xn[NSEG];
                if (c + 1 < 14) fetch(c + 1, wn, xn);
                __builtin_amdgcn_sched_barrier(0);   // keep the reads of c + 1 ahead of these MFMAs
#pragma unroll
                for (int s = 0; s < NSEG; s++)
#pragma unroll
                    for (int nb = 0; nb < NBB; nb++)
                        acc[s][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[nb], xb[s], acc[s][nb], 0, 0, 0);
                if (c + 1 < 14) {
#pragma unroll
                    for (int nb = 0; nb < NBB; nb++) wa[nb] = wn[nb];
#pragma unroll
                    for (int s = 0; s < NSEG; s++) xb[s] = xn[s];
                }
            }
        }
        __syncthreads();
    }
    if (y >= H) return;
    const long long vrow = (static_cast<long long>(z) * H + y) * W;
#pragma unroll
    for (int s = 0; s < NSEG; s++) {
        const int v = x0 + 16 * s + i;
        if (v >= W) continue;
#pragma unroll
        for (int nb = 0; nb < NBB; nb++) {
            float o[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                o[j] = acc[s][nb][j];
                if (RELU) o[j] = o[j] > 0.f ? o[j] : 0.f;
            }
            const int cob = grp * NBB + nb;
            if (OUT == 0) {
                uint16_t b[4];
#pragma unroll
                for (int j = 0; j < 4; j++) b[j] = f2bf(o[j]);
                const long long at = (vrow + v) * COUT + cob * 16 + kq * 4;
                uint2 pk;
                pk.x = static_cast<uint32_t>(b[0]) | (static_cast<uint32_t>(b[1]) << 16);
                pk.y = static_cast<uint32_t>(b[2]) | (static_cast<uint32_t>(b[3]) << 16);
                *reinterpret_cast<uint2*>(Y + at) = pk;
                if constexpr ((MODE & kConvAdd) != 0) {
                    const uint2 ev = *reinterpret_cast<const uint2*>(E + at);
                    const uint16_t e[4] = {static_cast<uint16_t>(ev.x & 0xffffu), static_cast<uint16_t>(ev.x >> 16),
                                           static_cast<uint16_t>(ev.y & 0xffffu), static_cast<uint16_t>(ev.y >> 16)};
                    uint16_t t[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) t[j] = f2bf(bf2f(b[j]) + bf2f(e[j]));
                    uint2 sk;
                    sk.x = static_cast<uint32_t>(t[0]) | (static_cast<uint32_t>(t[1]) << 16);
                    sk.y = static_cast<uint32_t>(t[2]) | (static_cast<uint32_t>(t[3]) << 16);
                    *reinterpret_cast<uint2*>(Y2 + at) = sk;
                }
            } else if (cob == 0 && kq == 0) {
                Yf[vrow + v] = o[0];
            }
        }
    }
}

// z-marching form for 16- and 32-channel inputs: a block owns a 16 NSEG (x) x 4 (y) tile, 16
// output channels (group grp), and the planes z0 .. z0+ZB-1. Its weights (all taps and input
// channels of its 16 outputs) stay resident for the whole march: with 32 input channels in
// registers (each lane holds the 28 A-operand fragments it feeds, so the LDS serves only the
// input operands; 224 VGPRs, the occupancy the LDS allowed anyway), with 16 in LDS (in
// registers they would cost a wave per SIMD, which loses more than the LDS reads they save).
// The input planes live in a ring of three LDS slots: plane z+2 is loaded into registers while plane z is computed and
// written into the slot of plane z-1 afterwards, so each input plane is staged once per block
// instead of three times and its loads overlap the MFMAs. MODE as k_conv (upsampled input,
// skip-sum epilogue). Same operand layout, K order and accumulation order as k_conv
// (bit-identical results).
template <int CIN, int COUT, bool RELU, int OUT, int NSEG, int MODE = 0>
__global__ __launch_bounds__(kNetThreads) void k_conv_zm(const uint16_t* __restrict__ X, const uint16_t* __restrict__ Wt,
                                                        uint16_t* __restrict__ Y, float* __restrict__ Yf, int D, int H,
                                                        int W, int ZB, const uint16_t* __restrict__ E,
                                                        uint16_t* __restrict__ Y2) {
    static_assert(CIN == 8 || CIN == 16 || CIN == 32, "ring of full-channel planes: 8, 16 or 32 input channels");
    constexpr int SX = 16 * NSEG + 2, SY = 6, NPL = SX * SY;    // voxels of one staged plane
    constexpr int NPS = (NPL + 15) / 16 * 16;                    // LDS stride of one 8-channel piece
    constexpr int CH8 = CIN / 8;                                 // 16-byte pieces per voxel
    constexpr int NGRP = COUT / 16;                              // 16 outputs per block
    constexpr bool WREG = CIN == 32;                             // weights in registers
    constexpr int NWP = WREG ? 1 : CH8 * 27 * 16;                // LDS weight pieces
    constexpr int PPL = CH8 * NPL;                               // pieces of one plane
    constexpr int SLOT = CH8 * NPS;                              // LDS pieces of one ring slot
    constexpr int RPL = (PPL + kNetThreads - 1) / kNetThreads;   // staged pieces per thread
    // K = 32 steps: two taps x 16 channels (CIN 16, 32), or four taps x 8 channels (CIN 8)
    constexpr int NSTEP = CIN == 8 ? 7 : (CIN / 16) * 14;
    // piece strides of 256 bytes multiples (NPS): the lane groups of one operand read that
    // differ only in the channel piece start on the same bank, free of bank conflicts
    __shared__ uint4 ring[3 * SLOT];                             // [slot][piece8][voxel]
    __shared__ uint4 wl[NWP];                                    // [piece8][tap][co] (CIN 16)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int grp = blockIdx.y % NGRP;
    const int x0 = blockIdx.x * (16 * NSEG), y0 = (blockIdx.y / NGRP) * 4;
    const int z0 = blockIdx.z * ZB;
    const int z1 = (z0 + ZB < D) ? z0 + ZB : D;
    const int y = y0 + wave;
    const int i = lane & 15, kq = lane >> 4, half = kq & 1;
    // one staged plane: piece t = (voxel, p8), p8 fastest (adjacent lanes load adjacent
    // 16-byte pieces of global memory)
    auto load_plane = [&](int gz, uint4* u) {
#pragma unroll
        for (int r = 0; r < RPL; r++) {
            const int t = r * kNetThreads + threadIdx.x;
            const int v = t / CH8, p8 = t % CH8;
            u[r] = uint4{0u, 0u, 0u, 0u};
            if (t < PPL && gz >= 0 && gz < D) {
                const int gx = x0 - 1 + v % SX, gy = y0 - 1 + v / SX;
                if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
                    const long long src = (MODE & kConvUp)
                        ? (static_cast<long long>(gz >> 1) * (H >> 1) + (gy >> 1)) * (W >> 1) + (gx >> 1)
                        : (static_cast<long long>(gz) * H + gy) * W + gx;
                    u[r] = *reinterpret_cast<const uint4*>(X + src * CIN + 8 * p8);
                }
            }
        }
    };
    auto store_plane = [&](int slot, const uint4* u) {
#pragma unroll
        for (int r = 0; r < RPL; r++) {
            const int t = r * kNetThreads + threadIdx.x;
            if (t < PPL) ring[slot * SLOT + (t % CH8) * NPS + t / CH8] = u[r];
        }
    };
    // this lane's weight fragments: step q = 16-channel slice q / 14, taps 2 (q % 14) and
    // 2 (q % 14) + 1; lane (i, kq) holds output channel grp 16 + i, tap 2 (q % 14) + kq / 2,
    // channels 8 (2 (q / 14) + kq % 2) .. +7 (zero for the 28th tap)
    bf16x8 wr[WREG ? NSTEP : 1];
    if constexpr (WREG) {
#pragma unroll
        for (int q = 0; q < NSTEP; q++) {
            const int p8 = 2 * (q / 14) + half;
            const int tap = 2 * (q % 14) + (kq >> 1);
            wr[q] = load8(Wt + (static_cast<long long>(tap) * COUT + grp * 16 + i) * CIN + 8 * p8, tap < 27);
        }
    } else {
        for (int t = threadIdx.x; t < NWP; t += kNetThreads) {
            const int co = t % 16, tap = (t / 16) % 27, p8 = t / (16 * 27);
            wl[t] = *reinterpret_cast<const uint4*>(Wt + (static_cast<long long>(tap) * COUT + grp * 16 + co) * CIN +
                                                    8 * p8);
        }
    }
    // the first three planes (z0-1, z0, z0+1 -> slots by plane mod 3)
    {
        uint4 u[RPL];
        for (int q = -1; q <= 1; q++) {
            load_plane(z0 + q, u);
            store_plane((z0 + q + 3) % 3, u);
        }
    }
    __syncthreads();
    for (int z = z0; z < z1; z++) {
        uint4 nxt[RPL];
        if (z + 1 < z1) load_plane(z + 2, nxt);               // block-uniform condition
        floatx4 acc[NSEG];
#pragma unroll
        for (int s = 0; s < NSEG; s++) acc[s] = floatx4{0.f, 0.f, 0.f, 0.f};
        if (y < H) {
            // operands of step q + 1 are read from LDS before the MFMAs of step q (the LDS
            // latency overlaps the matrix work of the same wave); step q = 16-channel slice
            // q / 14, taps 2 (q % 14) and 2 (q % 14) + 1
            bf16x8 wa{}, xb[NSEG];
            auto fetch = [&](int q, bf16x8& w, bf16x8* x) {
                const int p8 = CIN == 8 ? 0 : 2 * (q / 14) + half;
                const int tap = CIN == 8 ? 4 * q + kq : 2 * (q % 14) + (kq >> 1);
                const bool tap_ok = tap < 27;
                const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
                if constexpr (!WREG)
                    w = tap_ok ? __builtin_bit_cast(bf16x8, wl[(p8 * 27 + tap) * 16 + i]) : bf16x8{};
                const int slot = (z + dz + 3) % 3;
                const int base = slot * SLOT + p8 * NPS + (wave + dy + 1) * SX + (i + dx + 1);
#pragma unroll
                for (int s = 0; s < NSEG; s++)
                    x[s] = tap_ok ? __builtin_bit_cast(bf16x8, ring[base + 16 * s]) : bf16x8{};
            };
            fetch(0, wa, xb);
#pragma unroll
            for (int q = 0; q < NSTEP; q++) {
                bf16x8 wn{}, xn[NSEG];
                if (q + 1 < NSTEP) fetch(q + 1, wn, xn);
                __builtin_amdgcn_sched_barrier(0);   // keep the reads of q + 1 ahead of these MFMAs
                const bf16x8 a = WREG ? wr[WREG ? q : 0] : wa;
#pragma unroll
                for (int s = 0; s < NSEG; s++) acc[s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, xb[s], acc[s], 0, 0, 0);
                if (q + 1 < NSTEP) {
                    wa = wn;
#pragma unroll
                    for (int s = 0; s < NSEG; s++) xb[s] = xn[s];
                }
            }
            const long long vrow = (static_cast<long long>(z) * H + y) * W;
#pragma unroll
            for (int s = 0; s < NSEG; s++) {
                const int v = x0 + 16 * s + i;
                if (v >= W) continue;
                float o[4];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    o[j] = acc[s][j];
                    if (RELU) o[j] = o[j] > 0.f ? o[j] : 0.f;
                }
                if (OUT == 0) {
                    uint16_t b[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) b[j] = f2bf(o[j]);
                    const long long at = (vrow + v) * COUT + grp * 16 + kq * 4;
                    uint2 pk;
                    pk.x = static_cast<uint32_t>(b[0]) | (static_cast<uint32_t>(b[1]) << 16);
                    pk.y = static_cast<uint32_t>(b[2]) | (static_cast<uint32_t>(b[3]) << 16);
                    *reinterpret_cast<uint2*>(Y + at) = pk;
                    if constexpr ((MODE & kConvAdd) != 0) {
                        const uint2 ev = *reinterpret_cast<const uint2*>(E + at);
                        const uint16_t e[4] = {static_cast<uint16_t>(ev.x & 0xffffu), static_cast<uint16_t>(ev.x >> 16),
                                               static_cast<uint16_t>(ev.y & 0xffffu), static_cast<uint16_t>(ev.y >> 16)};
                        uint16_t t[4];
#pragma unroll
                        for (int j = 0; j < 4; j++) t[j] = f2bf(bf2f(b[j]) + bf2f(e[j]));
                        uint2 sk;
                        sk.x = static_cast<uint32_t>(t[0]) | (static_cast<uint32_t>(t[1]) << 16);
                        sk.y = static_cast<uint32_t>(t[2]) | (static_cast<uint32_t>(t[3]) << 16);
                        *reinterpret_cast<uint2*>(Y2 + at) = sk;
                    }
                } else if (grp == 0 && kq == 0) {
                    Yf[vrow + v] = o[0];
                }
            }
        }
        __syncthreads();                                         // plane z-1 no longer read
        if (z + 1 < z1) store_plane((z + 2) % 3, nxt);
        __syncthreads();
    }
}

// Nearest 2x upsampling followed by the 3x3x3 conv, in sub-pixel form: a fine voxel
// f = 2p + r (parity r per axis) reads through tap d in {-1, 0, 1} the coarse voxel
// p + floor((r + d) / 2), so for each of the 8 parity classes the 27 taps fold into a 2x2x2
// conv on the coarse tensor -- coarse tap o in {0, 1}^3 at offset o - 1 + r, with weight
// Ws[r][o] = the sum of the taps d in S(r, o) per axis, S(0,0) = {-1}, S(0,1) = {0, 1},
// S(1,0) = {-1, 0}, S(1,1) = {1}. 8 taps instead of 27 per output, the same sums up to the
// rounding of the folded weights (summed in FP32 from the BF16 taps, rounded once).
// Ws layout: [r = 4 rz + 2 ry + rx][o = 4 oz + 2 oy + ox][co][ci], BF16.
template <int CIN, int COUT>
__global__ __launch_bounds__(kNetThreads) void k_sub_weights(const uint16_t* __restrict__ Wt,
                                                            uint16_t* __restrict__ Ws) {
    const int n = 64 * COUT * CIN;
    for (int idx = blockIdx.x * kNetThreads + threadIdx.x; idx < n; idx += gridDim.x * kNetThreads) {
        const int ci = idx % CIN, co = (idx / CIN) % COUT, o = (idx / (CIN * COUT)) % 8, r = idx / (8 * CIN * COUT);
        int lo[3], hi[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const int ra = (r >> (2 - a)) & 1, oa = (o >> (2 - a)) & 1;
            lo[a] = oa == 0 ? -1 : (ra == 0 ? 0 : 1);
            hi[a] = oa == 0 ? (ra == 0 ? -1 : 0) : 1;
        }
        float acc = 0.f;
        for (int dz = lo[0]; dz <= hi[0]; dz++)
            for (int dy = lo[1]; dy <= hi[1]; dy++)
                for (int dx = lo[2]; dx <= hi[2]; dx++) {
                    const int tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
                    acc += bf2f(Wt[(tap * COUT + co) * CIN + ci]);
                }
        Ws[idx] = f2bf(acc);
    }
}

// The sub-pixel conv: Y (fine D x H x W, COUT channels) from X (coarse, CIN channels). A block
// owns one fine z parity rz, the fine planes z = 2k + rz for k in [k0, k0 + ZB), four fine rows
// y0..y0+3 (one per wave, so its y parity is fixed), 32 NSEG fine voxels along x (16 NSEG
// coarse columns, both x parities) and 16 output channels. Each lane loads its weight
// fragments once (2 x-parities x 8 coarse taps x CIN/32 slices, kept in registers). Fine plane
// 2k + rz reads coarse planes k - 1 + rz and k + rz, so the march keeps the coarse planes in a
// ring of three LDS slots (4 rows x (16 NSEG + 2) columns each, zero outside the grid, stored
// as [8-channel piece][voxel] with a 256-byte multiple piece stride): coarse plane k + 1 + rz,
// loaded into registers during step k - 1, is written after step k into the slot nobody reads
// in step k; one barrier per step. A K = 32 step is one coarse tap (oz, oy) x 32 channels; its
// B fragments at coarse x offsets -1, 0, +1 serve the four (rx, ox) products.
// ADD: the epilogue also writes Y2 = bf16(Y + E), with E loaded one step ahead. Y may be NULL
// (inference: only Y2 is stored).
template <int CIN, int COUT, bool RELU, int NSEG, bool ADD>
__global__ __launch_bounds__(kNetThreads) void k_conv_sub(const uint16_t* __restrict__ X, const uint16_t* __restrict__ Ws,
                                                         uint16_t* __restrict__ Y, int D, int H, int W, int ZB,
                                                         const uint16_t* __restrict__ E, uint16_t* __restrict__ Y2) {
    constexpr int CX = 16 * NSEG + 2, NPL = CX * 4;              // coarse voxels of one staged plane
    constexpr int NPS = (NPL + 15) / 16 * 16;                    // LDS stride of one 8-channel piece
    constexpr int CH8 = CIN / 8, KS = CIN / 32;
    constexpr int SLOT = CH8 * NPS;                              // LDS pieces of one ring slot
    constexpr int RPL = (SLOT + kNetThreads - 1) / kNetThreads;
    constexpr int NGRP = COUT / 16;
    __shared__ uint4 ring[3 * SLOT];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int grp = blockIdx.y % NGRP;
    const int y0 = (blockIdx.y / NGRP) * 4;
    const int rz = blockIdx.z & 1;
    const int Dc = D >> 1, Hc = H >> 1, Wc = W >> 1;
    const int k0 = (blockIdx.z >> 1) * ZB;
    const int k1 = (k0 + ZB < Dc) ? k0 + ZB : Dc;
    const int px0 = blockIdx.x * (16 * NSEG);
    const int y = y0 + wave, ry = wave & 1;
    const int i = lane & 15, kq = lane >> 4;
    const int cy0 = (y0 >> 1) - 1;
    bf16x8 wr[2][8][KS];
#pragma unroll
    for (int rx = 0; rx < 2; rx++)
#pragma unroll
        for (int o = 0; o < 8; o++)
#pragma unroll
            for (int c = 0; c < KS; c++)
                wr[rx][o][c] = load8(Ws + ((((rz * 4 + ry * 2 + rx) * 8 + o) * COUT + grp * 16 + i) * CIN + 32 * c + 8 * kq),
                                     true);
    // one coarse plane cz: 16 consecutive threads move 16 consecutive voxels of one piece
    auto load_plane = [&](int cz, uint4* u) {
#pragma unroll
        for (int r = 0; r < RPL; r++) {
            const int t = r * kNetThreads + threadIdx.x;
            const int g = t >> 4, p8 = g % CH8, v = (g / CH8) * 16 + (t & 15);
            u[r] = uint4{0u, 0u, 0u, 0u};
            if (v < NPL && cz >= 0 && cz < Dc) {
                const int gx = px0 - 1 + v % CX, gy = cy0 + v / CX;
                if (gx >= 0 && gx < Wc && gy >= 0 && gy < Hc)
                    u[r] = *reinterpret_cast<const uint4*>(
                        X + ((static_cast<long long>(cz) * Hc + gy) * Wc + gx) * CIN + 8 * p8);
            }
        }
    };
    auto store_plane = [&](int cz, const uint4* u) {
        const int slot = (cz + 3) % 3;
#pragma unroll
        for (int r = 0; r < RPL; r++) {
            const int t = r * kNetThreads + threadIdx.x;
            const int g = t >> 4, p8 = g % CH8, v = (g / CH8) * 16 + (t & 15);
            if (t < SLOT) ring[slot * SLOT + p8 * NPS + v] = u[r];
        }
    };
    {
        uint4 u[RPL];
        for (int q = 0; q < 2; q++) {                            // coarse planes k0 - 1 + rz, k0 + rz
            load_plane(k0 - 1 + rz + q, u);
            store_plane(k0 - 1 + rz + q, u);
        }
    }
    __syncthreads();
    // E of output plane z = 2k + rz (the skip operand), loaded before the MFMAs that precede
    // its use
    auto load_e = [&](int k, uint2 (&ev)[2][NSEG]) {
        if constexpr (ADD) {
            const long long vrow = (static_cast<long long>(2 * k + rz) * H + (y < H ? y : H - 1)) * W;
#pragma unroll
            for (int s = 0; s < NSEG; s++)
#pragma unroll
                for (int rx = 0; rx < 2; rx++) {
                    const int x = 2 * (px0 + 16 * s + i) + rx;
                    ev[rx][s] = (y < H && x < W)
                        ? *reinterpret_cast<const uint2*>(E + (vrow + x) * COUT + grp * 16 + kq * 4)
                        : uint2{0u, 0u};
                }
        }
    };
    // output plane z = 2k + rz from the ring's coarse planes k - 1 + rz and k + rz
    auto step = [&](int k, const uint2 (&ev)[2][NSEG]) {
        const int z = 2 * k + rz;
        const long long vrow = (static_cast<long long>(z) * H + (y < H ? y : H - 1)) * W;
        if (y < H) {                                             // wave-uniform
            floatx4 acc[2][NSEG];
#pragma unroll
            for (int rx = 0; rx < 2; rx++)
#pragma unroll
                for (int s = 0; s < NSEG; s++) acc[rx][s] = floatx4{0.f, 0.f, 0.f, 0.f};
            constexpr int NST = 4 * KS;                          // (oz, oy, c) steps
            bf16x8 bm[NSEG], b0[NSEG], bp[NSEG];
            auto fetch = [&](int q, bf16x8* m, bf16x8* c0, bf16x8* p) {
                const int oz = q / (2 * KS), oy = (q / KS) % 2, c = q % KS;
                const int slot = (k - 1 + rz + oz + 3) % 3;
                const int base = slot * SLOT + (4 * c + kq) * NPS + ((wave >> 1) + oy + ry) * CX + i + 1;
#pragma unroll
                for (int s = 0; s < NSEG; s++) {
                    m[s] = __builtin_bit_cast(bf16x8, ring[base + 16 * s - 1]);
                    c0[s] = __builtin_bit_cast(bf16x8, ring[base + 16 * s]);
                    p[s] = __builtin_bit_cast(bf16x8, ring[base + 16 * s + 1]);
                }
            };
            fetch(0, bm, b0, bp);
#pragma unroll
            for (int q = 0; q < NST; q++) {
                bf16x8 nm[NSEG], n0[NSEG], np[NSEG];
                if (q + 1 < NST) fetch(q + 1, nm, n0, np);
                __builtin_amdgcn_sched_barrier(0);   // keep the reads of q + 1 ahead of these MFMAs
                const int oz = q / (2 * KS), oy = (q / KS) % 2, c = q % KS;
                const int o0 = oz * 4 + oy * 2;                  // ox = 0; ox = 1 is o0 + 1
#pragma unroll
                for (int s = 0; s < NSEG; s++) {
                    acc[0][s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wr[0][o0][c], bm[s], acc[0][s], 0, 0, 0);
                    acc[1][s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wr[1][o0][c], b0[s], acc[1][s], 0, 0, 0);
                    acc[0][s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wr[0][o0 + 1][c], b0[s], acc[0][s], 0, 0, 0);
                    acc[1][s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wr[1][o0 + 1][c], bp[s], acc[1][s], 0, 0, 0);
                }
                if (q + 1 < NST) {
#pragma unroll
                    for (int s = 0; s < NSEG; s++) {
                        bm[s] = nm[s];
                        b0[s] = n0[s];
                        bp[s] = np[s];
                    }
                }
            }
#pragma unroll
            for (int s = 0; s < NSEG; s++)
#pragma unroll
                for (int rx = 0; rx < 2; rx++) {
                    const int x = 2 * (px0 + 16 * s + i) + rx;
                    if (x >= W) continue;
                    uint16_t b[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        float o = acc[rx][s][j];
                        if (RELU) o = o > 0.f ? o : 0.f;
                        b[j] = f2bf(o);
                    }
                    const long long at = (vrow + x) * COUT + grp * 16 + kq * 4;
                    uint2 pk;
                    pk.x = static_cast<uint32_t>(b[0]) | (static_cast<uint32_t>(b[1]) << 16);
                    pk.y = static_cast<uint32_t>(b[2]) | (static_cast<uint32_t>(b[3]) << 16);
                    if (Y) *reinterpret_cast<uint2*>(Y + at) = pk;     // NULL: inference, sum only
                    if constexpr (ADD) {
                        const uint2 e2 = ev[rx][s];
                        const uint16_t e[4] = {static_cast<uint16_t>(e2.x & 0xffffu), static_cast<uint16_t>(e2.x >> 16),
                                               static_cast<uint16_t>(e2.y & 0xffffu), static_cast<uint16_t>(e2.y >> 16)};
                        uint16_t t[4];
#pragma unroll
                        for (int j = 0; j < 4; j++) t[j] = f2bf(bf2f(b[j]) + bf2f(e[j]));
                        uint2 sk;
                        sk.x = static_cast<uint32_t>(t[0]) | (static_cast<uint32_t>(t[1]) << 16);
                        sk.y = static_cast<uint32_t>(t[2]) | (static_cast<uint32_t>(t[3]) << 16);
                        *reinterpret_cast<uint2*>(Y2 + at) = sk;
                    }
                }
        }
    };
    // Two steps per loop iteration with two register sets (pa/ea, pb/eb): the coarse plane a
    // step stores into the ring and the E operand of the next step are loaded one step ahead,
    // so their latency overlaps a whole step, and no register copy has to wait for them. The
    // slot of coarse plane k + 1 + rz, written at the end of step k, held plane k - 2 + rz,
    // last read in step k - 1. Every condition below is block-uniform.
    uint4 pa[RPL], pb[RPL];
    uint2 ea[2][NSEG], eb[2][NSEG];
    if (k0 + 1 < k1) load_plane(k0 + 1 + rz, pa);
    load_e(k0, ea);
    for (int k = k0; k < k1; k += 2) {
        if (k + 2 < k1) load_plane(k + 2 + rz, pb);
        if (k + 1 < k1) load_e(k + 1, eb);
        step(k, ea);
        if (k + 1 < k1) store_plane(k + 1 + rz, pa);
        __syncthreads();
        if (k + 1 >= k1) break;
        if (k + 3 < k1) load_plane(k + 3 + rz, pa);
        if (k + 2 < k1) load_e(k + 2, ea);
        step(k + 1, eb);
        if (k + 2 < k1) store_plane(k + 2 + rz, pb);
        __syncthreads();
    }
}

// Weight gradient partials: P[c][tap][co][ci] = sum over the voxels of chunk c of
// G[v][co] * X[v + off(tap)][ci]. Wave = one chunk of (z, y) rows for one 16x16 (co, ci)
// block; A[co][voxel] = G, B[voxel][ci] = shifted X; 27 accumulators, one per tap.
template <int CIN, int COUT>
__global__ __launch_bounds__(kNetThreads) void k_wgrad(const uint16_t* __restrict__ X, const uint16_t* __restrict__ G,
                                                      int D, int H, int W, int rows_per_chunk, int nchunks,
                                                      float* __restrict__ P) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int chunk = blockIdx.x * 4 + wave;
    if (chunk >= nchunks) return;
    constexpr int CB = CIN >= 16 ? CIN / 16 : 1;                // an 8-channel input is one half block
    const int cob = blockIdx.y / CB, cib = blockIdx.y % CB;
    const int i = lane & 15, kq = lane >> 4;
    const bool ci_ok = cib * 16 + i < CIN;
    floatx4 acc[27];
#pragma unroll
    for (int t = 0; t < 27; t++) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
    const int rows = D * H;
    const int r0 = chunk * rows_per_chunk;
    const int r1 = (r0 + rows_per_chunk < rows) ? r0 + rows_per_chunk : rows;
    for (int r = r0; r < r1; r++) {
        const int z = r / H, y = r % H;
        const long long vrow = static_cast<long long>(r) * W;
        for (int xs = 0; xs < W; xs += 16) {
            bf16x4 a;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int xv = xs + kq * 4 + e;
                a[e] = xv < W ? static_cast<short>(G[(vrow + xv) * COUT + cob * 16 + i]) : short(0);
            }
#pragma unroll
            for (int dz = -1; dz <= 1; dz++) {
                const int zz = z + dz;
                if (zz < 0 || zz >= D) continue;
#pragma unroll
                for (int dy = -1; dy <= 1; dy++) {
                    const int yy = y + dy;
                    if (yy < 0 || yy >= H) continue;
                    const long long srow = (static_cast<long long>(zz) * H + yy) * W;
#pragma unroll
                    for (int dx = -1; dx <= 1; dx++) {
                        bf16x4 b;
#pragma unroll
                        for (int e = 0; e < 4; e++) {
                            const int xv = xs + kq * 4 + e;
                            const int xx = xv + dx;
                            const bool ok = ci_ok && xv < W && xx >= 0 && xx < W;
                            b[e] = ok ? static_cast<short>(X[(srow + xx) * CIN + cib * 16 + i]) : short(0);
                        }
                        const int tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
                        acc[tap] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[tap], 0, 0, 0);
                    }
                }
            }
        }
    }
    if (!ci_ok) return;
#pragma unroll
    for (int t = 0; t < 27; t++)
#pragma unroll
        for (int e = 0; e < 4; e++)
            P[((static_cast<long long>(chunk) * 27 + t) * COUT + cob * 16 + kq * 4 + e) * CIN + cib * 16 + i] = acc[t][e];
}

// grad[idx] = sum over chunks in chunk order (deterministic)
__global__ __launch_bounds__(kNetThreads) void k_chunk_sum(const float* __restrict__ P, int nchunks, int n,
                                                          float* __restrict__ grad) {
    const int idx = blockIdx.x * kNetThreads + threadIdx.x;
    if (idx >= n) return;
    float s = 0.f;
    for (int c = 0; c < nchunks; c++) s += P[static_cast<long long>(c) * n + idx];
    grad[idx] = s;
}

// ------------------------------------------------------------------------------ resampling
// Element-wise kernels over channels-last BF16 tensors, eight channels (16 bytes) per work item
// (C is 16, 32 or 64); per channel the arithmetic and its order are those of one element at a
// time. Voxel counts fit 32 bits (a 1024^3 box has 2^30 voxels).
union Bf8 {
    uint4 v;
    uint16_t h[8];
};

__device__ inline Bf8 ld8(const uint16_t* p, long long i8) { return Bf8{reinterpret_cast<const uint4*>(p)[i8]}; }
__device__ inline void st8(uint16_t* p, long long i8, const Bf8& b) { reinterpret_cast<uint4*>(p)[i8] = b.v; }

#define DARE_EW8_LOOP(n8)                                                                             \
    for (long long q = blockIdx.x * (long long)kNetThreads + threadIdx.x; q < (n8);                   \
         q += (long long)gridDim.x * kNetThreads)

// (z, y, x, channel group) of work item q over a D x H x W box with C/8 groups per voxel
__device__ inline void ew_coords(long long q, int C8, int H, int W, int* z, int* y, int* x, int* cg) {
    *cg = static_cast<int>(q % C8);
    const long long v = q / C8;
    *x = static_cast<int>(v % W);
    const long long r = v / W;
    *y = static_cast<int>(r % H);
    *z = static_cast<int>(r / H);
}

// out (coarse, dims D x H x W) = mean of the 8 children of the fine grid (2D x 2H x 2W)
__global__ __launch_bounds__(kNetThreads) void k_pool(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, int C,
                                                     int D, int H, int W) {
    const int C8 = C / 8;
    const long long n8 = static_cast<long long>(D) * H * W * C8;
    DARE_EW8_LOOP(n8) {
        int z, y, x, cg;
        ew_coords(q, C8, H, W, &z, &y, &x, &cg);
        float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                for (int d = 0; d < 2; d++) {
                    const long long f = ((static_cast<long long>(2 * z + a) * (2 * H) + (2 * y + b)) * (2 * W) + (2 * x + d));
                    const Bf8 c = ld8(in, f * C8 + cg);
#pragma unroll
                    for (int e = 0; e < 8; e++) s[e] += bf2f(c.h[e]);
                }
        Bf8 o;
#pragma unroll
        for (int e = 0; e < 8; e++) o.h[e] = f2bf(s[e] * 0.125f);
        st8(out, q, o);
    }
}

// out (fine, D x H x W) = parent value (nearest upsampling)
__global__ __launch_bounds__(kNetThreads) void k_up(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, int C,
                                                   int D, int H, int W) {
    const int C8 = C / 8;
    const long long n8 = static_cast<long long>(D) * H * W * C8;
    DARE_EW8_LOOP(n8) {
        int z, y, x, cg;
        ew_coords(q, C8, H, W, &z, &y, &x, &cg);
        const long long p = (static_cast<long long>(z / 2) * (H / 2) + y / 2) * (W / 2) + x / 2;
        st8(out, q, ld8(in, p * C8 + cg));
    }
}

// out[parent] = (y[parent] > 0) ? sum of the 8 children of du : 0   (up-sampling backward + ReLU)
__global__ __launch_bounds__(kNetThreads) void k_up_bwd_mask(const uint16_t* __restrict__ du, const uint16_t* __restrict__ y,
                                                            uint16_t* __restrict__ out, int C, int D, int H, int W) {
    const int C8 = C / 8;
    const long long n8 = static_cast<long long>(D) * H * W * C8;
    DARE_EW8_LOOP(n8) {
        int z, yy, x, cg;
        ew_coords(q, C8, H, W, &z, &yy, &x, &cg);
        float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                for (int d = 0; d < 2; d++) {
                    const long long f = ((static_cast<long long>(2 * z + a) * (2 * H) + (2 * yy + b)) * (2 * W) + (2 * x + d));
                    const Bf8 c = ld8(du, f * C8 + cg);
#pragma unroll
                    for (int e = 0; e < 8; e++) s[e] += bf2f(c.h[e]);
                }
        const Bf8 yv = ld8(y, q);
        Bf8 o;
#pragma unroll
        for (int e = 0; e < 8; e++) o.h[e] = bf2f(yv.h[e]) > 0.f ? f2bf(s[e]) : uint16_t(0);
        st8(out, q, o);
    }
}

// out[child] = (y[child] > 0) ? extra[child] + dp[parent] / 8 : 0   (pooling backward + skip + ReLU)
__global__ __launch_bounds__(kNetThreads) void k_pool_bwd_add_mask(const uint16_t* __restrict__ dp,
                                                                  const uint16_t* __restrict__ extra,
                                                                  const uint16_t* __restrict__ y,
                                                                  uint16_t* __restrict__ out, int C, int D, int H, int W) {
    const int C8 = C / 8;
    const long long n8 = static_cast<long long>(D) * H * W * C8;
    DARE_EW8_LOOP(n8) {
        int z, yy, x, cg;
        ew_coords(q, C8, H, W, &z, &yy, &x, &cg);
        const long long p = (static_cast<long long>(z / 2) * (H / 2) + yy / 2) * (W / 2) + x / 2;
        const Bf8 ex = ld8(extra, q), pv = ld8(dp, p * C8 + cg), yv = ld8(y, q);
        Bf8 o;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const float g = bf2f(ex.h[e]) + 0.125f * bf2f(pv.h[e]);
            o.h[e] = bf2f(yv.h[e]) > 0.f ? f2bf(g) : uint16_t(0);
        }
        st8(out, q, o);
    }
}

// n: elements (a multiple of 8)
__global__ __launch_bounds__(kNetThreads) void k_add(const uint16_t* __restrict__ a, const uint16_t* __restrict__ b,
                                                    uint16_t* __restrict__ out, long long n) {
    DARE_EW8_LOOP(n / 8) {
        const Bf8 x = ld8(a, q), z = ld8(b, q);
        Bf8 o;
#pragma unroll
        for (int e = 0; e < 8; e++) o.h[e] = f2bf(bf2f(x.h[e]) + bf2f(z.h[e]));
        st8(out, q, o);
    }
}

__global__ __launch_bounds__(kNetThreads) void k_mask(const uint16_t* __restrict__ g, const uint16_t* __restrict__ y,
                                                     uint16_t* __restrict__ out, long long n) {
    DARE_EW8_LOOP(n / 8) {
        const Bf8 gv = ld8(g, q), yv = ld8(y, q);
        Bf8 o;
#pragma unroll
        for (int e = 0; e < 8; e++) o.h[e] = bf2f(yv.h[e]) > 0.f ? gv.h[e] : uint16_t(0);
        st8(out, q, o);
    }
}

// ------------------------------------------------------------------------------ solver side
// Normalisation scales: s[0] = max|phi^n|, s[1] = max|phi^n - phi^{n-1}|,
// s[2] = max|resid * resid_scale| over the cells of the box (block maxima, then one block);
// s[2] is also the scale of the output: the guess is phi^n + s[2] * out.
// The network may see a sub-box (crop) of the grid: n0..n2 are its sizes, o0..o2 its origin
// among the internal cells, m1, m2 the grid's internal sizes (row order of the solver vectors).
// With coarsening cf (1 or 2) the network sees the box at 1/cf resolution: network voxel
// (z, y, x) stands for the cf^3 cells (cf z + a, cf y + b, cf x + c) of the box, c0..c2 = n/cf.
struct FieldGeo {
    int n0, n1, n2, g;
    int o0, o1, o2, m1, m2;
    long long s0, s1;    // padded strides of i and j
    int cf, c0, c1, c2;
};

__device__ inline long long pad_index(const FieldGeo& F, int z, int y, int x) {
    return (z + F.o0 + F.g) * F.s0 + (y + F.o1 + F.g) * F.s1 + (x + F.o2 + F.g);
}

__device__ inline long long row_index(const FieldGeo& F, int z, int y, int x) {
    return (static_cast<long long>(z + F.o0) * F.m1 + (y + F.o1)) * F.m2 + (x + F.o2);
}

// (z, y, x) of linear index q over an a x b x c box, c fastest, in 32-bit arithmetic (boxes
// hold fewer than 2^32 cells: check_grid)
__device__ inline void box_zyx(long long q, int b, int c, int& z, int& y, int& x) {
    const unsigned u = static_cast<unsigned>(q);
    const unsigned t = u / static_cast<unsigned>(c);
    x = static_cast<int>(u - t * static_cast<unsigned>(c));
    const unsigned zz = t / static_cast<unsigned>(b);
    y = static_cast<int>(t - zz * static_cast<unsigned>(b));
    z = static_cast<int>(zz);
}

__global__ __launch_bounds__(kNetThreads) void k_scales(FieldGeo F, const double* __restrict__ phi,
                                                       const double* __restrict__ phim, const double* __restrict__ resid,
                                                       double resid_scale, double* __restrict__ part) {
    __shared__ double red[3][kNetThreads / 64];
    double m0 = 0., m1 = 0., m2 = 0.;
    const long long n = static_cast<long long>(F.n0) * F.n1 * F.n2;
    for (long long q = blockIdx.x * (long long)kNetThreads + threadIdx.x; q < n; q += (long long)gridDim.x * kNetThreads) {
        int z, y, x;
        box_zyx(q, F.n1, F.n2, z, y, x);
        const double p = phi[pad_index(F, z, y, x)];
        const double pm = phim[pad_index(F, z, y, x)];
        m0 = fmax(m0, fabs(p));
        m1 = fmax(m1, fabs(p - pm));
        m2 = fmax(m2, fabs(resid[row_index(F, z, y, x)] * resid_scale));
    }
    for (int off = 32; off > 0; off >>= 1) {
        m0 = fmax(m0, __shfl_down(m0, off));
        m1 = fmax(m1, __shfl_down(m1, off));
        m2 = fmax(m2, __shfl_down(m2, off));
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = m0;
        red[1][threadIdx.x >> 6] = m1;
        red[2][threadIdx.x >> 6] = m2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double m = 0.;
        for (int w = 0; w < kNetThreads / 64; w++) m = fmax(m, red[threadIdx.x][w]);
        part[threadIdx.x * 1024 + blockIdx.x] = m;
    }
}

// one wave per scale: lanes stride over the block maxima, then a wave maximum (max is exact in
// any order)
__global__ __launch_bounds__(192) void k_scales_finish(const double* __restrict__ part, int nblocks, double* __restrict__ s) {
    const int which = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double m = 0.;
    for (int b = lane; b < nblocks; b += 64) m = fmax(m, part[which * 1024 + b]);
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off));
    // floors keep the channels finite when a field is flat (first step of a trajectory)
    if (lane == 0) s[which] = fmax(m, 1e-300);
}

// x0[v][c]: c0 = phi^n / s0, c1 = (phi^n - phi^{n-1}) / s1, c2 = resid * resid_scale / s2,
// c3.. = (extra_e - offset_e) * scale_e (static fields), the rest 0.
struct Extras {
    const double* f[8];
    double scale[8], offset[8];
    int n;
};

template <int CF>
__global__ __launch_bounds__(kNetThreads) void k_pack(FieldGeo F, const double* __restrict__ phi,
                                                     const double* __restrict__ phim, const double* __restrict__ resid,
                                                     double resid_scale, Extras ex, const double* __restrict__ s,
                                                     uint16_t* __restrict__ x0) {
    // channel values of a voxel: their mean over its CF^3 cells (CF = 1: the cell's values)
    const long long n = static_cast<long long>(F.c0) * F.c1 * F.c2;
    const double i0 = 1. / s[0], i1 = 1. / s[1], i2 = 1. / s[2];
    const double avg = 1. / (CF * CF * CF);
    for (long long q = blockIdx.x * (long long)kNetThreads + threadIdx.x; q < n; q += (long long)gridDim.x * kNetThreads) {
        int z, y, x;
        box_zyx(q, F.c1, F.c2, z, y, x);
        double acc[kInChannels];
        for (int c = 0; c < kInChannels; c++) acc[c] = 0.;
        for (int a = 0; a < CF; a++)
            for (int b = 0; b < CF; b++)
                for (int d = 0; d < CF; d++) {
                    const int zz = CF * z + a, yy = CF * y + b, xx = CF * x + d;
                    const long long pi = pad_index(F, zz, yy, xx);
                    const double p = phi[pi];
                    acc[0] += p * i0;
                    acc[1] += (p - phim[pi]) * i1;
                    acc[2] += resid[row_index(F, zz, yy, xx)] * resid_scale * i2;
                    for (int c = 3; c < kInChannels; c++) {
                        const int e = c - 3;
                        if (e < ex.n) acc[c] += (ex.f[e][pi] - ex.offset[e]) * ex.scale[e];
                    }
                }
        uint16_t ch[kInChannels];
        for (int c = 0; c < kInChannels; c++) ch[c] = f2bf(static_cast<float>(acc[c] * avg));
        for (int c = 0; c < kInChannels; c++) x0[q * kInChannels + c] = ch[c];
    }
}

// x[row] = base[row] + s2 * out (the guess, solver row order; base = the classical guess
// whose residual the network saw). CF = 2: out is read at the cell through the trilinear
// interpolation of the voxel values (cell centres: weights 3/4 and 1/4 of the two nearest
// voxels per axis, the nearest one repeated at the box faces).
template <int CF>
__global__ __launch_bounds__(kNetThreads) void k_unpack(FieldGeo F, const double* __restrict__ base,
                                                       const float* __restrict__ out, const double* __restrict__ s,
                                                       double* __restrict__ xg) {
    const long long n = static_cast<long long>(F.n0) * F.n1 * F.n2;
    const double s2 = s[2];
    for (long long q = blockIdx.x * (long long)kNetThreads + threadIdx.x; q < n; q += (long long)gridDim.x * kNetThreads) {
        int z, y, x;
        box_zyx(q, F.n1, F.n2, z, y, x);
        const long long r = row_index(F, z, y, x);
        double o;
        if (CF == 1) {
            o = static_cast<double>(out[q]);
        } else {
            const int p[3] = {z, y, x}, nc[3] = {F.c0, F.c1, F.c2};
            int v0[3], v1[3];
            for (int d = 0; d < 3; d++) {
                v0[d] = p[d] >> 1;
                const int other = v0[d] + ((p[d] & 1) ? 1 : -1);
                v1[d] = other < 0 ? 0 : (other >= nc[d] ? nc[d] - 1 : other);
            }
            o = 0.;
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++)
                    for (int c = 0; c < 2; c++) {
                        const double w = (a ? 0.25 : 0.75) * (b ? 0.25 : 0.75) * (c ? 0.25 : 0.75);
                        const long long vz = a ? v1[0] : v0[0], vy = b ? v1[1] : v0[1], vx = c ? v1[2] : v0[2];
                        o += w * static_cast<double>(out[(vz * F.c1 + vy) * F.c2 + vx]);
                    }
        }
        xg[r] = base[r] + s2 * o;
    }
}

// loss = mean((out - t)^2) over the voxels, t = the mean over the voxel's cells of
// (x_true - base) / s2; dout[v][0] = 2 (out - t) / n, other channels 0.
template <int CF>
__global__ __launch_bounds__(kNetThreads) void k_mse(FieldGeo F, const double* __restrict__ base,
                                                    const double* __restrict__ xt, const float* __restrict__ out,
                                                    const double* __restrict__ s, uint16_t* __restrict__ dout,
                                                    double* __restrict__ part) {
    __shared__ double red[kNetThreads / 64];
    const long long n = static_cast<long long>(F.c0) * F.c1 * F.c2;
    const double inv = 1. / s[2];
    const double avg = 1. / (CF * CF * CF);
    double acc = 0.;
    for (long long q = blockIdx.x * (long long)kNetThreads + threadIdx.x; q < n; q += (long long)gridDim.x * kNetThreads) {
        int z, y, x;
        box_zyx(q, F.c1, F.c2, z, y, x);
        double tsum = 0.;
        for (int a = 0; a < CF; a++)
            for (int b = 0; b < CF; b++)
                for (int c = 0; c < CF; c++) {
                    const long long r = row_index(F, CF * z + a, CF * y + b, CF * x + c);
                    tsum += (xt[r] - base[r]) * inv;
                }
        const double t = tsum * avg;
        const double d = static_cast<double>(out[q]) - t;
        acc += d * d;
        dout[q * kOutWidth] = f2bf(static_cast<float>(2. * d / static_cast<double>(n)));
        for (int c = 1; c < kOutWidth; c++) dout[q * kOutWidth + c] = 0;
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.;
        for (int w = 0; w < kNetThreads / 64; w++) t += red[w];
        part[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(64) void k_mse_finish(const double* __restrict__ part, int nblocks, long long n,
                                                  double* __restrict__ loss) {
    if (threadIdx.x != 0) return;
    double t = 0.;
    for (int b = 0; b < nblocks; b++) t += part[b];
    loss[0] = t / static_cast<double>(n);
}

// ------------------------------------------------------------------------------ parameters
// Adam (Kingma & Ba 2015) on the FP32 master weights of one layer, then the two BF16 copies:
// forward [tap][co][ci] and data-backward [26 - tap][ci][co].
__global__ __launch_bounds__(kNetThreads) void k_adam_cast(float* __restrict__ w, const float* __restrict__ g,
                                                          float* __restrict__ m, float* __restrict__ v,
                                                          uint16_t* __restrict__ wf, uint16_t* __restrict__ wb,
                                                          int cin, int cout, int update, float lr, float b1, float b2,
                                                          float eps, float c1, float c2) {
    const int n = 27 * cin * cout;
    const int idx = blockIdx.x * kNetThreads + threadIdx.x;
    if (idx >= n) return;
    float wv = w[idx];
    if (update) {
        const float gv = g[idx];
        const float mv = b1 * m[idx] + (1.f - b1) * gv;
        const float vv = b2 * v[idx] + (1.f - b2) * gv * gv;
        m[idx] = mv;
        v[idx] = vv;
        wv -= lr * (mv / c1) / (sqrtf(vv / c2) + eps);
        w[idx] = wv;
    }
    const int ci = idx % cin;
    const int co = (idx / cin) % cout;
    const int tap = idx / (cin * cout);
    const uint16_t h = f2bf(wv);
    wf[idx] = h;
    wb[((26 - tap) * cin + ci) * cout + co] = h;
}

inline unsigned ew_grid(long long n) {
    long long g = (n + kNetThreads - 1) / kNetThreads;
    if (g > 8192) g = 8192;
    return static_cast<unsigned>(g < 1 ? 1 : g);
}

}  // namespace
}  // namespace dare_hip

using namespace dare_hip;

struct dare_convnet {
    int device = 0;
    int in_channels = 3;
    int max_dims[3] = {0, 0, 0};
    int dims[3] = {0, 0, 0};              // current full-resolution dims (divisible by 4)
    int cf = 1;                           // the network sees the grid at 1/cf resolution
    long long max_v0 = 0;
    // parameters
    long long off[kLayers + 1] = {};
    long long nparams = 0;
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *wf = nullptr, *wb = nullptr;
    uint16_t *ws6 = nullptr, *ws8 = nullptr;   // folded sub-pixel weights of layers 6 and 8
    bool folded_stale = true;                   // ws6 / ws8 not yet folded from the current wf
    int adam_t = 0;
    bool inference = false;    // forward skips a7, a9 (read only by the backward pass)
    bool acts_for_bwd = false; // the last forward pass wrote every activation bwd reads
    // forward activations
    uint16_t *x0 = nullptr, *a1 = nullptr, *e1 = nullptr, *p1 = nullptr, *a3 = nullptr, *e2 = nullptr, *p2 = nullptr,
             *a5 = nullptr, *e3 = nullptr, *u2 = nullptr, *a7 = nullptr, *s2 = nullptr, *d2 = nullptr, *u1 = nullptr,
             *a9 = nullptr, *s1 = nullptr, *d1 = nullptr;
    float* out = nullptr;
    // backward buffers (allocated on first use)
    bool bwd_ready = false;
    uint16_t *dout = nullptr, *t0a = nullptr, *t0b = nullptr, *t0c = nullptr, *t0w = nullptr, *h1a = nullptr,
             *h1b = nullptr, *h1c = nullptr, *h1w = nullptr, *q2a = nullptr, *q2b = nullptr;
    float* wpart = nullptr;
    long long wpart_floats = 0;
    // scalars
    double* scales = nullptr;             // 3
    double* part = nullptr;               // 3 * 1024
    double* loss = nullptr;
    double* h_pinned = nullptr;           // 4
    // dare_convnet_predict: pack -> forward (inference) -> unpack captured once as a HIP graph
    // on a private stream and replayed while every argument stays the same
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t predict_exec = nullptr;
    std::vector<unsigned char> predict_key;
};

namespace {

constexpr long long kWpartFloats = 16LL << 20;

template <typename T>
int dalloc(T** p, long long n) {
    if (hipMalloc(p, sizeof(T) * (n > 0 ? n : 1)) != hipSuccess) {
        (void)hipGetLastError();
        set_error("convnet allocation failed");
        return DARE_ERR_OOM;
    }
    return DARE_OK;
}

long long level_voxels(const int* d, int level) {
    return static_cast<long long>(d[0] >> level) * (d[1] >> level) * (d[2] >> level);
}

// compute units of the current device and resident 256-thread blocks of a kernel (queried
// once per instantiation; function-local statics are initialised once, thread-safely)
inline int device_cus() {
    static const int n = [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            return v;
        return 256;
    }();
    return n;
}

template <auto K>   // one cached value per kernel (the kernels share one function-pointer type)
int resident_blocks() {
    static const int n = [] {
        int v = 0;
        return (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, K, kNetThreads, 0) == hipSuccess && v > 0) ? v : 1;
    }();
    return n;
}

template <int CIN, int COUT, bool RELU, int OUT, int MODE = 0>
void conv(hipStream_t sm, const uint16_t* X, const uint16_t* Wt, uint16_t* Y, float* Yf, const int* d, int level,
          const uint16_t* E = nullptr, uint16_t* Y2 = nullptr) {
    const int D = d[0] >> level, H = d[1] >> level, W = d[2] >> level;
    const dim3 blk(kNetThreads);
    const unsigned gy = static_cast<unsigned>((H + 3) / 4) * (COUT / (16 * conv_nbb<COUT>()));
    if constexpr (CIN == 8 || CIN == 16 || CIN == 32) {
        // z-marching form: the planes per block are chosen so that the whole grid is resident at
        // once (no tail round): z-blocks = resident blocks / (x, y, channel-group) columns. 32-
        // channel rings use tiles of at most 32 voxels along x (registers).
        const int nseg = (W >= 64 && CIN <= 16) ? 4 : (W >= 32 ? 2 : 1);
        const unsigned gx = static_cast<unsigned>((W + 16 * nseg - 1) / (16 * nseg));
        const unsigned gzm = static_cast<unsigned>((H + 3) / 4) * (COUT / 16);
        int per_cu = 1;
        if (nseg == 4) per_cu = resident_blocks<k_conv_zm<CIN, COUT, RELU, OUT, (CIN <= 16 ? 4 : 2), MODE>>();
        else if (nseg == 2) per_cu = resident_blocks<k_conv_zm<CIN, COUT, RELU, OUT, 2, MODE>>();
        else per_cu = resident_blocks<k_conv_zm<CIN, COUT, RELU, OUT, 1, MODE>>();
        const long long cols = static_cast<long long>(gx) * gzm;
        const long long nzb = std::max<long long>(1, static_cast<long long>(device_cus()) * per_cu / cols);
        const int zb = static_cast<int>((D + nzb - 1) / nzb);
        const dim3 grid(gx, gzm, static_cast<unsigned>((D + zb - 1) / zb));
        if (nseg == 4)
            hipLaunchKernelGGL((k_conv_zm<CIN, COUT, RELU, OUT, (CIN <= 16 ? 4 : 2), MODE>), grid, blk, 0, sm, X, Wt, Y,
                               Yf, D, H, W, zb, E, Y2);
        else if (nseg == 2)
            hipLaunchKernelGGL((k_conv_zm<CIN, COUT, RELU, OUT, 2, MODE>), grid, blk, 0, sm, X, Wt, Y, Yf, D, H, W, zb,
                               E, Y2);
        else
            hipLaunchKernelGGL((k_conv_zm<CIN, COUT, RELU, OUT, 1, MODE>), grid, blk, 0, sm, X, Wt, Y, Yf, D, H, W, zb,
                               E, Y2);
    } else {
        if (W >= 64) {
            hipLaunchKernelGGL((k_conv<CIN, COUT, RELU, OUT, 4, MODE>), dim3((W + 63) / 64, gy, D), blk, 0, sm, X, Wt,
                               Y, Yf, D, H, W, E, Y2);
        } else if (W >= 32) {
            hipLaunchKernelGGL((k_conv<CIN, COUT, RELU, OUT, 2, MODE>), dim3((W + 31) / 32, gy, D), blk, 0, sm, X, Wt,
                               Y, Yf, D, H, W, E, Y2);
        } else {
            hipLaunchKernelGGL((k_conv<CIN, COUT, RELU, OUT, 1, MODE>), dim3((W + 15) / 16, gy, D), blk, 0, sm, X, Wt,
                               Y, Yf, D, H, W, E, Y2);
        }
    }
}

// decoder layer: ReLU(conv(upsample(X))) and the skip sum Y2 = Y + E, in sub-pixel form (the
// 27 taps folded into 8 per output parity, Ws refreshed from the current BF16 weights first).
// Y may be NULL (inference: only the sum is written).
template <int CIN, int COUT>
void conv_up(hipStream_t sm, const uint16_t* X, const uint16_t* Ws, uint16_t* Y, const int* d, int level,
             const uint16_t* E, uint16_t* Y2) {
    static_assert(CIN % 32 == 0 && COUT % 16 == 0, "sub-pixel conv: 32-channel K slices, 16-channel groups");
    const int D = d[0] >> level, H = d[1] >> level, W = d[2] >> level;
    const dim3 blk(kNetThreads);
    const unsigned gy = static_cast<unsigned>((H + 3) / 4) * (COUT / 16);
    // tile width by registers (2 x 8 x CIN/32 weight fragments + the operands of two K steps):
    // two waves per SIMD at 32 (CIN 32) or 16 (CIN 64) coarse columns
    constexpr int NMAX = CIN == 32 ? 2 : 1;
    const int nseg = std::min(NMAX, W >= 128 ? 4 : (W >= 64 ? 2 : 1));
    const unsigned gx = static_cast<unsigned>((W + 32 * nseg - 1) / (32 * nseg));
    // z-march per parity: blocks per parity column sized so that the whole grid is resident
    int per_cu = 1;
    if (nseg == 2) per_cu = resident_blocks<k_conv_sub<CIN, COUT, true, 2, true>>();
    else per_cu = resident_blocks<k_conv_sub<CIN, COUT, true, 1, true>>();
    const int Dc = D >> 1;
    const long long cols = 2LL * gx * gy;
    const long long nzb = std::max<long long>(1, static_cast<long long>(device_cus()) * per_cu / cols);
    const int zb = static_cast<int>((Dc + nzb - 1) / nzb);
    const dim3 grid(gx, gy, 2u * static_cast<unsigned>((Dc + zb - 1) / zb));
    if (nseg == 2)
        hipLaunchKernelGGL((k_conv_sub<CIN, COUT, true, 2, true>), grid, blk, 0, sm, X, Ws, Y, D, H, W, zb, E, Y2);
    else
        hipLaunchKernelGGL((k_conv_sub<CIN, COUT, true, 1, true>), grid, blk, 0, sm, X, Ws, Y, D, H, W, zb, E, Y2);
}

// the sub-pixel weights of layers 6 and 8 folded from the current BF16 weights, once after
// every change of the weights (set_params, Adam step), outside any captured prediction graph
void fold_sub_weights(hipStream_t sm, dare_convnet* net) {
    if (!net->folded_stale) return;
    const dim3 blk(kNetThreads);
    constexpr int n6 = 64 * kDefs[6].cin * kDefs[6].cout, n8 = 64 * kDefs[8].cin * kDefs[8].cout;
    hipLaunchKernelGGL((k_sub_weights<kDefs[6].cin, kDefs[6].cout>), dim3((n6 + kNetThreads - 1) / kNetThreads), blk,
                       0, sm, net->wf + net->off[6], net->ws6);
    hipLaunchKernelGGL((k_sub_weights<kDefs[8].cin, kDefs[8].cout>), dim3((n8 + kNetThreads - 1) / kNetThreads), blk,
                       0, sm, net->wf + net->off[8], net->ws8);
    net->folded_stale = false;
}

// dispatch on (cin, cout) of the network's layers
void conv_any(hipStream_t sm, int cin, int cout, bool relu, const uint16_t* X, const uint16_t* Wt, uint16_t* Y,
              const int* d, int level) {
#define DARE_CONV_CASE(CI, CO)                                                     \
    if (cin == CI && cout == CO) {                                                 \
        if (relu) conv<CI, CO, true, 0>(sm, X, Wt, Y, nullptr, d, level);          \
        else conv<CI, CO, false, 0>(sm, X, Wt, Y, nullptr, d, level);              \
        return;                                                                    \
    }
    DARE_CONV_CASE(16, 16)
    DARE_CONV_CASE(16, 32)
    DARE_CONV_CASE(32, 16)
    DARE_CONV_CASE(32, 32)
    DARE_CONV_CASE(32, 64)
    DARE_CONV_CASE(64, 32)
    DARE_CONV_CASE(64, 64)
#undef DARE_CONV_CASE
}

int wgrad_any(dare_convnet* net, hipStream_t sm, int layer, const uint16_t* X, const uint16_t* G) {
    const int cin = kDefs[layer].cin, cout = kDefs[layer].cout, level = kDefs[layer].level;
    const int D = net->dims[0] >> level, H = net->dims[1] >> level, W = net->dims[2] >> level;
    const int rows = D * H;
    const int pairs = std::max(1, cin / 16) * (cout / 16);
    long long cap = std::min<long long>(512, kWpartFloats / (27LL * cin * cout));
    cap = std::min<long long>(cap, std::max(1, 4096 / pairs));
    const int chunks0 = static_cast<int>(std::max<long long>(1, std::min<long long>(rows, cap)));
    const int rpc = (rows + chunks0 - 1) / chunks0;
    const int nchunks = (rows + rpc - 1) / rpc;
    dim3 grid((nchunks + 3) / 4, pairs);
#define DARE_WG_CASE(CI, CO)                                                                              \
    if (cin == CI && cout == CO)                                                                          \
        hipLaunchKernelGGL((k_wgrad<CI, CO>), grid, dim3(kNetThreads), 0, sm, X, G, D, H, W, rpc, nchunks, \
                           net->wpart);
    DARE_WG_CASE(kInChannels, 16)
    DARE_WG_CASE(16, 16)
    DARE_WG_CASE(16, 32)
    DARE_WG_CASE(32, 16)
    DARE_WG_CASE(32, 32)
    DARE_WG_CASE(32, 64)
    DARE_WG_CASE(64, 32)
    DARE_WG_CASE(64, 64)
#undef DARE_WG_CASE
    const int n = 27 * cin * cout;
    hipLaunchKernelGGL(k_chunk_sum, dim3((n + kNetThreads - 1) / kNetThreads), dim3(kNetThreads), 0, sm, net->wpart,
                       nchunks, n, net->g + net->off[layer]);
    return DARE_OK;
}

FieldGeo field_geo(const dare_grid* g, const int32_t* crop, int cf) {
    FieldGeo F;
    F.n0 = crop ? crop[3] : g->n[0];
    F.n1 = crop ? crop[4] : g->n[1];
    F.n2 = crop ? crop[5] : g->n[2];
    F.o0 = crop ? crop[0] : 0;
    F.o1 = crop ? crop[1] : 0;
    F.o2 = crop ? crop[2] : 0;
    F.m1 = g->n[1];
    F.m2 = g->n[2];
    F.g = g->ghost;
    const long long p1 = g->n[1] + 2 * g->ghost, p2 = g->n[2] + 2 * g->ghost;
    F.s1 = p2;
    F.s0 = p1 * p2;
    F.cf = cf;
    F.c0 = F.n0 / cf;
    F.c1 = F.n1 / cf;
    F.c2 = F.n2 / cf;
    return F;
}

int check_grid(const dare_convnet* net, const dare_grid* g, const int32_t* crop) {
    DARE_CHECK_ARG(g && g->dim == 3 && g->ncomp == 1, "the warm-start network takes 3-D scalar grids");
    int box[3];
    for (int d = 0; d < 3; d++) {
        box[d] = crop ? crop[3 + d] : g->n[d];
        if (crop)
            DARE_CHECK_ARG(crop[d] >= 0 && crop[3 + d] >= 1 && crop[d] + crop[3 + d] <= g->n[d],
                           "crop outside the grid");
        DARE_CHECK_ARG(box[d] % (4 * net->cf) == 0 && box[d] >= 4 * net->cf,
                       net->cf == 1 ? "network box sizes must be multiples of 4"
                                    : "network box sizes must be multiples of 8 (coarsening 2)");
        DARE_CHECK_ARG(box[d] / net->cf <= net->max_dims[d], "box larger than the network was created for");
    }
    DARE_CHECK_ARG(static_cast<long long>(box[0] / net->cf) * (box[1] / net->cf) * (box[2] / net->cf) <= net->max_v0,
                   "box too large");
    DARE_CHECK_ARG(static_cast<long long>(box[0]) * box[1] * box[2] < (1LL << 32), "box too large");
    return DARE_OK;
}

}  // namespace

extern "C" {

int dare_convnet_create(dare_ctx* ctx, int32_t in_channels, const int32_t* max_dims, dare_convnet** out) {
    DARE_CHECK_ARG(ctx && max_dims && out, "NULL argument");
    DARE_CHECK_ARG(in_channels >= 3 && in_channels <= kInChannels, "in_channels must be in [3, 8]");
    for (int d = 0; d < 3; d++) DARE_CHECK_ARG(max_dims[d] >= 4 && max_dims[d] % 4 == 0, "dims must be multiples of 4");
    *out = nullptr;
    DARE_HIP_TRY(hipSetDevice(ctx->device));
    dare_convnet* net = new dare_convnet();
    net->device = ctx->device;
    net->in_channels = in_channels;
    for (int d = 0; d < 3; d++) net->max_dims[d] = net->dims[d] = max_dims[d];
    const long long V0 = static_cast<long long>(max_dims[0]) * max_dims[1] * max_dims[2];
    net->max_v0 = V0;
    const long long V1 = V0 / 8, V2 = V0 / 64;
    for (int l = 0; l < kLayers; l++) net->off[l + 1] = net->off[l] + 27LL * kDefs[l].cin * kDefs[l].cout;
    net->nparams = net->off[kLayers];
    int rc = DARE_OK;
    auto A = [&](auto** p, long long n) {
        if (rc == DARE_OK) rc = dalloc(p, n);
    };
    A(&net->w, net->nparams);
    A(&net->g, net->nparams);
    A(&net->m, net->nparams);
    A(&net->v, net->nparams);
    A(&net->wf, net->nparams);
    A(&net->wb, net->nparams);
    A(&net->ws6, 64LL * kDefs[6].cin * kDefs[6].cout);
    A(&net->ws8, 64LL * kDefs[8].cin * kDefs[8].cout);
    A(&net->x0, V0 * kInChannels);
    A(&net->a1, V0 * 16);
    A(&net->e1, V0 * 16);
    A(&net->p1, V1 * 16);
    A(&net->a3, V1 * 32);
    A(&net->e2, V1 * 32);
    A(&net->p2, V2 * 32);
    A(&net->a5, V2 * 64);
    A(&net->e3, V2 * 64);
    A(&net->u2, V1 * 64);
    A(&net->a7, V1 * 32);
    A(&net->s2, V1 * 32);
    A(&net->d2, V1 * 32);
    A(&net->u1, V0 * 32);
    A(&net->a9, V0 * 16);
    A(&net->s1, V0 * 16);
    A(&net->d1, V0 * 16);
    A(&net->out, V0);
    A(&net->scales, 3);
    A(&net->part, 3 * 1024);
    A(&net->loss, 1);
    if (rc == DARE_OK && hipHostMalloc(reinterpret_cast<void**>(&net->h_pinned), 4 * sizeof(double), 0) != hipSuccess) {
        (void)hipGetLastError();
        set_error("pinned allocation failed");
        rc = DARE_ERR_OOM;
    }
    if (rc == DARE_OK) {
        // zero parameters and Adam moments (kernels, so that nothing here is a memset node)
        std::vector<float> z(net->nparams, 0.f);
        if (hipMemcpy(net->w, z.data(), sizeof(float) * net->nparams, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(net->m, z.data(), sizeof(float) * net->nparams, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(net->v, z.data(), sizeof(float) * net->nparams, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(net->g, z.data(), sizeof(float) * net->nparams, hipMemcpyHostToDevice) != hipSuccess) {
            set_error("convnet initialisation failed");
            rc = DARE_ERR_HIP;
        }
    }
    if (rc != DARE_OK) {
        dare_convnet_destroy(net);
        return rc;
    }
    *out = net;
    return DARE_OK;
}

int dare_convnet_destroy(dare_convnet* net) {
    if (!net) return DARE_OK;
    (void)hipSetDevice(net->device);
    (void)hipDeviceSynchronize();
    if (net->predict_exec) (void)hipGraphExecDestroy(net->predict_exec);
    if (net->cap_stream) (void)hipStreamDestroy(net->cap_stream);
    void* ptrs[] = {net->w, net->g, net->m, net->v, net->wf, net->wb, net->ws6, net->ws8, net->x0, net->a1, net->e1, net->p1, net->a3,
                    net->e2, net->p2, net->a5, net->e3, net->u2, net->a7, net->s2, net->d2, net->u1, net->a9,
                    net->s1, net->d1, net->out, net->dout, net->t0a, net->t0b, net->t0c, net->t0w, net->h1a,
                    net->h1b, net->h1c, net->h1w, net->q2a, net->q2b, net->wpart, net->scales, net->part, net->loss};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (net->h_pinned) (void)hipHostFree(net->h_pinned);
    delete net;
    return DARE_OK;
}

int dare_convnet_num_params(const dare_convnet* net, int64_t* n, int64_t* layer_offsets) {
    DARE_CHECK_ARG(net && n, "NULL argument");
    *n = net->nparams;
    if (layer_offsets)
        for (int l = 0; l <= kLayers; l++) layer_offsets[l] = net->off[l];
    return DARE_OK;
}

// host FP32 master weights in, both BF16 copies refreshed; Adam state reset
int dare_convnet_set_params(dare_ctx* ctx, dare_convnet* net, const float* host) {
    DARE_CHECK_ARG(ctx && net && host, "NULL argument");
    DARE_HIP_TRY(hipMemcpyAsync(net->w, host, sizeof(float) * net->nparams, hipMemcpyHostToDevice, ctx->stream));
    DARE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::vector<float> z(net->nparams, 0.f);
    DARE_HIP_TRY(hipMemcpy(net->m, z.data(), sizeof(float) * net->nparams, hipMemcpyHostToDevice));
    DARE_HIP_TRY(hipMemcpy(net->v, z.data(), sizeof(float) * net->nparams, hipMemcpyHostToDevice));
    net->adam_t = 0;
    for (int l = 0; l < kLayers; l++) {
        const int n = 27 * kDefs[l].cin * kDefs[l].cout;
        hipLaunchKernelGGL(k_adam_cast, dim3((n + kNetThreads - 1) / kNetThreads), dim3(kNetThreads), 0, ctx->stream,
                           net->w + net->off[l], net->g + net->off[l], net->m + net->off[l], net->v + net->off[l],
                           net->wf + net->off[l], net->wb + net->off[l], kDefs[l].cin, kDefs[l].cout, 0, 0.f, 0.f, 0.f,
                           0.f, 1.f, 1.f);
    }
    net->folded_stale = true;
    DARE_LAUNCH_CHECK();
    return DARE_OK;
}

int dare_convnet_get_params(dare_ctx* ctx, const dare_convnet* net, float* host_w, float* host_grad) {
    DARE_CHECK_ARG(ctx && net, "NULL argument");
    if (host_w)
        DARE_HIP_TRY(hipMemcpyAsync(host_w, net->w, sizeof(float) * net->nparams, hipMemcpyDeviceToHost, ctx->stream));
    if (host_grad)
        DARE_HIP_TRY(hipMemcpyAsync(host_grad, net->g, sizeof(float) * net->nparams, hipMemcpyDeviceToHost, ctx->stream));
    DARE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return DARE_OK;
}

// Input tensor from the solver's fields (channels as declared in dare_hip.h); extras: up to
// in_channels - 3 static padded fields, (f - offset) * scale.
int dare_convnet_pack(dare_ctx* ctx, dare_convnet* net, const dare_grid* g, const double* phi, const double* phi_prev,
                      const double* resid, double resid_scale, int32_t nextra, const double* const* extras,
                      const double* extra_scale, const double* extra_offset, const int32_t* crop) {
    DARE_CHECK_ARG(ctx && net && phi && phi_prev && resid, "NULL argument");
    int rc = check_grid(net, g, crop);
    if (rc) return rc;
    DARE_CHECK_ARG(nextra >= 0 && nextra <= net->in_channels - 3 && nextra <= 8, "too many extra channels");
    const FieldGeo F = field_geo(g, crop, net->cf);
    net->dims[0] = F.c0;
    net->dims[1] = F.c1;
    net->dims[2] = F.c2;
    const long long n = static_cast<long long>(F.n0) * F.n1 * F.n2;
    const long long nv = static_cast<long long>(F.c0) * F.c1 * F.c2;
    Extras ex{};
    ex.n = nextra;
    for (int e = 0; e < nextra; e++) {
        DARE_CHECK_ARG(extras && extras[e] && extra_scale && extra_offset, "NULL extra field");
        ex.f[e] = extras[e];
        ex.scale[e] = extra_scale[e];
        ex.offset[e] = extra_offset[e];
    }
    const unsigned gs = std::min<unsigned>(ew_grid(n), 1024);
    hipLaunchKernelGGL(k_scales, dim3(gs), dim3(kNetThreads), 0, ctx->stream, F, phi, phi_prev, resid, resid_scale,
                       net->part);
    hipLaunchKernelGGL(k_scales_finish, dim3(1), dim3(192), 0, ctx->stream, net->part, static_cast<int>(gs), net->scales);
    if (net->cf == 2)
        hipLaunchKernelGGL(k_pack<2>, dim3(ew_grid(nv)), dim3(kNetThreads), 0, ctx->stream, F, phi, phi_prev, resid,
                           resid_scale, ex, net->scales, net->x0);
    else
        hipLaunchKernelGGL(k_pack<1>, dim3(ew_grid(nv)), dim3(kNetThreads), 0, ctx->stream, F, phi, phi_prev, resid,
                           resid_scale, ex, net->scales, net->x0);
    DARE_LAUNCH_CHECK();
    return DARE_OK;
}

int dare_conv_net_fwd(dare_ctx* ctx, dare_convnet* net) {
    DARE_CHECK_ARG(ctx && net, "NULL argument");
    hipStream_t sm = ctx->stream;
    fold_sub_weights(sm, net);
    const int* d = net->dims;
    const long long V1 = level_voxels(d, 1);
    const long long V0 = level_voxels(d, 0);
    const uint16_t* W = net->wf;
    const long long* o = net->off;
    conv<kInChannels, 16, true, 0>(sm, net->x0, W + o[0], net->a1, nullptr, d, 0);
    conv<16, 16, true, 0>(sm, net->a1, W + o[1], net->e1, nullptr, d, 0);
    hipLaunchKernelGGL(k_pool, dim3(ew_grid(V1 * 16)), dim3(kNetThreads), 0, sm, net->e1, net->p1, 16, d[0] >> 1,
                       d[1] >> 1, d[2] >> 1);
    conv<16, 32, true, 0>(sm, net->p1, W + o[2], net->a3, nullptr, d, 1);
    conv<32, 32, true, 0>(sm, net->a3, W + o[3], net->e2, nullptr, d, 1);
    hipLaunchKernelGGL(k_pool, dim3(ew_grid(level_voxels(d, 2) * 32)), dim3(kNetThreads), 0, sm, net->e2, net->p2, 32,
                       d[0] >> 2, d[1] >> 2, d[2] >> 2);
    conv<32, 64, true, 0>(sm, net->p2, W + o[4], net->a5, nullptr, d, 2);
    conv<64, 64, true, 0>(sm, net->a5, W + o[5], net->e3, nullptr, d, 2);
    // decoder: the upsampling is read inside the conv (u2, u1 are written by the backward pass,
    // which needs them) and the skip sums s2 = a7 + e2, s1 = a9 + e1 come out of its epilogue
    // in inference mode the ReLU outputs a7, a9 (the backward pass's masks) are not stored
    conv_up<64, 32>(sm, net->e3, net->ws6, net->inference ? nullptr : net->a7, d, 1, net->e2, net->s2);
    conv<32, 32, true, 0>(sm, net->s2, W + o[7], net->d2, nullptr, d, 1);
    conv_up<32, 16>(sm, net->d2, net->ws8, net->inference ? nullptr : net->a9, d, 0, net->e1, net->s1);
    conv<16, 16, true, 0>(sm, net->s1, W + o[9], net->d1, nullptr, d, 0);
    conv<16, 16, false, 1>(sm, net->d1, W + o[10], nullptr, net->out, d, 0);
    DARE_LAUNCH_CHECK();
    net->acts_for_bwd = !net->inference;
    return DARE_OK;
}

int dare_convnet_set_inference(dare_convnet* net, int32_t on) {
    DARE_CHECK_ARG(net, "NULL argument");
    net->inference = on != 0;
    return DARE_OK;
}

int dare_convnet_set_coarsening(dare_convnet* net, int32_t factor) {
    DARE_CHECK_ARG(net, "NULL argument");
    DARE_CHECK_ARG(factor == 1 || factor == 2, "coarsening factor must be 1 or 2");
    net->cf = factor;
    return DARE_OK;
}

// x_guess[row] = base[row] + S * out[row]
int dare_convnet_unpack(dare_ctx* ctx, dare_convnet* net, const dare_grid* g, const double* base, double* x_guess,
                        const int32_t* crop) {
    DARE_CHECK_ARG(ctx && net && base && x_guess, "NULL argument");
    int rc = check_grid(net, g, crop);
    if (rc) return rc;
    const FieldGeo F = field_geo(g, crop, net->cf);
    DARE_CHECK_ARG(F.c0 == net->dims[0] && F.c1 == net->dims[1] && F.c2 == net->dims[2],
                   "box differs from the one of the last forward pass");
    const long long n = static_cast<long long>(F.n0) * F.n1 * F.n2;
    if (net->cf == 2)
        hipLaunchKernelGGL(k_unpack<2>, dim3(ew_grid(n)), dim3(kNetThreads), 0, ctx->stream, F, base, net->out,
                           net->scales, x_guess);
    else
        hipLaunchKernelGGL(k_unpack<1>, dim3(ew_grid(n)), dim3(kNetThreads), 0, ctx->stream, F, base, net->out,
                           net->scales, x_guess);
    DARE_LAUNCH_CHECK();
    return DARE_OK;
}

// MSE against the solver's solution x_true (row order) and the output gradient; loss to host.
int dare_convnet_loss(dare_ctx* ctx, dare_convnet* net, const dare_grid* g, const double* base, const double* x_true,
                      double* loss_host, const int32_t* crop) {
    DARE_CHECK_ARG(ctx && net && base && x_true, "NULL argument");
    int rc = check_grid(net, g, crop);
    if (rc) return rc;
    if (!net->dout) {
        rc = dalloc(&net->dout, net->max_v0 * 16);
        if (rc) return rc;
    }
    const FieldGeo F = field_geo(g, crop, net->cf);
    DARE_CHECK_ARG(F.c0 == net->dims[0] && F.c1 == net->dims[1] && F.c2 == net->dims[2],
                   "box differs from the one of the last forward pass");
    const long long n = static_cast<long long>(F.c0) * F.c1 * F.c2;
    const unsigned gs = std::min<unsigned>(ew_grid(n), 1024);
    if (net->cf == 2)
        hipLaunchKernelGGL(k_mse<2>, dim3(gs), dim3(kNetThreads), 0, ctx->stream, F, base, x_true, net->out,
                           net->scales, net->dout, net->part);
    else
        hipLaunchKernelGGL(k_mse<1>, dim3(gs), dim3(kNetThreads), 0, ctx->stream, F, base, x_true, net->out,
                           net->scales, net->dout, net->part);
    hipLaunchKernelGGL(k_mse_finish, dim3(1), dim3(64), 0, ctx->stream, net->part, static_cast<int>(gs), n, net->loss);
    DARE_LAUNCH_CHECK();
    if (loss_host) {
        DARE_HIP_TRY(hipMemcpyAsync(net->h_pinned, net->loss, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        DARE_HIP_TRY(hipStreamSynchronize(ctx->stream));
        *loss_host = net->h_pinned[0];
    }
    return DARE_OK;
}

// Back-propagation of the last loss through the last forward pass: weight gradients (FP32).
int dare_conv_net_bwd(dare_ctx* ctx, dare_convnet* net) {
    DARE_CHECK_ARG(ctx && net && net->dout, "call dare_convnet_loss first");
    DARE_CHECK_ARG(net->acts_for_bwd, "backward needs a forward pass outside inference mode");
    if (!net->bwd_ready) {
        const long long V0 = net->max_v0, V1 = V0 / 8, V2 = V0 / 64;
        int rc = DARE_OK;
        auto A = [&](auto** p, long long n) {
            if (rc == DARE_OK) rc = dalloc(p, n);
        };
        A(&net->t0a, V0 * 16);
        A(&net->t0b, V0 * 16);
        A(&net->t0c, V0 * 16);
        A(&net->t0w, V0 * 32);
        A(&net->h1a, V1 * 32);
        A(&net->h1b, V1 * 32);
        A(&net->h1c, V1 * 32);
        A(&net->h1w, V1 * 64);
        A(&net->q2a, V2 * 64);
        A(&net->q2b, V2 * 64);
        A(&net->wpart, kWpartFloats);
        if (rc) return rc;
        net->bwd_ready = true;
    }
    hipStream_t sm = ctx->stream;
    const int* d = net->dims;
    const long long V0 = level_voxels(d, 0), V1 = level_voxels(d, 1), V2 = level_voxels(d, 2);
    const uint16_t* WB = net->wb;
    const long long* o = net->off;
    // L11 (no ReLU): dd1 = conv_bwd(dout)
    wgrad_any(net, sm, 10, net->d1, net->dout);
    conv_any(sm, 16, 16, false, net->dout, WB + o[10], net->t0a, d, 0);            // dd1
    hipLaunchKernelGGL(k_mask, dim3(ew_grid(V0 * 16)), dim3(kNetThreads), 0, sm, net->t0a, net->d1, net->t0b, V0 * 16);
    // L10: g10 = t0b
    wgrad_any(net, sm, 9, net->s1, net->t0b);
    conv_any(sm, 16, 16, false, net->t0b, WB + o[9], net->t0c, d, 0);              // ds1 (kept)
    hipLaunchKernelGGL(k_mask, dim3(ew_grid(V0 * 16)), dim3(kNetThreads), 0, sm, net->t0c, net->a9, net->t0a, V0 * 16);
    // L9: g9 = t0a (u1 = up(d2), read fused by the forward pass, materialized here)
    hipLaunchKernelGGL(k_up, dim3(ew_grid(V0 * 32)), dim3(kNetThreads), 0, sm, net->d2, net->u1, 32, d[0], d[1], d[2]);
    wgrad_any(net, sm, 8, net->u1, net->t0a);
    conv_any(sm, 16, 32, false, net->t0a, WB + o[8], net->t0w, d, 0);              // du1 (32 ch)
    hipLaunchKernelGGL(k_up_bwd_mask, dim3(ew_grid(V1 * 32)), dim3(kNetThreads), 0, sm, net->t0w, net->d2, net->h1a,
                       32, d[0] >> 1, d[1] >> 1, d[2] >> 1);                       // g8
    // L8
    wgrad_any(net, sm, 7, net->s2, net->h1a);
    conv_any(sm, 32, 32, false, net->h1a, WB + o[7], net->h1b, d, 1);              // ds2 (kept)
    hipLaunchKernelGGL(k_mask, dim3(ew_grid(V1 * 32)), dim3(kNetThreads), 0, sm, net->h1b, net->a7, net->h1c, V1 * 32);
    // L7: g7 = h1c (u2 = up(e3), as u1)
    hipLaunchKernelGGL(k_up, dim3(ew_grid(V1 * 64)), dim3(kNetThreads), 0, sm, net->e3, net->u2, 64, d[0] >> 1,
                       d[1] >> 1, d[2] >> 1);
    wgrad_any(net, sm, 6, net->u2, net->h1c);
    conv_any(sm, 32, 64, false, net->h1c, WB + o[6], net->h1w, d, 1);              // du2 (64 ch)
    hipLaunchKernelGGL(k_up_bwd_mask, dim3(ew_grid(V2 * 64)), dim3(kNetThreads), 0, sm, net->h1w, net->e3, net->q2a,
                       64, d[0] >> 2, d[1] >> 2, d[2] >> 2);                       // g6
    // L6
    wgrad_any(net, sm, 5, net->a5, net->q2a);
    conv_any(sm, 64, 64, false, net->q2a, WB + o[5], net->q2b, d, 2);              // da5
    hipLaunchKernelGGL(k_mask, dim3(ew_grid(V2 * 64)), dim3(kNetThreads), 0, sm, net->q2b, net->a5, net->q2a, V2 * 64);
    // L5: g5 = q2a
    wgrad_any(net, sm, 4, net->p2, net->q2a);
    conv_any(sm, 64, 32, false, net->q2a, WB + o[4], net->q2b, d, 2);              // dp2 (32 ch, quarter)
    hipLaunchKernelGGL(k_pool_bwd_add_mask, dim3(ew_grid(V1 * 32)), dim3(kNetThreads), 0, sm, net->q2b, net->h1b,
                       net->e2, net->h1a, 32, d[0] >> 1, d[1] >> 1, d[2] >> 1);     // g4
    // L4
    wgrad_any(net, sm, 3, net->a3, net->h1a);
    conv_any(sm, 32, 32, false, net->h1a, WB + o[3], net->h1c, d, 1);              // da3
    hipLaunchKernelGGL(k_mask, dim3(ew_grid(V1 * 32)), dim3(kNetThreads), 0, sm, net->h1c, net->a3, net->h1a, V1 * 32);
    // L3: g3 = h1a
    wgrad_any(net, sm, 2, net->p1, net->h1a);
    conv_any(sm, 32, 16, false, net->h1a, WB + o[2], net->h1c, d, 1);              // dp1 (16 ch, half)
    hipLaunchKernelGGL(k_pool_bwd_add_mask, dim3(ew_grid(V0 * 16)), dim3(kNetThreads), 0, sm, net->h1c, net->t0c,
                       net->e1, net->t0a, 16, d[0], d[1], d[2]);                   // g2
    // L2
    wgrad_any(net, sm, 1, net->a1, net->t0a);
    conv_any(sm, 16, 16, false, net->t0a, WB + o[1], net->t0b, d, 0);              // da1
    hipLaunchKernelGGL(k_mask, dim3(ew_grid(V0 * 16)), dim3(kNetThreads), 0, sm, net->t0b, net->a1, net->t0a, V0 * 16);
    // L1: g1 = t0a
    wgrad_any(net, sm, 0, net->x0, net->t0a);
    DARE_LAUNCH_CHECK();
    (void)V2;
    return DARE_OK;
}

// One Adam step on every layer with the gradients of the last backward pass.
int dare_conv_net_step(dare_ctx* ctx, dare_convnet* net, float lr, float beta1, float beta2, float eps) {
    DARE_CHECK_ARG(ctx && net, "NULL argument");
    DARE_CHECK_ARG(lr >= 0.f && beta1 >= 0.f && beta1 < 1.f && beta2 >= 0.f && beta2 < 1.f && eps > 0.f,
                   "invalid Adam parameters");
    net->adam_t++;
    const float c1 = 1.f - std::pow(beta1, static_cast<float>(net->adam_t));
    const float c2 = 1.f - std::pow(beta2, static_cast<float>(net->adam_t));
    for (int l = 0; l < kLayers; l++) {
        const int n = 27 * kDefs[l].cin * kDefs[l].cout;
        hipLaunchKernelGGL(k_adam_cast, dim3((n + kNetThreads - 1) / kNetThreads), dim3(kNetThreads), 0, ctx->stream,
                           net->w + net->off[l], net->g + net->off[l], net->m + net->off[l], net->v + net->off[l],
                           net->wf + net->off[l], net->wb + net->off[l], kDefs[l].cin, kDefs[l].cout, 1, lr, beta1,
                           beta2, eps, c1, c2);
    }
    net->folded_stale = true;
    DARE_LAUNCH_CHECK();
    return DARE_OK;
}

// Device pointers of the current tensors (tests): x0, out, dout, and the forward activations
// in network order (a1, e1, p1, a3, e2, p2, a5, e3, u2, a7, s2, d2, u1, a9, s1, d1).
int dare_convnet_tensors(const dare_convnet* net, void** ptrs /* 19 */, double* scales_host) {
    DARE_CHECK_ARG(net && ptrs, "NULL argument");
    void* p[19] = {net->x0, net->out, net->dout, net->a1, net->e1, net->p1, net->a3, net->e2, net->p2, net->a5,
                   net->e3, net->u2, net->a7, net->s2, net->d2, net->u1, net->a9, net->s1, net->d1};
    for (int q = 0; q < 19; q++) ptrs[q] = p[q];
    if (scales_host) DARE_HIP_TRY(hipMemcpy(scales_host, net->scales, 3 * sizeof(double), hipMemcpyDeviceToHost));
    return DARE_OK;
}

// pack -> forward in inference mode -> unpack, replayed from a HIP graph: the kernels of the
// three steps are captured on the network's own stream at the first call and whenever an
// argument (a pointer, the grid, a scale) differs from the captured call's, then launched as
// one graph on the context's stream. The graph holds kernels only.
int dare_convnet_predict(dare_ctx* ctx, dare_convnet* net, const dare_grid* g, const double* phi,
                         const double* phi_prev, const double* resid, double resid_scale, int32_t nextra,
                         const double* const* extras, const double* extra_scale, const double* extra_offset,
                         const double* x_base, double* x_guess) {
    DARE_CHECK_ARG(ctx && net && g && phi && phi_prev && resid && x_base && x_guess, "NULL argument");
    DARE_CHECK_ARG(nextra >= 0 && nextra <= 8 && (nextra == 0 || (extras && extra_scale && extra_offset)),
                   "invalid extra fields");
    std::vector<unsigned char> key;
    auto put = [&](const void* p, size_t n) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        key.insert(key.end(), c, c + n);
    };
    put(g, sizeof(dare_grid));
    const void* ptrs[5] = {phi, phi_prev, resid, x_base, x_guess};
    put(ptrs, sizeof(ptrs));
    put(&resid_scale, sizeof(double));
    put(&nextra, sizeof(nextra));
    for (int e = 0; e < nextra; e++) {
        put(&extras[e], sizeof(void*));
        put(&extra_scale[e], sizeof(double));
        put(&extra_offset[e], sizeof(double));
    }
    put(&net->cf, sizeof(int));
    put(&ctx->device, sizeof(int));
    // the folded sub-pixel weights are refreshed on the context stream, never inside the graph,
    // so a graph replays correctly across weight updates
    fold_sub_weights(ctx->stream, net);
    if (!net->predict_exec || key != net->predict_key) {
        if (net->predict_exec) (void)hipGraphExecDestroy(net->predict_exec);
        net->predict_exec = nullptr;
        net->predict_key.clear();
        if (!net->cap_stream) DARE_HIP_TRY(hipStreamCreateWithFlags(&net->cap_stream, hipStreamNonBlocking));
        dare_ctx cc = *ctx;          // the same device, the network's capture stream
        cc.stream = net->cap_stream;
        const bool was_inference = net->inference;
        net->inference = true;
        DARE_HIP_TRY(hipStreamBeginCapture(net->cap_stream, hipStreamCaptureModeThreadLocal));
        int rc = dare_convnet_pack(&cc, net, g, phi, phi_prev, resid, resid_scale, nextra, extras, extra_scale,
                                   extra_offset, nullptr);
        if (rc == DARE_OK) rc = dare_conv_net_fwd(&cc, net);
        if (rc == DARE_OK) rc = dare_convnet_unpack(&cc, net, g, x_base, x_guess, nullptr);
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(net->cap_stream, &graph);
        net->inference = was_inference;
        if (rc != DARE_OK || ec != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipGetLastError();
            if (rc == DARE_OK) {
                set_error("graph capture of the prediction failed");
                rc = DARE_ERR_HIP;
            }
            return rc;
        }
        const hipError_t ei = hipGraphInstantiate(&net->predict_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            (void)hipGetLastError();
            net->predict_exec = nullptr;
            set_error("graph instantiation of the prediction failed");
            return DARE_ERR_HIP;
        }
        net->predict_key = key;
    }
    DARE_HIP_TRY(hipGraphLaunch(net->predict_exec, ctx->stream));
    net->acts_for_bwd = false;
    return DARE_OK;
}

}  // extern "C"
