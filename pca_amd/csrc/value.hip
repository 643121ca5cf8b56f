// value.hip -- BF16 MFMA kernels of the value network (include/pcavalue.h).
//
// GEMM: one 256-thread workgroup computes a 128 x 128 tile of C, each wave
// 64 x 64 as 2 x 2 blocks of v_mfma_f32_32x32x16_bf16 (gfx950), FP32
// accumulators in registers.  Each K step stages a 128 x 32 slab of A and
// of B in LDS, both k-contiguous per row/column so that a lane's MFMA
// operand (8 consecutive k of one row of A or one column of B) is one
// 16-byte LDS read; the slab is gathered from HBM in 8-byte pieces along
// whichever dimension is contiguous in memory (the transposed storage modes
// only change that).  Out-of-range rows, columns and k are zero-filled, so
// any M, N, K works.  Few output tiles with a long K (weight gradients) are
// split along K across workgroups, each range's partial tile stored to a
// workspace slice and the slices summed by a second kernel.  Workgroups that
// read the same rows of A run on one XCD (its L2 serves the slab to all of
// them).  Interior slabs issue all their global loads before the first LDS
// store waits on one; a row-contiguous slab is transposed in registers in
// 4 x 4 blocks.  The epilogue adds the bias (staged in LDS), applies ReLU
// and stores FP32 and/or BF16, four adjacent columns per store; a BF16-only
// tile goes through LDS and out in 16-byte row runs.
//
// Operand layout of v_mfma_f32_32x32x16_bf16 (lane l of 64): A: row l % 32,
// k = 8 * (l / 32) .. +7; B: column l % 32, the same k; the 16 accumulators
// hold C rows 8 * (r / 4) + 4 * (l / 32) + r % 4 of column l % 32.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <utility>

#include "../../include/pcagpu.h"
#include "../../include/pcavalue.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int GM = 128, GN = 128, GK = 32;
constexpr int LDS_K = GK + 8;   // row stride in halves: 80 bytes, 16-byte aligned rows

__device__ __forceinline__ uint16_t to_bf16(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even (inputs are finite)
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Four consecutive elements p[0..3] of a row that holds `avail` more valid
// elements from p: one 8-byte load when all four are valid and p is 8-byte
// aligned, else element by element (zero past the end).
__device__ __forceinline__ void load4(const uint16_t *p, int avail, uint16_t v[4])
{
    if (avail >= 4 && (reinterpret_cast<uintptr_t>(p) & 7u) == 0) {
        const uint2 q = *reinterpret_cast<const uint2 *>(p);
        v[0] = (uint16_t)q.x;
        v[1] = (uint16_t)(q.x >> 16);
        v[2] = (uint16_t)q.y;
        v[3] = (uint16_t)(q.y >> 16);
    } else {
#pragma unroll
        for (int j = 0; j < 4; j++) v[j] = j < avail ? p[j] : (uint16_t)0;
    }
}

// Stages a 128 x 32 slab into S[row][k] (k contiguous) from a matrix whose
// storage is KCONTIG: X[row * ld + k] (k contiguous in memory) or else
// X[k * ld + row] (rows contiguous); rows >= R or k >= K read as zero.
template <bool KCONTIG>
__device__ __forceinline__ void stage(uint16_t (*S)[LDS_K], const uint16_t *X, int ld, int r0, int R, int k0, int K,
                                      int tid)
{
    constexpr int CHUNKS = GM * GK / 4;   // chunks of 4 elements
    constexpr int PER_ROW = GK / 4;
    constexpr int PER_THREAD = CHUNKS / 256;
    // interior slab, 8-byte aligned rows: every load is issued before the
    // first LDS store waits on one (the per-chunk bounds checks below would
    // serialise the loads, one memory round trip each)
    if (!KCONTIG && r0 + GM <= R && k0 + GK <= K && (ld & 3) == 0 && (reinterpret_cast<uintptr_t>(X) & 7u) == 0) {
        // rows contiguous in memory: each thread loads a 4 (k) x 4 (row)
        // block as four 8-byte reads along rows, transposes it in registers
        // and stores four 8-byte k-runs, one per row
        constexpr int BLOCKS = (GK / 4) * (GM / 4) / 256;
#pragma unroll
        for (int i = 0; i < BLOCKS; i++) {
            const int blk = tid + i * 256;
            const int kb = blk / (GM / 4), rb = blk % (GM / 4);
            uint2 q[4];
#pragma unroll
            for (int j = 0; j < 4; j++)
                q[j] = *reinterpret_cast<const uint2 *>(X + (int64_t)(k0 + 4 * kb + j) * ld + r0 + 4 * rb);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                // element r of each k-row j: low / high half of word r / 2
                uint32_t e[4];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const uint32_t wv = (r < 2) ? q[j].x : q[j].y;
                    e[j] = (r & 1) ? (wv >> 16) : (wv & 0xFFFFu);
                }
                uint2 o;
                o.x = e[0] | (e[1] << 16);
                o.y = e[2] | (e[3] << 16);
                *reinterpret_cast<uint2 *>(&S[4 * rb + r][4 * kb]) = o;
            }
        }
        return;
    }
    if (r0 + GM <= R && k0 + GK <= K && (ld & 3) == 0 && (reinterpret_cast<uintptr_t>(X) & 7u) == 0) {
        uint2 q[PER_THREAD];
#pragma unroll
        for (int i = 0; i < PER_THREAD; i++) {
            const int c = tid + i * 256;
            const int64_t off = KCONTIG ? (int64_t)(r0 + c / PER_ROW) * ld + k0 + (c % PER_ROW) * 4
                                        : (int64_t)(k0 + (c >> 5)) * ld + r0 + (c & 31) * 4;
            q[i] = *reinterpret_cast<const uint2 *>(X + off);
        }
#pragma unroll
        for (int i = 0; i < PER_THREAD; i++) {
            const int c = tid + i * 256;
            if (KCONTIG) {
                *reinterpret_cast<uint2 *>(&S[c / PER_ROW][(c % PER_ROW) * 4]) = q[i];
            } else {
                const int k = c >> 5, rc = (c & 31) * 4;
                S[rc][k] = (uint16_t)q[i].x;
                S[rc + 1][k] = (uint16_t)(q[i].x >> 16);
                S[rc + 2][k] = (uint16_t)q[i].y;
                S[rc + 3][k] = (uint16_t)(q[i].y >> 16);
            }
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < PER_THREAD; i++) {
        const int c = tid + i * 256;
        uint16_t v[4];
        if (KCONTIG) {
            const int r = c / PER_ROW, kc = (c % PER_ROW) * 4;
            const int gr = r0 + r, gk = k0 + kc;
            if (gr < R && gk < K) load4(X + (int64_t)gr * ld + gk, K - gk, v);
            else v[0] = v[1] = v[2] = v[3] = 0;
            // the four values are adjacent in the LDS row: one 8-byte store
            // (rows are 80 bytes, kc a multiple of 4: 8-byte aligned)
            uint2 q;
            q.x = (uint32_t)v[0] | ((uint32_t)v[1] << 16);
            q.y = (uint32_t)v[2] | ((uint32_t)v[3] << 16);
            *reinterpret_cast<uint2 *>(&S[r][kc]) = q;
        } else {
            const int k = c >> 5, rc = (c & 31) * 4;
            const int gr = r0 + rc, gk = k0 + k;
            if (gr < R && gk < K) load4(X + (int64_t)gk * ld + gr, R - gr, v);
            else v[0] = v[1] = v[2] = v[3] = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) S[rc + j][k] = v[j];
        }
    }
}

// The fast path of stage() split in two, so that a slab's loads can be
// issued a K step ahead (into registers) and stored after the MFMAs of the
// current step.  slab_fast: 8-byte aligned rows and row / k bounds that are
// multiples of 4, so that every 4-element piece is wholly inside or wholly
// outside the matrix; load_slab issues this thread's four 8-byte loads
// (zeros for pieces outside), store_slab writes them to LDS (row-contiguous
// slabs transposed in registers first).
template <bool KCONTIG>
__device__ __forceinline__ bool slab_fast(const uint16_t *X, int ld, int R, int k0, int K)
{
    static_assert(GM * GK / 4 / 256 == 4 && (GK / 4) * (GM / 4) / 256 == 1, "one 4-load unit per thread");
    return k0 < K && (ld & 3) == 0 && (R & 3) == 0 && (K & 3) == 0 && (reinterpret_cast<uintptr_t>(X) & 7u) == 0;
}

template <bool KCONTIG>
__device__ __forceinline__ void load_slab(uint2 q[4], const uint16_t *X, int ld, int r0, int R, int k0, int K,
                                          int tid)
{
#pragma unroll
    for (int i = 0; i < 4; i++) {
        int r, k;
        if (KCONTIG) {
            const int c = tid + i * 256;
            r = r0 + c / (GK / 4);
            k = k0 + (c % (GK / 4)) * 4;
        } else {
            r = r0 + 4 * (tid % (GM / 4));
            k = k0 + 4 * (tid / (GM / 4)) + i;
        }
        const int64_t off = KCONTIG ? (int64_t)r * ld + k : (int64_t)k * ld + r;
        q[i] = (r < R && k < K) ? *reinterpret_cast<const uint2 *>(X + off) : make_uint2(0u, 0u);
    }
}

template <bool KCONTIG>
__device__ __forceinline__ void store_slab(uint16_t (*S)[LDS_K], const uint2 q[4], int tid)
{
    if (KCONTIG) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int c = tid + i * 256;
            *reinterpret_cast<uint2 *>(&S[c / (GK / 4)][(c % (GK / 4)) * 4]) = q[i];
        }
    } else {
        const int kb = tid / (GM / 4), rb = tid % (GM / 4);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            uint32_t e[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t wv = (r < 2) ? q[j].x : q[j].y;
                e[j] = (r & 1) ? (wv >> 16) : (wv & 0xFFFFu);
            }
            uint2 o;
            o.x = e[0] | (e[1] << 16);
            o.y = e[2] | (e[3] << 16);
            *reinterpret_cast<uint2 *>(&S[4 * rb + r][4 * kb]) = o;
        }
    }
}

// C tile of 128 x 128 per workgroup, 64 x 64 per wave (2 x 2 MFMA blocks).
// One-dimensional grid: workgroup b runs on XCD b % 8 (the hardware deals
// workgroups out round-robin), so b is mapped to (split z, row tile m,
// column tile n) with every column tile of one row tile -- the workgroups
// that read the same 128-row slab of A -- on one XCD at consecutive slots,
// and the slab comes from HBM once into that XCD's L2 instead of once per
// column tile.  Split K: z selects a K range; each range's partial tile
// goes to a workspace and a second kernel adds them.
//
// Register budget: 4 waves per SIMD (128 VGPRs, the accumulators in VGPRs,
// no spills), against 3 waves with the accumulators in AGPRs (152
// registers): the layer-1 GEMM (M = 131,072, N = 512, K = 180) took 65
// against 73 us.
//
// The MFMA takes B's slab as its first operand and A's as its second, so it
// produces C^T blocks: lane l holds, for each group g of four accumulators,
// C[m0 + l % 32][n0 + 8 g + 4 (l / 32) + 0..3] -- four adjacent columns of
// one row, stored as one 8-byte (BF16) or 16-byte (FP32) vector.
template <bool AT, bool BT>
__global__ __launch_bounds__(256, 4) void gemm_kernel(int M, int N, int K, int k_per_split, int tiles_m, int tiles_n,
                                                   const uint16_t *__restrict__ A, int lda,
                                                   const uint16_t *__restrict__ B, int ldb,
                                                   const float *__restrict__ bias, int relu, float *C32, int ldc,
                                                   int accumulate, uint16_t *C16, int ldc16, int64_t split_stride)
{
    // the K loop's A and B slabs, then (BF16-only outputs) the C tile
    constexpr int CS_LD = GN + 8;   // C tile row stride in halves: 272 bytes, 16-byte aligned
    constexpr int SMEM = (2 * GM * LDS_K > GM * CS_LD) ? 2 * GM * LDS_K : GM * CS_LD;
    __shared__ __attribute__((aligned(16))) uint16_t smem[SMEM];
    uint16_t (*As)[LDS_K] = reinterpret_cast<uint16_t (*)[LDS_K]>(smem);
    uint16_t (*Bs)[LDS_K] = reinterpret_cast<uint16_t (*)[LDS_K]>(smem + GM * LDS_K);
    __shared__ float Bias[GN];   // the tile's bias, read once (0 past N or without bias)
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w >> 1, wn = w & 1;
    // XCD-aware tile of this workgroup: the row slabs (split z, row tile
    // tm), numbered z * tiles_m + tm, are dealt to the XCDs round-robin, and
    // all column tiles of one row slab sit on its XCD at consecutive slots
    const int xcd = (int)blockIdx.x & 7, slot = (int)blockIdx.x >> 3;
    const int mz = (slot / tiles_n) * 8 + xcd, tn = slot % tiles_n;
    const int n_slabs = tiles_m * (K > k_per_split ? (K + k_per_split - 1) / k_per_split : 1);
    if (mz >= n_slabs) return;
    const int bz = mz / tiles_m, tm = mz - bz * tiles_m;
    const int m0 = tm * GM, n0 = tn * GN;
    if (tid < GN) Bias[tid] = (bias && n0 + tid < N) ? bias[n0 + tid] : 0.f;
    __syncthreads();
    const int kz0 = bz * k_per_split;
    const int kz1 = min(K, kz0 + k_per_split);
    const bool split = k_per_split < K;
    f32x16 acc[2][2] = {};
    // A: M x K, k contiguous unless transposed; B: K x N, n contiguous
    // unless transposed.  qa / qb hold the next slab, loaded one K step
    // ahead when both operands take the fast path; other slabs go through
    // stage() when their step comes.
    uint2 qa[4], qb[4];
    auto fast = [&](int k) { return slab_fast<!AT>(A, lda, M, k, kz1) && slab_fast<BT>(B, ldb, N, k, kz1); };
    bool ahead = fast(kz0);
    if (ahead) {
        load_slab<!AT>(qa, A, lda, m0, M, kz0, kz1, tid);
        load_slab<BT>(qb, B, ldb, n0, N, kz0, kz1, tid);
    }
    for (int k0 = kz0; k0 < kz1; k0 += GK) {
        if (ahead) {
            store_slab<!AT>(As, qa, tid);
            store_slab<BT>(Bs, qb, tid);
        } else {
            stage<!AT>(As, A, lda, m0, M, k0, kz1, tid);
            stage<BT>(Bs, B, ldb, n0, N, k0, kz1, tid);
        }
        __syncthreads();
        ahead = fast(k0 + GK);
        if (ahead) {   // in flight during this step's MFMAs
            load_slab<!AT>(qa, A, lda, m0, M, k0 + GK, kz1, tid);
            load_slab<BT>(qb, B, ldb, n0, N, k0 + GK, kz1, tid);
        }
#pragma unroll
        for (int kk = 0; kk < GK; kk += 16) {
            const int ko = kk + 8 * (lane >> 5);
            bf16x8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                __builtin_memcpy(&a[i], &As[wm * 64 + i * 32 + (lane & 31)][ko], 16);
                __builtin_memcpy(&b[i], &Bs[wn * 64 + i * 32 + (lane & 31)][ko], 16);
            }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[j], a[i], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    if (!split && !C32) {
        // BF16 output only (the forward activations): the tile goes through
        // LDS so that global stores are whole 16-byte runs of one row
        uint16_t (*Cs)[CS_LD] = reinterpret_cast<uint16_t (*)[CS_LD]>(smem);   // the K loop ended on a barrier
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int g = 0; g < 4; g++) {
                    const int rl = wm * 64 + i * 32 + (lane & 31);
                    const int cl = wn * 64 + j * 32 + 8 * g + 4 * (lane >> 5);
                    float v[4];
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        v[e] = acc[i][j][4 * g + e] + Bias[cl + e];
                        if (relu) v[e] = v[e] > 0.f ? v[e] : 0.f;
                    }
                    uint2 o;
                    o.x = (uint32_t)to_bf16(v[0]) | ((uint32_t)to_bf16(v[1]) << 16);
                    o.y = (uint32_t)to_bf16(v[2]) | ((uint32_t)to_bf16(v[3]) << 16);
                    *reinterpret_cast<uint2 *>(&Cs[rl][cl]) = o;
                }
        __syncthreads();
        const bool vec = (ldc16 & 7) == 0 && (reinterpret_cast<uintptr_t>(C16) & 15u) == 0 && n0 + GN <= N;
#pragma unroll
        for (int i = 0; i < GM * GN / 8 / 256; i++) {
            const int c = tid + i * 256;
            const int rl = c / (GN / 8), cl = (c % (GN / 8)) * 8;
            const int row = m0 + rl;
            if (row >= M) continue;
            uint16_t *p = C16 + (int64_t)row * ldc16 + n0 + cl;
            if (vec) {
                *reinterpret_cast<uint4 *>(p) = *reinterpret_cast<const uint4 *>(&Cs[rl][cl]);
            } else {
                for (int e = 0; e < 8; e++)
                    if (n0 + cl + e < N) p[e] = Cs[rl][cl + e];
            }
        }
        return;
    }
    // split K: this range's partial tile goes to its own slice of the
    // workspace (plain stores), summed by a second kernel
    if (split) {
        C32 += bz * split_stride;
        accumulate = 0;
    }
    const bool vec32 = C32 && (ldc & 3) == 0 && (reinterpret_cast<uintptr_t>(C32) & 15u) == 0;
    const bool vec16 = C16 && (ldc16 & 3) == 0 && (reinterpret_cast<uintptr_t>(C16) & 7u) == 0;
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int row = m0 + wm * 64 + i * 32 + (lane & 31);
        if (row >= M) continue;
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const int col = n0 + wn * 64 + j * 32 + 8 * g + 4 * (lane >> 5);
                if (col >= N) continue;
                float v[4];
#pragma unroll
                for (int e = 0; e < 4; e++) v[e] = acc[i][j][4 * g + e];
                const bool full = col + 3 < N;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    v[e] += Bias[col - n0 + e];
                    if (relu) v[e] = v[e] > 0.f ? v[e] : 0.f;
                }
                if (C32) {
                    float *p = C32 + (int64_t)row * ldc + col;
                    if (full && vec32) {
                        float4 o = make_float4(v[0], v[1], v[2], v[3]);
                        if (accumulate) {
                            const float4 q = *reinterpret_cast<const float4 *>(p);
                            o.x += q.x;
                            o.y += q.y;
                            o.z += q.z;
                            o.w += q.w;
                        }
                        *reinterpret_cast<float4 *>(p) = o;
                    } else {
                        for (int e = 0; e < 4; e++)
                            if (col + e < N) p[e] = accumulate ? p[e] + v[e] : v[e];
                    }
                }
                if (C16) {
                    uint16_t *p = C16 + (int64_t)row * ldc16 + col;
                    if (full && vec16) {
                        uint2 o;
                        o.x = (uint32_t)to_bf16(v[0]) | ((uint32_t)to_bf16(v[1]) << 16);
                        o.y = (uint32_t)to_bf16(v[2]) | ((uint32_t)to_bf16(v[3]) << 16);
                        *reinterpret_cast<uint2 *>(p) = o;
                    } else {
                        for (int e = 0; e < 4; e++)
                            if (col + e < N) p[e] = to_bf16(v[e]);
                    }
                }
            }
    }
}

// C[m][n] (+)= sum over s of W[s][m][n] (split-K partial tiles, ld N)
__global__ void reduce_splits_kernel(int M, int N, int splits, const float *__restrict__ W, float *C, int ldc,
                                     int accumulate)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t MN = (int64_t)M * N;
    if (i >= MN) return;
    // eight slices per round: their loads are all in flight together
    float v = 0.f;
    int z = 0;
    for (; z + 8 <= splits; z += 8) {
        float t[8];
#pragma unroll
        for (int u = 0; u < 8; u++) t[u] = W[(z + u) * MN + i];
#pragma unroll
        for (int u = 0; u < 8; u++) v += t[u];
    }
    for (; z < splits; z++) v += W[z * MN + i];
    float *p = C + (i / N) * ldc + (i % N);
    *p = accumulate ? *p + v : v;
}

__global__ void relu_backward_kernel(int64_t n, const uint16_t *act, float *grad, uint16_t *out)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const bool on = act[i] != 0 && !(act[i] & 0x8000u);
        const float g = on ? grad[i] : 0.f;
        grad[i] = g;
        out[i] = to_bf16(g);
    }
}

// one thread per (state, test, feature word): 180 features from 16 words
__global__ void features_kernel(int B, const uint16_t *rec, uint16_t *X)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;   // (b*32 + t) * 180 + f
    if (i >= (int64_t)B * 32 * 180) return;
    const int64_t row = i / 180;
    const int f = (int)(i % 180);
    const uint16_t *r = rec + row * 16;
    // f < 176: bit f % 16 of word f / 16 (input0, input1, target, r0..r7);
    // 176..179: flag bits 0..3
    const uint32_t word = f < 176 ? r[f >> 4] : r[11];
    const uint32_t bit = f < 176 ? (uint32_t)(f & 15) : (uint32_t)(f - 176);
    X[i] = ((word >> bit) & 1u) ? (uint16_t)0x3F80u : (uint16_t)0;   // bf16 1.0 / 0.0
}

// one workgroup per state, one thread per channel (C + E <= 1024)
__global__ void pool_forward_kernel(int T, int C, int E, const uint16_t *H, const uint8_t *nvalid,
                                    const int32_t *depth, const float *emb, uint16_t *Z)
{
    const int b = blockIdx.x, c = threadIdx.x;
    const int nv = nvalid[b];
    uint16_t *z = Z + (int64_t)b * (C + E);
    if (c < C) {
        float s = 0.f;
        for (int t = 0; t < nv; t++) s += from_bf16(H[((int64_t)b * T + t) * C + c]);
        z[c] = to_bf16(nv ? s / (float)nv : 0.f);
    } else if (c < C + E) {
        z[c] = to_bf16(emb[depth[b] * E + (c - C)]);
    }
}

__global__ void pool_backward_kernel(int T, int C, int E, const float *dZ, const uint8_t *nvalid,
                                     const int32_t *depth, float *dH, float *demb)
{
    const int b = blockIdx.x, c = threadIdx.x;
    const int nv = nvalid[b];
    const float *dz = dZ + (int64_t)b * (C + E);
    if (c < C) {
        const float g = nv ? dz[c] / (float)nv : 0.f;
        for (int t = 0; t < T; t++) dH[((int64_t)b * T + t) * C + c] = t < nv ? g : 0.f;
    } else if (c < C + E) {
        atomicAdd(&demb[depth[b] * E + (c - C)], dz[c]);
    }
}

// Column sums (bias gradients), two passes without atomics.  Pass 1: a
// block sums CS_ROWS rows of up to 256 columns, each lane four adjacent
// columns (one 16-byte load per row when aligned), its four waves on
// interleaved rows; one partial row per block goes to the workspace.
constexpr int CS_ROWS = 64;

__global__ __launch_bounds__(256) void colsum_partial_kernel(int M, int N, const float *__restrict__ G, int ldg,
                                                             int vec, float *__restrict__ P)
{
    __shared__ float red[4][4][64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int c = ((int)blockIdx.x * 64 + lane) * 4;
    const int r0 = (int)blockIdx.y * CS_ROWS, r1 = min(M, r0 + CS_ROWS);
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    if (c < N) {
        const bool full = vec && c + 3 < N;
#pragma unroll 4
        for (int r = r0 + w; r < r1; r += 4) {
            const float *row = G + (int64_t)r * ldg + c;
            if (full) {
                const float4 v = *reinterpret_cast<const float4 *>(row);
                s0 += v.x;
                s1 += v.y;
                s2 += v.z;
                s3 += v.w;
            } else {
                s0 += row[0];
                if (c + 1 < N) s1 += row[1];
                if (c + 2 < N) s2 += row[2];
                if (c + 3 < N) s3 += row[3];
            }
        }
    }
    red[w][0][lane] = s0;
    red[w][1][lane] = s1;
    red[w][2][lane] = s2;
    red[w][3][lane] = s3;
    __syncthreads();
    if (w == 0 && c < N) {
        float *p = P + (int64_t)blockIdx.y * N + c;
        for (int j = 0; j < 4 && c + j < N; j++) p[j] = red[0][j][lane] + red[1][j][lane] + red[2][j][lane] + red[3][j][lane];
    }
}

// Pass 2: out[c] = sum of the partial rows (four waves on interleaved rows).
__global__ __launch_bounds__(256) void colsum_reduce_kernel(int N, int parts, const float *__restrict__ P, float *out)
{
    __shared__ float red[4][64];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int c = (int)blockIdx.x * 64 + lane;
    float s = 0.f;
    if (c < N) {
#pragma unroll 8
        for (int q = w; q < parts; q += 4) s += P[(int64_t)q * N + c];
    }
    red[w][lane] = s;
    __syncthreads();
    if (w == 0 && c < N) out[c] = red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane];
}

__global__ void fill_f32_kernel(int n, float *p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// one workgroup: loss reduction in LDS
__global__ void bce_kernel(int B, const float *y, const float *label, float *dy, float *loss)
{
    __shared__ float part[256];
    float s = 0.f;
    for (int b = threadIdx.x; b < B; b += blockDim.x) {
        const float x = y[b], t = label[b];
        // softplus(x) - t*x, written to stay finite for large |x|
        s += fmaxf(x, 0.f) - t * x + log1pf(expf(-fabsf(x)));
        dy[b] = (1.f / (1.f + expf(-x)) - t) / (float)B;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[0] = part[0] / (float)B;
}

// The two-head loss of python/train.py:44-60: column 0 of each row of y
// (B x nh) is the reach logit -- binary cross-entropy against label, mean
// over the batch; columns 1..nh-1 are the remaining-steps logits --
// cross-entropy against remaining (clamped to 0..nh-2) over the rows whose
// label is positive, mean over those n_pos rows, times w.  dy gets the
// gradient of the total, loss[0] the total.  One block, fixed reduction
// order (deterministic).
__global__ void reach_dist_loss_kernel(int B, int nh, const float *y, const float *label, const int32_t *remaining,
                                       int n_pos, float w, float *dy, float *loss)
{
    __shared__ float part[256];
    float s = 0.f;
    for (int b = threadIdx.x; b < B; b += blockDim.x) {
        const float *yb = y + (int64_t)b * nh;
        float *db = dy + (int64_t)b * nh;
        const float x = yb[0], t = label[b];
        s += (fmaxf(x, 0.f) - t * x + log1pf(expf(-fabsf(x)))) / (float)B;
        db[0] = (1.f / (1.f + expf(-x)) - t) / (float)B;
        const bool pos = t > 0.5f && n_pos > 0 && w != 0.f;
        if (pos) {
            const int r = min(max(remaining[b], 0), nh - 2);
            float m = yb[1];
            for (int j = 2; j < nh; j++) m = fmaxf(m, yb[j]);
            float se = 0.f;
            for (int j = 1; j < nh; j++) se += expf(yb[j] - m);
            const float lse = m + logf(se);
            for (int j = 1; j < nh; j++)
                db[j] = w * (expf(yb[j] - lse) - (j - 1 == r ? 1.f : 0.f)) / (float)n_pos;
            s += w * (lse - yb[1 + r]) / (float)n_pos;
        } else {
            for (int j = 1; j < nh; j++) db[j] = 0.f;
        }
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[0] = part[0];
}

__global__ void adam_kernel(int64_t n, float *p, const float *g, float *m, float *v, uint16_t *p16, float lr,
                            float b1, float b2, float eps, float c1, float c2)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        p[i] = pi;
        if (p16) p16[i] = to_bf16(pi);
    }
}

__global__ void to_bf16_kernel(int64_t n, const float *x, uint16_t *y)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = to_bf16(x[i]);
}

inline hipStream_t st(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Split-K workspace of the current device, grown on demand and kept until
// the process ends.  Growing waits for the device first: launches still
// queued on any stream may read the old buffer.
float *split_workspace(size_t floats)
{
    static std::mutex mu;
    static std::map<int, std::pair<float *, size_t>> ws;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    auto &e = ws[dev];
    if (e.second >= floats) return e.first;
    if (e.first) {
        if (hipDeviceSynchronize() != hipSuccess) return nullptr;
        (void)hipFree(e.first);
        e = {nullptr, 0};
    }
    float *p = nullptr;
    if (hipMalloc((void **)&p, floats * sizeof(float)) != hipSuccess) return nullptr;
    e = {p, floats};
    return p;
}
inline int grid1d(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 65535); }
// error state of the launches just made; an entry point first clears any
// error a previous, unrelated runtime call of this thread left behind
inline int status() { return hipGetLastError() == hipSuccess ? PCAGPU_OK : PCAGPU_EHIP; }
inline void clear_error() { (void)hipGetLastError(); }

}  // namespace

extern "C" {

int pcav_gemm(int M, int N, int K, const uint16_t *A, int lda, int a_trans, const uint16_t *B, int ldb,
              int b_trans, const float *bias, int relu, float *C32, int ldc, int accumulate, uint16_t *C16,
              int ldc16, void *stream)
{
    if (M < 0 || N < 0 || K < 0 || !A || !B || (!C32 && !C16)) return PCAGPU_EINVAL;
    if (M == 0 || N == 0) return PCAGPU_OK;
    const int tiles_m = (M + GM - 1) / GM, tiles_n = (N + GN - 1) / GN;
    // split K when few output tiles face a long K (weight gradients): about
    // 512 workgroups, each K range at least 256 long; the partial tiles go
    // to a workspace and are summed into C32, so only a plain FP32 output
    // can take it
    int splits = 1;
    const int tiles = tiles_m * tiles_n;
    if (C32 && !C16 && !bias && !relu && tiles < 256 && K >= 1024)
        splits = std::min(std::min((512 + tiles - 1) / tiles, K / 256), 64);
    const int kps = std::max(((K + splits - 1) / splits + GK - 1) / GK * GK, GK);   // K = 0: one empty range
    splits = std::max(1, (K + kps - 1) / kps);
    float *ws = nullptr;
    if (splits > 1 && !(ws = split_workspace((size_t)splits * M * N))) return PCAGPU_ENOMEM;
    clear_error();
    // one-dimensional grid, XCD-aware tile order inside the kernel
    const int64_t blocks = ((int64_t)splits * tiles_m + 7) / 8 * 8 * tiles_n;
    if (blocks > 0x7FFFFFFF) return PCAGPU_EINVAL;
    const dim3 grid((unsigned)blocks);
    float *Cout = ws ? ws : C32;
    const int ldo = ws ? N : ldc;
    const int64_t stride = ws ? (int64_t)M * N : 0;
#define PCAV_GEMM(AT, BT)                                                                                            \
    hipLaunchKernelGGL((gemm_kernel<AT, BT>), grid, dim3(256), 0, st(stream), M, N, K, kps, tiles_m, tiles_n, A, lda, \
                       B, ldb, bias, relu, Cout, ldo, accumulate, C16, ldc16, stride)
    if (a_trans && b_trans) PCAV_GEMM(true, true);
    else if (a_trans) PCAV_GEMM(true, false);
    else if (b_trans) PCAV_GEMM(false, true);
    else PCAV_GEMM(false, false);
#undef PCAV_GEMM
    if (ws) {
        const int64_t n = (int64_t)M * N;
        hipLaunchKernelGGL(reduce_splits_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st(stream), M, N,
                           splits, ws, C32, ldc, accumulate);
    }
    return status();
}

int pcav_features(int B, const uint16_t *rec, uint16_t *X, void *stream)
{
    if (B < 0 || !rec || !X) return PCAGPU_EINVAL;
    if (B == 0) return PCAGPU_OK;
    const int64_t n = (int64_t)B * 32 * 180;
    clear_error();
    hipLaunchKernelGGL(features_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st(stream), B, rec, X);
    return status();
}

int pcav_relu_backward(int64_t n, const uint16_t *act16, float *grad, uint16_t *out16, void *stream)
{
    if (n < 0 || !act16 || !grad || !out16) return PCAGPU_EINVAL;
    if (n == 0) return PCAGPU_OK;
    clear_error();
    hipLaunchKernelGGL(relu_backward_kernel, dim3(grid1d(n)), dim3(256), 0, st(stream), n, act16, grad, out16);
    return status();
}

int pcav_pool_forward(int B, int T, int C, int E, const uint16_t *H, const uint8_t *nvalid, const int32_t *depth,
                      const float *emb, uint16_t *Z, void *stream)
{
    if (B < 0 || T <= 0 || C <= 0 || E < 0 || C + E > 1024 || !H || !nvalid || !Z || (E && (!depth || !emb)))
        return PCAGPU_EINVAL;
    if (B == 0) return PCAGPU_OK;
    const int threads = ((C + E + 63) / 64) * 64;
    clear_error();
    hipLaunchKernelGGL(pool_forward_kernel, dim3(B), dim3(threads), 0, st(stream), T, C, E, H, nvalid, depth, emb,
                       Z);
    return status();
}

int pcav_pool_backward(int B, int T, int C, int E, const float *dZ, const uint8_t *nvalid, const int32_t *depth,
                       float *dH, float *demb, void *stream)
{
    if (B < 0 || T <= 0 || C <= 0 || E < 0 || C + E > 1024 || !dZ || !nvalid || !dH || (E && (!depth || !demb)))
        return PCAGPU_EINVAL;
    if (B == 0) return PCAGPU_OK;
    const int threads = ((C + E + 63) / 64) * 64;
    clear_error();
    hipLaunchKernelGGL(pool_backward_kernel, dim3(B), dim3(threads), 0, st(stream), T, C, E, dZ, nvalid, depth, dH,
                       demb);
    return status();
}

int pcav_colsum(int M, int N, const float *G, int ldg, float *out, void *stream)
{
    if (M < 0 || N <= 0 || !G || !out || ldg < N) return PCAGPU_EINVAL;
    clear_error();
    if (M == 0) {
        hipLaunchKernelGGL(fill_f32_kernel, dim3((N + 255) / 256), dim3(256), 0, st(stream), N, out);
        return status();
    }
    const int parts = (M + CS_ROWS - 1) / CS_ROWS;
    if (parts > 65535) return PCAGPU_EINVAL;
    float *ws = split_workspace((size_t)parts * N);
    if (!ws) return PCAGPU_ENOMEM;
    const int vec = ((uintptr_t)G % 16 == 0 && ldg % 4 == 0) ? 1 : 0;
    clear_error();
    hipLaunchKernelGGL(colsum_partial_kernel, dim3((unsigned)((N + 255) / 256), (unsigned)parts), dim3(256), 0,
                       st(stream), M, N, G, ldg, vec, ws);
    hipLaunchKernelGGL(colsum_reduce_kernel, dim3((unsigned)((N + 63) / 64)), dim3(256), 0, st(stream), N, parts, ws,
                       out);
    return status();
}

int pcav_bce(int B, const float *logit, const float *label, float *dlogit, float *loss, void *stream)
{
    if (B <= 0 || !logit || !label || !dlogit || !loss) return PCAGPU_EINVAL;
    clear_error();
    hipLaunchKernelGGL(bce_kernel, dim3(1), dim3(256), 0, st(stream), B, logit, label, dlogit, loss);
    return status();
}

int pcav_reach_dist_loss(int B, int nh, const float *y, const float *label, const int32_t *remaining, int n_pos,
                         float w, float *dy, float *loss, void *stream)
{
    if (B <= 0 || nh < 2 || !y || !label || !dy || !loss || (n_pos > 0 && !remaining) || n_pos < 0 || n_pos > B)
        return PCAGPU_EINVAL;
    clear_error();
    hipLaunchKernelGGL(reach_dist_loss_kernel, dim3(1), dim3(256), 0, st(stream), B, nh, y, label, remaining, n_pos,
                       w, dy, loss);
    return status();
}

int pcav_adam(int64_t n, float *p, const float *g, float *m, float *v, uint16_t *p16, float lr, float beta1,
              float beta2, float eps, int step, void *stream)
{
    if (n < 0 || !p || !g || !m || !v || step < 1) return PCAGPU_EINVAL;
    if (n == 0) return PCAGPU_OK;
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    clear_error();
    hipLaunchKernelGGL(adam_kernel, dim3(grid1d(n)), dim3(256), 0, st(stream), n, p, g, m, v, p16, lr, beta1, beta2,
                       eps, c1, c2);
    return status();
}

void *pcav_malloc(int device, size_t bytes)
{
    void *p = nullptr;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) return nullptr;
    return p;
}

void pcav_free(void *p)
{
    if (p) (void)hipFree(p);
}

int pcav_copy_to_device(void *dst, const void *src, size_t bytes)
{
    if (!bytes) return PCAGPU_OK;
    if (!dst || !src) return PCAGPU_EINVAL;
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? PCAGPU_OK : PCAGPU_EHIP;
}

int pcav_copy_to_host(void *dst, const void *src, size_t bytes)
{
    if (!bytes) return PCAGPU_OK;
    if (!dst || !src) return PCAGPU_EINVAL;
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? PCAGPU_OK : PCAGPU_EHIP;
}

int pcav_clear(void *dst, size_t bytes)
{
    if (!bytes) return PCAGPU_OK;
    if (!dst) return PCAGPU_EINVAL;
    return hipMemset(dst, 0, bytes) == hipSuccess ? PCAGPU_OK : PCAGPU_EHIP;
}

int pcav_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? PCAGPU_OK : PCAGPU_EHIP; }

int pcav_to_bf16(int64_t n, const float *x, uint16_t *y16, void *stream)
{
    if (n < 0 || !x || !y16) return PCAGPU_EINVAL;
    if (n == 0) return PCAGPU_OK;
    clear_error();
    hipLaunchKernelGGL(to_bf16_kernel, dim3(grid1d(n)), dim3(256), 0, st(stream), n, x, y16);
    return status();
}

}  // extern "C"
