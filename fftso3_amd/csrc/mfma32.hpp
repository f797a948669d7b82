// mfma32.hpp -- batch-wide beta contraction on FP32 MFMA (v_mfma_f32_32x32x2f32), N = 2B <= 128.
//
// One workgroup per (m,n) group g for the whole batch.  The FP64 recursion writes the group's
// Wigner-d (or Legendre) block D[l][k] (M = B - m0 <= 64 rows, N <= 128 columns) to LDS once; the
// batch then passes through in blocks of 128 spectrum columns (64/NP functions x NP pairs x re/im,
// 512 contiguous bytes of every workspace row).  Wave w multiplies the 32 columns of pair
// p = w % NP of 16 of the block's functions, so each wave's coefficient rows are one pair's: 16
// functions x 8 bytes = 128 contiguous bytes of a row of PT[e][function], the transpose of the
// reference's packed coefficient array (e = packed index of (l,m,n)).
//
// Forward: the next block's spectrum rows load (16-byte loads, whole 512-byte row segments) into
// registers while the current block multiplies out of LDS; they are then written to LDS with the
// columns regrouped by wave.  Inverse: the next block's coefficient rows load while the current
// block multiplies; results go through LDS so the workspace rows are written as whole 512-byte
// segments.  Stores never branch (lanes with nothing to store write to a dummy slot), so the
// compiler's wait counts stay exact and no wait ever covers more than one block's stores.
//
// Operand layouts of v_mfma_f32_32x32x2f32 (lane i): A[i%32][i/32], B[i/32][i%32], result
// D[8(j/4) + 4(i/32) + j%4][i%32] for j < 16.
#pragma once

#include <type_traits>

#include "kernels.hpp"

typedef float f32x16m_t __attribute__((ext_vector_type(16)));

#define MX_DPITCH 129        // LDS pitch of D rows (floats): conflict-free column reads
#define MX_LPITCH 132        // LDS pitch of spectrum-block rows (floats)
#define MX_LDS_BYTES ((64 * MX_DPITCH + 128 * MX_LPITCH) * 4)
// Block and d rows of the grouped kernels (contract_fwd_mx16g, contract_inv_mx32g) are 128
// floats long, with columns XOR-swizzled by row bits so gfx950's 64 LDS banks see no conflict (a
// 128-float pitch also keeps every operand offset within the 16-bit LDS immediate):
//   forward block: column c of row r at c ^ 16 (r & 3) -- the 4 rows of a 16 x 16 x 4 operand
//     read, and the 2 rows of one 8-byte staging access, take disjoint 16-bank ranges;
//   inverse d: column k of row r at k ^ 32 (r & 1) -- the 2 rows of a 32 x 32 x 2 operand read;
//   inverse output block: column c of row k at c ^ 32 ((k >> 2) & 1) ^ 16 (k & 1) -- the
//     accumulator writes (lane halves 4 rows apart) and the 8-byte reads of two adjacent rows.
#define MXG_LPITCH 128
#define MXG_DPITCH 128
// the forward's block region first holds the two groups' D staging ([2][64][MX_DPITCH])
#define MX16G_BLOCK_BYTES (2 * 64 * MX_DPITCH * 4 > 128 * MXG_LPITCH * 4 ? 2 * 64 * MX_DPITCH * 4 : 128 * MXG_LPITCH * 4)
__host__ __device__ inline size_t mx16g_lds_bytes(int B) { return (size_t)MX16G_BLOCK_BYTES + 2 * (size_t)B * sizeof(RecCoef); }
__host__ __device__ inline size_t mx32g_inv_lds_bytes(int B)
{
    return (size_t)(2 * 64 * MXG_DPITCH + 128 * MXG_LPITCH) * 4 + 2 * (size_t)B * sizeof(RecCoef);
}

struct MxArgs {
    ChunkArgs ca;            // ca.C = functions of the class (workspace stride)
    RecTables rt;
    const int *items;        // groups (largest first)
    int nitems;
    int nsplit;              // workgroups per group, each taking a contiguous range of column blocks
    int gpw;                 // groups per workgroup (grouped kernels: items are sets of gpw groups)
    int g0;                  // grouped kernels: first group of item 0
    const void *src;         // forward: spectrum workspace; inverse: PT
    void *dst;               // forward: PT; inverse: spectrum workspace
    float *dummy;            // 16-byte slot that absorbs the stores of lanes with nothing to store
    int mrows;               // K-chunked forward: degree rows the launch's LDS holds (0: the bandlimit)
};

// Work unit of a batch-wide kernel: group a.items[u / nsplit] and its column blocks [fb0, fb1)
// (nfb blocks of 128 spectrum columns in all).  False for units past the grid or with no blocks;
// either is uniform over the workgroup.
// The unit a workgroup of a batch-wide kernel launch takes: groups come largest first, dealt
// round-robin over the XCDs (xcd_deal) in runs of the nsplit units of a group times the 2^gil
// groups whose workspace records share lines.
__device__ __forceinline__ int mx_block_unit(const MxArgs &a)
{
    return xcd_deal(a.gpw > 1 ? a.nsplit : a.nsplit << a.ca.gil);
}

__device__ __forceinline__ bool mx_unit(const MxArgs &a, int nfb, int unit, int &g, int &fb0, int &fb1)
{
    const int n = a.nitems * a.nsplit;
    if (unit < 0 || unit >= n) return false;
    const int it = unit / a.nsplit, part = unit - it * a.nsplit;
    g = a.gpw > 1 ? a.g0 + it * a.gpw : a.items[it];
    fb0 = (int)((long)part * nfb / a.nsplit);
    fb1 = (int)((long)(part + 1) * nfb / a.nsplit);
    return fb0 < fb1;
}

template <int NP>
__device__ __forceinline__ void mx_group(int g, int &m0, int &n0)
{
    if (NP == 2) { m0 = g; n0 = 0; } else so3_group(g, m0, n0);
}

template <int NP>
__device__ __forceinline__ int mx_pair_sign(int m0, int n0, int p)
{
    return NP == 2 ? (p ? parity_sign(m0) : 1) : so3_pair_sign(m0, n0, p);
}

// pair p of the group has a workspace slot of its own (duplicated pairs are never written by the FFT)
template <int NP>
__device__ __forceinline__ bool mx_pair_valid(int m0, int n0, int p)
{
    return NP == 2 ? (p == 0 || m0 > 0) : so3_pair_valid(m0, n0, p);
}

// (m, n) of pair p of the group, and the packed index of degree l of that pair
template <int NP>
struct MxPair {
    int m, n;
    __device__ __forceinline__ MxPair(int m0, int n0, int p)
    {
        if (NP == 2) { m = p ? -m0 : m0; n = 0; } else so3_pair(m0, n0, p, m, n);
    }
    __device__ __forceinline__ int packed(int l) const
    {
        return NP == 2 ? l * l + m + l : so3_off(l) + (m + l) + (n + l) * (2 * l + 1);
    }
};

// LDS column of block column c = (f*NP + p)*2 + ri (f < 64/NP): wave (f/16)*NP + p, lane (f%16)*2 + ri
template <int NP>
__device__ __forceinline__ int mx_lcol(int c)
{
    const int cc = c >> 1, f = cc / NP, p = cc - f * NP;
    return (((f >> 4) * NP + p) << 5) + ((f & 15) << 1) + (c & 1);
}

// Forward: PT(packed(l, pair); fn) = s_p sum_k w_k d^l_{m0 n0}(beta_k) S_k(fn, pair).
// FULL: N == 128; the lane's operands of D (64 per row tile) then stay in registers for the whole
// batch and the multiply is unrolled over all 64 steps, with two independent accumulators.
template <int NP, bool FULL>
__device__ __forceinline__ void contract_fwd_mx32_body(const MxArgs &a, int unit)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *D = reinterpret_cast<float *>(smem_raw);     // [64][MX_DPITCH], rows l - m0
    float *L = D + 64 * MX_DPITCH;                       // [128][MX_LPITCH], spectrum block
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, C = ca.C, CW = C * NP * 2;
    const int nfb = (ca.nfc * NP * 2 + 127) / 128;          // column blocks of 128
    int g, fb0, fb1;
    if (!mx_unit(a, nfb, unit, g, fb0, fb1)) return;
    int m0, n0;
    mx_group<NP>(g, m0, n0);
    const int M = B - m0, nrt = (M + 31) / 32;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kh = lane >> 5, cl = lane & 31;
    constexpr int FB = 64 / NP;                          // functions per block
    // load role: rows r0 + 8i (i < 16) of the block, columns 4q..4q+3
    const int q = tid & 31, r0 = tid >> 5;
    const int fq = (2 * q) / NP;                         // function (within the block) of the 4 columns
    const int lc0 = mx_lcol<NP>(4 * q), lc1 = mx_lcol<NP>(4 * q + 2);
    const float *S = reinterpret_cast<const float *>(a.src);
    // record of function fq of the block for this group (ws_rec: 2^gil groups interleaved)
    const int gm = (1 << ca.gil) - 1;
    const int qoff = (g & gm) * NP * 2 + (fq << ca.gil) * NP * 2 + (4 * q - fq * NP * 2);
    long roff[16];
#pragma unroll
    for (int i = 0; i < 16; i++) roff[i] = ws_row(ca, g & ~gm, min(r0 + 8 * i, N - 1)) * CW + qoff;
    float4 pf[16];
    auto load_block = [&](int fb) {
        const long co = fb * FB + fq < ca.nfc ? ((long)fb * 128 << ca.gil) : 0;
#pragma unroll
        for (int i = 0; i < 16; i++) pf[i] = *reinterpret_cast<const float4 *>(S + roff[i] + co);
    };
    load_block(fb0);
    const int N8 = (N + 7) & ~7;                         // beta extent multiplied (zero columns past N)
    RecCoef *rc = reinterpret_cast<RecCoef *>(smem_raw + MX_LDS_BYTES);
    group_coefs(a.rt, NP == 2, B, m0, n0, rc, tid, 256);
    __syncthreads();
    for (int k = tid; k < N8; k += 256) {
        if (k < N) {
            const double wk = a.rt.w[k];
            group_recursion_lds(a.rt, rc, NP == 2, g, m0, n0, k, B, [&](int l, double v) {
                D[(l - m0) * MX_DPITCH + k] = (float)(wk * v);
            });
            for (int r = M; r < 32 * nrt; r++) D[r * MX_DPITCH + k] = 0.f;
        } else {
            for (int r = 0; r < 32 * nrt; r++) D[r * MX_DPITCH + k] = 0.f;
        }
    }
    // output role of this lane: pair p (uniform over the wave), function fo of the block, re/im cl & 1
    const int p = __builtin_amdgcn_readfirstlane(wave) % NP, fo = (wave / NP) * 16 + (cl >> 1);
    const float sg = (float)mx_pair_sign<NP>(m0, n0, p);
    const bool pvalid = mx_pair_valid<NP>(m0, n0, p);
    const MxPair<NP> pr(m0, n0, p);
    float *PT = reinterpret_cast<float *>(a.dst);
    const float *Lw = L + kh * MX_LPITCH + wave * 32 + cl;
    const float *D0 = D + cl * MX_DPITCH + kh;
    float dA[FULL ? 64 : 1], dB[FULL ? 64 : 1];
    if constexpr (FULL) {
        __syncthreads();                                 // D written
#pragma unroll
        for (int s = 0; s < 64; s++) {
            dA[s] = D0[2 * s];
            dB[s] = nrt > 1 ? D0[32 * MX_DPITCH + 2 * s] : 0.f;
        }
    }
    for (int fb = fb0; fb < fb1; fb++) {
        __syncthreads();                                 // D written / previous block consumed
#pragma unroll
        for (int i = 0; i < 16; i++) {
            float *row = L + (r0 + 8 * i) * MX_LPITCH;
            *reinterpret_cast<float2 *>(row + lc0) = make_float2(pf[i].x, pf[i].y);
            *reinterpret_cast<float2 *>(row + lc1) = make_float2(pf[i].z, pf[i].w);
        }
        __syncthreads();
        if (fb + 1 < fb1) load_block(fb + 1);
        f32x16m_t acc0, acc1;
#pragma unroll
        for (int j = 0; j < 16; j++) { acc0[j] = 0.f; acc1[j] = 0.f; }
        if constexpr (FULL) {
            if (nrt > 1) {
#pragma unroll
                for (int s0 = 0; s0 < 64; s0 += 8) {
                    float b[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) b[u] = Lw[2 * (s0 + u) * MX_LPITCH];
                    __builtin_amdgcn_sched_barrier(0);   // keep the 8 reads ahead of the MFMAs
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(dA[s0 + u], b[u], acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(dB[s0 + u], b[u], acc1, 0, 0, 0);
                    }
                }
            } else {
                // even and odd steps in separate accumulators (independent MFMA chains), summed
#pragma unroll
                for (int s = 0; s < 64; s += 2) {
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(dA[s], Lw[2 * s * MX_LPITCH], acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(dA[s + 1], Lw[(2 * s + 2) * MX_LPITCH], acc1, 0, 0, 0);
                }
#pragma unroll
                for (int j = 0; j < 16; j++) { acc0[j] += acc1[j]; acc1[j] = 0.f; }
            }
        } else if (nrt > 1) {
            // rows k >= N of the block hold finite copies of row N-1 and meet zero columns of D
            for (int s4 = 0; s4 < N8 / 8; s4++) {
                float b[4], d0[4], d1[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int s = 4 * s4 + u;
                    b[u] = Lw[2 * s * MX_LPITCH];
                    d0[u] = D0[2 * s];
                    d1[u] = D0[32 * MX_DPITCH + 2 * s];
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(d0[u], b[u], acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(d1[u], b[u], acc1, 0, 0, 0);
                }
            }
        } else {
            for (int s4 = 0; s4 < N8 / 8; s4++) {
                float b[4], d0[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int s = 4 * s4 + u;
                    b[u] = Lw[2 * s * MX_LPITCH];
                    d0[u] = D0[2 * s];
                }
#pragma unroll
                for (int u = 0; u < 4; u++) acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(d0[u], b[u], acc0, 0, 0, 0);
            }
        }
        const int fn = fb * FB + fo;
        const bool ok = pvalid && fn < ca.nfc;
        float *o = PT + 2 * fn + (cl & 1);
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int r = 8 * (j >> 2) + 4 * kh + (j & 3);
            const long e0 = pr.packed(m0 + min(r, M - 1)), e1 = pr.packed(m0 + min(r + 32, M - 1));
            float *d0 = ok && r < M ? o + 2 * e0 * C : a.dummy;
            float *d1 = ok && r + 32 < M ? o + 2 * e1 * C : a.dummy;
            *d0 = sg * acc0[j];
            *d1 = sg * acc1[j];
        }
    }
}

template <int NP, bool FULL>
__global__ void __launch_bounds__(256) contract_fwd_mx32(MxArgs a)
{
    contract_fwd_mx32_body<NP, FULL>(a, mx_block_unit(a));
}

// Forward, N = 128, SO(3), workspace with 2^gil groups interleaved (GPW = 2^gil): one workgroup
// takes the GPW groups whose records share each workspace line, so a block (16 / GPW functions of
// all GPW groups) is again 512 contiguous bytes of every beta row.  Wave w works on group
// w / (4 / GPW) and on GPW of its pairs, 32 columns as before; each wave keeps its own group's
// D in registers, staged two groups at a time through the LDS that later holds the blocks.
template <int GPW>
__device__ __forceinline__ void contract_fwd_mx32g_body(const MxArgs &a, int unit)
{
    constexpr int NP = 4, FB = 16 / GPW, WPG = 4 / GPW;  // functions per block, waves per group
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *L = reinterpret_cast<float *>(smem_raw);     // [128][MX_LPITCH], spectrum block
    float *Dst = L;                                      // [2][64][MX_DPITCH] staging, before L is used
    RecCoef *rcs = reinterpret_cast<RecCoef *>(smem_raw + 128 * MX_LPITCH * 4);   // [2][B]
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, C = ca.C, CW = C * NP * 2;
    const int nfb = (ca.nfc + FB - 1) / FB;
    int gb, fb0, fb1;
    if (!mx_unit(a, nfb, unit, gb, fb0, fb1)) return;         // gb: first group of the interleaved set
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kh = lane >> 5, cl = lane & 31;
    const int gg = wave / WPG, pg = wave % WPG;          // this wave's group and pair set
    const int g = gb + gg;
    int m0, n0;
    so3_group(g, m0, n0);
    const int M = B - m0, nrt = (M + 31) / 32;
    // load role: rows r0 + 8i (i < 16) of the block, columns 4q..4q+3 = record (fl, gq), pairs
    // 2(q & 1), 2(q & 1) + 1
    const int q = tid & 31, r0 = tid >> 5;
    const int fl_q = (q >> 1) / GPW, gq = (q >> 1) % GPW;
    auto lcol = [](int fl, int gq_, int p) {            // LDS column of (function, group, pair), re part
        return (gq_ * WPG + p / GPW) * 32 + (p % GPW) * FB * 2 + fl * 2;
    };
    const int lc0 = lcol(fl_q, gq, 2 * (q & 1)), lc1 = lcol(fl_q, gq, 2 * (q & 1) + 1);
    const float *S = reinterpret_cast<const float *>(a.src);
    long roff[16];
#pragma unroll
    for (int i = 0; i < 16; i++) roff[i] = ws_row(ca, gb, min(r0 + 8 * i, N - 1)) * CW + 4 * q;
    float4 pf[16];
    auto load_block = [&](int fb) {
        const long co = fb * FB + fl_q < ca.nfc ? (long)fb * 128 : 0;
#pragma unroll
        for (int i = 0; i < 16; i++) pf[i] = *reinterpret_cast<const float4 *>(S + roff[i] + co);
    };
    load_block(fb0);
    // D of the GPW groups, two at a time: threads 0-127 recurse for the first, 128-255 the second
    float dA[64], dB[64];
    for (int ph = 0; ph < (GPW + 1) / 2; ph++) {
        const int h = tid >> 7, gs = 2 * ph + h;        // staged group (index within the set)
        const bool active = gs < GPW;
        int sm0 = 0, sn0 = 0;
        if (active) so3_group(gb + gs, sm0, sn0);
        RecCoef *rc = rcs + h * B;
        float *Dh = Dst + h * 64 * MX_DPITCH;
        if (active) group_coefs(a.rt, 0, B, sm0, sn0, rc, tid & 127, 128);
        __syncthreads();
        if (active) {
            const int k = tid & 127, sM = B - sm0, snrt = (sM + 31) / 32;
            const double wk = a.rt.w[k];
            group_recursion_lds(a.rt, rc, 0, gb + gs, sm0, sn0, k, B, [&](int l, double v) {
                Dh[(l - sm0) * MX_DPITCH + k] = (float)(wk * v);
            });
            for (int r = sM; r < 32 * snrt; r++) Dh[r * MX_DPITCH + k] = 0.f;
        }
        __syncthreads();
        if (gg / 2 == ph) {
            const float *D0 = Dst + (gg & 1) * 64 * MX_DPITCH + cl * MX_DPITCH + kh;
#pragma unroll
            for (int s = 0; s < 64; s++) {
                dA[s] = D0[2 * s];
                dB[s] = nrt > 1 ? D0[32 * MX_DPITCH + 2 * s] : 0.f;
            }
        }
        __syncthreads();                                 // staging read before it is overwritten
    }
    // output role of this lane: pair p, function fo of the block, re/im cl & 1
    const int p = pg * GPW + cl / (FB * 2), fo = (cl % (FB * 2)) >> 1;
    const float sg = (float)so3_pair_sign(m0, n0, p);
    const bool pvalid = so3_pair_valid(m0, n0, p);
    const MxPair<NP> pr(m0, n0, p);
    float *PT = reinterpret_cast<float *>(a.dst);
    const float *Lw = L + kh * MX_LPITCH + wave * 32 + cl;
    for (int fb = fb0; fb < fb1; fb++) {
        __syncthreads();                                 // previous block consumed
#pragma unroll
        for (int i = 0; i < 16; i++) {
            float *row = L + (r0 + 8 * i) * MX_LPITCH;
            *reinterpret_cast<float2 *>(row + lc0) = make_float2(pf[i].x, pf[i].y);
            *reinterpret_cast<float2 *>(row + lc1) = make_float2(pf[i].z, pf[i].w);
        }
        __syncthreads();
        if (fb + 1 < fb1) load_block(fb + 1);
        f32x16m_t acc0, acc1;
#pragma unroll
        for (int j = 0; j < 16; j++) { acc0[j] = 0.f; acc1[j] = 0.f; }
        if (nrt > 1) {
#pragma unroll
            for (int s0 = 0; s0 < 64; s0 += 8) {
                float b[8];
#pragma unroll
                for (int u = 0; u < 8; u++) b[u] = Lw[2 * (s0 + u) * MX_LPITCH];
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(dA[s0 + u], b[u], acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(dB[s0 + u], b[u], acc1, 0, 0, 0);
                }
            }
        } else {
#pragma unroll
            for (int s = 0; s < 64; s += 2) {
                acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(dA[s], Lw[2 * s * MX_LPITCH], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(dA[s + 1], Lw[(2 * s + 2) * MX_LPITCH], acc1, 0, 0, 0);
            }
#pragma unroll
            for (int j = 0; j < 16; j++) { acc0[j] += acc1[j]; acc1[j] = 0.f; }
        }
        const int fn = fb * FB + fo;
        const bool ok = pvalid && fn < ca.nfc;
        float *o = PT + 2 * fn + (cl & 1);
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int r = 8 * (j >> 2) + 4 * kh + (j & 3);
            const long e0 = pr.packed(m0 + min(r, M - 1)), e1 = pr.packed(m0 + min(r + 32, M - 1));
            float *d0 = ok && r < M ? o + 2 * e0 * C : a.dummy;
            float *d1 = ok && r + 32 < M ? o + 2 * e1 * C : a.dummy;
            *d0 = sg * acc0[j];
            *d1 = sg * acc1[j];
        }
    }
}

template <int GPW>
__global__ void __launch_bounds__(256) contract_fwd_mx32g(MxArgs a)
{
    contract_fwd_mx32g_body<GPW>(a, mx_block_unit(a));
}

typedef float f32x4m_t __attribute__((ext_vector_type(4)));

// Forward as contract_fwd_mx32g<2>, on 16 x 16 x 4 MFMA tiles: a group's degree rows are padded to
// a multiple of 16 instead of 32, which halves the MFMA work of the groups with at most 16 degrees
// (44% of the groups at l_max = 63).  Tile layouts (lane i): A[i%16][i/16] = D(l, k), B[i/16][i%16]
// = spectrum(k, column), result row 4(i/16) + j, column i%16.  Wave w takes group w / 2 and pairs
// 2(w % 2), 2(w % 2) + 1: one 16-column tile per pair (8 functions, re/im).
// NH = 2: eight waves, waves w and w + 4 splitting the group's row tiles (w + 4 takes tiles NRT ..
// 2 NRT - 1), so a group with up to 2 NRT x 16 degrees runs at two waves per SIMD.
template <int NRT, bool PAIR = true, int NH = 1>
__device__ __forceinline__ void contract_fwd_mx16g_body(const MxArgs &a, int unit)
{
    constexpr int GPW = 2, NP = 4, FB = 8, WPG = 2, RPI = 8 * NH;   // RPI: block rows per load step
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *L = reinterpret_cast<float *>(smem_raw);     // [128][MXG_LPITCH], spectrum block
    float *Dst = L;                                      // [2][64][MX_DPITCH] staging, before L is used
    RecCoef *rcs = reinterpret_cast<RecCoef *>(smem_raw + MX16G_BLOCK_BYTES);   // [2][B]
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, C = ca.C, CW = C * NP * 2;
    const int nfb = (ca.nfc + FB - 1) / FB;
    int gb, fb0, fb1;
    if (!mx_unit(a, nfb, unit, gb, fb0, fb1)) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kq = lane >> 4, cx = lane & 15;
    const int wq = wave & 3, th = wave >> 2;             // th: row-tile half (NH = 2)
    const int gg = wq / WPG, pg = wq % WPG;
    int m0, n0;
    so3_group(gb + gg, m0, n0);
    const int M = B - m0, nrt = (M + 15) / 16;
    const int t0 = NRT * th, nrw = min(max(nrt - t0, 0), NRT);   // this wave's row tiles t0 + [0, nrw)
    // load role (as contract_fwd_mx32g<2>): rows r0 + RPI i of the block, columns 4q..4q+3
    const int q = tid & 31, r0 = tid >> 5;
    const int fl_q = (q >> 1) / GPW, gq = (q >> 1) % GPW;
    auto lcol = [](int fl, int gq_, int p) { return (gq_ * WPG + p / GPW) * 32 + (p % GPW) * FB * 2 + fl * 2; };
    const int lc0 = lcol(fl_q, gq, 2 * (q & 1)), lc1 = lcol(fl_q, gq, 2 * (q & 1) + 1);
    const float *S = reinterpret_cast<const float *>(a.src);
    // rows r0 + RPI i (< N = 128): the workspace rows are [k][g] (no k blocking with the
    // interleave), so their offsets step by a constant
    const long rbase = ws_row(ca, gb, r0) * CW + 4 * q, rstep = (ws_row(ca, gb, r0 + RPI) - ws_row(ca, gb, r0)) * CW;
    float4 pf[16 / NH];
    auto load_block = [&](int fb) {
        const long co = fb * FB + fl_q < ca.nfc ? (long)fb * 128 : 0;
#pragma unroll
        for (int i = 0; i < 16 / NH; i++) pf[i] = *reinterpret_cast<const float4 *>(S + rbase + i * rstep + co);
    };
    load_block(fb0);
    // D of the two groups: threads 0-127 recurse for the first, 128-255 the second (with NH = 2
    // threads 256-511 only meet the barriers: the recursion has none of its own)
    float dR[NRT][32];
    {
        const int h = min(tid >> 7, 1);
        const bool rec = tid < 256;
        int sm0, sn0;
        so3_group(gb + h, sm0, sn0);
        RecCoef *rc = rcs + h * B;
        float *Dh = Dst + h * 64 * MX_DPITCH;
        if (rec) group_coefs(a.rt, 0, B, sm0, sn0, rc, tid & 127, 128);
        __syncthreads();
        if (rec) {
            const int k = tid & 127, sM = B - sm0;
            const double wk = a.rt.w[k];
            group_recursion_lds(a.rt, rc, 0, gb + h, sm0, sn0, k, B, [&](int l, double v) {
                Dh[(l - sm0) * MX_DPITCH + k] = (float)(wk * v);
            });
            for (int r = sM; r < 16 * ((sM + 15) / 16); r++) Dh[r * MX_DPITCH + k] = 0.f;
        }
        __syncthreads();
        const float *D0 = Dst + gg * 64 * MX_DPITCH + cx * MX_DPITCH + kq;
#pragma unroll
        for (int t = 0; t < NRT; t++)
#pragma unroll
            for (int s = 0; s < 32; s++) dR[t][s] = t < nrw ? D0[16 * (t0 + t) * MX_DPITCH + 4 * s] : 0.f;
        __syncthreads();                                 // staging read before it is overwritten
    }
    // output role of this lane: pairs pg*2 + {0, 1} (column tiles 0, 1), function fo, re/im cx & 1
    const int fo = cx >> 1;
    const MxPair<NP> pr0(m0, n0, pg * 2), pr1(m0, n0, pg * 2 + 1);
    const float sg0 = (float)so3_pair_sign(m0, n0, pg * 2), sg1 = (float)so3_pair_sign(m0, n0, pg * 2 + 1);
    const bool pv0 = so3_pair_valid(m0, n0, pg * 2), pv1 = so3_pair_valid(m0, n0, pg * 2 + 1);
    float *PT = reinterpret_cast<float *>(a.dst);
    // operand rows 4 s + kq: columns swizzled by 16 kq
    f32x4m_t accp[NRT][2];   // results of the first block of a pair
    const float *Lw0 = L + kq * MXG_LPITCH + ((wq * 32 + cx) ^ (16 * kq));
    const float *Lw1 = L + kq * MXG_LPITCH + ((wq * 32 + 16 + cx) ^ (16 * kq));
    const int lsw = 16 * (r0 & 3);
    for (int fb = fb0; fb < fb1; fb++) {
        __syncthreads();                                 // previous block consumed
#pragma unroll
        for (int i = 0; i < 16 / NH; i++) {
            float *row = L + (r0 + RPI * i) * MXG_LPITCH;   // (r0 + RPI i) & 3 = r0 & 3
            *reinterpret_cast<float2 *>(row + (lc0 ^ lsw)) = make_float2(pf[i].x, pf[i].y);
            *reinterpret_cast<float2 *>(row + (lc1 ^ lsw)) = make_float2(pf[i].z, pf[i].w);
        }
        __syncthreads();
        if (fb + 1 < fb1) load_block(fb + 1);
        f32x4m_t acc[NRT][2];
#pragma unroll
        for (int t = 0; t < NRT; t++)
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int j = 0; j < 4; j++) acc[t][c][j] = 0.f;
        auto run = [&](auto R) {
            constexpr int NR = decltype(R)::value;
#pragma unroll
            for (int s0 = 0; s0 < 32; s0 += 4) {
                float b[4][2];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    b[u][0] = Lw0[4 * (s0 + u) * MXG_LPITCH];
                    b[u][1] = Lw1[4 * (s0 + u) * MXG_LPITCH];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int t = 0; t < NR; t++) {
                        acc[t][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(dR[t][s0 + u], b[u][0], acc[t][0], 0, 0, 0);
                        acc[t][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(dR[t][s0 + u], b[u][1], acc[t][1], 0, 0, 0);
                    }
            }
        };
        if (nrw == 1) run(std::integral_constant<int, 1>());
        else if (NRT == 2 || nrw == 2) run(std::integral_constant<int, 2>());
        else if (nrw == 3) run(std::integral_constant<int, NRT >= 3 ? 3 : 2>());
        else if (nrw > 0) run(std::integral_constant<int, NRT>());
        // blocks 2i and 2i + 1 fill the two 64-byte halves of each 128-byte line of PT (16
        // functions): the first block's results wait in registers so both halves are stored by
        // back-to-back instructions
        const bool first = PAIR && !(fb & 1) && fb + 1 < fb1;
        if (first) {
#pragma unroll
            for (int t = 0; t < NRT; t++)
#pragma unroll
                for (int c = 0; c < 2; c++) accp[t][c] = acc[t][c];
            continue;
        }
        const bool paired = PAIR && (fb & 1) && fb > fb0;
        const int fn = fb * FB + fo;
        const bool okf = fn < ca.nfc;
        float *o = PT + 2 * fn + (cx & 1);
#pragma unroll
        for (int t = 0; t < NRT; t++) {
            if (t < nrw) {
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int r = 16 * (t0 + t) + 4 * kq + j;
                    const long e0 = pr0.packed(m0 + min(r, M - 1)), e1 = pr1.packed(m0 + min(r, M - 1));
                    const bool ok0 = pv0 && r < M, ok1 = pv1 && r < M;
                    if (paired) {   // the previous block: functions 8 lower, always < nfc
                        float *p0 = ok0 ? o - 2 * FB + 2 * e0 * C : a.dummy;
                        float *p1 = ok1 ? o - 2 * FB + 2 * e1 * C : a.dummy;
                        *p0 = sg0 * accp[t][0][j];
                        *p1 = sg1 * accp[t][1][j];
                    }
                    float *d0 = okf && ok0 ? o + 2 * e0 * C : a.dummy;
                    float *d1 = okf && ok1 ? o + 2 * e1 * C : a.dummy;
                    *d0 = sg0 * acc[t][0][j];
                    *d1 = sg1 * acc[t][1][j];
                }
            }
        }
    }
}

template <int NRT>
__global__ void __launch_bounds__(256, NRT == 2 ? 2 : 1) contract_fwd_mx16g(MxArgs a)
{
    contract_fwd_mx16g_body<NRT>(a, mx_block_unit(a));
}

// the groups with more than 32 degrees on eight waves, two row tiles each
__global__ void __launch_bounds__(512) contract_fwd_mx16g8(MxArgs a)
{
    contract_fwd_mx16g_body<2, true, 2>(a, mx_block_unit(a));
}

// Inverse: G_k(fn, pair) = s_p sum_l c_l d^l_{m0 n0}(beta_k) PT(packed(l, pair); fn) (c_l = 2l+1
// for SO(3), 1 for S^2).
template <int NP>
__device__ __forceinline__ void contract_inv_mx32_body(const MxArgs &a, int unit)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *D = reinterpret_cast<float *>(smem_raw);     // [64][MX_DPITCH], rows l - m0, columns k
    float *O = D + 64 * MX_DPITCH;                       // [128][MX_LPITCH], output block
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, C = ca.C, CW = C * NP * 2;
    const int nfb = (ca.nfc * NP * 2 + 127) / 128;          // column blocks of 128
    int g, fb0, fb1;
    if (!mx_unit(a, nfb, unit, g, fb0, fb1)) return;
    int m0, n0;
    mx_group<NP>(g, m0, n0);
    const int M = B - m0, M4 = (M + 3) & ~3, nkt = (N + 31) / 32;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kh = lane >> 5, cl = lane & 31;
    constexpr int FB = 64 / NP;
    // operand role of this lane: pair p, function fo of the block, re/im part cl & 1
    const int p = __builtin_amdgcn_readfirstlane(wave) % NP, fo = (wave / NP) * 16 + (cl >> 1);
    const float *PT = reinterpret_cast<const float *>(a.src);
    const MxPair<NP> pr(m0, n0, p);
    int eoff[32];
#pragma unroll
    for (int s = 0; s < 32; s++) eoff[s] = pr.packed(m0 + min(2 * s + kh, M - 1));
    float bA[32], bB[32];
    auto load_block = [&](float *buf, int fb) {
        const int fn = fb * FB + fo;
        const float *src = PT + 2 * (fn < ca.nfc ? fn : 0) + (cl & 1);
#pragma unroll
        for (int s = 0; s < 32; s++) buf[s] = src[2L * eoff[s] * C];
    };
    load_block(bA, fb0);
    RecCoef *rc = reinterpret_cast<RecCoef *>(smem_raw + MX_LDS_BYTES);
    group_coefs(a.rt, NP == 2, B, m0, n0, rc, tid, 256);
    __syncthreads();
    for (int k = tid; k < 32 * nkt; k += 256) {
        if (k < N) {
            group_recursion_lds(a.rt, rc, NP == 2, g, m0, n0, k, B, [&](int l, double v) {
                D[(l - m0) * MX_DPITCH + k] = (float)(NP == 2 ? v : (double)(2 * l + 1) * v);
            });
            for (int r = M; r < M4; r++) D[r * MX_DPITCH + k] = 0.f;
        } else {
            for (int r = 0; r < M4; r++) D[r * MX_DPITCH + k] = 0.f;
        }
    }
    const float sg = (float)mx_pair_sign<NP>(m0, n0, p);
    // store role: rows r0 + 8i (i < 16) of the block, columns 4q..4q+3
    const int q = tid & 31, r0 = tid >> 5;
    const int fq = (2 * q) / NP;
    const int lc0 = mx_lcol<NP>(4 * q), lc1 = mx_lcol<NP>(4 * q + 2);
    float *G = reinterpret_cast<float *>(a.dst);
    const int gm = (1 << ca.gil) - 1;
    const int qoff = (g & gm) * NP * 2 + (fq << ca.gil) * NP * 2 + (4 * q - fq * NP * 2);
    long roff[16];
#pragma unroll
    for (int i = 0; i < 16; i++) roff[i] = r0 + 8 * i < N ? ws_row(ca, g & ~gm, r0 + 8 * i) * CW + qoff : -1;
    auto multiply_store = [&](const float *buf, int fb) {
        f32x16m_t acc[4];
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int j = 0; j < 16; j++) acc[t][j] = 0.f;
        // degrees l >= M read finite coefficients (row M-1) against zero rows of D up to M4
        if (nkt == 4) {
#pragma unroll
            for (int s = 0; s < 32; s += 2) {
                if (4 * (s >> 1) < M4) {
                    const float *Dr = D + (2 * s + kh) * MX_DPITCH + cl;
                    float d[2][4];
#pragma unroll
                    for (int u = 0; u < 2; u++)
#pragma unroll
                        for (int t = 0; t < 4; t++) d[u][t] = Dr[2 * u * MX_DPITCH + 32 * t];
#pragma unroll
                    for (int u = 0; u < 2; u++)
#pragma unroll
                        for (int t = 0; t < 4; t++)
                            acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(d[u][t], buf[s + u], acc[t], 0, 0, 0);
                }
            }
        } else {
#pragma unroll
            for (int s = 0; s < 32; s += 2) {
                if (4 * (s >> 1) < M4) {
                    const float *Dr = D + (2 * s + kh) * MX_DPITCH + cl;
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        if (t < nkt) {
                            acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(Dr[32 * t], buf[s], acc[t], 0, 0, 0);
                            acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(Dr[2 * MX_DPITCH + 32 * t], buf[s + 1],
                                                                          acc[t], 0, 0, 0);
                        }
                    }
                }
            }
        }
        __syncthreads();                                 // previous block's rows stored
#pragma unroll
        for (int t = 0; t < 4; t++) {
            if (t < nkt) {
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    const int k = t * 32 + 8 * (j >> 2) + 4 * kh + (j & 3);
                    O[k * MX_LPITCH + wave * 32 + cl] = sg * acc[t][j];
                }
            }
        }
        __syncthreads();
        const bool ok = fb * FB + fq < ca.nfc;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const float *row = O + (r0 + 8 * i) * MX_LPITCH;
            const float2 u = *reinterpret_cast<const float2 *>(row + lc0);
            const float2 v = *reinterpret_cast<const float2 *>(row + lc1);
            float *d = ok && roff[i] >= 0 ? G + roff[i] + ((long)fb * 128 << ca.gil) : a.dummy;
            *reinterpret_cast<float4 *>(d) = make_float4(u.x, u.y, v.x, v.y);
        }
    };
    __syncthreads();                                     // D written
    for (int fb = fb0; fb < fb1; fb += 2) {
        if (fb + 1 < fb1) load_block(bB, fb + 1);
        multiply_store(bA, fb);
        if (fb + 1 >= fb1) break;
        if (fb + 2 < fb1) load_block(bA, fb + 2);
        multiply_store(bB, fb + 1);
    }
}

template <int NP>
__global__ void __launch_bounds__(256) contract_inv_mx32(MxArgs a)
{
    contract_inv_mx32_body<NP>(a, mx_block_unit(a));
}

// Inverse, N = 128, SO(3), workspace with two groups interleaved (GPW = 2): one workgroup takes
// the two groups whose records share each workspace line, so a block (8 functions of both groups)
// is one contiguous 512-byte run of every beta row and each 128-byte line is written whole by one
// store instruction.  Waves 0-1 work on the first group, 2-3 on the second, two pairs each; the
// two groups' c_l d^l blocks sit side by side in LDS.
template <int GPW>
__device__ __forceinline__ void contract_inv_mx32g_body(const MxArgs &a, int unit)
{
    static_assert(GPW == 2, "the two groups of one interleaved workspace line");
    constexpr int NP = 4, FB = 16 / GPW, WPG = 4 / GPW;  // functions per block, waves per group
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // D: [GPW][64][MXG_DPITCH], rows l - m0, column k stored at k ^ 32 (l - m0 odd)
    // O: [128][MXG_LPITCH], output block, column c of row k at c ^ 32 ((k >> 2) & 1) ^ 16 (k & 1)
    float *D = reinterpret_cast<float *>(smem_raw);
    float *O = D + GPW * 64 * MXG_DPITCH;
    RecCoef *rcs = reinterpret_cast<RecCoef *>(O + 128 * MXG_LPITCH);   // [GPW][B]
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, C = ca.C, CW = C * NP * 2;
    const int nfb = (ca.nfc + FB - 1) / FB;
    int gb, fb0, fb1;
    if (!mx_unit(a, nfb, unit, gb, fb0, fb1)) return;         // gb: first (even) group of the set
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kh = lane >> 5, cl = lane & 31;
    const int gw = wave / WPG;                           // this wave's group within the set
    int m0, n0;
    so3_group(gb + gw, m0, n0);
    const int M = B - m0, M4 = (M + 3) & ~3;
    // operand role of this lane: pair p, function fo of the block, re/im part cl & 1
    const int p = (wave % WPG) * GPW + (cl >> 4), fo = (cl & 15) >> 1;
    const float *PT = reinterpret_cast<const float *>(a.src);
    const MxPair<NP> pr(m0, n0, p);
    int eoff[32];
#pragma unroll
    for (int s = 0; s < 32; s++) eoff[s] = pr.packed(m0 + min(2 * s + kh, M - 1));
    float bA[32], bB[32];
    auto load_block = [&](float *buf, int fb) {
        const int fn = fb * FB + fo;
        const float *src = PT + 2 * (fn < ca.nfc ? fn : 0) + (cl & 1);
#pragma unroll
        for (int s = 0; s < 32; s++) buf[s] = src[2L * eoff[s] * C];
    };
    load_block(bA, fb0);
    {
        // threads 0-127 recurse for the first group, 128-255 for the second: one beta row each
        const int h = tid >> 7, k = tid & 127;
        int hm0, hn0;
        so3_group(gb + h, hm0, hn0);
        RecCoef *rc = rcs + h * B;
        float *Dh = D + h * 64 * MXG_DPITCH;
        group_coefs(a.rt, 0, B, hm0, hn0, rc, k, 128);
        __syncthreads();
        const int hM = B - hm0, hM4 = (hM + 3) & ~3;
        group_recursion_lds(a.rt, rc, 0, gb + h, hm0, hn0, k, B, [&](int l, double v) {
            Dh[(l - hm0) * MXG_DPITCH + (k ^ (((l - hm0) & 1) << 5))] = (float)((double)(2 * l + 1) * v);
        });
        for (int r = hM; r < hM4; r++) Dh[r * MXG_DPITCH + (k ^ ((r & 1) << 5))] = 0.f;
    }
    const float sg = (float)so3_pair_sign(m0, n0, p);
    // store role: rows r0 + 8i (i < 16) of the block, columns 4q..4q+3 = function q / 4, group
    // (q / 2) & 1, pairs 2(q & 1) and 2(q & 1) + 1; computed by wave wq, lanes 2fq + {0,1} and 16 + 2fq + {0,1}
    const int q = tid & 31, r0 = tid >> 5;
    const int fq = q >> 2, wq = ((q >> 1) & 1) * WPG + (q & 1);
    const int osw = 32 * ((r0 >> 2) & 1) + 16 * (r0 & 1);      // rows r0 + 8i share bits 0 and 2
    const int lc0 = (wq * 32 + 2 * fq) ^ osw, lc1 = (wq * 32 + 16 + 2 * fq) ^ osw;
    float *G = reinterpret_cast<float *>(a.dst);
    long roff[16];
#pragma unroll
    for (int i = 0; i < 16; i++) roff[i] = ws_row(ca, gb, r0 + 8 * i) * CW + 4 * q;
    // the lane's rows 2s + kh have parity kh: column 32t + cl sits at 32 (t ^ kh) + cl
    const float *Dg = D + gw * 64 * MXG_DPITCH + kh * MXG_DPITCH + cl;
    const float *DgE = Dg + 32 * kh, *DgO = Dg - 32 * kh;   // even / odd t
    auto multiply_store = [&](const float *buf, int fb) {
        f32x16m_t acc[4];
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int j = 0; j < 16; j++) acc[t][j] = 0.f;
        // degrees l >= M read finite coefficients (row M-1) against zero rows of D up to M4
#pragma unroll
        for (int s = 0; s < 32; s += 2) {
            if (4 * (s >> 1) < M4) {
                float d[2][4];
#pragma unroll
                for (int u = 0; u < 2; u++)
#pragma unroll
                    for (int t = 0; t < 4; t++) d[u][t] = (t & 1 ? DgO : DgE)[(2 * s + 2 * u) * MXG_DPITCH + 32 * t];
#pragma unroll
                for (int u = 0; u < 2; u++)
#pragma unroll
                    for (int t = 0; t < 4; t++)
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(d[u][t], buf[s + u], acc[t], 0, 0, 0);
            }
        }
        __syncthreads();                                 // previous block's rows stored
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int j = 0; j < 16; j++) {
                const int k = t * 32 + 8 * (j >> 2) + 4 * kh + (j & 3);
                O[k * MXG_LPITCH + ((wave * 32 + cl) ^ (32 * kh) ^ (16 * (j & 1)))] = sg * acc[t][j];   // k: bit 2 kh, bit 0 j
            }
        __syncthreads();
        const bool ok = fb * FB + fq < ca.nfc;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const float *row = O + (r0 + 8 * i) * MXG_LPITCH;
            const float2 u = *reinterpret_cast<const float2 *>(row + lc0);
            const float2 v = *reinterpret_cast<const float2 *>(row + lc1);
            float *d = ok ? G + roff[i] + (long)fb * 128 : a.dummy;
            *reinterpret_cast<float4 *>(d) = make_float4(u.x, u.y, v.x, v.y);
        }
    };
    __syncthreads();                                     // D written
    for (int fb = fb0; fb < fb1; fb += 2) {
        if (fb + 1 < fb1) load_block(bB, fb + 1);
        multiply_store(bA, fb);
        if (fb + 1 >= fb1) break;
        if (fb + 2 < fb1) load_block(bA, fb + 2);
        multiply_store(bB, fb + 1);
    }
}

// contract_inv_mx32g with blocks of 16 functions: a block is two 32-column tiles (functions 0-7
// and 8-15) per group and pair couple.  With NW = 4 waves one wave computes both tiles, so a lane's
// two operand loads of one degree read the two 64-byte halves of a 128-byte coefficient line back
// to back (with blocks of 8 functions the other half was read a block later).  With NW = 8 (the
// form launched) waves w and w + 4 take one tile each: the two halves of a line are loaded by two
// waves of one SIMD a few cycles apart, and a wave holds half the accumulators and operands (227
// registers, two waves per SIMD, where NW = 4 needs 467 and runs one).  The 256-column output
// block is staged in LDS 64 beta rows at a time (the same 64 KB as the 128 x 128 block of the
// 8-function form): column c of staged row r at c ^ 8 ((c >> 6) & 1) ^ 4 ((r >> 2) & 1), which
// keeps the accumulator writes free of bank conflicts and float4 groups whole, and each row goes
// out as one contiguous 1 KB segment.  The unified multi-class kernel keeps the 8-function form
// (fewer registers).
template <int NW>
__device__ __forceinline__ void contract_inv_mx32g16_body(const MxArgs &a, int unit)
{
    constexpr int GPW = 2, NP = 4, FB = 16, WPG = 2, OP = 256, NC = NW == 8 ? 1 : 2;   // NC: tiles per wave
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // D: [GPW][64][MXG_DPITCH], rows l - m0, column k stored at k ^ 32 (l - m0 odd)
    float *D = reinterpret_cast<float *>(smem_raw);
    float *O = D + GPW * 64 * MXG_DPITCH;                            // [64][OP]
    RecCoef *rcs = reinterpret_cast<RecCoef *>(O + 64 * OP);          // [GPW][B]
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, C = ca.C, CW = C * NP * 2;
    const int nfb = (ca.nfc + FB - 1) / FB;
    int gb, fb0, fb1;
    if (!mx_unit(a, nfb, unit, gb, fb0, fb1)) return;         // gb: first (even) group of the set
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kh = lane >> 5, cl = lane & 31;
    const int wq = wave & 3, ch = NW == 8 ? wave >> 2 : 0;   // ch: the wave's first tile
    const int gw = wq / WPG, pw = wq % WPG;                  // group and pair couple of this wave
    int m0, n0;
    so3_group(gb + gw, m0, n0);
    const int M = B - m0, M4 = (M + 3) & ~3;
    // operand role of this lane: pair p, function fo of each tile, re/im part cl & 1
    const int p = pw * GPW + (cl >> 4), fo = (cl & 15) >> 1;
    const float *PT = reinterpret_cast<const float *>(a.src);
    const MxPair<NP> pr(m0, n0, p);
    int eoff[32];
#pragma unroll
    for (int s = 0; s < 32; s++) eoff[s] = pr.packed(m0 + min(2 * s + kh, M - 1));
    float bA[NC][32], bB[NC][32];
    auto load_block = [&](float (&buf)[NC][32], int fb) {
        const int f0 = fb * FB + 8 * ch + fo, f1 = f0 + 8;
        const float *s0 = PT + 2 * (f0 < ca.nfc ? f0 : 0) + (cl & 1);
        const float *s1 = PT + 2 * (f1 < ca.nfc ? f1 : 0) + (cl & 1);
#pragma unroll
        for (int s = 0; s < 32; s++) {
            buf[0][s] = s0[2L * eoff[s] * C];
            if (NC == 2) buf[NC - 1][s] = s1[2L * eoff[s] * C];
        }
    };
    load_block(bA, fb0);
    {
        // threads 0-127 recurse for the first group, 128-255 for the second: one beta row each
        // (with NW = 8 threads 256-511 only meet the barrier: the recursion has none of its own)
        const int h = min(tid >> 7, 1), k = tid & 127;
        const bool rec = tid < 256;
        int hm0, hn0;
        so3_group(gb + h, hm0, hn0);
        RecCoef *rc = rcs + h * B;
        float *Dh = D + h * 64 * MXG_DPITCH;
        if (rec) group_coefs(a.rt, 0, B, hm0, hn0, rc, k, 128);
        __syncthreads();
        if (rec) {
            const int hM = B - hm0, hM4 = (hM + 3) & ~3;
            group_recursion_lds(a.rt, rc, 0, gb + h, hm0, hn0, k, B, [&](int l, double v) {
                Dh[(l - hm0) * MXG_DPITCH + (k ^ (((l - hm0) & 1) << 5))] = (float)((double)(2 * l + 1) * v);
            });
            for (int r = hM; r < hM4; r++) Dh[r * MXG_DPITCH + (k ^ ((r & 1) << 5))] = 0.f;
        }
    }
    const float sg = (float)so3_pair_sign(m0, n0, p);
    // block column of tile c, lane cl: function 8c + fo, group gw, pair p, re/im
    const int colb = 16 * fo + 8 * gw + 2 * p + (cl & 1);
    float *G = reinterpret_cast<float *>(a.dst);
    // the lane's rows 2s + kh have parity kh: column 32t + cl sits at 32 (t ^ kh) + cl
    const float *Dg = D + gw * 64 * MXG_DPITCH + kh * MXG_DPITCH + cl;
    const float *DgE = Dg + 32 * kh, *DgO = Dg - 32 * kh;   // even / odd t
    auto multiply_store = [&](const float (&buf)[NC][32], int fb) {
        f32x16m_t acc[4][NC];
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int j = 0; j < 16; j++) acc[t][c][j] = 0.f;
        // degrees l >= M read finite coefficients (row M-1) against zero rows of D up to M4
#pragma unroll
        for (int s = 0; s < 32; s += 2) {
            if (4 * (s >> 1) < M4) {
                float d[2][4];
#pragma unroll
                for (int u = 0; u < 2; u++)
#pragma unroll
                    for (int t = 0; t < 4; t++) d[u][t] = (t & 1 ? DgO : DgE)[(2 * s + 2 * u) * MXG_DPITCH + 32 * t];
#pragma unroll
                for (int u = 0; u < 2; u++)
#pragma unroll
                    for (int t = 0; t < 4; t++)
#pragma unroll
                        for (int c = 0; c < NC; c++)
                            acc[t][c] = __builtin_amdgcn_mfma_f32_32x32x2f32(d[u][t], buf[c][s + u], acc[t][c], 0, 0, 0);
            }
        }
        const bool okb = fb * FB < ca.nfc;
#pragma unroll
        for (int hh = 0; hh < 2; hh++) {
            __syncthreads();                             // previous half's rows stored
#pragma unroll
            for (int tt = 0; tt < 2; tt++)
#pragma unroll
                for (int c = 0; c < NC; c++)
#pragma unroll
                    for (int j = 0; j < 16; j++) {
                        const int r = tt * 32 + 8 * (j >> 2) + 4 * kh + (j & 3);   // row of the half
                        const int col = 128 * (ch + c) + colb;
                        O[r * OP + (col ^ (8 * ((col >> 6) & 1)) ^ (4 * kh))] = sg * acc[2 * hh + tt][c][j];
                    }
            __syncthreads();
            // rows w + NW i of the half, lanes 16-byte groups of one 1 KB row segment
            int rr = wave;
            asm volatile("" : "+v"(rr));   // row offsets formed at the stores
#pragma unroll
            for (int i = 0; i < 64 / NW; i++) {
                const int r = rr + NW * i, pos = 4 * lane;
                const int col = pos ^ (8 * ((pos >> 6) & 1)) ^ (4 * ((r >> 2) & 1));   // logical column of pos
                const float4 v = *reinterpret_cast<const float4 *>(O + r * OP + pos);
                const int k = 64 * hh + r;
                const bool ok = okb && fb * FB + (col >> 4) < ca.nfc;
                float *dst = ok ? G + ws_row(ca, gb, k) * CW + (long)fb * OP + col : a.dummy;
                *reinterpret_cast<float4 *>(dst) = v;
            }
        }
    };
    __syncthreads();                                     // D written
    for (int fb = fb0; fb < fb1; fb += 2) {
        if (fb + 1 < fb1) load_block(bB, fb + 1);
        multiply_store(bA, fb);
        if (fb + 1 >= fb1) break;
        if (fb + 2 < fb1) load_block(bA, fb + 2);
        multiply_store(bB, fb + 1);
    }
}

__global__ void __launch_bounds__(512) contract_inv_mx32g16w8(MxArgs a)
{
    contract_inv_mx32g16_body<8>(a, mx_block_unit(a));
}

template <int GPW>
__global__ void __launch_bounds__(256) contract_inv_mx32g(MxArgs a)
{
    contract_inv_mx32g_body<GPW>(a, mx_block_unit(a));
}

// ---------------------------------------------------------------------------------------------
// Coefficient transpose between the reference's packed layout (FFTSO3_matrix / FFTS2_matrix
// storage, one function after another: [function][e]) and PT[e][function].  Tiles of 64 packed
// indices x 32 functions go through LDS so both sides move in contiguous 512- and 256-byte runs.
// DIR 0: PT -> packed, DIR 1: packed -> PT.
// ---------------------------------------------------------------------------------------------
template <typename T2> constexpr int coef_transpose_te() { return sizeof(T2) == 8 ? 128 : 64; }

// launch grid of coef_transpose<., T2> for ncoef packed indices x nfc functions
template <typename T2> inline int coef_transpose_grid(long ncoef, int nfc)
{
    constexpr int TE = coef_transpose_te<T2>();
    return xcd_grid((int)(((ncoef + TE - 1) / TE) * ((nfc + 31) / 32)));
}

template <int DIR, typename T2 = float2>
__device__ __forceinline__ void coef_transpose_body(const ChunkArgs &ca, int ncoef, const void *src, void *dst, int unit)
{
    constexpr int TE = coef_transpose_te<T2>(), TF = 32;  // tile: TE packed indices x 32 functions
    constexpr int RUNS = TE / 64, NJ = 8 * RUNS;          // 64-index runs per row; items per thread
    __shared__ T2 tile[TF][TE + 1];
    __shared__ long poff[TF];
    const int nte = (ncoef + TE - 1) / TE, ntf = (ca.nfc + TF - 1) / TF, nunit = nte * ntf;
    if (unit < 0 || unit >= nunit) return;
    const int et = unit / ntf, ft = unit - et * ntf;
    const int e0 = et * TE, f0 = ft * TF;
    const int t = threadIdx.x;
    if (t < TF && f0 + t < ca.nfc) poff[t] = ca.coef_off[ca.fn[f0 + t]];
    __syncthreads();
    // packed side: wave w moves function rows w + 4j, its 64 lanes runs of 64 consecutive packed
    // indices; PT side: half-wave h moves element rows h + 8j, its 32 lanes 32 consecutive
    // functions.  Every instruction moves whole contiguous runs on both sides.
    const int pe = t & 63, pw = t >> 6;
    const int qf = t & 31, qh = t >> 5;
    T2 v[NJ];
#pragma unroll
    for (int j = 0; j < NJ; j++) v[j] = T2{};
    if (DIR == 1) {
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int fl = pw + 4 * (j / RUNS), e = e0 + pe + 64 * (j % RUNS);
            if (f0 + fl < ca.nfc && e < ncoef) v[j] = reinterpret_cast<const T2 *>(src)[poff[fl] + e];
        }
#pragma unroll
        for (int j = 0; j < NJ; j++) tile[pw + 4 * (j / RUNS)][pe + 64 * (j % RUNS)] = v[j];
        __syncthreads();
        if (f0 + qf < ca.nfc) {
            T2 *Q = reinterpret_cast<T2 *>(dst) + f0 + qf;
#pragma unroll
            for (int j = 0; j < TE / 8; j++) {
                const int el = qh + 8 * j;
                if (e0 + el < ncoef) Q[(long)(e0 + el) * ca.C] = tile[qf][el];
            }
        }
    } else {
        if (f0 + qf < ca.nfc) {
            const T2 *Q = reinterpret_cast<const T2 *>(src) + f0 + qf;
#pragma unroll
            for (int j = 0; j < TE / 8; j++) {
                const int el = qh + 8 * j;
                if (e0 + el < ncoef) v[j] = Q[(long)(e0 + el) * ca.C];
            }
        }
#pragma unroll
        for (int j = 0; j < TE / 8; j++) tile[qf][qh + 8 * j] = v[j];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int fl = pw + 4 * (j / RUNS), e = e0 + pe + 64 * (j % RUNS);
            if (f0 + fl < ca.nfc && e < ncoef) reinterpret_cast<T2 *>(dst)[poff[fl] + e] = tile[fl][pe + 64 * (j % RUNS)];
        }
    }
}

// tiles of coef_transpose for ncoef packed indices x nfc functions
template <typename T2> __host__ __device__ inline int coef_transpose_units(long ncoef, int nfc)
{
    constexpr int TE = coef_transpose_te<T2>();
    return (int)(((ncoef + TE - 1) / TE) * ((nfc + 31) / 32));
}

template <int DIR, typename T2 = float2>
__global__ void __launch_bounds__(256) coef_transpose(ChunkArgs ca, int ncoef, const void *src, void *dst)
{
    coef_transpose_body<DIR, T2>(ca, ncoef, src, dst, xcd_unit(coef_transpose_units<T2>(ncoef, ca.nfc)));
}

// ---------------------------------------------------------------------------------------------
// Batch-wide contraction for 128 < N with the group's Wigner-d block resident in LDS
// (32 * ceil(M/32) rows x N columns, pitch N+1: up to M = 128, N = 256 -- S^2 at l_max = 127,
// SO(3) up to l_max = 127).  Eight waves per workgroup; wave w multiplies block columns
// [32w, 32w+32) of a 256-column block in their workspace order (function, pair, re/im), so its
// coefficient rows are 64 contiguous bytes per pair.  The forward loads each lane's spectrum
// operands straight from HBM into registers, eight MFMA steps ahead, with no LDS staging and no
// barrier in the block loop (an earlier form staged 32-row chunks of the block through LDS with
// four waves: S^2 l_max = 127 x 4096 forward 1.39-1.53 ms).  A lane's pair varies with its column,
// so the packed index of each (degree, pair) comes from a per-group table in LDS.
// ---------------------------------------------------------------------------------------------
#define MXK_MAXROWS 128

// forward: D [32 nrt][N32 + 1] + index table [NP][MXK_MAXROWS]; inverse: D [M4][N + 1] + index table
__host__ __device__ inline size_t mxk_lds_bytes(int N, int B, int NP, bool fwd)
{
    const int N32 = (N + 31) & ~31;
    const size_t table = (size_t)NP * MXK_MAXROWS * 4;
    if (fwd) return (size_t)32 * ((B + 31) / 32) * (N32 + 1) * 4 + table;
    return (size_t)((B + 15) & ~15) * (N + 1) * 4 + table;
}

template <int NP, int NW>
__device__ __forceinline__ void contract_fwd_mxk_body(const MxArgs &a, int unit)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, C = ca.C, CW = C * NP * 2, NCOL = ca.nfc * NP * 2;
    const int nfb = (NCOL + 32 * NW - 1) / (32 * NW);       // column blocks of 32 NW
    int g, fb0, fb1;
    if (!mx_unit(a, nfb, unit, g, fb0, fb1)) return;
    int m0, n0;
    mx_group<NP>(g, m0, n0);
    const int N32 = (N + 31) & ~31, NS = N32 / 2;
    const int M = B - m0, nrt = (M + 31) / 32, DP = N32 + 1;
    float *D = reinterpret_cast<float *>(smem_raw);                 // [32 nrt][DP]
    int *etab = reinterpret_cast<int *>(D + (size_t)32 * nrt * DP); // [NP][MXK_MAXROWS]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kh = lane >> 5, cl = lane & 31;
    RecCoef *rc = reinterpret_cast<RecCoef *>(smem_raw + mxk_lds_bytes(N, a.mrows ? a.mrows : B, NP, true));
    group_coefs(a.rt, NP == 2, B, m0, n0, rc, tid, 64 * NW);
    __syncthreads();
    for (int k = tid; k < N32; k += 64 * NW) {
        if (k < N) {
            const double wk = a.rt.w[k];
            group_recursion_lds(a.rt, rc, NP == 2, g, m0, n0, k, B, [&](int l, double v) {
                D[(l - m0) * DP + k] = (float)(wk * v);
            });
            for (int r = M; r < 32 * nrt; r++) D[r * DP + k] = 0.f;
        } else {
            for (int r = 0; r < 32 * nrt; r++) D[r * DP + k] = 0.f;
        }
    }
    for (int i = tid; i < NP * MXK_MAXROWS; i += 64 * NW) {
        const int p = i / MXK_MAXROWS, r = i - p * MXK_MAXROWS;
        etab[i] = MxPair<NP>(m0, n0, p).packed(m0 + min(r, M - 1));
    }
    __syncthreads();
    // block column 32 wave + cl -> (function, pair, re/im); the lane's B operand of step s is the
    // spectrum at beta 2 s + kh, loaded straight into registers eight steps ahead (rows past N - 1
    // read row N - 1: finite values that meet zero columns of D)
    const int cb = 32 * wave + cl;
    const int p = (cb >> 1) % NP;
    const float sg = (float)mx_pair_sign<NP>(m0, n0, p);
    const bool pvalid = mx_pair_valid<NP>(m0, n0, p);
    const float *S = reinterpret_cast<const float *>(a.src);
    float *PT = reinterpret_cast<float *>(a.dst);
    const float *Dw = D + cl * DP + kh;
    for (int fb = fb0; fb < fb1; fb++) {
        const int c = fb * 32 * NW + cb, cv = c < NCOL ? c : 0;
        const float *src = S + cv;
        auto load = [&](float *x, int s0) {
#pragma unroll
            for (int u = 0; u < 8; u++) x[u] = src[ws_row(ca, g, min(2 * (s0 + u) + kh, N - 1)) * CW];
        };
        f32x16m_t acc[4];
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int j = 0; j < 16; j++) acc[t][j] = 0.f;
        float x[8], xn[8];
        load(x, 0);
        for (int s0 = 0; s0 < NS; s0 += 8) {
            if (s0 + 8 < NS) load(xn, s0 + 8);
            // the group's 8 x nrt operands of D read from LDS ahead of their MFMAs (NS is a multiple
            // of 16, so every group is whole)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                if (t < nrt) {
                    float d[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) d[u] = Dw[32 * t * DP + 2 * (s0 + u)];
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int u = 0; u < 8; u++) acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(d[u], x[u], acc[t], 0, 0, 0);
                }
            }
#pragma unroll
            for (int u = 0; u < 8; u++) x[u] = xn[u];
        }
        const int fn = c / (2 * NP);
        const bool ok = pvalid && c < NCOL;
        float *o = PT + 2 * fn + (c & 1);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            if (t < nrt) {
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    const int r = 32 * t + 8 * (j >> 2) + 4 * kh + (j & 3);
                    float *d = ok && r < M ? o + 2L * etab[p * MXK_MAXROWS + r] * C : a.dummy;
                    *d = sg * acc[t][j];
                }
            }
        }
    }
}

template <int NP>
__global__ void __launch_bounds__(512) contract_fwd_mxk(MxArgs a)
{
    contract_fwd_mxk_body<NP, 8>(a, mx_block_unit(a));
}


// Inverse with the group's c_l d^l block resident in LDS ([M16][N+1], rows l - m0 zero-padded to a
// multiple of 16, so each half of a group of 16 MFMA steps runs whole); wave w produces block
// columns [32w, 32w+32) for every beta row.
// NW waves per workgroup, each owning 32 columns of a block: 8 waves (2 per SIMD) share the
// group's d block so one wave's MFMAs cover another's LDS and HBM waits.
template <int NP, int NW, bool PF = true>
__device__ __forceinline__ void contract_inv_mxk_body(const MxArgs &a, int unit)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, C = ca.C, CW = C * NP * 2, NCOL = ca.nfc * NP * 2;
    const int nfb = (NCOL + 32 * NW - 1) / (32 * NW);       // column blocks of 32 NW
    int g, fb0, fb1;
    if (!mx_unit(a, nfb, unit, g, fb0, fb1)) return;
    int m0, n0;
    mx_group<NP>(g, m0, n0);
    const int M = B - m0, M16 = (M + 15) & ~15, DP = N + 1, nkt = (N + 31) / 32;
    float *D = reinterpret_cast<float *>(smem_raw);                 // [M16][DP], rows >= M zero
    int *etab = reinterpret_cast<int *>(D + (size_t)M16 * DP);      // [NP][MXK_MAXROWS]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kh = lane >> 5, cl = lane & 31;
    for (int i = tid; i < NP * MXK_MAXROWS; i += 64 * NW) {
        const int p = i / MXK_MAXROWS, r = i - p * MXK_MAXROWS;
        etab[i] = MxPair<NP>(m0, n0, p).packed(m0 + min(r, M - 1));
    }
    RecCoef *rc = reinterpret_cast<RecCoef *>(smem_raw + mxk_lds_bytes(N, a.mrows ? a.mrows : B, NP, false));
    group_coefs(a.rt, NP == 2, B, m0, n0, rc, tid, 64 * NW);
    __syncthreads();
    for (int k = tid; k < 32 * nkt; k += 64 * NW) {
        if (k < N) {
            group_recursion_lds(a.rt, rc, NP == 2, g, m0, n0, k, B, [&](int l, double v) {
                D[(l - m0) * DP + k] = (float)(NP == 2 ? v : (double)(2 * l + 1) * v);
            });
            for (int r = M; r < M16; r++) D[r * DP + k] = 0.f;
        } else if (k < DP) {
            for (int r = 0; r < M16; r++) D[r * DP + k] = 0.f;
        }
    }
    __syncthreads();
    const int cb = 32 * wave + cl, p = (cb >> 1) % NP;
    const float sg = (float)mx_pair_sign<NP>(m0, n0, p);
    const float *PT = reinterpret_cast<const float *>(a.src);
    float *G = reinterpret_cast<float *>(a.dst);
    const int *et = etab + p * MXK_MAXROWS;
    // coefficient rows l = m0 + 2s + kh of block fb's column, 16 steps (32 degrees) at a time up to
    // M; rows past M - 1 read row M - 1 (finite values against the zero rows of D).  The next
    // block's are loaded while the current block multiplies.
    auto load = [&](float *x, int fb) {
        const int c = fb * 32 * NW + cb, cv = c < NCOL ? c : 0;
        const float *src = PT + 2 * (cv / (2 * NP)) + (cv & 1);
        int kb = kh;   // the row offsets are read from the table at each load, not held across blocks
        asm volatile("" : "+v"(kb));
#pragma unroll
        for (int s0 = 0; s0 < 64; s0 += 16)
            if (2 * s0 < M)
#pragma unroll
                for (int u = 0; u < 16; u++) x[s0 + u] = src[2L * et[2 * (s0 + u) + kb] * C];
    };
    float buf[64] = {}, nbuf[64] = {};   // steps past the group's last 16-step group stay unused
    if (PF) load(buf, fb0);
    for (int fb = fb0; fb < fb1; fb++) {
        const int c = fb * 32 * NW + cb;
        if (!PF) load(buf, fb);
        else if (fb + 1 < fb1) load(nbuf, fb + 1);
        for (int t = 0; t < nkt; t++) {
            f32x16m_t acc;
#pragma unroll
            for (int j = 0; j < 16; j++) acc[j] = 0.f;
            // when N is not a multiple of 32 the last tile also reads columns k > N, which run into
            // the next row (or the index table): finite or not, they only reach rows k >= N of
            // the result, which are never stored
            const float *Dt = D + kh * DP + 32 * t + cl;
            // 16 operand reads from LDS ahead of their 16 MFMAs, so no MFMA waits on its own read;
            // the second 8 steps (16 degrees) of a group only where the group's degrees reach them
#pragma unroll
            for (int s0 = 0; s0 < 64; s0 += 16) {
                if (2 * s0 < M) {
                    const bool hi = 2 * (s0 + 8) < M;
                    float d[16];
#pragma unroll
                    for (int u = 0; u < 8; u++) d[u] = Dt[2 * (s0 + u) * DP];
                    if (hi)
#pragma unroll
                        for (int u = 8; u < 16; u++) d[u] = Dt[2 * (s0 + u) * DP];
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int u = 0; u < 8; u++) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(d[u], buf[s0 + u], acc, 0, 0, 0);
                    if (hi)
#pragma unroll
                        for (int u = 8; u < 16; u++) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(d[u], buf[s0 + u], acc, 0, 0, 0);
                }
            }
            const bool ok = c < NCOL;
            // the row offsets are formed at the stores rather than held across the block loop
            int kt = 32 * t + 4 * kh;
            asm volatile("" : "+v"(kt));
#pragma unroll
            for (int j = 0; j < 16; j++) {
                const int k = kt + 8 * (j >> 2) + (j & 3);
                float *d = ok && k < N ? G + ws_row(ca, g, k) * CW + c : a.dummy;
                *d = sg * acc[j];
            }
        }
        if (PF)
#pragma unroll
            for (int s = 0; s < 64; s++) buf[s] = nbuf[s];
    }
}

template <int NP, int NW>
__global__ void __launch_bounds__(64 * NW) contract_inv_mxk(MxArgs a)
{
    contract_inv_mxk_body<NP, NW>(a, mx_block_unit(a));
}

// the groups of at most 64 degrees: a 64-row d block and no operand prefetch, so that two
// workgroups share a CU
template <int NP>
__global__ void __launch_bounds__(512, 2) contract_inv_mxk64(MxArgs a)
{
    contract_inv_mxk_body<NP, 8, false>(a, mx_block_unit(a));
}

// ---------------------------------------------------------------------------------------------
// Inverse contraction of the chunked path at large bandlimits (FP32, Wigner-d table present):
// workgroup = (group, 128 beta values); the chunk's coefficient block of the group (M rows x up to
// 64 columns (function, pair, re/im), pair signs and the SO(3) factor 2l+1 applied) is staged in
// LDS once, the d table rows stream from HBM as the MFMA A operand (32 consecutive k per row:
// 128-byte reads, 16 steps per batch with the next batch in flight), wave w owns beta rows
// [32w, 32w+32) of the tile.  Columns follow the workspace row order, so each output row is one
// contiguous segment.
// ---------------------------------------------------------------------------------------------
// Forward on the d table (chunked path above l_max = 127, FP32): one workgroup per group.  The
// group's spectrum block (N beta rows x the chunk's <= 64 columns, scaled by w_k) is staged in LDS
// once; each wave then takes 32-degree row tiles and multiplies table rows (A operand, FP32 from
// HBM/L2) by the block on FP32 MFMA.  Step s of a tile covers beta rows s (lanes 0-31) and
// s + N/2 (lanes 32-63); the second half of the rows is stored 32 columns rotated so the two
// halves' operand reads fall in opposite LDS banks.
#define TABF_COLS 64

template <int NP>
__global__ void __launch_bounds__(256) contract_fwd_tab(ConArgs<float> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *Sl = reinterpret_cast<float *>(smem_raw);    // [N][64]
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, C = ca.C, CW = C * NP * 2, NCOL = ca.nfc * NP * 2, H = N / 2;
    const int g = xcd_deal(1);
    if (g >= ca.ngrp) return;
    int m0, n0;
    mx_group<NP>(g, m0, n0);
    const int M = B - m0, nt = (M + 31) / 32;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kh = lane >> 5, cl = lane & 31;
    // stage w_k S_k: CW / 4 float4 per beta row; row k >= H rotated by 32 columns
    {
        const float4 *S4 = reinterpret_cast<const float4 *>(reinterpret_cast<const float *>(a.src) + ws_row(ca, g, 0) * CW);
        const int q4 = CW / 4, ne = N * q4;
        for (int e0 = 0; e0 < ne; e0 += 256 * 8) {
            float4 v[8];
            double w[8];
#pragma unroll
            for (int i = 0; i < 8; i++) {                 // 8 loads in flight per thread
                const int e = e0 + tid + 256 * i;
                if (e < ne) {
                    v[i] = S4[e];
                    w[i] = a.rt.w[e / q4];
                }
            }
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int e = e0 + tid + 256 * i;
                if (e < ne) {
                    const int k = e / q4, c = 4 * (e - k * q4);
                    const float4 u = make_float4((float)(w[i] * v[i].x), (float)(w[i] * v[i].y), (float)(w[i] * v[i].z),
                                                 (float)(w[i] * v[i].w));
                    *reinterpret_cast<float4 *>(Sl + k * TABF_COLS + ((c + (k >= H ? 32 : 0)) & (TABF_COLS - 1))) = u;
                }
            }
        }
    }
    __syncthreads();
    const float *tab = reinterpret_cast<const float *>(a.dtab) + a.dtab_off[g];
    const int nct = (ca.hsplit ? CW : NCOL) > 32 ? 2 : 1;   // column tiles holding functions of the chunk
    // this lane's output column: function fi, pair p, re/im
    const int c0 = cl, c1 = 32 + cl;
    const float *Sb0 = Sl + kh * H * TABF_COLS + ((c0 + 32 * kh) & (TABF_COLS - 1));
    const float *Sb1 = Sl + kh * H * TABF_COLS + ((c1 + 32 * kh) & (TABF_COLS - 1));
    float *dst = reinterpret_cast<float *>(a.dst);
    const long es = a.pt ? 2L * C : 2L;
    for (int t = wave; t < nt; t += 4) {
        const int lr = 32 * t + cl;                       // this lane's A row (degree m0 + lr)
        const float *trow = tab + (long)min(lr, M - 1) * N + kh * H;   // rows >= M: finite, never stored
        f32x16m_t acc0, acc1;
#pragma unroll
        for (int j = 0; j < 16; j++) { acc0[j] = 0.f; acc1[j] = 0.f; }
        int s0 = 0;
        if ((H & 31) == 0 && (N & 3) == 0) {
            // table rows as float4, 32 steps per round with the next round's 8 loads in flight
            const float4 *t4 = reinterpret_cast<const float4 *>(trow);
            float4 cur[8], nxt[8];
#pragma unroll
            for (int i = 0; i < 8; i++) cur[i] = t4[i];
            for (; s0 < H; s0 += 32) {
                if (s0 + 32 < H) {
#pragma unroll
                    for (int i = 0; i < 8; i++) nxt[i] = t4[(s0 + 32) / 4 + i];
                }
#pragma unroll
                for (int u = 0; u < 32; u++) {
                    const float4 &q = cur[u >> 2];
                    const float av = (u & 3) == 0 ? q.x : (u & 3) == 1 ? q.y : (u & 3) == 2 ? q.z : q.w;
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(av, Sb0[(s0 + u) * TABF_COLS], acc0, 0, 0, 0);
                    if (nct > 1) acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(av, Sb1[(s0 + u) * TABF_COLS], acc1, 0, 0, 0);
                }
#pragma unroll
                for (int i = 0; i < 8; i++) cur[i] = nxt[i];
            }
        }
        for (; s0 + 8 <= H; s0 += 8) {
            float av[8];
#pragma unroll
            for (int u = 0; u < 8; u++) av[u] = trow[s0 + u];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(av[u], Sb0[(s0 + u) * TABF_COLS], acc0, 0, 0, 0);
                if (nct > 1) acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(av[u], Sb1[(s0 + u) * TABF_COLS], acc1, 0, 0, 0);
            }
        }
        for (; s0 < H; s0++) {
            const float av = trow[s0];
            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(av, Sb0[s0 * TABF_COLS], acc0, 0, 0, 0);
            if (nct > 1) acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(av, Sb1[s0 * TABF_COLS], acc1, 0, 0, 0);
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int c = h ? c1 : c0;
            if (c >= CW || (h && nct < 2)) continue;
            int fi, p, ri;
            ws_col(ca, NP, c, fi, p, ri);
            if (fi >= ca.nfc || !mx_pair_valid<NP>(m0, n0, p)) continue;
            const float sg = (float)mx_pair_sign<NP>(m0, n0, p);
            const MxPair<NP> pr(m0, n0, p);
            float *out = dst + (a.pt ? 2L * fi : 2L * ca.coef_off[ca.fn[fi]]) + ri;
#pragma unroll
            for (int j = 0; j < 16; j++) {
                const int r = 32 * t + 8 * (j >> 2) + 4 * kh + (j & 3);
                if (r < M) out[es * pr.packed(m0 + r)] = sg * (h ? acc1[j] : acc0[j]);
            }
        }
    }
}

// Forward on the d table with the spectrum block staged one half of the beta range at a time:
// 64 KB of LDS at l_max = 255, so two workgroups share a CU and one's staging overlaps the
// other's MFMAs.  Within a half, step u of block j covers beta rows 8j + 4kh + u (kh = lane / 32),
// so a lane reads its table row as one float4 per block; rows with bit 2 set are stored 32
// columns rotated so the two lane halves' operand reads fall in opposite banks.  A wave keeps the
// accumulators of up to two 32-degree tiles (t = wave, wave + 4) across both halves: B <= 256,
// and B % 8 == 0 so that every block is whole.
#define TABF2_TPW 2
template <int NP>
__global__ void __launch_bounds__(256) contract_fwd_tab2(ConArgs<float> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *Sl = reinterpret_cast<float *>(smem_raw);    // [H][TABF_COLS]
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, C = ca.C, CW = C * NP * 2, NCOL = ca.nfc * NP * 2, H = N / 2, nb = H / 8;
    const int g = xcd_deal(1);
    if (g >= ca.ngrp) return;
    int m0, n0;
    mx_group<NP>(g, m0, n0);
    const int M = B - m0, nt = (M + 31) / 32;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kh = lane >> 5, cl = lane & 31;
    const float *tab = reinterpret_cast<const float *>(a.dtab) + a.dtab_off[g];
    const int nct = (ca.hsplit ? CW : NCOL) > 32 ? 2 : 1;   // column tiles holding functions of the chunk
    const int c0 = cl, c1 = 32 + cl;
    const float *Sb0 = Sl + 4 * kh * TABF_COLS + ((c0 + 32 * kh) & (TABF_COLS - 1));
    const float *Sb1 = Sl + 4 * kh * TABF_COLS + ((c1 + 32 * kh) & (TABF_COLS - 1));
    f32x16m_t acc[TABF2_TPW][2];
#pragma unroll
    for (int i = 0; i < TABF2_TPW; i++)
#pragma unroll
        for (int j = 0; j < 16; j++) { acc[i][0][j] = 0.f; acc[i][1][j] = 0.f; }
    const float4 *S4 = reinterpret_cast<const float4 *>(reinterpret_cast<const float *>(a.src) + ws_row(ca, g, 0) * CW);
    const int q4 = CW / 4, ne = H * q4;
    for (int h = 0; h < 2; h++) {
        const int k0 = h * H;
        __syncthreads();                                 // the previous half's operand reads are done
        // stage w_k S_k of rows k0 .. k0 + H - 1, 8 loads in flight per thread
        for (int e0 = 0; e0 < ne; e0 += 256 * 8) {
            float4 v[8];
            double w[8];
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int e = min(e0 + tid + 256 * i, ne - 1);
                v[i] = S4[(long)k0 * q4 + e];
                w[i] = a.rt.w[k0 + e / q4];
            }
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int e = e0 + tid + 256 * i;
                if (e < ne) {
                    const int kk = e / q4, c = 4 * (e - kk * q4);
                    const float4 u = make_float4((float)(w[i] * v[i].x), (float)(w[i] * v[i].y), (float)(w[i] * v[i].z),
                                                 (float)(w[i] * v[i].w));
                    *reinterpret_cast<float4 *>(Sl + kk * TABF_COLS + ((c + ((kk >> 2) & 1) * 32) & (TABF_COLS - 1))) = u;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < TABF2_TPW; i++) {
            const int t = wave + 4 * i;
            if (t >= nt) break;
            const int lr = 32 * t + cl;                   // this lane's A row (degree m0 + lr)
            const float4 *t4 = reinterpret_cast<const float4 *>(tab + (long)min(lr, M - 1) * N + k0) + kh;
            // rounds of 8 blocks (64 beta rows), the next round's 8 loads in flight
            float4 cur[8], nxt[8];
#pragma unroll
            for (int b = 0; b < 8; b++) cur[b] = t4[2 * min(b, nb - 1)];
            for (int j0 = 0; j0 < nb; j0 += 8) {
                if (j0 + 8 < nb) {
#pragma unroll
                    for (int b = 0; b < 8; b++) nxt[b] = t4[2 * min(j0 + 8 + b, nb - 1)];
                }
#pragma unroll
                for (int b = 0; b < 8; b++) {
                    if (j0 + b < nb) {
                        const float q[4] = {cur[b].x, cur[b].y, cur[b].z, cur[b].w};
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            const int row = (8 * (j0 + b) + u) * TABF_COLS;
                            acc[i][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(q[u], Sb0[row], acc[i][0], 0, 0, 0);
                            if (nct > 1) acc[i][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(q[u], Sb1[row], acc[i][1], 0, 0, 0);
                        }
                    }
                }
#pragma unroll
                for (int b = 0; b < 8; b++) cur[b] = nxt[b];
            }
        }
    }
    float *dst = reinterpret_cast<float *>(a.dst);
    const long es = a.pt ? 2L * C : 2L;
#pragma unroll
    for (int i = 0; i < TABF2_TPW; i++) {
        const int t = wave + 4 * i;
        if (t >= nt) break;
#pragma unroll
        for (int hc = 0; hc < 2; hc++) {
            const int c = hc ? c1 : c0;
            if (c >= CW || (hc && nct < 2)) continue;
            int fi, p, ri;
            ws_col(ca, NP, c, fi, p, ri);
            if (fi >= ca.nfc || !mx_pair_valid<NP>(m0, n0, p)) continue;
            const float sg = (float)mx_pair_sign<NP>(m0, n0, p);
            const MxPair<NP> pr(m0, n0, p);
            float *out = dst + (a.pt ? 2L * fi : 2L * ca.coef_off[ca.fn[fi]]) + ri;
#pragma unroll
            for (int j = 0; j < 16; j++) {
                const int r = 32 * t + 8 * (j >> 2) + 4 * kh + (j & 3);
                if (r < M) out[es * pr.packed(m0 + r)] = sg * acc[i][hc][j];
            }
        }
    }
}

#define TAB_CPITCH 65

template <int NP>
__global__ void __launch_bounds__(256) contract_inv_tab(ConArgs<float> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *Cb = reinterpret_cast<float *>(smem_raw);   // [M4][TAB_CPITCH]
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, C = ca.C, CW = C * NP * 2, NCOL = ca.nfc * NP * 2;
    const int unit = xcd_deal((a.ca.N + 127) / 128);
    if (unit >= a.nitems) return;
    const int2 it = a.items[unit];
    const int g = it.x, k0 = it.y;
    int m0, n0;
    mx_group<NP>(g, m0, n0);
    const int M = B - m0, M4 = (M + 3) & ~3;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kh = lane >> 5, cl = lane & 31;
    // stage the coefficient block: rows l - m0 < M4 (zero beyond M), columns c < 64 (zero beyond NCOL)
    const float2 *Fc = reinterpret_cast<const float2 *>(a.src);
    for (int e = tid; e < M4 * 64; e += 256) {
        const int lr = e >> 6, c = e & 63;
        float v = 0.f;
        int fi, p, ri;
        ws_col(ca, NP, c, fi, p, ri);
        if (lr < M && c < CW && fi < ca.nfc) {
            if (mx_pair_valid<NP>(m0, n0, p)) {
                const long idx = MxPair<NP>(m0, n0, p).packed(m0 + lr);
                const float2 f = a.pt ? Fc[idx * C + fi] : Fc[ca.coef_off[ca.fn[fi]] + idx];
                const float cf = NP == 2 ? 1.f : (float)(2 * (m0 + lr) + 1);   // SO(3): (2l+1)
                v = (float)mx_pair_sign<NP>(m0, n0, p) * cf * (ri ? f.y : f.x);
            }
        }
        Cb[lr * TAB_CPITCH + c] = v;
    }
    __syncthreads();
    const float *tab = reinterpret_cast<const float *>(a.dtab) + a.dtab_off[g];
    const int k = k0 + 32 * wave + cl;           // this lane's A row (beta index)
    const int kc = k < N ? k : N - 1;            // clamped: rows k >= N are computed, never stored
    const int nct = (ca.hsplit ? CW : NCOL) > 32 ? 2 : 1;   // column tiles holding functions of the chunk
    f32x16m_t acc0, acc1;
#pragma unroll
    for (int j = 0; j < 16; j++) { acc0[j] = 0.f; acc1[j] = 0.f; }
    // 16 steps per batch, the next batch's table loads in flight during this batch's MFMAs.
    // Degree rows past M read row M - 1 (the address is clamped, not the value) against the zero
    // rows of the coefficient block.
    constexpr int SB = 16;
    const int nst = M4 / 2;
    const float *trow = tab + kc + (long)kh * N;
    auto load_batch = [&](float *av, int s0) {
#pragma unroll
        for (int u = 0; u < SB; u++) av[u] = trow[(long)min(2 * (s0 + u), M - 1 - kh) * N];
    };
    auto mult_batch = [&](const float *av, int s0) {
#pragma unroll
        for (int u = 0; u < SB; u++) {
            if (s0 + u < nst) {
                const float *cb = Cb + (2 * (s0 + u) + kh) * TAB_CPITCH + cl;
                acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(av[u], cb[0], acc0, 0, 0, 0);
                if (nct > 1) acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(av[u], cb[32], acc1, 0, 0, 0);
            }
        }
    };
    float aA[SB], aB[SB];
    load_batch(aA, 0);
    for (int s0 = 0; s0 < nst; s0 += 2 * SB) {
        if (s0 + SB < nst) load_batch(aB, s0 + SB);
        mult_batch(aA, s0);
        if (s0 + SB >= nst) break;
        if (s0 + 2 * SB < nst) load_batch(aA, s0 + 2 * SB);
        mult_batch(aB, s0 + SB);
    }
    float *G = reinterpret_cast<float *>(a.dst);
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int kk = k0 + 32 * wave + 8 * (j >> 2) + 4 * kh + (j & 3);
        if (kk < N) {
            float *row = G + ws_row(ca, g, kk) * CW;
            if (cl < CW) row[cl] = acc0[j];
            if (nct > 1 && 32 + cl < CW) row[32 + cl] = acc1[j];
        }
    }
}
