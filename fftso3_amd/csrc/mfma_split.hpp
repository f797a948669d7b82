// mfma_split.hpp -- batch-wide beta contractions in split-BF16 on MFMA (v_mfma_f32_32x32x16_bf16):
// every product formed from bf16 pieces of both operands and accumulated in FP32 (BF16X3: 2 pieces,
// 3 products; BF16X6: 3 pieces, 6 products; see mfma.hpp).  The Wigner-d values come from the FP64
// recursion, split once per work item into pieces in LDS.
//
// contract_fwd_mxs_body / contract_inv_mxs_body (N <= 128) are the units of the unified schedule
// of uneven plans (multi.hpp): the work decomposition, workspace layout and coefficient staging
// (PT) of the FP32 kernels contract_fwd_mx32 / contract_inv_mx32 (one workgroup per (m,n) group
// for the whole batch, the batch in blocks of 128 spectrum columns).  The reduction index of one
// instruction is 16 values: lane i supplies 8 of them for row (A) or column (B) i % 32; the beta
// (forward) or degree (inverse) index r of a lane's j-th value in instruction t is
// r = 2(8t + j) + i/32, the pairing the FP32 kernels use for their 2-deep instructions, so both
// operands read the same LDS rows as there.  The pieces of D are stored in instruction order,
// [piece][row][t][i/32][j], so a lane reads its 8 values as 16 bytes.
//
// contract_fwd_mxks / contract_inv_mxks (N <= 256, below) are the launches of every batch-wide
// split plan of its own: each lane loads its spectrum (forward) or coefficient (inverse) operands
// straight into registers, one step ahead, and the block loop has no barrier.
#pragma once

#include "mfma.hpp"
#include "mfma32.hpp"

#define MXS_FPITCH 136   // bf16 per D row in the forward (8 instructions x 16 + 8): conflict-free 16-byte reads
#define MXS_IPITCH 72    // bf16 per beta row in the inverse (4 instructions x 16 + 8)

__host__ __device__ constexpr size_t mxs_fwd_lds_bytes(int npc)
{
    return (size_t)128 * MX_LPITCH * 4 + (size_t)npc * 64 * MXS_FPITCH * 2;
}
__host__ __device__ constexpr size_t mxs_inv_lds_bytes(int npc)
{
    return (size_t)128 * MX_LPITCH * 4 + (size_t)npc * 128 * MXS_IPITCH * 2;
}

// pieces of 8 consecutive FP32 values as bf16x8 vectors (the rounding of split_f; the hardware
// conversion forms two pieces per instruction)
template <int NPC>
__device__ __forceinline__ void split8(const float *x, bf16x8_t *out)
{
    float r[8];
#pragma unroll
    for (int j = 0; j < 8; j++) r[j] = x[j];
#pragma unroll
    for (int c = 0; c < NPC; c++) {
        bf16x8_t v;
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = (__bf16)r[j];
        out[c] = v;
        if (c + 1 < NPC) {
#pragma unroll
            for (int j = 0; j < 8; j++) r[j] -= (float)v[j];
        }
    }
}

// Forward: PT(packed(l, pair); fn) = s_p sum_k w_k d^l_{m0 n0}(beta_k) S_k(fn, pair), split-BF16.
template <int NP, int NPC>
__device__ __forceinline__ void contract_fwd_mxs_body(const MxArgs &a, int unit)
{
    typedef SplitProducts<NPC> SP;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *L = reinterpret_cast<float *>(smem_raw);                           // [128][MX_LPITCH]
    unsigned short *Dp = reinterpret_cast<unsigned short *>(smem_raw + (size_t)128 * MX_LPITCH * 4);
    RecCoef *rc = reinterpret_cast<RecCoef *>(smem_raw + mxs_fwd_lds_bytes(NPC));
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, C = ca.C, CW = C * NP * 2;
    const int nfb = (ca.nfc * NP * 2 + 127) / 128;
    int g, fb0, fb1;
    if (!mx_unit(a, nfb, unit, g, fb0, fb1)) return;
    int m0, n0;
    mx_group<NP>(g, m0, n0);
    const int M = B - m0, nrt = (M + 31) / 32, nt = (N + 15) / 16;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kh = lane >> 5, cl = lane & 31;
    constexpr int FB = 64 / NP;
    // load role (as contract_fwd_mx32): rows r0 + 8i of the block, columns 4q..4q+3
    const int q = tid & 31, r0 = tid >> 5;
    const int fq = (2 * q) / NP;
    const int lc0 = mx_lcol<NP>(4 * q), lc1 = mx_lcol<NP>(4 * q + 2);
    const float *S = reinterpret_cast<const float *>(a.src);
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
    // split D rows: degree row r, beta k = 2(8t + j) + kh at Dp[c][r][t][kh][j]
    auto dp_at = [&](int c, int r, int k) -> unsigned short & {
        return Dp[(c * 64 + r) * MXS_FPITCH + (k >> 4) * 16 + (k & 1) * 8 + ((k >> 1) & 7)];
    };
    group_coefs(a.rt, NP == 2, B, m0, n0, rc, tid, 256);
    __syncthreads();
    for (int k = tid; k < 16 * nt; k += 256) {
        if (k < N) {
            const double wk = a.rt.w[k];
            group_recursion_lds(a.rt, rc, NP == 2, g, m0, n0, k, B, [&](int l, double v) {
                unsigned short p[NPC];
                split_d<NPC>(wk * v, p);
#pragma unroll
                for (int c = 0; c < NPC; c++) dp_at(c, l - m0, k) = p[c];
            });
            for (int r = M; r < 32 * nrt; r++)
#pragma unroll
                for (int c = 0; c < NPC; c++) dp_at(c, r, k) = 0;
        } else {
            for (int r = 0; r < 32 * nrt; r++)
#pragma unroll
                for (int c = 0; c < NPC; c++) dp_at(c, r, k) = 0;
        }
    }
    const int p = __builtin_amdgcn_readfirstlane(wave) % NP, fo = (wave / NP) * 16 + (cl >> 1);
    const float sg = (float)mx_pair_sign<NP>(m0, n0, p);
    const bool pvalid = mx_pair_valid<NP>(m0, n0, p);
    const MxPair<NP> pr(m0, n0, p);
    float *PT = reinterpret_cast<float *>(a.dst);
    const float *Lw = L + kh * MX_LPITCH + wave * 32 + cl;
    const unsigned short *Dl = Dp + cl * MXS_FPITCH + kh * 8;
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
        f32x16_t acc0, acc1;
#pragma unroll
        for (int j = 0; j < 16; j++) { acc0[j] = 0.f; acc1[j] = 0.f; }
        for (int t = 0; t < nt; t++) {
            float b[8];
#pragma unroll
            for (int j = 0; j < 8; j++) b[j] = Lw[2 * (8 * t + j) * MX_LPITCH];
            bf16x8_t bp[NPC], a0[NPC], a1[NPC];
            split8<NPC>(b, bp);
#pragma unroll
            for (int c = 0; c < NPC; c++) a0[c] = lds_bf16x8(Dl + c * 64 * MXS_FPITCH + t * 16);
#pragma unroll
            for (int i = 0; i < SP::n; i++)
                acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0[SP::a(i)], bp[SP::b(i)], acc0, 0, 0, 0);
            if (nrt > 1) {
#pragma unroll
                for (int c = 0; c < NPC; c++) a1[c] = lds_bf16x8(Dl + (c * 64 + 32) * MXS_FPITCH + t * 16);
#pragma unroll
                for (int i = 0; i < SP::n; i++)
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1[SP::a(i)], bp[SP::b(i)], acc1, 0, 0, 0);
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


// Inverse: G_k(fn, pair) = s_p sum_l c_l d^l_{m0 n0}(beta_k) PT(packed(l, pair); fn), split-BF16
// (c_l = 2l+1 for SO(3), 1 for S^2).
template <int NP, int NPC>
__device__ __forceinline__ void contract_inv_mxs_body(const MxArgs &a, int unit)
{
    typedef SplitProducts<NPC> SP;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *O = reinterpret_cast<float *>(smem_raw);                           // [128][MX_LPITCH] output block
    unsigned short *Dq = reinterpret_cast<unsigned short *>(smem_raw + (size_t)128 * MX_LPITCH * 4);
    RecCoef *rc = reinterpret_cast<RecCoef *>(smem_raw + mxs_inv_lds_bytes(NPC));
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, C = ca.C, CW = C * NP * 2;
    const int nfb = (ca.nfc * NP * 2 + 127) / 128;
    int g, fb0, fb1;
    if (!mx_unit(a, nfb, unit, g, fb0, fb1)) return;
    int m0, n0;
    mx_group<NP>(g, m0, n0);
    const int M = B - m0, nu = (M + 15) / 16, nkt = (N + 31) / 32;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kh = lane >> 5, cl = lane & 31;
    constexpr int FB = 64 / NP;
    const int p = __builtin_amdgcn_readfirstlane(wave) % NP, fo = (wave / NP) * 16 + (cl >> 1);
    const float *PT = reinterpret_cast<const float *>(a.src);
    const MxPair<NP> pr(m0, n0, p);
    // this lane's degrees l = 2s + kh, s < 32 (rows past M - 1 read row M - 1: finite values that
    // meet zero pieces of D)
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
    // split D^T: beta row k, degree row r = 2(8u + j) + kh at Dq[c][k][u][kh][j]
    auto dq_at = [&](int c, int k, int r) -> unsigned short & {
        return Dq[(c * 128 + k) * MXS_IPITCH + (r >> 4) * 16 + (r & 1) * 8 + ((r >> 1) & 7)];
    };
    group_coefs(a.rt, NP == 2, B, m0, n0, rc, tid, 256);
    __syncthreads();
    for (int k = tid; k < 32 * nkt; k += 256) {
        if (k < N) {
            group_recursion_lds(a.rt, rc, NP == 2, g, m0, n0, k, B, [&](int l, double v) {
                unsigned short pc[NPC];
                split_d<NPC>(NP == 2 ? v : (double)(2 * l + 1) * v, pc);
#pragma unroll
                for (int c = 0; c < NPC; c++) dq_at(c, k, l - m0) = pc[c];
            });
            for (int r = M; r < 16 * nu; r++)
#pragma unroll
                for (int c = 0; c < NPC; c++) dq_at(c, k, r) = 0;
        } else {
            for (int r = 0; r < 16 * nu; r++)
#pragma unroll
                for (int c = 0; c < NPC; c++) dq_at(c, k, r) = 0;
        }
    }
    const float sg = (float)mx_pair_sign<NP>(m0, n0, p);
    // store role: rows r0 + 8i of the block, columns 4q..4q+3 (as contract_inv_mx32)
    const int q = tid & 31, r0 = tid >> 5;
    const int fq = (2 * q) / NP;
    const int lc0 = mx_lcol<NP>(4 * q), lc1 = mx_lcol<NP>(4 * q + 2);
    float *G = reinterpret_cast<float *>(a.dst);
    const int gm = (1 << ca.gil) - 1;
    const int qoff = (g & gm) * NP * 2 + (fq << ca.gil) * NP * 2 + (4 * q - fq * NP * 2);
    long roff[16];
#pragma unroll
    for (int i = 0; i < 16; i++) roff[i] = r0 + 8 * i < N ? ws_row(ca, g & ~gm, r0 + 8 * i) * CW + qoff : -1;
    const unsigned short *Dl = Dq + cl * MXS_IPITCH + kh * 8;
    auto multiply_store = [&](const float *buf, int fb) {
        f32x16_t acc[4];
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int j = 0; j < 16; j++) acc[t][j] = 0.f;
        for (int u = 0; u < nu; u++) {
            bf16x8_t bp[NPC];
            split8<NPC>(buf + 8 * u, bp);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                if (t < nkt) {
                    bf16x8_t ap[NPC];
#pragma unroll
                    for (int c = 0; c < NPC; c++) ap[c] = lds_bf16x8(Dl + (c * 128 + 32 * t) * MXS_IPITCH + u * 16);
#pragma unroll
                    for (int i = 0; i < SP::n; i++)
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[SP::a(i)], bp[SP::b(i)], acc[t], 0, 0, 0);
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


// ---------------------------------------------------------------------------------------------
// Batch-wide split-BF16 contractions for N <= 256 (S^2 up to l_max = 127, SO(3) up to l_max = 127;
// the unified schedule of uneven plans keeps contract_fwd_mxs / contract_inv_mxs above): the FP32
// K-chunked kernels' scheme (contract_fwd_mxk / contract_inv_mxk) with bf16 pieces.  A group's whole d block in pieces (6 bytes per value in bf16x6) does not fit in LDS
// at B = 128, so each work item covers part of it: the forward a range of up to 64 degrees (the
// item's spectrum blocks are read once per range), the inverse a range of up to 128 beta values
// (its coefficient blocks are read once per range).  Within an item the reduction index of one
// instruction is 16 consecutive values, lane half kh taking values 8 kh .. 8 kh + 7.
// ---------------------------------------------------------------------------------------------
#define MXKS_LR 64     // degrees per forward item
#define MXKS_KR 128    // beta values per inverse item

__host__ __device__ inline int mxks_dpitch(int N) { return ((N + 31) & ~31) + 8; }
__host__ __device__ inline size_t mxks_fwd_lds_bytes(int N, int npc, int NP)
{
    return (size_t)npc * MXKS_LR * mxks_dpitch(N) * 2 + (size_t)NP * MXK_MAXROWS * 4;
}
// inverse: D^T pieces [npc][128][M16 + 8] for the largest group (M16 = B rounded up to 16)
__host__ __device__ inline int mxks_inv_pitch(int B) { return ((B + 15) & ~15) + 8; }
__host__ __device__ inline size_t mxks_inv_lds_bytes(int npc, int NP, int B)
{
    return (size_t)npc * MXKS_KR * mxks_inv_pitch(B) * 2 + (size_t)NP * MXK_MAXROWS * 4;
}
// forward items 2 g + h (degrees m0 + 64 h ..), inverse items g * ceil(N / 128) + kr (beta 128 kr ..)

// offset within a workspace row of block column c = (function, pair, re/im) of group g: with group
// interleave (ws_rec) the records of 2^gil consecutive groups sit side by side per function
template <int NP>
__device__ __forceinline__ long mxks_col(const ChunkArgs &ca, int g, int c)
{
    const int fn = c / (2 * NP), rem = c - fn * 2 * NP;
    return (long)((fn << ca.gil) + (g & ((1 << ca.gil) - 1))) * (2 * NP) + rem;
}

// seed d^{m0}_{m0 n0}(beta_k) (S^2: nP^{m0}_{m0}) of group g, as group_recursion_lds
__device__ __forceinline__ double mxks_seed(const RecTables &t, int s2, int g, int m0, int n0, int k)
{
    if (s2) {
        const double s = exp(t.lnbin[g] + (m0 ? m0 * t.lns[k] : 0.0));
        return (m0 & 1) ? -s : s;
    }
    const int a = m0 + n0, b = m0 - n0;
    const double s = exp(t.lnbin[g] + (a ? a * t.lnc[k] : 0.0) + (b ? b * t.lns[k] : 0.0));
    return (b & 1) ? -s : s;
}

template <int NP, int NPC, int NW>
__device__ __forceinline__ void contract_fwd_mxks_body(const MxArgs &a, int unit)
{
    typedef SplitProducts<NPC> SP;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, C = ca.C, CW = C * NP * 2, NCOL = ca.nfc * NP * 2;
    const int nfb = (NCOL + 32 * NW - 1) / (32 * NW);
    int item, fb0, fb1;
    if (!mx_unit(a, nfb, unit, item, fb0, fb1)) return;
    const int g = item >> 1, rb = MXKS_LR * (item & 1);
    int m0, n0;
    mx_group<NP>(g, m0, n0);
    const int M = B - m0, LR = min(MXKS_LR, M - rb), nrt = (LR + 31) / 32;
    const int N32 = (N + 31) & ~31, DPK = N32 + 8, nst = N32 / 16;
    unsigned short *Dp = reinterpret_cast<unsigned short *>(smem_raw);   // [NPC][64][DPK]: w_k d, beta contiguous
    int *etab = reinterpret_cast<int *>(Dp + (size_t)NPC * MXKS_LR * DPK); // [NP][MXK_MAXROWS]
    RecCoef *rc = reinterpret_cast<RecCoef *>(smem_raw + mxks_fwd_lds_bytes(N, NPC, NP));
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kh = lane >> 5, cl = lane & 31;
    group_coefs(a.rt, NP == 2, B, m0, n0, rc, tid, 64 * NW);
    __syncthreads();
    // the item's degree rows rb .. rb + LR - 1: thread k runs the chain of beta k from m0
    for (int k = tid; k < N32; k += 64 * NW) {
        if (k < N) {
            const double wk = a.rt.w[k];
            group_recursion_lds(a.rt, rc, NP == 2, g, m0, n0, k, m0 + rb + LR, [&](int l, double v) {
                const int r = l - m0 - rb;
                if (r >= 0) {
                    unsigned short pc[NPC];
                    split_d<NPC>(wk * v, pc);
#pragma unroll
                    for (int c = 0; c < NPC; c++) Dp[(c * MXKS_LR + r) * DPK + k] = pc[c];
                }
            });
            for (int r = LR; r < 32 * nrt; r++)
#pragma unroll
                for (int c = 0; c < NPC; c++) Dp[(c * MXKS_LR + r) * DPK + k] = 0;
        } else {
            for (int r = 0; r < 32 * nrt; r++)
#pragma unroll
                for (int c = 0; c < NPC; c++) Dp[(c * MXKS_LR + r) * DPK + k] = 0;
        }
    }
    for (int i = tid; i < NP * MXK_MAXROWS; i += 64 * NW) {
        const int p = i / MXK_MAXROWS, r = i - p * MXK_MAXROWS;
        etab[i] = MxPair<NP>(m0, n0, p).packed(m0 + min(r, M - 1));
    }
    __syncthreads();
    // block column 32 wave + cl -> (function, pair, re/im); the lane's spectrum octet of step s is
    // beta 16 s + 8 kh + j, loaded straight into registers one step ahead (rows past N - 1 read row
    // N - 1: finite values that meet zero columns of D)
    const int cb = 32 * wave + cl;
    const int p = (cb >> 1) % NP;
    const float sg = (float)mx_pair_sign<NP>(m0, n0, p);
    const bool pvalid = mx_pair_valid<NP>(m0, n0, p);
    const float *S = reinterpret_cast<const float *>(a.src);
    float *PT = reinterpret_cast<float *>(a.dst);
    const unsigned short *Da = Dp + cl * DPK + 8 * kh;
    const int gr = g & ~((1 << ca.gil) - 1);   // the workspace row group of g's interleaved records
    for (int fb = fb0; fb < fb1; fb++) {
        const int c = fb * 32 * NW + cb, cv = c < NCOL ? c : 0;
        const float *src = S + mxks_col<NP>(ca, g, cv);
        auto load = [&](float *x, int st) {
#pragma unroll
            for (int j = 0; j < 8; j++) x[j] = src[ws_row(ca, gr, min(16 * st + 8 * kh + j, N - 1)) * CW];
        };
        f32x16_t acc0, acc1;
#pragma unroll
        for (int j = 0; j < 16; j++) { acc0[j] = 0.f; acc1[j] = 0.f; }
        float x[8], xn[8];
        load(x, 0);
        for (int st = 0; st < nst; st++) {
            if (st + 1 < nst) load(xn, st + 1);
            bf16x8_t bv[NPC], av[NPC];
            split8<NPC>(x, bv);
#pragma unroll
            for (int cc = 0; cc < NPC; cc++) av[cc] = lds_bf16x8(Da + cc * MXKS_LR * DPK + 16 * st);
#pragma unroll
            for (int i = 0; i < SP::n; i++)
                acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[SP::a(i)], bv[SP::b(i)], acc0, 0, 0, 0);
            if (nrt > 1) {
#pragma unroll
                for (int cc = 0; cc < NPC; cc++) av[cc] = lds_bf16x8(Da + (cc * MXKS_LR + 32) * DPK + 16 * st);
#pragma unroll
                for (int i = 0; i < SP::n; i++)
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[SP::a(i)], bv[SP::b(i)], acc1, 0, 0, 0);
            }
#pragma unroll
            for (int j = 0; j < 8; j++) x[j] = xn[j];
        }
        const int fn = c / (2 * NP);
        const bool ok = pvalid && c < NCOL;
        float *o = PT + 2 * fn + (c & 1);
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int r = 8 * (j >> 2) + 4 * kh + (j & 3);
            float *d0 = ok && r < LR ? o + 2L * etab[p * MXK_MAXROWS + rb + r] * C : a.dummy;
            float *d1 = ok && r + 32 < LR ? o + 2L * etab[p * MXK_MAXROWS + rb + r + 32] * C : a.dummy;
            *d0 = sg * acc0[j];
            *d1 = sg * acc1[j];
        }
    }
}

template <int NP, int NPC>
__global__ void __launch_bounds__(512) contract_fwd_mxks(MxArgs a)
{
    contract_fwd_mxks_body<NP, NPC, 8>(a, mx_block_unit(a));
}

// Inverse: the item's beta rows k0 .. k0 + 127 of c_l d^l (c_l = 2l+1 for SO(3), 1 for S^2) in
// pieces in LDS, [piece][beta][degree] with degrees contiguous; wave w produces block columns
// [32w, 32w + 32) of a 256-column block for every beta row of the item, the block's coefficient
// octets loading one step ahead of their MFMAs.
template <int NP, int NPC, int NW>
__device__ __forceinline__ void contract_inv_mxks_body(const MxArgs &a, int unit)
{
    typedef SplitProducts<NPC> SP;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, C = ca.C, CW = C * NP * 2, NCOL = ca.nfc * NP * 2;
    const int nfb = (NCOL + 32 * NW - 1) / (32 * NW);
    int item, fb0, fb1;
    if (!mx_unit(a, nfb, unit, item, fb0, fb1)) return;
    const int nkr = (N + MXKS_KR - 1) / MXKS_KR;
    const int g = item / nkr, k0 = MXKS_KR * (item - (item / nkr) * nkr);
    int m0, n0;
    mx_group<NP>(g, m0, n0);
    const int M = B - m0, M16 = (M + 15) & ~15, LP = M16 + 8, nu = M16 / 16;
    const int kn = min(MXKS_KR, N - k0), nkt = (kn + 31) / 32;
    unsigned short *Dt = reinterpret_cast<unsigned short *>(smem_raw);                          // [NPC][128][LP]
    int *etab = reinterpret_cast<int *>(smem_raw + (size_t)NPC * MXKS_KR * mxks_inv_pitch(B) * 2); // [NP][MXK_MAXROWS]
    RecCoef *rc = reinterpret_cast<RecCoef *>(smem_raw + mxks_inv_lds_bytes(NPC, NP, B));
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int kh = lane >> 5, cl = lane & 31;
    for (int i = tid; i < NP * MXK_MAXROWS; i += 64 * NW) {
        const int p = i / MXK_MAXROWS, r = i - p * MXK_MAXROWS;
        etab[i] = MxPair<NP>(m0, n0, p).packed(m0 + min(r, M - 1));
    }
    group_coefs(a.rt, NP == 2, B, m0, n0, rc, tid, 64 * NW);
    __syncthreads();
    if (tid < MXKS_KR) {
        // the chain of beta k0 + tid, eight degrees per store; zero past the last degree and for
        // beta values past N - 1
        const int k = k0 + tid;
        const bool live = k < N;
        const double cb = live ? a.rt.cb[k] : 0.0;
        double d1 = live ? mxks_seed(a.rt, NP == 2, g, m0, n0, k) : 0.0, d2 = 0.0;
        for (int l8 = 0; l8 < M16; l8 += 8) {
            u16x8_t pcs[NPC];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int l = m0 + l8 + j;
                double v = 0.0;
                if (live && l < B) {
                    if (l > m0) {
                        const RecCoef r = rc[l];
                        const double nv = (r.a * cb - r.b) * d1 - r.c * d2;
                        d2 = d1;
                        d1 = nv;
                    }
                    v = NP == 2 ? d1 : (double)(2 * l + 1) * d1;
                }
                unsigned short pc[NPC];
                split_d<NPC>(v, pc);
#pragma unroll
                for (int c = 0; c < NPC; c++) pcs[c][j] = pc[c];
            }
#pragma unroll
            for (int c = 0; c < NPC; c++) *reinterpret_cast<u16x8_t *>(Dt + (c * MXKS_KR + tid) * LP + l8) = pcs[c];
        }
    }
    __syncthreads();
    const int cb = 32 * wave + cl, p = (cb >> 1) % NP;
    const float sg = (float)mx_pair_sign<NP>(m0, n0, p);
    const float *PT = reinterpret_cast<const float *>(a.src);
    float *G = reinterpret_cast<float *>(a.dst);
    const int *et = etab + p * MXK_MAXROWS + 8 * kh;
    const int gr = g & ~((1 << ca.gil) - 1);   // the workspace row group of g's interleaved records
    for (int fb = fb0; fb < fb1; fb++) {
        const int c = fb * 32 * NW + cb, cv = c < NCOL ? c : 0;
        const long gcol = mxks_col<NP>(ca, g, cv);
        const float *src = PT + 2 * (cv / (2 * NP)) + (cv & 1);
        f32x16_t acc[4];
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int j = 0; j < 16; j++) acc[t][j] = 0.f;
        // coefficient octet of step u: degrees 16 u + 8 kh + j (rows past M - 1 read row M - 1:
        // finite values that meet zero degree columns of D)
        float bx[8], bn[8];
#pragma unroll
        for (int j = 0; j < 8; j++) bx[j] = src[2L * et[j] * C];
        for (int u = 0; u < nu; u++) {
            if (u + 1 < nu) {
#pragma unroll
                for (int j = 0; j < 8; j++) bn[j] = src[2L * et[16 * (u + 1) + j] * C];
            }
            bf16x8_t bp[NPC];
            split8<NPC>(bx, bp);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                if (t < nkt) {
                    bf16x8_t ap[NPC];
#pragma unroll
                    for (int cc = 0; cc < NPC; cc++) ap[cc] = lds_bf16x8(Dt + (cc * MXKS_KR + 32 * t + cl) * LP + 16 * u + 8 * kh);
#pragma unroll
                    for (int i = 0; i < SP::n; i++)
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ap[SP::a(i)], bp[SP::b(i)], acc[t], 0, 0, 0);
                }
            }
#pragma unroll
            for (int j = 0; j < 8; j++) bx[j] = bn[j];
        }
        const bool ok = c < NCOL;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            if (t < nkt) {
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    int kk = 32 * t + 8 * (j >> 2) + 4 * kh + (j & 3);
                    // opaque here, so the 64 row addresses are formed at the stores instead of being
                    // hoisted above the degree loop and held in registers across it
                    asm volatile("" : "+v"(kk));
                    float *d = ok && kk < kn ? G + ws_row(ca, gr, k0 + kk) * CW + gcol : a.dummy;
                    *d = sg * acc[t][j];
                }
            }
        }
    }
}

template <int NP, int NPC>
__global__ void __launch_bounds__(512) contract_inv_mxks(MxArgs a)
{
    contract_inv_mxks_body<NP, NPC, 8>(a, mx_block_unit(a));
}
