// mfma_split.hpp -- batch-wide beta contractions in split-BF16 on MFMA (v_mfma_f32_32x32x16_bf16),
// N <= 256 (SO(3) and S^2 up to l_max = 127): every product formed from bf16 pieces of both
// operands and accumulated in FP32 (BF16X3: 2 pieces, 3 products; BF16X6: 3 pieces, 6 products;
// see mfma.hpp).  The Wigner-d values come from the FP64 recursion, split once per work item into
// pieces in LDS; each lane loads its other operand -- the spectrum (forward) or the coefficients
// (inverse) -- straight into registers one step ahead and splits it there, so the block loop has
// no barrier.  A first version of the N <= 128 kernels staged the spectrum blocks through LDS
// (l_max = 63 x 256 bf16x6: 31.9k transforms/s; these give 34-35k).
#pragma once

#include "mfma.hpp"
#include "mfma32.hpp"

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

// ---------------------------------------------------------------------------------------------
// The FP32 K-chunked kernels' scheme (contract_fwd_mxk / contract_inv_mxk) with bf16 pieces.  A
// group's whole d block in pieces (6 bytes per value in bf16x6) does not fit in LDS at B = 128, so
// each work item covers part of it: the forward a range of up to 64 degrees (the
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
