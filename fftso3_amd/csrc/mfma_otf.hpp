// mfma_otf.hpp -- split-BF16 beta contractions of complex SO(3) at N = 2B = 512 (l_max = 255) with
// the Wigner-d values regenerated on the fly by the FP64 recursion, for chunks of 8 functions whose
// workspace rows are half records (fft_lines.hpp): 64 real columns per (group, beta) row, and whose
// coefficients are staged in PT[e][function] (8 complex values per row).
//
// One workgroup of 512 threads per (m, n) group g runs the group's recursion once for all 512 beta
// values and all its degrees, at FP64, with the group's recursion coefficients in LDS
// (group_coefs), and forms every product from bf16 pieces on v_mfma_f32_32x32x16_bf16 with FP32
// accumulation (BF16X3 / BF16X6 as in mfma.hpp).  This replaces the chunked split kernels at this
// size, which restarted the recursion per (degree tile, beta chunk) on a quarter of their lanes.
//
// Inverse (G_k = sum_l (2l+1) F^l d^l(beta_k)): wave w owns the beta rows 64w .. 64w + 63 (two
// 32-row tiles) and all 64 columns.  Lane (cl, kh) runs the chain of beta index 64w + 32kh + cl;
// the MFMA's A operand of row tile t wants 8 consecutive degrees of chain 64w + 32t + cl in lane
// half kh = (degree / 8) % 2, so for every 16 degrees a lane keeps the octet of its own chain that
// its half needs and swaps the other octet with lane cl + 32 (one cross-half shuffle per value).
// The coefficient block, scaled by (2l+1) and the pair sign, is split into pieces in LDS 64
// degrees at a time, the next 64 degrees loading into registers meanwhile.  No LDS holds d.
//
// Forward (F^l = s_p sum_k w_k d^l(beta_k) S_k): the reduction runs over beta, which the chains
// own, so d goes through LDS: per 32 degrees, 256 threads run two chains each (adjacent beta
// values, stored as one 4-byte word per piece) and write w_k d split into pieces, rows of degree,
// beta contiguous.  Wave w takes column tile w % 2 and the beta quarter w / 2, keeps that quarter
// of the group's spectrum block split into pieces in registers for all the group's degrees, and
// the four quarters' partial tiles are summed through LDS and written to PT as whole 64-byte row
// pieces.
//
// Operand layouts of v_mfma_f32_32x32x16_bf16 (lane i): A[i%32][8(i/32)+j], B[8(i/32)+j][i%32]
// for j < 8; result D[8(j/4) + 4(i/32) + j%4][i%32] for j < 16.
#pragma once

#include "mfma.hpp"
#include "mfma32.hpp"

#define OTF_N 512
#define OTF_COLS 64           // real columns of a workspace row: 8 functions x 4 pairs x (re, im)
#define OTF_THREADS 512
#define OTF_LC 64             // degrees per staged coefficient chunk (inverse)
#define OTF_CP (OTF_LC + 8)   // bf16 pitch of a staged coefficient column (inverse)
#define OTF_DL 32             // degrees per d chunk (forward)
#define OTF_DP (OTF_N + 8)    // bf16 pitch of a d row (forward)

__host__ __device__ constexpr size_t otf_inv_lds_bytes(int npc, int B)
{
    return (size_t)npc * OTF_COLS * OTF_CP * 2 + (size_t)B * sizeof(RecCoef);
}
__host__ __device__ constexpr size_t otf_fwd_lds_bytes(int npc, int B)
{
    return (size_t)npc * OTF_DL * OTF_DP * 2 + (size_t)8 * 16 * 64 * 4 + (size_t)B * sizeof(RecCoef);
}

// pieces of 8 FP32 values (hardware round-to-nearest-even conversion; each residual is exact)
template <int NPC>
__device__ __forceinline__ void otf_split8(const float *x, bf16x8_t *out)
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

// seed d^{m0}_{m0 n0}(beta_k) of group g (as group_recursion_lds)
__device__ __forceinline__ double otf_seed(const RecTables &t, int g, int m0, int n0, int k)
{
    const int a = m0 + n0, b = m0 - n0;
    const double s = exp(t.lnbin[g] + (a ? a * t.lnc[k] : 0.0) + (b ? b * t.lns[k] : 0.0));
    return (b & 1) ? -s : s;
}

template <int NPC>
__global__ void __launch_bounds__(OTF_THREADS) contract_inv_otf(ConArgs<float> a)
{
    typedef SplitProducts<NPC> SP;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    unsigned short *Cs = reinterpret_cast<unsigned short *>(smem_raw);   // [NPC][64 columns][OTF_CP]
    RecCoef *rc = reinterpret_cast<RecCoef *>(smem_raw + (size_t)NPC * OTF_COLS * OTF_CP * 2);
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, C = ca.C;
    const int g = xcd_deal(1);
    if (g >= ca.ngrp) return;
    int m0, n0;
    so3_group(g, m0, n0);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, cl = lane & 31, kh = lane >> 5;
    // staging role: degree slr of the chunk, pair sp, functions 4 sh .. 4 sh + 3 (two float4 of PT)
    const int slr = tid >> 3, sp = (tid >> 1) & 3, sh = tid & 1;
    const bool spv = mx_pair_valid<4>(m0, n0, sp);
    const float ssg = (float)mx_pair_sign<4>(m0, n0, sp);
    const MxPair<4> spr(m0, n0, sp);
    const float4 *PT4 = reinterpret_cast<const float4 *>(a.src);
    float4 st0, st1;
    auto load_chunk = [&](int lb) {
        const int l = lb + slr;
        st0 = st1 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (spv && l < B) {
            const long base = ((long)spr.packed(l) * C + 4 * sh) >> 1;   // C = 8: float4 aligned
            st0 = PT4[base];
            st1 = PT4[base + 1];
        }
    };
    auto store_chunk = [&](int lb) {
        const float cf = ssg * (float)(2 * (lb + slr) + 1);
        const float v[8] = {st0.x, st0.y, st0.z, st0.w, st1.x, st1.y, st1.z, st1.w};
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int fi = 4 * sh + (q >> 1), ri = q & 1;
            const int col = 32 * (sp >> 1) + 4 * fi + 2 * (sp & 1) + ri;
            float r = fi < ca.nfc ? cf * v[q] : 0.f;
#pragma unroll
            for (int c = 0; c < NPC; c++) {
                const __bf16 b = (__bf16)r;
                Cs[(c * OTF_COLS + col) * OTF_CP + slr] = __builtin_bit_cast(unsigned short, b);
                r -= (float)b;
            }
        }
    };
    // this lane's chain: beta index k
    const int k = 64 * wave + 32 * kh + cl;
    const double cb = a.rt.cb[k];
    double cur = otf_seed(a.rt, g, m0, n0, k), prev = 0.0;
    f32x16_t acc[2][2];
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
            for (int j = 0; j < 16; j++) acc[t][u][j] = 0.f;
    group_coefs(a.rt, 0, B, m0, n0, rc, tid, OTF_THREADS);
    load_chunk(m0);
    for (int lb = m0; lb < B; lb += OTF_LC) {
        __syncthreads();   // recursion coefficients written / the previous chunk's operand reads done
        store_chunk(lb);
        __syncthreads();
        if (lb + OTF_LC < B) load_chunk(lb + OTF_LC);
#pragma unroll 1
        for (int s = 0; s < OTF_LC / 16; s++) {
            const int l0 = lb + 16 * s;
            if (l0 >= B) break;
            float dv[16];
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int l = l0 + i;
                double v = 0.0;
                if (l < B) {
                    if (l > m0) {
                        const RecCoef r = rc[l];
                        const double nv = (r.a * cb - r.b) * cur - r.c * prev;
                        prev = cur;
                        cur = nv;
                    }
                    v = cur;
                }
                dv[i] = (float)v;
            }
            float A0[8], A1[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const float snd = kh ? dv[j] : dv[8 + j];
                const float rcv = __shfl_xor(snd, 32);
                A0[j] = kh ? rcv : dv[j];
                A1[j] = kh ? dv[8 + j] : rcv;
            }
            bf16x8_t a0[NPC], a1[NPC], b0[NPC], b1[NPC];
            otf_split8<NPC>(A0, a0);
            otf_split8<NPC>(A1, a1);
#pragma unroll
            for (int c = 0; c < NPC; c++) {
                b0[c] = lds_bf16x8(Cs + (c * OTF_COLS + cl) * OTF_CP + 16 * s + 8 * kh);
                b1[c] = lds_bf16x8(Cs + (c * OTF_COLS + 32 + cl) * OTF_CP + 16 * s + 8 * kh);
            }
#pragma unroll
            for (int i = 0; i < SP::n; i++) {
                acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0[SP::a(i)], b0[SP::b(i)], acc[0][0], 0, 0, 0);
                acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0[SP::a(i)], b1[SP::b(i)], acc[0][1], 0, 0, 0);
                acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1[SP::a(i)], b0[SP::b(i)], acc[1][0], 0, 0, 0);
                acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1[SP::a(i)], b1[SP::b(i)], acc[1][1], 0, 0, 0);
            }
        }
    }
    float *G = reinterpret_cast<float *>(a.dst);
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int kk = 64 * wave + 32 * t + 8 * (j >> 2) + 4 * kh + (j & 3);
            float *row = G + ws_row(ca, g, kk) * OTF_COLS;
            row[cl] = acc[t][0][j];
            row[32 + cl] = acc[t][1][j];
        }
}

template <int NPC>
__global__ void __launch_bounds__(OTF_THREADS) contract_fwd_otf(ConArgs<float> a)
{
    typedef SplitProducts<NPC> SP;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    unsigned short *Dp = reinterpret_cast<unsigned short *>(smem_raw);                        // [NPC][32][OTF_DP]
    float *R = reinterpret_cast<float *>(smem_raw + (size_t)NPC * OTF_DL * OTF_DP * 2);      // [wave][16][64]
    RecCoef *rc = reinterpret_cast<RecCoef *>(R + 8 * 16 * 64);
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, C = ca.C;
    const int g = xcd_deal(1);
    if (g >= ca.ngrp) return;
    int m0, n0;
    so3_group(g, m0, n0);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, cl = lane & 31, kh = lane >> 5;
    const int ct = wave & 1, kq = wave >> 1;
    // the wave's quarter of the spectrum block as the B operand, split once: step u, value j is
    // beta 128 kq + 16 u + 8 kh + j, column 32 ct + cl
    bf16x8_t sb[8][NPC];
    {
        const float *W = reinterpret_cast<const float *>(a.src) + 32 * ct + cl;
        float x[8][8];
#pragma unroll
        for (int u = 0; u < 8; u++)
#pragma unroll
            for (int j = 0; j < 8; j++) x[u][j] = W[ws_row(ca, g, 128 * kq + 16 * u + 8 * kh + j) * OTF_COLS];
#pragma unroll
        for (int u = 0; u < 8; u++) otf_split8<NPC>(x[u], sb[u]);
    }
    // recursion role: threads 0..255 run the chains of beta 2t and 2t + 1
    const bool rrole = tid < 256;
    const int kr = 2 * (tid & 255);
    const double cbA = a.rt.cb[kr], cbB = a.rt.cb[kr + 1], wA = a.rt.w[kr], wB = a.rt.w[kr + 1];
    double curA = 0.0, curB = 0.0, prvA = 0.0, prvB = 0.0;
    if (rrole) {
        curA = otf_seed(a.rt, g, m0, n0, kr);
        curB = otf_seed(a.rt, g, m0, n0, kr + 1);
    }
    group_coefs(a.rt, 0, B, m0, n0, rc, tid, OTF_THREADS);
    // reduction role: degree rlr of the chunk, function rfi, re/im rri; pairs 0..3 in turn
    const int rlr = tid >> 4, rfi = (tid >> 1) & 7, rri = tid & 1;
    const int rj = 4 * (rlr >> 3) + (rlr & 3), rkh = (rlr >> 2) & 1;
    float *PT = reinterpret_cast<float *>(a.dst);
    for (int lb = m0; lb < B; lb += OTF_DL) {
        __syncthreads();   // recursion coefficients written / the previous chunk's partial sums read
        if (rrole) {
            unsigned *Dw = reinterpret_cast<unsigned *>(Dp) + (kr >> 1);
#pragma unroll 4
            for (int i = 0; i < OTF_DL; i++) {
                const int l = lb + i;
                double vA = 0.0, vB = 0.0;
                if (l < B) {
                    if (l > m0) {
                        const RecCoef r = rc[l];
                        const double nA = (r.a * cbA - r.b) * curA - r.c * prvA;
                        const double nB = (r.a * cbB - r.b) * curB - r.c * prvB;
                        prvA = curA;
                        curA = nA;
                        prvB = curB;
                        curB = nB;
                    }
                    vA = wA * curA;
                    vB = wB * curB;
                }
                float fA = (float)vA, fB = (float)vB;
#pragma unroll
                for (int c = 0; c < NPC; c++) {
                    const __bf16 pA = (__bf16)fA, pB = (__bf16)fB;
                    Dw[(c * OTF_DL + i) * (OTF_DP / 2)] = (unsigned)__builtin_bit_cast(unsigned short, pA) |
                                                          ((unsigned)__builtin_bit_cast(unsigned short, pB) << 16);
                    fA -= (float)pA;
                    fB -= (float)pB;
                }
            }
        }
        __syncthreads();
        f32x16_t acc;
#pragma unroll
        for (int j = 0; j < 16; j++) acc[j] = 0.f;
        const unsigned short *Da = Dp + cl * OTF_DP + 128 * kq + 8 * kh;
#pragma unroll
        for (int u = 0; u < 8; u++) {
            bf16x8_t av[NPC];
#pragma unroll
            for (int c = 0; c < NPC; c++) av[c] = lds_bf16x8(Da + c * OTF_DL * OTF_DP + 16 * u);
#pragma unroll
            for (int i = 0; i < SP::n; i++)
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[SP::a(i)], sb[u][SP::b(i)], acc, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 16; j++) R[(wave * 16 + j) * 64 + lane] = acc[j];
        __syncthreads();
        const int l = lb + rlr;
        if (l < B && rfi < ca.nfc) {
#pragma unroll
            for (int p = 0; p < 4; p++) {
                if (!mx_pair_valid<4>(m0, n0, p)) continue;
                const int pct = p >> 1, pcl = 4 * rfi + 2 * (p & 1) + rri;
                const float *Rp = R + rj * 64 + 32 * rkh + pcl;
                float s = 0.f;
#pragma unroll
                for (int q = 0; q < 4; q++) s += Rp[(2 * q + pct) * 16 * 64];
                PT[((long)MxPair<4>(m0, n0, p).packed(l) * C + rfi) * 2 + rri] = (float)mx_pair_sign<4>(m0, n0, p) * s;
            }
        }
    }
}
