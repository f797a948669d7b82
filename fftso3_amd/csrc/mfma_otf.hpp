// mfma_otf.hpp -- split-BF16 beta contractions of complex SO(3) at N = 2B = 512 (l_max = 255) with
// the Wigner-d values regenerated on the fly by the FP64 recursion, for chunks of 8 functions whose
// workspace rows are half records (fft_lines.hpp): 64 real columns per (group, beta) row, and whose
// coefficients are staged in PT[e][function] (8 complex values per row).
//
// One workgroup of 1024 threads per (m, n) group g runs the group's recursion once for all 512 beta
// values and all its degrees, at FP64, with the group's recursion coefficients in LDS (otf_coefs:
// branch-free chains), and forms every product from bf16 pieces on MFMA with FP32 accumulation
// (BF16X3 / BF16X6 as in mfma.hpp).  This replaces the chunked split kernels at this size, which
// restarted the recursion per (degree tile, beta chunk) on a quarter of their lanes (28.7 -> 279
// transforms/s for 8 functions in bf16x6).
//
// Inverse (G_k = sum_l (2l+1) F^l d^l(beta_k)), v_mfma_f32_32x32x16_bf16: wave w owns the beta
// rows 32w .. 32w + 31 and all 64 columns (two 32-column tiles).  Lanes cl and cl + 32 both run the
// chain of beta 32w + cl and keep the octet of each 16 degrees that their half of the A operand
// takes, so d never leaves registers.  The coefficient block, scaled by (2l+1) and the pair sign,
// is split into pieces in LDS 64 degrees at a time, the next 64 degrees loading into registers
// meanwhile.  Sixteen waves (four per SIMD) let one wave's MFMAs run beside another's recursion.
//
// Forward (F^l = s_p sum_k w_k d^l(beta_k) S_k), v_mfma_f32_16x16x32_bf16: the reduction runs over
// beta, which the chains own, so d goes through LDS in chunks of 16 degrees, double-buffered.
// Waves 0-7 produce: thread k runs the chain of beta k and writes w_k d split into pieces, rows of
// degree, beta contiguous, for chunk c + 1, then sums the partial tiles of chunk c - 1 and writes
// them to PT as 64-byte row pieces.  Waves 8-15 consume chunk c: wave 8 + 2 kq + ct takes columns
// 32 ct .. 32 ct + 31 and the beta quarter kq, with that quarter of the group's spectrum block
// split into pieces in registers for all the group's degrees.  Each SIMD holds two producers and
// two consumers.
//
// Measured at l_max = 255 x 8 functions, bf16x6 (ms per launch): inverse 4.9, forward 6.7.  Both
// are held by a per-group cost that does not shrink with the degree count: every group writes
// (inverse) or reads (forward) a whole 128 KB workspace block (8.6 GB per launch).  With the MFMAs
// and the recursion both skipped the kernels still took 2.8 and 4.0 ms.
//
// Operand layouts (lane i): 32x32x16: A[i%32][8(i/32)+j], B[8(i/32)+j][i%32] (j < 8), result
// D[8(j/4) + 4(i/32) + j%4][i%32] (j < 16); 16x16x32: A[i%16][8(i/16)+j], B[8(i/16)+j][i%16],
// result D[4(i/16) + j][i%16] (j < 4).
#pragma once

#include "mfma.hpp"
#include "mfma32.hpp"

#define OTF_N 512
#define OTF_COLS 64           // real columns of a workspace row: 8 functions x 4 pairs x (re, im)
#define OTFF_THREADS 1024     // forward
#define OTFI_THREADS 1024     // inverse
#define OTF_LC 64             // degrees per staged coefficient chunk (inverse)
#define OTF_CP (OTF_LC + 8)   // bf16 pitch of a staged coefficient column (inverse)
#define OTF_DL 16             // degrees per d chunk (forward; two chunks in LDS)
#define OTF_DP (OTF_N + 8)    // bf16 pitch of a d row (forward)
#define OTFF_RWAVE (2 * 4 * 64)              // partial tiles of one consumer wave (forward)
#define OTFF_RFLOATS (2 * 8 * OTFF_RWAVE)    // ... of the 8 consumer waves, two chunks

typedef float f32x4_t __attribute__((ext_vector_type(4)));

// recursion coefficient rows: B + 16 (the degrees of a last partial 16-degree step)
__host__ __device__ constexpr size_t otf_inv_lds_bytes(int npc, int B)
{
    return (size_t)npc * OTF_COLS * OTF_CP * 2 + (size_t)(B + 16) * sizeof(RecCoef);
}
__host__ __device__ constexpr size_t otf_fwd_lds_bytes(int npc, int B)
{
    return (size_t)2 * npc * OTF_DL * OTF_DP * 2 + (size_t)OTFF_RFLOATS * 4 + (size_t)(B + 16) * sizeof(RecCoef);
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

// The group's recursion coefficients for a branch-free chain: the chain starts at (d^{m0-1}, d^{m0}) =
// (0, seed) one step early, row m0 = (0, -1, 0) makes its first step return the seed (and row m0 + 1
// has c = 0, so the duplicated seed in the other slot is never used), and rows B .. B + 15 are zero,
// so the chain returns 0 past the last degree.
__device__ __forceinline__ void otf_coefs(const RecTables &t, int B, int m0, int n0, RecCoef *rc, int tid, int nthr)
{
    group_coefs(t, 0, B, m0, n0, rc, tid, nthr);
    if (tid == 0) rc[m0] = RecCoef{0.0, -1.0, 0.0};
    if (tid < 16) rc[B + tid] = RecCoef{0.0, 0.0, 0.0};
}

template <int NPC>
__global__ void __launch_bounds__(OTFI_THREADS) contract_inv_otf(ConArgs<float> a)
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
    // staging role: degree slr of the chunk, pair sp, functions 2 sq and 2 sq + 1 (one float4 of PT)
    const int slr = tid >> 4, sp = (tid >> 2) & 3, sq = tid & 3;
    const bool spv = mx_pair_valid<4>(m0, n0, sp);
    const float ssg = (float)mx_pair_sign<4>(m0, n0, sp);
    const MxPair<4> spr(m0, n0, sp);
    const float4 *PT4 = reinterpret_cast<const float4 *>(a.src);
    float4 st;
    auto load_chunk = [&](int lb) {
        const int l = lb + slr;
        st = make_float4(0.f, 0.f, 0.f, 0.f);
        if (spv && l < B) st = PT4[((long)spr.packed(l) * C + 2 * sq) >> 1];   // C = 8: float4 aligned
    };
    auto store_chunk = [&](int lb) {
        const float cf = ssg * (float)(2 * (lb + slr) + 1);
        const float v[4] = {st.x, st.y, st.z, st.w};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int fi = 2 * sq + (q >> 1), ri = q & 1;
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
    // this lane's chain: beta index k (lanes cl and cl + 32 run the same chain and keep the octet
    // of each 16 degrees that their half of the A operand takes)
    const int k = 32 * wave + cl;
    const double cb = a.rt.cb[k];
    double cur = otf_seed(a.rt, g, m0, n0, k), prev = 0.0;
    f32x16_t acc0, acc1;
#pragma unroll
    for (int j = 0; j < 16; j++) { acc0[j] = 0.f; acc1[j] = 0.f; }
    otf_coefs(a.rt, B, m0, n0, rc, tid, OTFI_THREADS);
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
            float x[8];
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const RecCoef r = rc[l0 + i];
                const double nv = (r.a * cb - r.b) * cur - r.c * prev;
                prev = cur;
                cur = nv;
                if (i < 8) x[i] = (float)nv;
                else if (kh) x[i - 8] = (float)nv;
            }
            bf16x8_t av[NPC], b0[NPC], b1[NPC];
            otf_split8<NPC>(x, av);
#pragma unroll
            for (int c = 0; c < NPC; c++) {
                b0[c] = lds_bf16x8(Cs + (c * OTF_COLS + cl) * OTF_CP + 16 * s + 8 * kh);
                b1[c] = lds_bf16x8(Cs + (c * OTF_COLS + 32 + cl) * OTF_CP + 16 * s + 8 * kh);
            }
#pragma unroll
            for (int i = 0; i < SP::n; i++) {
                acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[SP::a(i)], b0[SP::b(i)], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[SP::a(i)], b1[SP::b(i)], acc1, 0, 0, 0);
            }
        }
    }
    float *G = reinterpret_cast<float *>(a.dst);
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int kk = 32 * wave + 8 * (j >> 2) + 4 * kh + (j & 3);
        float *row = G + ws_row(ca, g, kk) * OTF_COLS;
        row[cl] = acc0[j];
        row[32 + cl] = acc1[j];
    }
}

template <int NPC>
__global__ void __launch_bounds__(OTFF_THREADS) contract_fwd_otf(ConArgs<float> a)
{
    typedef SplitProducts<NPC> SP;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    unsigned short *Dp = reinterpret_cast<unsigned short *>(smem_raw);                   // [2][NPC][16][OTF_DP]
    float *R = reinterpret_cast<float *>(smem_raw + (size_t)2 * NPC * OTF_DL * OTF_DP * 2);   // [2][8][2][4][64]
    RecCoef *rc = reinterpret_cast<RecCoef *>(R + OTFF_RFLOATS);
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, C = ca.C;
    const int g = xcd_deal(1);
    if (g >= ca.ngrp) return;
    int m0, n0;
    so3_group(g, m0, n0);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool producer = wave < 8;
    const int nch = (B - m0 + OTF_DL - 1) / OTF_DL;
    // producer: the chain of beta k = tid, w_k d^l(beta_k) split into pieces, rows of degree
    const int k = tid & 511;
    const double cb = a.rt.cb[k], wk = a.rt.w[k];
    double cur = 0.0, prev = 0.0;
    // consumer wave cw: columns 32 ct .. 32 ct + 31 (two 16-column tiles), beta quarter kq
    const int cw = wave & 7, ct = cw & 1, kq = cw >> 1, ci = lane & 15, kg = lane >> 4;
    bf16x8_t sb[4][2][NPC];   // [32-beta step][column tile][piece]: beta 128 kq + 32 u + 8 kg + j
    if (producer) {
        cur = otf_seed(a.rt, g, m0, n0, k);
    } else {
        const float *W = reinterpret_cast<const float *>(a.src) + 32 * ct + ci;
        float x[4][2][8];
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int t = 0; t < 2; t++)
#pragma unroll
                for (int j = 0; j < 8; j++) x[u][t][j] = W[ws_row(ca, g, 128 * kq + 32 * u + 8 * kg + j) * OTF_COLS + 16 * t];
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int t = 0; t < 2; t++) otf_split8<NPC>(x[u][t], sb[u][t]);
    }
    otf_coefs(a.rt, B, m0, n0, rc, tid, OTFF_THREADS);
    __syncthreads();
    auto produce = [&](int c) {
        unsigned short *D = Dp + (size_t)(c & 1) * NPC * OTF_DL * OTF_DP + k;
        const int l0 = m0 + OTF_DL * c;
#pragma unroll 4
        for (int i = 0; i < OTF_DL; i++) {
            const RecCoef r = rc[l0 + i];
            const double nv = (r.a * cb - r.b) * cur - r.c * prev;
            prev = cur;
            cur = nv;
            float f = (float)(wk * nv);
#pragma unroll
            for (int pc = 0; pc < NPC; pc++) {
                const __bf16 b = (__bf16)f;
                D[(pc * OTF_DL + i) * OTF_DP] = __builtin_bit_cast(unsigned short, b);
                f -= (float)b;
            }
        }
    };
    // reduction of chunk c by the producers: thread -> degree row, pair, function; re and im
    float *PT = reinterpret_cast<float *>(a.dst);
    auto reduce = [&](int c) {
        const int lr = tid >> 5, p = (tid >> 3) & 3, fi = tid & 7, l = m0 + OTF_DL * c + lr;
        if (l >= B || fi >= ca.nfc || !mx_pair_valid<4>(m0, n0, p)) return;
        const int cc = 4 * fi + 2 * (p & 1), st = cc >> 4, ln = 16 * (lr >> 2) + (cc & 15), j = lr & 3;
        const float *Rp = R + (size_t)(c & 1) * (OTFF_RFLOATS / 2) + ((((p >> 1) * 2 + st) * 4 + j) * 64 + ln);
        float re = 0.f, im = 0.f;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            re += Rp[q * 2 * OTFF_RWAVE];   // consumer wave 2 q + ct
            im += Rp[q * 2 * OTFF_RWAVE + 1];
        }
        const float sg = (float)mx_pair_sign<4>(m0, n0, p);
        *reinterpret_cast<float2 *>(PT + ((long)MxPair<4>(m0, n0, p).packed(l) * C + fi) * 2) = make_float2(sg * re, sg * im);
    };
    if (producer) produce(0);
    __syncthreads();
    for (int c = 0; c < nch; c++) {
        if (producer) {
            if (c > 0) reduce(c - 1);
            if (c + 1 < nch) produce(c + 1);
        } else {
            const unsigned short *Da = Dp + (size_t)(c & 1) * NPC * OTF_DL * OTF_DP + ci * OTF_DP + 128 * kq + 8 * kg;
            f32x4_t acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int u = 0; u < 4; u++) {
                bf16x8_t av[NPC];
#pragma unroll
                for (int pc = 0; pc < NPC; pc++) av[pc] = lds_bf16x8(Da + pc * OTF_DL * OTF_DP + 32 * u);
#pragma unroll
                for (int i = 0; i < SP::n; i++) {
                    acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[SP::a(i)], sb[u][0][SP::b(i)], acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[SP::a(i)], sb[u][1][SP::b(i)], acc1, 0, 0, 0);
                }
            }
            float *Rw = R + (size_t)(c & 1) * (OTFF_RFLOATS / 2) + (size_t)cw * OTFF_RWAVE + lane;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                Rw[j * 64] = acc0[j];
                Rw[(4 + j) * 64] = acc1[j];
            }
        }
        __syncthreads();
    }
    if (producer) reduce(nch - 1);
}
