// mfma.hpp -- split-BF16 MFMA contraction kernels (gfx950, v_mfma_f32_32x32x16_bf16).
//
// Each FP32 (or FP64) operand x is split into bf16 pieces x = x0 + x1 (+ x2), every residual
// exact in FP32, and the product is formed from the leading cross terms on the matrix cores
// with FP32 accumulation:
//   BF16X3: a0 b0 + a0 b1 + a1 b0                           (2 pieces, 3 products)
//   BF16X6: a0 b0 + a0 b1 + a1 b0 + a0 b2 + a1 b1 + a2 b0   (3 pieces, 6 products)
// (Henry, Tang & Heinecke, "Leveraging the bfloat16 artificial intelligence datatype for
// higher-precision computations", ARITH 2019.)  The Wigner-d / Legendre values are generated in
// FP64 and split from the FP64 value.
//
// Operand layouts of v_mfma_f32_32x32x16_bf16 (64 lanes, lane i): A[i%32][8(i/32)+j] and
// B[8(i/32)+j][i%32] for j < 8; result D[8(j/4) + 4(i/32) + j%4][i%32] for j < 16.
// Both operands are kept in LDS with the reduction index contiguous (16-byte lane reads).
#pragma once

#include "kernels.hpp"

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef unsigned short u16x8_t __attribute__((ext_vector_type(8)));

// bf16 pieces: the hardware conversion (v_cvt_pk_bf16_f32) rounds to nearest even; each residual
// is exact
template <int NPC> __device__ __forceinline__ void split_d(double v, unsigned short *out)
{
    double r = v;
#pragma unroll
    for (int i = 0; i < NPC; i++) {
        const __bf16 b = (__bf16)(float)r;
        out[i] = __builtin_bit_cast(unsigned short, b);
        r -= (double)(float)b;
    }
}
template <int NPC> __device__ __forceinline__ void split_f(float v, unsigned short *out)
{
    float r = v;
#pragma unroll
    for (int i = 0; i < NPC; i++) {
        const __bf16 b = (__bf16)r;
        out[i] = __builtin_bit_cast(unsigned short, b);
        r -= (float)b;
    }
}

// product list (piece of A, piece of B), smallest terms first
template <int NPC> struct SplitProducts;
template <> struct SplitProducts<2> {
    static constexpr int n = 3;
    __device__ static constexpr int a(int i) { return i == 0 ? 1 : 0; }
    __device__ static constexpr int b(int i) { return i == 1 ? 1 : 0; }
};
template <> struct SplitProducts<3> {
    static constexpr int n = 6;
    // (2,0) (1,1) (0,2) (1,0) (0,1) (0,0)
    __device__ static constexpr int a(int i) { return i == 0 ? 2 : (i == 1 ? 1 : (i == 3 ? 1 : 0)); }
    __device__ static constexpr int b(int i) { return i == 1 ? 1 : (i == 2 ? 2 : (i == 4 ? 1 : 0)); }
};

__device__ __forceinline__ bf16x8_t lds_bf16x8(const unsigned short *p)
{
    u16x8_t v = *reinterpret_cast<const u16x8_t *>(p);
    return __builtin_bit_cast(bf16x8_t, v);
}

// Forward contraction on MFMA.  Item = (g, l-tile of 64 rows); all C = 8 functions of the chunk;
// NP pairs -> NCOL = 16 NP real columns.  k is streamed in chunks of KC = 64: lanes 0..63 run the
// FP64 recursion for their beta_k and write the split D rows, the spectrum chunk (prefetched in
// registers during the previous chunk's products) is split and stored transposed, then each wave
// multiplies its 32x32 tiles.
template <int NP, int NPC>
__global__ void __launch_bounds__(256) contract_fwd_bf16(ConArgs<float> a)
{
    constexpr int LT = 64, KC = 64, KP = KC + 8, NCOL = 16 * NP, NT = 256;
    constexpr int NCQ = NCOL / 4;                 // column quads
    typedef SplitProducts<NPC> SP;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    unsigned short *Dp = reinterpret_cast<unsigned short *>(smem_raw);     // [NPC][LT][KP]
    unsigned short *Sp = Dp + NPC * LT * KP;                                // [NPC][NCOL][KP]
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N;
    const int unit = xcd_unit(a.nitems);
    if (unit >= a.nitems) return;
    const int2 it = a.items[unit];
    const int g = it.x, l0 = it.y;
    int m0, n0;
    if (NP == 2) { m0 = g; n0 = 0; } else so3_group(g, m0, n0);
    const int lend = min(B, l0 + LT);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nrt = (lend - l0 + 31) / 32, nct = NCOL / 32, ntile = nrt * nct;
    // spectrum staging role: column quad cq, k octet ko
    const int cq = tid % NCQ, ko = tid / NCQ;
    const bool s_role = ko < KC / 8;
    const float4 *S4 = reinterpret_cast<const float4 *>(reinterpret_cast<const float *>(a.src) + (size_t)g * N * NCOL);
    float4 sreg[8];
    auto load_s = [&](int kc) {
        if (!s_role) return;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            int k = kc + 8 * ko + j;
            sreg[j] = k < N ? S4[((size_t)k * NCOL) / 4 + cq] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    };
    f32x16_t acc[2];
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int j = 0; j < 16; j++) acc[t][j] = 0.f;
    load_s(0);
    for (int kc = 0; kc < N; kc += KC) {
        __syncthreads();
        if (tid < KC) {
            const int k = kc + tid;
            if (k < N) {
                const double wk = a.rt.w[k];
                group_recursion(a.rt, NP == 2, B, g, m0, n0, k, lend, [&](int l, double v) {
                    if (l >= l0) {
                        unsigned short pc[NPC];
                        split_d<NPC>(wk * v, pc);
#pragma unroll
                        for (int q = 0; q < NPC; q++) Dp[(q * LT + (l - l0)) * KP + tid] = pc[q];
                    }
                });
                for (int l = max(lend, l0); l < l0 + LT; l++)
#pragma unroll
                    for (int q = 0; q < NPC; q++) Dp[(q * LT + (l - l0)) * KP + tid] = 0;
            } else {
                for (int r = 0; r < LT; r++)
#pragma unroll
                    for (int q = 0; q < NPC; q++) Dp[(q * LT + r) * KP + tid] = 0;
            }
        }
        if (s_role) {
#pragma unroll
            for (int c4 = 0; c4 < 4; c4++) {
                u16x8_t pcs[NPC];
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    float v = c4 == 0 ? sreg[j].x : c4 == 1 ? sreg[j].y : c4 == 2 ? sreg[j].z : sreg[j].w;
                    unsigned short pc[NPC];
                    split_f<NPC>(v, pc);
#pragma unroll
                    for (int q = 0; q < NPC; q++) pcs[q][j] = pc[q];
                }
#pragma unroll
                for (int q = 0; q < NPC; q++)
                    *reinterpret_cast<u16x8_t *>(&Sp[(q * NCOL + 4 * cq + c4) * KP + 8 * ko]) = pcs[q];
            }
        }
        __syncthreads();
        if (kc + KC < N) load_s(kc + KC);
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int tile = wave + 4 * t;
            if (tile < ntile) {
                const int rt = tile / nct, ct = tile - rt * nct;
                const unsigned short *Ar = Dp + (rt * 32 + (lane & 31)) * KP + 8 * (lane >> 5);
                const unsigned short *Br = Sp + (ct * 32 + (lane & 31)) * KP + 8 * (lane >> 5);
#pragma unroll
                for (int ks = 0; ks < KC / 16; ks++) {
#pragma unroll
                    for (int pr = 0; pr < SP::n; pr++) {
                        bf16x8_t av = lds_bf16x8(Ar + SP::a(pr) * LT * KP + ks * 16);
                        bf16x8_t bv = lds_bf16x8(Br + SP::b(pr) * NCOL * KP + ks * 16);
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[t], 0, 0, 0);
                    }
                }
            }
        }
    }
    // store
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const int tile = wave + 4 * t;
        if (tile >= ntile) continue;
        const int rt = tile / nct, ct = tile - rt * nct;
        const int col = ct * 32 + (lane & 31);
        const int ri = col & 1, p = (col >> 1) % NP, fi = col / (2 * NP);
        if (fi >= ca.nfc) continue;
        int m, n, sp;
        bool valid;
        if (NP == 2) {
            m = p ? -m0 : m0; n = 0;
            sp = p ? parity_sign(m0) : 1;
            valid = (p == 0) || (m0 > 0);
        } else {
            so3_pair(m0, n0, p, m, n);
            sp = so3_pair_sign(m0, n0, p);
            valid = so3_pair_valid(m0, n0, p);
        }
        if (!valid) continue;
        float *out = reinterpret_cast<float *>(a.dst) + 2 * ca.coef_off[ca.fn[fi]];
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int l = l0 + rt * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
            if (l < lend) {
                long idx = NP == 2 ? s2_idx(l, m) : so3_idx(l, m, n);
                out[2 * idx + ri] = (float)sp * acc[t][j];
            }
        }
    }
}

// Inverse contraction on MFMA.  Item = (g, k-tile of 128); l is streamed in chunks of LC = 64:
// the coefficient gather for the chunk is issued first, lanes 0..127 continue their FP64
// recursion through the chunk and write D^T rows (l contiguous), then each wave multiplies its
// 32x32 (k, column) tiles.
template <int NP, int NPC>
__global__ void __launch_bounds__(256) contract_inv_bf16(ConArgs<float> a)
{
    constexpr int KT = 128, LC = 64, LP = LC + 8, NCOL = 16 * NP, NCC = NCOL / 2, NT = 256;
    constexpr int PF = LC * NCC / NT;             // gathered complex values per thread per chunk
    typedef SplitProducts<NPC> SP;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    unsigned short *Dt = reinterpret_cast<unsigned short *>(smem_raw);     // [NPC][KT][LP]
    unsigned short *Fp = Dt + NPC * KT * LP;                                // [NPC][NCOL][LP]
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N;
    const int unit = xcd_unit(a.nitems);
    if (unit >= a.nitems) return;
    const int2 it = a.items[unit];
    const int g = it.x, k0 = it.y;
    int m0, n0;
    if (NP == 2) { m0 = g; n0 = 0; } else so3_group(g, m0, n0);
    const int kn = min(KT, N - k0);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nrt = (kn + 31) / 32, nct = NCOL / 32, ntile = nrt * nct;
    float2 fr[PF];
    auto gather = [&](int lb) {
#pragma unroll
        for (int i = 0; i < PF; i++) {
            int e = tid + i * NT;
            int lr = e / NCC, cc = e - lr * NCC, l = lb + lr;
            int p = cc % NP, fi = cc / NP;
            float2 val = make_float2(0.f, 0.f);
            if (l < B && fi < ca.nfc) {
                int m, n, sp;
                if (NP == 2) { m = p ? -m0 : m0; n = 0; sp = p ? parity_sign(m0) : 1; }
                else { so3_pair(m0, n0, p, m, n); sp = so3_pair_sign(m0, n0, p); }
                const float2 *F = reinterpret_cast<const float2 *>(a.src) + ca.coef_off[ca.fn[fi]];
                float2 f = F[NP == 2 ? s2_idx(l, m) : so3_idx(l, m, n)];
                val = make_float2((float)sp * f.x, (float)sp * f.y);
            }
            fr[i] = val;
        }
    };
    f32x16_t acc[2];
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int j = 0; j < 16; j++) acc[t][j] = 0.f;
    // recursion state of lane k (tid < KT)
    WignerRec wr;
    LegendreRec lr_;
    double dv = 0.0;
    const int kk = k0 + tid;
    const bool rec_role = tid < KT && kk < N;
    if (rec_role) {
        if (NP == 2) { lr_.init(a.rt, B, g, kk); dv = lr_.d1; }
        else { wr.init(a.rt, B, g, m0, n0, kk); dv = wr.d1; }
    }
    for (int lb = m0; lb < B; lb += LC) {
        gather(lb);
        __syncthreads();
        if (tid < KT) {
            unsigned short *row = Dt + tid * LP;
#pragma unroll 1
            for (int l8 = 0; l8 < LC; l8 += 8) {
                u16x8_t pcs[NPC];
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int l = lb + l8 + j;
                    double v = 0.0;
                    if (rec_role && l < B) {
                        if (l > m0) dv = NP == 2 ? lr_.step(l) : wr.step(l);
                        v = NP == 2 ? dv : (double)(2 * l + 1) * dv;
                    }
                    unsigned short pc[NPC];
                    split_d<NPC>(v, pc);
#pragma unroll
                    for (int q = 0; q < NPC; q++) pcs[q][j] = pc[q];
                }
#pragma unroll
                for (int q = 0; q < NPC; q++) *reinterpret_cast<u16x8_t *>(row + q * KT * LP + l8) = pcs[q];
            }
        }
#pragma unroll
        for (int i = 0; i < PF; i++) {
            int e = tid + i * NT;
            int lr = e / NCC, cc = e - lr * NCC;
            unsigned short pr[NPC], pi[NPC];
            split_f<NPC>(fr[i].x, pr);
            split_f<NPC>(fr[i].y, pi);
#pragma unroll
            for (int q = 0; q < NPC; q++) {
                Fp[(q * NCOL + 2 * cc) * LP + lr] = pr[q];
                Fp[(q * NCOL + 2 * cc + 1) * LP + lr] = pi[q];
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int tile = wave + 4 * t;
            if (tile < ntile) {
                const int rt = tile / nct, ct = tile - rt * nct;
                const unsigned short *Ar = Dt + (rt * 32 + (lane & 31)) * LP + 8 * (lane >> 5);
                const unsigned short *Br = Fp + (ct * 32 + (lane & 31)) * LP + 8 * (lane >> 5);
#pragma unroll
                for (int ks = 0; ks < LC / 16; ks++) {
#pragma unroll
                    for (int pr = 0; pr < SP::n; pr++) {
                        bf16x8_t av = lds_bf16x8(Ar + SP::a(pr) * KT * LP + ks * 16);
                        bf16x8_t bv = lds_bf16x8(Br + SP::b(pr) * NCOL * LP + ks * 16);
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[t], 0, 0, 0);
                    }
                }
            }
        }
    }
    float *G = reinterpret_cast<float *>(a.dst) + (size_t)g * N * NCOL;
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const int tile = wave + 4 * t;
        if (tile >= ntile) continue;
        const int rt = tile / nct, ct = tile - rt * nct;
        const int col = ct * 32 + (lane & 31);
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int kr = rt * 32 + 8 * (j >> 2) + 4 * (lane >> 5) + (j & 3);
            if (kr < kn) G[(size_t)(k0 + kr) * NCOL + col] = acc[t][j];
        }
    }
}
