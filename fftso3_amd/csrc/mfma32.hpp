// mfma32.hpp -- batch-wide beta contraction on FP32 MFMA (v_mfma_f32_32x32x2f32), N = 2B <= 128.
//
// One workgroup per (group g, block of FPI functions): the FP64 recursion writes the group's
// whole Wigner-d (or Legendre) block D[l][k] (M <= 64 rows, N <= 128 columns) to LDS once, and
// the four waves then stream the spectrum of FPI functions straight from HBM into MFMA operand
// registers: every spectrum value is read exactly once per transform and the recursion is shared
// by FPI functions instead of one chunk.
//
// Operand layouts of v_mfma_f32_32x32x2f32 (lane i): A[i%32][i/32], B[i/32][i%32], result
// D[8(j/4) + 4(i/32) + j%4][i%32] for j < 16.
#pragma once

#include "kernels.hpp"

typedef float f32x16m_t __attribute__((ext_vector_type(16)));

#define MX_FPI 32            // functions per work item
#define MX_DPITCH 129        // LDS pitch of D rows (floats), conflict-free column reads

// Forward: F^l(fn, pair) = s_p sum_k w_k d^l_{m0 n0}(beta_k) S_k(fn, pair).
// Workspace row k of group g holds CW = C*NP*2 reals (function, pair, re/im).
template <int NP>
__global__ void __launch_bounds__(256) contract_fwd_mx32(ConArgs<float> a)
{
    constexpr int NCOL = MX_FPI * NP * 2, NCT = NCOL / 32;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *D = reinterpret_cast<float *>(smem_raw);     // [64][MX_DPITCH], rows l - m0
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, CW = ca.C * NP * 2;
    const int unit = xcd_unit(a.nitems);
    if (unit >= a.nitems) return;
    const int2 it = a.items[unit];
    const int g = it.x, fb = it.y;                       // fb: first function of the block
    int m0, n0;
    if (NP == 2) { m0 = g; n0 = 0; } else so3_group(g, m0, n0);
    const int M = B - m0, nrt = (M + 31) / 32;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int k = tid; k < N; k += 256) {
        const double wk = a.rt.w[k];
        group_recursion(a.rt, NP == 2, B, g, m0, n0, k, B, [&](int l, double v) {
            D[(l - m0) * MX_DPITCH + k] = (float)(wk * v);
        });
        for (int r = M; r < 32 * nrt; r++) D[r * MX_DPITCH + k] = 0.f;
    }
    __syncthreads();
    const float *S = reinterpret_cast<const float *>(a.src) + (size_t)g * N * CW;
    const int kh = lane >> 5, cl = lane & 31;
    for (int ct = wave; ct < NCT; ct += 4) {
        const int colb = fb * NP * 2 + ct * 32;          // first workspace column of the tile
        if (colb >= ca.nfc * NP * 2) break;
        const float *Sc = S + colb + cl;
        f32x16m_t acc0, acc1;
#pragma unroll
        for (int j = 0; j < 16; j++) { acc0[j] = 0.f; acc1[j] = 0.f; }
        for (int k0 = 0; k0 < N; k0 += 16) {
            float b[8];
#pragma unroll
            for (int s = 0; s < 8; s++) {
                const int k = k0 + 2 * s + kh;
                b[s] = k < N ? Sc[(size_t)k * CW] : 0.f;
            }
#pragma unroll
            for (int s = 0; s < 8; s++) {
                const int k = k0 + 2 * s + kh;
                const float a0 = k < N ? D[cl * MX_DPITCH + k] : 0.f;
                acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b[s], acc0, 0, 0, 0);
                if (nrt > 1) {
                    const float a1 = k < N ? D[(32 + cl) * MX_DPITCH + k] : 0.f;
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b[s], acc1, 0, 0, 0);
                }
            }
        }
        // store: column -> (function, pair, re/im)
        const int col = colb + cl;
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
            const int r = 8 * (j >> 2) + 4 * kh + (j & 3);
            const int l0 = m0 + r, l1 = m0 + 32 + r;
            if (l0 < B) out[2 * (long)(NP == 2 ? s2_idx(l0, m) : so3_idx(l0, m, n)) + ri] = (float)sp * acc0[j];
            if (nrt > 1 && l1 < B) out[2 * (long)(NP == 2 ? s2_idx(l1, m) : so3_idx(l1, m, n)) + ri] = (float)sp * acc1[j];
        }
    }
}

// Inverse: G_k(fn, pair) = s_p sum_l c_l d^l_{m0 n0}(beta_k) F^l(fn, pair) (c_l = 2l+1 for SO(3),
// 1 for S^2).  Output tile per wave: all k (up to 4 tiles of 32) x 32 columns; reduction over l.
template <int NP>
__global__ void __launch_bounds__(256) contract_inv_mx32(ConArgs<float> a)
{
    constexpr int NCOL = MX_FPI * NP * 2, NCT = NCOL / 32;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *D = reinterpret_cast<float *>(smem_raw);     // [64][MX_DPITCH], rows l - m0
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, CW = ca.C * NP * 2;
    const int unit = xcd_unit(a.nitems);
    if (unit >= a.nitems) return;
    const int2 it = a.items[unit];
    const int g = it.x, fb = it.y;
    int m0, n0;
    if (NP == 2) { m0 = g; n0 = 0; } else so3_group(g, m0, n0);
    const int M = B - m0, M2 = (M + 1) & ~1, nkt = (N + 31) / 32;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int k = tid; k < 32 * nkt; k += 256) {
        if (k < N) {
            group_recursion(a.rt, NP == 2, B, g, m0, n0, k, B, [&](int l, double v) {
                D[(l - m0) * MX_DPITCH + k] = (float)(NP == 2 ? v : (double)(2 * l + 1) * v);
            });
            if (M2 > M) D[M * MX_DPITCH + k] = 0.f;
        } else {
            for (int r = 0; r < M2; r++) D[r * MX_DPITCH + k] = 0.f;
        }
    }
    __syncthreads();
    const int kh = lane >> 5, cl = lane & 31;
    float *G = reinterpret_cast<float *>(a.dst) + (size_t)g * N * CW;
    for (int ct = wave; ct < NCT; ct += 4) {
        const int colb = fb * NP * 2 + ct * 32;
        if (colb >= ca.nfc * NP * 2) break;
        const int col = colb + cl;
        const int ri = col & 1, p = (col >> 1) % NP, fi = col / (2 * NP);
        int m, n, sp;
        if (NP == 2) { m = p ? -m0 : m0; n = 0; sp = p ? parity_sign(m0) : 1; }
        else { so3_pair(m0, n0, p, m, n); sp = so3_pair_sign(m0, n0, p); }
        const float *F = fi < ca.nfc ? reinterpret_cast<const float *>(a.src) + 2 * ca.coef_off[ca.fn[fi]] : nullptr;
        f32x16m_t acc[4];
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int j = 0; j < 16; j++) acc[t][j] = 0.f;
        for (int lb = 0; lb < M2; lb += 16) {
            float b[8];
#pragma unroll
            for (int s = 0; s < 8; s++) {
                const int lr = lb + 2 * s + kh, l = m0 + lr;
                float v = 0.f;
                if (F && lr < M) v = (float)sp * F[2 * (long)(NP == 2 ? s2_idx(l, m) : so3_idx(l, m, n)) + ri];
                b[s] = v;
            }
#pragma unroll
            for (int s = 0; s < 8; s++) {
                const int lr = lb + 2 * s + kh;
                if (lb + 2 * s < M2) {
#pragma unroll
                    for (int t = 0; t < 4; t++)
                        if (t < nkt) {
                            const float av = D[lr * MX_DPITCH + t * 32 + cl];
                            acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(av, b[s], acc[t], 0, 0, 0);
                        }
                }
            }
        }
        if (fi >= ca.nfc) continue;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            if (t >= nkt) break;
#pragma unroll
            for (int j = 0; j < 16; j++) {
                const int k = t * 32 + 8 * (j >> 2) + 4 * kh + (j & 3);
                if (k < N) G[(size_t)k * CW + col] = acc[t][j];
            }
        }
    }
}
