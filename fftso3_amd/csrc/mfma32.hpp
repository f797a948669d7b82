// mfma32.hpp -- batch-wide beta contraction on FP32 MFMA (v_mfma_f32_32x32x2f32), N = 2B <= 128.
//
// One workgroup per (m,n) group g for the whole batch: the FP64 recursion writes the group's
// Wigner-d (or Legendre) block D[l][k] (M = B - m0 <= 64 rows, N <= 128 columns) to LDS once and
// the four waves stream 32-column tiles of the batch through MFMA, the next tile's operands
// loading while the current tile multiplies.  Columns are (function, pair, re/im) as in the
// spectrum workspace [g][k][function][pair].  Coefficients enter and leave through PT[e][function],
// the transpose of the reference's packed coefficient array (e = packed index of (l,m,n)): the
// four waves of a workgroup multiply four consecutive column tiles at the same time, so together
// they read or write whole 128-byte segments of every PT row they touch.
//
// Operand layouts of v_mfma_f32_32x32x2f32 (lane i): A[i%32][i/32], B[i/32][i%32], result
// D[8(j/4) + 4(i/32) + j%4][i%32] for j < 16.  A lane holds N/2 <= 64 forward and M/2 <= 32
// inverse operands per tile (B <= 64); offsets within one group's block fit in 32 bits.
#pragma once

#include "kernels.hpp"

typedef float f32x16m_t __attribute__((ext_vector_type(16)));

#define MX_DPITCH 129        // LDS pitch of D rows (floats): conflict-free column reads

struct MxArgs {
    ChunkArgs ca;            // ca.C = functions of the class (workspace stride)
    RecTables rt;
    const int *items;        // group per workgroup (largest first)
    int nitems;
    const void *src;         // forward: spectrum workspace; inverse: PT
    void *dst;               // forward: PT; inverse: spectrum workspace
};

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

// packed index of degree l of pair p of group (m0, n0)
template <int NP>
__device__ __forceinline__ int mx_packed(int l, int m0, int n0, int p)
{
    if (NP == 2) return s2_idx(l, p ? -m0 : m0);
    int m, n;
    so3_pair(m0, n0, p, m, n);
    return so3_idx(l, m, n);
}

// Forward: PT(l, pair; fn) = s_p sum_k w_k d^l_{m0 n0}(beta_k) S_k(fn, pair).
template <int NP>
__global__ void __launch_bounds__(256) contract_fwd_mx32(MxArgs a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *D = reinterpret_cast<float *>(smem_raw);     // [64][MX_DPITCH], rows l - m0
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, CW = ca.C * NP * 2, NCOL = ca.nfc * NP * 2;
    const int unit = xcd_unit(a.nitems);
    if (unit >= a.nitems) return;
    const int g = a.items[unit];
    int m0, n0;
    mx_group<NP>(g, m0, n0);
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
    float *PT = reinterpret_cast<float *>(a.dst);
    const int kh = lane >> 5, cl = lane & 31;
    const int NCT = (NCOL + 31) / 32;
    float bA[64], bB[64];
    // Loads go straight into the operand registers: nothing is applied to a loaded value, so the
    // loads of a tile stay in flight together.  An invalid column (c >= NCOL) reads column 0 and
    // rows k >= N read row 0; MFMA columns are independent and invalid ones are never stored, and
    // tile rows k = 2s + kh with 2s >= N are skipped by the multiply (N is even).
#define MX_FWD_LOAD(buf, ct_)                                                      \
    {                                                                              \
        const int c_ = (ct_) * 32 + cl;                                            \
        const float *p_ = S + kh * CW + (c_ < NCOL ? c_ : 0);                      \
        _Pragma("unroll") for (int s = 0; s < 64; s++)                             \
            buf[s] = p_[(2 * s < N ? s : 0) * 2 * CW];                             \
    }
#define MX_FWD_TILE(buf, ct_)                                                                    \
    {                                                                                            \
        f32x16m_t acc0, acc1;                                                                    \
        _Pragma("unroll") for (int j = 0; j < 16; j++) { acc0[j] = 0.f; acc1[j] = 0.f; }         \
        _Pragma("unroll") for (int s = 0; s < 64; s++)                                           \
        {                                                                                        \
            if (2 * s < N) {                                                                     \
                const float a0 = D[cl * MX_DPITCH + 2 * s + kh];                                 \
                acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, buf[s], acc0, 0, 0, 0);          \
                if (nrt > 1) {                                                                   \
                    const float a1 = D[(32 + cl) * MX_DPITCH + 2 * s + kh];                      \
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, buf[s], acc1, 0, 0, 0);      \
                }                                                                                \
            }                                                                                    \
        }                                                                                        \
        const int c_ = (ct_) * 32 + cl;                                                          \
        if (c_ < NCOL && mx_pair_valid<NP>(m0, n0, (c_ >> 1) % NP)) {                            \
            const int p_ = (c_ >> 1) % NP;                                                       \
            const float sg = (float)mx_pair_sign<NP>(m0, n0, p_);                                \
            float *o_ = PT + 2 * (c_ / (2 * NP)) + (c_ & 1);                                     \
            _Pragma("unroll") for (int j = 0; j < 16; j++)                                       \
            {                                                                                    \
                const int r = 8 * (j >> 2) + 4 * kh + (j & 3);                                   \
                if (r < M) o_[2L * mx_packed<NP>(m0 + r, m0, n0, p_) * ca.C] = sg * acc0[j];     \
                if (nrt > 1 && r + 32 < M)                                                       \
                    o_[2L * mx_packed<NP>(m0 + r + 32, m0, n0, p_) * ca.C] = sg * acc1[j];       \
            }                                                                                    \
        }                                                                                        \
    }
    int ct = wave;
    if (ct < NCT) MX_FWD_LOAD(bA, ct)
#pragma unroll 1
    for (; ct < NCT; ct += 8) {
        if (ct + 4 < NCT) MX_FWD_LOAD(bB, ct + 4)
        MX_FWD_TILE(bA, ct)
        if (ct + 4 >= NCT) break;
        if (ct + 8 < NCT) MX_FWD_LOAD(bA, ct + 8)
        MX_FWD_TILE(bB, ct + 4)
    }
#undef MX_FWD_LOAD
#undef MX_FWD_TILE
}

// Inverse: G_k(fn, pair) = s_p sum_l c_l d^l_{m0 n0}(beta_k) PT(l, pair; fn) (c_l = 2l+1 for SO(3),
// 1 for S^2).  Output tile per wave: all k (up to 4 tiles of 32) x 32 columns.
template <int NP>
__global__ void __launch_bounds__(256) contract_inv_mx32(MxArgs a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float *D = reinterpret_cast<float *>(smem_raw);     // [64][MX_DPITCH], rows l - m0
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, CW = ca.C * NP * 2, NCOL = ca.nfc * NP * 2;
    const int unit = xcd_unit(a.nitems);
    if (unit >= a.nitems) return;
    const int g = a.items[unit];
    int m0, n0;
    mx_group<NP>(g, m0, n0);
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
    const float *PT = reinterpret_cast<const float *>(a.src);
    float *G = reinterpret_cast<float *>(a.dst) + (size_t)g * N * CW;
    const int NCT = (NCOL + 31) / 32;
    float bA[32], bB[32];
    // As in the forward kernel, loads go straight into operand registers: an invalid column reads
    // column 0 and rows l >= M read row m0 (finite coefficients) against zero rows of D (rows up to
    // M2 are zeroed, later ones skipped).  The pair sign is applied to the result.
#define MX_INV_LOAD(buf, ct_)                                                        \
    {                                                                                \
        const int c_ = (ct_) * 32 + cl;                                              \
        const int cv_ = c_ < NCOL ? c_ : 0;                                          \
        const int p_ = (cv_ >> 1) % NP;                                              \
        const float *i_ = PT + 2 * (cv_ / (2 * NP)) + (cv_ & 1);                     \
        _Pragma("unroll") for (int s = 0; s < 32; s++)                               \
        {                                                                            \
            const int lr = 2 * s + kh;                                               \
            buf[s] = i_[2L * mx_packed<NP>(m0 + (lr < M ? lr : 0), m0, n0, p_) * ca.C]; \
        }                                                                            \
    }
#define MX_INV_TILE(buf, ct_)                                                                    \
    {                                                                                            \
        f32x16m_t acc[4];                                                                        \
        _Pragma("unroll") for (int t = 0; t < 4; t++)                                            \
            _Pragma("unroll") for (int j = 0; j < 16; j++) acc[t][j] = 0.f;                      \
        _Pragma("unroll") for (int s = 0; s < 32; s++)                                           \
        {                                                                                        \
            if (2 * s < M2) {                                                                    \
                const int lr = 2 * s + kh;                                                       \
                _Pragma("unroll") for (int t = 0; t < 4; t++) if (t < nkt)                       \
                {                                                                                \
                    const float av = D[lr * MX_DPITCH + t * 32 + cl];                            \
                    acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(av, buf[s], acc[t], 0, 0, 0);  \
                }                                                                                \
            }                                                                                    \
        }                                                                                        \
        const int c_ = (ct_) * 32 + cl;                                                          \
        if (c_ < NCOL) {                                                                         \
            const float sg = (float)mx_pair_sign<NP>(m0, n0, (c_ >> 1) % NP);                    \
            _Pragma("unroll") for (int t = 0; t < 4; t++) if (t < nkt)                           \
            {                                                                                    \
                _Pragma("unroll") for (int j = 0; j < 16; j++)                                   \
                {                                                                                \
                    const int k = t * 32 + 8 * (j >> 2) + 4 * kh + (j & 3);                      \
                    if (k < N) G[k * CW + c_] = sg * acc[t][j];                                  \
                }                                                                                \
            }                                                                                    \
        }                                                                                        \
    }
    int ct = wave;
    if (ct < NCT) MX_INV_LOAD(bA, ct)
#pragma unroll 1
    for (; ct < NCT; ct += 8) {
        if (ct + 4 < NCT) MX_INV_LOAD(bB, ct + 4)
        MX_INV_TILE(bA, ct)
        if (ct + 4 >= NCT) break;
        if (ct + 8 < NCT) MX_INV_LOAD(bA, ct + 8)
        MX_INV_TILE(bB, ct + 4)
    }
#undef MX_INV_LOAD
#undef MX_INV_TILE
}

// ---------------------------------------------------------------------------------------------
// Coefficient transpose between the reference's packed layout (FFTSO3_matrix / FFTS2_matrix
// storage, one function after another: [function][e]) and PT[e][function].  Tiles of 64 packed
// indices x 32 functions go through LDS so both sides move in contiguous 512- and 256-byte runs.
// DIR 0: PT -> packed, DIR 1: packed -> PT.
// ---------------------------------------------------------------------------------------------
template <int DIR>
__global__ void __launch_bounds__(256) coef_transpose(ChunkArgs ca, int ncoef, const void *src, void *dst)
{
    __shared__ float2 tile[32][65];
    const int nte = (ncoef + 63) / 64, ntf = (ca.nfc + 31) / 32, nunit = nte * ntf;
    const int unit = xcd_unit(nunit);
    if (unit >= nunit) return;
    const int et = unit / ntf, ft = unit - et * ntf;
    const int e0 = et * 64, f0 = ft * 32;
    const int t = threadIdx.x;
    const int fl = t >> 3, eo = (t & 7) * 8;        // packed side: function row, 8 elements
    const int el = t >> 2, fo = (t & 3) * 8;        // PT side: element row, 8 functions
    const int fi = f0 + fl;
    if (DIR == 1) {
        if (fi < ca.nfc) {
            const float2 *P = reinterpret_cast<const float2 *>(src) + ca.coef_off[ca.fn[fi]] + e0 + eo;
#pragma unroll
            for (int j = 0; j < 8; j++)
                if (e0 + eo + j < ncoef) tile[fl][eo + j] = P[j];
        }
        __syncthreads();
        if (e0 + el < ncoef) {
            float2 *Q = reinterpret_cast<float2 *>(dst) + (long)(e0 + el) * ca.C + f0 + fo;
#pragma unroll
            for (int j = 0; j < 8; j++)
                if (f0 + fo + j < ca.nfc) Q[j] = tile[fo + j][el];
        }
    } else {
        if (e0 + el < ncoef) {
            const float2 *Q = reinterpret_cast<const float2 *>(src) + (long)(e0 + el) * ca.C + f0 + fo;
#pragma unroll
            for (int j = 0; j < 8; j++)
                if (f0 + fo + j < ca.nfc) tile[fo + j][el] = Q[j];
        }
        __syncthreads();
        if (fi < ca.nfc) {
            float2 *P = reinterpret_cast<float2 *>(dst) + ca.coef_off[ca.fn[fi]] + e0 + eo;
#pragma unroll
            for (int j = 0; j < 8; j++)
                if (e0 + eo + j < ncoef) P[j] = tile[fl][eo + j];
        }
    }
}
