// fft_r16.hpp -- 2-D DFT of one N x N complex-float beta slice for N = 16R (R = 1, 2, 3, 4, 6),
// gfx950: the SO(3) slices of bandlimits B = 8R (l_max = 7, 15, 23, 31, 47).
//
// The same two-stage decomposition as the 128-point kernels, N = R x 16:
//   X[s + R q] = sum_j W16^(j q) [ W_N^(j s) sum_r x[j + 16 r] W_R^(r s) ]
// stage A is a radix-R DFT over x[j + 16 r] (r < R) and the twiddle W_N^(j s), written back in
// place at 16 s + j; stage B is a 16-point DFT in registers of the 16 contiguous values of one s,
// written to s + R q (natural order).  Rows (gamma) first, then columns (alpha); the slice lives
// in LDS with row pitch N + 1, so every stage's lane pattern is conflict free.
//
// T = N^2 / 16 threads work on one slice (stage B: one 16-point DFT per thread), and a workgroup
// of NTHR threads works on NTHR / T slices side by side, each thread group in its own LDS region,
// so the small slices of low bandlimits keep the whole workgroup busy.  A thread group without a
// slice of its own (slices past the class's end, or a partial last group when NTHR is not a
// multiple of T) runs the same barriers and LDS work on its own region and skips every global
// load and store.  The class's own launches use whole thread groups only (r16_threads: 576
// threads and one 74 KB region at R = 6, so two workgroups share a CU); the unified multi-class
// FFT kernel runs 1024 threads for every R.
#pragma once

#include "fft128.hpp"

template <int R, int NTHR = 1024> struct R16 {
    static constexpr int N = 16 * R, P = N + 1, T = N * N / 16;
    static constexpr int SUB = NTHR / T;                      // whole thread groups per workgroup
    static constexpr int REGIONS = (NTHR + T - 1) / T;        // LDS regions (partial group included)
    static constexpr size_t SLICE_BYTES = (size_t)N * P * sizeof(float2);
};

// threads of a class's own R16 launch: the whole thread groups that fit 1024
template <int R> __host__ __device__ constexpr int r16_threads() { return (1024 / R16<R>::T) * R16<R>::T; }

// LDS of an NTHR-thread workgroup running R16 slices: the regions plus the twiddle table W_N^t
template <int R, int NTHR = 1024> __host__ __device__ constexpr size_t r16_lds_bytes()
{
    return R16<R, NTHR>::REGIONS * R16<R, NTHR>::SLICE_BYTES + (size_t)R16<R>::N * sizeof(float2);
}

// one pass (rows or columns) of N-point DFTs over the slice at buf
template <int R, bool ROWS>
__device__ __forceinline__ void r16_pass(float2 *buf, const float2 *twl, int tl)
{
    constexpr int N = R16<R>::N, P = R16<R>::P, T = R16<R>::T;
    auto at = [&](int line, int pos) -> float2 & { return ROWS ? buf[line * P + pos] : buf[pos * P + line]; };
    if constexpr (R > 1) {
        for (int it = tl; it < 16 * N; it += T) {
            const int line = ROWS ? it >> 4 : it % N, j = ROWS ? it & 15 : it / N;
            float2 v[R];
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = at(line, j + 16 * r);
            dft_small<float2, R>(v, twl, N);
#pragma unroll
            for (int s = 1; s < R; s++) v[s] = cmul(v[s], twl[j * s]);   // j s < N
#pragma unroll
            for (int s = 0; s < R; s++) at(line, j + 16 * s) = v[s];
        }
    }
    __syncthreads();
    {
        const int line = tl % N, s = tl / N;   // tl < T = R N: one 16-point DFT per thread
        float2 v[16];
#pragma unroll
        for (int q = 0; q < 16; q++) v[q] = at(line, 16 * s + q);
        __syncthreads();   // the strided writes below land in other threads' inputs
        dft16(v);
#pragma unroll
        for (int q = 0; q < 16; q++) at(line, s + R * q) = v[q];
    }
    __syncthreads();
}

// Forward: slice u (beta index k = u / nfc, function fi = u % nfc) of the class, grid -> the
// 4-pair records of its B^2 groups in the workspace (ws_rec, two float4 per record).  Real
// functions go in pairs (k = u / nfp, functions fi = 2 (u % nfp) and fi + 1): z = a + i b is one
// complex slice, and since X(-p) = conj X(p) for a real slice, X_a(p) = (Z(p) + conj Z(-p)) / 2 and
// X_b(p) = -i (Z(p) - conj Z(-p)) / 2; a record holds the bins p and -p of both its couples.
template <int R, bool REAL>
__device__ __forceinline__ void r16_fwd_slice(const FftArgs<float> &a, int u, bool active, float2 *buf,
                                              const float2 *twl, int tl)
{
    constexpr int N = R16<R>::N, P = R16<R>::P, T = R16<R>::T, NN = N * N, B = N / 2;
    constexpr int PER = NN / T;                                  // 16 elements per thread
    const ChunkArgs &ca = a.ca;
    const int nf = REAL ? (ca.nfc + 1) / 2 : ca.nfc;
    const int k = u / nf, fi = REAL ? 2 * (u - k * nf) : u - k * nf;
    const bool two = REAL && fi + 1 < ca.nfc;
    if (active) {
        const long base = ca.grid_off[ca.fn[fi]] + (long)k * NN;
        const long base2 = two ? ca.grid_off[ca.fn[fi + 1]] + (long)k * NN : base;
        float2 v[PER];
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int e = tl + i * T;
            if (REAL) {
                const float *g = reinterpret_cast<const float *>(a.src);
                v[i] = make_float2(g[base + e], two ? g[base2 + e] : 0.f);
            } else v[i] = reinterpret_cast<const float2 *>(a.src)[base + e];
        }
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int e = tl + i * T;
            buf[(e / N) * P + e % N] = v[i];
        }
    }
    __syncthreads();
    r16_pass<R, true>(buf, twl, tl);
    r16_pass<R, false>(buf, twl, tl);
    if (active) {
        float4 *S4 = reinterpret_cast<float4 *>(a.dst);
        for (int g = tl; g < B * B; g += T) {
            int m0, n0;
            so3_group(g, m0, n0);
            // S_k(m, n) = X(-m mod N, -n mod N) for the pairs (m0,n0), (-m0,-n0), (-n0,-m0), (n0,m0)
            const int um = m0 ? N - m0 : 0, un = n0 > 0 ? N - n0 : -n0, vm = m0, vn = n0 >= 0 ? n0 : N + n0;
            const float2 p0 = buf[um * P + un], p1 = buf[vm * P + vn];
            const float2 p2 = buf[vn * P + vm], p3 = buf[un * P + um];
            const long o = ws_rec(ca, g, k, fi) * 2;
            if (REAL) {
                auto sep = [](float2 z, float2 w, float2 &xa, float2 &xb) {   // w = Z(-p)
                    xa = make_float2(0.5f * (z.x + w.x), 0.5f * (z.y - w.y));
                    xb = make_float2(0.5f * (z.y + w.y), 0.5f * (w.x - z.x));
                };
                float2 a0, b0, a1, b1, a2, b2, a3, b3;
                sep(p0, p1, a0, b0);
                sep(p1, p0, a1, b1);
                sep(p2, p3, a2, b2);
                sep(p3, p2, a3, b3);
                S4[o] = make_float4(a0.x, a0.y, a1.x, a1.y);
                S4[o + 1] = make_float4(a2.x, a2.y, a3.x, a3.y);
                if (two) {
                    const long o2 = ws_rec(ca, g, k, fi + 1) * 2;
                    S4[o2] = make_float4(b0.x, b0.y, b1.x, b1.y);
                    S4[o2 + 1] = make_float4(b2.x, b2.y, b3.x, b3.y);
                }
            } else {
                S4[o] = make_float4(p0.x, p0.y, p1.x, p1.y);
                S4[o + 1] = make_float4(p2.x, p2.y, p3.x, p3.y);
            }
        }
    }
}

// Inverse: gather the records of slice u into bins (m mod N, n mod N), 2-D DFT, store the slice.
// Real functions go in pairs as in the forward: the bins of X_a + i X_b transform to the grid of
// function fi in the real part and that of fi + 1 in the imaginary part.
template <int R, bool REAL>
__device__ __forceinline__ void r16_inv_slice(const FftArgs<float> &a, int u, bool active, float2 *buf,
                                              const float2 *twl, int tl)
{
    constexpr int N = R16<R>::N, P = R16<R>::P, T = R16<R>::T, NN = N * N, B = N / 2;
    constexpr int PER = NN / T;
    const ChunkArgs &ca = a.ca;
    const int nf = REAL ? (ca.nfc + 1) / 2 : ca.nfc;
    const int k = u / nf, fi = REAL ? 2 * (u - k * nf) : u - k * nf;
    const bool two = REAL && fi + 1 < ca.nfc;
    // the Nyquist row and column (m or n = +-B) are zero; every other bin holds one pair
    for (int i = tl; i < N; i += T) {
        buf[B * P + i] = make_float2(0.f, 0.f);
        buf[i * P + B] = make_float2(0.f, 0.f);
    }
    const float4 *G4 = reinterpret_cast<const float4 *>(a.src);
    for (int g = tl; g < B * B; g += T) {
        int m0, n0;
        so3_group(g, m0, n0);
        float4 q0 = make_float4(0.f, 0.f, 0.f, 0.f), q1 = q0;
        if (active) {
            const long o = ws_rec(ca, g, k, fi) * 2;
            q0 = G4[o];
            q1 = G4[o + 1];
            if (two) {   // z = x_a + i x_b
                const long o2 = ws_rec(ca, g, k, fi + 1) * 2;
                const float4 r0 = G4[o2], r1 = G4[o2 + 1];
                q0 = make_float4(q0.x - r0.y, q0.y + r0.x, q0.z - r0.w, q0.w + r0.z);
                q1 = make_float4(q1.x - r1.y, q1.y + r1.x, q1.z - r1.w, q1.w + r1.z);
            }
        }
        const int vm = m0, vn = n0 >= 0 ? n0 : N + n0, um = m0 ? N - m0 : 0, un = n0 > 0 ? N - n0 : -n0;
        buf[vm * P + vn] = make_float2(q0.x, q0.y);                    // (m0, n0)
        if (m0 > 0) {
            buf[um * P + un] = make_float2(q0.z, q0.w);                // (-m0, -n0)
            if (n0 != m0 && n0 != -m0) {
                buf[un * P + um] = make_float2(q1.x, q1.y);            // (-n0, -m0)
                buf[vn * P + vm] = make_float2(q1.z, q1.w);            // (n0, m0)
            }
        }
    }
    __syncthreads();
    r16_pass<R, true>(buf, twl, tl);
    r16_pass<R, false>(buf, twl, tl);
    if (active) {
        const long base = ca.grid_off[ca.fn[fi]] + (long)k * NN;
        const long base2 = two ? ca.grid_off[ca.fn[fi + 1]] + (long)k * NN : base;
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int e = tl + i * T;
            const float2 v = buf[(e / N) * P + e % N];
            if (REAL) {
                reinterpret_cast<float *>(a.dst)[base + e] = v.x;
                if (two) reinterpret_cast<float *>(a.dst)[base2 + e] = v.y;
            } else reinterpret_cast<float2 *>(a.dst)[base + e] = v;
        }
    }
}

// the thread group, region and slice of thread tid for an item whose first slice is u0
template <int R, int NTHR>
__device__ __forceinline__ void r16_role(int tid, int u0, int nslice, float2 *regions, float2 *&buf, int &tl, int &u,
                                         bool &active)
{
    const int sub = tid / R16<R>::T;
    tl = tid - sub * R16<R>::T;
    buf = regions + (size_t)sub * (R16<R>::SLICE_BYTES / sizeof(float2));
    u = u0 + sub;
    active = sub < R16<R, NTHR>::SUB && u < nslice;
}

// Slices of a class: N per function, or per pair of functions for real functions (two real slices
// run as one complex DFT).
template <int R> __host__ __device__ inline int r16_slices(int nfc, bool real)
{
    return (real ? (nfc + 1) / 2 : nfc) * R16<R>::N;
}

// One class's slices through NTHR-thread R16 workgroups: item i covers slices [i SUB, (i+1) SUB).
template <int R, int NTHR = 1024> __host__ __device__ inline int r16_units(int nfc, bool real)
{
    return (r16_slices<R>(nfc, real) + R16<R, NTHR>::SUB - 1) / R16<R, NTHR>::SUB;
}

template <int R, bool FWD, bool REAL, int NTHR = 1024>
__device__ __forceinline__ void r16_item(const FftArgs<float> &a, int item, unsigned char *smem)
{
    float2 *regions = reinterpret_cast<float2 *>(smem);
    float2 *twl = regions + R16<R, NTHR>::REGIONS * (R16<R>::SLICE_BYTES / sizeof(float2));
    // opaque thread index: in a persistent loop over items the compiler must not hoist the
    // per-lane index arithmetic of every R out of the loop (it would hold it all in registers)
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    if (tid < R16<R>::N) twl[tid] = a.tw[tid];
    __syncthreads();
    float2 *buf;
    int tl, u;
    bool active;
    r16_role<R, NTHR>(tid, item * R16<R, NTHR>::SUB, r16_slices<R>(a.ca.nfc, REAL), regions, buf, tl, u, active);
    if (FWD) r16_fwd_slice<R, REAL>(a, u, active, buf, twl, tl);
    else r16_inv_slice<R, REAL>(a, u, active, buf, twl, tl);
}

template <int R, bool FWD, bool REAL>
__global__ void __launch_bounds__(r16_threads<R>()) so3_fft_r16(FftArgs<float> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n = r16_units<R, r16_threads<R>()>(a.ca.nfc, REAL);
    const int item = xcd_unit(n);
    if (item >= n) return;
    r16_item<R, FWD, REAL, r16_threads<R>()>(a, item, smem_raw);
}
