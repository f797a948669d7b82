// fft128.hpp -- 2-D DFT of one 128 x 128 complex-float beta slice (the l_max = 63 grid), gfx950.
//
// 128 = 8 x 16: stage A is a radix-8 butterfly over elements j + 16 r (r < 8) followed by the
// twiddle W_128^(j s), written back in place; stage B is a 16-point DFT of the 16 contiguous
// results y[16 s + j], done entirely in registers, which leaves X[s + 8 q] in natural order (a
// barrier separates stage B's reads from its strided writes, which land in other threads' input).  Rows (gamma) first, then
// columns (alpha).  The slice lives in LDS with a row pitch of 129 elements so that every stage's
// lane pattern is conflict free; the forward kernel's row stage A reads the grid straight from
// HBM.  One workgroup of 1024 threads per slice.
#pragma once

#include "kernels.hpp"

#define F128 128
#define P128 129  // LDS row pitch (elements)

// 16-point DFT in registers: v[q] <- sum_j v[j] W16^(j q), as 4 x 4 (radix-4, twiddle, radix-4)
__device__ __forceinline__ void dft16(float2 *v)
{
    const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f, h = 0.70710678118654752440f;
    // stage 1: four radix-4 DFTs over j = j1 + 4 j2 (j2 summed), index j1
    float2 a[16];
#pragma unroll
    for (int j1 = 0; j1 < 4; j1++) {
        float2 t[4] = {v[j1], v[j1 + 4], v[j1 + 8], v[j1 + 12]};
        dft4(t);
#pragma unroll
        for (int q2 = 0; q2 < 4; q2++) a[j1 * 4 + q2] = t[q2];  // a[j1][q2]
    }
    // twiddle W16^(j1 q2)
    const float2 w[10] = {
        {1.f, 0.f},  {c1, -s1}, {h, -h},   {s1, -c1},  // W^0..W^3
        {0.f, -1.f}, {-s1, -c1}, {-h, -h}, {-c1, -s1}, // W^4..W^7
        {-1.f, 0.f}, {-c1, s1}                          // W^8, W^9
    };
#pragma unroll
    for (int j1 = 1; j1 < 4; j1++)
#pragma unroll
        for (int q2 = 1; q2 < 4; q2++) a[j1 * 4 + q2] = cmul(a[j1 * 4 + q2], w[j1 * q2]);
    // stage 2: radix-4 over j1 for each q2; q = q2 + 4 q1
#pragma unroll
    for (int q2 = 0; q2 < 4; q2++) {
        float2 t[4] = {a[q2], a[4 + q2], a[8 + q2], a[12 + q2]};
        dft4(t);
#pragma unroll
        for (int q1 = 0; q1 < 4; q1++) v[q2 + 4 * q1] = t[q1];
    }
}

// radix-8 stage A on 8 values x[j + 16 r], then twiddle W128^(j s) from the LDS table
__device__ __forceinline__ void stageA(float2 *v, int j, const float2 *twl)
{
    dft8(v);
#pragma unroll
    for (int s = 1; s < 8; s++) v[s] = cmul(v[s], twl[(s * j) & 127]);
}

// Row pass over a slice already in LDS (stage A reads LDS).
__device__ __forceinline__ void rows_from_lds(float2 *buf, const float2 *twl, int tid)
{
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int t = tid + i * 1024, row = t >> 4, j = t & 15;
        float2 v[8];
#pragma unroll
        for (int r = 0; r < 8; r++) v[r] = buf[row * P128 + j + 16 * r];
        stageA(v, j, twl);
#pragma unroll
        for (int s = 0; s < 8; s++) buf[row * P128 + j + 16 * s] = v[s];
    }
}

__device__ __forceinline__ void rows_stageB(float2 *buf, int tid)
{
    const int row = tid & 127, s = tid >> 7;
    float2 v[16];
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = buf[row * P128 + 16 * s + q];
    __syncthreads();  // the strided write-back below overlaps other threads' contiguous reads
    dft16(v);
#pragma unroll
    for (int q = 0; q < 16; q++) buf[row * P128 + s + 8 * q] = v[q];
}

// Column pass whose stage B hands X[s + 8q][col] to out(row, col, value) instead of writing LDS
// (lanes of a wave hold 64 consecutive columns of one row: the inverse writes them straight to
// the grid).
template <typename F>
__device__ __forceinline__ void cols_pass_out(float2 *buf, const float2 *twl, int tid, F out)
{
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int t = tid + i * 1024, col = t & 127, j = t >> 7;
        float2 v[8];
#pragma unroll
        for (int r = 0; r < 8; r++) v[r] = buf[(j + 16 * r) * P128 + col];
        stageA(v, j, twl);
#pragma unroll
        for (int s = 0; s < 8; s++) buf[(j + 16 * s) * P128 + col] = v[s];
    }
    __syncthreads();
    const int col = tid & 127, s = tid >> 7;
    float2 v[16];
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = buf[(16 * s + q) * P128 + col];
    dft16(v);
#pragma unroll
    for (int q = 0; q < 16; q++) out(s + 8 * q, col, v[q]);
}

__device__ __forceinline__ void cols_pass(float2 *buf, const float2 *twl, int tid)
{
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int t = tid + i * 1024, col = t & 127, j = t >> 7;
        float2 v[8];
#pragma unroll
        for (int r = 0; r < 8; r++) v[r] = buf[(j + 16 * r) * P128 + col];
        stageA(v, j, twl);
#pragma unroll
        for (int s = 0; s < 8; s++) buf[(j + 16 * s) * P128 + col] = v[s];
    }
    __syncthreads();
    {
        const int col = tid & 127, s = tid >> 7;
        float2 v[16];
#pragma unroll
        for (int q = 0; q < 16; q++) v[q] = buf[(16 * s + q) * P128 + col];
        __syncthreads();  // as in rows_stageB: reads and strided writes overlap across threads
        dft16(v);
#pragma unroll
        for (int q = 0; q < 16; q++) buf[(s + 8 * q) * P128 + col] = v[q];
    }
    __syncthreads();
}

// 2-D forward DFT of slice (k, fi) into LDS (so3_fft128_fwd, so3_fft128_fwd2), in three parts:
// the grid loads into registers, row stage A into LDS, and the remaining passes.
// This thread's two stage A butterflies.  Complex: row tid / 8, columns j0 + 16r and j0 + 1 + 16r
// (j0 = 2 (tid % 8)), so each 16-byte load brings both butterflies' inputs of one r (measured
// 1.97 -> 1.89 ms at l_max = 63 x 256).  Real: rows tid / 16 and 64 + tid / 16, column tid % 16
// (4-byte loads).
template <bool REAL>
__device__ __forceinline__ void fft128_load(const FftArgs<float> &a, int k, int fi, int tid, float2 (&v)[2][8])
{
    const ChunkArgs &ca = a.ca;
    const long base = ca.grid_off[ca.fn[fi]] + (long)k * F128 * F128;
    const int row0 = REAL ? tid >> 4 : tid >> 3, j0 = REAL ? tid & 15 : 2 * (tid & 7);
    const int drow = REAL ? 64 : 0;
#pragma unroll
    for (int r = 0; r < 8; r++) {
        if (REAL) {
            const float *src = reinterpret_cast<const float *>(a.src) + base + j0 + 16 * r;
            v[0][r] = make_float2(src[row0 * F128], 0.f);
            v[1][r] = make_float2(src[(row0 + drow) * F128], 0.f);
        } else {
            const float4 x = *reinterpret_cast<const float4 *>(reinterpret_cast<const float2 *>(a.src) + base +
                                                               row0 * F128 + j0 + 16 * r);
            v[0][r] = make_float2(x.x, x.y);
            v[1][r] = make_float2(x.z, x.w);
        }
    }
}

template <bool REAL>
__device__ __forceinline__ void fft128_stageA_lds(float2 (&v)[2][8], float2 *buf, const float2 *twl, int tid)
{
    const int row0 = REAL ? tid >> 4 : tid >> 3, j0 = REAL ? tid & 15 : 2 * (tid & 7);
    const int drow = REAL ? 64 : 0, dj = REAL ? 0 : 1;
    __syncthreads();  // twiddle table / previous slice's LDS reads
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int row = row0 + i * drow, j = j0 + i * dj;
        stageA(v[i], j, twl);
#pragma unroll
        for (int s = 0; s < 8; s++) buf[row * P128 + j + 16 * s] = v[i][s];
    }
    __syncthreads();
}

__device__ __forceinline__ void fft128_rest(float2 *buf, const float2 *twl, int tid)
{
    rows_stageB(buf, tid);
    __syncthreads();
    cols_pass(buf, twl, tid);
}

template <bool REAL>
__device__ __forceinline__ void fft128_fwd_slice(const FftArgs<float> &a, float2 *buf, const float2 *twl, int k,
                                                 int fi, int tid)
{
    float2 v[2][8];
    fft128_load<REAL>(a, k, fi, tid, v);
    fft128_stageA_lds<REAL>(v, buf, twl, tid);
    fft128_rest(buf, twl, tid);
}

// the 4 pairs of group g of the spectrum in LDS, as two float4 (one 32-byte record)
__device__ __forceinline__ void fft128_record(const float2 *buf, int g, float4 &h0, float4 &h1)
{
    int m0, n0;
    so3_group(g, m0, n0);
    const int um = (-m0) & 127, un = (-n0) & 127, vm = m0 & 127, vn = n0 & 127;
    const float2 p0 = buf[um * P128 + un], p1 = buf[vm * P128 + vn];
    const float2 p2 = buf[vn * P128 + vm], p3 = buf[un * P128 + um];
    h0 = make_float4(p0.x, p0.y, p1.x, p1.y);
    h1 = make_float4(p2.x, p2.y, p3.x, p3.y);
}

template <bool REAL>
__global__ void __launch_bounds__(1024) so3_fft128_fwd(FftArgs<float> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2 *buf = reinterpret_cast<float2 *>(smem_raw);
    float2 *twl = buf + F128 * P128;
    const ChunkArgs &ca = a.ca;
    const int nslice = ca.nfc * F128;
    const int unit = xcd_unit(nslice);
    if (unit >= nslice) return;
    const int k = unit / ca.nfc, fi = unit - k * ca.nfc;
    const int tid = threadIdx.x;
    if (tid < F128) twl[tid] = a.tw[tid];
    fft128_fwd_slice<REAL>(a, buf, twl, k, fi, tid);
    // emit S_k(m,n) = X(-m mod 128, -n mod 128) for the 4 pairs of every group (4096 groups):
    // the 4 pairs are contiguous, two 16-byte stores
    float4 *S4 = reinterpret_cast<float4 *>(a.dst);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int g = tid + i * 1024;
        float4 h0, h1;
        fft128_record(buf, g, h0, h1);
        const long o = ws_rec(ca, g, k, fi) * 2;
        S4[o] = h0;
        S4[o + 1] = h1;
    }
}

template <bool REAL>
__global__ void __launch_bounds__(1024) so3_fft128_inv(FftArgs<float> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2 *buf = reinterpret_cast<float2 *>(smem_raw);
    float2 *twl = buf + F128 * P128;
    const ChunkArgs &ca = a.ca;
    const int nslice = ca.nfc * F128;
    const int unit = xcd_unit(nslice);
    if (unit >= nslice) return;
    const int k = unit / ca.nfc, fi = unit - k * ca.nfc;
    const int tid = threadIdx.x;
    if (tid < F128) twl[tid] = a.tw[tid];
    // gather the 4 pairs of every group, scatter to bins (m mod 128, n mod 128)
    float4 q[4][2];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int g = tid + i * 1024;
        const float4 *G4 = reinterpret_cast<const float4 *>(a.src);
        const long o = ws_rec(ca, g, k, fi) * 2;
        q[i][0] = G4[o];
        q[i][1] = G4[o + 1];
    }
    // Nyquist row and column (m or n = +-64) are zero
    if (tid < F128) {
        buf[64 * P128 + tid] = make_float2(0.f, 0.f);
        buf[tid * P128 + 64] = make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int g = tid + i * 1024;
        int m0, n0;
        so3_group(g, m0, n0);
        const int vm = m0 & 127, vn = n0 & 127, um = (-m0) & 127, un = (-n0) & 127;
        buf[vm * P128 + vn] = make_float2(q[i][0].x, q[i][0].y);                    // (m0, n0)
        if (m0 > 0) {
            buf[um * P128 + un] = make_float2(q[i][0].z, q[i][0].w);                // (-m0, -n0)
            if (n0 != m0 && n0 != -m0) {
                buf[un * P128 + um] = make_float2(q[i][1].x, q[i][1].y);            // (-n0, -m0)
                buf[vn * P128 + vm] = make_float2(q[i][1].z, q[i][1].w);            // (n0, m0)
            }
        }
    }
    __syncthreads();
    rows_from_lds(buf, twl, tid);
    __syncthreads();
    rows_stageB(buf, tid);
    __syncthreads();
    // last stage straight to the grid: each store instruction covers 64 consecutive gamma columns
    const long base = ca.grid_off[ca.fn[fi]] + (long)k * F128 * F128;
    if (REAL) {
        float *out = reinterpret_cast<float *>(a.dst) + base;
        cols_pass_out(buf, twl, tid, [&](int row, int col, float2 v) { out[row * F128 + col] = v.x; });
    } else {
        float2 *out = reinterpret_cast<float2 *>(a.dst) + base;
        cols_pass_out(buf, twl, tid, [&](int row, int col, float2 v) { out[row * F128 + col] = v; });
    }
}

// ---------------------------------------------------------------------------------------------
// S^2, N = 256 (l_max = 127), FP32: the length-256 DFT of each ring as 16 x 16.  A workgroup of
// 256 threads takes 16 rings: 2 beta rows k of 8 consecutive functions, 16 threads per ring, so
// the DFT's exchanges stay inside a wave.  Stage A: thread j holds x[j + 16t] (t < 16), a
// 16-point DFT in registers and the twiddle W_256^(j s); stage B: thread s takes the 16 values
// a_s(j) from LDS (row pitch 17: no bank conflicts), a 16-point DFT, and holds X[s + 16q].
// The 8 functions of a workgroup own whole 128-byte lines of every workspace row, so the
// spectrum moves as full lines: the forward stages X in LDS and writes (X[g], X[256 - g])
// records 8 functions at a time; the inverse reads records the same way into LDS.
// ---------------------------------------------------------------------------------------------
#define S2F_ROWS 16
#define S2F_PITCH 17
#define S2F_FNS 8                      // functions per workgroup
#define S2F_KS (S2F_ROWS / S2F_FNS)    // beta rows per workgroup

// stage A from registers v (x[j + 16t]), then stage B: v[q] = X[s + 16q] with s = j
__device__ __forceinline__ void s2f_256(float2 *v, float2 *rowbuf, const float2 *twl, int j)
{
    dft16(v);
#pragma unroll
    for (int s = 1; s < 16; s++) v[s] = cmul(v[s], twl[(j * s) & 255]);
#pragma unroll
    for (int s = 0; s < 16; s++) rowbuf[s * S2F_PITCH + j] = v[s];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; q++) v[q] = rowbuf[j * S2F_PITCH + q];
    dft16(v);
}

// workgroup -> (first beta row k0, first function f0); consecutive workgroups: consecutive functions
__device__ __forceinline__ bool s2f_unit(const ChunkArgs &ca, int &k0, int &f0)
{
    const int nfb = (ca.nfc + S2F_FNS - 1) / S2F_FNS, nunit = nfb * (256 / S2F_KS);
    const int unit = xcd_unit(nunit);
    if (unit >= nunit) return false;
    f0 = (unit % nfb) * S2F_FNS;
    k0 = (unit / nfb) * S2F_KS;
    return true;
}

template <bool REAL>
__global__ void __launch_bounds__(256) s2_fft256_fwd(FftArgs<float> a)
{
    __shared__ float2 twl[256];
    __shared__ float2 buf[S2F_ROWS * 16 * S2F_PITCH];
    const ChunkArgs &ca = a.ca;
    int k0, f0;
    if (!s2f_unit(ca, k0, f0)) return;
    const int tid = threadIdx.x, rr = tid >> 4, j = tid & 15;
    const int k = k0 + rr / S2F_FNS, fi = f0 + rr % S2F_FNS;   // this thread's ring
    const bool live = fi < ca.nfc;
    twl[tid] = a.tw[tid];
    const long base = live ? ca.grid_off[ca.fn[fi]] + (long)k * 256 : 0;
    float2 v[16];
#pragma unroll
    for (int t = 0; t < 16; t++) {
        if (!live) v[t] = make_float2(0.f, 0.f);
        else if (REAL) v[t] = make_float2(reinterpret_cast<const float *>(a.src)[base + j + 16 * t], 0.f);
        else v[t] = reinterpret_cast<const float2 *>(a.src)[base + j + 16 * t];
    }
    __syncthreads();  // twiddle table
    float2 *rowbuf = buf + rr * 16 * S2F_PITCH;
    s2f_256(v, rowbuf, twl, j);
    __syncthreads();  // stage B's reads of the row buffers
#pragma unroll
    for (int q = 0; q < 16; q++) rowbuf[q * S2F_PITCH + j] = v[q];   // X[j + 16q]
    __syncthreads();
    // S_k(m) = X(m mod 256), |m| <= 127 (src/fdcl_FFTS2.cpp:159-166): group g = |m| holds the
    // record (X[g], X[256 - g]); lanes 8i..8i+7 write the 8 functions of one (k, g): a whole line
    float4 *S4 = reinterpret_cast<float4 *>(a.dst);
    const int fl = tid % S2F_FNS;
#pragma unroll
    for (int i = 0; i < S2F_KS * 128 / 32; i++) {
        const int idx = tid / S2F_FNS + 32 * i, kl = idx >> 7, g = idx & 127, h = (256 - g) & 255;
        const float2 *rb = buf + (kl * S2F_FNS + fl) * 16 * S2F_PITCH;
        const float2 x0 = rb[(g >> 4) * S2F_PITCH + (g & 15)], x1 = rb[(h >> 4) * S2F_PITCH + (h & 15)];
        if (f0 + fl < ca.nfc)
            S4[ws_at(ca, g, k0 + kl, 0, f0 + fl, 2) >> 1] = make_float4(x0.x, x0.y, x1.x, x1.y);
    }
}

template <bool REAL>
__global__ void __launch_bounds__(256) s2_fft256_inv(FftArgs<float> a)
{
    __shared__ float2 twl[256];
    __shared__ float2 buf[S2F_ROWS * 16 * S2F_PITCH];
    const ChunkArgs &ca = a.ca;
    int k0, f0;
    if (!s2f_unit(ca, k0, f0)) return;
    const int tid = threadIdx.x, rr = tid >> 4, j = tid & 15;
    twl[tid] = a.tw[tid];
    // records (G(g), G(-g)) of the 8 functions of each (k, g), whole lines; bin u of a ring holds
    // G(m) with -m = u (mod 256): G(g) at bin (256 - g) & 255, G(-g) at bin g; bin 128 is empty
    const float4 *G4 = reinterpret_cast<const float4 *>(a.src);
    const int fl = tid % S2F_FNS;
#pragma unroll
    for (int i = 0; i < S2F_KS * 128 / 32; i++) {
        const int idx = tid / S2F_FNS + 32 * i, kl = idx >> 7, g = idx & 127;
        const float4 r = f0 + fl < ca.nfc ? G4[ws_at(ca, g, k0 + kl, 0, f0 + fl, 2) >> 1] : make_float4(0.f, 0.f, 0.f, 0.f);
        float2 *rb = buf + (kl * S2F_FNS + fl) * 16 * S2F_PITCH;
        const int u0 = (256 - g) & 255;
        rb[(u0 >> 4) * S2F_PITCH + (u0 & 15)] = make_float2(r.x, r.y);
        if (g > 0) rb[(g >> 4) * S2F_PITCH + (g & 15)] = make_float2(r.z, r.w);
    }
    if (tid < S2F_ROWS) buf[tid * 16 * S2F_PITCH + 8 * S2F_PITCH] = make_float2(0.f, 0.f);   // bin 128
    __syncthreads();
    float2 *rowbuf = buf + rr * 16 * S2F_PITCH;
    float2 v[16];
#pragma unroll
    for (int t = 0; t < 16; t++) v[t] = rowbuf[t * S2F_PITCH + j];   // bin j + 16t
    __syncthreads();  // every ring's bins read before stage A overwrites the buffers
    s2f_256(v, rowbuf, twl, j);
    const int k = k0 + rr / S2F_FNS, fi = f0 + rr % S2F_FNS;
    if (fi < ca.nfc) {
        const long base = ca.grid_off[ca.fn[fi]] + (long)k * 256;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            if (REAL) reinterpret_cast<float *>(a.dst)[base + j + 16 * q] = v[q].x;
            else reinterpret_cast<float2 *>(a.dst)[base + j + 16 * q] = v[q];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// SO(3), N = 128 forward, two functions per workgroup (workspace with 2 groups interleaved,
// gil = 1): the two slices (k, fi) and (k, fi + 1) share every 128-byte workspace line (64 bytes
// each), so one workgroup transforms both in turn, keeps the first slice's group records in
// registers until the second's are ready, and writes each line whole.
// ---------------------------------------------------------------------------------------------

template <bool REAL, bool OPAQUE = false>
__device__ __forceinline__ void so3_fft128_fwd2_body(const FftArgs<float> &a, int unit)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2 *buf = reinterpret_cast<float2 *>(smem_raw);
    float2 *twl = buf + F128 * P128;
    const ChunkArgs &ca = a.ca;
    const int nfp = (ca.nfc + 1) / 2, nunit = nfp * F128;
    if (unit < 0 || unit >= nunit) return;
    const int k = unit / nfp, fi = 2 * (unit - k * nfp);
    const bool two = fi + 1 < ca.nfc;
    int tid = threadIdx.x;
    // in a persistent loop over units (OPAQUE) the per-lane index arithmetic stays inside the
    // iteration instead of being hoisted and held in registers across it
    if (OPAQUE) asm volatile("" : "+v"(tid));
    if (tid < F128) twl[tid] = a.tw[tid];
    float2 vA[2][8], vB[2][8];
    if (REAL && two) {
        // two real slices in one complex DFT: z = a + i b, Z = X_a + i X_b, and since X(-p) =
        // conj X(p) for a real slice, X_a(p) = (Z(p) + conj Z(-p)) / 2 and X_b(p) = -i (Z(p) -
        // conj Z(-p)) / 2.  A record holds the bins p and -p of both its pair couples, so its four
        // reads of Z give both functions' records.
        fft128_load<true>(a, k, fi, tid, vA);
        fft128_load<true>(a, k, fi + 1, tid, vB);
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int r = 0; r < 8; r++) vA[i][r].y = vB[i][r].x;
        fft128_stageA_lds<true>(vA, buf, twl, tid);
        fft128_rest(buf, twl, tid);
        float4 *S4 = reinterpret_cast<float4 *>(a.dst);
        auto sep = [](float2 z, float2 w, float2 &xa, float2 &xb) {   // w = Z(-p)
            xa = make_float2(0.5f * (z.x + w.x), 0.5f * (z.y - w.y));
            xb = make_float2(0.5f * (z.y + w.y), 0.5f * (w.x - z.x));
        };
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int g = tid + i * 1024;
            int m0, n0;
            so3_group(g, m0, n0);
            const int um = (-m0) & 127, un = (-n0) & 127, vm = m0 & 127, vn = n0 & 127;
            const float2 p0 = buf[um * P128 + un], p1 = buf[vm * P128 + vn];
            const float2 p2 = buf[vn * P128 + vm], p3 = buf[un * P128 + um];
            float2 a0, b0, a1, b1, a2, b2, a3, b3;
            sep(p0, p1, a0, b0);
            sep(p1, p0, a1, b1);
            sep(p2, p3, a2, b2);
            sep(p3, p2, a3, b3);
            const long o = ws_rec(ca, g, k, fi) * 2;
            S4[o] = make_float4(a0.x, a0.y, a1.x, a1.y);
            S4[o + 1] = make_float4(a2.x, a2.y, a3.x, a3.y);
            S4[o + 4] = make_float4(b0.x, b0.y, b1.x, b1.y);
            S4[o + 5] = make_float4(b2.x, b2.y, b3.x, b3.y);
        }
        return;
    }
    fft128_load<REAL>(a, k, fi, tid, vA);
    fft128_stageA_lds<REAL>(vA, buf, twl, tid);
    // the second slice's loads go out once the first slice's inputs are in LDS, and are in flight
    // during its remaining passes (1.89 -> 1.83 ms at l_max = 63 x 256; issuing them together with
    // the first slice's loads was slower)
    if (two) fft128_load<REAL>(a, k, fi + 1, tid, vB);
    fft128_rest(buf, twl, tid);
    float4 r0[4][2];
#pragma unroll
    for (int i = 0; i < 4; i++) fft128_record(buf, tid + i * 1024, r0[i][0], r0[i][1]);
    float4 *S4 = reinterpret_cast<float4 *>(a.dst);
    if (!two) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const long o = ws_rec(ca, tid + i * 1024, k, fi) * 2;
            S4[o] = r0[i][0];
            S4[o + 1] = r0[i][1];
        }
        return;
    }
    fft128_stageA_lds<REAL>(vB, buf, twl, tid);
    fft128_rest(buf, twl, tid);
    int tl = tid;
    asm volatile("" : "+v"(tl));  // decode the groups again rather than keep them live across the second slice
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int g = tl + i * 1024;
        float4 h0, h1;
        fft128_record(buf, g, h0, h1);
        // function fi's record of g, then function fi + 1's two records later (2 groups per function)
        const long o = ws_rec(ca, g, k, fi) * 2;
        S4[o] = r0[i][0];
        S4[o + 1] = r0[i][1];
        S4[o + 4] = h0;
        S4[o + 5] = h1;
    }
}

// units of so3_fft128_fwd2: (beta index, pair of functions)
__host__ __device__ inline int so3_fft128_fwd2_units(int nfc) { return (nfc + 1) / 2 * F128; }

template <bool REAL>
__global__ void __launch_bounds__(1024) so3_fft128_fwd2(FftArgs<float> a)
{
    so3_fft128_fwd2_body<REAL>(a, xcd_unit(so3_fft128_fwd2_units(a.ca.nfc)));
}

// Inverse with the same pairing: both slices' records are gathered at once (whole lines), then
// each slice is scattered to bins, transformed and written in turn.
// the bins of one slice from its records (4 groups per thread); row and column 64 are zero
__device__ __forceinline__ void fft128_inv_scatter(float2 *buf, const float4 (&q)[4][2], int tid)
{
    if (tid < F128) {
        buf[64 * P128 + tid] = make_float2(0.f, 0.f);
        buf[tid * P128 + 64] = make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int g = tid + i * 1024;
        int m0, n0;
        so3_group(g, m0, n0);
        const int vm = m0 & 127, vn = n0 & 127, um = (-m0) & 127, un = (-n0) & 127;
        buf[vm * P128 + vn] = make_float2(q[i][0].x, q[i][0].y);
        if (m0 > 0) {
            buf[um * P128 + un] = make_float2(q[i][0].z, q[i][0].w);
            if (n0 != m0 && n0 != -m0) {
                buf[un * P128 + um] = make_float2(q[i][1].x, q[i][1].y);
                buf[vn * P128 + vm] = make_float2(q[i][1].z, q[i][1].w);
            }
        }
    }
}

template <bool REAL>
__device__ __forceinline__ void fft128_inv_slice(const FftArgs<float> &a, float2 *buf, const float2 *twl,
                                                 const float4 (&q)[4][2], int k, int fi, int tid)
{
    const ChunkArgs &ca = a.ca;
    fft128_inv_scatter(buf, q, tid);
    __syncthreads();
    rows_from_lds(buf, twl, tid);
    __syncthreads();
    rows_stageB(buf, tid);
    __syncthreads();
    const long base = ca.grid_off[ca.fn[fi]] + (long)k * F128 * F128;
    if (REAL) {
        float *out = reinterpret_cast<float *>(a.dst) + base;
        cols_pass_out(buf, twl, tid, [&](int row, int col, float2 v) { out[row * F128 + col] = v.x; });
    } else {
        float2 *out = reinterpret_cast<float2 *>(a.dst) + base;
        cols_pass_out(buf, twl, tid, [&](int row, int col, float2 v) { out[row * F128 + col] = v; });
    }
}

template <bool REAL, bool OPAQUE = false>
__device__ __forceinline__ void so3_fft128_inv2_body(const FftArgs<float> &a, int unit)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2 *buf = reinterpret_cast<float2 *>(smem_raw);
    float2 *twl = buf + F128 * P128;
    const ChunkArgs &ca = a.ca;
    const int nfp = (ca.nfc + 1) / 2, nunit = nfp * F128;
    if (unit < 0 || unit >= nunit) return;
    const int k = unit / nfp, fi = 2 * (unit - k * nfp);
    const bool two = fi + 1 < ca.nfc;
    int tid = threadIdx.x;
    // in a persistent loop over units (OPAQUE) the per-lane index arithmetic stays inside the
    // iteration instead of being hoisted and held in registers across it
    if (OPAQUE) asm volatile("" : "+v"(tid));
    if (tid < F128) twl[tid] = a.tw[tid];
    float4 q0[4][2], q1[4][2];
    const float4 *G4 = reinterpret_cast<const float4 *>(a.src);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const long o = ws_rec(ca, tid + i * 1024, k, fi) * 2, o1 = two ? o + 4 : o;   // select the address, not the value
        q0[i][0] = G4[o];
        q0[i][1] = G4[o + 1];
        q1[i][0] = G4[o1];
        q1[i][1] = G4[o1 + 1];
    }
    if (REAL && two) {
        // two real slices in one complex DFT: the bins of Z = X_a + i X_b, whose transform has
        // the grid of function fi in its real part and that of function fi + 1 in its imaginary part
        float4 z[4][2];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int h = 0; h < 2; h++)
                z[i][h] = make_float4(q0[i][h].x - q1[i][h].y, q0[i][h].y + q1[i][h].x, q0[i][h].z - q1[i][h].w,
                                      q0[i][h].w + q1[i][h].z);
        fft128_inv_scatter(buf, z, tid);
        __syncthreads();
        rows_from_lds(buf, twl, tid);
        __syncthreads();
        rows_stageB(buf, tid);
        __syncthreads();
        float *outa = reinterpret_cast<float *>(a.dst) + ca.grid_off[ca.fn[fi]] + (long)k * F128 * F128;
        float *outb = reinterpret_cast<float *>(a.dst) + ca.grid_off[ca.fn[fi + 1]] + (long)k * F128 * F128;
        cols_pass_out(buf, twl, tid, [&](int row, int col, float2 v) {
            outa[row * F128 + col] = v.x;
            outb[row * F128 + col] = v.y;
        });
        return;
    }
    fft128_inv_slice<REAL>(a, buf, twl, q0, k, fi, tid);
    if (!two) return;
    __syncthreads();  // the first slice's LDS reads before the second slice's scatter
    fft128_inv_slice<REAL>(a, buf, twl, q1, k, fi + 1, tid);
}

// units of so3_fft128_inv2: (beta index, pair of functions)
__host__ __device__ inline int so3_fft128_inv2_units(int nfc) { return (nfc + 1) / 2 * F128; }

template <bool REAL>
__global__ void __launch_bounds__(1024) so3_fft128_inv2(FftArgs<float> a)
{
    so3_fft128_inv2_body<REAL>(a, xcd_unit(so3_fft128_inv2_units(a.ca.nfc)));
}
