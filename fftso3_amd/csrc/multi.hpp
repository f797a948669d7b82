// multi.hpp -- one launch per pipeline stage for every bandlimit class of an SO(3) batch (FP32,
// batch-wide kernels, gfx950).
//
// Each stage is one persistent kernel over a static work queue of (class, unit) items that the
// plan builds once: for every class, largest bandlimit first, the units its own launch would
// have given each XCD (xcd_unit / xcd_deal), appended to that XCD's queue, and the 8 queues
// interleaved so that queue position p belongs to XCD p % 8.  The grid is a multiple of 8 and
// workgroup b takes positions b, b + grid, b + 2 grid, ...: it stays on XCD b % 8 and drains that
// XCD's queue in order, so every XCD works through its even share of every class, long units
// first, and the units that share workspace lines still meet in one L2.  A unit runs the same
// device code as the per-class kernels (the *_body functions); a barrier separates consecutive
// units of a workgroup, whose LDS layouts differ between classes.
#pragma once

#include "fft_r16.hpp"
#include "mfma32.hpp"

// the queue entry at position p, in scalar registers: every lane of the workgroup reads the same
// item, and its class index then selects kernel arguments with scalar loads
__device__ __forceinline__ int2 uniform_item(const int2 *items, int p)
{
    const int2 it = items[p];
    return make_int2(__builtin_amdgcn_readfirstlane(it.x), __builtin_amdgcn_readfirstlane(it.y));
}

#define MULTI_MAX_FFT 8     // classes of one FFT stage launch
#define MULTI_MAX_MX 12     // contraction entries (a class may take two: grouped forward at N = 128)

// FFT item kinds: R of the R16 slices (1, 2, 3, 4, 6) or 8 for the N = 128 paired-slice kernels
struct MultiFft {
    FftArgs<float> c[MULTI_MAX_FFT];
    int kind[MULTI_MAX_FFT];
    const int2 *items;     // (entry, unit); entry < 0: an empty slot of the queue
    int nitems;
};

template <bool FWD, bool REAL>
__global__ void __launch_bounds__(1024) so3_fft_multi(MultiFft m)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    for (int p = blockIdx.x; p < m.nitems; p += gridDim.x) {
        const int2 it = uniform_item(m.items, p);
        __syncthreads();                      // the previous unit's LDS reads
        if (it.x < 0) continue;
        const FftArgs<float> &a = m.c[it.x];
        switch (m.kind[it.x]) {
        case 1: r16_item<1, FWD, REAL>(a, it.y, smem_raw); break;
        case 2: r16_item<2, FWD, REAL>(a, it.y, smem_raw); break;
        case 3: r16_item<3, FWD, REAL>(a, it.y, smem_raw); break;
        case 4: r16_item<4, FWD, REAL>(a, it.y, smem_raw); break;
        case 6: r16_item<6, FWD, REAL>(a, it.y, smem_raw); break;
        default:
            if (FWD) so3_fft128_fwd2_body<REAL, true>(a, it.y);
            else so3_fft128_inv2_body<REAL, true>(a, it.y);
        }
    }
}

// contraction kinds: forward 0 = contract_fwd_mx32<4, false> (N < 128), 2 = contract_fwd_mx16g<4>,
// 3 = contract_fwd_mx16g<2> (N = 128, group interleave 1); inverse 0 = contract_inv_mx32<4>,
// 1 = contract_inv_mx32g<2>
struct MultiMx {
    MxArgs c[MULTI_MAX_MX];
    int kind[MULTI_MAX_MX];
    const int2 *items;
    int nitems;
};

__global__ void __launch_bounds__(256) contract_fwd_multi(MultiMx m)
{
    for (int p = blockIdx.x; p < m.nitems; p += gridDim.x) {
        const int2 it = uniform_item(m.items, p);
        __syncthreads();
        if (it.x < 0) continue;
        const MxArgs &a = m.c[it.x];
        switch (m.kind[it.x]) {
        case 2: contract_fwd_mx16g_body<4>(a, it.y); break;
        case 3: contract_fwd_mx16g_body<2>(a, it.y); break;
        default: contract_fwd_mx32_body<4, false>(a, it.y);
        }
    }
}

__global__ void __launch_bounds__(256) contract_inv_multi(MultiMx m)
{
    for (int p = blockIdx.x; p < m.nitems; p += gridDim.x) {
        const int2 it = uniform_item(m.items, p);
        __syncthreads();
        if (it.x < 0) continue;
        const MxArgs &a = m.c[it.x];
        if (m.kind[it.x] == 1) contract_inv_mx32g_body<2>(a, it.y);
        else contract_inv_mx32_body<4>(a, it.y);
    }
}

// coefficient layout passes (PT <-> packed complex) of every class
struct MultiTr {
    ChunkArgs ca[MULTI_MAX_FFT];
    int ncoef[MULTI_MAX_FFT];
    const void *src[MULTI_MAX_FFT];
    void *dst[MULTI_MAX_FFT];
    const int2 *items;
    int nitems;
};

template <int DIR>
__global__ void __launch_bounds__(256) coef_transpose_multi(MultiTr m)
{
    for (int p = blockIdx.x; p < m.nitems; p += gridDim.x) {
        const int2 it = uniform_item(m.items, p);
        __syncthreads();
        if (it.x < 0) continue;
        const int c = it.x;
        coef_transpose_body<DIR, float2>(m.ca[c], m.ncoef[c], m.src[c], m.dst[c], it.y);
    }
}

// real <-> complex coefficient bases of every class: elements [start[c], start[c+1]) are class c's
struct MultiTmix {
    ChunkArgs ca[MULTI_MAX_FFT];
    const void *src[MULTI_MAX_FFT];
    void *dst[MULTI_MAX_FFT];
    long ws_stride[MULTI_MAX_FFT];
    long start[MULTI_MAX_FFT + 1];
    int ncls;
};

template <int DIR>
__global__ void so3_tmix_multi(MultiTmix m)
{
    const long total = m.start[m.ncls];
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        int c = 0;
        while (c + 1 < m.ncls && e >= m.start[c + 1]) c++;
        so3_tmix_elem<float, DIR>(m.ca[c], m.src[c], m.dst[c], m.ws_stride[c], e - m.start[c]);
    }
}
