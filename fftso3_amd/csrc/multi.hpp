// multi.hpp -- one launch per pipeline stage for every bandlimit class of an SO(3) batch (FP32,
// batch-wide kernels, gfx950).
//
// Each stage is one persistent kernel over a work queue of (class, unit) items that the plan
// builds once: for every class, largest bandlimit first, the units its own launch would have given
// each XCD (xcd_unit / xcd_deal) are appended to that XCD's queue; queue x holds its q-th item at
// position q * 8 + x.  Workgroup b belongs to XCD b % 8 (the hardware deals workgroups round-robin
// over the XCDs) and claims the next item of that XCD's queue with an atomic counter until the
// queue is empty, so every XCD works through its share of every class, long units first, the
// units that share workspace lines still meet in one L2, and a slow workgroup takes fewer items.
// The last workgroup to finish resets the counters for the next launch (stream order makes the
// reset visible to it).  A unit runs the same device code as the per-class kernels (the *_body
// functions); the claim's barrier also separates consecutive units, whose LDS layouts differ.
#pragma once

#include "fft_r16.hpp"
#include "mfma32.hpp"
#include "mfma_split.hpp"

// the queue entry at position p, in scalar registers: every lane of the workgroup reads the same
// item, and its class index then selects kernel arguments with scalar loads
__device__ __forceinline__ int2 uniform_item(const int2 *items, int p)
{
    const int2 it = items[p];
    return make_int2(__builtin_amdgcn_readfirstlane(it.x), __builtin_amdgcn_readfirstlane(it.y));
}

// counters of one stage's queues: [XCD_COUNT] next position per XCD, then the finished workgroups
#define MULTI_CTRS (XCD_COUNT + 1)

// the next position of XCD x's queue, the same for the whole workgroup
__device__ __forceinline__ int multi_claim(unsigned *ctr, int x)
{
    __shared__ int s_q;
    __syncthreads();                                  // the previous claim read, the previous unit done
    if (threadIdx.x == 0) s_q = (int)atomicAdd(ctr + x, 1u);
    __syncthreads();
    return s_q;
}

// every workgroup calls this once, after its last claim; the last one resets the counters
__device__ __forceinline__ void multi_finish(unsigned *ctr)
{
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(ctr + XCD_COUNT, 1u) == gridDim.x - 1) {
            for (int x = 0; x <= XCD_COUNT; x++) atomicExch(ctr + x, 0u);
        }
    }
}

struct MultiQueue {
    const int2 *items;     // (entry, unit) at position q * XCD_COUNT + x
    unsigned *ctr;         // MULTI_CTRS counters, zero before the launch
    int len[XCD_COUNT];    // items of each XCD's queue
};

// drain XCD (blockIdx.x % 8)'s queue, calling f(entry, unit) for every item
template <typename F>
__device__ __forceinline__ void multi_drain(const MultiQueue &q, F f)
{
    const int x = blockIdx.x % XCD_COUNT;
    for (;;) {
        const int p = multi_claim(q.ctr, x);
        if (p >= q.len[x]) break;
        const int2 it = uniform_item(q.items, p * XCD_COUNT + x);
        if (it.x >= 0) f(it.x, it.y);
    }
    multi_finish(q.ctr);
}

#define MULTI_MAX_FFT 8     // classes of one FFT stage launch
#define MULTI_MAX_MX 12     // contraction entries (a class may take two: grouped forward at N = 128)

// FFT item kinds: R of the R16 slices (1, 2, 3, 4, 6) or 8 for the N = 128 paired-slice kernels
struct MultiFft {
    FftArgs<float> c[MULTI_MAX_FFT];
    int kind[MULTI_MAX_FFT];
    MultiQueue q;
};

template <bool FWD, bool REAL>
__global__ void __launch_bounds__(1024) so3_fft_multi(MultiFft m)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    multi_drain(m.q, [&](int e, int unit) {
        const FftArgs<float> &a = m.c[e];
        switch (m.kind[e]) {
        case 1: r16_item<1, FWD, REAL>(a, unit, smem_raw); break;
        case 2: r16_item<2, FWD, REAL>(a, unit, smem_raw); break;
        case 3: r16_item<3, FWD, REAL>(a, unit, smem_raw); break;
        case 4: r16_item<4, FWD, REAL>(a, unit, smem_raw); break;
        case 6: r16_item<6, FWD, REAL>(a, unit, smem_raw); break;
        default:
            if (FWD) so3_fft128_fwd2_body<REAL, true>(a, unit);
            else so3_fft128_inv2_body<REAL, true>(a, unit);
        }
    });
}

// contraction kinds: forward 0 = contract_fwd_mx32<4, false> (N < 128), 2 = the contract_fwd_mx16g
// body with four row tiles per wave, 3 = with two (N = 128, group interleave 1); inverse 0 =
// contract_inv_mx32<4>, 1 = contract_inv_mx32g<2>; split plans 4 = contract_fwd_mxks /
// contract_inv_mxks
struct MultiMx {
    MxArgs c[MULTI_MAX_MX];
    int kind[MULTI_MAX_MX];
    MultiQueue q;
};

__global__ void __launch_bounds__(256) contract_fwd_multi(MultiMx m)
{
    multi_drain(m.q, [&](int e, int unit) {
        const MxArgs &a = m.c[e];
        switch (m.kind[e]) {
        // without the paired-block stores, whose extra registers would make this combined kernel spill
        case 2: contract_fwd_mx16g_body<4, false>(a, unit); break;
        case 3: contract_fwd_mx16g_body<2, false>(a, unit); break;
        default: contract_fwd_mx32_body<4, false>(a, unit);
        }
    });
}

__global__ void __launch_bounds__(256) contract_inv_multi(MultiMx m)
{
    multi_drain(m.q, [&](int e, int unit) {
        const MxArgs &a = m.c[e];
        if (m.kind[e] == 1) contract_inv_mx32g_body<2>(a, unit);
        else contract_inv_mx32_body<4>(a, unit);
    });
}

// split-BF16 plans (NPC pieces per operand): every class's contraction is the register-operand
// split kernel of its own launches (contract_fwd_mxks / contract_inv_mxks, 8 waves)
template <int NPC>
__global__ void __launch_bounds__(512) contract_fwd_multi_s(MultiMx m)
{
    multi_drain(m.q, [&](int e, int unit) { contract_fwd_mxks_body<4, NPC, 8>(m.c[e], unit); });
}

template <int NPC>
__global__ void __launch_bounds__(512) contract_inv_multi_s(MultiMx m)
{
    multi_drain(m.q, [&](int e, int unit) { contract_inv_mxks_body<4, NPC, 8>(m.c[e], unit); });
}

// coefficient layout passes (PT <-> packed complex) of every class
struct MultiTr {
    ChunkArgs ca[MULTI_MAX_FFT];
    int ncoef[MULTI_MAX_FFT];
    const void *src[MULTI_MAX_FFT];
    void *dst[MULTI_MAX_FFT];
    MultiQueue q;
};

// The layout tiles are small and of equal cost: workgroups stride through their XCD's queue
// statically instead of claiming tiles one at a time (the counters are not used).
template <int DIR>
__global__ void __launch_bounds__(256) coef_transpose_multi(MultiTr m)
{
    const int x = blockIdx.x % XCD_COUNT, per = gridDim.x / XCD_COUNT;
    for (int q = blockIdx.x / XCD_COUNT; q < m.q.len[x]; q += per) {
        const int2 it = uniform_item(m.q.items, q * XCD_COUNT + x);
        __syncthreads();                      // the previous tile's LDS reads
        if (it.x >= 0) coef_transpose_body<DIR, float2>(m.ca[it.x], m.ncoef[it.x], m.src[it.x], m.dst[it.x], it.y);
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
