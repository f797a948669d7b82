// fft.hpp -- batched length-N DFTs on data resident in LDS (gfx950, one workgroup).
//
// The reference computes every alpha/gamma transform with Eigen::FFT<double>::fwd (kissfft,
// src/fdcl_FFTSO3.cpp:689-698, src/fdcl_FFTS2.cpp:153): the unnormalised forward DFT
//     X[f] = sum_j x[j] exp(-2 pi i f j / N).
// Here a tile of `nb` transforms (element stride es, transform stride bs, in LDS) is
// transformed in place by mixed-radix decimation in frequency (radices 8,4,2,3,5,7).  In place
// DIF needs no barrier between the reads and writes of a stage, so a thread only holds one
// radix-R butterfly at a time.  The output is left in mixed-radix digit-reversed order: the
// frequency f is found at position rev[f] (table built on the host).  Lengths with a prime
// factor above 7 use a direct O(N^2) DFT instead (all outputs of a thread held in registers
// across one barrier), whose output is in natural order.
#pragma once

#include "common.hpp"

#define FFT_MAX_STAGES 16
#define FFT_EPT 16  // max elements per thread in one tile (direct-DFT register bound)

struct FftDesc {
    int N;
    int lgN;                   // log2 N when N is a power of two, else -1
    int nst;                   // number of DIF stages (0 => direct DFT)
    int R[FFT_MAX_STAGES];     // radix of stage t
    int M[FFT_MAX_STAGES];     // M_t = L_t / R_t, L_t = span of stage t (L_0 = N)
    int lgR[FFT_MAX_STAGES];   // log2 R_t (power-of-two N only)
    int lgM[FFT_MAX_STAGES];   // log2 M_t (power-of-two N only)
};

__device__ __forceinline__ int ilog2_or_neg(int x) { return (x > 0 && (x & (x - 1)) == 0) ? __ffs(x) - 1 : -1; }

// small in-register DFTs: v[s] <- sum_r v[r] exp(-2 pi i r s / R)
template <typename T2> __device__ __forceinline__ void dft2(T2 *v)
{
    T2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <typename T2> __device__ __forceinline__ T2 mul_mi(T2 a) { return cmk<T2>(a.y, -a.x); } // * (-i)

template <typename T2> __device__ __forceinline__ void dft4(T2 *v)
{
    T2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    T2 b0 = cadd(v[1], v[3]), b1 = mul_mi(csub(v[1], v[3]));
    v[0] = cadd(a0, b0);
    v[2] = csub(a0, b0);
    v[1] = cadd(a1, b1);
    v[3] = csub(a1, b1);
}

template <typename T2> __device__ __forceinline__ void dft8(T2 *v)
{
    typedef decltype(T2::x) T;
    const T h = (T)0.70710678118654752440084436210485;
    // radix-2 first level: e_r = v[r] + v[r+4], o_r = (v[r] - v[r+4]) * W8^r
    T2 e[4], o[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        e[r] = cadd(v[r], v[r + 4]);
        o[r] = csub(v[r], v[r + 4]);
    }
    // W8^1 = (h, -h), W8^2 = -i, W8^3 = (-h, -h)
    o[1] = cmk<T2>(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    o[2] = mul_mi(o[2]);
    o[3] = cmk<T2>(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
    dft4(e);
    dft4(o);
#pragma unroll
    for (int s = 0; s < 4; s++) {
        v[2 * s] = e[s];
        v[2 * s + 1] = o[s];
    }
}

// generic small odd radix from the twiddle table W_N (R | N)
template <typename T2, int R> __device__ __forceinline__ void dft_gen(T2 *v, const T2 *tw, int N)
{
    T2 out[R];
    const int step = N / R;
#pragma unroll
    for (int s = 0; s < R; s++) {
        T2 acc = v[0];
#pragma unroll
        for (int r = 1; r < R; r++) acc = cadd(acc, cmul(v[r], tw[((r * s) % R) * step]));
        out[s] = acc;
    }
#pragma unroll
    for (int s = 0; s < R; s++) v[s] = out[s];
}

template <typename T2, int R> __device__ __forceinline__ void dft_small(T2 *v, const T2 *tw, int N)
{
    if constexpr (R == 2) dft2(v);
    else if constexpr (R == 4) dft4(v);
    else if constexpr (R == 8) dft8(v);
    else dft_gen<T2, R>(v, tw, N);
}

// one DIF stage of radix R over a tile in LDS
template <typename T2, int R>
__device__ __forceinline__ void dif_stage(T2 *buf, const FftDesc &fd, int t, int es, int bs, int nb, bool jfast,
                                          const T2 *__restrict__ tw, int tid, int nthr)
{
    const int N = fd.N;
    const int M = fd.M[t];
    const int Lt = M * R;
    const int per = N / R;           // butterflies per transform
    const int total = per * nb;
    const int twstep = N / Lt;
    const int lgnb = ilog2_or_neg(nb);
    const bool p2 = fd.lgN >= 0 && lgnb >= 0;
    const int lgper = fd.lgN - fd.lgR[t], lgM = fd.lgM[t];
    for (int beta = tid; beta < total; beta += nthr) {
        int b, w, blk, j;
        if (p2) {
            if (jfast) { b = beta >> lgper; w = beta & (per - 1); }
            else { w = beta >> lgnb; b = beta & (nb - 1); }
            blk = w >> lgM;
            j = w & (M - 1);
        } else {
            if (jfast) { b = beta / per; w = beta - b * per; }
            else { w = beta / nb; b = beta - w * nb; }
            blk = w / M;
            j = w - blk * M;
        }
        int base = b * bs + (blk * Lt + j) * es;
        T2 v[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = buf[base + r * M * es];
        dft_small<T2, R>(v, tw, N);
        // twiddle W_Lt^(s j) = W_N^(s j N/Lt)
#pragma unroll
        for (int s = 1; s < R; s++) {
            int e = s * j * twstep;
            e = p2 ? (e & (N - 1)) : (e % N);
            v[s] = cmul(v[s], tw[e]);
        }
#pragma unroll
        for (int s = 0; s < R; s++) buf[base + s * M * es] = v[s];
    }
}

// Direct DFT of every transform of the tile (natural-order output).
template <typename T2>
__device__ __forceinline__ void direct_dft(T2 *buf, const FftDesc &fd, int es, int bs, int nb, bool jfast,
                                           const T2 *__restrict__ tw, int tid, int nthr)
{
    const int N = fd.N;
    const int total = N * nb;
    T2 out[FFT_EPT];
#pragma unroll
    for (int i = 0; i < FFT_EPT; i++) {
        int e = tid + i * nthr;
        if (e < total) {
            int b, f;
            if (jfast) { b = e / N; f = e - b * N; }
            else { f = e / nb; b = e - f * nb; }
            T2 acc = cmk<T2>(0, 0);
            int idx = 0;
            for (int j = 0; j < N; j++) {
                acc = cadd(acc, cmul(buf[b * bs + j * es], tw[idx]));
                idx += f;
                if (idx >= N) idx -= N;
            }
            out[i] = acc;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < FFT_EPT; i++) {
        int e = tid + i * nthr;
        if (e < total) {
            int b, f;
            if (jfast) { b = e / N; f = e - b * N; }
            else { f = e / nb; b = e - f * nb; }
            buf[b * bs + f * es] = out[i];
        }
    }
}

// Full transform of a tile; ends with a barrier.  jfast: consecutive threads walk the
// element index (use for es == 1) rather than the transform index (use for es > 1).
template <typename T2>
__device__ void fft_tile(T2 *buf, const FftDesc &fd, int es, int bs, int nb, bool jfast, const T2 *__restrict__ tw,
                         int tid, int nthr)
{
    if (fd.nst == 0) {
        direct_dft(buf, fd, es, bs, nb, jfast, tw, tid, nthr);
        __syncthreads();
        return;
    }
    for (int t = 0; t < fd.nst; t++) {
        switch (fd.R[t]) {
        case 8: dif_stage<T2, 8>(buf, fd, t, es, bs, nb, jfast, tw, tid, nthr); break;
        case 4: dif_stage<T2, 4>(buf, fd, t, es, bs, nb, jfast, tw, tid, nthr); break;
        case 2: dif_stage<T2, 2>(buf, fd, t, es, bs, nb, jfast, tw, tid, nthr); break;
        case 3: dif_stage<T2, 3>(buf, fd, t, es, bs, nb, jfast, tw, tid, nthr); break;
        case 5: dif_stage<T2, 5>(buf, fd, t, es, bs, nb, jfast, tw, tid, nthr); break;
        default: dif_stage<T2, 7>(buf, fd, t, es, bs, nb, jfast, tw, tid, nthr); break;
        }
        __syncthreads();
    }
}
