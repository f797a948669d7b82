// fft_lines.hpp -- 2-D DFT of SO(3) beta slices too large for LDS (FP64 N = 128, FP32 N = 512),
// in two passes through a scratch grid, with the spectrum workspace in half records.
//
// Pass A transforms along alpha (j1): a workgroup takes CB gamma columns (j2) of one slice, all N
// alpha rows, and writes rows u of the result to the scratch grid [fi][k][u][j2].  Pass B
// transforms along gamma: a workgroup takes the row pair {r, N - r} of one beta index for a block
// of 8 functions.  The four pairs of a group (m0, n0) come from rows -m0, m0 (pairs 0, 1) and
// n0, -n0 (pairs 2, 3) of the spectrum X(-m mod N, -n mod N), so the row pair {r, N - r} holds
// pairs 0 and 1 of the groups (r, n0) and pairs 2 and 3 of the groups (m0, +-r): the workspace
// row of (group, beta index) is split into two halves, [half][function][2 pairs], and the
// workgroup writes whole half rows (8 functions x 2 pairs: 128 bytes in FP32, 256 in FP64) of
// every group it serves.  (One slice per workgroup, as in the single-pass kernels, would write
// 16- or 32-byte pieces of records that other workgroups complete.)  The inverse runs the same
// passes in the other order: pass B gathers the half rows into the bins of its two rows
// (G(m, n) at (m mod N, n mod N)), pass A writes the grid.
//
// The 1-D DFTs are in-LDS Stockham passes (radix 8, then 4 or 2 for what remains), unnormalised
// forward DFTs X[f] = sum_j x[j] W_N^(f j), as every FFT of this library; element e of a line
// sits at e + e / 8 (one pad per 8 elements keeps the strided accesses of the early passes off
// a few banks), and lines are N + N/8 + 1 elements apart, so that lanes reading the same element
// of consecutive lines (pass A's column staging, pass B's records) fall in different banks.
#pragma once

#include "kernels.hpp"

__host__ __device__ constexpr int pr_pad(int e) { return e + (e >> 3); }
__host__ __device__ constexpr int pr_pitch(int N) { return N + N / 8 + 1; }

// one Stockham pass of radix R over L lines of N points (Ns: the length of the sub-transforms
// already done); PER butterflies per thread at most
template <int R, int PER, typename T2>
__device__ __forceinline__ void pr_pass(T2 *buf, int L, int N, int Ns, const T2 *tw, int tid, int nthr)
{
    const int nb = N / R, tot = L * nb, P = pr_pitch(N);
    const int tstep = N / (Ns * R);   // W_{Ns R}^(r k) = W_N^(r k tstep)
    T2 v[PER][R];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int b = tid + i * nthr;
        if (b < tot) {
            const int line = b / nb, j = b - line * nb, k = j % Ns;
            const T2 *x = buf + line * P;
#pragma unroll
            for (int r = 0; r < R; r++) v[i][r] = x[pr_pad(j + r * nb)];
#pragma unroll
            for (int r = 1; r < R; r++) v[i][r] = cmul(v[i][r], tw[r * k * tstep]);
            if (R == 8) dft8(v[i]);
            else if (R == 4) dft4(v[i]);
            else dft2(v[i]);
        }
    }
    __syncthreads();   // every read of the pass before the first write (in place)
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int b = tid + i * nthr;
        if (b < tot) {
            const int line = b / nb, j = b - line * nb, k = j % Ns;
            T2 *x = buf + line * P;
            const int d = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) x[pr_pad(d + r * Ns)] = v[i][r];
        }
    }
    __syncthreads();
}

// forward DFTs of L lines of N points (N = 8^a 2^b) in LDS, natural order in and out
template <int N, int L, int NTHR, typename T2>
__device__ __forceinline__ void pr_lines_dft(T2 *buf, const T2 *tw, int tid)
{
    constexpr int P8 = (L * (N / 8) + NTHR - 1) / NTHR;
    constexpr int P4 = (L * (N / 4) + NTHR - 1) / NTHR;
    constexpr int P2 = (L * (N / 2) + NTHR - 1) / NTHR;
    int Ns = 1;
#pragma unroll
    for (int rem = N; rem >= 8; rem /= 8) {
        pr_pass<8, P8>(buf, L, N, Ns, tw, tid, NTHR);
        Ns *= 8;
    }
    if (N / Ns == 4) pr_pass<4, P4>(buf, L, N, Ns, tw, tid, NTHR);
    else if (N / Ns == 2) pr_pass<2, P2>(buf, L, N, Ns, tw, tid, NTHR);
}

// half record h (0: pairs 0, 1; 1: pairs 2, 3) of (group g, beta index k, function fi): complex
// index of its first pair (rows [g][k], [half][function][2 pairs])
__device__ __forceinline__ long pr_half(const ChunkArgs &ca, int g, int k, int h, int fi)
{
    return 2 * ((((long)g * ca.N + k) * 2 + h) * ca.C + fi);
}

template <typename T, int N> struct PrCfg {
    static constexpr int F = 8;                         // functions per pass-B workgroup
    static constexpr int LB = 2 * F;                    // pass-B lines
    static constexpr int NTHR = LB * N / 8;             // one radix-8 butterfly per thread
    static constexpr int P = pr_pitch(N);
};
// pass-A columns per workgroup: 128-byte row segments of the grid
template <typename T, int N, bool REAL> struct PrColsCfg {
    static constexpr int CB = 128 / (REAL ? (int)sizeof(T) : 2 * (int)sizeof(T));
    static constexpr int NTHR = PrCfg<T, N>::NTHR;
};

template <typename T, int N>
__host__ __device__ constexpr size_t pr_rows_lds() { return ((size_t)PrCfg<T, N>::LB * pr_pitch(N) + N) * 2 * sizeof(T); }
template <typename T, int N, bool REAL>
__host__ __device__ constexpr size_t pr_cols_lds() { return ((size_t)PrColsCfg<T, N, REAL>::CB * pr_pitch(N) + N) * 2 * sizeof(T); }

// Pass A.  Forward: grid columns [cb CB, cb CB + CB) of slice (k, fi) -> DFT along alpha ->
// scratch rows.  Inverse: scratch rows (row N/2, the Nyquist row, is zero) -> DFT along u -> grid.
// Units (fi, k, cb), column blocks fastest.
template <typename T, int N, bool REAL, bool FWD>
__global__ void __launch_bounds__((PrColsCfg<T, N, REAL>::NTHR)) so3_pr_cols(FftArgs<T> a)
{
    typedef typename cplx_t<T>::type T2;
    constexpr int CB = PrColsCfg<T, N, REAL>::CB, NTHR = PrColsCfg<T, N, REAL>::NTHR, P = pr_pitch(N);
    constexpr int NCB = N / CB;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T2 *buf = reinterpret_cast<T2 *>(smem_raw);   // [CB][P]
    T2 *twl = buf + CB * P;
    const ChunkArgs &ca = a.ca;
    const int nunit = ca.nfc * N * NCB;
    const int unit = xcd_unit(nunit);
    if (unit >= nunit) return;
    const int cb = unit % NCB, k = (unit / NCB) % N, fi = unit / (NCB * N);
    const int tid = threadIdx.x;
    for (int t = tid; t < N; t += NTHR) twl[t] = a.tw[t];
    const long gbase = ca.grid_off[ca.fn[fi]] + (long)k * N * N + cb * CB;   // (j1, j2) at j1 N + j2
    T2 *sc = a.scratch + ((long)fi * N + k) * N * N + cb * CB;                 // (u, j2) at u N + j2
    // element e = (row, c): consecutive threads take consecutive columns of one row
    for (int e = tid; e < N * CB; e += NTHR) {
        const int row = e / CB, c = e - row * CB;
        T2 v;
        if (FWD) {
            if (REAL) v = cmk<T2>(reinterpret_cast<const T *>(a.src)[gbase + (long)row * N + c], (T)0);
            else v = reinterpret_cast<const T2 *>(a.src)[gbase + (long)row * N + c];
        } else {
            v = row == N / 2 ? cmk<T2>((T)0, (T)0) : sc[(long)row * N + c];
        }
        buf[c * P + pr_pad(row)] = v;
    }
    __syncthreads();
    pr_lines_dft<N, CB, NTHR>(buf, twl, tid);
    for (int e = tid; e < N * CB; e += NTHR) {
        const int row = e / CB, c = e - row * CB;
        const T2 v = buf[c * P + pr_pad(row)];
        if (FWD) sc[(long)row * N + c] = v;
        else if (REAL) reinterpret_cast<T *>(a.dst)[gbase + (long)row * N + c] = v.x;
        else reinterpret_cast<T2 *>(a.dst)[gbase + (long)row * N + c] = v;
    }
}

// Items of the row pair {r, N - r} of pass B: half 0 of the groups (r, n0), |n0| <= r (2r + 1),
// then half 1 of the groups (m0, s r), m0 in [r, B), s = +1, -1 (one sign when r = 0).
__device__ __forceinline__ int pr_nitems(int r, int B) { return 2 * r + 1 + (r ? 2 : 1) * (B - r); }
__device__ __forceinline__ void pr_item(int it, int r, int B, int &h, int &m0, int &n0)
{
    if (it < 2 * r + 1) {
        h = 0;
        m0 = r;
        n0 = it - r;
        return;
    }
    it -= 2 * r + 1;
    h = 1;
    const int ns = r ? 2 : 1;
    m0 = r + it / ns;
    n0 = (it % ns) ? -r : r;
}

// Pass B.  Forward: scratch rows {r, N - r} of 8 functions -> DFT along gamma -> half records.
// Inverse: half records -> bins of the two rows -> DFT -> scratch rows.  Units (fb, k, r),
// r in [0, N/2) (row N/2 carries no pair), functions of block fb: [8 fb, 8 fb + 8).
template <typename T, int N, bool FWD>
__global__ void __launch_bounds__((PrCfg<T, N>::NTHR)) so3_pr_rows(FftArgs<T> a)
{
    typedef typename cplx_t<T>::type T2;
    constexpr int F = PrCfg<T, N>::F, LB = PrCfg<T, N>::LB, NTHR = PrCfg<T, N>::NTHR, P = pr_pitch(N), H = N / 2;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T2 *buf = reinterpret_cast<T2 *>(smem_raw);   // line 2 f + s: function f, row r (s = 0) or N - r (s = 1)
    T2 *twl = buf + LB * P;
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, nfb = (ca.nfc + F - 1) / F;
    const int nunit = nfb * N * H;
    const int unit = xcd_unit(nunit);
    if (unit >= nunit) return;
    const int r = unit % H, k = (unit / H) % N, fb = unit / (H * N);
    const int tid = threadIdx.x;
    for (int t = tid; t < N; t += NTHR) twl[t] = a.tw[t];
    const int rows[2] = {r, r ? N - r : 0};
    const int nrow = r ? 2 : 1;
    const int nit = pr_nitems(r, B);
    auto line_of = [&](int u) { return u == r ? 0 : 1; };   // row u of the pair
    if (FWD) {
        for (int e = tid; e < F * nrow * N; e += NTHR) {
            const int j = e % N, ls = e / N, f = ls / nrow, s = ls - f * nrow, fi = F * fb + f;
            T2 v = cmk<T2>((T)0, (T)0);
            if (fi < ca.nfc) v = a.scratch[(((long)fi * N + k) * N + rows[s]) * N + j];
            buf[(2 * f + s) * P + pr_pad(j)] = v;
        }
        __syncthreads();
        pr_lines_dft<N, LB, NTHR>(buf, twl, tid);
        // half record (p_a, p_b) of item it for function f: lanes f of one item write one half row
        T2 *W = reinterpret_cast<T2 *>(a.dst);
        for (int e = tid; e < nit * F; e += NTHR) {
            const int f = e % F, it = e / F, fi = F * fb + f;
            int h, m0, n0;
            pr_item(it, r, B, h, m0, n0);
            const T2 *lf = buf + 2 * f * P;
            T2 pa, pb;
            if (h == 0) {   // p0 = X(-m0, -n0), p1 = X(m0, n0), m0 = r
                pa = lf[line_of(wrapN(-m0, N)) * P + pr_pad(wrapN(-n0, N))];
                pb = lf[line_of(m0) * P + pr_pad(wrapN(n0, N))];
            } else {        // p2 = X(n0, m0), p3 = X(-n0, -m0), n0 = +-r
                pa = lf[line_of(wrapN(n0, N)) * P + pr_pad(m0)];
                pb = lf[line_of(wrapN(-n0, N)) * P + pr_pad(wrapN(-m0, N))];
            }
            if (fi < ca.nfc) {
                const long o = pr_half(ca, m0 * m0 + n0 + m0, k, h, fi);
                W[o] = pa;
                W[o + 1] = pb;
            }
        }
    } else {
        for (int e = tid; e < LB * N; e += NTHR) buf[(e / N) * P + pr_pad(e % N)] = cmk<T2>((T)0, (T)0);
        __syncthreads();
        // G(m, n) of every valid pair into bin (m mod N, n mod N) of its row's line
        const T2 *G = reinterpret_cast<const T2 *>(a.src);
        for (int e = tid; e < nit * F; e += NTHR) {
            const int f = e % F, it = e / F, fi = F * fb + f;
            if (fi >= ca.nfc) continue;
            int h, m0, n0;
            pr_item(it, r, B, h, m0, n0);
            const long o = pr_half(ca, m0 * m0 + n0 + m0, k, h, fi);
            const T2 ga = G[o], gb = G[o + 1];
            T2 *lf = buf + 2 * f * P;
            if (h == 0) {   // p0 = (m0, n0), p1 = (-m0, -n0) (not for (0, 0))
                lf[line_of(m0) * P + pr_pad(wrapN(n0, N))] = ga;
                if (m0 > 0) lf[line_of(wrapN(-m0, N)) * P + pr_pad(wrapN(-n0, N))] = gb;
            } else if (n0 != m0 && n0 != -m0) {   // p2 = (-n0, -m0), p3 = (n0, m0)
                lf[line_of(wrapN(-n0, N)) * P + pr_pad(wrapN(-m0, N))] = ga;
                lf[line_of(wrapN(n0, N)) * P + pr_pad(m0)] = gb;
            }
        }
        __syncthreads();
        pr_lines_dft<N, LB, NTHR>(buf, twl, tid);
        for (int e = tid; e < F * nrow * N; e += NTHR) {
            const int j = e % N, ls = e / N, f = ls / nrow, s = ls - f * nrow, fi = F * fb + f;
            if (fi < ca.nfc) a.scratch[(((long)fi * N + k) * N + rows[s]) * N + j] = buf[(2 * f + s) * P + pr_pad(j)];
        }
    }
}
