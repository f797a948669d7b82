// kernels.hpp -- HIP kernels (gfx950) for the SO(3) and S^2 harmonic transforms.
//
// Pipeline per chunk of C functions of one bandlimit B (N = 2B):
//   forward:  grid --fft kernels--> spectrum workspace S --contraction--> coefficients
//   inverse:  coefficients --contraction^T--> workspace G --fft kernels--> grid
// The workspace is group-major: S[g][k][fi][p] (complex), g the canonical (m0,n0) group,
// k the beta index, fi the function within the chunk, p the symmetry pair (4 for SO(3),
// 2 for S^2).  The contraction of one group is a small real-times-complex GEMM whose
// Wigner-d (or Legendre) matrix is regenerated on the fly in FP64 by the three-term recursion
// in l and kept in LDS; nothing of the d-table ever goes to HBM.
#pragma once

#include "common.hpp"
#include "fft.hpp"

#define XCD_COUNT 8

// Map blockIdx.x so that consecutive work units land on the same XCD (each XCD has its own
// L2; the hardware deals workgroups round-robin over the 8 XCDs).
__device__ __forceinline__ int xcd_unit(int nunits)
{
    int q = (nunits + XCD_COUNT - 1) / XCD_COUNT;
    int b = blockIdx.x;
    return (b % XCD_COUNT) * q + b / XCD_COUNT;
}
static inline int xcd_grid(int nunits) { return ((nunits + XCD_COUNT - 1) / XCD_COUNT) * XCD_COUNT; }

// Map blockIdx.x for units of unequal cost (groups come largest first): runs of `run`
// consecutive units are dealt round-robin over the XCDs, so every XCD gets an even share of the
// long units while each run (units that share cache lines) stays under one L2.  The grid is
// deal_grid(n, run); units >= n exit.
__device__ __forceinline__ int xcd_deal(int run)
{
    const int x = blockIdx.x % XCD_COUNT, i = blockIdx.x / XCD_COUNT;
    return ((i / run) * XCD_COUNT + x) * run + (i % run);
}
static inline int deal_grid(int nunits, int run)
{
    const int q = XCD_COUNT * run;
    return ((nunits + q - 1) / q) * q;
}

// ---------------------------------------------------------------------------------------------
// Per-chunk addressing shared by all kernels
// ---------------------------------------------------------------------------------------------
struct ChunkArgs {
    int B, N;
    int C;        // workspace stride in functions (chunk capacity)
    int nfc;      // functions in this chunk (<= C)
    const int *fn;          // chunk-local index -> global function index (device)
    const long *grid_off;   // global function -> grid offset (elements)
    const long *coef_off;   // global function -> coefficient offset (elements)
    int kmajor;             // workspace rows: 0 = [g][k] (chunked kernels), 1 = blocked (batch-wide)
    int klo;                // blocked rows: [k >> klo][g][k & (2^klo - 1)]
    int ngrp;               // (m,n) groups of the bandlimit: B^2 (SO(3)) or B (S^2)
    int gil;                // log2 of the groups whose records interleave per function (ws_rec)
    int hsplit;             // SO(3) rows [half][function][2 pairs] (fft_lines.hpp) instead of [function][4 pairs]
};

// column c of a chunk's workspace row (real values) -> function, pair, re/im
__device__ __forceinline__ void ws_col(const ChunkArgs &ca, int NP, int c, int &fi, int &p, int &ri)
{
    ri = c & 1;
    if (ca.hsplit) {
        const int h = c / (4 * ca.C), q = c - h * 4 * ca.C;
        fi = q >> 2;
        p = 2 * h + ((q >> 1) & 1);
    } else {
        p = (c >> 1) % NP;
        fi = c / (2 * NP);
    }
}

// Workspace row of (group g, beta index k); a row holds C functions x NP pairs of complex values.
// The blocked order keeps the rows one FFT slice (one k) touches within 2^klo rows of stride of
// each other, and a group's rows in runs of 2^klo contiguous rows, so neither the FFT nor the
// contraction spreads its accesses over many pages.
__device__ __forceinline__ long ws_row(const ChunkArgs &ca, int g, int k)
{
    if (!ca.kmajor) return (long)g * ca.N + k;
    return ((((long)(k >> ca.klo) * ca.ngrp + g)) << ca.klo) + (k & ((1 << ca.klo) - 1));
}

// workspace rows of a bandlimit class (the blocked order pads N up to a multiple of 2^klo)
__host__ __device__ inline long ws_rows(int kmajor, int klo, int ngrp, int N)
{
    return kmajor ? (long)(((N + (1 << klo) - 1) >> klo) * ngrp) << klo : (long)ngrp * N;
}

// Record (NP pairs) of (group g, beta index k, function fi).  With gil > 0 (batch-wide rows,
// klo = 0) the records of 2^gil consecutive groups sit side by side for each function, so one
// FFT slice writes (reads) whole 128-byte lines of its own instead of 32-byte pieces that only
// neighbouring slices complete; the contraction's row segments then stride by 2^gil records.
__device__ __forceinline__ long ws_rec(const ChunkArgs &ca, int g, int k, int fi)
{
    const int gm = (1 << ca.gil) - 1;
    return ws_row(ca, g & ~gm, k) * ca.C + ((long)fi << ca.gil) + (g & gm);
}

// workspace element of (group g, beta index k, pair p, function fi)
__device__ __forceinline__ long ws_at(const ChunkArgs &ca, int g, int k, int p, int fi, int NP)
{
    return (ws_row(ca, g, k) * ca.C + fi) * NP + p;
}

// ---------------------------------------------------------------------------------------------
// SO(3) FFT kernels
// ---------------------------------------------------------------------------------------------
template <typename T> struct FftArgs {
    ChunkArgs ca;
    FftDesc fd;
    const typename cplx_t<T>::type *tw;  // W_N^t, t < N
    const int *rev;                      // frequency -> position in the DIF output
    const void *src;                     // grid (forward) or workspace (inverse)
    void *dst;                           // workspace (forward) or grid (inverse)
    typename cplx_t<T>::type *scratch;   // two-pass fallback: [fi][k][j1][v] complex
    int tile;                            // two-pass fallback: rows (or columns) per workgroup
};

// Copy n elements into LDS with U loads in flight per thread before the first store.
template <int U, typename T2, typename F>
__device__ __forceinline__ void fill_lds(T2 *dst, int n, int tid, int nthr, F get)
{
    for (int base = 0; base < n; base += U * nthr) {
        T2 r[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            int e = base + tid + u * nthr;
            if (e < n) r[u] = get(e);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            int e = base + tid + u * nthr;
            if (e < n) dst[e] = r[u];
        }
    }
}

// twiddle and digit-reversal tables follow the data in LDS
template <typename T2>
__device__ __forceinline__ void load_fft_tables(T2 *twl, int *revl, const T2 *tw, const int *rev, int N, int tid, int nthr)
{
    for (int i = tid; i < N; i += nthr) {
        twl[i] = tw[i];
        revl[i] = rev[i];
    }
}

// Fused forward: one (fi, k) slice per workgroup, whole N x N slice in LDS.
template <typename T, bool REAL>
__global__ void __launch_bounds__(1024) so3_fft_fwd_slice(FftArgs<T> a)
{
    typedef typename cplx_t<T>::type T2;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const ChunkArgs &ca = a.ca;
    const int N = ca.N, B = ca.B, NN = N * N, nslice = ca.nfc * N;
    T2 *buf = reinterpret_cast<T2 *>(smem_raw);
    T2 *twl = buf + NN;
    int *revl = reinterpret_cast<int *>(twl + N);
    const int unit = xcd_unit(nslice);
    if (unit >= nslice) return;
    const int k = unit / ca.nfc, fi = unit - k * ca.nfc;
    const int tid = threadIdx.x, nthr = blockDim.x;
    load_fft_tables(twl, revl, a.tw, a.rev, N, tid, nthr);
    const long base = ca.grid_off[ca.fn[fi]] + (long)k * NN;
    if (REAL) {
        const T *g = reinterpret_cast<const T *>(a.src) + base;
        fill_lds<8>(buf, NN, tid, nthr, [&](int e) { return cmk<T2>(g[e], (T)0); });
    } else {
        const T2 *g = reinterpret_cast<const T2 *>(a.src) + base;
        fill_lds<8>(buf, NN, tid, nthr, [&](int e) { return g[e]; });
    }
    __syncthreads();
    fft_tile(buf, a.fd, 1, N, N, true, twl, tid, nthr);   // over j2 (gamma)
    fft_tile(buf, a.fd, N, 1, N, false, twl, tid, nthr);  // over j1 (alpha)
    // emit S_k(m,n) = X(-m mod N, -n mod N) for the 4 pairs of every group
    T2 *S = reinterpret_cast<T2 *>(a.dst);
    for (int g = tid; g < B * B; g += nthr) {
        int m0, n0;
        so3_group(g, m0, n0);
        const int um = revl[wrapN(-m0, N)], un = revl[wrapN(-n0, N)], vm = revl[wrapN(m0, N)], vn = revl[wrapN(n0, N)];
        S[ws_at(ca, g, k, 0, fi, 4)] = buf[um * N + un];
        S[ws_at(ca, g, k, 1, fi, 4)] = buf[vm * N + vn];
        S[ws_at(ca, g, k, 2, fi, 4)] = buf[vn * N + vm];
        S[ws_at(ca, g, k, 3, fi, 4)] = buf[un * N + um];
    }
}

// Fused inverse: gather G_k(m,n) into bins (m mod N, n mod N), 2-D DFT, store the slice.
template <typename T, bool REAL>
__global__ void __launch_bounds__(1024) so3_fft_inv_slice(FftArgs<T> a)
{
    typedef typename cplx_t<T>::type T2;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const ChunkArgs &ca = a.ca;
    const int N = ca.N, B = ca.B, NN = N * N, nslice = ca.nfc * N;
    T2 *buf = reinterpret_cast<T2 *>(smem_raw);
    T2 *twl = buf + NN;
    int *revl = reinterpret_cast<int *>(twl + N);
    const int unit = xcd_unit(nslice);
    if (unit >= nslice) return;
    const int k = unit / ca.nfc, fi = unit - k * ca.nfc;
    const int tid = threadIdx.x, nthr = blockDim.x;
    load_fft_tables(twl, revl, a.tw, a.rev, N, tid, nthr);
    const T2 *G = reinterpret_cast<const T2 *>(a.src);
    // Nyquist row and column are zero; every other bin holds exactly one pair of one group
    for (int i = tid; i < N; i += nthr) {
        buf[B * N + i] = cmk<T2>((T)0, (T)0);
        buf[i * N + B] = cmk<T2>((T)0, (T)0);
    }
    for (int g = tid; g < B * B; g += nthr) {
        int m0, n0;
        so3_group(g, m0, n0);
        T2 q0 = G[ws_at(ca, g, k, 0, fi, 4)], q1 = G[ws_at(ca, g, k, 1, fi, 4)];
        T2 q2 = G[ws_at(ca, g, k, 2, fi, 4)], q3 = G[ws_at(ca, g, k, 3, fi, 4)];
        const int vm = wrapN(m0, N), vn = wrapN(n0, N), um = wrapN(-m0, N), un = wrapN(-n0, N);
        buf[vm * N + vn] = q0;
        if (m0 > 0) {
            buf[um * N + un] = q1;
            if (n0 != m0 && n0 != -m0) {
                buf[un * N + um] = q2;
                buf[vn * N + vm] = q3;
            }
        }
    }
    __syncthreads();
    fft_tile(buf, a.fd, 1, N, N, true, twl, tid, nthr);
    fft_tile(buf, a.fd, N, 1, N, false, twl, tid, nthr);
    const long base = ca.grid_off[ca.fn[fi]] + (long)k * NN;
    for (int e = tid; e < NN; e += nthr) {
        int j1 = e / N, j2 = e - j1 * N;
        T2 val = buf[revl[j1] * N + revl[j2]];
        if (REAL) reinterpret_cast<T *>(a.dst)[base + e] = val.x;
        else reinterpret_cast<T2 *>(a.dst)[base + e] = val;
    }
}

// Two-pass fallback for slices larger than LDS.
// Rows pass: `tile` rows of one (fi,k) slice; DFT over the contiguous axis; natural order out.
// MODE 0: grid -> scratch (forward first pass), MODE 1: scratch -> grid (inverse second pass).
template <typename T, bool REAL, int MODE>
__global__ void __launch_bounds__(1024) so3_fft_rows(FftArgs<T> a)
{
    typedef typename cplx_t<T>::type T2;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const ChunkArgs &ca = a.ca;
    const int N = ca.N, R = a.tile, nblk = (N + R - 1) / R, nunit = ca.nfc * N * nblk;
    T2 *buf = reinterpret_cast<T2 *>(smem_raw);
    T2 *twl = buf + R * N;
    int *revl = reinterpret_cast<int *>(twl + N);
    const int unit = xcd_unit(nunit);
    if (unit >= nunit) return;
    const int rb = unit % nblk, k = (unit / nblk) % N, fi = unit / (nblk * N);
    // rows of this block: the last block may be partial; the rows past it are transformed along
    // with the rest (each row's DFT is independent) and never stored
    const int nrow = min(R, N - rb * R);
    const int fn = ca.fn[fi];
    const int tid = threadIdx.x, nthr = blockDim.x;
    load_fft_tables(twl, revl, a.tw, a.rev, N, tid, nthr);
    const long gbase = ca.grid_off[fn] + ((long)k * N + (long)rb * R) * N;
    const long sbase = (((long)fi * N + k) * N + (long)rb * R) * N;
    if (MODE == 0) {
        if (REAL) {
            const T *g = reinterpret_cast<const T *>(a.src) + gbase;
            fill_lds<8>(buf, nrow * N, tid, nthr, [&](int e) { return cmk<T2>(g[e], (T)0); });
        } else {
            const T2 *g = reinterpret_cast<const T2 *>(a.src) + gbase;
            fill_lds<8>(buf, nrow * N, tid, nthr, [&](int e) { return g[e]; });
        }
    } else {
        const T2 *sc = a.scratch + sbase;
        fill_lds<8>(buf, nrow * N, tid, nthr, [&](int e) { return sc[e]; });
    }
    __syncthreads();
    fft_tile(buf, a.fd, 1, N, R, true, twl, tid, nthr);
    for (int e = tid; e < nrow * N; e += nthr) {
        int r = e / N, v = e - r * N;
        T2 val = buf[r * N + revl[v]];
        if (MODE == 0) a.scratch[sbase + e] = val;
        else if (REAL) reinterpret_cast<T *>(a.dst)[gbase + e] = val.x;
        else reinterpret_cast<T2 *>(a.dst)[gbase + e] = val;
    }
}

// Columns pass over `tile` columns of one (fi,k) slice.
// MODE 0 (forward): scratch -> DFT over j1 -> emit to the group-major workspace.
// MODE 1 (inverse): gather from the workspace -> DFT over the m bins -> scratch.
template <typename T, int MODE>
__global__ void __launch_bounds__(1024) so3_fft_cols(FftArgs<T> a)
{
    typedef typename cplx_t<T>::type T2;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const ChunkArgs &ca = a.ca;
    const int N = ca.N, B = ca.B, TC = a.tile, nblk = (N + TC - 1) / TC, nunit = ca.nfc * N * nblk;
    T2 *buf = reinterpret_cast<T2 *>(smem_raw);  // [row][tc]; columns past N (last block) stay zero
    T2 *twl = buf + TC * N;
    int *revl = reinterpret_cast<int *>(twl + N);
    const int unit = xcd_unit(nunit);
    if (unit >= nunit) return;
    // Unit order.  Functions fastest, then column blocks: the workgroups that read (inverse) or
    // write (forward, FP32) the same workspace lines -- the functions and pairs of one beta row's
    // groups -- run together under one L2.  An FP64 function's 4 pairs fill a 64-byte segment on
    // their own, which the column blocks of one slice complete together, so the FP64 forward
    // keeps slices whole (measured: the function-fastest order made it 7% slower at l_max = 63).
    constexpr bool fn_fast = MODE == 1 || sizeof(T) == 4;
    const int fi = fn_fast ? unit % ca.nfc : unit / (nblk * N);
    const int cb = fn_fast ? (unit / ca.nfc) % nblk : unit % nblk;
    const int k = fn_fast ? unit / (ca.nfc * nblk) : (unit / nblk) % N;
    const int tid = threadIdx.x, nthr = blockDim.x;
    load_fft_tables(twl, revl, a.tw, a.rev, N, tid, nthr);
    const long sbase = ((long)fi * N + k) * N * N;
    T2 *W = reinterpret_cast<T2 *>(MODE == 0 ? a.dst : const_cast<void *>(a.src));
    if (MODE == 0) {
        const T2 *sc = a.scratch + sbase;
        fill_lds<8>(buf, N * TC, tid, nthr, [&](int e) {
            int r = e / TC, c = e - r * TC, v = cb * TC + c;
            return v < N ? sc[(long)r * N + v] : cmk<T2>((T)0, (T)0);
        });
    } else {
        fill_lds<8>(buf, N * TC, tid, nthr, [&](int e) {
            int u = e / TC, c = e - u * TC, v = cb * TC + c;
            T2 val = cmk<T2>((T)0, (T)0);
            if (u != B && v != B && v < N) {
                int m = u < B ? u : u - N, n = v < B ? v : v - N, g, p;
                so3_slot(m, n, g, p);
                val = W[ws_at(ca, g, k, p, fi, 4)];
            }
            return val;
        });
    }
    __syncthreads();
    fft_tile(buf, a.fd, TC, 1, TC, false, twl, tid, nthr);
    if (MODE == 0) {
        for (int e = tid; e < N * TC; e += nthr) {
            int u = e / TC, c = e - u * TC, v = cb * TC + c;
            if (u == B || v == B || v >= N) continue;
            int m = wrapN(-u, N), n = wrapN(-v, N);
            m = m < B ? m : m - N;
            n = n < B ? n : n - N;
            int g, p;
            so3_slot(m, n, g, p);
            // duplicated slots of orbits of size 1 or 2 are left unwritten: the contraction's
            // columns are independent and it never stores results for invalid pairs
            W[ws_at(ca, g, k, p, fi, 4)] = buf[revl[u] * TC + c];
        }
    } else {
        for (int e = tid; e < N * TC; e += nthr) {
            int j1 = e / TC, c = e - j1 * TC;
            if (cb * TC + c < N) a.scratch[sbase + (long)j1 * N + cb * TC + c] = buf[revl[j1] * TC + c];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// S^2 FFT kernels: `tile` rings of one function per workgroup
// ---------------------------------------------------------------------------------------------
template <typename T, bool REAL>
__global__ void __launch_bounds__(1024) s2_fft_fwd(FftArgs<T> a)
{
    typedef typename cplx_t<T>::type T2;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const ChunkArgs &ca = a.ca;
    const int N = ca.N, B = ca.B, R = a.tile, nblk = N / R, nunit = ca.nfc * nblk;
    T2 *buf = reinterpret_cast<T2 *>(smem_raw);
    T2 *twl = buf + R * N;
    int *revl = reinterpret_cast<int *>(twl + N);
    const int unit = xcd_unit(nunit);
    if (unit >= nunit) return;
    // consecutive units: the same rows of consecutive functions, whose workspace records share lines
    const int fi = unit % ca.nfc, rb = unit / ca.nfc;
    const int fn = ca.fn[fi];
    const int tid = threadIdx.x, nthr = blockDim.x;
    load_fft_tables(twl, revl, a.tw, a.rev, N, tid, nthr);
    const long gbase = ca.grid_off[fn] + (long)rb * R * N;
    if (REAL) {
        const T *g = reinterpret_cast<const T *>(a.src) + gbase;
        fill_lds<8>(buf, R * N, tid, nthr, [&](int e) { return cmk<T2>(g[e], (T)0); });
    } else {
        const T2 *g = reinterpret_cast<const T2 *>(a.src) + gbase;
        fill_lds<8>(buf, R * N, tid, nthr, [&](int e) { return g[e]; });
    }
    __syncthreads();
    fft_tile(buf, a.fd, 1, N, R, true, twl, tid, nthr);
    // S_k(m) = X(m mod N) (src/fdcl_FFTS2.cpp:159-166); slot (|m|, m<0)
    T2 *S = reinterpret_cast<T2 *>(a.dst);
    const int nq = 2 * B;
    for (int e = tid; e < R * nq; e += nthr) {
        int r = e / nq, q = e - r * nq, g = q >> 1, p = q & 1;
        int m = p ? -g : g;
        int k = rb * R + r;
        S[ws_at(ca, g, k, p, fi, 2)] = buf[r * N + revl[wrapN(m, N)]];
    }
}

template <typename T, bool REAL>
__global__ void __launch_bounds__(1024) s2_fft_inv(FftArgs<T> a)
{
    typedef typename cplx_t<T>::type T2;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const ChunkArgs &ca = a.ca;
    const int N = ca.N, B = ca.B, R = a.tile, nblk = N / R, nunit = ca.nfc * nblk;
    T2 *buf = reinterpret_cast<T2 *>(smem_raw);
    T2 *twl = buf + R * N;
    int *revl = reinterpret_cast<int *>(twl + N);
    const int unit = xcd_unit(nunit);
    if (unit >= nunit) return;
    // consecutive units: the same rows of consecutive functions, whose workspace records share lines
    const int fi = unit % ca.nfc, rb = unit / ca.nfc;
    const int fn = ca.fn[fi];
    const int tid = threadIdx.x, nthr = blockDim.x;
    load_fft_tables(twl, revl, a.tw, a.rev, N, tid, nthr);
    const T2 *G = reinterpret_cast<const T2 *>(a.src);
    // G_k(m) at bin (-m mod N)
    fill_lds<4>(buf, R * N, tid, nthr, [&](int e) {
        int r = e / N, u = e - r * N, k = rb * R + r;
        T2 val = cmk<T2>((T)0, (T)0);
        if (u != B) {
            int m = -(u < B ? u : u - N);   // bin u holds m with -m = u (mod N)
            int g = iabs(m), p = m < 0;
            val = G[ws_at(ca, g, k, p, fi, 2)];
        }
        return val;
    });
    __syncthreads();
    fft_tile(buf, a.fd, 1, N, R, true, twl, tid, nthr);
    const long gbase = ca.grid_off[fn] + (long)rb * R * N;
    for (int e = tid; e < R * N; e += nthr) {
        int r = e / N, j = e - r * N;
        T2 val = buf[r * N + revl[j]];
        if (REAL) reinterpret_cast<T *>(a.dst)[gbase + e] = val.x;
        else reinterpret_cast<T2 *>(a.dst)[gbase + e] = val;
    }
}

// ---------------------------------------------------------------------------------------------
// FP64 recursions (one lane per beta_k), shared by the contraction kernels
// ---------------------------------------------------------------------------------------------
struct RecTables {
    const double *cb;      // cos(beta_k)                        [N]
    const double *lnc;     // ln cos(beta_k/2)                   [N]
    const double *lns;     // ln sin(beta_k/2)                   [N]
    const double *w;       // quadrature weight w_k              [N]
    const double *lnbin;   // SO(3): 0.5 ln C(2m0, m0+n0) per g  [B^2]; S^2: seed log per m0 [B]
    const double *sq;      // SO(3): sqrt(l^2 - m^2)             [B*B] (l-major); S^2: fact table
    const double *isq;     // SO(3): 1/sqrt(l^2 - m^2);          S^2: 1/fact
    const double *lcoef;   // SO(3): per l {l(2l-1), (2l-1)/(l-1), l/(l-1)} [3B]
};

// Wigner d^l_{m0 n0}(beta_k) for l = m0, m0+1, ...: seed d^{m0}_{m0 n0} =
// sqrt(C(2m0, m0+n0)) cos(b/2)^(m0+n0) (-sin(b/2))^(m0-n0), then the three-term recursion (64)
// of src/fdcl_FFTSO3.cpp:154-161:
//   d^l = [ (l(2l-1) cb - (2l-1) m n/(l-1)) d^{l-1} - l/(l-1) sqrt(((l-1)^2-m^2)((l-1)^2-n^2)) d^{l-2} ]
//         / sqrt((l^2-m^2)(l^2-n^2)).
struct WignerRec {
    int B, m0, n0, an0;
    double cb, d1, d2;  // d1 = d^{l-1}, d2 = d^{l-2}
    const double *sq, *isq, *lc;
    __device__ __forceinline__ void init(const RecTables &t, int B_, int g, int m0_, int n0_, int k)
    {
        B = B_; m0 = m0_; n0 = n0_; an0 = iabs(n0_);
        cb = t.cb[k];
        sq = t.sq; isq = t.isq; lc = t.lcoef;
        int a = m0 + n0, b = m0 - n0;
        double x = t.lnbin[g] + (a ? a * t.lnc[k] : 0.0) + (b ? b * t.lns[k] : 0.0);
        double s = exp(x);
        d1 = (b & 1) ? -s : s;
        d2 = 0.0;
    }
    // value at l = m0 is d1 after init; advance to l (l > m0) and return d^l
    __device__ __forceinline__ double step(int l)
    {
        double v;
        if (l == 1) {
            v = cb * d1;
        } else {
            const double ir = isq[l * B + m0] * isq[l * B + an0];
            const double A = lc[3 * l] * ir;
            const double Bc = lc[3 * l + 1] * (double)(m0 * n0) * ir;
            const double Cc = lc[3 * l + 2] * sq[(l - 1) * B + m0] * sq[(l - 1) * B + an0] * ir;
            v = (A * cb - Bc) * d1 - Cc * d2;
        }
        d2 = d1;
        d1 = v;
        return v;
    }
};

// Normalised associated Legendre nP^l_{m0}(cos theta_k) (src/fdcl_FFTS2.cpp:48-100, NR 6.7):
//   nP(m,m) = (-1)^m sqrt((2m+1)/(4 pi) prod_{i<=m} (2i-1)/(2i)) sin^m(theta)
//   nP(l,m) = fact_l (x nP(l-1,m) - nP(l-2,m)/fact_{l-1}),  fact_l = sqrt((4l^2-1)/(l^2-m^2))
struct LegendreRec {
    int B, m0;
    double x, d1, d2;
    const double *fact, *ifact;
    __device__ __forceinline__ void init(const RecTables &t, int B_, int g, int k)
    {
        B = B_; m0 = g;
        x = t.cb[k];
        fact = t.sq; ifact = t.isq;
        double s = exp(t.lnbin[g] + (m0 ? m0 * t.lns[k] : 0.0));
        d1 = (m0 & 1) ? -s : s;
        d2 = 0.0;
    }
    __device__ __forceinline__ double step(int l)
    {
        double f = fact[l * B + m0];
        double v = f * (x * d1 - (l - 2 >= m0 ? d2 * ifact[(l - 1) * B + m0] : 0.0));
        d2 = d1;
        d1 = v;
        return v;
    }
};

// Recursion with its coefficients staged once per group.  The per-l coefficients depend on
// (m0, n0, l) only, so the workgroup computes them in parallel into LDS (one l per thread) and every
// beta lane then runs d^l = (a_l cos(beta) - b_l) d^{l-1} - c_l d^{l-2} reading broadcast LDS
// values instead of chasing table loads through the cache on its sequential path:
//   SO(3): a_l = l(2l-1) r, b_l = (2l-1)/(l-1) m0 n0 r, c_l = l/(l-1) sqrt((l-1)^2-m0^2) sqrt((l-1)^2-n0^2) r,
//          r = 1/sqrt((l^2-m0^2)(l^2-n0^2)); l = 1 (m0 = 0): d^1 = cos(beta) d^0.
//   S^2:   a_l = fact_l, b_l = 0, c_l = fact_l / fact_{l-1} (l - 2 >= m0), as LegendreRec.
struct RecCoef {
    double a, b, c;
};

__device__ __forceinline__ void group_coefs(const RecTables &t, int s2, int B, int m0, int n0, RecCoef *rc, int tid,
                                            int nthr)
{
    const int an0 = iabs(n0);
    for (int l = m0 + 1 + tid; l < B; l += nthr) {
        RecCoef r;
        if (s2) {
            const double f = t.sq[l * B + m0];
            r.a = f;
            r.b = 0.0;
            r.c = l - 2 >= m0 ? f * t.isq[(l - 1) * B + m0] : 0.0;
        } else if (l == 1) {
            r.a = 1.0;
            r.b = 0.0;
            r.c = 0.0;
        } else {
            const double ir = t.isq[l * B + m0] * t.isq[l * B + an0];
            r.a = t.lcoef[3 * l] * ir;
            r.b = t.lcoef[3 * l + 1] * (double)(m0 * n0) * ir;
            r.c = t.lcoef[3 * l + 2] * t.sq[(l - 1) * B + m0] * t.sq[(l - 1) * B + an0] * ir;
        }
        rc[l] = r;
    }
}

template <class F>
__device__ __forceinline__ void group_recursion_lds(const RecTables &t, const RecCoef *rc, int s2, int g, int m0,
                                                    int n0, int k, int lend, F f)
{
    const double cb = t.cb[k];
    double d1;
    if (s2) {
        const double s = exp(t.lnbin[g] + (m0 ? m0 * t.lns[k] : 0.0));
        d1 = (m0 & 1) ? -s : s;
    } else {
        const int a = m0 + n0, b = m0 - n0;
        const double s = exp(t.lnbin[g] + (a ? a * t.lnc[k] : 0.0) + (b ? b * t.lns[k] : 0.0));
        d1 = (b & 1) ? -s : s;
    }
    double d2 = 0.0;
    if (m0 < lend) f(m0, d1);
    for (int l = m0 + 1; l < lend; l++) {
        const RecCoef r = rc[l];
        const double v = (r.a * cb - r.b) * d1 - r.c * d2;
        d2 = d1;
        d1 = v;
        f(l, v);
    }
}

// ---------------------------------------------------------------------------------------------
// Contraction kernels (vector FMA path; FP32 or FP64 arithmetic)
// ---------------------------------------------------------------------------------------------
template <typename T> struct ConArgs {
    ChunkArgs ca;
    RecTables rt;
    const int2 *items;    // forward: (g, l_start); inverse: (g, k_start)
    int nitems;
    int tile;             // forward: rows per item (LT); inverse: k per item (KT)
    const void *src;      // forward: workspace S; inverse: coefficients
    void *dst;            // forward: coefficients; inverse: workspace G
    int s2;               // 1 for S^2 (two pairs, Legendre recursion)
    int pt;               // coefficients in PT[e][fi] (rows of C complex values) instead of packed
    const void *dtab;     // precomputed d^l_{m0 n0}(beta_k) per group ([l - m0][k], storage type), or null
    const long *dtab_off; // first element of each group's block in dtab
    int nsplit;           // FP64 MFMA kernels: workgroups sharing an item's blocks of 8 functions
};

// Table of d^l_{m0 n0}(beta_k) (S^2: nP^l_{m0}(cos theta_k)) for every group, k and l >= m0, laid out
// [g][l - m0][k]: one thread per (g, k) runs the recursion; consecutive threads write consecutive k.
template <typename T>
__global__ void build_dtab(RecTables rt, int B, int N, int s2, const long *goff, T *tab)
{
    const int ngrp = s2 ? B : B * B;
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= (long)ngrp * N) return;
    const int g = (int)(e / N), k = (int)(e - (long)g * N);
    int m0, n0;
    if (s2) { m0 = g; n0 = 0; } else so3_group(g, m0, n0);
    T *row = tab + goff[g] + k;
    group_recursion(rt, s2, B, g, m0, n0, k, B, [&](int l, double v) { row[(long)(l - m0) * N] = (T)v; });
}

// 16-byte native vectors (clang ext_vector_type; HIP's float4 class copies lower to memcpy and
// would keep register arrays of them in scratch)
typedef float nf4 __attribute__((ext_vector_type(4)));
typedef double nd2 __attribute__((ext_vector_type(2)));
template <typename T> struct vec16;
template <> struct vec16<float> { typedef nf4 type; static constexpr int n = 4; };
template <> struct vec16<double> { typedef nd2 type; static constexpr int n = 2; };

// Run the recursion of group g for one beta_k and hand every (l, value) with l in [m0, lend) to f.
template <typename F>
__device__ __forceinline__ void group_recursion(const RecTables &rt, int s2, int B, int g, int m0, int n0, int k,
                                                int lend, F f)
{
    if (s2) {
        LegendreRec r;
        r.init(rt, B, g, k);
        double v = r.d1;
        for (int l = m0; l < lend; l++) {
            if (l > m0) v = r.step(l);
            f(l, v);
        }
    } else {
        WignerRec r;
        r.init(rt, B, g, m0, n0, k);
        double v = r.d1;
        for (int l = m0; l < lend; l++) {
            if (l > m0) v = r.step(l);
            f(l, v);
        }
    }
}

// Forward: F^l_{pair} = s_p sum_k w_k d^l_{m0 n0}(beta_k) S_k(pair), l in [l_start, l_start+LT).
// Block = 32*C threads: ncol real columns (fi, p, re/im) x row groups of 4 rows (LT = 4 * groups).
// The spectrum tile of the group is streamed through LDS in chunks of up to 128 beta values; the
// first chunk's loads are issued before the FP64 recursion so their latency hides behind it, and
// each further chunk is loaded while the previous one is multiplied.
template <typename T, int C>
__global__ void __launch_bounds__(32 * C) contract_fwd(ConArgs<T> a)
{
    typedef typename vec16<T>::type VT;
    constexpr int V = vec16<T>::n;
    constexpr int NT = 32 * C;
    constexpr int KCMAX = 128;
    constexpr int PF = KCMAX * 8 * C / (NT * V);   // vector loads per thread per chunk (SO(3) bound)
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, LT = a.tile;
    const int unit = xcd_unit(a.nitems);
    if (unit >= a.nitems) return;
    const int2 it = a.items[unit];
    const int g = it.x, l0 = it.y;
    const int np = a.s2 ? 2 : 4;
    const int ncol = 2 * np * C;             // real columns
    int m0, n0;
    if (a.s2) { m0 = g; n0 = 0; } else so3_group(g, m0, n0);
    const int lend = min(B, l0 + LT);
    const int KC = min(N, KCMAX);
    T *D = reinterpret_cast<T *>(smem_raw);          // [k][LT]
    T *Ss = D + (size_t)N * LT;                      // [KC][ncol]
    const int tid = threadIdx.x;
    const VT *Sv = reinterpret_cast<const VT *>(reinterpret_cast<const T *>(a.src) + (size_t)g * N * ncol);
    VT r[PF];
#define CONTRACT_FWD_LOAD(kc_)                                               \
    {                                                                        \
        const int nv_ = min(KC, N - (kc_)) * ncol / V;                       \
        const VT *src_ = Sv + (size_t)(kc_) * ncol / V;                      \
        _Pragma("unroll") for (int i = 0; i < PF; i++)                       \
        {                                                                    \
            int e = tid + i * NT;                                            \
            if (e < nv_) r[i] = src_[e];                                     \
        }                                                                    \
    }
    CONTRACT_FWD_LOAD(0)
    if (a.dtab) {
        // D[k][l-l0] = w_k d^l(beta_k) from the precomputed table (reads contiguous in k)
        const T *tab = reinterpret_cast<const T *>(a.dtab) + a.dtab_off[g];
        for (int e = tid; e < LT * N; e += NT) {
            const int lr = e / N, k = e - lr * N, l = l0 + lr;
            D[(size_t)k * LT + lr] = l < lend ? (T)(a.rt.w[k] * (double)tab[(long)(l - m0) * N + k]) : (T)0;
        }
    } else {
        // FP64 recursion, one lane per k; D[k][l-l0] = w_k d^l(beta_k)
        for (int k = tid; k < N; k += NT) {
            const double wk = a.rt.w[k];
            T *Dk = D + (size_t)k * LT;
            group_recursion(a.rt, a.s2, B, g, m0, n0, k, lend, [&](int l, double v) {
                if (l >= l0) Dk[l - l0] = (T)(wk * v);
            });
            for (int l = max(lend, l0); l < l0 + LT; l++) Dk[l - l0] = (T)0;
        }
    }
    const int rows_per_thread = 4;
    const int ngroups = NT / ncol;
    const int col = tid % ncol, rg = tid / ncol;
    T acc[rows_per_thread] = {};
    for (int kc = 0; kc < N; kc += KC) {
        const int kn = min(KC, N - kc);
        __syncthreads();
        {
            const int nv = kn * ncol / V;
            VT *dst = reinterpret_cast<VT *>(Ss);
#pragma unroll
            for (int i = 0; i < PF; i++) {
                int e = tid + i * NT;
                if (e < nv) dst[e] = r[i];
            }
        }
        __syncthreads();
        if (kc + KC < N) CONTRACT_FWD_LOAD(kc + KC)
        if (rg < ngroups) {
            for (int kk = 0; kk < kn; kk++) {
                T s = Ss[kk * ncol + col];
                const T *dk = D + (size_t)(kc + kk) * LT + rg * rows_per_thread;
#pragma unroll
                for (int q = 0; q < rows_per_thread; q++) acc[q] += dk[q] * s;
            }
        }
    }
#undef CONTRACT_FWD_LOAD
    if (rg >= ngroups) return;
    const int ri = col & 1, p = (col >> 1) % np, fi = col / (2 * np);
    if (fi >= ca.nfc) return;
    int m, n, sp;
    bool valid;
    if (a.s2) {
        m = p ? -m0 : m0; n = 0;
        sp = p ? parity_sign(m0) : 1;
        valid = (p == 0) || (m0 > 0);
    } else {
        so3_pair(m0, n0, p, m, n);
        sp = so3_pair_sign(m0, n0, p);
        valid = so3_pair_valid(m0, n0, p);
    }
    if (!valid) return;
    T *out = reinterpret_cast<T *>(a.dst) + (a.pt ? 2 * fi : 2 * ca.coef_off[ca.fn[fi]]);
    const long es = a.pt ? 2L * C : 2L;     // PT rows hold the chunk's C functions
#pragma unroll
    for (int q = 0; q < rows_per_thread; q++) {
        int l = l0 + rg * rows_per_thread + q;
        if (l < lend) {
            long idx = a.s2 ? s2_idx(l, m) : so3_idx(l, m, n);
            out[es * idx + ri] = (T)sp * acc[q];
        }
    }
}

// Inverse: G_k(pair) = s_p sum_l c_l d^l_{m0 n0}(beta_k) F^l(pair), c_l = 2l+1 (SO(3)) or 1 (S^2),
// k in [k_start, k_start+KT).  Block = 32*C threads: ncc complex columns x k groups of 4.
// The coefficient gather (scattered complex loads) is issued before the FP64 recursion.
template <typename T, int C>
__global__ void __launch_bounds__(32 * C) contract_inv(ConArgs<T> a)
{
    typedef typename cplx_t<T>::type T2;
    constexpr int NT = 32 * C;
    constexpr int PF = 16;                      // gathered complex values per thread per pass
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, KT = a.tile;
    const int unit = xcd_deal((a.ca.N + a.tile - 1) / a.tile);  // runs: the k tiles of one group
    if (unit >= a.nitems) return;
    const int2 it = a.items[unit];
    const int g = it.x, k0 = it.y;
    const int np = a.s2 ? 2 : 4;
    const int ncc = np * C;                     // complex columns (fi, p)
    int m0, n0;
    if (a.s2) { m0 = g; n0 = 0; } else so3_group(g, m0, n0);
    const int M = B - m0;
    const int kn = min(KT, N - k0);
    T *D = reinterpret_cast<T *>(smem_raw);                          // [l-m0][KT]
    T2 *Fv = reinterpret_cast<T2 *>(D + (size_t)B * KT);             // [l-m0][ncc]
    const int tid = threadIdx.x;
    const int total = M * ncc;
    T2 r[PF];
    auto gather = [&](int base) {
#pragma unroll
        for (int i = 0; i < PF; i++) {
            int e = base + tid + i * NT;
            if (e < total) {
                int lr = e / ncc, cc = e - lr * ncc, l = m0 + lr;
                int p = cc % np, fi = cc / np;
                T2 val = cmk<T2>((T)0, (T)0);
                if (fi < ca.nfc) {
                    int m, n, sp;
                    if (a.s2) { m = p ? -m0 : m0; n = 0; sp = p ? parity_sign(m0) : 1; }
                    else { so3_pair(m0, n0, p, m, n); sp = so3_pair_sign(m0, n0, p); }
                    const long idx = a.s2 ? s2_idx(l, m) : so3_idx(l, m, n);
                    const T2 f = a.pt ? reinterpret_cast<const T2 *>(a.src)[idx * C + fi]
                                      : reinterpret_cast<const T2 *>(a.src)[ca.coef_off[ca.fn[fi]] + idx];
                    val = cmk<T2>((T)sp * f.x, (T)sp * f.y);
                }
                r[i] = val;
            }
        }
    };
    auto deposit = [&](int base) {
#pragma unroll
        for (int i = 0; i < PF; i++) {
            int e = base + tid + i * NT;
            if (e < total) Fv[e] = r[i];
        }
    };
    gather(0);
    if (a.dtab) {
        const T *tab = reinterpret_cast<const T *>(a.dtab) + a.dtab_off[g];
        for (int e = tid; e < M * KT; e += NT) {
            const int lr = e / KT, kk = e - lr * KT;
            const double cl = a.s2 ? 1.0 : (double)(2 * (m0 + lr) + 1);
            D[(size_t)lr * KT + kk] = kk < kn ? (T)(cl * (double)tab[(long)lr * N + k0 + kk]) : (T)0;
        }
    } else {
        for (int kk = tid; kk < KT; kk += NT) {
            if (kk >= kn) {
                for (int l = m0; l < B; l++) D[(size_t)(l - m0) * KT + kk] = (T)0;
                continue;
            }
            group_recursion(a.rt, a.s2, B, g, m0, n0, k0 + kk, B, [&](int l, double v) {
                D[(size_t)(l - m0) * KT + kk] = (T)(a.s2 ? v : (double)(2 * l + 1) * v);
            });
        }
    }
    deposit(0);
    for (int base = PF * NT; base < total; base += PF * NT) {
        gather(base);
        deposit(base);
    }
    __syncthreads();
    const int ngroups = NT / ncc;
    const int cc = tid % ncc, kg = tid / ncc;
    if (kg >= ngroups) return;
    T2 *G = reinterpret_cast<T2 *>(a.dst) + (size_t)g * N * ncc;
    for (int kb = kg * 4; kb < kn; kb += ngroups * 4) {
        T2 acc[4];
#pragma unroll
        for (int j = 0; j < 4; j++) acc[j] = cmk<T2>((T)0, (T)0);
        for (int lr = 0; lr < M; lr++) {
            T2 f = Fv[lr * ncc + cc];
            const T *dl = D + (size_t)lr * KT + kb;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                acc[j].x += dl[j] * f.x;
                acc[j].y += dl[j] * f.y;
            }
        }
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (kb + j < kn) G[(size_t)(k0 + kb + j) * ncc + cc] = acc[j];
    }
}

// ---------------------------------------------------------------------------------------------
// Real <-> complex coefficient bases (per-l unitary T, at most 2 nonzeros per row)
// ---------------------------------------------------------------------------------------------
// SO(3) forward epilogue: Fr(l,m,n) = Re sum_{a=+-m, b=+-n} T(m,a) Fc(a,b) conj T(n,b)
//   (src/fdcl_FFTSO3.cpp:1169-1181).   Inverse prologue: Fc(a,b) = sum_{m=+-a,n=+-b} conj T(m,a) Fr(m,n) T(n,b).
// orbits {(+-a, +-b)} of every degree of a bandlimit, per function
__host__ __device__ inline long so3_tmix_orbits(int B) { return (long)B * (B + 1) * (2 * B + 1) / 6; }

// element e (an orbit of one function) of so3_tmix
template <typename T, int DIR>
__device__ __forceinline__ void so3_tmix_elem(const ChunkArgs &ca, const void *src, void *dst, long ws_stride, long e)
{
    typedef typename cplx_t<T>::type T2;
    // one thread per orbit {(+-a, +-b)} of a degree l (a, b in [0, l]): its (up to) 4 inputs are
    // read once and its (up to) 4 outputs written; orbits of l start at l (l+1) (2l+1) / 6
    const int B = ca.B;
    const long no = so3_tmix_orbits(B);
    auto orb_off = [](int l) { return l * (l + 1) * (2 * l + 1) / 6; };
    {
        const int fi = (int)(e / no);
        const int idx = (int)(e - (long)fi * no);
        int l = (int)cbrtf((float)idx * 3.0f);
        while (orb_off(l + 1) <= idx) l++;
        while (orb_off(l) > idx) l--;
        const int r = idx - orb_off(l), a = r % (l + 1), b = r / (l + 1);
        const int na = a ? 2 : 1, nb = b ? 2 : 1;
        const long fo = ca.coef_off[ca.fn[fi]];
        if (DIR == 0) {  // complex workspace -> real output
            const T2 *Fc = reinterpret_cast<const T2 *>(src) + (long)fi * ws_stride;
            T2 f[2][2];
            for (int ia = 0; ia < na; ia++)
                for (int ib = 0; ib < nb; ib++) f[ia][ib] = Fc[so3_idx(l, ia ? -a : a, ib ? -b : b)];
            T *out = reinterpret_cast<T *>(dst) + fo;
            for (int io = 0; io < na; io++) {
                const int m = io ? -a : a;
                for (int jo = 0; jo < nb; jo++) {
                    const int n = jo ? -b : b;
                    T acc = (T)0;
                    for (int ia = 0; ia < na; ia++) {
                        const T2 tm = tmat<T2>(m, ia ? -a : a);
                        for (int ib = 0; ib < nb; ib++) {
                            const T2 tn = tmat<T2>(n, ib ? -b : b);
                            const T2 x = cmul(cmul(tm, f[ia][ib]), cmk<T2>(tn.x, -tn.y));
                            acc += x.x;
                        }
                    }
                    out[so3_idx(l, m, n)] = acc;
                }
            }
        } else {         // real input -> complex workspace
            const T *Fr = reinterpret_cast<const T *>(src) + fo;
            T f[2][2];
            for (int ia = 0; ia < na; ia++)
                for (int ib = 0; ib < nb; ib++) f[ia][ib] = Fr[so3_idx(l, ia ? -a : a, ib ? -b : b)];
            T2 *out = reinterpret_cast<T2 *>(dst) + (long)fi * ws_stride;
            for (int io = 0; io < na; io++) {
                const int m = io ? -a : a;
                for (int jo = 0; jo < nb; jo++) {
                    const int n = jo ? -b : b;
                    // Fc(m,n) = sum over rows mm in {m,-m}, cols nn in {n,-n} of conj T(mm,m) Fr(mm,nn) T(nn,n)
                    T2 acc = cmk<T2>((T)0, (T)0);
                    for (int ia = 0; ia < na; ia++) {
                        const T2 tm = tmat<T2>(ia ? -a : a, m);
                        for (int ib = 0; ib < nb; ib++) {
                            const T2 tn = tmat<T2>(ib ? -b : b, n);
                            const T fv = f[ia][ib];
                            acc = cadd(acc, cmul(cmk<T2>(tm.x * fv, -tm.y * fv), tn));
                        }
                    }
                    out[so3_idx(l, m, n)] = acc;
                }
            }
        }
    }
}

template <typename T, int DIR>
__global__ void so3_tmix(ChunkArgs ca, const void *src, void *dst, long ws_stride)
{
    const long total = so3_tmix_orbits(ca.B) * ca.nfc;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x)
        so3_tmix_elem<T, DIR>(ca, src, dst, ws_stride, e);
}

// Inverse prologue of real SO(3) plans straight into the batch-wide coefficient staging
// PT[e][function] (row stride C): the real mixing fused with the layout pass, so the complex packed
// coefficients never go to memory.  Unit = (degree l, row pair n = +-b) x a tile of 16 functions
// (one 128-byte PT line per packed index): Fc(l, m, n) = sum conj T(a, m) Fr(a, c) T(c, n) over
// a = +-|m|, c = +-|n| (src/fdcl_FFTSO3.cpp:1169-1181, as so3_tmix) takes only the rows n = b and
// n = -b of the degree block.  Each thread reads (row, m) positions of the pair for the tile's
// functions into LDS; then 16 lanes per packed index form Fc and write whole PT lines.  (The
// forward direction keeps coef_transpose + so3_tmix: a fused form wrote each function's real rows
// as 508-byte runs whose partial end lines other units complete at other times, and was slower.)
#define SCR_TF 16
__host__ __device__ inline int so3_coef_real_units(int B, int nfc) { return B * (B + 1) / 2 * ((nfc + SCR_TF - 1) / SCR_TF); }
__host__ __device__ inline size_t so3_coef_real_lds(int B) { return (size_t)2 * (2 * B - 1) * (SCR_TF + 1) * sizeof(float); }

__global__ void __launch_bounds__(256) so3_coef_real_inv(ChunkArgs ca, const float *in, float2 *PT)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int TF = SCR_TF, TP = TF + 1;
    const int ntf = (ca.nfc + TF - 1) / TF, nunit = so3_coef_real_units(ca.B, ca.nfc);
    const int u = xcd_unit(nunit);
    if (u >= nunit) return;
    const int lb = u / ntf, f0 = (u - lb * ntf) * TF;
    int l = (int)((sqrtf(8.f * (float)lb + 1.f) - 1.f) * 0.5f);
    while (l * (l + 1) / 2 > lb) l--;
    while ((l + 1) * (l + 2) / 2 <= lb) l++;
    const int b = lb - l * (l + 1) / 2, W = 2 * l + 1, R = b ? 2 : 1, nb = b ? 2 : 1;
    const int nf = min(TF, ca.nfc - f0);
    const long base = so3_off(l);
    const int t = threadIdx.x;
    // element (row r: n = b or -b, column mi = m + l) of the pair for function f0 + fl
    auto tix = [&](int r, int mi, int fl) { return (r * W + mi) * TP + fl; };
    float *tile = reinterpret_cast<float *>(smem_raw);
    for (int x = t; x < R * W; x += 256) {
        const int tr = x >= W ? 1 : 0, tmi = x - tr * W, n = tr ? -b : b;
        const long epos = base + tmi + (long)(n + l) * W;   // packed index of (l, m, n)
        for (int fl = 0; fl < nf; fl++) tile[tix(tr, tmi, fl)] = in[ca.coef_off[ca.fn[f0 + fl]] + epos];
    }
    __syncthreads();
    for (int i = t; i < R * W * TF; i += 256) {
        const int fl = i & (TF - 1), x = i / TF, r = x >= W ? 1 : 0, mi = x - r * W;
        if (fl >= nf) continue;
        const int m = mi - l, n = r ? -b : b, a = m < 0 ? -m : m, na = a ? 2 : 1;
        float2 acc = make_float2(0.f, 0.f);
        for (int ia = 0; ia < na; ia++) {
            const int xa = ia ? -a : a;
            const float2 tm = tmat<float2>(xa, m);
            for (int ib = 0; ib < nb; ib++) {
                const float2 tn = tmat<float2>(ib ? -b : b, n);
                const float fv = tile[tix(ib, xa + l, fl)];
                acc = cadd(acc, cmul(make_float2(tm.x * fv, -tm.y * fv), tn));
            }
        }
        const long e = base + mi + (long)(n + l) * W;
        PT[e * ca.C + f0 + fl] = acc;
    }
}

// S^2: forward Fr(l,a) = Re sum_{c=+-a} conj T(a,c) Fc(l,c) (src/fdcl_FFTS2.cpp:424-436);
// inverse Fc(l,a) = sum_{c=+-a} T(c,a) Fr(l,c).
template <typename T, int DIR>
__global__ void s2_tmix(ChunkArgs ca, const void *src, void *dst, long ws_stride)
{
    typedef typename cplx_t<T>::type T2;
    const int B = ca.B;
    const long nc = (long)B * B;
    const long total = nc * ca.nfc;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        int fi = (int)(e / nc);
        int idx = (int)(e - (long)fi * nc);
        int l = (int)sqrtf((float)idx);
        while ((l + 1) * (l + 1) <= idx) l++;
        while (l * l > idx) l--;
        int a = idx - l * l - l;
        const long fo = ca.coef_off[ca.fn[fi]];
        if (DIR == 0) {
            const T2 *Fc = reinterpret_cast<const T2 *>(src) + (long)fi * ws_stride;
            T acc = (T)0;
            for (int ic = 0; ic < (a ? 2 : 1); ic++) {
                int c = ic ? -a : a;
                T2 t = tmat<T2>(a, c);
                T2 f = Fc[s2_idx(l, c)];
                acc += t.x * f.x + t.y * f.y;  // Re(conj(t) f)
            }
            reinterpret_cast<T *>(dst)[fo + idx] = acc;
        } else {
            const T *Fr = reinterpret_cast<const T *>(src) + fo;
            T2 acc = cmk<T2>((T)0, (T)0);
            for (int ic = 0; ic < (a ? 2 : 1); ic++) {
                int c = ic ? -a : a;
                T2 t = tmat<T2>(c, a);
                T f = Fr[s2_idx(l, c)];
                acc = cadd(acc, cmk<T2>(t.x * f, t.y * f));
            }
            reinterpret_cast<T2 *>(dst)[(long)fi * ws_stride + idx] = acc;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Pointwise evaluation at arbitrary rotations (SO(3)) or directions (S^2): one workgroup per point,
// threads over the (m,n) groups, FP64 recursion at the point's beta and FP64 accumulation
//   SO(3): f = sum_{m,n} exp(-i(m alpha + n gamma)) sum_l (2l+1) F^l_mn d^l_mn(beta)
//          (src/fdcl_FFTSO3.cpp:498-531)
//   S^2:   f = sum_m exp(i m phi) sum_l F^l_m nP^l_|m|(cos theta) (-1)^m for m < 0
//          (src/fdcl_FFTS2.cpp:194-230)
// beta (theta) is reduced to [0, pi] first: d(beta) = d(beta + 2 pi), d(2 pi - b) = (-1)^(m-n) d(b)
// (S^2: nP^l_m(cos(2 pi - t)) carries (-1)^m through sin^m).
// ---------------------------------------------------------------------------------------------
template <typename T> struct EvalArgs {
    int B;
    RecTables rt;
    const void *F;        // complex coefficients of the function (packed), storage type T
    const double *ang;    // SO(3): (alpha, beta, gamma) per point; S^2: (theta, phi)
    long npts;
    void *out;            // complex values, or their real parts when real != 0
    int real;
};

template <typename T, int NP>
__global__ void __launch_bounds__(256) eval_points(EvalArgs<T> a)
{
    typedef typename cplx_t<T>::type T2;
    __shared__ double sh[3];
    __shared__ double red[2][4];
    const int B = a.B, ngrp = NP == 2 ? B : B * B, tid = threadIdx.x;
    const T2 *F = reinterpret_cast<const T2 *>(a.F);
    for (long pt = blockIdx.x; pt < a.npts; pt += gridDim.x) {
        double al, be, ga;
        if (NP == 2) { be = a.ang[2 * pt]; al = a.ang[2 * pt + 1]; ga = 0.0; }
        else { al = a.ang[3 * pt]; be = a.ang[3 * pt + 1]; ga = a.ang[3 * pt + 2]; }
        const double tp = 2.0 * M_PI;
        be = fmod(be, tp);
        if (be < 0) be += tp;
        const bool fold = be > M_PI;
        if (fold) be = tp - be;
        if (tid == 0) {
            sh[0] = cos(be);
            sh[1] = log(cos(0.5 * be));
            sh[2] = NP == 2 ? log(sin(be)) : log(sin(0.5 * be));  // the tables' S^2 seed uses sin(theta)
        }
        __syncthreads();
        RecTables t = a.rt;
        t.cb = sh;
        t.lnc = sh + 1;
        t.lns = sh + 2;
        double re = 0.0, im = 0.0;
        for (int g = tid; g < ngrp; g += 256) {
            int m0, n0;
            if (NP == 2) { m0 = g; n0 = 0; } else so3_group(g, m0, n0);
            double ar[NP], ai[NP];
#pragma unroll
            for (int p = 0; p < NP; p++) { ar[p] = 0.0; ai[p] = 0.0; }
            int pm[NP], pn[NP];
#pragma unroll
            for (int p = 0; p < NP; p++) {
                if (NP == 2) { pm[p] = p ? -m0 : m0; pn[p] = 0; }
                else so3_pair(m0, n0, p, pm[p], pn[p]);
            }
            group_recursion(t, NP == 2, B, g, m0, n0, 0, B, [&](int l, double v) {
                const double c = NP == 2 ? v : (double)(2 * l + 1) * v;
#pragma unroll
                for (int p = 0; p < NP; p++) {
                    const T2 f = F[NP == 2 ? s2_idx(l, pm[p]) : so3_idx(l, pm[p], pn[p])];
                    ar[p] += c * (double)f.x;
                    ai[p] += c * (double)f.y;
                }
            });
#pragma unroll
            for (int p = 0; p < NP; p++) {
                const bool valid = NP == 2 ? (p == 0 || m0 > 0) : so3_pair_valid(m0, n0, p);
                if (!valid) continue;
                double sgn = NP == 2 ? (p ? (double)parity_sign(m0) : 1.0) : (double)so3_pair_sign(m0, n0, p);
                if (fold) sgn *= (double)parity_sign(NP == 2 ? pm[p] : pm[p] - pn[p]);
                // SO(3): exp(-i(m alpha + n gamma)); S^2: exp(+i m phi)
                const double ph = NP == 2 ? (double)pm[p] * al : -((double)pm[p] * al + (double)pn[p] * ga);
                double sp, cp;
                sincos(ph, &sp, &cp);
                re += sgn * (ar[p] * cp - ai[p] * sp);
                im += sgn * (ar[p] * sp + ai[p] * cp);
            }
        }
        // workgroup sum: waves by shuffles, then across the 4 waves through LDS
        for (int o = 32; o > 0; o >>= 1) {
            re += __shfl_down(re, o);
            im += __shfl_down(im, o);
        }
        if ((tid & 63) == 0) { red[0][tid >> 6] = re; red[1][tid >> 6] = im; }
        __syncthreads();
        if (tid == 0) {
            const double r = red[0][0] + red[0][1] + red[0][2] + red[0][3];
            const double i = red[1][0] + red[1][1] + red[1][2] + red[1][3];
            if (a.real) reinterpret_cast<T *>(a.out)[pt] = (T)r;
            else reinterpret_cast<T2 *>(a.out)[pt] = cmk<T2>((T)r, (T)i);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------------
// Rotation of spherical-harmonic coefficients by many rotations (FP64):
//   complex: G^l_a = sum_b D^l_ab(R) F^l_b, D^l_ab = exp(-i a alpha) d^l_ab(beta) exp(-i b gamma)
//            (wigner_D, src/fdcl_FFTSO3.cpp:201-224)
//   real:    G^l = U^l(R) F^l with U^l = conj(T^l) D^l (T^l)^T (real_harmonics,
//            src/fdcl_FFTSO3.cpp:1062-1105; example/example3.cpp:50-53)
// One workgroup per rotation.  Threads take the sets (m0, n0), m0 >= n0 >= 0, largest first, and
// run the recursions of the canonical groups (m0, n0) and (m0, -n0), which give d^l at the 8
// index pairs (+-m0, +-n0) and (+-n0, +-m0) (the symmetries of common.hpp).  Each quad
// {(+-a, +-b)} closes under T, so its U (or D) entries are l-independent combinations of its 4 d
// values, formed once per set; G^l accumulates in LDS with FP64 atomic adds, for the degrees
// [l0, l1) of the launch (the LDS holds their coefficients).
// ---------------------------------------------------------------------------------------------
struct RotArgs {
    int B;
    RecTables rt;         // SO(3) tables of the plan's class (cb / lnc / lns replaced per rotation)
    const double *F;      // B^2 coefficients, real or complex interleaved
    const double *ang;    // (alpha, beta, gamma) per rotation
    long nrot;
    double *G;            // nrot x B^2 coefficients
    int l0, l1;           // degrees of this launch
};

// the (up to) 4 index pairs of a quad {(+-a, +-b)}, a, b >= 0: index i over ms = {a, -a}, j over
// ns = {b, -b}; na = 1 when a = 0 (one row), nb = 1 when b = 0
struct RotQuad {
    int a, b, na, nb;
    double kr[2][2][2][2];   // real: U(ms[i], ns[j]) = sum_{i',j'} kr[i][j][i'][j'] d(ms[i'], ns[j'])
    double pr[2][2], pi[2][2];  // complex: D(ms[i], ns[j]) = (pr + i pi)[i][j] d(ms[i], ns[j])
    __device__ __forceinline__ void init(int a_, int b_, double al, double ga)
    {
        a = a_; b = b_; na = a ? 2 : 1; nb = b ? 2 : 1;
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) {
                const int mi = i ? -a : a, nj = j ? -b : b;
                double s, c;
                sincos(-((double)mi * al + (double)nj * ga), &s, &c);
                pr[i][j] = c;
                pi[i][j] = s;
            }
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) {
                const int mi = i ? -a : a, nj = j ? -b : b;
                for (int i2 = 0; i2 < 2; i2++)
                    for (int j2 = 0; j2 < 2; j2++) {
                        const int mc = i2 ? -a : a, ne = j2 ? -b : b;
                        // conj(T(mi, mc)) T(nj, ne) exp(-i(mc alpha + ne gamma)): real part
                        const double2 t1 = tmat<double2>(mi, mc), t2 = tmat<double2>(nj, ne);
                        const double xr = t1.x * t2.x + t1.y * t2.y, xi = t1.x * t2.y - t1.y * t2.x;
                        kr[i][j][i2][j2] = (i2 < na && j2 < nb) ? xr * pr[i2][j2] - xi * pi[i2][j2] : 0.0;
                    }
            }
    }
    // accumulate the quad's contributions at degree l: dv[i][j] = d^l(ms[i], ns[j])
    template <bool REAL>
    __device__ __forceinline__ void add(double *acc, int e0, int l, const double (&dv)[2][2], const double *F) const
    {
        const int base = l * l + l;
        for (int i = 0; i < na; i++) {
            const int mi = i ? -a : a;
            double gr = 0.0, gi = 0.0;
            for (int j = 0; j < nb; j++) {
                const int nj = j ? -b : b;
                if (REAL) {
                    double u = 0.0;
                    for (int i2 = 0; i2 < na; i2++)
                        for (int j2 = 0; j2 < nb; j2++) u += kr[i][j][i2][j2] * dv[i2][j2];
                    gr += u * F[base + nj];
                } else {
                    const double dr = pr[i][j] * dv[i][j], di = pi[i][j] * dv[i][j];
                    const double fr = F[2 * (base + nj)], fi = F[2 * (base + nj) + 1];
                    gr += dr * fr - di * fi;
                    gi += dr * fi + di * fr;
                }
            }
            if (REAL) atomicAdd(acc + (base + mi - e0), gr);
            else {
                atomicAdd(acc + 2 * (base + mi - e0), gr);
                atomicAdd(acc + 2 * (base + mi - e0) + 1, gi);
            }
        }
    }
};

template <bool REAL>
__global__ void __launch_bounds__(256) s2_rotate(RotArgs a)
{
    extern __shared__ double acc[];   // coefficients l0^2 .. l1^2 - 1 (complex: re, im)
    __shared__ double sh[3];
    constexpr int W = REAL ? 1 : 2;
    const int tid = threadIdx.x, B = a.B;
    const int e0 = a.l0 * a.l0, ne = W * (a.l1 * a.l1 - e0);
    const int nsets = a.l1 * (a.l1 + 1) / 2;          // (m0, n0), 0 <= n0 <= m0 < l1
    for (long r = blockIdx.x; r < a.nrot; r += gridDim.x) {
        const double tp = 2.0 * M_PI;
        const double al = a.ang[3 * r], ga = a.ang[3 * r + 2];
        double be = fmod(a.ang[3 * r + 1], tp);
        if (be < 0) be += tp;
        const bool fold = be > M_PI;           // d(2 pi - b) = (-1)^(m-n) d(b)
        if (fold) be = tp - be;
        for (int e = tid; e < ne; e += 256) acc[e] = 0.0;
        if (tid == 0) {
            sh[0] = cos(be);
            sh[1] = log(cos(0.5 * be));
            sh[2] = log(sin(0.5 * be));
        }
        __syncthreads();
        RecTables t = a.rt;
        t.cb = sh;
        t.lnc = sh + 1;
        t.lns = sh + 2;
        for (int u = tid; u < nsets; u += 256) {
            int m0 = (int)((sqrt(8.0 * u + 1.0) - 1.0) * 0.5);
            while (m0 * (m0 + 1) / 2 > u) m0--;
            while ((m0 + 1) * (m0 + 2) / 2 <= u) m0++;
            const int n0 = u - m0 * (m0 + 1) / 2;
            const double s = (double)parity_sign(m0 + n0);
            const double f = fold ? s : 1.0;    // (-1)^(m-n) for every pair of both quads
            RotQuad q1, q2;
            q1.init(m0, n0, al, ga);
            const bool two = n0 != m0;
            if (two) q2.init(n0, m0, al, ga);
            WignerRec rp, rm;
            rp.init(t, B, m0 * m0 + n0 + m0, m0, n0, 0);
            if (n0 > 0) rm.init(t, B, m0 * m0 - n0 + m0, m0, -n0, 0);
            for (int l = m0; l < a.l1; l++) {
                const double dp = l == m0 ? rp.d1 : rp.step(l);
                const double dm = n0 > 0 ? (l == m0 ? rm.d1 : rm.step(l)) : dp;
                if (l < a.l0) continue;
                // quad 1 over (+-m0) x (+-n0); quad 2 over (+-n0) x (+-m0)
                const double d1[2][2] = {{f * dp, f * dm}, {f * s * dm, f * s * dp}};
                q1.add<REAL>(acc, e0, l, d1, a.F);
                if (two) {
                    const double d2[2][2] = {{f * s * dp, f * dm}, {f * s * dm, f * dp}};
                    q2.add<REAL>(acc, e0, l, d2, a.F);
                }
            }
        }
        __syncthreads();
        double *G = a.G + r * (long)W * B * B + (long)W * e0;
        for (int e = tid; e < ne; e += 256) G[e] = acc[e];
        __syncthreads();
    }
}
