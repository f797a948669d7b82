// mfma64.hpp -- FP64 beta contractions of SO(3) on v_mfma_f64_16x16x4f64 (gfx950), batch-wide:
// the drop-in precision of the class layer, for batches (l_max <= 63).
//
// The workspace holds the whole class block-major: for each block fb of 8 functions, rows [g][k]
// of 8 functions x 4 pairs x (re, im) = 64 doubles, so the FFTs of 8 functions write (read) one
// contiguous region and a group's rows of a block are 64 KB in one piece. One workgroup takes a work item -- a group's 64 degrees (forward) or a group's tile of up to 64
// beta values (inverse) -- generates that block of Wigner-d values once (FP64 recursion with the
// group's recursion coefficients in LDS, or the d table), and then runs every block of 8
// functions (64 real columns) of its share of the batch through 16x16x4 FP64 MFMA tiles.
// Products, signs and epilogues are those of contract_fwd / contract_inv (kernels.hpp).
//
// Operand layout of v_mfma_f64_16x16x4f64, lane i: A[i % 16][i / 16] (16 rows x 4 reduction
// steps), B[i / 16][i % 16] (4 steps x 16 columns); the accumulator's element j is row
// 4 j + i / 16, column i % 16 (measured with one-hot operands; unlike the FP32 16x16 tiles, whose
// element j is row 4 (i / 16) + j).
//
// Workgroups of 512 threads (8 waves): 64 real columns are 4 column tiles, and wave w takes
// column tile w % 4 and the row tiles w / 4 and w / 4 + 2, so two waves share each SIMD. LDS
// pitches are chosen so that each MFMA operand read (16 consecutive doubles from each of 4 rows,
// half a wave per cycle) and each recursion write (consecutive lanes, consecutive k) is free of
// bank conflicts.
#pragma once

#include "kernels.hpp"

typedef double f64x4_t __attribute__((ext_vector_type(4)));

#define MX64_FB 8           // functions per column block
#define MX64_COLS 64        // real columns of a block (8 functions x 4 pairs x re/im)
#define MX64_THREADS 512
#define MX64_KC 64          // beta rows of the spectrum staged per step (forward)
#define MX64_SP 80          // LDS pitch (doubles) of spectrum / coefficient rows: 2 x 80 = 32 mod 64 banks
#define MX64_RT 4           // row tiles per work item (64 rows)

// Forward D block [l - l0][k], pitch N + 2 doubles (2 (N + 2) = 4 mod 64 banks for N = 0 mod 32;
// the two extra columns are the zero padding of N = 2 mod 4), then the spectrum step.  The group's
// recursion coefficients (B x 24 bytes) use the spectrum step's space before its first deposit.
__host__ __device__ inline int mx64_dpitch_fwd(int N) { return N + 2; }
__host__ __device__ inline size_t mx64_fwd_lds_bytes(int N)
{
    return ((size_t)16 * MX64_RT * mx64_dpitch_fwd(N) + (size_t)MX64_KC * MX64_SP) * sizeof(double);
}
// Inverse D block [l - m0][k - k0] (pitch KT + 16) and coefficient block [l - m0][col] (pitch 80),
// degree rows padded to a multiple of 4; the recursion coefficients use the coefficient block's
// space (at least 4 rows of 640 bytes: enough for B <= 106) before its first deposit.  80 KB at
// B = 64, KT = 64: two workgroups per CU.
__host__ __device__ inline size_t mx64_inv_lds_bytes(int B, int KT)
{
    const size_t M4 = ((size_t)B + 3) & ~(size_t)3;
    return (M4 * (KT + 16) + M4 * MX64_SP) * sizeof(double);
}

// blocks [fb0, fb1) of 8 functions for split part `part` of nsplit
__device__ __forceinline__ void mx64_blocks(int nfc, int nsplit, int part, int &fb0, int &fb1)
{
    const int nfb = (nfc + MX64_FB - 1) / MX64_FB;
    fb0 = (int)((long)part * nfb / nsplit);
    fb1 = (int)((long)(part + 1) * nfb / nsplit);
}

// Forward: F^l_{pair} = s_p sum_k w_k d^l_{m0 n0}(beta_k) S_k(pair) for l in [l0, l0 + 64); units
// (item (g, l0), split part), a.tile = 64.
__global__ void __launch_bounds__(MX64_THREADS) contract_fwd_mx64(ConArgs<double> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, LT = 16 * MX64_RT, DP = mx64_dpitch_fwd(N);
    const int unit = xcd_unit(a.nitems * a.nsplit);
    if (unit >= a.nitems * a.nsplit) return;
    const int item = unit / a.nsplit, part = unit - item * a.nsplit;
    int fb0, fb1;
    mx64_blocks(ca.nfc, a.nsplit, part, fb0, fb1);
    if (fb0 >= fb1) return;
    const int2 it = a.items[item];
    const int g = it.x, l0 = it.y;
    int m0, n0;
    so3_group(g, m0, n0);
    const int lend = min(B, l0 + LT), nrt = (lend - l0 + 15) / 16;
    double *D = reinterpret_cast<double *>(smem_raw);    // [LT][DP]
    double *Ss = D + (size_t)LT * DP;                      // [KC][SP]
    RecCoef *rc = reinterpret_cast<RecCoef *>(Ss);         // until the first deposit
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // block fb of group g: N rows of 32 double2 at ((fb B^2 + g) N) * 32
    const size_t blk2 = (size_t)B * B * N * (MX64_COLS / 2);
    const double2 *S2 = reinterpret_cast<const double2 *>(a.src) + (size_t)g * N * (MX64_COLS / 2);
    // step (fb, kc): KC beta rows x 32 double2 of block fb, four per thread, held in registers
    // until deposited; the next step's loads are in flight during the current step's MFMAs
    constexpr int NR = MX64_KC * (MX64_COLS / 2) / MX64_THREADS;
    double2 r[NR];
    auto load = [&](int fb, int kc) {
        const int nv = min(MX64_KC, N - kc) * (MX64_COLS / 2);
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const int e = tid + i * MX64_THREADS, row = e >> 5, c2 = e & 31;
            r[i] = e < nv ? S2[fb * blk2 + (size_t)(kc + row) * (MX64_COLS / 2) + c2] : make_double2(0.0, 0.0);
        }
    };
    load(fb0, 0);
    // D[l - l0][k] = w_k d^l(beta_k); rows past lend and columns N, N + 1 are zero
    for (int e = tid; e < LT * 2; e += MX64_THREADS) D[(e >> 1) * DP + N + (e & 1)] = 0.0;
    if (a.dtab) {
        const double *tab = reinterpret_cast<const double *>(a.dtab) + a.dtab_off[g];
        for (int e = tid; e < LT * N; e += MX64_THREADS) {
            const int lr = e / N, k = e - lr * N, l = l0 + lr;
            D[lr * DP + k] = l < lend ? a.rt.w[k] * tab[(long)(l - m0) * N + k] : 0.0;
        }
    } else {
        group_coefs(a.rt, 0, B, m0, n0, rc, tid, MX64_THREADS);
        __syncthreads();
        for (int k = tid; k < N; k += MX64_THREADS) {
            const double wk = a.rt.w[k];
            group_recursion_lds(a.rt, rc, 0, g, m0, n0, k, lend, [&](int l, double v) {
                if (l >= l0) D[(l - l0) * DP + k] = wk * v;
            });
            for (int l = max(lend, l0); l < l0 + LT; l++) D[(l - l0) * DP + k] = 0.0;
        }
    }
    const int ct = wave & 3, rt0 = wave >> 2;   // column tile; row tiles rt0, rt0 + 2
    const int rq = lane >> 4, cq = lane & 15;
    // epilogue constants: column 16 ct + cq of a block is (function 8 fb + fl, pair p, re/im ri);
    // rows [function][4 pairs], or [half][function][2 pairs] (half records, fft_lines.hpp)
    const int col = 16 * ct + cq, ri = col & 1;
    const int p = ca.hsplit ? 2 * (col >> 5) + ((col >> 1) & 1) : (col >> 1) & 3;
    const int fl = ca.hsplit ? (col >> 2) & 7 : col >> 3;
    const bool pvalid = so3_pair_valid(m0, n0, p);
    int m, n;
    so3_pair(m0, n0, p, m, n);
    const double sp = (double)so3_pair_sign(m0, n0, p);
    const long es = a.pt ? 2L * ca.C : 2L;
    for (int fb = fb0; fb < fb1; fb++) {
        f64x4_t acc[2];
#pragma unroll
        for (int t = 0; t < 2; t++) acc[t] = f64x4_t{0.0, 0.0, 0.0, 0.0};
        for (int kc = 0; kc < N; kc += MX64_KC) {
            const int kn = min(MX64_KC, N - kc);
            __syncthreads();   // D written / the previous step's operand reads done
#pragma unroll
            for (int i = 0; i < NR; i++) {
                const int e = tid + i * MX64_THREADS, row = e >> 5, c2 = e & 31;
                *reinterpret_cast<double2 *>(Ss + row * MX64_SP + 2 * c2) = r[i];
            }
            __syncthreads();
            if (kc + MX64_KC < N) load(fb, kc + MX64_KC);
            else if (fb + 1 < fb1) load(fb + 1, 0);
            const int ksteps = (kn + 3) >> 2;
            for (int k4 = 0; k4 < ksteps; k4++) {
                const double b = Ss[(4 * k4 + rq) * MX64_SP + 16 * ct + cq];
#pragma unroll
                for (int t = 0; t < 2; t++) {
                    const int rt = rt0 + 2 * t;
                    if (rt < nrt) {
                        const double av = D[(16 * rt + cq) * DP + kc + 4 * k4 + rq];
                        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b, acc[t], 0, 0, 0);
                    }
                }
            }
        }
        const int fi = MX64_FB * fb + fl;
        if (fi < ca.nfc && pvalid) {
            double *out = reinterpret_cast<double *>(a.dst) + (a.pt ? 2L * fi : 2 * ca.coef_off[ca.fn[fi]]);
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const int rt = rt0 + 2 * t;
                if (rt >= nrt) continue;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int l = l0 + 16 * rt + 4 * j + rq;
                    if (l < lend) out[es * so3_idx(l, m, n) + ri] = sp * acc[t][j];
                }
            }
        }
    }
}

// Inverse: G_k(pair) = s_p sum_l (2l+1) d^l_{m0 n0}(beta_k) F^l(pair) for k in [k0, k0 + KT);
// items (g, k0) with a.tile = KT (at most 64), nkt = N / KT tiles per group; units run
// (group, split part, k tile) with the k tiles fastest, dealt to the XCDs in runs of nkt so the
// tiles that gather the same coefficients share an L2.
__global__ void __launch_bounds__(MX64_THREADS) contract_inv_mx64(ConArgs<double> a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const ChunkArgs &ca = a.ca;
    const int B = ca.B, N = ca.N, KT = a.tile, DP = KT + 16, nkt = (N + KT - 1) / KT;
    const int unit = xcd_deal(nkt);
    if (unit >= a.nitems * a.nsplit) return;
    const int gi = unit / (a.nsplit * nkt), rem = unit - gi * a.nsplit * nkt;
    const int part = rem / nkt, item = gi * nkt + (rem - part * nkt);
    int fb0, fb1;
    mx64_blocks(ca.nfc, a.nsplit, part, fb0, fb1);
    if (fb0 >= fb1) return;
    const int2 it = a.items[item];
    const int g = it.x, k0 = it.y;
    int m0, n0;
    so3_group(g, m0, n0);
    const int M = B - m0, M4 = (M + 3) & ~3, kn = min(KT, N - k0), nrt = (kn + 15) / 16;
    double *D = reinterpret_cast<double *>(smem_raw);      // [M4][DP]
    double *Fv = D + (size_t)M4 * DP;                       // [M4][SP]: column 2 (fl 4 + p) + re/im
    RecCoef *rc = reinterpret_cast<RecCoef *>(Fv);         // until the first deposit
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // coefficient gather of block fb (scattered complex loads), held in registers until deposited;
    // M4 <= 64 rows x 32 complex columns = at most 4 per thread
    constexpr int NCC = MX64_FB * 4;
    constexpr int PF = 64 * NCC / MX64_THREADS;
    const int total = M4 * NCC;
    double2 q[PF];
    auto gather = [&](int fb) {
#pragma unroll
        for (int i = 0; i < PF; i++) {
            const int e = tid + i * MX64_THREADS;
            double2 v = make_double2(0.0, 0.0);
            if (e < total) {
                const int lr = e / NCC, cc = e - lr * NCC;
                const int p = ca.hsplit ? 2 * (cc >> 4) + (cc & 1) : cc & 3;
                const int fi = MX64_FB * fb + (ca.hsplit ? (cc >> 1) & 7 : cc >> 2);
                if (lr < M && fi < ca.nfc) {
                    int m, n;
                    so3_pair(m0, n0, p, m, n);
                    const double sp = (double)so3_pair_sign(m0, n0, p);
                    const long idx = so3_idx(m0 + lr, m, n);
                    const double2 f = a.pt ? reinterpret_cast<const double2 *>(a.src)[idx * ca.C + fi]
                                           : reinterpret_cast<const double2 *>(a.src)[ca.coef_off[ca.fn[fi]] + idx];
                    v = make_double2(sp * f.x, sp * f.y);
                }
            }
            q[i] = v;
        }
    };
    gather(fb0);
    // D[l - m0][k - k0] = (2l+1) d^l(beta_k); degree rows past M and k past kn are zero
    for (int e = tid; e < (M4 - M) * KT; e += MX64_THREADS) D[(M + e / KT) * DP + e % KT] = 0.0;
    if (a.dtab) {
        const double *tab = reinterpret_cast<const double *>(a.dtab) + a.dtab_off[g];
        for (int e = tid; e < M * KT; e += MX64_THREADS) {
            const int lr = e / KT, kk = e - lr * KT;
            D[lr * DP + kk] = kk < kn ? (double)(2 * (m0 + lr) + 1) * tab[(long)lr * N + k0 + kk] : 0.0;
        }
    } else {
        group_coefs(a.rt, 0, B, m0, n0, rc, tid, MX64_THREADS);
        __syncthreads();
        for (int kk = tid; kk < KT; kk += MX64_THREADS) {
            if (kk >= kn) {
                for (int lr = 0; lr < M; lr++) D[lr * DP + kk] = 0.0;
                continue;
            }
            group_recursion_lds(a.rt, rc, 0, g, m0, n0, k0 + kk, B, [&](int l, double v) {
                D[(l - m0) * DP + kk] = (double)(2 * l + 1) * v;
            });
        }
    }
    const int ct = wave & 3, rt0 = wave >> 2;   // column tile; k tiles rt0, rt0 + 2
    const int rq = lane >> 4, cq = lane & 15, col = 16 * ct + cq;
    const size_t blk = (size_t)B * B * N * MX64_COLS;   // doubles per block of 8 functions
    double *G = reinterpret_cast<double *>(a.dst) + (size_t)g * N * MX64_COLS;
    for (int fb = fb0; fb < fb1; fb++) {
        __syncthreads();   // D written / the previous block's operand reads done
#pragma unroll
        for (int i = 0; i < PF; i++) {
            const int e = tid + i * MX64_THREADS;
            if (e < total) {
                const int lr = e / NCC, cc = e - lr * NCC;
                *reinterpret_cast<double2 *>(Fv + lr * MX64_SP + 2 * cc) = q[i];
            }
        }
        __syncthreads();
        if (fb + 1 < fb1) gather(fb + 1);   // in flight during this block's MFMAs
        f64x4_t acc[2];
#pragma unroll
        for (int t = 0; t < 2; t++) acc[t] = f64x4_t{0.0, 0.0, 0.0, 0.0};
        for (int l4 = 0; l4 < M4 / 4; l4++) {
            const double b = Fv[(4 * l4 + rq) * MX64_SP + 16 * ct + cq];
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const int rt = rt0 + 2 * t;
                if (rt < nrt) {
                    const double av = D[(4 * l4 + rq) * DP + 16 * rt + cq];
                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b, acc[t], 0, 0, 0);
                }
            }
        }
        // G[k][block columns]: 16 lanes write 128 contiguous bytes of a beta row of the block
        const bool live = MX64_FB * fb + (ca.hsplit ? (col >> 2) & 7 : col >> 3) < ca.nfc;
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int rt = rt0 + 2 * t;
            if (rt >= nrt) continue;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int kk = 16 * rt + 4 * j + rq;
                if (kk < kn && live) G[fb * blk + (size_t)(k0 + kk) * MX64_COLS + col] = acc[t][j];
            }
        }
    }
}
