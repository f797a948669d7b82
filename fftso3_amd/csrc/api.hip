// api.hip -- plans, chunked execution and the C ABI of libfftso3_amd.so (gfx950).
#include "kernels.hpp"
#include "mfma.hpp"
#include "fft128.hpp"
#include "mfma32.hpp"
#include "mfma_split.hpp"
#include "multi.hpp"
#include "mfma64.hpp"
#include "fft_lines.hpp"
#include "mfma_otf.hpp"

#include "../../include/fftso3_amd.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

namespace {

thread_local std::string g_last_error;

int fail(int status, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return status;
}

#define HIP_TRY(expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(FFTSO3_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                  \
    } while (0)

constexpr int kMaxLds = 160 * 1024;  // LDS per workgroup on gfx950
constexpr int kFftThreads = 1024;
constexpr int kNumStages = 6;

// fftso3_debug_fail_alloc: the countdown to the device allocation of this thread that fails
thread_local int g_fail_alloc_after = 0;

// device buffer owned by a plan; n is the size actually held (0 when nothing is)
struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int alloc(size_t bytes)
    {
        release();
        if (bytes == 0) return FFTSO3_OK;
        if (g_fail_alloc_after > 0 && --g_fail_alloc_after == 0)
            return fail(FFTSO3_ENOMEM, "hipMalloc(%zu) failed (forced by fftso3_debug_fail_alloc)", bytes);
        if (hipMalloc(&p, bytes) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return fail(FFTSO3_ENOMEM, "hipMalloc(%zu) failed", bytes);
        }
        n = bytes;
        return FFTSO3_OK;
    }
    // hold exactly `bytes` (keeps the current buffer when it already has that size)
    int resize(size_t bytes)
    {
        if (n == bytes && (p || bytes == 0)) return FFTSO3_OK;
        return alloc(bytes);
    }
    template <typename V> int upload(const std::vector<V> &v)
    {
        int st = alloc(v.size() * sizeof(V));
        if (st) return st;
        if (!v.empty()) HIP_TRY(hipMemcpy(p, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice));
        return FFTSO3_OK;
    }
};

// Device memory of fftso3_inverse_noalloc: the spectrum workspace and coefficient staging live
// here instead of in buffers of the plan.  Every GPU that loads the library holds its own copy
// until the process ends; host code must never pass the symbol itself to a kernel or a copy (on
// the host it is only a stand-in), only the per-device address from hipGetSymbolAddress.
constexpr size_t kNoallocScratchBytes = (size_t)128 << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

// Looks up this device's address of `scratch` (once per device, with the device current).
int scratch_setup(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) return fail(FFTSO3_EUNSUPPORTED, "device %d beyond %d", dev, kMaxDevices);
    if (g_scratch_done[dev]) return FFTSO3_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(FFTSO3_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
    g_scratch_done[dev] = true;
    return FFTSO3_OK;
}

// One bandlimit within a plan
struct Class {
    int B = 0, N = 0;
    std::vector<int> fns;           // global function indices with this bandlimit
    DevBuf d_fn, d_ident, d_ws_off;
    DevBuf d_tw32, d_tw64, d_rev;   // twiddles and DIF digit reversal
    FftDesc fd{};
    // recursion tables
    DevBuf d_cb, d_lnc, d_lns, d_w, d_lnbin, d_sq, d_isq, d_lcoef;
    DevBuf d_items_fwd, d_items_inv;
    bool mx64 = false;                       // batch-wide FP64 MFMA contractions (mfma64.hpp)
    bool prfft = false;                      // two-pass FFTs of fft_lines.hpp (workspace in half records)
    bool otf = false;                        // split-BF16 contractions of mfma_otf.hpp (l_max = 255, chunks of 8)
    DevBuf d_items_fwd64, d_items_inv64;
    int nitems_fwd64 = 0, nitems_inv64 = 0, kt64 = 0;
    int nitems_fwd = 0, nitems_inv = 0, lt_fwd = 0, kt_inv = 0;
    DevBuf d_items_fwd_mx, d_items_inv_mx;   // MFMA kernels: 64-row and 128-k tiles
    int nitems_fwd_mx = 0, nitems_inv_mx = 0;
    bool fused = false;             // whole slice fits in LDS
    int row_tile = 0, col_tile = 0; // two-pass fallback tiles (S^2 rings: divide N)
    int so3_tile = 0;               // SO(3) row / column passes (partial last block)
    bool mx = false;                // batch-wide FP32-MFMA contraction (whole class in one pass)
    int r16 = 0;                    // R of the N = 16R slice FFTs (fft_r16.hpp) of this class, or 0
    size_t ws_off = 0, fc_off = 0, tmp_off = 0;   // the class's region of the plan's buffers (unified plans)
    bool mxk = false;               // ... with the K-chunked kernels (128 < N, B <= 128)
    bool mxks = false;              // ... split-BF16 K-chunked kernels (128 < N, B <= 128)
    DevBuf d_items_mxs_f, d_items_mxs_i;   // their items: (group, degree range), (group, beta range)
    int nitems_mxs_f = 0, nitems_mxs_i = 0;
    bool pt = false;                // the contraction reads / writes transposed coefficients PT
    DevBuf d_dtab, d_dtab_off;      // precomputed d table of the chunked path at large B (or empty)
    int chunk = 8;                  // functions per pipeline pass for this class
    DevBuf d_items_mx;              // groups of the batch-wide path (largest first)
    int nitems_mx = 0;
};

}  // namespace

struct fftso3_plan_s {
    int domain = 0, kind = 0, prec = 0, n_fn = 0;
    int C = 8;                         // chunk capacity (functions per pipeline pass)
    size_t esz = 8;                    // bytes per complex element of storage (8 or 16)
    std::vector<int> Bs;
    std::vector<long> grid_off, coef_off;
    long grid_total = 0, coef_total = 0;
    DevBuf d_grid_off, d_coef_off;
    std::vector<Class *> classes;
    bool force_chunked = false;        // set_chunk() selects the chunked FP32 kernels
    DevBuf d_ws;        // spectrum workspace (rows of (group, beta index))
    DevBuf d_scratch;   // two-pass FFT scratch
    DevBuf d_fc;        // complex coefficients of the real path / transposed coefficients PT
    DevBuf d_dummy;     // store sink of the batch-wide contraction (lanes with nothing to store)
    DevBuf d_tmp;       // packed complex coefficients of the batch-wide real path
    // false while the plan's tables and workspaces are not all in place (a failed finalize_plan,
    // e.g. out of memory in set_chunk or autotune): transforms refuse to run on such a plan
    bool valid = false;
    int device = 0;                    // the HIP device current when the plan was created
    int klo = 0;                       // row blocking of the batch-wide workspace (ws_row)
    int gil = 1;                       // group interleave of the batch-wide workspace (ws_rec)
    int split_mul = 1;                 // multiplier of the automatic group split (batch-wide kernels)
    // unified schedule: every bandlimit class in one launch per stage (multi.hpp), each class in
    // its own region of the workspaces; the queues of (entry, unit) items per stage
    bool multi_on = true;              // allowed (autotune may switch it off)
    bool uni_pc = false;               // unified FFT and layout stages, contractions launched per class
    bool unified = false;              // in use for the current schedule
    DevBuf d_q_fft, d_q_cf, d_q_ci, d_q_tr;
    int nq_fft = 0, nq_cf = 0, nq_ci = 0, nq_tr = 0;
    int ql_fft[XCD_COUNT] = {}, ql_cf[XCD_COUNT] = {}, ql_ci[XCD_COUNT] = {}, ql_tr[XCD_COUNT] = {};
    DevBuf d_ctr;                      // queue counters of the four stage kernels (MULTI_CTRS each, 64 B apart)
    std::vector<Class *> uorder;                       // classes, largest bandlimit first
    std::vector<std::pair<Class *, int>> ent_cf, ent_ci;   // contraction entries (class, kind)
    // graph mode: one instantiated graph per (direction, input, output, stream)
    bool graph = false;
    struct GraphEntry {
        bool fwd;
        const void *in;
        void *out;
        hipStream_t s;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    void drop_graphs()
    {
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.exec);
        graphs.clear();
    }
    // timing
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
    std::vector<std::pair<int, int>> ev_used;   // (stage, pool index)
    double acc_ms[kNumStages] = {};
    long acc_n[kNumStages] = {};
    ~fftso3_plan_s()
    {
        drop_graphs();
        for (auto *c : classes) delete c;
        for (auto &e : ev_pool) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    }
};

namespace {

// ---------------------------------------------------------------------------------------------
// tables
// ---------------------------------------------------------------------------------------------
FftDesc make_fft_desc(int N, std::vector<int> &rev)
{
    FftDesc fd{};
    fd.N = N;
    fd.lgN = (N & (N - 1)) == 0 ? __builtin_ctz(N) : -1;
    int n = N, e = 0;
    while (n % 2 == 0) { n /= 2; e++; }
    std::vector<int> rad;
    int n8 = e / 3, r = e % 3;
    if (r == 1 && n8 >= 1) { n8 -= 1; for (int i = 0; i < n8; i++) rad.push_back(8); rad.push_back(4); rad.push_back(4); }
    else {
        for (int i = 0; i < n8; i++) rad.push_back(8);
        if (r == 2) rad.push_back(4);
        if (r == 1) rad.push_back(2);
    }
    for (int p : {3, 5, 7})
        while (n % p == 0) { n /= p; rad.push_back(p); }
    rev.assign(N, 0);
    if (n != 1 || (int)rad.size() > FFT_MAX_STAGES) {  // prime factor > 7: direct DFT
        fd.nst = 0;
        for (int f = 0; f < N; f++) rev[f] = f;
        return fd;
    }
    fd.nst = (int)rad.size();
    int L = N;
    for (int t = 0; t < fd.nst; t++) {
        fd.R[t] = rad[t];
        fd.M[t] = L / rad[t];
        L = fd.M[t];
        fd.lgR[t] = (fd.R[t] & (fd.R[t] - 1)) == 0 ? __builtin_ctz(fd.R[t]) : -1;
        fd.lgM[t] = (fd.M[t] & (fd.M[t] - 1)) == 0 ? __builtin_ctz(fd.M[t]) : -1;
    }
    for (int f = 0; f < N; f++) {
        int rem = f, pos = 0;
        for (int t = 0; t < fd.nst; t++) {
            pos += (rem % fd.R[t]) * fd.M[t];
            rem /= fd.R[t];
        }
        rev[f] = pos;
    }
    return fd;
}

int build_class_tables(fftso3_plan_s *pl, Class *c)
{
    const int B = c->B, N = c->N;
    const bool s2 = pl->domain == FFTSO3_DOMAIN_S2;
    const long double PI = 3.141592653589793238462643383279502884L;
    int st;
    // twiddles W_N^t
    std::vector<float2> tw32(N);
    std::vector<double2> tw64(N);
    for (int t = 0; t < N; t++) {
        long double a = -2.0L * PI * (long double)t / (long double)N;
        tw64[t] = make_double2((double)cosl(a), (double)sinl(a));
        tw32[t] = make_float2((float)cosl(a), (float)sinl(a));
    }
    std::vector<int> rev;
    c->fd = make_fft_desc(N, rev);
    if ((st = c->d_tw32.upload(tw32)) || (st = c->d_tw64.upload(tw64)) || (st = c->d_rev.upload(rev))) return st;

    // grid and recursion tables (FP64, evaluated in long double)
    std::vector<double> cb(N), lnc(N), lns(N), w(N);
    for (int k = 0; k < N; k++) {
        long double beta = (long double)(2 * k + 1) * PI / (4.0L * B);  // src/fdcl_FFTSO3.cpp:578-581
        cb[k] = (double)cosl(beta);
        lnc[k] = (double)logl(cosl(beta / 2));
        lns[k] = s2 ? (double)logl(sinl(beta)) : (double)logl(sinl(beta / 2));
        // weights: src/fdcl_FFTSO3.cpp:236-256 (SO(3)), src/fdcl_FFTS2.cpp:102-123 (S^2)
        long double sum = 0;
        for (int i = 0; i < B; i++)
            sum += sinl((long double)((2 * k + 1) * (2 * i + 1)) * PI / (4.0L * B)) / (long double)(2 * i + 1);
        long double pref = s2 ? 2.0L * PI / ((long double)B * B) : 1.0L / (4.0L * B * B * B);
        w[k] = (double)(pref * sinl(beta) * sum);
    }
    std::vector<double> sq((size_t)B * B, 0.0), isq((size_t)B * B, 0.0);
    std::vector<double> lnbin;
    if (!s2) {
        for (int l = 0; l < B; l++)
            for (int m = 0; m <= l; m++) {
                long double v = sqrtl((long double)l * l - (long double)m * m);
                sq[(size_t)l * B + m] = (double)v;
                isq[(size_t)l * B + m] = m < l ? (double)(1.0L / v) : 0.0;
            }
        lnbin.resize((size_t)B * B);
        for (int m0 = 0; m0 < B; m0++)
            for (int n0 = -m0; n0 <= m0; n0++)
                lnbin[(size_t)m0 * m0 + n0 + m0] =
                    (double)(0.5L * (lgammal(2.0L * m0 + 1) - lgammal((long double)(m0 + n0) + 1) -
                                     lgammal((long double)(m0 - n0) + 1)));
    } else {
        for (int l = 0; l < B; l++)
            for (int m = 0; m < l; m++) {
                long double f = sqrtl((4.0L * l * l - 1.0L) / ((long double)l * l - (long double)m * m));
                sq[(size_t)l * B + m] = (double)f;
                isq[(size_t)l * B + m] = (double)(1.0L / f);
            }
        lnbin.resize(B);
        long double lp = 0;  // ln prod_{i<=m} (2i-1)/(2i)
        for (int m = 0; m < B; m++) {
            if (m > 0) lp += logl((2.0L * m - 1.0L) / (2.0L * m));
            lnbin[m] = (double)(0.5L * (logl((2.0L * m + 1.0L) / (4.0L * PI)) + lp));
        }
    }
    std::vector<double> lcoef((size_t)3 * B, 0.0);
    for (int l = 2; l < B; l++) {
        lcoef[3 * l] = (double)l * (2.0 * l - 1.0);
        lcoef[3 * l + 1] = (double)((2.0L * l - 1.0L) / (long double)(l - 1));
        lcoef[3 * l + 2] = (double)((long double)l / (long double)(l - 1));
    }
    if ((st = c->d_lcoef.upload(lcoef))) return st;
    if ((st = c->d_cb.upload(cb)) || (st = c->d_lnc.upload(lnc)) || (st = c->d_lns.upload(lns)) ||
        (st = c->d_w.upload(w)) || (st = c->d_lnbin.upload(lnbin)) || (st = c->d_sq.upload(sq)) ||
        (st = c->d_isq.upload(isq)))
        return st;
    return FFTSO3_OK;
}

RecTables rec_tables(Class *c)
{
    RecTables rt;
    rt.cb = (const double *)c->d_cb.p;
    rt.lnc = (const double *)c->d_lnc.p;
    rt.lns = (const double *)c->d_lns.p;
    rt.w = (const double *)c->d_w.p;
    rt.lnbin = (const double *)c->d_lnbin.p;
    rt.sq = (const double *)c->d_sq.p;
    rt.isq = (const double *)c->d_isq.p;
    rt.lcoef = (const double *)c->d_lcoef.p;
    return rt;
}

// FP64 SO(3) batches (a plan of 5 or more functions per class, so 8 per pass by default) at
// B <= 64 run batch-wide: the whole class per pass, contractions on FP64 MFMA (mfma64.hpp; the D
// and coefficient blocks of larger B exceed the LDS its kernels use).  set_chunk() keeps the
// chunked vector kernels.
static bool mx64_eligible(const fftso3_plan_s *pl, const Class *c)
{
    return pl->prec == FFTSO3_FP64 && pl->domain == FFTSO3_DOMAIN_SO3 && c->B <= 64 && pl->C >= MX64_FB &&
           !pl->force_chunked;
}

int build_class_work(fftso3_plan_s *pl, Class *c)
{
    const int B = c->B, N = c->N;
    // batch-wide FP32 MFMA contraction: N <= 128 (D in registers / LDS), or up to B = 128 with the
    // group's d block resident in LDS (K-chunked kernels)
    {
        const int np = pl->domain == FFTSO3_DOMAIN_S2 ? 2 : 4;
        const size_t rcb = (size_t)B * sizeof(RecCoef);
        const bool fits = B <= MXK_MAXROWS && mxk_lds_bytes(N, B, np, true) + rcb <= (size_t)kMaxLds &&
                          mxk_lds_bytes(N, B, np, false) + rcb <= (size_t)kMaxLds;
        const bool split = pl->prec == FFTSO3_BF16X3 || pl->prec == FFTSO3_BF16X6;
        c->mxk = pl->prec == FFTSO3_FP32 && N > 128 && fits && !pl->force_chunked;
        const int npc = pl->prec == FFTSO3_BF16X6 ? 3 : 2;
        c->mxks = split && N > 128 && B <= MXK_MAXROWS && !pl->force_chunked &&
                  mxks_fwd_lds_bytes(N, npc, np) + rcb <= (size_t)kMaxLds &&
                  mxks_inv_lds_bytes(npc, np, B) + rcb <= (size_t)kMaxLds;
        c->mx = ((pl->prec == FFTSO3_FP32 || split) && N <= 128 && !pl->force_chunked) || c->mxk || c->mxks;
    }
    c->mx64 = mx64_eligible(pl, c);
    c->chunk = c->mx || c->mx64 ? (int)c->fns.size() : pl->C;
    // split precisions at l_max = 255 (complex SO(3), chunks of 8): on-the-fly recursion with
    // MFMA products (mfma_otf.hpp), over half-record workspace rows and PT coefficients
    c->otf = (pl->prec == FFTSO3_BF16X3 || pl->prec == FFTSO3_BF16X6) && pl->domain == FFTSO3_DOMAIN_SO3 &&
             pl->kind == FFTSO3_COMPLEX && N == OTF_N && c->chunk == 8 && !c->mx && !pl->force_chunked;
    // large bandlimits on the chunked kernels, and the batch-wide FP64 contractions: coefficients
    // go through PT[e][fi] so the contraction moves whole rows of the chunk's functions instead of
    // scattered single values (batch-wide FP64 at l_max = 63 x 256: the inverse read 7.4x its
    // coefficient bytes as 16-byte pieces of the packed arrays and the forward wrote 1.8x; with PT
    // and the two layout passes 10,530 -> 11,390 transforms/s).  Real FP64 batches stage the
    // packed complex coefficients of the real mixing in d_tmp, between the mixing and the layout
    // pass.
    c->pt = (!c->mx && pl->kind == FFTSO3_COMPLEX && B > 64 && pl->prec != FFTSO3_BF16X3 &&
             pl->prec != FFTSO3_BF16X6) || c->otf || c->mx64;
    // Above l_max = 127 the chunked contraction would spend its time in the sequential recursion
    // (one lane per k, the rest of the workgroup idle): the plan computes d once, at full
    // occupancy, into a table it keeps ([g][l - m0][k] in the storage type; 11.5 GB at l_max = 255
    // in FP32), if that stays within 64 GB.
    if (!c->mx && B > 128 && pl->prec != FFTSO3_BF16X3 && pl->prec != FFTSO3_BF16X6) {
        const bool s2d = pl->domain == FFTSO3_DOMAIN_S2;
        const int ngrp = s2d ? B : B * B;
        const size_t tsz = pl->esz / 2;
        std::vector<long> goff(ngrp);
        long rows = 0;
        for (int g = 0; g < ngrp; g++) {
            int m0 = s2d ? g : (int)std::sqrt((double)g + 0.5);
            if (!s2d) { while (m0 * m0 > g) m0--; while ((m0 + 1) * (m0 + 1) <= g) m0++; }
            goff[g] = rows * N;
            rows += B - m0;
        }
        const size_t bytes = (size_t)rows * N * tsz;
        if (bytes <= ((size_t)64 << 30)) {
            int st;
            if ((st = c->d_dtab_off.upload(goff)) || (st = c->d_dtab.alloc(bytes))) return st;
            const long total = (long)ngrp * N;
            const int blocks = (int)((total + 255) / 256);
            RecTables rt = rec_tables(c);
            if (tsz == 8) hipLaunchKernelGGL((build_dtab<double>), dim3(blocks), dim3(256), 0, 0, rt, B, N, s2d ? 1 : 0,
                                             (const long *)c->d_dtab_off.p, (double *)c->d_dtab.p);
            else hipLaunchKernelGGL((build_dtab<float>), dim3(blocks), dim3(256), 0, 0, rt, B, N, s2d ? 1 : 0,
                                    (const long *)c->d_dtab_off.p, (float *)c->d_dtab.p);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipDeviceSynchronize());
        }
    }
    const int C = c->mx ? 8 : c->chunk;  // tiling of the chunked kernels (unused by the MX path)
    const bool s2 = pl->domain == FFTSO3_DOMAIN_S2;
    const size_t tsz = pl->esz / 2;  // real element size
    int st;
    // contraction work lists
    const int np = s2 ? 2 : 4;
    const int ncol = 2 * np * C;
    const int ngroups = 32 * C / ncol;
    c->lt_fwd = 4 * ngroups;
    const int ngrp = s2 ? B : B * B;
    std::vector<int2> itf, iti;
    for (int g = 0; g < ngrp; g++) {
        int m0 = s2 ? g : (int)std::sqrt((double)g + 0.5);
        if (!s2) { while (m0 * m0 > g) m0--; while ((m0 + 1) * (m0 + 1) <= g) m0++; }
        for (int l = m0; l < B; l += c->lt_fwd) itf.push_back(make_int2(g, l));
    }
    // inverse: k tile sized so that D[B][KT] + Fv[B][ncol] fits the LDS (every k tile re-reads
    // the group's coefficients, so larger tiles at large B matter more than occupancy)
    int kt = N;
    const size_t inv_lds_cap = B > 64 ? (size_t)kMaxLds - 8 * 1024 : (size_t)64 * 1024;
    while (kt > 4 && (size_t)B * (kt + ncol) * tsz > inv_lds_cap) kt /= 2;
    c->kt_inv = kt;
    for (int g = 0; g < ngrp; g++)
        for (int k = 0; k < N; k += kt) iti.push_back(make_int2(g, k));
    c->nitems_fwd = (int)itf.size();
    c->nitems_inv = (int)iti.size();
    if ((st = c->d_items_fwd.upload(itf)) || (st = c->d_items_inv.upload(iti))) return st;
    if (c->mx64) {
        // FP64 MFMA contractions: 64-degree forward items (a whole group at B <= 64), inverse k
        // tiles of up to 64
        std::vector<int2> f64, i64;
        c->kt64 = std::min(64, N);
        for (int g = 0; g < ngrp; g++) {
            int m0 = (int)std::sqrt((double)g + 0.5);
            while (m0 * m0 > g) m0--;
            while ((m0 + 1) * (m0 + 1) <= g) m0++;
            for (int l = m0; l < B; l += 16 * MX64_RT) f64.push_back(make_int2(g, l));
            for (int k = 0; k < N; k += c->kt64) i64.push_back(make_int2(g, k));
        }
        c->nitems_fwd64 = (int)f64.size();
        c->nitems_inv64 = (int)i64.size();
        if ((st = c->d_items_fwd64.upload(f64)) || (st = c->d_items_inv64.upload(i64))) return st;
    }
    std::vector<int2> mf, mi;
    for (int g = 0; g < ngrp; g++) {
        int m0 = s2 ? g : (int)std::sqrt((double)g + 0.5);
        if (!s2) { while (m0 * m0 > g) m0--; while ((m0 + 1) * (m0 + 1) <= g) m0++; }
        for (int l = m0; l < B; l += 64) mf.push_back(make_int2(g, l));
        for (int k = 0; k < N; k += 128) mi.push_back(make_int2(g, k));
    }
    c->nitems_fwd_mx = (int)mf.size();
    c->nitems_inv_mx = (int)mi.size();
    if ((st = c->d_items_fwd_mx.upload(mf)) || (st = c->d_items_inv_mx.upload(mi))) return st;
    // batch-wide MFMA path: one workgroup per group; group g has m0 = floor(sqrt(g)) (SO(3)) or g
    // (S^2), so ascending g is descending row count (largest first)
    std::vector<int> mx(ngrp);
    for (int g = 0; g < ngrp; g++) mx[g] = g;
    c->nitems_mx = ngrp;
    if ((st = c->d_items_mx.upload(mx))) return st;
    if (c->mx && (pl->prec == FFTSO3_BF16X3 || pl->prec == FFTSO3_BF16X6)) {
        // split batch-wide kernels (mxks): forward items 2 g + h (degrees m0 + 64 h ..), inverse
        // items g * ceil(N / 128) + kr (beta values 128 kr ..), groups largest first
        std::vector<int> f, i;
        const int nkr = (N + MXKS_KR - 1) / MXKS_KR;
        for (int g = 0; g < ngrp; g++) {
            int m0 = s2 ? g : (int)std::sqrt((double)g + 0.5);
            if (!s2) { while (m0 * m0 > g) m0--; while ((m0 + 1) * (m0 + 1) <= g) m0++; }
            for (int h = 0; h < 2 && m0 + MXKS_LR * h < B; h++) f.push_back(2 * g + h);
            for (int kr = 0; kr < nkr; kr++) i.push_back(g * nkr + kr);
        }
        c->nitems_mxs_f = (int)f.size();
        c->nitems_mxs_i = (int)i.size();
        if ((st = c->d_items_mxs_f.upload(f)) || (st = c->d_items_mxs_i.upload(i))) return st;
    }
    const int CC = c->chunk;
    std::vector<int> ident(CC);
    for (int i = 0; i < CC; i++) ident[i] = i;
    std::vector<long> wsoff(CC);
    const long nc = s2 ? (long)B * B : so3_ncoef(B);
    for (int i = 0; i < CC; i++) wsoff[i] = (long)i * nc;
    if ((st = c->d_ident.upload(ident)) || (st = c->d_ws_off.upload(wsoff))) return st;
    // FFT mode
    const size_t tables = (size_t)N * (pl->esz + sizeof(int));
    size_t slice = (size_t)N * N * pl->esz + tables;
    c->fused = !s2 && slice <= (size_t)kMaxLds && (size_t)N * N <= (size_t)FFT_EPT * kFftThreads;
    const size_t cap = std::min(((size_t)kMaxLds - tables) / pl->esz, (size_t)FFT_EPT * kFftThreads);
    int rt = 1;
    while (rt * 2 <= N && N % (rt * 2) == 0 && (size_t)(rt * 2) * N <= cap) rt *= 2;
    c->row_tile = rt;
    c->col_tile = rt;
    // SO(3) row / column passes: the largest power-of-two tile that fits, whether or not it
    // divides N (the last block is partial): N = 200 gets 64 rows instead of 8
    int rt3 = 1;
    while (rt3 * 2 <= N && (size_t)(rt3 * 2) * N <= cap) rt3 *= 2;
    c->so3_tile = rt3;
    // slices larger than LDS whose contraction reads half records: FP64 l_max = 63 (batch-wide MFMA)
    // and FP32 l_max = 255 (d table)
    c->prfft = !s2 && ((pl->prec == FFTSO3_FP64 && c->mx64 && N == 128) ||
                       (pl->prec == FFTSO3_FP32 && !c->mx && c->d_dtab.p && N == 512) || c->otf);
    return FFTSO3_OK;
}
int build_unified(fftso3_plan_s *pl);
bool unified_eligible(const fftso3_plan_s *pl);

int finalize_plan(fftso3_plan_s *pl)
{
    pl->drop_graphs();  // captured graphs hold the old buffers
    pl->valid = false;
    int st;
    for (auto *c : pl->classes) delete c;
    pl->classes.clear();
    std::map<int, Class *> byB;
    for (int i = 0; i < pl->n_fn; i++) {
        int B = pl->Bs[i];
        if (!byB.count(B)) {
            Class *c = new (std::nothrow) Class();
            if (!c) return fail(FFTSO3_ENOMEM, "host allocation failed");
            c->B = B;
            c->N = 2 * B;
            byB[B] = c;
            pl->classes.push_back(c);
        }
        byB[B]->fns.push_back(i);
    }
    size_t ws = 0, scr = 0, fc = 0, tmpb = 0;
    const bool s2 = pl->domain == FFTSO3_DOMAIN_S2;
    for (auto *c : pl->classes) {
        if ((st = build_class_tables(pl, c)) || (st = build_class_work(pl, c))) return st;
        if ((st = c->d_fn.upload(c->fns))) return st;
        // N = 16R slice FFTs (fft_r16.hpp) for the float batch-wide SO(3) classes of B = 8R, R <= 6
        const int R = c->B / 8;
        c->r16 = (pl->prec != FFTSO3_FP64 && !s2 && c->mx && !c->mxk && c->B % 8 == 0 &&
                  (R == 1 || R == 2 || R == 3 || R == 4 || R == 6)) ? R : 0;
    }
    pl->unified = unified_eligible(pl);
    for (auto *c : pl->classes) {
        const size_t B = c->B, N = c->N, C = c->chunk;
        const int klo = pl->klo;
        // (batch-wide FP64: whole blocks of 8 functions)
        const size_t Cw = c->mx64 ? (C + MX64_FB - 1) / MX64_FB * MX64_FB : C;
        const size_t w = (size_t)ws_rows(c->mx, klo, s2 ? (int)B : (int)(B * B), (int)N) * Cw * (s2 ? 2 : 4) * pl->esz;
        const size_t cf = C * (size_t)(s2 ? B * B : so3_ncoef((int)B)) * pl->esz;
        if (pl->unified) {
            // every class in a region of its own, 256-byte aligned
            c->ws_off = ws;
            c->fc_off = fc;
            c->tmp_off = tmpb;
            ws += (w + 255) & ~(size_t)255;
            fc += (cf + 255) & ~(size_t)255;
            if (pl->kind == FFTSO3_REAL) tmpb += (cf + 255) & ~(size_t)255;
            continue;
        }
        c->ws_off = c->fc_off = c->tmp_off = 0;
        ws = std::max(ws, w);
        // two-pass FFT scratch: a pass's functions (batch-wide FP64: 8 at a time, fft_pass)
        if (!s2 && !c->fused) scr = std::max(scr, (c->mx64 ? std::min<size_t>(C, MX64_FB) : C) * N * N * N * pl->esz);
        if (pl->kind == FFTSO3_REAL) fc = std::max(fc, cf);
        if (c->mx || c->pt) {
            fc = std::max(fc, cf);  // PT
            if (pl->kind == FFTSO3_REAL) tmpb = std::max(tmpb, cf);
        }
    }
    // ws and PT are read up to 512 bytes (ws with 4 groups interleaved: 2 KB) past a row start
    // (clamped columns of a partial block)
    ws += 4096;
    if (fc) fc += 512;
    if ((st = pl->d_dummy.resize(256)) || (st = pl->d_ws.resize(ws)) || (st = pl->d_scratch.resize(scr)) ||
        (st = pl->d_fc.resize(fc)) || (st = pl->d_tmp.resize(tmpb)))
        return st;
    if (pl->unified && (st = build_unified(pl))) return st;
    pl->valid = true;
    return FFTSO3_OK;
}

// ---------------------------------------------------------------------------------------------
// launching
// ---------------------------------------------------------------------------------------------
struct StageTimer {
    fftso3_plan_s *pl;
    hipStream_t s;
    int stage;
    int idx = -1;
    StageTimer(fftso3_plan_s *p, hipStream_t st, int sg) : pl(p), s(st), stage(sg)
    {
        if (!pl->timing) return;
        idx = (int)pl->ev_used.size();
        if (idx >= (int)pl->ev_pool.size()) {
            hipEvent_t a, b;
            if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) { idx = -1; return; }
            pl->ev_pool.push_back({a, b});
        }
        pl->ev_used.push_back({stage, idx});
        (void)hipEventRecord(pl->ev_pool[idx].first, s);
    }
    ~StageTimer()
    {
        if (idx >= 0) (void)hipEventRecord(pl->ev_pool[idx].second, s);
    }
};

int g_num_cus = 0;

// persistent grid: every CU filled to its occupancy for this kernel and LDS size, multiple of 8
template <typename K>
int persistent_grid(K kernel, size_t lds, int nunits, int threads = kFftThreads)
{
    if (g_num_cus == 0) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&g_num_cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) g_num_cus = 256;
    }
    int per = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, lds) != hipSuccess || per < 1) per = 1;
    int grid = g_num_cus * per;
    grid = std::min(grid, xcd_grid(nunits));
    return std::max(XCD_COUNT, (grid / XCD_COUNT) * XCD_COUNT);
}

template <typename T> bool use_fft128(const Class *c) { return sizeof(T) == 4 && c->N == 128; }
// register-based ring FFTs of the S^2 grid at l_max = 127
template <typename T> bool use_s2fft256(const Class *c) { return sizeof(T) == 4 && c->N == 256; }

// Group interleave of the workspace (ws_rec) for the N = 128 and N = 16R FFT kernels with the
// batch-wide MFMA contractions and unblocked rows; 0 everywhere else.  Two groups side by side
// (gil = 1) let each FFT slice fill 64-byte halves of lines by itself (the paired-function kernels
// fill whole lines) while the contraction's row segments still cover half of every line they
// touch; four (gil = 2) make the FFT's lines whole but quarter the contraction's.  At N = 16R,
// interleave 1 took the uneven real batch's FFT stages from 1.05 + 1.04 to 0.94 + 0.97 ms.
static int ws_gil(const fftso3_plan_s *pl, const Class *c)
{
    const bool ok = pl->prec != FFTSO3_FP64 && c->mx && !c->mxk && (c->N == 128 || c->r16) && pl->klo == 0 &&
                    pl->domain == FFTSO3_DOMAIN_SO3;
    return ok ? pl->gil : 0;
}

template <typename T>
ChunkArgs chunk_args(fftso3_plan_s *pl, Class *c, int i0, int nfc)
{
    ChunkArgs ca;
    ca.B = c->B;
    ca.N = c->N;
    ca.C = c->chunk;
    ca.nfc = nfc;
    ca.fn = reinterpret_cast<const int *>(c->d_fn.p) + i0;
    ca.grid_off = reinterpret_cast<const long *>(pl->d_grid_off.p);
    ca.coef_off = reinterpret_cast<const long *>(pl->d_coef_off.p);
    ca.kmajor = c->mx ? 1 : 0;
    ca.klo = pl->klo;
    ca.ngrp = pl->domain == FFTSO3_DOMAIN_S2 ? c->B : c->B * c->B;
    ca.gil = ws_gil(pl, c);
    ca.hsplit = c->prfft ? 1 : 0;
    return ca;
}

template <typename T>
FftArgs<T> fft_args(fftso3_plan_s *pl, Class *c, const ChunkArgs &ca, const void *src, void *dst)
{
    typedef typename cplx_t<T>::type T2;
    FftArgs<T> a;
    a.ca = ca;
    a.fd = c->fd;
    a.tw = reinterpret_cast<const T2 *>(sizeof(T) == 4 ? c->d_tw32.p : c->d_tw64.p);
    a.rev = reinterpret_cast<const int *>(c->d_rev.p);
    a.src = src;
    a.dst = dst;
    a.scratch = reinterpret_cast<T2 *>(pl->d_scratch.p);
    a.tile = 0;
    return a;
}


template <typename T, int C>
int launch_contract(fftso3_plan_s *pl, Class *c, const ChunkArgs &ca, bool fwd, const void *src, void *dst,
                    hipStream_t s)
{
    ConArgs<T> a;
    a.ca = ca;
    a.rt = rec_tables(c);
    a.s2 = pl->domain == FFTSO3_DOMAIN_S2;
    a.pt = c->pt ? 1 : 0;
    a.dtab = c->d_dtab.p;
    a.dtab_off = (const long *)c->d_dtab_off.p;
    a.src = src;
    a.dst = dst;
    const int np = a.s2 ? 2 : 4, ncol = 2 * np * C;
    if (fwd && sizeof(T) == 4 && c->d_dtab.p && ncol <= TABF_COLS && (size_t)c->N * TABF_COLS * 4 <= (size_t)kMaxLds) {
        // FP32 with the d table: MFMA over 32-degree tiles, one workgroup per group
        ConArgs<float> af;
        std::memcpy(&af, &a, sizeof(af));
        const int ngrp = a.s2 ? c->B : c->B * c->B;
        if (c->B <= 32 * 4 * TABF2_TPW && c->B % 8 == 0) {
            // the spectrum block staged half a beta range at a time: two workgroups per CU
            const size_t lds = (size_t)(c->N / 2) * TABF_COLS * sizeof(float);
            if (a.s2) hipLaunchKernelGGL((contract_fwd_tab2<2>), dim3(deal_grid(ngrp, 1)), dim3(256), lds, s, af);
            else hipLaunchKernelGGL((contract_fwd_tab2<4>), dim3(deal_grid(ngrp, 1)), dim3(256), lds, s, af);
        } else {
            const size_t lds = (size_t)c->N * TABF_COLS * sizeof(float);
            if (a.s2) hipLaunchKernelGGL((contract_fwd_tab<2>), dim3(deal_grid(ngrp, 1)), dim3(256), lds, s, af);
            else hipLaunchKernelGGL((contract_fwd_tab<4>), dim3(deal_grid(ngrp, 1)), dim3(256), lds, s, af);
        }
    } else if (fwd) {
        a.items = (const int2 *)c->d_items_fwd.p;
        a.nitems = c->nitems_fwd;
        a.tile = c->lt_fwd;
        // D[N][LT] + spectrum chunk [min(N,128)][ncol] (contract_fwd's KCMAX)
        size_t lds = ((size_t)c->N * a.tile + (size_t)std::min(c->N, 128) * ncol) * sizeof(T);
        hipLaunchKernelGGL((contract_fwd<T, C>), dim3(xcd_grid(a.nitems)), dim3(32 * C), lds, s, a);
    } else if (sizeof(T) == 4 && c->d_dtab.p) {
        // FP32 with the d table: MFMA over 128-beta tiles
        ConArgs<float> af;
        std::memcpy(&af, &a, sizeof(af));
        af.tile = 128;
        // Items run largest group first.  The coefficient block takes M4 x TAB_CPITCH floats of
        // LDS, so the groups with at most 128 degrees (75% of them at l_max = 255) go in a launch
        // of their own whose smaller LDS lets 4 workgroups share a CU instead of 2.
        const int nk = (c->N + 127) / 128, MT = 128;
        const int gsplit = c->B > MT ? (a.s2 ? c->B - MT : (c->B - MT) * (c->B - MT)) : (a.s2 ? c->B : c->B * c->B);
        const int isplit = gsplit * nk;
        for (int part = 0; part < 2; part++) {
            const int i0 = part ? isplit : 0, i1 = part ? c->nitems_inv_mx : isplit;
            if (i1 <= i0) continue;
            af.items = (const int2 *)c->d_items_inv_mx.p + i0;
            af.nitems = i1 - i0;
            const int mmax = part ? MT : c->B;
            const size_t lds = (size_t)((mmax + 3) & ~3) * TAB_CPITCH * sizeof(float);
            const int tgrid = deal_grid(af.nitems, nk);
            if (a.s2) hipLaunchKernelGGL((contract_inv_tab<2>), dim3(tgrid), dim3(256), lds, s, af);
            else hipLaunchKernelGGL((contract_inv_tab<4>), dim3(tgrid), dim3(256), lds, s, af);
        }
    } else {
        a.items = (const int2 *)c->d_items_inv.p;
        a.nitems = c->nitems_inv;
        a.tile = c->kt_inv;
        size_t lds = ((size_t)c->B * a.tile + (size_t)c->B * ncol) * sizeof(T);
        hipLaunchKernelGGL((contract_inv<T, C>), dim3(deal_grid(a.nitems, (c->N + a.tile - 1) / a.tile)), dim3(32 * C),
                           lds, s, a);
    }
    HIP_TRY(hipGetLastError());
    return FFTSO3_OK;
}

template <int NP, int NPC>
int launch_contract_bf16(Class *c, const ConArgs<float> &a0, bool fwd, hipStream_t s)
{
    ConArgs<float> a = a0;
    if (fwd) {
        a.items = (const int2 *)c->d_items_fwd_mx.p;
        a.nitems = c->nitems_fwd_mx;
        a.tile = 64;
        size_t lds = (size_t)NPC * (64 + 16 * NP) * 72 * 2;
        hipLaunchKernelGGL((contract_fwd_bf16<NP, NPC>), dim3(xcd_grid(a.nitems)), dim3(256), lds, s, a);
    } else {
        a.items = (const int2 *)c->d_items_inv_mx.p;
        a.nitems = c->nitems_inv_mx;
        a.tile = 128;
        size_t lds = (size_t)NPC * (128 + 16 * NP) * 72 * 2;
        hipLaunchKernelGGL((contract_inv_bf16<NP, NPC>), dim3(xcd_grid(a.nitems)), dim3(256), lds, s, a);
    }
    HIP_TRY(hipGetLastError());
    return FFTSO3_OK;
}

// Threads of a whole-slice FFT workgroup: N^2 / 16 rounded up to whole waves (16 slice elements
// per thread, the direct DFT's register bound FFT_EPT), so the small slices of low bandlimits do
// not tie up 1024-thread workgroups with most threads idle and more of them share a CU (measured:
// 16 elements per thread beat 8, which beat a fixed 1024).
static int fused_threads(int N)
{
    const int t = (N * N / FFT_EPT + 63) / 64 * 64;
    return std::max(64, std::min(kFftThreads, t));
}

// The batch-wide contraction of one class (forward: spectrum workspace src -> PT dst; inverse: PT
// src -> workspace dst), with the kernels and work split its own launches use; run_mx and the
// unified schedule's per-class contraction stages call it.
template <int NP>
int launch_mx_contract(fftso3_plan_s *pl, Class *c, const ChunkArgs &ca, bool fwd, const void *src, void *dst,
                       hipStream_t s)
{
    MxArgs a;
    a.mrows = 0;
    a.ca = ca;
    a.rt = rec_tables(c);
    a.items = (const int *)c->d_items_mx.p;
    a.nitems = c->nitems_mx;
    a.gpw = 1;
    a.g0 = 0;
    {
        // split each group's column blocks over several workgroups when there are few groups
        const int nfb = (ca.nfc * NP * 2 + 127) / 128;
        a.nsplit = std::max(1, std::min(nfb, pl->split_mul * ((2048 + a.nitems - 1) / a.nitems)));
    }
    const int mgrid = deal_grid(a.nitems * a.nsplit, a.nsplit << ca.gil);
    const size_t rcb = (size_t)c->B * sizeof(RecCoef);   // per-group recursion coefficients
    const size_t lds = MX_LDS_BYTES + rcb;
    a.dummy = reinterpret_cast<float *>(pl->d_dummy.p);
    a.src = src;
    a.dst = dst;
    if (fwd) {
        if (pl->prec == FFTSO3_BF16X6 || pl->prec == FFTSO3_BF16X3) {
            MxArgs af = a;
            af.items = (const int *)c->d_items_mxs_f.p;
            af.nitems = c->nitems_mxs_f;
            const int nfb8 = (ca.nfc * NP * 2 + 255) / 256;
            af.nsplit = std::max(1, std::min(nfb8, pl->split_mul * ((1024 + af.nitems - 1) / af.nitems)));
            const int fgrid = deal_grid(af.nitems * af.nsplit, af.nsplit << ca.gil);
            if (pl->prec == FFTSO3_BF16X6)
                hipLaunchKernelGGL((contract_fwd_mxks<NP, 3>), dim3(fgrid), dim3(512),
                                   mxks_fwd_lds_bytes(c->N, 3, NP) + rcb, s, af);
            else
                hipLaunchKernelGGL((contract_fwd_mxks<NP, 2>), dim3(fgrid), dim3(512),
                                   mxks_fwd_lds_bytes(c->N, 2, NP) + rcb, s, af);
        } else if (c->mxk) {
            // 8 waves per workgroup over blocks of 256 columns: half as many units.  The groups of
            // at most 64 degrees (items come largest first) run in a second launch whose LDS holds
            // 64 d rows, so two workgroups share a CU.
            const int nbig = c->B <= 64 ? a.nitems : NP == 2 ? c->B - 64 : (c->B - 64) * (c->B - 64);
            const int nfb8 = (ca.nfc * NP * 2 + 255) / 256;
            for (int part = 0; part < 2; part++) {
                MxArgs a8 = a;
                if (part == 1) {
                    a8.items += nbig;
                    a8.mrows = 64;
                }
                a8.nitems = part == 0 ? nbig : a.nitems - nbig;
                if (a8.nitems <= 0) continue;
                a8.nsplit = std::max(1, std::min(nfb8, pl->split_mul * ((1024 + a8.nitems - 1) / a8.nitems)));
                hipLaunchKernelGGL((contract_fwd_mxk<NP>), dim3(deal_grid(a8.nitems * a8.nsplit, a8.nsplit << ca.gil)),
                                   dim3(512), mxk_lds_bytes(c->N, part ? 64 : c->B, NP, true) + rcb, s, a8);
            }
        }
        else if (c->N == 128 && NP == 4 && ca.gil > 0 && ca.gil <= 2) {
            // the 2^gil groups that share workspace lines in one workgroup
            MxArgs ag = a;
            ag.gpw = 1 << ca.gil;
            ag.nitems = c->nitems_mx / ag.gpw;
            const int fbw = 16 / ag.gpw, nfbg = (ca.nfc + fbw - 1) / fbw;
            ag.nsplit = std::max(1, std::min(nfbg, pl->split_mul * ((2048 + ag.nitems - 1) / ag.nitems)));
            const int ggrid = deal_grid(ag.nitems * ag.nsplit, ag.nsplit);
            const size_t glds = mx16g_lds_bytes(c->B);
            if (ag.gpw == 2) {
                // every wave keeps two 16-row tiles of D in registers: the groups with more than 32
                // degrees run on eight waves (two per group and pair couple, one per tile half, at
                // two waves per SIMD), the groups with at most 32 (m0 >= B - 32; 3/4 of them at
                // l_max = 63) on four, two workgroups per CU
                const int gsplit = c->B > 32 ? (c->B - 32) * (c->B - 32) : 0;
                MxArgs a4 = ag, a2 = ag;
                a4.nitems = gsplit / 2;
                a2.g0 = gsplit;
                a2.nitems = ag.nitems - gsplit / 2;
                if (a4.nitems > 0)
                    hipLaunchKernelGGL(contract_fwd_mx16g8, dim3(deal_grid(a4.nitems * a4.nsplit, a4.nsplit)),
                                       dim3(512), glds, s, a4);
                if (a2.nitems > 0)
                    hipLaunchKernelGGL((contract_fwd_mx16g<2>), dim3(deal_grid(a2.nitems * a2.nsplit, a2.nsplit)),
                                       dim3(256), glds, s, a2);
            }
            else hipLaunchKernelGGL((contract_fwd_mx32g<4>), dim3(ggrid), dim3(256),
                                    (size_t)128 * MX_LPITCH * sizeof(float) + 2 * rcb, s, ag);
        } else if (c->N == 128) hipLaunchKernelGGL((contract_fwd_mx32<NP, true>), dim3(mgrid), dim3(256), lds, s, a);
        else hipLaunchKernelGGL((contract_fwd_mx32<NP, false>), dim3(mgrid), dim3(256), lds, s, a);
    } else {
        if (pl->prec == FFTSO3_BF16X6 || pl->prec == FFTSO3_BF16X3) {
            MxArgs ai = a;
            ai.items = (const int *)c->d_items_mxs_i.p;
            ai.nitems = c->nitems_mxs_i;
            const int nfb8 = (ca.nfc * NP * 2 + 255) / 256;
            ai.nsplit = std::max(1, std::min(nfb8, pl->split_mul * ((1024 + ai.nitems - 1) / ai.nitems)));
            const int igrid = deal_grid(ai.nitems * ai.nsplit, ai.nsplit << ca.gil);
            if (pl->prec == FFTSO3_BF16X6)
                hipLaunchKernelGGL((contract_inv_mxks<NP, 3>), dim3(igrid), dim3(512), mxks_inv_lds_bytes(3, NP, c->B) + rcb, s, ai);
            else
                hipLaunchKernelGGL((contract_inv_mxks<NP, 2>), dim3(igrid), dim3(512), mxks_inv_lds_bytes(2, NP, c->B) + rcb, s, ai);
        } else if (c->mxk) {
            // 8 waves per workgroup over blocks of 256 columns: half as many units.  The groups of
            // at most 64 degrees run in a second launch (contract_inv_mxk64: a 64-row d block, two
            // workgroups per CU).
            const int nbig = c->B <= 64 ? a.nitems : NP == 2 ? c->B - 64 : (c->B - 64) * (c->B - 64);
            const int nfb8 = (ca.nfc * NP * 2 + 255) / 256;
            for (int part = 0; part < 2; part++) {
                MxArgs a8 = a;
                if (part == 1) {
                    a8.items += nbig;
                    a8.mrows = 64;
                }
                a8.nitems = part == 0 ? nbig : a.nitems - nbig;
                if (a8.nitems <= 0) continue;
                a8.nsplit = std::max(1, std::min(nfb8, pl->split_mul * ((1024 + a8.nitems - 1) / a8.nitems)));
                const int grid = deal_grid(a8.nitems * a8.nsplit, a8.nsplit << ca.gil);
                if (part == 0)
                    hipLaunchKernelGGL((contract_inv_mxk<NP, 8>), dim3(grid), dim3(512), mxk_lds_bytes(c->N, c->B, NP, false) + rcb, s, a8);
                else
                    hipLaunchKernelGGL((contract_inv_mxk64<NP>), dim3(grid), dim3(512), mxk_lds_bytes(c->N, 64, NP, false) + rcb, s, a8);
            }
        }
        else if (c->N == 128 && NP == 4 && ca.gil == 1) {
            // the two groups that share workspace lines in one workgroup
            MxArgs ag = a;
            ag.gpw = 2;
            ag.nitems = c->nitems_mx / 2;
            const int nfbg = (ca.nfc + 15) / 16;   // blocks of 16 functions (contract_inv_mx32g16w8)
            ag.nsplit = std::max(1, std::min(nfbg, pl->split_mul * ((2048 + ag.nitems - 1) / ag.nitems)));
            const size_t glds = mx32g_inv_lds_bytes(c->B);
            hipLaunchKernelGGL(contract_inv_mx32g16w8, dim3(deal_grid(ag.nitems * ag.nsplit, ag.nsplit)), dim3(512),
                               glds, s, ag);
        } else hipLaunchKernelGGL((contract_inv_mx32<NP>), dim3(mgrid), dim3(256), lds, s, a);
    }
    HIP_TRY(hipGetLastError());
    return FFTSO3_OK;
}

template <int NP>
int run_mx(fftso3_plan_s *pl, Class *c, const ChunkArgs &ca, bool fwd, const void *in, void *out, hipStream_t s)
{
    const bool real = pl->kind == FFTSO3_REAL;
    const bool s2 = NP == 2;
    const int ncp = s2 ? c->B * c->B : (int)so3_ncoef(c->B);
    const int tgrid = coef_transpose_grid<float2>(ncp, ca.nfc);
    // the real path stages packed complex coefficients in d_tmp, addressed per class-local function
    ChunkArgs ct = ca;
    if (real) {
        ct.fn = reinterpret_cast<const int *>(c->d_ident.p);
        ct.coef_off = reinterpret_cast<const long *>(c->d_ws_off.p);
    }
    const long nc = ncp;
    const int tmix_blocks = (int)std::min<long>(4096, (nc * ca.nfc + 255) / 256);
    int st;
    if (fwd) {
        {
            StageTimer t(pl, s, 1);
            if ((st = launch_mx_contract<NP>(pl, c, ca, true, pl->d_ws.p, pl->d_fc.p, s))) return st;
        }
        StageTimer t(pl, s, 2);
        hipLaunchKernelGGL((coef_transpose<0, float2>), dim3(tgrid), dim3(256), 0, s, ct, ncp, pl->d_fc.p,
                           real ? pl->d_tmp.p : out);
        HIP_TRY(hipGetLastError());
        if (real) {
            if (s2) hipLaunchKernelGGL((s2_tmix<float, 0>), dim3(tmix_blocks), dim3(256), 0, s, ca, pl->d_tmp.p, out, nc);
            else hipLaunchKernelGGL((so3_tmix<float, 0>), dim3(tmix_blocks), dim3(256), 0, s, ca, pl->d_tmp.p, out, nc);
            HIP_TRY(hipGetLastError());
        }
    } else {
        {
            StageTimer t(pl, s, 3);
            if (real && !s2) {
                // real SO(3): the real mixing and the layout pass in one (so3_coef_real_inv)
                hipLaunchKernelGGL(so3_coef_real_inv, dim3(xcd_grid(so3_coef_real_units(c->B, ca.nfc))), dim3(256),
                                   so3_coef_real_lds(c->B), s, ca, reinterpret_cast<const float *>(in),
                                   reinterpret_cast<float2 *>(pl->d_fc.p));
            } else {
                if (real) {
                    hipLaunchKernelGGL((s2_tmix<float, 1>), dim3(tmix_blocks), dim3(256), 0, s, ca, in, pl->d_tmp.p, nc);
                    HIP_TRY(hipGetLastError());
                }
                hipLaunchKernelGGL((coef_transpose<1, float2>), dim3(tgrid), dim3(256), 0, s, ct, ncp, real ? pl->d_tmp.p : in,
                                   pl->d_fc.p);
            }
            HIP_TRY(hipGetLastError());
        }
        StageTimer t(pl, s, 4);
        if ((st = launch_mx_contract<NP>(pl, c, ca, false, pl->d_fc.p, pl->d_ws.p, s))) return st;
    }
    return FFTSO3_OK;
}

// batch-wide FP64 MFMA contractions (mfma64.hpp)
int launch_contract_mx64(fftso3_plan_s *pl, Class *c, const ChunkArgs &ca, bool fwd, const void *src, void *dst,
                         hipStream_t s)
{
    ConArgs<double> a;
    std::memset(&a, 0, sizeof(a));
    a.ca = ca;
    a.rt = rec_tables(c);
    a.s2 = 0;
    a.pt = c->pt ? 1 : 0;
    a.dtab = c->d_dtab.p;
    a.dtab_off = (const long *)c->d_dtab_off.p;
    a.src = src;
    a.dst = dst;
    a.items = (const int2 *)(fwd ? c->d_items_fwd64.p : c->d_items_inv64.p);
    a.nitems = fwd ? c->nitems_fwd64 : c->nitems_inv64;
    a.tile = fwd ? 16 * MX64_RT : c->kt64;
    // several workgroups share an item's blocks of 8 functions when items are few
    const int nfb = (ca.nfc + MX64_FB - 1) / MX64_FB;
    a.nsplit = std::max(1, std::min(nfb, pl->split_mul * ((2048 + a.nitems - 1) / a.nitems)));
    if (fwd)
        hipLaunchKernelGGL(contract_fwd_mx64, dim3(xcd_grid(a.nitems * a.nsplit)), dim3(MX64_THREADS),
                           mx64_fwd_lds_bytes(c->N), s, a);
    else
        hipLaunchKernelGGL(contract_inv_mx64, dim3(deal_grid(a.nitems * a.nsplit, (c->N + c->kt64 - 1) / c->kt64)),
                           dim3(MX64_THREADS), mx64_inv_lds_bytes(c->B, c->kt64), s, a);
    HIP_TRY(hipGetLastError());
    return FFTSO3_OK;
}

template <typename T>
int contract(fftso3_plan_s *pl, Class *c, const ChunkArgs &ca, bool fwd, const void *src, void *dst, hipStream_t s)
{
    if constexpr (sizeof(T) == 8) {
        if (c->mx64) return launch_contract_mx64(pl, c, ca, fwd, src, dst, s);
    }
    if constexpr (sizeof(T) == 4) {
        if (pl->prec == FFTSO3_BF16X3 || pl->prec == FFTSO3_BF16X6) {
            ConArgs<float> a;
            a.ca = ca;
            a.rt = rec_tables(c);
            a.s2 = pl->domain == FFTSO3_DOMAIN_S2;
            a.pt = 0;
            a.dtab = nullptr;
            a.dtab_off = nullptr;
            a.src = src;
            a.dst = dst;
            const bool x6 = pl->prec == FFTSO3_BF16X6;
            if (c->otf) {
                a.pt = 1;
                // one workgroup per group, largest first (a persistent form that claimed groups in a
                // loop was twice as slow: each group's first load waits also for the previous group's
                // 128 KB of workspace stores, which a new workgroup does not inherit)
#define OTF_LAUNCH(K, T, L) hipLaunchKernelGGL((K), dim3(deal_grid(ca.ngrp, 1)), dim3(T), L, s, a)
                if (fwd) {
                    if (x6) OTF_LAUNCH(contract_fwd_otf<3>, OTFF_THREADS, otf_fwd_lds_bytes(3, c->B));
                    else OTF_LAUNCH(contract_fwd_otf<2>, OTFF_THREADS, otf_fwd_lds_bytes(2, c->B));
                } else {
                    if (x6) OTF_LAUNCH(contract_inv_otf<3>, OTFI_THREADS, otf_inv_lds_bytes(3, c->B));
                    else OTF_LAUNCH(contract_inv_otf<2>, OTFI_THREADS, otf_inv_lds_bytes(2, c->B));
                }
#undef OTF_LAUNCH
                HIP_TRY(hipGetLastError());
                return FFTSO3_OK;
            }
            if (a.s2) return x6 ? launch_contract_bf16<2, 3>(c, a, fwd, s) : launch_contract_bf16<2, 2>(c, a, fwd, s);
            return x6 ? launch_contract_bf16<4, 3>(c, a, fwd, s) : launch_contract_bf16<4, 2>(c, a, fwd, s);
        }
    }
    switch (c->chunk) {
    case 1: return launch_contract<T, 1>(pl, c, ca, fwd, src, dst, s);
    case 2: return launch_contract<T, 2>(pl, c, ca, fwd, src, dst, s);
    case 4: return launch_contract<T, 4>(pl, c, ca, fwd, src, dst, s);
    default: return launch_contract<T, 8>(pl, c, ca, fwd, src, dst, s);
    }
}

// the N = 16R slice FFTs of one class (fft_r16.hpp)
template <int R>
int launch_r16_R(bool fwd, bool real, const FftArgs<float> &a, hipStream_t s)
{
    constexpr int nt = r16_threads<R>();
    const int grid = xcd_grid(r16_units<R, nt>(a.ca.nfc, real));
    const size_t lds = r16_lds_bytes<R, nt>();
    if (fwd && real) hipLaunchKernelGGL((so3_fft_r16<R, true, true>), dim3(grid), dim3(nt), lds, s, a);
    else if (fwd) hipLaunchKernelGGL((so3_fft_r16<R, true, false>), dim3(grid), dim3(nt), lds, s, a);
    else if (real) hipLaunchKernelGGL((so3_fft_r16<R, false, true>), dim3(grid), dim3(nt), lds, s, a);
    else hipLaunchKernelGGL((so3_fft_r16<R, false, false>), dim3(grid), dim3(nt), lds, s, a);
    HIP_TRY(hipGetLastError());
    return FFTSO3_OK;
}

int launch_r16(int R, bool fwd, bool real, const FftArgs<float> &a, hipStream_t s)
{
    switch (R) {
    case 1: return launch_r16_R<1>(fwd, real, a, s);
    case 2: return launch_r16_R<2>(fwd, real, a, s);
    case 3: return launch_r16_R<3>(fwd, real, a, s);
    case 4: return launch_r16_R<4>(fwd, real, a, s);
    case 6: return launch_r16_R<6>(fwd, real, a, s);
    default: return fail(FFTSO3_EINVAL, "no 16R slice FFT for R = %d", R);
    }
}

// The two passes of fft_lines.hpp for one bandlimit N (FP64 128, FP32 512)
template <typename T, int N>
int launch_prfft(const FftArgs<T> &a, bool fwd, bool real, hipStream_t s)
{
    const int nfc = a.ca.nfc, nfb = (nfc + PrCfg<T, N>::F - 1) / PrCfg<T, N>::F;
    const int grows = xcd_grid(nfb * N * (N / 2));
    constexpr size_t lrows = pr_rows_lds<T, N>();
#define PR_COLS(REAL, FWD)                                                                                        \
    hipLaunchKernelGGL((so3_pr_cols<T, N, REAL, FWD>), dim3(xcd_grid(nfc * N * (N / PrColsCfg<T, N, REAL>::CB))), \
                       dim3(PrColsCfg<T, N, REAL>::NTHR), (pr_cols_lds<T, N, REAL>()), s, a)
    if (fwd) {
        if (real) PR_COLS(true, true);
        else PR_COLS(false, true);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL((so3_pr_rows<T, N, true>), dim3(grows), dim3(PrCfg<T, N>::NTHR), lrows, s, a);
    } else {
        hipLaunchKernelGGL((so3_pr_rows<T, N, false>), dim3(grows), dim3(PrCfg<T, N>::NTHR), lrows, s, a);
        HIP_TRY(hipGetLastError());
        if (real) PR_COLS(true, false);
        else PR_COLS(false, false);
    }
#undef PR_COLS
    HIP_TRY(hipGetLastError());
    return FFTSO3_OK;
}

// The alpha/gamma FFTs of one pass: grid -> workspace (forward) or workspace -> grid (inverse)
template <typename T>
int fft_stage(fftso3_plan_s *pl, Class *c, const ChunkArgs &ca, bool fwd, const void *src, void *dst, hipStream_t s)
{
    typedef typename cplx_t<T>::type T2;
    const bool real = pl->kind == FFTSO3_REAL;
    const bool s2 = pl->domain == FFTSO3_DOMAIN_S2;
    const int N = c->N, nfc = ca.nfc;
    const size_t tables = (size_t)N * (sizeof(T2) + sizeof(int));
    int st = FFTSO3_OK;
    (void)st;
    if (fwd) {
        FftArgs<T> a = fft_args<T>(pl, c, ca, src, dst);
        if (s2 && use_s2fft256<T>(c)) {
            if constexpr (sizeof(T) == 4) {
                const int g256 = xcd_grid(((nfc + S2F_FNS - 1) / S2F_FNS) * (256 / S2F_KS));
                if (real) hipLaunchKernelGGL((s2_fft256_fwd<true>), dim3(g256), dim3(256), 0, s, a);
                else hipLaunchKernelGGL((s2_fft256_fwd<false>), dim3(g256), dim3(256), 0, s, a);
            }
        } else if (s2) {
            a.tile = c->row_tile;
            int nunit = nfc * (N / a.tile);
            size_t lds = (size_t)a.tile * N * sizeof(T2) + tables;
            if (real) hipLaunchKernelGGL((s2_fft_fwd<T, true>), dim3(xcd_grid(nunit)), dim3(kFftThreads), lds, s, a);
            else hipLaunchKernelGGL((s2_fft_fwd<T, false>), dim3(xcd_grid(nunit)), dim3(kFftThreads), lds, s, a);
        } else if (use_fft128<T>(c)) {
            if constexpr (sizeof(T) == 4) {
                size_t lds = (size_t)F128 * P128 * sizeof(float2) + F128 * sizeof(float2);
                if (ca.gil == 1) {
                    const int g2 = xcd_grid((nfc + 1) / 2 * N);
                    if (real) hipLaunchKernelGGL((so3_fft128_fwd2<true>), dim3(g2), dim3(1024), lds, s, a);
                    else hipLaunchKernelGGL((so3_fft128_fwd2<false>), dim3(g2), dim3(1024), lds, s, a);
                } else if (real) hipLaunchKernelGGL((so3_fft128_fwd<true>), dim3(xcd_grid(nfc * N)), dim3(1024), lds, s, a);
                else hipLaunchKernelGGL((so3_fft128_fwd<false>), dim3(xcd_grid(nfc * N)), dim3(1024), lds, s, a);
            }
        } else if (c->r16) {
            if constexpr (sizeof(T) == 4) {
                if ((st = launch_r16(c->r16, true, real, a, s))) return st;
            }
        } else if (c->fused) {
            size_t lds = (size_t)N * N * sizeof(T2) + tables;
            if (real) hipLaunchKernelGGL((so3_fft_fwd_slice<T, true>), dim3(xcd_grid(nfc * N)), dim3(fused_threads(N)), lds, s, a);
            else hipLaunchKernelGGL((so3_fft_fwd_slice<T, false>), dim3(xcd_grid(nfc * N)), dim3(fused_threads(N)), lds, s, a);
        } else if (c->prfft) {
            if constexpr (sizeof(T) == 8) return launch_prfft<T, 128>(a, true, real, s);
            else return launch_prfft<T, 512>(a, true, real, s);
        } else {
            a.tile = c->so3_tile;
            int nunit = nfc * N * ((N + a.tile - 1) / a.tile);
            size_t lds = (size_t)a.tile * N * sizeof(T2) + tables;
            if (real) hipLaunchKernelGGL((so3_fft_rows<T, true, 0>), dim3(xcd_grid(nunit)), dim3(kFftThreads), lds, s, a);
            else hipLaunchKernelGGL((so3_fft_rows<T, false, 0>), dim3(xcd_grid(nunit)), dim3(kFftThreads), lds, s, a);
            HIP_TRY(hipGetLastError());
            a.tile = c->so3_tile;
            nunit = nfc * N * ((N + a.tile - 1) / a.tile);
            lds = (size_t)a.tile * N * sizeof(T2) + tables;
            hipLaunchKernelGGL((so3_fft_cols<T, 0>), dim3(xcd_grid(nunit)), dim3(kFftThreads), lds, s, a);
        }
        HIP_TRY(hipGetLastError());
        return FFTSO3_OK;
    }
    FftArgs<T> a = fft_args<T>(pl, c, ca, src, dst);
    if (s2 && use_s2fft256<T>(c)) {
        if constexpr (sizeof(T) == 4) {
            const int g256 = xcd_grid(((nfc + S2F_FNS - 1) / S2F_FNS) * (256 / S2F_KS));
            if (real) hipLaunchKernelGGL((s2_fft256_inv<true>), dim3(g256), dim3(256), 0, s, a);
            else hipLaunchKernelGGL((s2_fft256_inv<false>), dim3(g256), dim3(256), 0, s, a);
        }
    } else if (s2) {
        a.tile = c->row_tile;
        int nunit = nfc * (N / a.tile);
        size_t lds = (size_t)a.tile * N * sizeof(T2) + tables;
        if (real) hipLaunchKernelGGL((s2_fft_inv<T, true>), dim3(xcd_grid(nunit)), dim3(kFftThreads), lds, s, a);
        else hipLaunchKernelGGL((s2_fft_inv<T, false>), dim3(xcd_grid(nunit)), dim3(kFftThreads), lds, s, a);
    } else if (use_fft128<T>(c)) {
        if constexpr (sizeof(T) == 4) {
            size_t lds = (size_t)F128 * P128 * sizeof(float2) + F128 * sizeof(float2);
            if (ca.gil == 1) {
                const int g2 = xcd_grid((nfc + 1) / 2 * N);
                if (real) hipLaunchKernelGGL((so3_fft128_inv2<true>), dim3(g2), dim3(1024), lds, s, a);
                else hipLaunchKernelGGL((so3_fft128_inv2<false>), dim3(g2), dim3(1024), lds, s, a);
            } else if (real) hipLaunchKernelGGL((so3_fft128_inv<true>), dim3(xcd_grid(nfc * N)), dim3(1024), lds, s, a);
            else hipLaunchKernelGGL((so3_fft128_inv<false>), dim3(xcd_grid(nfc * N)), dim3(1024), lds, s, a);
        }
    } else if (c->r16) {
        if constexpr (sizeof(T) == 4) {
            if ((st = launch_r16(c->r16, false, real, a, s))) return st;
        }
    } else if (c->fused) {
        size_t lds = (size_t)N * N * sizeof(T2) + tables;
        if (real) hipLaunchKernelGGL((so3_fft_inv_slice<T, true>), dim3(xcd_grid(nfc * N)), dim3(fused_threads(N)), lds, s, a);
        else hipLaunchKernelGGL((so3_fft_inv_slice<T, false>), dim3(xcd_grid(nfc * N)), dim3(fused_threads(N)), lds, s, a);
    } else if (c->prfft) {
        if constexpr (sizeof(T) == 8) return launch_prfft<T, 128>(a, false, real, s);
        else return launch_prfft<T, 512>(a, false, real, s);
    } else {
        a.tile = c->so3_tile;
        int nunit = nfc * N * ((N + a.tile - 1) / a.tile);
        size_t lds = (size_t)a.tile * N * sizeof(T2) + tables;
        hipLaunchKernelGGL((so3_fft_cols<T, 1>), dim3(xcd_grid(nunit)), dim3(kFftThreads), lds, s, a);
        HIP_TRY(hipGetLastError());
        a.tile = c->so3_tile;
        nunit = nfc * N * ((N + a.tile - 1) / a.tile);
        lds = (size_t)a.tile * N * sizeof(T2) + tables;
        if (real) hipLaunchKernelGGL((so3_fft_rows<T, true, 1>), dim3(xcd_grid(nunit)), dim3(kFftThreads), lds, s, a);
        else hipLaunchKernelGGL((so3_fft_rows<T, false, 1>), dim3(xcd_grid(nunit)), dim3(kFftThreads), lds, s, a);
    }
    HIP_TRY(hipGetLastError());
    return FFTSO3_OK;
}

// The FFTs of a pass.  A batch-wide FP64 class (mfma64.hpp) keeps its workspace block-major (rows
// of 8 functions, one region per block of 8) and runs its FFTs one block at a time, each into its
// own region exactly as a chunk of 8 would; the two-pass FFTs' scratch grid between the row and
// column passes (8 N^3 complex doubles, 268 MB at l_max = 63) is then also reused in the Infinity
// Cache instead of making a round trip to HBM.
template <typename T>
int fft_pass(fftso3_plan_s *pl, Class *c, const ChunkArgs &ca, bool fwd, const void *src, void *dst, hipStream_t s)
{
    typedef typename cplx_t<T>::type T2;
    if (!c->mx64) return fft_stage<T>(pl, c, ca, fwd, src, dst, s);
    for (int f0 = 0; f0 < ca.nfc; f0 += MX64_FB) {
        ChunkArgs cs = ca;
        cs.C = MX64_FB;
        cs.nfc = std::min(MX64_FB, ca.nfc - f0);
        cs.fn = ca.fn + f0;
        const size_t wo = (size_t)(f0 / MX64_FB) * c->B * c->B * c->N * MX64_FB * 4 * sizeof(T2);   // block's region
        int st = fwd ? fft_stage<T>(pl, c, cs, true, src, static_cast<char *>(dst) + wo, s)
                     : fft_stage<T>(pl, c, cs, false, static_cast<const char *>(src) + wo, dst, s);
        if (st) return st;
    }
    return FFTSO3_OK;
}

template <typename T>
int run_chunk(fftso3_plan_s *pl, Class *c, int i0, int nfc, bool fwd, const void *in, void *out, hipStream_t s)
{
    typedef typename cplx_t<T>::type T2;
    const bool real = pl->kind == FFTSO3_REAL;
    const bool s2 = pl->domain == FFTSO3_DOMAIN_S2;
    const int N = c->N;
    ChunkArgs ca = chunk_args<T>(pl, c, i0, nfc);
    // contraction-side addressing: the real path stages complex coefficients in d_fc
    ChunkArgs cc = ca;
    if (real) {
        cc.fn = reinterpret_cast<const int *>(c->d_ident.p);
        cc.coef_off = reinterpret_cast<const long *>(c->d_ws_off.p);
    }
    const long nc = s2 ? (long)c->B * c->B : so3_ncoef(c->B);
    int st;
    const int tmix_blocks = (int)std::min<long>(4096, (nc * nfc + 255) / 256);
    const size_t tables = (size_t)N * (sizeof(T2) + sizeof(int));
    if (fwd) {
        {
            StageTimer t(pl, s, 0);
            if ((st = fft_pass<T>(pl, c, ca, true, in, pl->d_ws.p, s))) return st;
            HIP_TRY(hipGetLastError());
        }
        if constexpr (sizeof(T) == 4) {
            if (c->mx) return s2 ? run_mx<2>(pl, c, ca, true, in, out, s) : run_mx<4>(pl, c, ca, true, in, out, s);
        }
        {
            StageTimer t(pl, s, 1);
            if ((st = contract<T>(pl, c, cc, true, pl->d_ws.p, (real || c->pt) ? pl->d_fc.p : out, s))) return st;
        }
        // PT -> packed complex: the output, or (real) the staging in d_tmp that the mixing reads
        void *packed = real && c->pt ? pl->d_tmp.p : real ? pl->d_fc.p : out;
        if (c->pt) {
            StageTimer t(pl, s, 2);
            const int tgrid = coef_transpose_grid<T2>(nc, nfc);
            hipLaunchKernelGGL((coef_transpose<0, T2>), dim3(tgrid), dim3(256), 0, s, cc, (int)nc, pl->d_fc.p, packed);
            HIP_TRY(hipGetLastError());
        }
        if (real) {
            StageTimer t(pl, s, 2);
            if (s2) hipLaunchKernelGGL((s2_tmix<T, 0>), dim3(tmix_blocks), dim3(256), 0, s, ca, packed, out, nc);
            else hipLaunchKernelGGL((so3_tmix<T, 0>), dim3(tmix_blocks), dim3(256), 0, s, ca, packed, out, nc);
            HIP_TRY(hipGetLastError());
        }
    } else {
        const void *csrc = in;
        bool contracted = false;
        if constexpr (sizeof(T) == 4) {
            if (c->mx) {
                if ((st = s2 ? run_mx<2>(pl, c, ca, false, in, out, s) : run_mx<4>(pl, c, ca, false, in, out, s))) return st;
                contracted = true;
            }
        }
        if (real && !contracted) {
            // the mixing's packed complex output: the contraction's input, or (PT) the layout
            // pass's, staged in d_tmp
            void *mixed = c->pt ? pl->d_tmp.p : pl->d_fc.p;
            StageTimer t(pl, s, 3);
            if (s2) hipLaunchKernelGGL((s2_tmix<T, 1>), dim3(tmix_blocks), dim3(256), 0, s, ca, in, mixed, nc);
            else hipLaunchKernelGGL((so3_tmix<T, 1>), dim3(tmix_blocks), dim3(256), 0, s, ca, in, mixed, nc);
            HIP_TRY(hipGetLastError());
            csrc = mixed;
        }
        if (!contracted && c->pt) {
            StageTimer t(pl, s, 3);
            const int tgrid = coef_transpose_grid<T2>(nc, nfc);
            hipLaunchKernelGGL((coef_transpose<1, T2>), dim3(tgrid), dim3(256), 0, s, cc, (int)nc, csrc, pl->d_fc.p);
            HIP_TRY(hipGetLastError());
            csrc = pl->d_fc.p;
        }
        if (!contracted) {
            StageTimer t(pl, s, 4);
            if ((st = contract<T>(pl, c, cc, false, csrc, pl->d_ws.p, s))) return st;
        }
        {
            StageTimer t(pl, s, 5);
            if ((st = fft_pass<T>(pl, c, ca, false, pl->d_ws.p, out, s))) return st;
            HIP_TRY(hipGetLastError());
        }
    }
    return FFTSO3_OK;
}

// ---------------------------------------------------------------------------------------------
// unified schedule (multi.hpp): every class of an uneven FP32 SO(3) plan in one launch per stage
// ---------------------------------------------------------------------------------------------
bool unified_eligible(const fftso3_plan_s *pl)
{
    if (!pl->multi_on || pl->prec == FFTSO3_FP64 || pl->domain != FFTSO3_DOMAIN_SO3 || pl->force_chunked ||
        pl->klo != 0 || pl->classes.size() < 2 || pl->classes.size() > MULTI_MAX_FFT)
        return false;
    for (const Class *c : pl->classes) {
        if (!c->mx || c->mxk || c->mxks) return false;
        if (c->N == 128) {
            if (ws_gil(pl, c) != 1) return false;
        } else if (!c->r16) {
            return false;
        }
    }
    return true;
}

// contraction arguments of a class as its own launch would use them; kind as in multi.hpp
// (forward: 0 one group per workgroup, 2 / 3 the grouped N = 128 kernels for groups with more /
// at most 32 degrees; inverse: 0 one group, 1 grouped)
MxArgs make_mx(fftso3_plan_s *pl, Class *c, const ChunkArgs &ca, bool fwd, int kind)
{
    MxArgs a;
    a.mrows = 0;
    a.ca = ca;
    a.rt = rec_tables(c);
    a.items = (const int *)c->d_items_mx.p;
    a.nitems = c->nitems_mx;
    a.gpw = 1;
    a.g0 = 0;
    a.dummy = reinterpret_cast<float *>(pl->d_dummy.p);
    a.src = a.dst = nullptr;
    const int nfb = (ca.nfc * 4 * 2 + 127) / 128;
    a.nsplit = std::max(1, std::min(nfb, pl->split_mul * ((2048 + a.nitems - 1) / a.nitems)));
    if (kind == 0) return a;
    if (kind == 4) {
        // split kernels (as run_mx launches them): items of (group, degree / beta range), blocks of
        // 256 columns
        a.items = (const int *)(fwd ? c->d_items_mxs_f.p : c->d_items_mxs_i.p);
        a.nitems = fwd ? c->nitems_mxs_f : c->nitems_mxs_i;
        const int nfb8 = (ca.nfc * 4 * 2 + 255) / 256;
        a.nsplit = std::max(1, std::min(nfb8, pl->split_mul * ((1024 + a.nitems - 1) / a.nitems)));
        return a;
    }
    a.gpw = 2;
    a.nitems = c->nitems_mx / 2;
    const int nfbg = (ca.nfc + 7) / 8;
    a.nsplit = std::max(1, std::min(nfbg, pl->split_mul * ((2048 + a.nitems - 1) / a.nitems)));
    if (!fwd) return a;
    const int gsplit = c->B > 32 ? (c->B - 32) * (c->B - 32) : 0;
    if (kind == 2) {
        a.nitems = gsplit / 2;
    } else {
        a.g0 = gsplit;
        a.nitems = c->nitems_mx / 2 - gsplit / 2;
    }
    return a;
}

// per-XCD queues, interleaved so that position p belongs to XCD p % 8
struct XcdQueue {
    std::vector<int2> x[XCD_COUNT];
    void add_units(int entry, int n)   // the units of an xcd_unit launch
    {
        const int q = (n + XCD_COUNT - 1) / XCD_COUNT;
        for (int u = 0; u < n; u++) x[u / q].push_back(make_int2(entry, u));
    }
    void add_dealt(int entry, int n, int run)   // the units of an xcd_deal launch
    {
        const int grid = deal_grid(n, run);
        for (int b = 0; b < grid; b++) {
            const int xx = b % XCD_COUNT, i = b / XCD_COUNT;
            const int u = ((i / run) * XCD_COUNT + xx) * run + i % run;
            if (u < n) x[xx].push_back(make_int2(entry, u));
        }
    }
    void lengths(int *len) const
    {
        for (int i = 0; i < XCD_COUNT; i++) len[i] = (int)x[i].size();
    }
    std::vector<int2> flat() const
    {
        size_t L = 0;
        for (const auto &v : x) L = std::max(L, v.size());
        std::vector<int2> out(L * XCD_COUNT, make_int2(-1, 0));
        for (size_t p = 0; p < out.size(); p++)
            if (p / XCD_COUNT < x[p % XCD_COUNT].size()) out[p] = x[p % XCD_COUNT][p / XCD_COUNT];
        return out;
    }
};

int fft_units(const fftso3_plan_s *pl, const Class *c)
{
    const int nf = (int)c->fns.size();
    const bool real = pl->kind == FFTSO3_REAL;
    if (c->N == 128) return so3_fft128_fwd2_units(nf);
    switch (c->r16) {
    case 1: return r16_units<1>(nf, real);
    case 2: return r16_units<2>(nf, real);
    case 3: return r16_units<3>(nf, real);
    case 4: return r16_units<4>(nf, real);
    default: return r16_units<6>(nf, real);
    }
}

size_t fft_lds(const Class *c)
{
    if (c->N == 128) return (size_t)F128 * P128 * sizeof(float2) + F128 * sizeof(float2);
    switch (c->r16) {
    case 1: return r16_lds_bytes<1>();
    case 2: return r16_lds_bytes<2>();
    case 3: return r16_lds_bytes<3>();
    case 4: return r16_lds_bytes<4>();
    default: return r16_lds_bytes<6>();
    }
}

int build_unified(fftso3_plan_s *pl)
{
    pl->uorder = pl->classes;
    std::stable_sort(pl->uorder.begin(), pl->uorder.end(), [](const Class *a, const Class *b) { return a->N > b->N; });
    pl->ent_cf.clear();
    pl->ent_ci.clear();
    XcdQueue qf, qcf, qci, qtr;
    for (int i = 0; i < (int)pl->uorder.size(); i++) {
        Class *c = pl->uorder[i];
        const int nfc = (int)c->fns.size();
        const ChunkArgs ca = chunk_args<float>(pl, c, 0, nfc);
        qf.add_units(i, fft_units(pl, c));
        qtr.add_units(i, coef_transpose_units<float2>(so3_ncoef(c->B), nfc));
        // split-BF16 plans run the split kernels (kind 4) at every bandlimit
        const bool split = pl->prec == FFTSO3_BF16X3 || pl->prec == FFTSO3_BF16X6;
        const bool grouped = c->N == 128 && pl->prec == FFTSO3_FP32;
        const std::vector<int> kf = split ? std::vector<int>{4} : grouped ? std::vector<int>{2, 3} : std::vector<int>{0};
        for (int kind : kf) {
            const MxArgs a = make_mx(pl, c, ca, true, kind);
            if (a.nitems <= 0) continue;
            qcf.add_dealt((int)pl->ent_cf.size(), a.nitems * a.nsplit, a.gpw > 1 ? a.nsplit : a.nsplit << ca.gil);
            pl->ent_cf.push_back({c, kind});
        }
        const int ki = split ? 4 : grouped ? 1 : 0;
        const MxArgs ai = make_mx(pl, c, ca, false, ki);
        qci.add_dealt((int)pl->ent_ci.size(), ai.nitems * ai.nsplit, ai.gpw > 1 ? ai.nsplit : ai.nsplit << ca.gil);
        pl->ent_ci.push_back({c, ki});
    }
    if (pl->ent_cf.size() > MULTI_MAX_MX || pl->ent_ci.size() > MULTI_MAX_MX)
        return fail(FFTSO3_EUNSUPPORTED, "too many contraction entries for one launch");
    int st;
    const std::vector<int2> f = qf.flat(), cf = qcf.flat(), ci = qci.flat(), tr = qtr.flat();
    if ((st = pl->d_q_fft.upload(f)) || (st = pl->d_q_cf.upload(cf)) || (st = pl->d_q_ci.upload(ci)) ||
        (st = pl->d_q_tr.upload(tr)))
        return st;
    pl->nq_fft = (int)f.size();
    pl->nq_cf = (int)cf.size();
    pl->nq_ci = (int)ci.size();
    pl->nq_tr = (int)tr.size();
    qf.lengths(pl->ql_fft);
    qcf.lengths(pl->ql_cf);
    qci.lengths(pl->ql_ci);
    qtr.lengths(pl->ql_tr);
    // the counters start at zero; each stage kernel's last workgroup resets its own
    if ((st = pl->d_ctr.resize(4 * 16 * sizeof(unsigned)))) return st;
    HIP_TRY(hipMemset(pl->d_ctr.p, 0, 4 * 16 * sizeof(unsigned)));
    return FFTSO3_OK;
}

int run_unified(fftso3_plan_s *pl, bool fwd, const void *in, void *out, hipStream_t s)
{
    const bool real = pl->kind == FFTSO3_REAL;
    const int ncls = (int)pl->uorder.size();
    char *ws = static_cast<char *>(pl->d_ws.p), *fc = static_cast<char *>(pl->d_fc.p);
    char *tmp = static_cast<char *>(pl->d_tmp.p);
    // per-class chunk arguments; the real path stages packed complex coefficients per class-local
    // function in the class's d_tmp region
    std::vector<ChunkArgs> ca(ncls), ct(ncls);
    for (int i = 0; i < ncls; i++) {
        Class *c = pl->uorder[i];
        ca[i] = chunk_args<float>(pl, c, 0, (int)c->fns.size());
        ct[i] = ca[i];
        if (real) {
            ct[i].fn = reinterpret_cast<const int *>(c->d_ident.p);
            ct[i].coef_off = reinterpret_cast<const long *>(c->d_ws_off.p);
        }
    }
    auto fft_stage = [&](bool f, const void *src, void *dst) -> int {
        MultiFft m;
        std::memset(&m, 0, sizeof(m));
        size_t lds = 0;
        for (int i = 0; i < ncls; i++) {
            Class *c = pl->uorder[i];
            m.c[i] = fft_args<float>(pl, c, ca[i], f ? src : ws + c->ws_off, f ? (void *)(ws + c->ws_off) : dst);
            m.kind[i] = c->N == 128 ? 8 : c->r16;
            lds = std::max(lds, fft_lds(c));
        }
        m.q.items = reinterpret_cast<const int2 *>(pl->d_q_fft.p);
        m.q.ctr = reinterpret_cast<unsigned *>(pl->d_ctr.p);
        std::memcpy(m.q.len, pl->ql_fft, sizeof(m.q.len));
        if (f && real) hipLaunchKernelGGL((so3_fft_multi<true, true>), dim3(persistent_grid(so3_fft_multi<true, true>, lds, pl->nq_fft)), dim3(1024), lds, s, m);
        else if (f) hipLaunchKernelGGL((so3_fft_multi<true, false>), dim3(persistent_grid(so3_fft_multi<true, false>, lds, pl->nq_fft)), dim3(1024), lds, s, m);
        else if (real) hipLaunchKernelGGL((so3_fft_multi<false, true>), dim3(persistent_grid(so3_fft_multi<false, true>, lds, pl->nq_fft)), dim3(1024), lds, s, m);
        else hipLaunchKernelGGL((so3_fft_multi<false, false>), dim3(persistent_grid(so3_fft_multi<false, false>, lds, pl->nq_fft)), dim3(1024), lds, s, m);
        HIP_TRY(hipGetLastError());
        return FFTSO3_OK;
    };
    // split-BF16 pieces per operand (0: FP32 MFMA kernels)
    const int npc = pl->prec == FFTSO3_BF16X6 ? 3 : pl->prec == FFTSO3_BF16X3 ? 2 : 0;
    auto contract_stage = [&](bool f) -> int {
        if (pl->uni_pc) {
            // each class's own contraction launch on its region of the workspaces: a persistent
            // kernel's next unit waits on the previous unit's stores (loads and stores share one
            // counter), which separate launches do not
            for (int i = 0; i < ncls; i++) {
                Class *c = pl->uorder[i];
                const void *src = f ? (const void *)(ws + c->ws_off) : (const void *)(fc + c->fc_off);
                void *dst = f ? (void *)(fc + c->fc_off) : (void *)(ws + c->ws_off);
                int st = launch_mx_contract<4>(pl, c, ca[i], f, src, dst, s);
                if (st) return st;
            }
            return FFTSO3_OK;
        }
        MultiMx m;
        std::memset(&m, 0, sizeof(m));
        const auto &ent = f ? pl->ent_cf : pl->ent_ci;
        size_t lds = 0;
        for (size_t e = 0; e < ent.size(); e++) {
            Class *c = ent[e].first;
            const int i = (int)(std::find(pl->uorder.begin(), pl->uorder.end(), c) - pl->uorder.begin());
            MxArgs a = make_mx(pl, c, ca[i], f, ent[e].second);
            a.src = f ? (const void *)(ws + c->ws_off) : (const void *)(fc + c->fc_off);
            a.dst = f ? (void *)(fc + c->fc_off) : (void *)(ws + c->ws_off);
            m.c[e] = a;
            m.kind[e] = ent[e].second;
            const size_t rcb = (size_t)c->B * sizeof(RecCoef);
            const size_t l = npc ? (f ? mxks_fwd_lds_bytes(c->N, npc, 4) : mxks_inv_lds_bytes(npc, 4, c->B)) + rcb
                           : ent[e].second == 0 ? MX_LDS_BYTES + rcb
                           : f ? mx16g_lds_bytes(c->B) : mx32g_inv_lds_bytes(c->B);
            lds = std::max(lds, l);
        }
        m.q.items = reinterpret_cast<const int2 *>((f ? pl->d_q_cf : pl->d_q_ci).p);
        m.q.ctr = reinterpret_cast<unsigned *>(pl->d_ctr.p) + (f ? 16 : 32);
        std::memcpy(m.q.len, f ? pl->ql_cf : pl->ql_ci, sizeof(m.q.len));
        const int nitems = f ? pl->nq_cf : pl->nq_ci;
#define MULTI_LAUNCH(K, T) hipLaunchKernelGGL(K, dim3(persistent_grid(K, lds, nitems, T)), dim3(T), lds, s, m)
        if (npc == 3) {
            if (f) MULTI_LAUNCH(contract_fwd_multi_s<3>, 512);
            else MULTI_LAUNCH(contract_inv_multi_s<3>, 512);
        } else if (npc == 2) {
            if (f) MULTI_LAUNCH(contract_fwd_multi_s<2>, 512);
            else MULTI_LAUNCH(contract_inv_multi_s<2>, 512);
        } else if (f) MULTI_LAUNCH(contract_fwd_multi, 256);
        else MULTI_LAUNCH(contract_inv_multi, 256);
#undef MULTI_LAUNCH
        HIP_TRY(hipGetLastError());
        return FFTSO3_OK;
    };
    auto transpose_stage = [&](int dir, const void *src_user, void *dst_user) -> int {
        MultiTr m;
        std::memset(&m, 0, sizeof(m));
        for (int i = 0; i < ncls; i++) {
            Class *c = pl->uorder[i];
            m.ca[i] = ct[i];
            m.ncoef[i] = (int)so3_ncoef(c->B);
            if (dir == 0) {
                m.src[i] = fc + c->fc_off;
                m.dst[i] = real ? (void *)(tmp + c->tmp_off) : dst_user;
            } else {
                m.src[i] = real ? (const void *)(tmp + c->tmp_off) : src_user;
                m.dst[i] = fc + c->fc_off;
            }
        }
        m.q.items = reinterpret_cast<const int2 *>(pl->d_q_tr.p);
        m.q.ctr = reinterpret_cast<unsigned *>(pl->d_ctr.p) + 48;
        std::memcpy(m.q.len, pl->ql_tr, sizeof(m.q.len));
        if (dir == 0) hipLaunchKernelGGL((coef_transpose_multi<0>), dim3(persistent_grid(coef_transpose_multi<0>, 0, pl->nq_tr, 256)), dim3(256), 0, s, m);
        else hipLaunchKernelGGL((coef_transpose_multi<1>), dim3(persistent_grid(coef_transpose_multi<1>, 0, pl->nq_tr, 256)), dim3(256), 0, s, m);
        HIP_TRY(hipGetLastError());
        return FFTSO3_OK;
    };
    auto tmix_stage = [&](int dir, const void *src_user, void *dst_user) -> int {
        MultiTmix m;
        std::memset(&m, 0, sizeof(m));
        long tot = 0;
        for (int i = 0; i < ncls; i++) {
            Class *c = pl->uorder[i];
            m.ca[i] = ca[i];
            m.src[i] = dir == 0 ? (const void *)(tmp + c->tmp_off) : src_user;
            m.dst[i] = dir == 0 ? dst_user : (void *)(tmp + c->tmp_off);
            m.ws_stride[i] = so3_ncoef(c->B);
            m.start[i] = tot;
            tot += so3_tmix_orbits(c->B) * (long)c->fns.size();
        }
        m.start[ncls] = tot;
        m.ncls = ncls;
        const int blocks = (int)std::min<long>(8192, (tot + 255) / 256);
        if (dir == 0) hipLaunchKernelGGL((so3_tmix_multi<0>), dim3(blocks), dim3(256), 0, s, m);
        else hipLaunchKernelGGL((so3_tmix_multi<1>), dim3(blocks), dim3(256), 0, s, m);
        HIP_TRY(hipGetLastError());
        return FFTSO3_OK;
    };
    int st;
    if (fwd) {
        {
            StageTimer t(pl, s, 0);
            if ((st = fft_stage(true, in, nullptr))) return st;
        }
        {
            StageTimer t(pl, s, 1);
            if ((st = contract_stage(true))) return st;
        }
        StageTimer t(pl, s, 2);
        if ((st = transpose_stage(0, nullptr, out))) return st;
        if (real && (st = tmix_stage(0, nullptr, out))) return st;
    } else {
        {
            StageTimer t(pl, s, 3);
            if (real && (st = tmix_stage(1, in, nullptr))) return st;
            if ((st = transpose_stage(1, in, nullptr))) return st;
        }
        {
            StageTimer t(pl, s, 4);
            if ((st = contract_stage(false))) return st;
        }
        StageTimer t(pl, s, 5);
        if ((st = fft_stage(false, nullptr, out))) return st;
    }
    return FFTSO3_OK;
}

template <typename T>
int run(fftso3_plan_s *pl, bool fwd, const void *in, void *out, hipStream_t s)
{
    if constexpr (sizeof(T) == 4) {
        if (pl->unified) return run_unified(pl, fwd, in, out, s);
    }
    int st;
    for (auto *c : pl->classes) {
        const int n = (int)c->fns.size();
        for (int i0 = 0; i0 < n; i0 += c->chunk)
            if ((st = run_chunk<T>(pl, c, i0, std::min(c->chunk, n - i0), fwd, in, out, s))) return st;
    }
    return FFTSO3_OK;
}

bool kernel_attrs_done = false;
int set_kernel_attrs()
{
    if (kernel_attrs_done) return FFTSO3_OK;
#define ALLOW_LDS(k) HIP_TRY(hipFuncSetAttribute((const void *)(k), hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLds))
    ALLOW_LDS((so3_fft_fwd_slice<float, true>));
    ALLOW_LDS((so3_fft_fwd_slice<float, false>));
    ALLOW_LDS((so3_fft_fwd_slice<double, true>));
    ALLOW_LDS((so3_fft_fwd_slice<double, false>));
    ALLOW_LDS((so3_fft_inv_slice<float, true>));
    ALLOW_LDS((so3_fft_inv_slice<float, false>));
    ALLOW_LDS((so3_fft_inv_slice<double, true>));
    ALLOW_LDS((so3_fft_inv_slice<double, false>));
    ALLOW_LDS((so3_fft_rows<float, true, 0>));
    ALLOW_LDS((so3_fft_rows<float, false, 0>));
    ALLOW_LDS((so3_fft_rows<float, true, 1>));
    ALLOW_LDS((so3_fft_rows<float, false, 1>));
    ALLOW_LDS((so3_fft_rows<double, true, 0>));
    ALLOW_LDS((so3_fft_rows<double, false, 0>));
    ALLOW_LDS((so3_fft_rows<double, true, 1>));
    ALLOW_LDS((so3_fft_rows<double, false, 1>));
    ALLOW_LDS((so3_fft_cols<float, 0>));
    ALLOW_LDS((so3_fft_cols<float, 1>));
    ALLOW_LDS((so3_fft_cols<double, 0>));
    ALLOW_LDS((so3_fft_cols<double, 1>));
    ALLOW_LDS((s2_fft_fwd<float, true>));
    ALLOW_LDS((s2_fft_fwd<float, false>));
    ALLOW_LDS((s2_fft_fwd<double, true>));
    ALLOW_LDS((s2_fft_fwd<double, false>));
    ALLOW_LDS((s2_fft_inv<float, true>));
    ALLOW_LDS((s2_fft_inv<float, false>));
    ALLOW_LDS((s2_fft_inv<double, true>));
    ALLOW_LDS((s2_fft_inv<double, false>));
    ALLOW_LDS((contract_fwd<float, 1>));
    ALLOW_LDS((contract_fwd<float, 2>));
    ALLOW_LDS((contract_fwd<float, 4>));
    ALLOW_LDS((contract_fwd<float, 8>));
    ALLOW_LDS((contract_fwd<double, 1>));
    ALLOW_LDS((contract_fwd<double, 2>));
    ALLOW_LDS((contract_fwd<double, 4>));
    ALLOW_LDS((contract_fwd<double, 8>));
    ALLOW_LDS((contract_inv<float, 1>));
    ALLOW_LDS((contract_inv<float, 2>));
    ALLOW_LDS((contract_inv<float, 4>));
    ALLOW_LDS((contract_inv<float, 8>));
    ALLOW_LDS((contract_inv<double, 1>));
    ALLOW_LDS((contract_inv<double, 2>));
    ALLOW_LDS((contract_inv<double, 4>));
    ALLOW_LDS((contract_inv<double, 8>));
    ALLOW_LDS((contract_fwd_mx32<4, true>));
    ALLOW_LDS((contract_fwd_mx32g<4>));
    ALLOW_LDS((contract_inv_mx32g<2>));
    ALLOW_LDS(contract_inv_mx32g16w8);
    ALLOW_LDS(contract_fwd_mx16g8);
    ALLOW_LDS((contract_fwd_mx16g<2>));
    ALLOW_LDS((contract_fwd_tab<4>));
    ALLOW_LDS((contract_fwd_tab<2>));
    ALLOW_LDS((contract_fwd_tab2<4>));
    ALLOW_LDS((contract_fwd_tab2<2>));
    ALLOW_LDS((contract_fwd_mx32<2, true>));
    ALLOW_LDS((contract_fwd_mx32<4, false>));
    ALLOW_LDS((contract_fwd_mx32<2, false>));
    ALLOW_LDS((contract_inv_mx32<4>));
    ALLOW_LDS((contract_inv_tab<4>));
    ALLOW_LDS((contract_inv_tab<2>));
    ALLOW_LDS((contract_fwd_mxk<4>));
    ALLOW_LDS((contract_fwd_mxk<2>));
    ALLOW_LDS((contract_inv_mxk<4, 8>));
    ALLOW_LDS((contract_inv_mxk<2, 8>));
    ALLOW_LDS((contract_inv_mxk64<4>));
    ALLOW_LDS((contract_inv_mxk64<2>));
    ALLOW_LDS(so3_coef_real_inv);
    ALLOW_LDS((contract_fwd_mxks<4, 2>));
    ALLOW_LDS((contract_fwd_mxks<4, 3>));
    ALLOW_LDS((contract_fwd_mxks<2, 2>));
    ALLOW_LDS((contract_fwd_mxks<2, 3>));
    ALLOW_LDS((contract_inv_mxks<4, 2>));
    ALLOW_LDS((contract_inv_mxks<4, 3>));
    ALLOW_LDS((contract_inv_mxks<2, 2>));
    ALLOW_LDS((contract_inv_mxks<2, 3>));
    ALLOW_LDS((contract_inv_mx32<2>));
    ALLOW_LDS((so3_fft128_fwd<true>));
    ALLOW_LDS((so3_fft128_fwd<false>));
    ALLOW_LDS((so3_fft128_inv<true>));
    ALLOW_LDS((so3_fft128_inv<false>));
    ALLOW_LDS((so3_pr_cols<float, 512, false, true>));
    ALLOW_LDS((so3_pr_cols<float, 512, false, false>));
    ALLOW_LDS((so3_pr_cols<float, 512, true, true>));
    ALLOW_LDS((so3_pr_cols<float, 512, true, false>));
    ALLOW_LDS((so3_pr_rows<float, 512, true>));
    ALLOW_LDS((so3_pr_rows<float, 512, false>));
    ALLOW_LDS(contract_fwd_mx64);
    ALLOW_LDS(contract_inv_mx64);
    ALLOW_LDS((so3_fft128_fwd2<true>));
    ALLOW_LDS((so3_fft128_fwd2<false>));
    ALLOW_LDS((so3_fft128_inv2<true>));
    ALLOW_LDS((so3_fft128_inv2<false>));
    ALLOW_LDS((contract_fwd_bf16<4, 2>));
    ALLOW_LDS((contract_fwd_bf16<4, 3>));
    ALLOW_LDS((contract_fwd_bf16<2, 2>));
    ALLOW_LDS((contract_fwd_bf16<2, 3>));
    ALLOW_LDS((contract_inv_bf16<4, 2>));
    ALLOW_LDS((contract_inv_bf16<4, 3>));
    ALLOW_LDS((contract_inv_bf16<2, 2>));
    ALLOW_LDS((contract_inv_bf16<2, 3>));
    ALLOW_LDS((contract_fwd_otf<2>));
    ALLOW_LDS((contract_fwd_otf<3>));
    ALLOW_LDS((contract_inv_otf<2>));
    ALLOW_LDS((contract_inv_otf<3>));
    // s2_rotate also has static LDS: its dynamic part is at most the 128 KB accumulator
    HIP_TRY(hipFuncSetAttribute((const void *)(s2_rotate<true>), hipFuncAttributeMaxDynamicSharedMemorySize, 128 * 1024));
    HIP_TRY(hipFuncSetAttribute((const void *)(s2_rotate<false>), hipFuncAttributeMaxDynamicSharedMemorySize, 128 * 1024));
#define ALLOW_R16(R)                                 \
    ALLOW_LDS((so3_fft_r16<R, true, true>));         \
    ALLOW_LDS((so3_fft_r16<R, true, false>));        \
    ALLOW_LDS((so3_fft_r16<R, false, true>));        \
    ALLOW_LDS((so3_fft_r16<R, false, false>))
    ALLOW_R16(1);
    ALLOW_R16(2);
    ALLOW_R16(3);
    ALLOW_R16(4);
    ALLOW_R16(6);
#undef ALLOW_R16
    // the queue kernels keep a few bytes of static LDS for their claims
#define ALLOW_LDS_Q(k) HIP_TRY(hipFuncSetAttribute((const void *)(k), hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLds - 1024))
    ALLOW_LDS_Q((so3_fft_multi<true, true>));
    ALLOW_LDS_Q((so3_fft_multi<true, false>));
    ALLOW_LDS_Q((so3_fft_multi<false, true>));
    ALLOW_LDS_Q((so3_fft_multi<false, false>));
    ALLOW_LDS_Q(contract_fwd_multi);
    ALLOW_LDS_Q(contract_inv_multi);
    ALLOW_LDS_Q(contract_fwd_multi_s<3>);
    ALLOW_LDS_Q(contract_inv_multi_s<3>);
    ALLOW_LDS_Q(contract_fwd_multi_s<2>);
    ALLOW_LDS_Q(contract_inv_multi_s<2>);
#undef ALLOW_LDS_Q
#undef ALLOW_LDS
    kernel_attrs_done = true;
    return FFTSO3_OK;
}

int create_common(fftso3_plan_t *out, int domain, const int *lmax, int n_fn, int kind, int precision, unsigned flags)
{
    if (flags & ~(unsigned)(FFTSO3_FLAG_PER_CLASS_LAUNCHES | FFTSO3_FLAG_PER_CLASS_CONTRACTIONS))
        return fail(FFTSO3_EINVAL, "unknown plan flags 0x%x", flags);
    if (!out) return fail(FFTSO3_EINVAL, "plan pointer is NULL");
    *out = nullptr;
    if (domain != FFTSO3_DOMAIN_SO3 && domain != FFTSO3_DOMAIN_S2) return fail(FFTSO3_EINVAL, "unknown domain %d", domain);
    if (kind != FFTSO3_COMPLEX && kind != FFTSO3_REAL) return fail(FFTSO3_EINVAL, "unknown kind %d", kind);
    if (precision < FFTSO3_FP64 || precision > FFTSO3_BF16X6) return fail(FFTSO3_EINVAL, "unknown precision %d", precision);
    if (n_fn <= 0) return fail(FFTSO3_EINVAL, "n_fn must be positive (got %d)", n_fn);
    const int maxB = domain == FFTSO3_DOMAIN_SO3 ? 256 : 1024;
    for (int i = 0; i < n_fn; i++)
        if (lmax[i] < 0 || lmax[i] + 1 > maxB)
            return fail(FFTSO3_EINVAL, "l_max[%d] = %d outside [0, %d]", i, lmax[i], maxB - 1);
    int st;
    if ((st = set_kernel_attrs())) return st;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if ((st = scratch_setup(dev))) return st;
    fftso3_plan_s *pl = new (std::nothrow) fftso3_plan_s();
    if (!pl) return fail(FFTSO3_ENOMEM, "host allocation failed");
    pl->device = dev;
    pl->multi_on = !(flags & FFTSO3_FLAG_PER_CLASS_LAUNCHES);
    pl->uni_pc = (flags & FFTSO3_FLAG_PER_CLASS_CONTRACTIONS) != 0;
    pl->domain = domain;
    pl->kind = kind;
    pl->prec = precision;
    pl->n_fn = n_fn;
    pl->esz = precision == FFTSO3_FP64 ? 16 : 8;
    pl->Bs.resize(n_fn);
    pl->grid_off.resize(n_fn);
    pl->coef_off.resize(n_fn);
    long go = 0, co = 0;
    for (int i = 0; i < n_fn; i++) {
        long B = lmax[i] + 1, N = 2 * B;
        pl->Bs[i] = (int)B;
        pl->grid_off[i] = go;
        pl->coef_off[i] = co;
        go += domain == FFTSO3_DOMAIN_SO3 ? N * N * N : N * N;
        co += domain == FFTSO3_DOMAIN_SO3 ? so3_ncoef((int)B) : B * B;
    }
    pl->grid_total = go;
    pl->coef_total = co;
    int maxfn = 1;
    {
        std::map<int, int> cnt;
        for (int B : pl->Bs) maxfn = std::max(maxfn, ++cnt[B]);
    }
    pl->C = 1;
    while (pl->C < 8 && pl->C < maxfn) pl->C *= 2;
    if (precision == FFTSO3_BF16X3 || precision == FFTSO3_BF16X6) pl->C = 8;  // MFMA tiles span 8 functions
    if ((st = pl->d_grid_off.upload(pl->grid_off)) || (st = pl->d_coef_off.upload(pl->coef_off)) ||
        (st = finalize_plan(pl))) {
        delete pl;
        return st;
    }
    *out = pl;
    return FFTSO3_OK;
}

}  // namespace

// =============================================================================================
// C ABI
// =============================================================================================
extern "C" {

int fftso3_plan_create(fftso3_plan_t *plan, int domain, int l_max, int n_fn, int kind, int precision, unsigned flags)
{
    if (n_fn <= 0) return fail(FFTSO3_EINVAL, "n_fn must be positive (got %d)", n_fn);
    std::vector<int> lm((size_t)n_fn, l_max);
    return create_common(plan, domain, lm.data(), n_fn, kind, precision, flags);
}

int fftso3_plan_create_uneven(fftso3_plan_t *plan, int domain, const int *l_max, int n_fn, int kind, int precision,
                              unsigned flags)
{
    if (!l_max) return fail(FFTSO3_EINVAL, "l_max array is NULL");
    return create_common(plan, domain, l_max, n_fn, kind, precision, flags);
}

int fftso3_plan_destroy(fftso3_plan_t plan)
{
    if (!plan) return fail(FFTSO3_EINVAL, "plan is NULL");
    delete plan;
    return FFTSO3_OK;
}

int fftso3_plan_set_chunk(fftso3_plan_t plan, int chunk)
{
    if (!plan) return fail(FFTSO3_EINVAL, "plan is NULL");
    if (chunk != 1 && chunk != 2 && chunk != 4 && chunk != 8)
        return fail(FFTSO3_EINVAL, "chunk must be 1, 2, 4 or 8 (got %d)", chunk);
    if ((plan->prec == FFTSO3_BF16X3 || plan->prec == FFTSO3_BF16X6) && chunk != 8)
        return fail(FFTSO3_EINVAL, "split-BF16 plans process 8 functions per chunk");
    if (chunk == plan->C && plan->force_chunked && plan->valid) return FFTSO3_OK;
    plan->C = chunk;
    plan->force_chunked = true;
    return finalize_plan(plan);
}

int fftso3_plan_set_schedule(fftso3_plan_t plan, int klo, int gil, int split)
{
    if (!plan) return fail(FFTSO3_EINVAL, "plan is NULL");
    if (klo < 0 || klo > 7) return fail(FFTSO3_EINVAL, "row blocking klo must be in [0, 7] (got %d)", klo);
    if (gil < 0 || gil > 2) return fail(FFTSO3_EINVAL, "group interleave must be 0, 1 or 2 (got %d)", gil);
    if (split != 1 && split != 2 && split != 4) return fail(FFTSO3_EINVAL, "split must be 1, 2 or 4 (got %d)", split);
    plan->klo = klo;
    plan->gil = gil;
    plan->split_mul = split;
    return finalize_plan(plan);
}

int fftso3_plan_describe(fftso3_plan_t plan, char *buf, int len)
{
    if (!plan || !buf || len < 1) return fail(FFTSO3_EINVAL, "NULL argument or empty buffer");
    std::string out;
    const char *prec = plan->prec == FFTSO3_FP64 ? "fp64" : plan->prec == FFTSO3_FP32 ? "fp32"
                     : plan->prec == FFTSO3_BF16X6 ? "bf16x6" : "bf16x3";
    for (const Class *c : plan->classes) {
        const char *path = c->mxk || c->mxks ? "batch-wide mfma (k-chunked)" : c->mx ? "batch-wide mfma" : "chunked";
        char line[256];
        ChunkArgs ca = chunk_args<float>(plan, const_cast<Class *>(c), 0, (int)c->fns.size());
        std::snprintf(line, sizeof(line), "l_max=%d fns=%d %s %s chunk=%d klo=%d gil=%d split=%d%s%s%s%s%s%s\n", c->B - 1,
                      (int)c->fns.size(), prec, path, c->chunk, c->mx ? plan->klo : 0, ca.gil, plan->split_mul,
                      c->r16 ? " fft16R" : "", plan->unified ? (plan->uni_pc ? " unified (contractions per class)" : " unified") : "", c->d_dtab.p ? " d-table" : "",
                      c->mx64 ? " mfma64" : "", c->otf ? " d-on-the-fly" : "", plan->graph ? " graph" : "");
        out += line;
    }
    std::snprintf(buf, (size_t)len, "%s", out.c_str());
    return FFTSO3_OK;
}

int fftso3_plan_get_chunk(fftso3_plan_t plan, int *chunk)
{
    if (!plan || !chunk) return fail(FFTSO3_EINVAL, "NULL argument");
    *chunk = plan->classes.empty() ? plan->C : plan->classes[0]->chunk;
    return FFTSO3_OK;
}

long fftso3_plan_grid_offset(fftso3_plan_t plan, int fn)
{
    if (!plan || fn < 0 || fn >= plan->n_fn) return -1;
    return plan->grid_off[fn];
}
long fftso3_plan_coef_offset(fftso3_plan_t plan, int fn)
{
    if (!plan || fn < 0 || fn >= plan->n_fn) return -1;
    return plan->coef_off[fn];
}
long fftso3_plan_grid_elems(fftso3_plan_t plan) { return plan ? plan->grid_total : -1; }
long fftso3_plan_coef_elems(fftso3_plan_t plan) { return plan ? plan->coef_total : -1; }
long fftso3_plan_workspace_bytes(fftso3_plan_t plan)
{
    if (!plan) return -1;
    return (long)(plan->d_ws.n + plan->d_scratch.n + plan->d_fc.n + plan->d_tmp.n);
}

static int check_valid(fftso3_plan_t plan)
{
    if (!plan) return fail(FFTSO3_EINVAL, "plan is NULL");
    if (!plan->valid)
        return fail(FFTSO3_EINVAL, "plan is not usable: its last reconfiguration failed (%s)", "set_chunk or autotune");
    return FFTSO3_OK;
}

static int run_dispatch(fftso3_plan_t plan, bool fwd, const void *in, void *out, void *stream)
{
    if (int st = check_valid(plan)) return st;
    if (!in || !out) return fail(FFTSO3_EINVAL, "NULL device pointer");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    auto eager = [&]() {
        return plan->prec == FFTSO3_FP64 ? run<double>(plan, fwd, in, out, s) : run<float>(plan, fwd, in, out, s);
    };
    // graph mode needs a capturable stream and no per-stage timing events
    if (!plan->graph || s == nullptr || plan->timing) return eager();
    for (auto &g : plan->graphs)
        if (g.fwd == fwd && g.in == in && g.out == out && g.s == s) {
            HIP_TRY(hipGraphLaunch(g.exec, s));
            return FFTSO3_OK;
        }
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int st = eager();
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &graph);
    if (st != FFTSO3_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    if (ec != hipSuccess) return fail(FFTSO3_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(ec));
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(FFTSO3_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
    if (plan->graphs.size() >= 16) {
        (void)hipGraphExecDestroy(plan->graphs.front().exec);
        plan->graphs.erase(plan->graphs.begin());
    }
    plan->graphs.push_back({fwd, in, out, s, exec});
    HIP_TRY(hipGraphLaunch(exec, s));
    return FFTSO3_OK;
}

extern "C++" {
template <typename T>
static int evaluate(fftso3_plan_s *pl, int fn, const void *coef, const double *ang, long npts, void *out, hipStream_t s)
{
    typedef typename cplx_t<T>::type T2;
    Class *c = nullptr;
    int local = -1;
    for (auto *k : pl->classes)
        for (size_t i = 0; i < k->fns.size(); i++)
            if (k->fns[i] == fn) { c = k; local = (int)i; }
    if (!c) return fail(FFTSO3_EINVAL, "function %d not in the plan", fn);
    const bool s2 = pl->domain == FFTSO3_DOMAIN_S2, real = pl->kind == FFTSO3_REAL;
    EvalArgs<T> a;
    a.B = c->B;
    a.rt = rec_tables(c);
    a.ang = ang;
    a.npts = npts;
    a.out = out;
    a.real = real ? 1 : 0;
    if (real) {
        // complex coefficients F_c = T^H F_r T (S^2: T^T F_r) of this function, in d_fc
        ChunkArgs ca = chunk_args<T>(pl, c, local, 1);
        const long nc = s2 ? (long)c->B * c->B : so3_ncoef(c->B);
        const int blocks = (int)std::min<long>(4096, (nc + 255) / 256);
        if (s2) hipLaunchKernelGGL((s2_tmix<T, 1>), dim3(blocks), dim3(256), 0, s, ca, coef, pl->d_fc.p, nc);
        else hipLaunchKernelGGL((so3_tmix<T, 1>), dim3(blocks), dim3(256), 0, s, ca, coef, pl->d_fc.p, nc);
        HIP_TRY(hipGetLastError());
        a.F = pl->d_fc.p;
    } else {
        a.F = reinterpret_cast<const T2 *>(coef) + pl->coef_off[fn];
    }
    const int grid = (int)std::min<long>(npts, 65536);
    if (s2) hipLaunchKernelGGL((eval_points<T, 2>), dim3(grid), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((eval_points<T, 4>), dim3(grid), dim3(256), 0, s, a);
    HIP_TRY(hipGetLastError());
    return FFTSO3_OK;
}
}  // extern "C++"

int fftso3_evaluate(fftso3_plan_t plan, int fn, const void *d_coef, const double *d_angles, long n_points,
                    void *d_values, void *stream)
{
    if (int st = check_valid(plan)) return st;
    if (fn < 0 || fn >= plan->n_fn) return fail(FFTSO3_EINVAL, "function index %d outside [0, %d)", fn, plan->n_fn);
    if (n_points < 0) return fail(FFTSO3_EINVAL, "n_points must be >= 0 (got %ld)", n_points);
    if (n_points == 0) return FFTSO3_OK;
    if (!d_coef || !d_angles || !d_values) return fail(FFTSO3_EINVAL, "NULL device pointer");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    return plan->prec == FFTSO3_FP64 ? evaluate<double>(plan, fn, d_coef, d_angles, n_points, d_values, s)
                                     : evaluate<float>(plan, fn, d_coef, d_angles, n_points, d_values, s);
}

int fftso3_rotate_s2(fftso3_plan_t plan, const double *d_F, const double *d_angles, long n_rot, double *d_G,
                     void *stream)
{
    if (int st = check_valid(plan)) return st;
    if (plan->domain != FFTSO3_DOMAIN_SO3) return fail(FFTSO3_EINVAL, "rotation needs an SO(3) plan (its Wigner tables)");
    if (plan->classes.size() != 1) return fail(FFTSO3_EINVAL, "rotation needs a plan of one bandlimit");
    if (n_rot < 0) return fail(FFTSO3_EINVAL, "n_rot must be >= 0 (got %ld)", n_rot);
    if (n_rot == 0) return FFTSO3_OK;
    if (!d_F || !d_angles || !d_G) return fail(FFTSO3_EINVAL, "NULL device pointer");
    Class *c = plan->classes[0];
    const bool real = plan->kind == FFTSO3_REAL;
    const int W = real ? 1 : 2, B = c->B;
    RotArgs a;
    a.B = B;
    a.rt = rec_tables(c);
    a.F = d_F;
    a.ang = d_angles;
    a.nrot = n_rot;
    a.G = d_G;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int grid = (int)std::min<long>(n_rot, 65536);
    // degree ranges whose coefficients fit the LDS accumulator (one range up to l_max = 127 real)
    const size_t cap = (size_t)128 * 1024;
    for (int l0 = 0; l0 < B;) {
        int l1 = l0 + 1;
        while (l1 < B && (size_t)((l1 + 1) * (l1 + 1) - l0 * l0) * W * sizeof(double) <= cap) l1++;
        a.l0 = l0;
        a.l1 = l1;
        const size_t lds = (size_t)(l1 * l1 - l0 * l0) * W * sizeof(double);
        if (real) hipLaunchKernelGGL((s2_rotate<true>), dim3(grid), dim3(256), lds, s, a);
        else hipLaunchKernelGGL((s2_rotate<false>), dim3(grid), dim3(256), lds, s, a);
        HIP_TRY(hipGetLastError());
        l0 = l1;
    }
    return FFTSO3_OK;
}

int fftso3_plan_set_graph(fftso3_plan_t plan, int on)
{
    if (!plan) return fail(FFTSO3_EINVAL, "plan is NULL");
    plan->graph = on != 0;
    if (!plan->graph) plan->drop_graphs();
    return FFTSO3_OK;
}

// Times inverse + forward (reps each after two warm-ups) under each candidate schedule and keeps
// the fastest: for batch-wide classes the workspace row blocking and the group split; for the
// chunked kernels the chunk size.  The two arrays serve as scratch.  With graph mode on the pairs
// are graph replays, as the plan's later calls will be.
int fftso3_plan_autotune(fftso3_plan_t plan, void *d_grid, void *d_coef, void *stream, int reps)
{
    if (int st = check_valid(plan)) return st;
    if (!d_grid || !d_coef) return fail(FFTSO3_EINVAL, "NULL device pointer");
    if (reps < 1 || reps > 100) return fail(FFTSO3_EINVAL, "reps must be in [1, 100] (got %d)", reps);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    struct Cand { bool chunked; int C, klo, split, gil; bool multi = true; bool pc = false; };
    std::vector<Cand> cands;
    const bool bf16 = plan->prec == FFTSO3_BF16X3 || plan->prec == FFTSO3_BF16X6;
    if (!bf16) {
        if (plan->prec == FFTSO3_FP32)
            for (int klo : {0, 3})
                for (int sp : {1, 2, 4})
                    for (int gil : {0, 1, 2})
                        if (klo == 0 || gil == 0) cands.push_back({false, plan->C, klo, sp, gil});
        // several bandlimit classes: also one launch per class and stage instead of the unified
        // launches, and unified FFT stages with per-class contraction launches
        if (plan->prec == FFTSO3_FP32 && plan->classes.size() > 1)
            for (int sp : {1, 2}) {
                cands.push_back({false, plan->C, 0, sp, 1, false});
                cands.push_back({false, plan->C, 0, sp, 1, true, true});
            }
        // FP64: the batch-wide MFMA contractions with groups split over 1, 2 or 4 workgroups
        if (plan->prec == FFTSO3_FP64)
            for (int sp : {1, 2, 4}) cands.push_back({false, plan->C, 0, sp, 0});
        for (int C : {1, 2, 4, 8}) cands.push_back({true, C, plan->klo, 1, plan->gil});
    } else {
        // split-BF16: the batch-wide kernels with each group split over 1, 2 or 4 workgroups, or
        // the chunked kernels (8 functions per pass)
        for (int sp : {1, 2, 4})
            for (int gil : {0, 1}) cands.push_back({false, 8, 0, sp, gil});
        // several bandlimit classes: also one launch per class and stage, and unified FFT stages
        // with per-class contraction launches
        if (plan->classes.size() > 1)
            for (int sp : {1, 2}) {
                cands.push_back({false, 8, 0, sp, 1, false});
                cands.push_back({false, 8, 0, sp, 1, true, true});
            }
        cands.push_back({true, 8, 0, 1, 0});
    }
    if (cands.empty()) return FFTSO3_OK;
    // candidates run as the plan will: with graph mode on, the warm-up pair captures each
    // candidate's graphs and the timed pairs replay them, so schedules with many small launches are
    // not charged launch costs that replays do not pay
    const bool timing = plan->timing, graph = plan->graph;
    plan->timing = false;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    float best = 1e30f;
    Cand win = cands[0];
    int st = FFTSO3_OK;
    for (const Cand &c : cands) {
        plan->force_chunked = c.chunked;
        plan->C = c.C;
        plan->klo = c.klo;
        plan->split_mul = c.split;
        plan->gil = c.gil;
        plan->multi_on = c.multi;
        plan->uni_pc = c.pc;
        if ((st = finalize_plan(plan))) break;
        // two warm-up pairs: with graph mode the first captures the graphs and the second is their
        // first replay, which also uploads them (a cost that grows with the launch count)
        for (int w = 0; w < 2 && st == FFTSO3_OK; w++)
            if ((st = run_dispatch(plan, false, d_coef, d_grid, s)) == FFTSO3_OK) st = run_dispatch(plan, true, d_grid, d_coef, s);
        if (st) break;
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps && st == FFTSO3_OK; r++) {
            st = run_dispatch(plan, false, d_coef, d_grid, s);
            if (st == FFTSO3_OK) st = run_dispatch(plan, true, d_grid, d_coef, s);
        }
        if (st) break;
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) { best = ms; win = c; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    plan->force_chunked = win.chunked;
    plan->C = win.C;
    plan->klo = win.klo;
    plan->split_mul = win.split;
    plan->gil = win.gil;
    plan->multi_on = win.multi;
    plan->uni_pc = win.pc;
    const int st2 = finalize_plan(plan);
    plan->timing = timing;
    plan->graph = graph;
    return st ? st : st2;
}

int fftso3_forward(fftso3_plan_t plan, const void *d_grid, void *d_coef, void *stream)
{
    return run_dispatch(plan, true, d_grid, d_coef, stream);
}

int fftso3_inverse(fftso3_plan_t plan, const void *d_coef, void *d_grid, void *stream)
{
    return run_dispatch(plan, false, d_coef, d_grid, stream);
}

int fftso3_inverse_noalloc(fftso3_plan_t plan, const void *d_coef, void *d_grid, void *stream)
{
    if (int st = check_valid(plan)) return st;
    if (!d_coef || !d_grid) return fail(FFTSO3_EINVAL, "NULL device pointer");
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != plan->device)
        return fail(FFTSO3_EINVAL, "plan was created on device %d, device %d is current", plan->device, dev);
    if (dev < 0 || dev >= kMaxDevices || !g_scratch_done[dev]) return fail(FFTSO3_EHIP, "no scratch address for device %d", dev);
    // the plan's workspaces, each rounded up to 256 bytes, laid out one after the other in scratch
    DevBuf *bufs[4] = {&plan->d_ws, &plan->d_scratch, &plan->d_fc, &plan->d_tmp};
    size_t off[4], total = 0;
    for (int i = 0; i < 4; i++) {
        off[i] = total;
        total += (bufs[i]->n + 255) & ~(size_t)255;
    }
    if (total > kNoallocScratchBytes)
        return fail(FFTSO3_EINVAL, "this plan's inverse needs %zu bytes of workspace; the static scratch holds %zu (%d MiB)",
                    total, kNoallocScratchBytes, (int)(kNoallocScratchBytes >> 20));
    // run the plan's own launch sequence eagerly (no graph cache, no timing events) with its
    // workspace pointers redirected into scratch; restored before returning
    unsigned char *base = static_cast<unsigned char *>(g_scratch_ptr[dev]);
    void *saved[4];
    for (int i = 0; i < 4; i++) {
        saved[i] = bufs[i]->p;
        bufs[i]->p = bufs[i]->n ? base + off[i] : nullptr;
    }
    const bool timing = plan->timing;
    plan->timing = false;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const int st = plan->prec == FFTSO3_FP64 ? run<double>(plan, false, d_coef, d_grid, s)
                                             : run<float>(plan, false, d_coef, d_grid, s);
    plan->timing = timing;
    for (int i = 0; i < 4; i++) bufs[i]->p = saved[i];
    return st;
}

const char *fftso3_status_string(int status)
{
    switch (status) {
    case FFTSO3_OK: return "FFTSO3_OK";
    case FFTSO3_EINVAL: return "FFTSO3_EINVAL: invalid argument";
    case FFTSO3_ENOMEM: return "FFTSO3_ENOMEM: allocation failed";
    case FFTSO3_EHIP: return "FFTSO3_EHIP: HIP runtime error";
    case FFTSO3_EUNSUPPORTED: return "FFTSO3_EUNSUPPORTED: not supported";
    default: return "unknown status";
    }
}

const char *fftso3_last_error(void) { return g_last_error.c_str(); }

int fftso3_debug_fail_alloc(int after)
{
    if (after < 0) return fail(FFTSO3_EINVAL, "after must be >= 0 (got %d)", after);
    g_fail_alloc_after = after;
    return FFTSO3_OK;
}

int fftso3_timing_enable(fftso3_plan_t plan, int on)
{
    if (!plan) return fail(FFTSO3_EINVAL, "plan is NULL");
    plan->timing = on != 0;
    plan->ev_used.clear();
    for (int i = 0; i < kNumStages; i++) { plan->acc_ms[i] = 0; plan->acc_n[i] = 0; }
    return FFTSO3_OK;
}

int fftso3_timing_read(fftso3_plan_t plan, double *ms, long *count)
{
    if (!plan) return fail(FFTSO3_EINVAL, "plan is NULL");
    for (auto &u : plan->ev_used) {
        auto &e = plan->ev_pool[u.second];
        HIP_TRY(hipEventSynchronize(e.second));
        float t = 0;
        HIP_TRY(hipEventElapsedTime(&t, e.first, e.second));
        plan->acc_ms[u.first] += t;
        plan->acc_n[u.first] += 1;
    }
    plan->ev_used.clear();
    for (int i = 0; i < kNumStages; i++) {
        if (ms) ms[i] = plan->acc_ms[i];
        if (count) count[i] = plan->acc_n[i];
        plan->acc_ms[i] = 0;
        plan->acc_n[i] = 0;
    }
    return FFTSO3_OK;
}

long fftso3_so3_ncoef(int l_max) { return l_max < 0 ? -1 : so3_ncoef(l_max + 1); }
long fftso3_s2_ncoef(int l_max) { return l_max < 0 ? -1 : (long)(l_max + 1) * (l_max + 1); }

}  // extern "C"
