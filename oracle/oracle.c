/*
 * oracle.c -- FP64 CPU restatement of fdcl-gwu/FFTSO3's SO(3) and S^2 harmonic transforms.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  It is loaded by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py, never by fftso3_amd/.
 *
 * Every function cites the reference file:line whose algorithm it restates.  The reference
 * itself cannot be built here (its Eigen 3 submodule is empty), so this restatement is pinned
 * by closed forms, identities, numpy's FFT and high-precision golden vectors (tests/).
 *
 * Parallelism follows the reference: OpenMP over the beta index k for the forward transform
 * (src/fdcl_FFTSO3.cpp:665-676) with per-thread partial sums.  The reduction is done in thread
 * order after the parallel region (the reference uses an omp critical section, whose order is
 * not deterministic); the sums are the same up to rounding.
 */
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;
typedef long double complex lcplx;

#define PI_L 3.141592653589793238462643383279502884L

/* ------------------------------------------------------------------------------------------ */
/* index maps                                                                                  */
/* ------------------------------------------------------------------------------------------ */

/* number of SO(3) coefficients for l = 0..B-1: sum (2l+1)^2 = B(4B^2-1)/3 */
long or_so3_ncoef(int B) { return (long)B * (4L * B * B - 1) / 3; }
long or_s2_ncoef(int B) { return (long)B * B; }

/* FFTSO3_matrix::operator()(l,m,n) = M[l](m+l, n+l), column-major (src/fdcl_FFTSO3_matrix.cpp:43-47) */
int or_so3_index(int l, int m, int n)
{
    return l * (4 * l * l - 1) / 3 + (m + l) + (n + l) * (2 * l + 1);
}
#define IX(l, m, n) or_so3_index((l), (m), (n))
/* FFTS2_matrix::operator()(l,m) = M[l](m+l) (src/fdcl_FFTS2_matrix.cpp:43-47) */
#define IX2(l, m) ((l) * (l) + (m) + (l))

static inline int wrap(int u, int N) { u %= N; return u < 0 ? u + N : u; }
static inline double sgn_pow(int e) { return (e & 1) ? -1.0 : 1.0; } /* (-1)^e, any sign of e */

static int nthreads_of(int req)
{
#ifdef _OPENMP
    return req > 0 ? req : omp_get_max_threads();
#else
    (void)req;
    return 1;
#endif
}

/* ------------------------------------------------------------------------------------------ */
/* quadrature weights                                                                          */
/* ------------------------------------------------------------------------------------------ */

/* src/fdcl_FFTSO3.cpp:236-256:
 *   w_j = sin(beta_j)/(4B^3) * sum_{k<B} sin((2j+1)(2k+1) pi/(4B)) / (2k+1) */
void or_so3_weights(int B, double *w)
{
    double factor = M_PI / (double)(4 * B);
    for (int j = 0; j < 2 * B; j++) {
        double s = 0.0;
        for (int k = 0; k < B; k++)
            s += 1.0 / (double)(2 * k + 1) * sin((double)((2 * j + 1) * (2 * k + 1)) * factor);
        s *= 1.0 / ((double)4 * B * B * B) * sin((double)(2 * j + 1) * factor);
        w[j] = s;
    }
}

/* src/fdcl_FFTS2.cpp:102-123: same sum, prefactor 2 pi / B^2 */
void or_s2_weights(int B, double *w)
{
    double factor = M_PI / (double)(4 * B);
    for (int j = 0; j < 2 * B; j++) {
        double s = 0.0;
        for (int k = 0; k < B; k++)
            s += 1.0 / (double)(2 * k + 1) * sin((double)((2 * j + 1) * (2 * k + 1)) * factor);
        s *= 2.0 * M_PI / ((double)B * B) * sin((double)(2 * j + 1) * factor);
        w[j] = s;
    }
}

static inline double beta_k(int B, int k) { return (double)(2 * k + 1) * M_PI / 4.0 / (double)B; } /* :578-581 */
static inline double alpha_j(int B, int j) { return (double)j * M_PI / (double)B; }                /* :583-586 */

/* ------------------------------------------------------------------------------------------ */
/* Wigner small-d                                                                              */
/* ------------------------------------------------------------------------------------------ */

/* src/fdcl_FFTSO3.cpp:127-194 (Blanco, Florez & Bermejo 1997, eqs. (64)-(68)):
 * lower triangle m >= |n| by the three-term recursion in l for m <= l-2, rows m = l and
 * m = l-1 from the tan(beta/2) seeds, the rest by the symmetries
 * d(l,m,n) = (-1)^(m+n) d(l,-m,-n) = d(l,-n,-m) = (-1)^(m+n) d(l,n,m). */
void or_wigner_d_ref(double beta, int L, double *d)
{
    double cb = cos(beta), sb = sin(beta);
    double sb2 = sin(beta / 2.), cb2 = cos(beta / 2.), tb2 = tan(beta / 2.);
    memset(d, 0, sizeof(double) * (size_t)or_so3_ncoef(L + 1));

    d[IX(0, 0, 0)] = 1.0;
    if (L >= 1) {
        d[IX(1, 0, 0)] = cb;
        d[IX(1, 1, -1)] = sb2 * sb2;
        d[IX(1, 1, 0)] = -sb / sqrt(2.0);
        d[IX(1, 1, 1)] = cb2 * cb2;
    }
    for (int l = 2; l <= L; l++) {
        double dl = (double)l;
        for (int m = 0; m <= l - 2; m++) {
            for (int n = -m; n <= m; n++) {
                double dm = m, dn = n;
                double den = sqrt((dl * dl - dm * dm) * (dl * dl - dn * dn));
                double a = (dl * (2. * dl - 1.) * cb - (2. * dl - 1.) * dm * dn / (dl - 1.)) * d[IX(l - 1, m, n)];
                double b = sqrt(((dl - 1.) * (dl - 1.) - dm * dm) * ((dl - 1.) * (dl - 1.) - dn * dn)) * dl / (dl - 1.) *
                           d[IX(l - 2, m, n)];
                d[IX(l, m, n)] = (a - b) / den;
            }
        }
        d[IX(l, l, l)] = d[IX(1, 1, 1)] * d[IX(l - 1, l - 1, l - 1)];              /* (65) */
        d[IX(l, l - 1, l - 1)] = (dl * cb - dl + 1.) * d[IX(l - 1, l - 1, l - 1)];   /* (66) */
        for (int n = l - 1; n >= -l; n--)                                          /* (67) */
            d[IX(l, l, n)] = -sqrt((dl + n + 1.) / (dl - n)) * tb2 * d[IX(l, l, n + 1)];
        for (int n = l - 2; n >= 1 - l; n--)                                       /* (68) */
            d[IX(l, l - 1, n)] = -(dl * cb - n) / (dl * cb - n - 1.) * sqrt((dl + n + 1.) / (dl - n)) * tb2 *
                                 d[IX(l, l - 1, n + 1)];
    }
    for (int l = 1; l <= L; l++) {
        for (int m = -l; m < 0; m++)          /* upper triangle */
            for (int n = m; n <= -m; n++)
                d[IX(l, m, n)] = sgn_pow(m + n) * d[IX(l, -m, -n)];
        for (int n = -l; n < 0; n++)          /* left triangle */
            for (int m = n + 1; m < -n; m++)
                d[IX(l, m, n)] = d[IX(l, -n, -m)];
        for (int n = 1; n <= l; n++)          /* right triangle */
            for (int m = -n + 1; m <= n - 1; m++)
                d[IX(l, m, n)] = sgn_pow(m + n) * d[IX(l, n, m)];
    }
}

/* Stable variant for large L (the reference's tan(beta/2) seeds underflow for B >~ 70).
 * Per canonical (m,n), m >= |n|: seed at l = m from the closed form
 *   d^m_{mn} = sqrt(C(2m, m+n)) cos(b/2)^(m+n) (-sin(b/2))^(m-n)
 * (agrees with src/fdcl_FFTSO3.cpp:146-148 at l=1), evaluated in log space in long double,
 * then the reference's three-term recursion (64) upward in l, also in long double (its 15-bit
 * exponent holds every seed for B <= 256), then the same symmetries as the reference. */
void or_wigner_d_stable(double beta, int L, double *d)
{
    long double cb = cosl((long double)beta);
    long double c2 = cosl((long double)beta / 2.L), s2 = sinl((long double)beta / 2.L);
    long double lc2 = logl(fabsl(c2)), ls2 = logl(fabsl(s2));
    memset(d, 0, sizeof(double) * (size_t)or_so3_ncoef(L + 1));
    long double *col = (long double *)malloc(sizeof(long double) * (size_t)(L + 2));
    for (int m = 0; m <= L; m++) {
        for (int n = -m; n <= m; n++) {
            int a = m + n, b = m - n;
            long double lg = 0.5L * (lgammal(2.L * m + 1) - lgammal((long double)a + 1) - lgammal((long double)b + 1));
            long double mag, sgn = 1.L;
            if ((a > 0 && c2 == 0.L) || (b > 0 && s2 == 0.L)) {
                mag = 0.L;
            } else {
                mag = expl(lg + (a ? a * lc2 : 0.L) + (b ? b * ls2 : 0.L));
            }
            if (c2 < 0 && (a & 1)) sgn = -sgn;
            if ((b & 1) && s2 > 0) sgn = -sgn;      /* (-sin)^b */
            if ((b & 1) && s2 < 0) sgn = sgn;       /* (-sin) > 0 */
            col[m] = sgn * mag;
            long double dm = m, dn = n;
            for (int l = m + 1; l <= L; l++) {
                long double dl = l;
                long double prev2 = (l - 2 >= m) ? col[l - 2] : 0.L;
                if (l == 1) { /* only (m,n) = (0,0) */
                    col[l] = cb * col[0];
                    continue;
                }
                long double den = sqrtl((dl * dl - dm * dm) * (dl * dl - dn * dn));
                long double A = (dl * (2.L * dl - 1.L) * cb - (2.L * dl - 1.L) * dm * dn / (dl - 1.L)) * col[l - 1];
                long double Bt = sqrtl(((dl - 1.L) * (dl - 1.L) - dm * dm) * ((dl - 1.L) * (dl - 1.L) - dn * dn)) * dl /
                                 (dl - 1.L) * prev2;
                col[l] = (A - Bt) / den;
            }
            /* scatter over the symmetry orbit of (m,n) */
            for (int l = m; l <= L; l++) {
                double v = (double)col[l];
                double s = sgn_pow(m + n);
                d[IX(l, m, n)] = v;
                d[IX(l, -m, -n)] = s * v;
                d[IX(l, -n, -m)] = v;
                d[IX(l, n, m)] = s * v;
            }
        }
    }
    free(col);
}

/* The oracle's grid transforms use the reference recursion where the reference is valid on the
 * grid (B <= 64; it agrees with the stable recursion to ~1e-13 there) and the stable recursion
 * beyond.  Pointwise evaluations (any beta) always use the stable recursion: the reference's seed
 * step (68) divides by (l cos(beta) - n - 1), which vanishes at beta = pi/2 (n = -1) and loses
 * all accuracy near it. */
static void wigner_d(double beta, int L, double *d)
{
    if (L <= 63)
        or_wigner_d_ref(beta, L, d);
    else
        or_wigner_d_stable(beta, L, d);
}

/* ------------------------------------------------------------------------------------------ */
/* normalised associated Legendre functions (src/fdcl_FFTS2.cpp:48-100; NR 3rd ed. sec 6.7)   */
/* ------------------------------------------------------------------------------------------ */
void or_legendre(double x, int L, double *nP)
{
    memset(nP, 0, sizeof(double) * (size_t)(L + 1) * (L + 1));
    nP[IX2(0, 0)] = sqrt(1. / (4.0 * M_PI));
    for (int l = 1; l <= L; l++) {
        double pmm = 1.0, omx2 = (1.0 - x) * (1.0 + x), fact = 1.0;
        for (int i = 1; i <= l; i++) {
            pmm *= omx2 * fact / (fact + 1.0);
            fact += 2.0;
        }
        pmm = sqrt((2 * l + 1) * pmm / (4.0 * M_PI));
        if (l & 1) pmm = -pmm;
        nP[IX2(l, l)] = pmm;
    }
    for (int l = 0; l < L; l++)
        nP[IX2(l + 1, l)] = x * sqrt(2.0 * l + 3.0) * nP[IX2(l, l)];
    for (int m = 0; m <= L; m++) {
        double oldfact = sqrt(2.0 * m + 3.0);
        for (int l = m + 2; l <= L; l++) {
            double fact = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
            nP[IX2(l, m)] = (x * nP[IX2(l - 1, m)] - nP[IX2(l - 2, m)] / oldfact) * fact;
            oldfact = fact;
        }
    }
    for (int l = 1; l <= L; l++)
        for (int m = 1; m <= l; m++)
            nP[IX2(l, -m)] = sgn_pow(m) * nP[IX2(l, m)];
}

/* ------------------------------------------------------------------------------------------ */
/* FFT: unnormalised forward DFT (what Eigen::FFT<double>::fwd computes; the reference calls  */
/* it at src/fdcl_FFTSO3.cpp:689-698 and src/fdcl_FFTS2.cpp:153).  Mixed-radix recursive     */
/* decimation in time for any N, twiddles evaluated in long double.                           */
/* ------------------------------------------------------------------------------------------ */
typedef struct { int N; cplx *W; cplx *tmp; } fftplan;

static void fftplan_init(fftplan *p, int N)
{
    p->N = N;
    p->W = (cplx *)malloc(sizeof(cplx) * (size_t)N);
    p->tmp = (cplx *)malloc(sizeof(cplx) * (size_t)N);
    for (int t = 0; t < N; t++) {
        long double ang = -2.L * PI_L * (long double)t / (long double)N;
        p->W[t] = (double)cosl(ang) + I * (double)sinl(ang);
    }
}
static void fftplan_free(fftplan *p) { free(p->W); free(p->tmp); }

static int smallest_factor(int n)
{
    if (n % 4 == 0) return 4;
    for (int f = 2; f * f <= n; f++)
        if (n % f == 0) return f;
    return n;
}

/* out[u] = sum_j in[j*is] W_n^{u j}, n | p->N */
static void fft_rec(const fftplan *p, int n, const cplx *in, long is, cplx *out)
{
    if (n == 1) { out[0] = in[0]; return; }
    int r = smallest_factor(n), m = n / r, step = p->N / n;
    for (int q = 0; q < r; q++)
        fft_rec(p, m, in + q * is, is * r, out + q * m);
    cplx buf[16];
    cplx *acc = (r <= 16) ? buf : (cplx *)malloc(sizeof(cplx) * (size_t)r);
    for (int k = 0; k < m; k++) {
        for (int s = 0; s < r; s++) {
            cplx a = 0;
            int u = k + s * m;
            for (int q = 0; q < r; q++)
                a += out[q * m + k] * p->W[(long)((long)q * u % n) * step];
            acc[s] = a;
        }
        for (int s = 0; s < r; s++) out[s * m + k] = acc[s];  /* in place: column k only */
    }
    if (acc != buf) free(acc);
}

static void fft_strided(const fftplan *p, const cplx *in, long is, cplx *out, long os, cplx *work)
{
    fft_rec(p, p->N, in, is, work);
    for (int u = 0; u < p->N; u++) out[u * os] = work[u];
}

void or_fft(int N, const double *in, double *out)
{
    fftplan p;
    fftplan_init(&p, N);
    cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)N);
    fft_strided(&p, (const cplx *)in, 1, (cplx *)out, 1, work);
    free(work);
    fftplan_free(&p);
}

/* 2-D DFT of an N x N slice [j1][j2]: rows (over j2) then columns (over j1), as the reference
 * does at src/fdcl_FFTSO3.cpp:689-698 */
static void fft2(const fftplan *p, const cplx *in, cplx *out, cplx *work, cplx *colbuf)
{
    int N = p->N;
    for (int j1 = 0; j1 < N; j1++)
        fft_strided(p, in + (long)j1 * N, 1, out + (long)j1 * N, 1, work);
    for (int j2 = 0; j2 < N; j2++) {
        for (int j1 = 0; j1 < N; j1++) colbuf[j1] = out[(long)j1 * N + j2];
        fft_strided(p, colbuf, 1, out + j2, N, work);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* SO(3) complex transforms                                                                    */
/* ------------------------------------------------------------------------------------------ */

/* src/fdcl_FFTSO3.cpp:659-763 (complex forward, is_real selects the Hermitian-half branch):
 *   S_k = 2-D DFT of the beta_k slice;  F(l,m,n) += w_k d^l_mn(beta_k) S_k(-m mod 2B, -n mod 2B) */
static void so3_forward_impl(int B, const cplx *grid, cplx *F, int is_real, int nthreads)
{
    int N = 2 * B, L = B - 1;
    long nc = or_so3_ncoef(B);
    double *w = (double *)malloc(sizeof(double) * (size_t)N);
    or_so3_weights(B, w);
    int nth = nthreads_of(nthreads);
    cplx *part = (cplx *)calloc((size_t)nth * (size_t)nc, sizeof(cplx));
#pragma omp parallel num_threads(nth)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        cplx *Fl = part + (size_t)t * nc;
        double *d = (double *)malloc(sizeof(double) * (size_t)nc);
        cplx *S = (cplx *)malloc(sizeof(cplx) * (size_t)N * N);
        cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)N);
        cplx *colbuf = (cplx *)malloc(sizeof(cplx) * (size_t)N);
        fftplan p;
        fftplan_init(&p, N);
#pragma omp for schedule(static)
        for (int k = 0; k < N; k++) {
            fft2(&p, grid + (long)k * N * N, S, work, colbuf);
            wigner_d(beta_k(B, k), L, d);
            for (int l = 0; l <= L; l++) {
                int mlo = is_real ? 0 : -l;
                for (int m = mlo; m <= l; m++) {
                    int nhi = (is_real && m == 0) ? 0 : l;
                    for (int n = -l; n <= nhi; n++)
                        Fl[IX(l, m, n)] += w[k] * d[IX(l, m, n)] * S[(long)wrap(-m, N) * N + wrap(-n, N)];
                }
            }
        }
        fftplan_free(&p);
        free(d); free(S); free(work); free(colbuf);
    }
    memset(F, 0, sizeof(cplx) * (size_t)nc);
    for (int t = 0; t < nth; t++)
        for (long i = 0; i < nc; i++) F[i] += part[(size_t)t * nc + i];
    free(part);
    free(w);
    if (is_real) { /* conjugate-symmetric fill, src/fdcl_FFTSO3.cpp:746-758 */
        for (int l = 0; l <= L; l++) {
            for (int m = -l; m < 0; m++)
                for (int n = -l; n <= l; n++)
                    F[IX(l, m, n)] = sgn_pow(m - n) * conj(F[IX(l, -m, -n)]);
            for (int n = 1; n <= l; n++)
                F[IX(l, 0, n)] = sgn_pow(-n) * conj(F[IX(l, 0, -n)]);
        }
    }
}

void or_so3_forward(int B, const double *grid, double *F, int nthreads)
{
    so3_forward_impl(B, (const cplx *)grid, (cplx *)F, 0, nthreads);
}

/* Grid inverse.  The reference evaluates the series only pointwise
 * (src/fdcl_FFTSO3.cpp:498-531); on the grid the same sum separates into
 *   G_k(m,n) = sum_{l >= max(|m|,|n|)} (2l+1) F(l,m,n) d^l_mn(beta_k)
 *   f(alpha_j1, beta_k, gamma_j2) = sum_{m,n} G_k(m,n) exp(-2 pi i (j1 m + j2 n)/(2B)),
 * an unnormalised forward 2-D DFT with G_k(m,n) at bin (m mod 2B, n mod 2B). */
void or_so3_inverse(int B, const double *Fd, double *gridd, int nthreads)
{
    const cplx *F = (const cplx *)Fd;
    cplx *grid = (cplx *)gridd;
    int N = 2 * B, L = B - 1;
    long nc = or_so3_ncoef(B);
    int nth = nthreads_of(nthreads);
#pragma omp parallel num_threads(nth)
    {
        double *d = (double *)malloc(sizeof(double) * (size_t)nc);
        cplx *X = (cplx *)malloc(sizeof(cplx) * (size_t)N * N);
        cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)N);
        cplx *colbuf = (cplx *)malloc(sizeof(cplx) * (size_t)N);
        fftplan p;
        fftplan_init(&p, N);
#pragma omp for schedule(static)
        for (int k = 0; k < N; k++) {
            wigner_d(beta_k(B, k), L, d);
            memset(X, 0, sizeof(cplx) * (size_t)N * N);
            for (int m = -L; m <= L; m++)
                for (int n = -L; n <= L; n++) {
                    cplx g = 0;
                    int l0 = abs(m) > abs(n) ? abs(m) : abs(n);
                    for (int l = l0; l <= L; l++) g += (double)(2 * l + 1) * F[IX(l, m, n)] * d[IX(l, m, n)];
                    X[(long)wrap(m, N) * N + wrap(n, N)] = g;
                }
            fft2(&p, X, grid + (long)k * N * N, work, colbuf);
        }
        fftplan_free(&p);
        free(d); free(X); free(work); free(colbuf);
    }
}

/* src/fdcl_FFTSO3.cpp:498-531: f = e(alpha)^T Fmn e(gamma),
 * Fmn(m,n) = sum_{l >= max(|m|,|n|)} (2l+1) F(l,m,n) d^l_mn(beta), e_m(x) = exp(-i m x) */
void or_so3_inverse_point(int B, const double *Fd, double alpha, double beta, double gamma, double *out2)
{
    const cplx *F = (const cplx *)Fd;
    int L = B - 1;
    double *d = (double *)malloc(sizeof(double) * (size_t)or_so3_ncoef(B));
    or_wigner_d_stable(beta, L, d);
    cplx f = 0;
    for (int m = -L; m <= L; m++) {
        cplx ea = cexp(-I * (double)m * alpha);
        for (int n = -L; n <= L; n++) {
            cplx g = 0;
            int l0 = abs(m) > abs(n) ? abs(m) : abs(n);
            for (int l = l0; l <= L; l++) g += (double)(2 * l + 1) * F[IX(l, m, n)] * d[IX(l, m, n)];
            f += ea * g * cexp(-I * (double)n * gamma);
        }
    }
    free(d);
    out2[0] = creal(f);
    out2[1] = cimag(f);
}

/* src/fdcl_FFTSO3.cpp:612-652 (forward_transform_0): the direct O(B^6) sum
 *   F(l,m,n) = sum_{k,j1,j2} w_k exp(i m alpha_j1) f exp(i n gamma_j2) d^l_mn(beta_k),
 * accumulated in long double; for tiny B only. */
void or_so3_forward_direct(int B, const double *gridd, double *Fd)
{
    const cplx *grid = (const cplx *)gridd;
    int N = 2 * B, L = B - 1;
    long nc = or_so3_ncoef(B);
    double *w = (double *)malloc(sizeof(double) * (size_t)N);
    double *d = (double *)malloc(sizeof(double) * (size_t)nc);
    lcplx *acc = (lcplx *)calloc((size_t)nc, sizeof(lcplx));
    or_so3_weights(B, w);
    for (int k = 0; k < N; k++) {
        or_wigner_d_stable(beta_k(B, k), L, d);
        for (int j1 = 0; j1 < N; j1++)
            for (int j2 = 0; j2 < N; j2++) {
                lcplx f = grid[((long)k * N + j1) * N + j2];
                for (int l = 0; l <= L; l++)
                    for (int m = -l; m <= l; m++)
                        for (int n = -l; n <= l; n++) {
                            long double ang = (long double)m * (long double)alpha_j(B, j1) +
                                              (long double)n * (long double)alpha_j(B, j2);
                            lcplx e = cosl(ang) + I * sinl(ang);
                            acc[IX(l, m, n)] += (long double)w[k] * e * f * (long double)d[IX(l, m, n)];
                        }
            }
    }
    cplx *F = (cplx *)Fd;
    for (long i = 0; i < nc; i++) F[i] = (cplx)acc[i];
    free(w); free(d); free(acc);
}

/* ------------------------------------------------------------------------------------------ */
/* SO(3) real transforms                                                                       */
/* ------------------------------------------------------------------------------------------ */

/* src/fdcl_FFTSO3.cpp:1107-1139 (matrix2rsph): unitary T^l with
 *   T(m,m) = i/sqrt2, T(m,-m) = -i(-1)^m/sqrt2 (m<0);  T(0,0) = 1;
 *   T(m,-m) = 1/sqrt2, T(m,m) = (-1)^m/sqrt2 (m>0) */
void or_so3_T(int L, double *Td)
{
    cplx *T = (cplx *)Td;
    memset(T, 0, sizeof(cplx) * (size_t)or_so3_ncoef(L + 1));
    for (int l = 0; l <= L; l++) {
        for (int m = -l; m < 0; m++) {
            T[IX(l, m, m)] = I / sqrt(2.);
            T[IX(l, m, -m)] = -I / sqrt(2.) * sgn_pow(m);
        }
        T[IX(l, 0, 0)] = 1.;
        for (int m = 1; m <= l; m++) {
            T[IX(l, m, -m)] = 1. / sqrt(2.);
            T[IX(l, m, m)] = sgn_pow(m) / sqrt(2.);
        }
    }
}

/* dense per-l products: Fr = Re(T Fc T^H) (src/fdcl_FFTSO3.cpp:1174-1180) */
void or_so3_c2r(int B, const double *Fcd, double *Fr)
{
    const cplx *Fc = (const cplx *)Fcd;
    int L = B - 1;
    long nc = or_so3_ncoef(B);
    cplx *T = (cplx *)malloc(sizeof(cplx) * (size_t)nc);
    or_so3_T(L, (double *)T);
    for (int l = 0; l <= L; l++) {
        int w = 2 * l + 1;
        cplx *tmp = (cplx *)calloc((size_t)w * w, sizeof(cplx)); /* T Fc */
        for (int a = -l; a <= l; a++)
            for (int b = -l; b <= l; b++) {
                cplx s = 0;
                for (int c = -l; c <= l; c++) s += T[IX(l, a, c)] * Fc[IX(l, c, b)];
                tmp[(a + l) + (b + l) * w] = s;
            }
        for (int a = -l; a <= l; a++)
            for (int b = -l; b <= l; b++) {
                cplx s = 0;
                for (int c = -l; c <= l; c++) s += tmp[(a + l) + (c + l) * w] * conj(T[IX(l, b, c)]);
                Fr[IX(l, a, b)] = creal(s);
            }
        free(tmp);
    }
    free(T);
}

/* inverse of the above: Fc = T^H Fr T (T unitary) */
void or_so3_r2c(int B, const double *Fr, double *Fcd)
{
    cplx *Fc = (cplx *)Fcd;
    int L = B - 1;
    long nc = or_so3_ncoef(B);
    cplx *T = (cplx *)malloc(sizeof(cplx) * (size_t)nc);
    or_so3_T(L, (double *)T);
    for (int l = 0; l <= L; l++) {
        int w = 2 * l + 1;
        cplx *tmp = (cplx *)calloc((size_t)w * w, sizeof(cplx)); /* T^H Fr */
        for (int a = -l; a <= l; a++)
            for (int b = -l; b <= l; b++) {
                cplx s = 0;
                for (int c = -l; c <= l; c++) s += conj(T[IX(l, c, a)]) * Fr[IX(l, c, b)];
                tmp[(a + l) + (b + l) * w] = s;
            }
        for (int a = -l; a <= l; a++)
            for (int b = -l; b <= l; b++) {
                cplx s = 0;
                for (int c = -l; c <= l; c++) s += tmp[(a + l) + (c + l) * w] * T[IX(l, c, b)];
                Fc[IX(l, a, b)] = s;
            }
        free(tmp);
    }
    free(T);
}

static int signum_i(int x) { return (x > 0) - (x < 0); } /* src/fdcl_FFTSO3.cpp:1771-1782 */

/* src/fdcl_FFTSO3.cpp:1431-1471 (compute_Psi) */
static void compute_psi(const double *d, int L, double *Psi)
{
    for (int l = 0; l <= L; l++) {
        Psi[IX(l, 0, 0)] = d[IX(l, 0, 0)];
        for (int n = 1; n <= l; n++) {
            double C = sqrt(2.) * d[IX(l, 0, n)];
            Psi[IX(l, 0, n)] = sgn_pow(-n) * C;
            Psi[IX(l, 0, -n)] = Psi[IX(l, 0, n)];
        }
        for (int m = 1; m <= l; m++) {
            double C = sqrt(2.) * d[IX(l, m, 0)];
            Psi[IX(l, m, 0)] = sgn_pow(m) * C;
            Psi[IX(l, -m, 0)] = Psi[IX(l, m, 0)];
            for (int n = 1; n <= l; n++) {
                double A = sgn_pow(m - n) * d[IX(l, m, n)];
                double Bv = sgn_pow(m) * signum_i(m) * d[IX(l, m, -n)];
                Psi[IX(l, m, n)] = A + Bv;
                Psi[IX(l, -m, n)] = A - Bv;
                Psi[IX(l, m, -n)] = Psi[IX(l, m, n)];
                Psi[IX(l, -m, -n)] = Psi[IX(l, -m, n)];
            }
        }
    }
}

/* src/fdcl_FFTSO3.cpp:774-850 (FFTSO3_real::forward_transform, the reference's public real
 * forward): cos/sin combinations of the 2-D DFT at bins (-m, +-n) against Psi(l,+-m,n). */
void or_so3_real_forward(int B, const double *grid, double *Fr, int nthreads)
{
    int N = 2 * B, L = B - 1;
    long nc = or_so3_ncoef(B);
    double *w = (double *)malloc(sizeof(double) * (size_t)N);
    or_so3_weights(B, w);
    int nth = nthreads_of(nthreads);
    double *part = (double *)calloc((size_t)nth * (size_t)nc, sizeof(double));
#pragma omp parallel num_threads(nth)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        double *Fl = part + (size_t)t * nc;
        double *d = (double *)malloc(sizeof(double) * (size_t)nc);
        double *Psi = (double *)malloc(sizeof(double) * (size_t)nc);
        cplx *X = (cplx *)malloc(sizeof(cplx) * (size_t)N * N);
        cplx *S = (cplx *)malloc(sizeof(cplx) * (size_t)N * N);
        cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)N);
        cplx *colbuf = (cplx *)malloc(sizeof(cplx) * (size_t)N);
        fftplan p;
        fftplan_init(&p, N);
#pragma omp for schedule(static)
        for (int k = 0; k < N; k++) {
            for (long i = 0; i < (long)N * N; i++) X[i] = grid[(long)k * N * N + i];
            fft2(&p, X, S, work, colbuf);
            wigner_d(beta_k(B, k), L, d);
            compute_psi(d, L, Psi);
            for (int l = 0; l <= L; l++)
                for (int m = -l; m <= l; m++) {
                    int im = wrap(-m, N);
                    for (int n = -l; n <= l; n++) {
                        int in = wrap(-n, N), i_n = wrap(n, N);
                        double cos_mang = creal(S[(long)im * N + in]);
                        double cos_ma_ng = creal(S[(long)im * N + i_n]);
                        double sin_mang = cimag(S[(long)im * N + in]);
                        double sin_ma_ng = cimag(S[(long)im * N + i_n]);
                        double ss = 0.5 * (cos_ma_ng - cos_mang);
                        double cc = 0.5 * (cos_ma_ng + cos_mang);
                        double sc = 0.5 * (sin_mang + sin_ma_ng);
                        double cs = 0.5 * (sin_mang - sin_ma_ng);
                        if ((m >= 0 && n >= 0) || (m < 0 && n < 0))
                            Fl[IX(l, m, n)] += w[k] * (-ss * Psi[IX(l, -m, n)] + cc * Psi[IX(l, m, n)]);
                        else
                            Fl[IX(l, m, n)] += w[k] * (-sc * Psi[IX(l, -m, n)] + cs * Psi[IX(l, m, n)]);
                    }
                }
        }
        fftplan_free(&p);
        free(d); free(Psi); free(X); free(S); free(work); free(colbuf);
    }
    memset(Fr, 0, sizeof(double) * (size_t)nc);
    for (int t = 0; t < nth; t++)
        for (long i = 0; i < nc; i++) Fr[i] += part[(size_t)t * nc + i];
    free(part);
    free(w);
}

/* src/fdcl_FFTSO3.cpp:1169-1181 (forward_transform_2): Fr = Re(T Fc T^H), Fc from the complex
 * forward with the real-input branch */
void or_so3_real_forward_T(int B, const double *grid, double *Fr, int nthreads)
{
    int N = 2 * B;
    long nc = or_so3_ncoef(B), ng = (long)N * N * N;
    cplx *g = (cplx *)malloc(sizeof(cplx) * (size_t)ng);
    for (long i = 0; i < ng; i++) g[i] = grid[i];
    cplx *Fc = (cplx *)malloc(sizeof(cplx) * (size_t)nc);
    so3_forward_impl(B, g, Fc, 1, nthreads);
    or_so3_c2r(B, (const double *)Fc, Fr);
    free(g);
    free(Fc);
}

/* Real grid inverse: f = sum (2l+1) sum Fr U with U = conj(T) D T^T (src/fdcl_FFTSO3.cpp:1015-1019,
 * 1141-1157) equals the complex inverse of Fc = T^H Fr T; the grid values are its real part. */
void or_so3_real_inverse(int B, const double *Fr, double *grid, int nthreads)
{
    int N = 2 * B;
    long nc = or_so3_ncoef(B), ng = (long)N * N * N;
    cplx *Fc = (cplx *)malloc(sizeof(cplx) * (size_t)nc);
    cplx *g = (cplx *)malloc(sizeof(cplx) * (size_t)ng);
    or_so3_r2c(B, Fr, (double *)Fc);
    or_so3_inverse(B, (const double *)Fc, (double *)g, nthreads);
    for (long i = 0; i < ng; i++) grid[i] = creal(g[i]);
    free(Fc);
    free(g);
}

/* src/fdcl_FFTSO3.cpp:1062-1105 (real_harmonics) */
void or_so3_real_harmonics(int L, double alpha, double beta, double gamma, double *U)
{
    long nc = or_so3_ncoef(L + 1);
    double *d = (double *)malloc(sizeof(double) * (size_t)nc);
    or_wigner_d_stable(beta, L, d);
    memset(U, 0, sizeof(double) * (size_t)nc);
    for (int l = 0; l <= L; l++) {
        U[IX(l, 0, 0)] = d[IX(l, 0, 0)];
        for (int m = 1; m <= l; m++) {
            U[IX(l, -m, 0)] = sqrt(2.) * d[IX(l, -m, 0)] * sin((double)m * alpha);
            U[IX(l, m, 0)] = sgn_pow(m) * sqrt(2.) * d[IX(l, m, 0)] * cos((double)m * alpha);
            for (int n = 1; n <= l; n++) {
                double cmg = cos((double)m * alpha + (double)n * gamma);
                double c_mg = cos((double)m * alpha - (double)n * gamma);
                double smg = sin((double)m * alpha + (double)n * gamma);
                double s_mg = sin((double)m * alpha - (double)n * gamma);
                U[IX(l, m, n)] = sgn_pow(m + n) * d[IX(l, m, n)] * cmg + sgn_pow(m) * d[IX(l, m, -n)] * c_mg;
                U[IX(l, m, -n)] = sgn_pow(m) * d[IX(l, m, -n)] * s_mg - sgn_pow(m + n) * d[IX(l, m, n)] * smg;
                U[IX(l, -m, n)] = -sgn_pow(n) * d[IX(l, -m, n)] * -s_mg + d[IX(l, -m, -n)] * smg;
                U[IX(l, -m, -n)] = d[IX(l, -m, -n)] * cmg - sgn_pow(n) * d[IX(l, -m, n)] * c_mg;
            }
        }
        for (int n = 1; n <= l; n++) {
            U[IX(l, 0, -n)] = -sqrt(2.) * d[IX(l, 0, -n)] * sin((double)n * gamma);
            U[IX(l, 0, n)] = sgn_pow(n) * sqrt(2.) * d[IX(l, 0, n)] * cos((double)n * gamma);
        }
    }
    free(d);
}

/* src/fdcl_FFTSO3.cpp:1141-1157 */
double or_so3_real_inverse_point(int B, const double *Fr, double alpha, double beta, double gamma)
{
    int L = B - 1;
    long nc = or_so3_ncoef(B);
    double *U = (double *)malloc(sizeof(double) * (size_t)nc);
    or_so3_real_harmonics(L, alpha, beta, gamma, U);
    double f = 0.;
    for (int l = 0; l <= L; l++)
        for (int m = -l; m <= l; m++)
            for (int n = -l; n <= l; n++) f += (double)(2 * l + 1) * Fr[IX(l, m, n)] * U[IX(l, m, n)];
    free(U);
    return f;
}

/* ------------------------------------------------------------------------------------------ */
/* S^2                                                                                         */
/* ------------------------------------------------------------------------------------------ */

/* src/fdcl_FFTS2.cpp:125-192: S_k = DFT of ring k over phi;
 *   F(l,m) += w_k nP^l_m(cos theta_k) S_k(m mod 2B) */
static void s2_forward_impl(int B, const cplx *grid, cplx *F, int is_real)
{
    int N = 2 * B, L = B - 1;
    long nc = or_s2_ncoef(B);
    double *w = (double *)malloc(sizeof(double) * (size_t)N);
    double *nP = (double *)malloc(sizeof(double) * (size_t)nc);
    cplx *S = (cplx *)malloc(sizeof(cplx) * (size_t)N);
    cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)N);
    fftplan p;
    fftplan_init(&p, N);
    or_s2_weights(B, w);
    memset(F, 0, sizeof(cplx) * (size_t)nc);
    for (int k = 0; k < N; k++) {
        fft_strided(&p, grid + (long)k * N, 1, S, 1, work);
        or_legendre(cos(beta_k(B, k)), L, nP);
        for (int l = 0; l <= L; l++) {
            for (int m = 0; m <= l; m++) F[IX2(l, m)] += w[k] * nP[IX2(l, m)] * S[m];
            if (!is_real)
                for (int m = -l; m < 0; m++) F[IX2(l, m)] += w[k] * nP[IX2(l, m)] * S[N + m];
        }
    }
    if (is_real)
        for (int l = 0; l <= L; l++)
            for (int m = -l; m < 0; m++) F[IX2(l, m)] = sgn_pow(m) * conj(F[IX2(l, -m)]);
    fftplan_free(&p);
    free(w); free(nP); free(S); free(work);
}

void or_s2_forward(int B, const double *grid, double *F, int nthreads)
{
    (void)nthreads;
    s2_forward_impl(B, (const cplx *)grid, (cplx *)F, 0);
}

/* Grid inverse (the reference is pointwise only, src/fdcl_FFTS2.cpp:194-230):
 *   G_k(m) = sum_{l>=|m|} F(l,m) nP^l_m(cos theta_k);  f(theta_k, phi_j) = sum_m G_k(m) e^{i m phi_j},
 * an unnormalised forward DFT with G_k(m) at bin (-m mod 2B). */
void or_s2_inverse(int B, const double *Fd, double *gridd, int nthreads)
{
    (void)nthreads;
    const cplx *F = (const cplx *)Fd;
    cplx *grid = (cplx *)gridd;
    int N = 2 * B, L = B - 1;
    double *nP = (double *)malloc(sizeof(double) * (size_t)or_s2_ncoef(B));
    cplx *X = (cplx *)malloc(sizeof(cplx) * (size_t)N);
    cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)N);
    fftplan p;
    fftplan_init(&p, N);
    for (int k = 0; k < N; k++) {
        or_legendre(cos(beta_k(B, k)), L, nP);
        memset(X, 0, sizeof(cplx) * (size_t)N);
        for (int m = -L; m <= L; m++) {
            cplx g = 0;
            for (int l = abs(m); l <= L; l++) g += F[IX2(l, m)] * nP[IX2(l, m)];
            X[wrap(-m, N)] = g;
        }
        fft_strided(&p, X, 1, grid + (long)k * N, 1, work);
    }
    fftplan_free(&p);
    free(nP); free(X); free(work);
}

/* src/fdcl_FFTS2.cpp:17-46 (spherical_harmonics) and :194-230 (inverse_transform) */
void or_s2_inverse_point(int B, const double *Fd, double theta, double phi, double *out2)
{
    const cplx *F = (const cplx *)Fd;
    int L = B - 1;
    double *nP = (double *)malloc(sizeof(double) * (size_t)or_s2_ncoef(B));
    or_legendre(cos(theta), L, nP);
    cplx y = 0;
    for (int l = 0; l <= L; l++)
        for (int m = 0; m <= l; m++) {
            cplx Y = nP[IX2(l, m)] * cexp(I * (double)m * phi);
            y += F[IX2(l, m)] * Y;
            if (m > 0) y += F[IX2(l, -m)] * sgn_pow(m) * conj(Y);
        }
    free(nP);
    out2[0] = creal(y);
    out2[1] = cimag(y);
}

/* src/fdcl_FFTS2.cpp:424-436: Fr[l] = Re(conj(T[l]) Fc[l]), Fc from the real-input forward;
 * T as at src/fdcl_FFTS2.cpp:324-369 (same entries as the SO(3) T). */
void or_s2_real_forward(int B, const double *grid, double *Fr, int nthreads)
{
    (void)nthreads;
    int N = 2 * B, L = B - 1;
    long ng = (long)N * N;
    cplx *g = (cplx *)malloc(sizeof(cplx) * (size_t)ng);
    for (long i = 0; i < ng; i++) g[i] = grid[i];
    cplx *Fc = (cplx *)malloc(sizeof(cplx) * (size_t)or_s2_ncoef(B));
    s2_forward_impl(B, g, Fc, 1);
    cplx *T = (cplx *)malloc(sizeof(cplx) * (size_t)or_so3_ncoef(B));
    or_so3_T(L, (double *)T);
    for (int l = 0; l <= L; l++)
        for (int a = -l; a <= l; a++) {
            cplx s = 0;
            for (int c = -l; c <= l; c++) s += conj(T[IX(l, a, c)]) * Fc[IX2(l, c)];
            Fr[IX2(l, a)] = creal(s);
        }
    free(g); free(Fc); free(T);
}

/* Real grid inverse: y = T Y (src/fdcl_FFTS2.cpp:383-422), so sum Fr y = sum (T^T Fr) Y;
 * the grid is the real part of the complex inverse of Fc = T^T Fr. */
void or_s2_real_inverse(int B, const double *Fr, double *grid, int nthreads)
{
    int N = 2 * B, L = B - 1;
    long ng = (long)N * N;
    cplx *T = (cplx *)malloc(sizeof(cplx) * (size_t)or_so3_ncoef(B));
    or_so3_T(L, (double *)T);
    cplx *Fc = (cplx *)malloc(sizeof(cplx) * (size_t)or_s2_ncoef(B));
    for (int l = 0; l <= L; l++)
        for (int a = -l; a <= l; a++) {
            cplx s = 0;
            for (int c = -l; c <= l; c++) s += T[IX(l, c, a)] * Fr[IX2(l, c)];
            Fc[IX2(l, a)] = s;
        }
    cplx *g = (cplx *)malloc(sizeof(cplx) * (size_t)ng);
    or_s2_inverse(B, (const double *)Fc, (double *)g, nthreads);
    for (long i = 0; i < ng; i++) grid[i] = creal(g[i]);
    free(T); free(Fc); free(g);
}

/* src/fdcl_FFTS2.cpp:383-422 (real spherical_harmonics) and :500-512 (inverse_transform) */
double or_s2_real_inverse_point(int B, const double *Fr, double theta, double phi)
{
    int L = B - 1;
    double *nP = (double *)malloc(sizeof(double) * (size_t)or_s2_ncoef(B));
    or_legendre(cos(theta), L, nP);
    double z = 0.;
    for (int l = 0; l <= L; l++) {
        z += Fr[IX2(l, 0)] * nP[IX2(l, 0)];
        for (int m = 1; m <= l; m++) {
            double tmp = sqrt(2.) * sgn_pow(m) * nP[IX2(l, m)];
            z += Fr[IX2(l, m)] * tmp * cos((double)m * phi);
            z += Fr[IX2(l, -m)] * tmp * sin((double)m * phi);
        }
    }
    free(nP);
    return z;
}

/* ------------------------------------------------------------------------------------------ */
/* rotations (src/misc_matrix_func.cpp:144-178)                                                */
/* ------------------------------------------------------------------------------------------ */
void or_euler323_to_R(double a, double b, double g, double *R)
{
    double ca = cos(a), sa = sin(a), cb = cos(b), sb = sin(b), cg = cos(g), sg = sin(g);
    R[0] = ca * cb * cg - sa * sg;  R[1] = -cg * sa - ca * cb * sg; R[2] = ca * sb;
    R[3] = ca * sg + cb * cg * sa;  R[4] = ca * cg - cb * sa * sg;  R[5] = sa * sb;
    R[6] = -cg * sb;                R[7] = sb * sg;                 R[8] = cb;
}

void or_R_to_euler323(const double *R, double *abg)
{
    double a, b, g;
    b = acos(R[8]);
    if (b != 0.) {
        g = atan2(R[7], -R[6]);
        a = atan2(R[5], R[2]);
    } else {
        a = acos(R[0]) / 2.;
        g = a;
    }
    abg[0] = a; abg[1] = b; abg[2] = g;
}
