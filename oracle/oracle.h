/*
 * oracle.h -- FP64 CPU restatement of the reference's SO(3) / S^2 harmonic transforms.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * Nothing in fftso3_amd/ links, loads or calls it.
 *
 * Conventions (reference fdcl-gwu/FFTSO3):
 *   B = l_max + 1, N = 2B                                  src/fdcl_FFTSO3.cpp:8-14
 *   alpha_j = gamma_j = j*pi/B, beta_k = (2k+1)*pi/(4B)     src/fdcl_FFTSO3.cpp:578-591
 *   D^l_mn(a,b,g) = exp(-i m a) d^l_mn(b) exp(-i n g)      src/fdcl_FFTSO3.cpp:201-224
 *   SO(3) packed coefficient index: off(l) + (m+l) + (n+l)(2l+1), off(l) = l(4l^2-1)/3
 *     (column-major (2l+1)^2 block per l, as FFTSO3_matrix::operator() at
 *      src/fdcl_FFTSO3_matrix.cpp:43-47)
 *   S^2 packed index: l^2 + m + l                            src/fdcl_FFTS2_matrix.cpp:43-47
 *   SO(3) grid layout: [k][j1][j2]  (beta slice major, gamma fastest)
 *   S^2 grid layout:   [k][j]       (theta ring major, phi fastest)
 *   complex arrays are interleaved (re, im) doubles.
 */
#ifndef FFTSO3_ORACLE_H
#define FFTSO3_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

long or_so3_ncoef(int B);
long or_s2_ncoef(int B);
int or_so3_index(int l, int m, int n);

/* quadrature weights, 2B values */
void or_so3_weights(int B, double *w);
void or_s2_weights(int B, double *w);

/* Wigner small-d for l = 0..L, packed SO(3) layout (ncoef(L+1) doubles) */
void or_wigner_d_ref(double beta, int L, double *d);     /* reference recursion (64)-(68) */
void or_wigner_d_stable(double beta, int L, double *d);  /* log-space seeds, long double */

/* normalised associated Legendre nP^l_m(x), packed S^2 layout ((L+1)^2 doubles) */
void or_legendre(double x, int L, double *nP);

/* unnormalised forward DFT, X_u = sum_j x_j exp(-2 pi i u j / N) (complex, interleaved) */
void or_fft(int N, const double *in, double *out);

/* SO(3) complex transforms, grid (2B)^3 complex, coefficients ncoef(B) complex */
void or_so3_forward(int B, const double *grid, double *F, int nthreads);
void or_so3_inverse(int B, const double *F, double *grid, int nthreads);
void or_so3_inverse_point(int B, const double *F, double alpha, double beta, double gamma,
                          double *out2);
void or_so3_forward_direct(int B, const double *grid, double *F); /* O(B^6) brute force */

/* SO(3) real transforms, grid (2B)^3 real, coefficients ncoef(B) real */
void or_so3_real_forward(int B, const double *grid, double *Fr, int nthreads);     /* Psi path */
void or_so3_real_forward_T(int B, const double *grid, double *Fr, int nthreads);   /* via T */
void or_so3_real_inverse(int B, const double *Fr, double *grid, int nthreads);
double or_so3_real_inverse_point(int B, const double *Fr, double alpha, double beta, double gamma);
void or_so3_real_harmonics(int L, double alpha, double beta, double gamma, double *U);
void or_so3_T(int L, double *T);            /* complex, packed SO(3) layout */
void or_so3_c2r(int B, const double *Fc, double *Fr);  /* Fr = Re(T Fc T^H) */
void or_so3_r2c(int B, const double *Fr, double *Fc);  /* Fc = T^H Fr T */

/* S^2 transforms, grid (2B)^2, coefficients B^2 */
void or_s2_forward(int B, const double *grid, double *F, int nthreads);
void or_s2_inverse(int B, const double *F, double *grid, int nthreads);
void or_s2_inverse_point(int B, const double *F, double theta, double phi, double *out2);
void or_s2_real_forward(int B, const double *grid, double *Fr, int nthreads);
void or_s2_real_inverse(int B, const double *Fr, double *grid, int nthreads);
double or_s2_real_inverse_point(int B, const double *Fr, double theta, double phi);

/* rotation matrix helpers (3-2-3 Euler), row-major 3x3 */
void or_euler323_to_R(double a, double b, double g, double *R);
void or_R_to_euler323(const double *R, double *abg);

#ifdef __cplusplus
}
#endif
#endif
