// test_sanitize.cpp -- host code run under AddressSanitizer + UndefinedBehaviorSanitizer and
// under ThreadSanitizer: the FP64 oracle (its OpenMP transforms at several thread counts) and
// the class layer's OpenMP sampling step (fdcl_sampling.cpp).  Exit status 0 iff every check
// passes; the sanitizers abort (or report, for ThreadSanitizer) on any error they find.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../oracle/oracle.h"
#include "fdcl_sampling.hpp"

static int failures = 0;

static void expect(const char *name, double err, double tol)
{
    const bool ok = std::isfinite(err) && err <= tol;
    std::printf("%s %s err=%.3e tol=%.1e\n", ok ? "PASS" : "FAIL", name, err, tol);
    if (!ok) failures++;
}

static double rel(const std::vector<double> &a, const std::vector<double> &b)
{
    double n = 0, d = 0;
    for (size_t i = 0; i < a.size(); i++) {
        n += (a[i] - b[i]) * (a[i] - b[i]);
        d += b[i] * b[i];
    }
    return std::sqrt(n / d);
}

static std::vector<double> random_vec(size_t n, unsigned seed)
{
    std::vector<double> v(n);
    unsigned s = seed;
    for (auto &x : v) {
        s = s * 1664525u + 1013904223u;
        x = (double)(s >> 8) / (double)(1u << 24) - 0.5;
    }
    return v;
}

int main()
{
    // oracle round trips, complex and real, SO(3) and S^2, at 1 and 4 threads
    for (int B : {2, 6, 11}) {
        for (int th : {1, 4}) {
            const long nc = or_so3_ncoef(B), N = 2 * B;
            std::vector<double> F = random_vec(2 * nc, 100 + B), g(2 * N * N * N), F2(2 * nc);
            or_so3_inverse(B, F.data(), g.data(), th);
            or_so3_forward(B, g.data(), F2.data(), th);
            char name[64];
            std::snprintf(name, sizeof name, "oracle_so3_complex_B%d_t%d", B, th);
            expect(name, rel(F2, F), 1e-12);
            std::vector<double> Fr = random_vec(nc, 200 + B), gr(N * N * N), Fr2(nc), Fr3(nc);
            or_so3_real_inverse(B, Fr.data(), gr.data(), th);
            or_so3_real_forward(B, gr.data(), Fr2.data(), th);
            or_so3_real_forward_T(B, gr.data(), Fr3.data(), th);
            std::snprintf(name, sizeof name, "oracle_so3_real_B%d_t%d", B, th);
            expect(name, std::max(rel(Fr2, Fr), rel(Fr3, Fr)), 1e-12);
            std::vector<double> S = random_vec(2 * B * B, 300 + B), gs(2 * N * N), S2(2 * B * B);
            or_s2_inverse(B, S.data(), gs.data(), th);
            or_s2_forward(B, gs.data(), S2.data(), th);
            std::snprintf(name, sizeof name, "oracle_s2_complex_B%d_t%d", B, th);
            expect(name, rel(S2, S), 1e-12);
        }
    }
    {
        std::vector<double> d(or_so3_ncoef(41)), U(or_so3_ncoef(13));
        or_wigner_d_stable(1.3, 40, d.data());
        or_so3_real_harmonics(12, 0.4, 2.2, -1.0, U.data());
        std::vector<double> x = random_vec(2 * 48, 7), X(2 * 48);
        or_fft(48, x.data(), X.data());
        double e = 0;   // Parseval of the unnormalised DFT
        double sx = 0, sX = 0;
        for (int i = 0; i < 96; i++) { sx += x[i] * x[i]; sX += X[i] * X[i]; }
        e = std::fabs(sX / 48 - sx) / sx;
        expect("oracle_fft48_parseval", e, 1e-13);
    }
    // the sampling step: every value written once, by whichever thread, in FP64 and FP32
    for (int fp64 : {1, 0}) {
        for (long n : {1L, 1023L, 1024L, 5000L, 70001L}) {
            const int w = 2;
            const long piece = 1024, c0 = 777;
            std::vector<double> out64(fp64 ? n * w : 0);
            std::vector<float> out32(fp64 ? 0 : n * w);
            auto fill = [&](long i0, long i1, double *o) {
                for (long i = i0; i < i1; i++) {
                    o[(i - i0) * w] = std::sin(0.001 * i);
                    o[(i - i0) * w + 1] = (double)i;
                }
            };
            fdcl::detail::sample_chunk(fill, c0, n, piece, w, fp64 != 0,
                                       fp64 ? (void *)out64.data() : (void *)out32.data());
            double e = 0;
            for (long i = 0; i < n; i++) {
                const double re = fp64 ? out64[i * w] : out32[i * w], im = fp64 ? out64[i * w + 1] : out32[i * w + 1];
                e = std::max(e, std::fabs(re - std::sin(0.001 * (c0 + i))) + std::fabs(im - (double)(c0 + i)) / (c0 + i));
            }
            char name[64];
            std::snprintf(name, sizeof name, "sample_chunk_%s_n%ld", fp64 ? "fp64" : "fp32", n);
            expect(name, e, fp64 ? 0.0 : 1e-6);
        }
    }
    return failures ? 1 : 0;
}
