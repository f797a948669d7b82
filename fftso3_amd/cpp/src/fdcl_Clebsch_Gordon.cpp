// fdcl_Clebsch_Gordon.cpp -- Clebsch-Gordan coefficients of the complex and the real harmonics
// (host, FP64).  Interface: the reference's include/fdcl_Clebsch_Gordon.hpp; values: the
// Condon-Shortley convention, pinned against sympy's CG in the tests.
#include "fdcl_Clebsch_Gordon.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace fdcl
{

Clebsch_Gordon_complex::Clebsch_Gordon_complex(int l1_, int l2_) { init(l1_, l2_); }

void Clebsch_Gordon_complex::init(int l1_, int l2_)
{
    if (l1_ < 0 || l2_ < 0) throw std::invalid_argument("Clebsch_Gordon: negative degree");
    l1 = l1_;
    l2 = l2_;
    const int n = (2 * l1 + 1) * (2 * l2 + 1);
    C.resize(n, n);
}

void Clebsch_Gordon_complex::check_index(int l, int m, int l1_, int m1, int l2_, int m2) const
{
    if (l1_ != l1 || l2_ != l2 || l < std::abs(l1 - l2) || l > l1 + l2 || std::abs(m) > l || std::abs(m1) > l1 ||
        std::abs(m2) > l2)
        throw std::out_of_range("Clebsch_Gordon: (l,m,l1,m1,l2,m2) = (" + std::to_string(l) + "," + std::to_string(m) +
                                "," + std::to_string(l1_) + "," + std::to_string(m1) + "," + std::to_string(l2_) + "," +
                                std::to_string(m2) + ") outside the computed set (l1,l2) = (" + std::to_string(l1) +
                                "," + std::to_string(l2) + ")");
}

double &Clebsch_Gordon_complex::operator()(int l, int m, int l1_, int m1, int l2_, int m2)
{
    check_index(l, m, l1_, m1, l2_, m2);
    return C(row(l1, m1, l2, m2), col(l, m, l1, l2));
}

// For fixed (l, m), x(m1) = <l1 m1 l2 m-m1 | l m> is the eigenvector of
// J^2 = J1^2 + J2^2 + 2 J1z J2z + J1+ J2- + J1- J2+ for l(l+1) in the product basis, a tridiagonal
// problem in m1 (the approach of Straub, "Efficient computation of Clebsch-Gordan coefficients",
// which src/fdcl_Clebsch_Gordon.cpp:42-86 also follows).  Recurse downward from the largest m1,
// normalise, and give the largest-m1 coefficient a positive sign (Condon-Shortley).
void Clebsch_Gordon_complex::compute(int l1_, int l2_)
{
    init(l1_, l2_);
    const double j1 = l1 * (l1 + 1.0), j2 = l2 * (l2 + 1.0);
    std::vector<double> x;
    for (int l = std::abs(l1 - l2); l <= l1 + l2; l++) {
        const double jl = l * (l + 1.0);
        for (int m = -l; m <= l; m++) {
            const int lo = std::max(-l1, m - l2), hi = std::min(l1, m + l2);
            const int n = hi - lo + 1;
            x.assign(n + 1, 0.0);   // x[i] = coefficient at m1 = lo + i; x[n] = 0 (above the top)
            // diagonal a(m1) and coupling b(m1) between m1 and m1 + 1
            auto a = [&](int m1) { return j1 + j2 + 2.0 * m1 * (m - m1); };
            auto b = [&](int m1) {
                return std::sqrt((j1 - (double)m1 * (m1 + 1)) * (j2 - (double)(m - m1) * (m - m1 - 1)));
            };
            x[n - 1] = 1.0;
            for (int i = n - 1; i >= 1; i--) {
                const int m1 = lo + i;
                x[i - 1] = ((jl - a(m1)) * x[i] - b(m1) * x[i + 1]) / b(m1 - 1);
            }
            double s = 0.0;
            for (int i = 0; i < n; i++) s += x[i] * x[i];
            s = 1.0 / std::sqrt(s);
            for (int i = 0; i < n; i++) C(row(l1, lo + i, l2, m - lo - i), col(l, m, l1, l2)) = x[i] * s;
        }
    }
}

Clebsch_Gordon_real::Clebsch_Gordon_real(int l1_, int l2_) { init(l1_, l2_); }

void Clebsch_Gordon_real::init(int l1_, int l2_)
{
    Clebsch_Gordon_complex::init(l1_, l2_);
    const int n = (2 * l1 + 1) * (2 * l2 + 1);
    c.resize(n, n);
}

std::complex<double> &Clebsch_Gordon_real::operator()(int l, int m, int l1_, int m1, int l2_, int m2)
{
    check_index(l, m, l1_, m1, l2_, m2);
    return c(row(l1, m1, l2, m2), col(l, m, l1, l2));
}

namespace
{
// nonzeros of row m of T^l (src/fdcl_FFTSO3.cpp:1107-1139): (column, value) pairs
int t_row(int m, std::complex<double> *val, int *colm)
{
    const double h = 1.0 / std::sqrt(2.0), sm = (m & 1) ? -1.0 : 1.0;
    const std::complex<double> I(0.0, 1.0);
    if (m == 0) {
        colm[0] = 0;
        val[0] = 1.0;
        return 1;
    }
    if (m < 0) {
        colm[0] = m;
        val[0] = I * h;
        colm[1] = -m;
        val[1] = -I * h * sm;
    } else {
        colm[0] = -m;
        val[0] = h;
        colm[1] = m;
        val[1] = sm * h;
    }
    return 2;
}
}  // namespace

// c^{l,m}_{l1,m1,l2,m2} = sum_{p1,p2,p} conj(T^{l1}_{m1 p1}) conj(T^{l2}_{m2 p2}) C^{l,p}_{l1,p1,l2,p2} T^l_{m p}
// with p = p1 + p2; T has at most two nonzeros per row, so each term is a short sum.
void Clebsch_Gordon_real::compute(int l1_, int l2_)
{
    Clebsch_Gordon_complex::compute(l1_, l2_);
    const int n = (2 * l1 + 1) * (2 * l2 + 1);
    c.resize(n, n);
    std::complex<double> v1[2], v2[2], v[2];
    int p1s[2], p2s[2], ps[2];
    for (int m1 = -l1; m1 <= l1; m1++) {
        const int n1 = t_row(m1, v1, p1s);
        for (int m2 = -l2; m2 <= l2; m2++) {
            const int n2 = t_row(m2, v2, p2s);
            for (int l = std::abs(l1 - l2); l <= l1 + l2; l++)
                for (int m = -l; m <= l; m++) {
                    const int nn = t_row(m, v, ps);
                    std::complex<double> s = 0.0;
                    for (int a = 0; a < n1; a++)
                        for (int b = 0; b < n2; b++)
                            for (int q = 0; q < nn; q++)
                                if (p1s[a] + p2s[b] == ps[q])
                                    s += std::conj(v1[a]) * std::conj(v2[b]) *
                                         C(row(l1, p1s[a], l2, p2s[b]), col(l, ps[q], l1, l2)) * v[q];
                    c(row(l1, m1, l2, m2), col(l, m, l1, l2)) = s;
                }
        }
    }
}

}  // namespace fdcl
