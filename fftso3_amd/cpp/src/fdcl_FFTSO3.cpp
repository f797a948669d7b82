// fdcl_FFTSO3.cpp -- host side of fdcl::FFTSO3_complex / fdcl::FFTSO3_real: FP64 Wigner functions,
// quadrature weights, pointwise inverses and self checks; transforms go to the GPU.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iostream>

#include "fdcl_FFTSO3.hpp"

namespace fdcl
{

// ---------------------------------------------------------------------------------------------
// Wigner small-d for every l <= L at one beta.
//
// The reference seeds each (m, n) column with closed forms and runs recursion (64)
// (src/fdcl_FFTSO3.cpp:127-194).  Here every group (m0, n0), m0 >= |n0|, starts at l = m0 from
//   d^{m0}_{m0 n0}(beta) = sqrt((2 m0)! / ((m0+n0)! (m0-n0)!)) cos(beta/2)^(m0+n0) (-sin(beta/2))^(m0-n0)
// evaluated in logarithms (no overflow or underflow of the factorials for large l), then the
// three-term recursion in l
//   (l-1) sqrt((l^2-m^2)(l^2-n^2)) d^l = (2l-1)(l(l-1) cos beta - m n) d^{l-1}
//                                        - l sqrt(((l-1)^2-m^2)((l-1)^2-n^2)) d^{l-2},
// and the other three members of the group follow from the symmetries of src/fdcl_FFTSO3.cpp:174-190:
//   d(l,m,n) = (-1)^(m+n) d(l,-m,-n) = d(l,-n,-m) = (-1)^(m+n) d(l,n,m).
// ---------------------------------------------------------------------------------------------
static void wigner_d_all(double beta, int L, FFTSO3_matrix_real &d)
{
    d.init(L);
    const double cb = std::cos(beta), ch = std::cos(0.5 * beta), sh = std::sin(0.5 * beta);
    const double lc = std::log(std::fabs(ch)), ls = std::log(std::fabs(sh));
    for (int m0 = 0; m0 <= L; m0++) {
        for (int n0 = -m0; n0 <= m0; n0++) {
            const int a = m0 + n0, b = m0 - n0;
            // seed magnitude and sign; x^0 = 1 also at x = 0
            double lg = 0.5 * (std::lgamma(2.0 * m0 + 1) - std::lgamma(a + 1.0) - std::lgamma(b + 1.0));
            if (a) lg += a * lc;
            if (b) lg += b * ls;
            double v = (a && ch == 0.0) || (b && sh == 0.0) ? 0.0 : std::exp(lg);
            if ((a & 1) && ch < 0) v = -v;
            if ((b & 1) && sh > 0) v = -v;  // (-sin(beta/2))^b
            double vm1 = 0.0;
            const double mn = (double)m0 * n0;
            const double sgn = ((m0 + n0) & 1) ? -1.0 : 1.0;
            for (int l = m0; l <= L; l++) {
                if (l > m0) {
                    double nv;
                    if (l == 1) {
                        nv = cb;  // d^1_00
                    } else {
                        const double l1 = l - 1.0;
                        const double c1 = (2.0 * l - 1.0) * (l * l1 * cb - mn);
                        const double c2 = l * std::sqrt((l1 * l1 - (double)m0 * m0) * (l1 * l1 - (double)n0 * n0));
                        const double den = l1 * std::sqrt(((double)l * l - (double)m0 * m0) * ((double)l * l - (double)n0 * n0));
                        nv = (c1 * v - c2 * vm1) / den;
                    }
                    vm1 = v;
                    v = nv;
                }
                d(l, m0, n0) = v;
                d(l, -m0, -n0) = sgn * v;
                d(l, -n0, -m0) = v;
                d(l, n0, m0) = sgn * v;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// FFTSO3_complex
// ---------------------------------------------------------------------------------------------
FFTSO3_complex::FFTSO3_complex(int l_max_) { init(l_max_); }

void FFTSO3_complex::init(int l_max_)
{
    if (l_max_ == l_max && (int)weight.size() == 2 * (l_max_ + 1)) return;
    l_max = l_max_;
    B = l_max + 1;
    weight.assign(2 * B, 0.0);
    compute_weight();
}

double FFTSO3_complex::alpha_j(int j) const { return (double)j * M_PI / (double)B; }
double FFTSO3_complex::gamma_j(int j) const { return (double)j * M_PI / (double)B; }
double FFTSO3_complex::beta_k(int k) const { return (double)(2 * k + 1) * M_PI / (4.0 * B); }

// w_k = sin(beta_k) / (4 B^3) sum_{i<B} sin((2k+1)(2i+1) pi / 4B) / (2i+1) (src/fdcl_FFTSO3.cpp:236-256)
std::vector<double> FFTSO3_complex::compute_weight()
{
    const double f = M_PI / (4.0 * B);
    weight.assign(2 * B, 0.0);
    for (int k = 0; k < 2 * B; k++) {
        double s = 0.0;
        for (int i = 0; i < B; i++) s += std::sin((double)((2 * k + 1) * (2 * i + 1)) * f) / (2 * i + 1);
        weight[k] = s * std::sin((2 * k + 1) * f) / (4.0 * B * B * B);
    }
    return weight;
}

FFTSO3_matrix_real FFTSO3_complex::wigner_d(double beta, int L)
{
    FFTSO3_matrix_real d;
    wigner_d_all(beta, L, d);
    return d;
}

FFTSO3_matrix_real FFTSO3_complex::wigner_d(double beta) { return wigner_d(beta, l_max); }

FFTSO3_matrix_complex FFTSO3_complex::wigner_D(double alpha, double beta, double gamma, int L)
{
    FFTSO3_matrix_real d = wigner_d(beta, L);
    FFTSO3_matrix_complex D(L);
    for (int l = 0; l <= L; l++)
        for (int m = -l; m <= l; m++)
            for (int n = -l; n <= l; n++)
                D(l, m, n) = std::exp(-I * ((double)m * alpha + (double)n * gamma)) * d(l, m, n);
    return D;
}

FFTSO3_matrix_complex FFTSO3_complex::wigner_D(double alpha, double beta, double gamma)
{
    return wigner_D(alpha, beta, gamma, l_max);
}

FFTSO3_matrix_complex FFTSO3_complex::wigner_D(const Matrix3d &R)
{
    const std::vector<double> abg = R2Euler323(R);
    return wigner_D(abg[0], abg[1], abg[2], l_max);
}

std::vector<double> FFTSO3_complex::character(double theta)
{
    std::vector<double> chi(l_max + 1);
    for (int l = 0; l <= l_max; l++) chi[l] = std::sin((2.0 * l + 1) / 2. * theta) / std::sin(theta / 2.);
    return chi;
}

// f is sampled on the host by OpenMP threads (as src/fdcl_FFTSO3.cpp:675-685 does), chunk by chunk
// into pinned buffers whose copies to the device overlap the sampling of the next chunk
FFTSO3_matrix_complex FFTSO3_complex::forward_transform(std::function<complex<double>(double, double, double)> func)
{
    const long N = 2 * B;
    FFTSO3_matrix_complex F(l_max);
    dev_->forward_fill(FFTSO3_DOMAIN_SO3, l_max, FFTSO3_COMPLEX, precision_, [&](long i0, long i1, double *out) {
        for (long i = i0; i < i1; i++) {
            const long k = i / (N * N), j1 = (i / N) % N, j2 = i % N;
            const complex<double> v = func(alpha_j((int)j1), beta_k((int)k), gamma_j((int)j2));
            out[2 * (i - i0)] = v.real();
            out[2 * (i - i0) + 1] = v.imag();
        }
    }, F.v.data());
    return F;
}

FFTSO3_matrix_complex FFTSO3_complex::forward_transform(std::function<complex<double>(Matrix3d)> func)
{
    return forward_transform([&](double a, double b, double g) { return func(Euler3232R(a, b, g)); });
}

FFTSO3_matrix_complex FFTSO3_complex::forward_transform(const std::vector<complex<double>> &grid)
{
    return forward_transform_batch(grid, 1)[0];
}

std::vector<FFTSO3_matrix_complex> FFTSO3_complex::forward_transform_batch(const std::vector<complex<double>> &grids,
                                                                           int n_fn)
{
    const long N = 2 * B, ng = N * N * N;
    if (n_fn < 1 || (long)grids.size() != ng * n_fn)
        throw std::invalid_argument("forward_transform_batch: expected n_fn grids of (2B)^3 values");
    std::vector<complex<double>> coef(FFTSO3_matrix_complex::size_for(l_max) * n_fn);
    dev_->forward(FFTSO3_DOMAIN_SO3, l_max, n_fn, FFTSO3_COMPLEX, precision_, grids.data(), coef.data());
    std::vector<FFTSO3_matrix_complex> out(n_fn);
    const long nc = FFTSO3_matrix_complex::size_for(l_max);
    for (int i = 0; i < n_fn; i++) {
        out[i].l_max = l_max;
        out[i].v.assign(coef.begin() + i * nc, coef.begin() + (i + 1) * nc);
    }
    return out;
}

complex<double> FFTSO3_complex::inverse_transform(FFTSO3_matrix_complex F, double alpha, double beta, double gamma)
{
    init(F.l_max);
    const FFTSO3_matrix_real d = wigner_d(beta, l_max);
    // f = sum_{m,n} exp(-i m alpha) [sum_l (2l+1) F d] exp(-i n gamma)
    complex<double> f = 0.0;
    for (int m = -l_max; m <= l_max; m++) {
        const complex<double> em = std::exp(-I * ((double)m * alpha));
        for (int n = -l_max; n <= l_max; n++) {
            complex<double> t = 0.0;
            for (int l = std::max(std::abs(m), std::abs(n)); l <= l_max; l++)
                t += (double)(2 * l + 1) * F(l, m, n) * d(l, m, n);
            f += em * t * std::exp(-I * ((double)n * gamma));
        }
    }
    return f;
}

complex<double> FFTSO3_complex::inverse_transform(FFTSO3_matrix_complex F, const Matrix3d &R)
{
    const std::vector<double> abg = R2Euler323(R);
    return inverse_transform(std::move(F), abg[0], abg[1], abg[2]);
}

std::vector<complex<double>> FFTSO3_complex::inverse_transform(const FFTSO3_matrix_complex &F)
{
    return inverse_transform_batch(std::vector<FFTSO3_matrix_complex>{F});
}

std::vector<complex<double>> FFTSO3_complex::inverse_transform_batch(const std::vector<FFTSO3_matrix_complex> &F)
{
    if (F.empty()) return {};
    init(F[0].l_max);
    const long nc = FFTSO3_matrix_complex::size_for(l_max), N = 2 * B;
    std::vector<complex<double>> coef(nc * F.size());
    for (size_t i = 0; i < F.size(); i++) {
        if (F[i].l_max != l_max) throw std::invalid_argument("inverse_transform_batch: mixed l_max");
        std::copy(F[i].v.begin(), F[i].v.end(), coef.begin() + i * nc);
    }
    std::vector<complex<double>> grid(N * N * N * F.size());
    dev_->inverse(FFTSO3_DOMAIN_SO3, l_max, (int)F.size(), FFTSO3_COMPLEX, precision_, coef.data(), grid.data());
    return grid;
}

std::vector<complex<double>> FFTSO3_complex::inverse_transform_points(const FFTSO3_matrix_complex &F,
                                                                      const std::vector<double> &angles)
{
    if (angles.size() % 3) throw std::invalid_argument("inverse_transform_points: 3 angles per point");
    init(F.l_max);
    const long n = (long)angles.size() / 3;
    std::vector<complex<double>> v(n);
    dev_->evaluate(FFTSO3_DOMAIN_SO3, l_max, FFTSO3_COMPLEX, precision_, F.data(), angles.data(), n, v.data());
    return v;
}

// sum_k w_k d^l_00(beta_k) 4B^2 = delta_{l0} for l < 2B (src/fdcl_FFTSO3.cpp:258-288)
double FFTSO3_complex::check_weight()
{
    compute_weight();
    std::vector<double> sum(2 * B, 0.0);
    for (int k = 0; k < 2 * B; k++) {
        const FFTSO3_matrix_real d = wigner_d(beta_k(k), 2 * B - 1);
        for (int l = 0; l < 2 * B; l++) sum[l] += d(l, 0, 0) * weight[k];
    }
    double err = std::fabs(sum[0] * 4.0 * B * B - 1.0);
    for (int l = 1; l < 2 * B; l++) err += std::fabs(sum[l] * 4.0 * B * B);
    if (check_verbose)
        for (int l = 0; l < 2 * B; l++) std::cout << "l=" << l << ": " << sum[l] * 4.0 * B * B << std::endl;
    std::cout << "fdcl::FFTSO3_complex::check_weight: error = " << err << std::endl;
    return err;
}

// Orthogonality d^l (d^l)^T = I at a random beta, and the closed forms of degree 1
// (src/fdcl_FFTSO3.cpp:16-120, 320-392)
double FFTSO3_complex::check_wigner_d()
{
    const int L = std::max(l_max, 1);
    const double beta = (double)std::rand() / RAND_MAX * M_PI;
    const FFTSO3_matrix_real d = wigner_d(beta, L);
    double e_orth = 0.0;
    for (int l = 0; l <= L; l++) {
        const auto b = d[l];
        for (int i = 0; i < 2 * l + 1; i++)
            for (int j = 0; j < 2 * l + 1; j++) {
                double s = 0.0;
                for (int q = 0; q < 2 * l + 1; q++) s += b(q, i) * b(q, j);
                e_orth = std::max(e_orth, std::fabs(s - (i == j ? 1.0 : 0.0)));
            }
    }
    const double c = std::cos(beta), s = std::sin(beta), r2 = std::sqrt(2.0);
    const double d1[3][3] = {{(1 + c) / 2, s / r2, (1 - c) / 2},     // m = -1
                             {-s / r2, c, s / r2},                  // m = 0
                             {(1 - c) / 2, -s / r2, (1 + c) / 2}};  // m = 1
    double e_exp = std::fabs(d(0, 0, 0) - 1.0);
    for (int m = -1; m <= 1; m++)
        for (int n = -1; n <= 1; n++) e_exp = std::max(e_exp, std::fabs(d(1, m, n) - d1[m + 1][n + 1]));
    const double err = std::max(e_orth, e_exp);
    if (check_verbose) std::cout << "orthogonality " << e_orth << ", closed form l=1 " << e_exp << std::endl;
    std::cout << "fdcl::FFTSO3_complex::check_wigner_d: error = " << err << std::endl;
    return err;
}

// f = inverse_transform(F) for random F, then forward_transform(f) must return F
// (src/fdcl_FFTSO3.cpp:532-545)
double FFTSO3_complex::check_transform()
{
    FFTSO3_matrix_complex F(l_max);
    F.setRandom();
    const FFTSO3_matrix_complex Fc = F;
    FFTSO3_matrix_complex F2 = forward_transform(
        [&](double a, double b, double g) { return inverse_transform(Fc, a, b, g); });
    const double err = (F2 - F).norm();
    std::cout << "fdcl::FFTSO3_complex::check_transform: error = " << err << std::endl;
    return err;
}

// u[i] = d/dt D(exp(t e_i)) at t = 0 = -i J_i in the |l m> basis: J_z = m, J_x = (J+ + J-)/2,
// J_y = (J+ - J-)/(2i) with <m+1|J+|m> = sqrt((l-m)(l+m+1)), <m-1|J-|m> = sqrt((l+m)(l-m+1))
std::vector<FFTSO3_matrix_complex> FFTSO3_complex::deriv_wigner_D()
{
    std::vector<FFTSO3_matrix_complex> u(4);
    for (int i = 1; i <= 3; i++) u[i].init(l_max);
    for (int l = 0; l <= l_max; l++)
        for (int n = -l; n <= l; n++) {
            u[3](l, n, n) = -I * (double)n;
            if (n < l) {   // <n+1| J+ |n>
                const double jp = std::sqrt((double)(l - n) * (l + n + 1));
                u[1](l, n + 1, n) = -0.5 * I * jp;
                u[2](l, n + 1, n) = -0.5 * jp;
            }
            if (n > -l) {  // <n-1| J- |n>
                const double jm = std::sqrt((double)(l + n) * (l - n + 1));
                u[1](l, n - 1, n) = -0.5 * I * jm;
                u[2](l, n - 1, n) = 0.5 * jm;
            }
        }
    return u;
}

namespace
{
// R = exp(hat(r)) by Rodrigues' formula
Matrix3d expm_so3(const double r[3])
{
    const double th = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    const double a = th > 0 ? std::sin(th) / th : 1.0, b = th > 0 ? (1.0 - std::cos(th)) / (th * th) : 0.5;
    const double K[3][3] = {{0, -r[2], r[1]}, {r[2], 0, -r[0]}, {-r[1], r[0], 0}};
    Matrix3d R;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double k2 = 0;
            for (int q = 0; q < 3; q++) k2 += K[i][q] * K[q][j];
            R(i, j) = (i == j ? 1.0 : 0.0) + a * K[i][j] + b * k2;
        }
    return R;
}

// exp(A) of an n x n column-major matrix: scaling and squaring of a 24-term Taylor sum
template <typename T> std::vector<T> expm(const std::vector<T> &A, int n)
{
    double nrm = 0;
    for (const T &x : A) nrm = std::max(nrm, std::abs(x));
    int sq = 0;
    while (nrm * n > 0.5) {
        nrm *= 0.5;
        sq++;
    }
    const double sc = std::ldexp(1.0, -sq);
    std::vector<T> X(A.size()), E(A.size(), T(0)), P(A.size(), T(0)), Q(A.size());
    for (size_t i = 0; i < A.size(); i++) X[i] = A[i] * sc;
    for (int i = 0; i < n; i++) E[i * n + i] = P[i * n + i] = T(1);
    for (int k = 1; k <= 24; k++) {
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                T t = 0;
                for (int q = 0; q < n; q++) t += P[q * n + i] * X[j * n + q];
                Q[j * n + i] = t / (double)k;
            }
        P.swap(Q);
        for (size_t i = 0; i < E.size(); i++) E[i] += P[i];
    }
    for (int r = 0; r < sq; r++) {
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                T t = 0;
                for (int q = 0; q < n; q++) t += E[q * n + i] * E[j * n + q];
                Q[j * n + i] = t;
            }
        E.swap(Q);
    }
    return E;
}
}  // namespace

// D is a representation, so D(exp(t e_i)) = exp(t u[i]) exactly; compared at t = 0.4 for every
// l <= l_max (relative Frobenius error, largest over l and i)
double FFTSO3_complex::check_deriv_wigner_D()
{
    const std::vector<FFTSO3_matrix_complex> u = deriv_wigner_D();
    const double t = 0.4;
    double err = 0.0;
    for (int i = 1; i <= 3; i++) {
        double r[3] = {0, 0, 0};
        r[i - 1] = t;
        const std::vector<double> abg = R2Euler323(expm_so3(r));
        const FFTSO3_matrix_complex D = wigner_D(abg[0], abg[1], abg[2], l_max);
        for (int l = 1; l <= l_max; l++) {
            const int n = 2 * l + 1;
            std::vector<complex<double>> A(n * n);
            for (int a = 0; a < n; a++)
                for (int b = 0; b < n; b++) A[b * n + a] = t * u[i](l, a - l, b - l);
            const std::vector<complex<double>> E = expm(A, n);
            double e = 0, d = 0;
            for (int a = 0; a < n; a++)
                for (int b = 0; b < n; b++) {
                    e += std::norm(E[b * n + a] - D(l, a - l, b - l));
                    d += std::norm(D(l, a - l, b - l));
                }
            err = std::max(err, std::sqrt(e / d));
        }
    }
    std::cout << "fdcl::FFTSO3_complex::check_deriv_wigner_D: error = " << err << std::endl;
    return err;
}

double FFTSO3_complex::check_Clebsch_Gordon()
{
    const int l1 = 2, l2 = 5;
    const double a = (double)std::rand() / RAND_MAX * 2 * M_PI, b = (double)std::rand() / RAND_MAX * M_PI,
                 g = (double)std::rand() / RAND_MAX * 2 * M_PI;
    const FFTSO3_matrix_complex D = wigner_D(a, b, g, l1 + l2);
    C.compute(l1, l2);
    double err = 0.0;
    for (int m1 = -l1; m1 <= l1; m1++)
        for (int n1 = -l1; n1 <= l1; n1++)
            for (int m2 = -l2; m2 <= l2; m2++)
                for (int n2 = -l2; n2 <= l2; n2++) {
                    complex<double> y = 0.0;
                    const int lo = std::max(std::abs(l1 - l2), std::max(std::abs(m1 + m2), std::abs(n1 + n2)));
                    for (int l = lo; l <= l1 + l2; l++)
                        y += C(l, m1 + m2, l1, m1, l2, m2) * C(l, n1 + n2, l1, n1, l2, n2) * D(l, m1 + m2, n1 + n2);
                    err = std::max(err, std::abs(D(l1, m1, n1) * D(l2, m2, n2) - y));
                }
    std::cout << "fdcl::FFTSO3_complex::check_Clebsch_Gordon: error = " << err << std::endl;
    return err;
}

void FFTSO3_complex::check_all()
{
    check_weight();
    check_wigner_d();
    check_deriv_wigner_D();
    check_Clebsch_Gordon();
    check_transform();
}

// ---------------------------------------------------------------------------------------------
// FFTSO3_real
// ---------------------------------------------------------------------------------------------
FFTSO3_real::FFTSO3_real(int l_max_) { init(l_max_); }

void FFTSO3_real::init(int l_max_) { FFTSO3_complex::init(l_max_); }

// T^l: row m < 0 has i/sqrt2 at n = m and -i(-1)^m/sqrt2 at n = -m; row 0 is e_0; row m > 0 has
// 1/sqrt2 at n = -m and (-1)^m/sqrt2 at n = m (src/fdcl_FFTSO3.cpp:1107-1139)
FFTSO3_matrix_complex FFTSO3_real::matrix2rsph(int L)
{
    FFTSO3_matrix_complex T(L);
    const double h = 1.0 / std::sqrt(2.0);
    for (int l = 0; l <= L; l++) {
        for (int m = -l; m < 0; m++) {
            const double sm = (m & 1) ? -1.0 : 1.0;
            T(l, m, m) = I * h;
            T(l, m, -m) = -I * h * sm;
        }
        T(l, 0, 0) = 1.0;
        for (int m = 1; m <= l; m++) {
            const double sm = (m & 1) ? -1.0 : 1.0;
            T(l, m, -m) = h;
            T(l, m, m) = sm * h;
        }
    }
    return T;
}

// U^l = conj(T^l) D^l (T^l)^T; T has two nonzeros per row, so each entry is a sum of four terms
FFTSO3_matrix_real FFTSO3_real::real_harmonics(double alpha, double beta, double gamma, int L)
{
    const FFTSO3_matrix_complex D = wigner_D(alpha, beta, gamma, L);
    const FFTSO3_matrix_complex T = matrix2rsph(L);
    FFTSO3_matrix_real U(L);
    for (int l = 0; l <= L; l++)
        for (int m = -l; m <= l; m++)
            for (int n = -l; n <= l; n++) {
                complex<double> u = 0.0;
                const int as[2] = {m, -m}, bs[2] = {n, -n};
                for (int ia = 0; ia < (m ? 2 : 1); ia++)
                    for (int ib = 0; ib < (n ? 2 : 1); ib++)
                        u += std::conj(T(l, m, as[ia])) * D(l, as[ia], bs[ib]) * T(l, n, bs[ib]);
                U(l, m, n) = u.real();
            }
    return U;
}

FFTSO3_matrix_real FFTSO3_real::real_harmonics(double alpha, double beta, double gamma)
{
    return real_harmonics(alpha, beta, gamma, l_max);
}

FFTSO3_matrix_real FFTSO3_real::real_harmonics(const Matrix3d &R)
{
    const std::vector<double> abg = R2Euler323(R);
    return real_harmonics(abg[0], abg[1], abg[2], l_max);
}

FFTSO3_matrix_real FFTSO3_real::forward_transform(std::function<double(double, double, double)> func)
{
    const long N = 2 * B;
    FFTSO3_matrix_real F(l_max);
    dev_->forward_fill(FFTSO3_DOMAIN_SO3, l_max, FFTSO3_REAL, precision_, [&](long i0, long i1, double *out) {
        for (long i = i0; i < i1; i++) {
            const long k = i / (N * N), j1 = (i / N) % N, j2 = i % N;
            out[i - i0] = func(alpha_j((int)j1), beta_k((int)k), gamma_j((int)j2));
        }
    }, F.v.data());
    return F;
}

FFTSO3_matrix_real FFTSO3_real::forward_transform(std::function<double(Matrix3d)> func)
{
    return forward_transform([&](double a, double b, double g) { return func(Euler3232R(a, b, g)); });
}

FFTSO3_matrix_real FFTSO3_real::forward_transform(const std::vector<double> &grid)
{
    return forward_transform_batch(grid, 1)[0];
}

std::vector<FFTSO3_matrix_real> FFTSO3_real::forward_transform_batch(const std::vector<double> &grids, int n_fn)
{
    const long N = 2 * B, ng = N * N * N;
    if (n_fn < 1 || (long)grids.size() != ng * n_fn)
        throw std::invalid_argument("forward_transform_batch: expected n_fn grids of (2B)^3 values");
    const long nc = FFTSO3_matrix_real::size_for(l_max);
    std::vector<double> coef(nc * n_fn);
    dev_->forward(FFTSO3_DOMAIN_SO3, l_max, n_fn, FFTSO3_REAL, precision_, grids.data(), coef.data());
    std::vector<FFTSO3_matrix_real> out(n_fn);
    for (int i = 0; i < n_fn; i++) {
        out[i].l_max = l_max;
        out[i].v.assign(coef.begin() + i * nc, coef.begin() + (i + 1) * nc);
    }
    return out;
}

double FFTSO3_real::inverse_transform(FFTSO3_matrix_real F, double alpha, double beta, double gamma)
{
    const FFTSO3_matrix_real U = real_harmonics(alpha, beta, gamma, l_max);
    double f = 0.0;
    for (int l = 0; l <= l_max; l++)
        for (int m = -l; m <= l; m++)
            for (int n = -l; n <= l; n++) f += (double)(2 * l + 1) * F(l, m, n) * U(l, m, n);
    return f;
}

double FFTSO3_real::inverse_transform(FFTSO3_matrix_real F, const Matrix3d &R)
{
    const std::vector<double> abg = R2Euler323(R);
    return inverse_transform(std::move(F), abg[0], abg[1], abg[2]);
}

std::vector<double> FFTSO3_real::inverse_transform(const FFTSO3_matrix_real &F)
{
    return inverse_transform_batch(std::vector<FFTSO3_matrix_real>{F});
}

std::vector<double> FFTSO3_real::inverse_transform_batch(const std::vector<FFTSO3_matrix_real> &F)
{
    if (F.empty()) return {};
    init(F[0].l_max);
    const long nc = FFTSO3_matrix_real::size_for(l_max), N = 2 * B;
    std::vector<double> coef(nc * F.size());
    for (size_t i = 0; i < F.size(); i++) {
        if (F[i].l_max != l_max) throw std::invalid_argument("inverse_transform_batch: mixed l_max");
        std::copy(F[i].v.begin(), F[i].v.end(), coef.begin() + i * nc);
    }
    std::vector<double> grid(N * N * N * F.size());
    dev_->inverse(FFTSO3_DOMAIN_SO3, l_max, (int)F.size(), FFTSO3_REAL, precision_, coef.data(), grid.data());
    return grid;
}

std::vector<double> FFTSO3_real::inverse_transform_points(const FFTSO3_matrix_real &F, const std::vector<double> &angles)
{
    if (angles.size() % 3) throw std::invalid_argument("inverse_transform_points: 3 angles per point");
    init(F.l_max);
    const long n = (long)angles.size() / 3;
    std::vector<double> v(n);
    dev_->evaluate(FFTSO3_DOMAIN_SO3, l_max, FFTSO3_REAL, precision_, F.data(), angles.data(), n, v.data());
    return v;
}

// the real harmonics are real, orthogonal (U U^T = I) and agree with the explicit d-based form of
// src/fdcl_FFTSO3.cpp:1062-1105 at degree 1
double FFTSO3_real::check_real_harmonics()
{
    const double a = (double)std::rand() / RAND_MAX * 2 * M_PI, b = (double)std::rand() / RAND_MAX * M_PI,
                 g = (double)std::rand() / RAND_MAX * 2 * M_PI;
    const FFTSO3_matrix_real U = real_harmonics(a, b, g, std::max(l_max, 1));
    double err = 0.0;
    for (int l = 0; l <= U.l_max; l++) {
        const auto u = U[l];
        for (int i = 0; i < 2 * l + 1; i++)
            for (int j = 0; j < 2 * l + 1; j++) {
                double s = 0.0;
                for (int q = 0; q < 2 * l + 1; q++) s += u(i, q) * u(j, q);
                err = std::max(err, std::fabs(s - (i == j ? 1.0 : 0.0)));
            }
    }
    const FFTSO3_matrix_real d = wigner_d(b, 1);
    const double r2 = std::sqrt(2.0);
    // U(1,1,0) = (-1) sqrt2 d(1,1,0) cos(alpha), U(1,0,-1) = -sqrt2 d(1,0,-1) sin(gamma)
    err = std::max(err, std::fabs(U(1, 1, 0) + r2 * d(1, 1, 0) * std::cos(a)));
    err = std::max(err, std::fabs(U(1, 0, -1) + r2 * d(1, 0, -1) * std::sin(g)));
    err = std::max(err, std::fabs(U(0, 0, 0) - 1.0));
    std::cout << "fdcl::FFTSO3_real::check_real_harmonics: error = " << err << std::endl;
    return err;
}

double FFTSO3_real::check_transform()
{
    FFTSO3_matrix_real F(l_max);
    F.setRandom();
    const FFTSO3_matrix_real Fc = F;
    FFTSO3_matrix_real F2 = forward_transform(
        std::function<double(double, double, double)>([&](double a, double b, double g) {
            return inverse_transform(Fc, a, b, g);
        }));
    const double err = (F2 - F).norm();
    std::cout << "fdcl::FFTSO3_real::check_transform: error = " << err << std::endl;
    return err;
}

FFTS2_matrix_real FFTSO3_real::rotate(const FFTS2_matrix_real &F, const Matrix3d &R)
{
    const std::vector<double> abg = R2Euler323(R);
    const FFTSO3_matrix_real U = real_harmonics(abg[0], abg[1], abg[2], F.l_max);
    FFTS2_matrix_real G(F.l_max);
    for (int l = 0; l <= F.l_max; l++)
        for (int m = -l; m <= l; m++) {
            double s = 0.0;
            for (int n = -l; n <= l; n++) s += U(l, m, n) * F(l, n);
            G(l, m) = s;
        }
    return G;
}

static std::vector<double> euler_angles_of(const std::vector<Matrix3d> &R)
{
    std::vector<double> ang;
    ang.reserve(3 * R.size());
    for (const Matrix3d &r : R) {
        const std::vector<double> abg = R2Euler323(r);
        ang.insert(ang.end(), abg.begin(), abg.end());
    }
    return ang;
}

std::vector<FFTS2_matrix_real> FFTSO3_real::rotate_batch(const FFTS2_matrix_real &F, const std::vector<Matrix3d> &R)
{
    if (F.l_max != l_max) throw std::runtime_error("rotate_batch: F.l_max must equal the object's l_max");
    const std::vector<double> ang = euler_angles_of(R);
    const long nc = (long)(l_max + 1) * (l_max + 1);
    std::vector<double> G((size_t)nc * R.size());
    dev_->rotate_s2(l_max, FFTSO3_REAL, precision_, F.data(), ang.data(), (long)R.size(), G.data());
    std::vector<FFTS2_matrix_real> out(R.size(), FFTS2_matrix_real(l_max));
    for (size_t i = 0; i < R.size(); i++) std::copy(G.begin() + i * nc, G.begin() + (i + 1) * nc, out[i].v.begin());
    return out;
}

std::vector<FFTS2_matrix_complex> FFTSO3_complex::rotate_batch(const FFTS2_matrix_complex &F,
                                                              const std::vector<Matrix3d> &R)
{
    if (F.l_max != l_max) throw std::runtime_error("rotate_batch: F.l_max must equal the object's l_max");
    const std::vector<double> ang = euler_angles_of(R);
    const long nc = (long)(l_max + 1) * (l_max + 1);
    std::vector<complex<double>> G((size_t)nc * R.size());
    dev_->rotate_s2(l_max, FFTSO3_COMPLEX, precision_, F.data(), ang.data(), (long)R.size(), G.data());
    std::vector<FFTS2_matrix_complex> out(R.size(), FFTS2_matrix_complex(l_max));
    for (size_t i = 0; i < R.size(); i++) std::copy(G.begin() + i * nc, G.begin() + (i + 1) * nc, out[i].v.begin());
    return out;
}

std::vector<double> FFTSO3_real::correlation(const FFTS2_matrix_real &F, const FFTS2_matrix_real &G)
{
    const int L = l_max;
    FFTSO3_matrix_real H(L);
    for (int l = 0; l <= std::min(L, std::min(F.l_max, G.l_max)); l++)
        for (int m = -l; m <= l; m++)
            for (int n = -l; n <= l; n++) H(l, m, n) = G(l, m) * F(l, n) / (2.0 * l + 1.0);
    return inverse_transform(H);
}

std::vector<FFTSO3_matrix_real> FFTSO3_real::deriv_real_harmonics()
{
    const std::vector<FFTSO3_matrix_complex> uc = deriv_wigner_D();
    const FFTSO3_matrix_complex T = matrix2rsph(l_max);
    std::vector<FFTSO3_matrix_real> u(4);
    for (int i = 1; i <= 3; i++) {
        u[i].init(l_max);
        for (int l = 0; l <= l_max; l++)
            for (int m = -l; m <= l; m++)
                for (int n = -l; n <= l; n++) {
                    complex<double> s = 0.0;
                    for (int p = -l; p <= l; p++) {
                        if (T(l, m, p) == 0.0) continue;
                        for (int q = -l; q <= l; q++)
                            if (T(l, n, q) != 0.0) s += std::conj(T(l, m, p)) * uc[i](l, p, q) * T(l, n, q);
                    }
                    u[i](l, m, n) = s.real();   // the imaginary part vanishes
                }
    }
    return u;
}

double FFTSO3_real::check_deriv_real_harmonics()
{
    const std::vector<FFTSO3_matrix_real> u = deriv_real_harmonics();
    const double t = 0.4;
    double err = 0.0;
    for (int i = 1; i <= 3; i++) {
        double r[3] = {0, 0, 0};
        r[i - 1] = t;
        const std::vector<double> abg = R2Euler323(expm_so3(r));
        const FFTSO3_matrix_real U = real_harmonics(abg[0], abg[1], abg[2], l_max);
        for (int l = 1; l <= l_max; l++) {
            const int n = 2 * l + 1;
            std::vector<double> A(n * n);
            for (int a = 0; a < n; a++)
                for (int b = 0; b < n; b++) A[b * n + a] = t * u[i](l, a - l, b - l);
            const std::vector<double> E = expm(A, n);
            double e = 0, d = 0;
            for (int a = 0; a < n; a++)
                for (int b = 0; b < n; b++) {
                    e += std::pow(E[b * n + a] - U(l, a - l, b - l), 2);
                    d += std::pow(U(l, a - l, b - l), 2);
                }
            err = std::max(err, std::sqrt(e / d));
        }
    }
    std::cout << "fdcl::FFTSO3_real::check_deriv_real_harmonics: error = " << err << std::endl;
    return err;
}

double FFTSO3_real::check_Clebsch_Gordon()
{
    const int l1 = 3, l2 = 5;
    const double a = (double)std::rand() / RAND_MAX * 2 * M_PI, b = (double)std::rand() / RAND_MAX * M_PI,
                 g = (double)std::rand() / RAND_MAX * 2 * M_PI;
    const FFTSO3_matrix_real U = real_harmonics(a, b, g, l1 + l2);
    c.compute(l1, l2);
    double err = 0.0;
    for (int m1 = -l1; m1 <= l1; m1++)
        for (int n1 = -l1; n1 <= l1; n1++)
            for (int m2 = -l2; m2 <= l2; m2++)
                for (int n2 = -l2; n2 <= l2; n2++) {
                    // the real harmonics of order m1 and m2 couple to orders +-m1 +-m2 only
                    int M[4] = {m1 + m2, m1 - m2, -m1 + m2, -m1 - m2}, Nn[4] = {n1 + n2, n1 - n2, -n1 + n2, -n1 - n2};
                    std::sort(M, M + 4);
                    std::sort(Nn, Nn + 4);
                    const int nm = (int)(std::unique(M, M + 4) - M), nn = (int)(std::unique(Nn, Nn + 4) - Nn);
                    double y = 0.0;
                    for (int x = 0; x < nm; x++)
                        for (int z = 0; z < nn; z++)
                            for (int l = std::max(std::abs(l1 - l2), std::max(std::abs(M[x]), std::abs(Nn[z])));
                                 l <= l1 + l2; l++)
                                y += std::real(c(l, M[x], l1, m1, l2, m2) * std::conj(c(l, Nn[z], l1, n1, l2, n2))) *
                                     U(l, M[x], Nn[z]);
                    err = std::max(err, std::fabs(U(l1, m1, n1) * U(l2, m2, n2) - y));
                }
    std::cout << "fdcl::FFTSO3_real::check_Clebsch_Gordon: error = " << err << std::endl;
    return err;
}

void FFTSO3_real::check_all()
{
    check_weight();
    check_wigner_d();
    check_real_harmonics();
    check_Clebsch_Gordon();
    check_deriv_real_harmonics();
    check_transform();
}

}  // namespace fdcl
