// fdcl_FFTS2.cpp -- host side of fdcl::FFTS2_complex / fdcl::FFTS2_real: FP64 Legendre functions and
// spherical harmonics, weights, pointwise inverses and self checks; transforms go to the GPU.
#include <cmath>
#include <cstdlib>
#include <iostream>

#include "fdcl_FFTS2.hpp"

namespace fdcl
{

FFTS2_complex::FFTS2_complex(int l_max_) { init(l_max_); }

void FFTS2_complex::init(int l_max_)
{
    if (l_max_ == l_max && (int)weight.size() == 2 * (l_max_ + 1)) return;
    l_max = l_max_;
    B = l_max + 1;
    compute_weight();
}

double FFTS2_complex::theta_k(int k) const { return (double)(2 * k + 1) * M_PI / (4.0 * B); }
double FFTS2_complex::phi_j(int j) const { return (double)j * M_PI / (double)B; }

// w_k = 2 pi / B^2 sin(theta_k) sum_{i<B} sin((2k+1)(2i+1) pi / 4B) / (2i+1) (src/fdcl_FFTS2.cpp:102-123)
std::vector<double> FFTS2_complex::compute_weight()
{
    const double f = M_PI / (4.0 * B);
    weight.assign(2 * B, 0.0);
    for (int k = 0; k < 2 * B; k++) {
        double s = 0.0;
        for (int i = 0; i < B; i++) s += std::sin((double)((2 * k + 1) * (2 * i + 1)) * f) / (2 * i + 1);
        weight[k] = s * 2.0 * M_PI / ((double)B * B) * std::sin((2 * k + 1) * f);
    }
    return weight;
}

// nP^l_m(x) = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) P^l_m(x) with the Condon-Shortley phase:
// diagonal nP^m_m = (-1)^m sqrt((2m+1)/(4 pi) prod_{i<=m} (2i-1)/(2i)) (1-x^2)^(m/2), first
// off-diagonal nP^{m+1}_m = x sqrt(2m+3) nP^m_m, then
// nP^l_m = sqrt((4l^2-1)/(l^2-m^2)) (x nP^{l-1}_m - sqrt(((l-1)^2-m^2)/(4(l-1)^2-1)) nP^{l-2}_m),
// and nP^l_{-m} = (-1)^m nP^l_m (src/fdcl_FFTS2.cpp:48-100)
FFTS2_matrix_real FFTS2_complex::nor_assoc_Legendre_poly(double x, int L)
{
    FFTS2_matrix_real nP(L);
    const double omx2 = (1.0 - x) * (1.0 + x);
    double pmm = 1.0 / (4.0 * M_PI);  // squared magnitude of the diagonal, built up in m
    for (int m = 0; m <= L; m++) {
        if (m > 0) pmm *= omx2 * (2.0 * m - 1.0) / (2.0 * m);
        double d = std::sqrt((2.0 * m + 1.0) * pmm);
        if (m & 1) d = -d;
        nP(m, m) = d;
        if (m + 1 <= L) nP(m + 1, m) = x * std::sqrt(2.0 * m + 3.0) * d;
        for (int l = m + 2; l <= L; l++) {
            const double a = std::sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
            const double b = std::sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
            nP(l, m) = a * (x * nP(l - 1, m) - b * nP(l - 2, m));
        }
    }
    for (int l = 1; l <= L; l++)
        for (int m = 1; m <= l; m++) nP(l, -m) = ((m & 1) ? -1.0 : 1.0) * nP(l, m);
    return nP;
}

FFTS2_matrix_complex FFTS2_complex::spherical_harmonics(double theta, double phi, int L)
{
    const FFTS2_matrix_real nP = nor_assoc_Legendre_poly(std::cos(theta), L);
    FFTS2_matrix_complex Y(L);
    for (int l = 0; l <= L; l++) {
        for (int m = 0; m <= l; m++) Y(l, m) = nP(l, m) * std::exp(I * ((double)m * phi));
        for (int m = 1; m <= l; m++) Y(l, -m) = ((m & 1) ? -1.0 : 1.0) * std::conj(Y(l, m));
    }
    return Y;
}

FFTS2_matrix_complex FFTS2_complex::forward_transform(std::function<complex<double>(double, double)> func)
{
    const long N = 2 * B;
    FFTS2_matrix_complex F(l_max);
    dev_->forward_fill(FFTSO3_DOMAIN_S2, l_max, FFTSO3_COMPLEX, precision_, [&](long i0, long i1, double *out) {
        for (long i = i0; i < i1; i++) {
            const complex<double> v = func(theta_k((int)(i / N)), phi_j((int)(i % N)));
            out[2 * (i - i0)] = v.real();
            out[2 * (i - i0) + 1] = v.imag();
        }
    }, F.v.data());
    return F;
}

FFTS2_matrix_complex FFTS2_complex::forward_transform(const std::vector<complex<double>> &grid)
{
    return forward_transform_batch(grid, 1)[0];
}

std::vector<FFTS2_matrix_complex> FFTS2_complex::forward_transform_batch(const std::vector<complex<double>> &grids,
                                                                         int n_fn)
{
    const long N = 2 * B;
    if (n_fn < 1 || (long)grids.size() != N * N * n_fn)
        throw std::invalid_argument("forward_transform_batch: expected n_fn grids of (2B)^2 values");
    const long nc = FFTS2_matrix_complex::size_for(l_max);
    std::vector<complex<double>> coef(nc * n_fn);
    dev_->forward(FFTSO3_DOMAIN_S2, l_max, n_fn, FFTSO3_COMPLEX, precision_, grids.data(), coef.data());
    std::vector<FFTS2_matrix_complex> out(n_fn);
    for (int i = 0; i < n_fn; i++) {
        out[i].l_max = l_max;
        out[i].v.assign(coef.begin() + i * nc, coef.begin() + (i + 1) * nc);
    }
    return out;
}

complex<double> FFTS2_complex::inverse_transform(FFTS2_matrix_complex F, double theta, double phi)
{
    const FFTS2_matrix_complex Y = spherical_harmonics(theta, phi, F.l_max);
    complex<double> f = 0.0;
    for (int l = 0; l <= F.l_max; l++)
        for (int m = -l; m <= l; m++) f += F(l, m) * Y(l, m);
    return f;
}

std::vector<complex<double>> FFTS2_complex::inverse_transform(const FFTS2_matrix_complex &F)
{
    return inverse_transform_batch(std::vector<FFTS2_matrix_complex>{F});
}

std::vector<complex<double>> FFTS2_complex::inverse_transform_batch(const std::vector<FFTS2_matrix_complex> &F)
{
    if (F.empty()) return {};
    init(F[0].l_max);
    const long nc = FFTS2_matrix_complex::size_for(l_max), N = 2 * B;
    std::vector<complex<double>> coef(nc * F.size());
    for (size_t i = 0; i < F.size(); i++) {
        if (F[i].l_max != l_max) throw std::invalid_argument("inverse_transform_batch: mixed l_max");
        std::copy(F[i].v.begin(), F[i].v.end(), coef.begin() + i * nc);
    }
    std::vector<complex<double>> grid(N * N * F.size());
    dev_->inverse(FFTSO3_DOMAIN_S2, l_max, (int)F.size(), FFTSO3_COMPLEX, precision_, coef.data(), grid.data());
    return grid;
}

std::vector<complex<double>> FFTS2_complex::inverse_transform_points(const FFTS2_matrix_complex &F,
                                                                     const std::vector<double> &angles)
{
    if (angles.size() % 2) throw std::invalid_argument("inverse_transform_points: 2 angles per point");
    init(F.l_max);
    const long n = (long)angles.size() / 2;
    std::vector<complex<double>> v(n);
    dev_->evaluate(FFTSO3_DOMAIN_S2, l_max, FFTSO3_COMPLEX, precision_, F.data(), angles.data(), n, v.data());
    return v;
}

// sum_k w_k Y^l_0(theta_k, 0) B / sqrt(pi) = delta_{l0} for l < 2B (src/fdcl_FFTS2.cpp:243-262)
double FFTS2_complex::check_weight()
{
    compute_weight();
    std::vector<complex<double>> sum(2 * B, 0.0);
    for (int k = 0; k < 2 * B; k++) {
        const FFTS2_matrix_complex Y = spherical_harmonics(theta_k(k), 0.0, 2 * B - 1);
        for (int l = 0; l < 2 * B; l++) sum[l] += Y(l, 0) * weight[k];
    }
    const double s = (double)B / std::sqrt(M_PI);
    double err = std::abs(sum[0] * s - 1.0);
    for (int l = 1; l < 2 * B; l++) err += std::abs(sum[l] * s);
    std::cout << "fdcl::FFTS2_complex::check_weight: error = " << err << std::endl;
    return err;
}

double FFTS2_complex::check_transform()
{
    FFTS2_matrix_complex F(l_max);
    F.setRandom();
    const FFTS2_matrix_complex Fc = F;
    FFTS2_matrix_complex F2 = forward_transform(
        std::function<complex<double>(double, double)>([&](double t, double p) { return inverse_transform(Fc, t, p); }));
    const double err = (F2 - F).norm();
    std::cout << "fdcl::FFTS2_complex::check_transform: error = " << err << std::endl;
    return err;
}

void FFTS2_complex::check_all()
{
    check_weight();
    check_transform();
}

// ---------------------------------------------------------------------------------------------
FFTS2_real::FFTS2_real(int l_max_) { init(l_max_); }

void FFTS2_real::init(int l_max_) { FFTS2_complex::init(l_max_); }

// y^l_0 = nP^l_0, y^l_m = sqrt2 (-1)^m nP^l_m cos(m phi), y^l_{-m} = sqrt2 (-1)^m nP^l_m sin(m phi)
// (src/fdcl_FFTS2.cpp:383-422)
FFTS2_matrix_real FFTS2_real::spherical_harmonics(double theta, double phi, int L)
{
    const FFTS2_matrix_real nP = nor_assoc_Legendre_poly(std::cos(theta), L);
    FFTS2_matrix_real y(L);
    const double r2 = std::sqrt(2.0);
    for (int l = 0; l <= L; l++) {
        y(l, 0) = nP(l, 0);
        for (int m = 1; m <= l; m++) {
            const double t = r2 * ((m & 1) ? -1.0 : 1.0) * nP(l, m);
            y(l, m) = t * std::cos((double)m * phi);
            y(l, -m) = t * std::sin((double)m * phi);
        }
    }
    return y;
}

FFTS2_matrix_real FFTS2_real::forward_transform(std::function<double(double, double)> func)
{
    const long N = 2 * B;
    FFTS2_matrix_real F(l_max);
    dev_->forward_fill(FFTSO3_DOMAIN_S2, l_max, FFTSO3_REAL, precision_, [&](long i0, long i1, double *out) {
        for (long i = i0; i < i1; i++) out[i - i0] = func(theta_k((int)(i / N)), phi_j((int)(i % N)));
    }, F.v.data());
    return F;
}

FFTS2_matrix_real FFTS2_real::forward_transform(const std::vector<double> &grid)
{
    return forward_transform_batch(grid, 1)[0];
}

std::vector<FFTS2_matrix_real> FFTS2_real::forward_transform_batch(const std::vector<double> &grids, int n_fn)
{
    const long N = 2 * B;
    if (n_fn < 1 || (long)grids.size() != N * N * n_fn)
        throw std::invalid_argument("forward_transform_batch: expected n_fn grids of (2B)^2 values");
    const long nc = FFTS2_matrix_real::size_for(l_max);
    std::vector<double> coef(nc * n_fn);
    dev_->forward(FFTSO3_DOMAIN_S2, l_max, n_fn, FFTSO3_REAL, precision_, grids.data(), coef.data());
    std::vector<FFTS2_matrix_real> out(n_fn);
    for (int i = 0; i < n_fn; i++) {
        out[i].l_max = l_max;
        out[i].v.assign(coef.begin() + i * nc, coef.begin() + (i + 1) * nc);
    }
    return out;
}

double FFTS2_real::inverse_transform(FFTS2_matrix_real F, double theta, double phi)
{
    const FFTS2_matrix_real y = spherical_harmonics(theta, phi, F.l_max);
    double f = 0.0;
    for (int l = 0; l <= F.l_max; l++)
        for (int m = -l; m <= l; m++) f += F(l, m) * y(l, m);
    return f;
}

std::vector<double> FFTS2_real::inverse_transform(const FFTS2_matrix_real &F)
{
    return inverse_transform_batch(std::vector<FFTS2_matrix_real>{F});
}

std::vector<double> FFTS2_real::inverse_transform_batch(const std::vector<FFTS2_matrix_real> &F)
{
    if (F.empty()) return {};
    init(F[0].l_max);
    const long nc = FFTS2_matrix_real::size_for(l_max), N = 2 * B;
    std::vector<double> coef(nc * F.size());
    for (size_t i = 0; i < F.size(); i++) {
        if (F[i].l_max != l_max) throw std::invalid_argument("inverse_transform_batch: mixed l_max");
        std::copy(F[i].v.begin(), F[i].v.end(), coef.begin() + i * nc);
    }
    std::vector<double> grid(N * N * F.size());
    dev_->inverse(FFTSO3_DOMAIN_S2, l_max, (int)F.size(), FFTSO3_REAL, precision_, coef.data(), grid.data());
    return grid;
}

std::vector<double> FFTS2_real::inverse_transform_points(const FFTS2_matrix_real &F, const std::vector<double> &angles)
{
    if (angles.size() % 2) throw std::invalid_argument("inverse_transform_points: 2 angles per point");
    init(F.l_max);
    const long n = (long)angles.size() / 2;
    std::vector<double> v(n);
    dev_->evaluate(FFTSO3_DOMAIN_S2, l_max, FFTSO3_REAL, precision_, F.data(), angles.data(), n, v.data());
    return v;
}

double FFTS2_real::check_transform()
{
    FFTS2_matrix_real F(l_max);
    F.setRandom();
    const FFTS2_matrix_real Fc = F;
    FFTS2_matrix_real F2 = forward_transform(
        std::function<double(double, double)>([&](double t, double p) { return inverse_transform(Fc, t, p); }));
    const double err = (F2 - F).norm();
    std::cout << "fdcl::FFTS2_real::check_transform: error = " << err << std::endl;
    return err;
}

void FFTS2_real::check_all()
{
    check_weight();
    check_transform();
}

}  // namespace fdcl
