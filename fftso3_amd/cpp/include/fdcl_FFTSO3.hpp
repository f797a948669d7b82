// fdcl_FFTSO3.hpp -- fdcl::FFTSO3_complex / fdcl::FFTSO3_real on MI355X.
//
// Drop-in for the reference's classes (include/fdcl_FFTSO3.hpp:27-269): same names, members and
// overloads for construction, Wigner functions, forward / inverse transforms and the self
// checks.  The forward transform runs on the GPU through the C ABI of libfftso3_amd.so
// (include/fftso3_amd.h); pointwise evaluations (inverse_transform at one rotation, wigner_d,
// wigner_D, real_harmonics) stay on the host in FP64, as they are O(B^3) per call.
//
// Additions over the reference (it has no grid or batch entry points):
//   forward_transform(grid)            grid [k][j1][j2] of (2B)^3 samples -> coefficients
//   inverse_transform(F)               coefficients -> the whole (2B)^3 grid
//   forward_transform_batch / inverse_transform_batch
//   set_precision(FFTSO3_FP64 | FFTSO3_FP32 | FFTSO3_BF16X3 | FFTSO3_BF16X6)
//
// Differences forced by the environment: there is no Eigen, so rotations are fdcl::Matrix3d
// (a 3x3 row-major value type with R(i,j)) and F[l] is a column-major view (fdcl::BlockView).
// Errors of the device path throw std::runtime_error carrying the library's message; the
// reference only asserts.
#ifndef FDCL_FFTSO3_HPP
#define FDCL_FFTSO3_HPP

#include <functional>
#include <memory>
#include <stdexcept>
#include <vector>

#include "fdcl_Clebsch_Gordon.hpp"
#include "fdcl_FFTS2_matrix.hpp"
#include "fdcl_FFTSO3_matrix.hpp"
#include "fftso3_amd.h"

// Callers that have Eigen get the reference's Eigen::Matrix3d overloads as inline adapters.
#if defined(__has_include)
#if __has_include(<Eigen/Dense>)
#include <Eigen/Dense>
#define FDCL_HAVE_EIGEN 1
#endif
#endif

namespace fdcl
{

// 3x3 rotation matrix, row-major: R(i,j) is row i, column j.
struct Matrix3d {
    double a[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double &operator()(int i, int j) { return a[3 * i + j]; }
    double operator()(int i, int j) const { return a[3 * i + j]; }
};

#ifdef FDCL_HAVE_EIGEN
inline Matrix3d to_matrix3d(const Eigen::Matrix3d &E)
{
    Matrix3d R;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) R(i, j) = E(i, j);
    return R;
}
inline Eigen::Matrix3d to_eigen(const Matrix3d &R)
{
    Eigen::Matrix3d E;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) E(i, j) = R(i, j);
    return E;
}
#endif

// 3-2-3 Euler angles <-> rotation matrix (src/misc_matrix_func.cpp:144-178)
std::vector<double> R2Euler323(const Matrix3d &R);
Matrix3d Euler3232R(double a, double b, double g);

namespace detail
{
// A device plan of the C ABI with its own device grid / coefficient buffers, created on first
// use and kept while the (domain, l_max, n_fn, kind, precision) it was made for is unchanged.
class DevicePlan
{
public:
    ~DevicePlan();
    // grid / coefficient arrays of n_fn functions, host side in double or complex<double>
    void forward(int domain, int l_max, int n_fn, int kind, int precision, const void *grid, void *coef);
    void inverse(int domain, int l_max, int n_fn, int kind, int precision, const void *coef, void *grid);
    // values of one function at n points (angles: 3 or 2 doubles per point), host arrays
    void evaluate(int domain, int l_max, int kind, int precision, const void *coef, const double *angles, long n,
                  void *values);
    // Forward transform of one function sampled on the host: fill(i0, i1, out) writes grid values
    // [i0, i1) (grid order; complex as re, im) as doubles and is called concurrently from OpenMP
    // threads.  Sampling runs chunk by chunk into pinned buffers while the previous chunk is copied
    // to the device on a stream of the plan's own.
    using Fill = std::function<void(long i0, long i1, double *out)>;
    void forward_fill(int domain, int l_max, int kind, int precision, const Fill &fill, void *coef);
    // S^2 coefficients F ((l_max+1)^2, double or complex) rotated by n rotations (angles: alpha,
    // beta, gamma per rotation) into G (n x (l_max+1)^2), through the SO(3) plan's tables
    void rotate_s2(int l_max, int kind, int precision, const void *F, const double *angles, long n, void *G);

private:
    void prepare(int domain, int l_max, int n_fn, int kind, int precision);
    fftso3_plan_t plan_ = nullptr;
    int key_[5] = {-1, -1, -1, -1, -1};
    void *d_grid_ = nullptr, *d_coef_ = nullptr;
    long grid_elems_ = 0, coef_elems_ = 0;
    std::vector<char> stage_;  // host staging for the single-precision plans
    void *d_ang_ = nullptr, *d_val_ = nullptr;
    long pts_cap_ = 0;
    void *pin_[2] = {nullptr, nullptr};  // pinned sampling buffers of forward_fill
    size_t pin_bytes_ = 0;
    void *stream_ = nullptr, *ev_[2] = {nullptr, nullptr};
};

// The device plan of one transform object.  Copies never share it: a copied (or assigned)
// object starts with a device plan of its own, created on first use, so copies may run from
// different threads just as the reference's by-value objects can.
class DeviceHandle
{
public:
    DeviceHandle() : p_(new DevicePlan) {}
    DeviceHandle(const DeviceHandle &) : p_(new DevicePlan) {}
    DeviceHandle &operator=(const DeviceHandle &o)
    {
        if (this != &o) p_.reset(new DevicePlan);
        return *this;
    }
    DevicePlan *operator->() const { return p_.get(); }
    const DevicePlan *get() const { return p_.get(); }

private:
    std::unique_ptr<DevicePlan> p_;
};
}  // namespace detail

class FFTSO3_complex
{
public:
    int l_max = 0;
    bool check_verbose = false;

    FFTSO3_complex() {}
    explicit FFTSO3_complex(int l_max);
    virtual ~FFTSO3_complex() {}
    void init(int l_max);

    // arithmetic of the device transforms (FFTSO3_FP64 by default: the exact drop-in)
    void set_precision(int precision) { precision_ = precision; }
    int precision() const { return precision_; }

    // Wigner small-d d^l_{mn}(beta), l <= l_max (src/fdcl_FFTSO3.cpp:127-194)
    FFTSO3_matrix_real wigner_d(double beta);
    // Wigner D^l_{mn} = exp(-i m alpha) d^l_{mn}(beta) exp(-i n gamma) (src/fdcl_FFTSO3.cpp:201-224)
    FFTSO3_matrix_complex wigner_D(const Matrix3d &R);
    FFTSO3_matrix_complex wigner_D(double alpha, double beta, double gamma);

    // forward transform of f sampled on the grid (src/fdcl_FFTSO3.cpp:654-763); the callback is
    // called concurrently from OpenMP threads, as in the reference
    FFTSO3_matrix_complex forward_transform(std::function<complex<double>(double, double, double)> func);
    FFTSO3_matrix_complex forward_transform(std::function<complex<double>(Matrix3d)> func);
    // grid [k][j1][j2] = f(alpha_j1, beta_k, gamma_j2), (2B)^3 values
    FFTSO3_matrix_complex forward_transform(const std::vector<complex<double>> &grid);
    std::vector<FFTSO3_matrix_complex> forward_transform_batch(const std::vector<complex<double>> &grids, int n_fn);

    // pointwise inverse; like the reference, re-initialises the object to F.l_max
    // (src/fdcl_FFTSO3.cpp:498-531)
    complex<double> inverse_transform(FFTSO3_matrix_complex F, double alpha, double beta, double gamma);
    complex<double> inverse_transform(FFTSO3_matrix_complex F, const Matrix3d &R);
    // inverse on the whole grid [k][j1][j2] (on the GPU)
    std::vector<complex<double>> inverse_transform(const FFTSO3_matrix_complex &F);
    std::vector<complex<double>> inverse_transform_batch(const std::vector<FFTSO3_matrix_complex> &F);
    // pointwise inverse at many rotations on the GPU: angles holds (alpha, beta, gamma) per point
    std::vector<complex<double>> inverse_transform_points(const FFTSO3_matrix_complex &F,
                                                          const std::vector<double> &angles);

    // S^2 coefficients rotated by many rotations on the GPU: G_i^l = D^l(R_i) F^l (wigner_D,
    // src/fdcl_FFTSO3.cpp:201-224), FP64; l_max of F is the object's
    std::vector<FFTS2_matrix_complex> rotate_batch(const FFTS2_matrix_complex &F, const std::vector<Matrix3d> &R);

    // character chi_l(theta) = sin((2l+1) theta / 2) / sin(theta / 2) (src/fdcl_FFTSO3.cpp:404-415)
    std::vector<double> character(double theta);

    // Clebsch-Gordan coefficients C^{l,m}_{l1,m1,l2,m2}: call C.compute(l1, l2), then read C(l,m,l1,m1,l2,m2)
    // (include/fdcl_FFTSO3.hpp:35-39 of the reference)
    Clebsch_Gordon_complex C;
    // derivatives of D^l at the identity along e_1, e_2, e_3: u[i] = d/dt D(exp(t e_i)) at t = 0,
    // i = 1..3 (u[0] is empty), i.e. -i times the angular momentum matrices
    // (src/fdcl_FFTSO3.cpp:417-453)
    std::vector<FFTSO3_matrix_complex> deriv_wigner_D();

    // self checks of the reference (src/fdcl_FFTSO3.cpp:258-415,455-493,532-545,943-986,1483-1571);
    // each returns an error measure that is ~0 (below 1e-10) when the implementation is right
    void check_all();
    double check_weight();
    double check_wigner_d();
    double check_transform();
    // D(exp(t e_i)) against the matrix exponential of t u[i] (an exact identity, no finite difference)
    double check_deriv_wigner_D();
    // D^{l1}_{m1 n1} D^{l2}_{m2 n2} = sum_l C^{l,m1+m2}_{l1,m1,l2,m2} C^{l,n1+n2}_{l1,n1,l2,n2} D^l at a random rotation
    double check_Clebsch_Gordon();

#ifdef FDCL_HAVE_EIGEN
    FFTSO3_matrix_complex wigner_D(const Eigen::Matrix3d &R) { return wigner_D(to_matrix3d(R)); }
    FFTSO3_matrix_complex forward_transform(std::function<complex<double>(Eigen::Matrix3d)> func)
    {
        return forward_transform(std::function<complex<double>(Matrix3d)>([&](Matrix3d R) { return func(to_eigen(R)); }));
    }
    complex<double> inverse_transform(FFTSO3_matrix_complex F, const Eigen::Matrix3d &R)
    {
        return inverse_transform(std::move(F), to_matrix3d(R));
    }
#endif

    // grid nodes
    double alpha_j(int j) const;
    double beta_k(int k) const;
    double gamma_j(int j) const;
    std::vector<double> weight;  // quadrature weights w_k, k < 2B (src/fdcl_FFTSO3.cpp:236-256)

protected:
    int B = 1;
    int precision_ = FFTSO3_FP64;
    std::vector<double> compute_weight();
    FFTSO3_matrix_real wigner_d(double beta, int L);
    FFTSO3_matrix_complex wigner_D(double alpha, double beta, double gamma, int L);
    detail::DeviceHandle dev_;

public:
    // identity of this object's device plan (distinct for every object, copies included)
    const void *device_plan_id() const { return dev_.get(); }
};

class FFTSO3_real : public FFTSO3_complex
{
public:
    FFTSO3_real() {}
    explicit FFTSO3_real(int l_max);
    void init(int l_max);

    // real harmonics U^l = conj(T^l) D^l (T^l)^T (src/fdcl_FFTSO3.cpp:1015-1105)
    FFTSO3_matrix_real real_harmonics(double alpha, double beta, double gamma);
    FFTSO3_matrix_real real_harmonics(const Matrix3d &R);

    // real forward transform (src/fdcl_FFTSO3.cpp:774-850, 1169-1181)
    FFTSO3_matrix_real forward_transform(std::function<double(double, double, double)> func);
    FFTSO3_matrix_real forward_transform(std::function<double(Matrix3d)> func);
    FFTSO3_matrix_real forward_transform(const std::vector<double> &grid);
    std::vector<FFTSO3_matrix_real> forward_transform_batch(const std::vector<double> &grids, int n_fn);

    // pointwise real inverse f = sum (2l+1) F_r U (src/fdcl_FFTSO3.cpp:1141-1167); uses l_max
    double inverse_transform(FFTSO3_matrix_real F, double alpha, double beta, double gamma);
    double inverse_transform(FFTSO3_matrix_real F, const Matrix3d &R);
    std::vector<double> inverse_transform(const FFTSO3_matrix_real &F);
    std::vector<double> inverse_transform_batch(const std::vector<FFTSO3_matrix_real> &F);
    std::vector<double> inverse_transform_points(const FFTSO3_matrix_real &F, const std::vector<double> &angles);

#ifdef FDCL_HAVE_EIGEN
    FFTSO3_matrix_real real_harmonics(const Eigen::Matrix3d &R) { return real_harmonics(to_matrix3d(R)); }
    FFTSO3_matrix_real forward_transform(std::function<double(Eigen::Matrix3d)> func)
    {
        return forward_transform(std::function<double(Matrix3d)>([&](Matrix3d R) { return func(to_eigen(R)); }));
    }
    double inverse_transform(FFTSO3_matrix_real F, const Eigen::Matrix3d &R)
    {
        return inverse_transform(std::move(F), to_matrix3d(R));
    }
#endif

    // T^l of src/fdcl_FFTSO3.cpp:1107-1139
    FFTSO3_matrix_complex matrix2rsph(int L);

    // Clebsch-Gordan coefficients of the real harmonics: c.compute(l1, l2), then c(l,m,l1,m1,l2,m2)
    Clebsch_Gordon_real c;

    // S^2 workflows of example/example3.cpp:50-53 and example/example3_SSM.hpp:62-70:
    // real spherical-harmonic coefficients of a rotated function, G^l = U^l(R) F^l
    FFTS2_matrix_real rotate(const FFTS2_matrix_real &F, const Matrix3d &R);
    // the same for many rotations at once, on the GPU (FP64); l_max of F is the object's
    std::vector<FFTS2_matrix_real> rotate_batch(const FFTS2_matrix_real &F, const std::vector<Matrix3d> &R);
    // the matching cost J(R) = sum_l (G^l)^T U^l(R) F^l at every rotation of the SO(3) grid
    // [k][j1][j2] = (alpha_j1, beta_k, gamma_j2), l <= l_max: one real inverse transform on the GPU
    // of H^l_mn = G^l_m F^l_n / (2l+1)
    std::vector<double> correlation(const FFTS2_matrix_real &F, const FFTS2_matrix_real &G);
    // derivatives of U^l at the identity along e_1, e_2, e_3 (u[0] empty): conj(T) u_D T^T with
    // the complex derivatives u_D (src/fdcl_FFTSO3.cpp:1577-1645)
    std::vector<FFTSO3_matrix_real> deriv_real_harmonics();

    void check_all();
    double check_transform();
    double check_real_harmonics();
    double check_deriv_real_harmonics();  // U(exp(t e_i)) against expm(t u[i])
    // U^{l1}_{m1 n1} U^{l2}_{m2 n2} = sum Re(c c*) U^l at a random rotation (src/fdcl_FFTSO3.cpp:1685-1768)
    double check_Clebsch_Gordon();

protected:
    FFTSO3_matrix_real real_harmonics(double alpha, double beta, double gamma, int L);
};

}  // namespace fdcl

#endif
