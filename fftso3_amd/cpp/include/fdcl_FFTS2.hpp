// fdcl_FFTS2.hpp -- fdcl::FFTS2_complex / fdcl::FFTS2_real on MI355X.
//
// Drop-in for the reference's S^2 classes (include/fdcl_FFTS2.hpp:28-160): spherical harmonics,
// normalised associated Legendre functions, forward / inverse transforms and the self checks.
// The forward transform runs on the GPU through libfftso3_amd.so; pointwise evaluations stay on
// the host in FP64.  Grid and batch entry points are additions, as in fdcl_FFTSO3.hpp.
#ifndef FDCL_FFTS2_HPP
#define FDCL_FFTS2_HPP

#include "fdcl_FFTS2_matrix.hpp"
#include "fdcl_FFTSO3.hpp"

namespace fdcl
{

class FFTS2_complex
{
public:
    int l_max = 0;
    bool check_verbose = false;

    FFTS2_complex() {}
    explicit FFTS2_complex(int l_max);
    virtual ~FFTS2_complex() {}
    void init(int l_max);

    void set_precision(int precision) { precision_ = precision; }
    int precision() const { return precision_; }

    // Y^l_m(theta, phi) = nP^l_m(cos theta) exp(i m phi), Y^l_{-m} = (-1)^m conj Y^l_m
    // (src/fdcl_FFTS2.cpp:17-46)
    FFTS2_matrix_complex spherical_harmonics(double theta, double phi, int L);
    FFTS2_matrix_complex spherical_harmonics(double theta, double phi) { return spherical_harmonics(theta, phi, l_max); }

    // forward transform (src/fdcl_FFTS2.cpp:125-192); the callback is called concurrently
    FFTS2_matrix_complex forward_transform(std::function<complex<double>(double, double)> func);
    // grid [k][j] = f(theta_k, phi_j), (2B)^2 values
    FFTS2_matrix_complex forward_transform(const std::vector<complex<double>> &grid);
    std::vector<FFTS2_matrix_complex> forward_transform_batch(const std::vector<complex<double>> &grids, int n_fn);

    // pointwise inverse f = sum F^l_m Y^l_m (src/fdcl_FFTS2.cpp:194-230)
    complex<double> inverse_transform(FFTS2_matrix_complex F, double theta, double phi);
    std::vector<complex<double>> inverse_transform(const FFTS2_matrix_complex &F);
    std::vector<complex<double>> inverse_transform_batch(const std::vector<FFTS2_matrix_complex> &F);
    // pointwise inverse at many directions on the GPU: angles holds (theta, phi) per point
    std::vector<complex<double>> inverse_transform_points(const FFTS2_matrix_complex &F,
                                                          const std::vector<double> &angles);

    void check_all();
    double check_weight();
    double check_transform();

    double theta_k(int k) const;
    double phi_j(int j) const;
    std::vector<double> weight;  // w_k, k < 2B (src/fdcl_FFTS2.cpp:102-123)

    // normalised associated Legendre nP^l_m(x), Condon-Shortley phase (src/fdcl_FFTS2.cpp:48-100)
    FFTS2_matrix_real nor_assoc_Legendre_poly(double x, int L);

protected:
    int B = 1;
    int precision_ = FFTSO3_FP64;
    std::vector<double> compute_weight();
    detail::DeviceHandle dev_;

public:
    // identity of this object's device plan (distinct for every object, copies included)
    const void *device_plan_id() const { return dev_.get(); }
};

class FFTS2_real : public FFTS2_complex
{
public:
    FFTS2_real() {}
    explicit FFTS2_real(int l_max);
    void init(int l_max);

    // real spherical harmonics y = T Y (src/fdcl_FFTS2.cpp:383-422)
    FFTS2_matrix_real spherical_harmonics(double theta, double phi, int L);
    FFTS2_matrix_real spherical_harmonics(double theta, double phi) { return spherical_harmonics(theta, phi, l_max); }

    // F_r = Re(conj(T) F_c) (src/fdcl_FFTS2.cpp:424-436)
    FFTS2_matrix_real forward_transform(std::function<double(double, double)> func);
    FFTS2_matrix_real forward_transform(const std::vector<double> &grid);
    std::vector<FFTS2_matrix_real> forward_transform_batch(const std::vector<double> &grids, int n_fn);

    double inverse_transform(FFTS2_matrix_real F, double theta, double phi);
    std::vector<double> inverse_transform(const FFTS2_matrix_real &F);
    std::vector<double> inverse_transform_batch(const std::vector<FFTS2_matrix_real> &F);
    std::vector<double> inverse_transform_points(const FFTS2_matrix_real &F, const std::vector<double> &angles);

    void check_all();
    double check_transform();
};

}  // namespace fdcl

#endif
