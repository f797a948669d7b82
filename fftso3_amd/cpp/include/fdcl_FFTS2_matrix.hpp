// fdcl_FFTS2_matrix.hpp -- coefficient container of the S^2 transforms, fdcl::FFTS2_matrix<T>.
//
// Same interface as the reference's container (include/fdcl_FFTS2_matrix.hpp:40-86,
// src/fdcl_FFTS2_matrix.cpp:10-47,163-171): F(l,m) is F^l_m and F[l] the (2l+1)-vector of
// degree l (entry m+l).  Storage is one packed array with F(l,m) at l^2 + m + l, the layout the
// device transforms use.
#ifndef FDCL_FFTS2_MATRIX_HPP
#define FDCL_FFTS2_MATRIX_HPP

#include "fdcl_FFTSO3_matrix.hpp"

namespace fdcl
{

template <class ScalarType>
class FFTS2_matrix
{
public:
    int l_max = 0;
    std::vector<ScalarType> v;  // packed coefficients, l^2 + m + l

    FFTS2_matrix() {}
    explicit FFTS2_matrix(int l_max) { init(l_max); }

    static long index(int l, int m) { return (long)l * l + m + l; }
    static long size_for(int l_max) { return (long)(l_max + 1) * (l_max + 1); }

    void init(int l_max_)
    {
        l_max = l_max_;
        v.assign(size_for(l_max), ScalarType(0));
    }

    BlockView<ScalarType> operator[](int l)
    {
        assert(l >= 0 && l <= l_max);
        return BlockView<ScalarType>(v.data() + (long)l * l, 2 * l + 1, 1);
    }
    const BlockView<ScalarType> operator[](int l) const
    {
        assert(l >= 0 && l <= l_max);
        return BlockView<ScalarType>(const_cast<ScalarType *>(v.data()) + (long)l * l, 2 * l + 1, 1);
    }
    ScalarType &operator()(int l, int m)
    {
        assert(l >= 0 && l <= l_max && std::abs(m) <= l);
        return v[index(l, m)];
    }
    const ScalarType &operator()(int l, int m) const { return v[index(l, m)]; }

    ScalarType *data() { return v.data(); }
    const ScalarType *data() const { return v.data(); }
    long size() const { return (long)v.size(); }

    void setRandom()
    {
        for (auto &x : v) detail::draw(x);
    }
    void setZero()
    {
        for (auto &x : v) x = ScalarType(0);
    }

    // sum over l of the 2-norm of F[l] (src/fdcl_FFTS2_matrix.cpp:163-171)
    double norm() const
    {
        double s = 0.0;
        for (int l = 0; l <= l_max; l++) s += (*this)[l].norm();
        return s;
    }

    FFTS2_matrix<double> real() const
    {
        FFTS2_matrix<double> R(l_max);
        for (long i = 0; i < size(); i++) R.v[i] = detail::real_part(v[i]);
        return R;
    }

    template <class U>
    auto operator+(const FFTS2_matrix<U> &o) const -> FFTS2_matrix<decltype(ScalarType() + U())>
    {
        assert(o.l_max == l_max);
        FFTS2_matrix<decltype(ScalarType() + U())> R(l_max);
        for (long i = 0; i < size(); i++) R.v[i] = v[i] + o.v[i];
        return R;
    }
    template <class U>
    auto operator-(const FFTS2_matrix<U> &o) const -> FFTS2_matrix<decltype(ScalarType() - U())>
    {
        assert(o.l_max == l_max);
        FFTS2_matrix<decltype(ScalarType() - U())> R(l_max);
        for (long i = 0; i < size(); i++) R.v[i] = v[i] - o.v[i];
        return R;
    }
    FFTS2_matrix<complex<double>> operator*(const complex<double> &c) const
    {
        FFTS2_matrix<complex<double>> R(l_max);
        for (long i = 0; i < size(); i++) R.v[i] = v[i] * c;
        return R;
    }
    FFTS2_matrix<ScalarType> operator*(const double &c) const
    {
        FFTS2_matrix<ScalarType> R(l_max);
        for (long i = 0; i < size(); i++) R.v[i] = v[i] * c;
        return R;
    }

    friend std::ostream &operator<<(std::ostream &os, const FFTS2_matrix &M)
    {
        for (int l = 0; l <= M.l_max; l++) {
            os << "l=" << l << std::endl;
            const auto b = M[l];
            for (int i = 0; i < b.rows(); i++) os << b(i, 0) << std::endl;
            os << std::endl;
        }
        return os;
    }
};

typedef FFTS2_matrix<double> FFTS2_matrix_real;
typedef FFTS2_matrix<complex<double>> FFTS2_matrix_complex;

}  // namespace fdcl

#endif
