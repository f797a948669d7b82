// fdcl_FFTSO3_matrix.hpp -- coefficient container of the SO(3) transforms, fdcl::FFTSO3_matrix<T>.
//
// Same interface as the reference's container (include/fdcl_FFTSO3_matrix.hpp:44-91,
// src/fdcl_FFTSO3_matrix.cpp:10-47,154-181): F(l,m,n) is the coefficient F^l_{m,n}, F[l] the
// (2l+1) x (2l+1) block of degree l (row m+l, column n+l), plus setRandom / setZero / norm /
// real / + / - / * and stream output.  Storage is one packed array: degree l's block starts at
// off(l) = l(4l^2-1)/3 and is column-major, i.e. F(l,m,n) sits at off(l) + (m+l) + (n+l)(2l+1) --
// the reference's per-degree column-major blocks laid end to end, and the layout the device
// transforms read and write, so a whole matrix moves to and from the GPU in one copy.
//
// The reference returns F[l] as an Eigen matrix; this build has no Eigen, so F[l] is a light
// column-major view (rows(), cols(), operator()(i,j), data(), norm()).
#ifndef FDCL_FFTSO3_MATRIX_HPP
#define FDCL_FFTSO3_MATRIX_HPP

#include <cassert>
#include <cmath>
#include <complex>
#include <cstdint>
#include <iostream>
#include <vector>

namespace fdcl
{
using std::complex;

#ifndef FDCL_IMAGINARY_UNIT
#define FDCL_IMAGINARY_UNIT
const complex<double> I(0.0, 1.0);
#endif

namespace detail
{
inline double abs2(double x) { return x * x; }
inline double abs2(const complex<double> &x) { return std::norm(x); }
inline double real_part(double x) { return x; }
inline double real_part(const complex<double> &x) { return x.real(); }

// uniform [-1, 1] draws for setRandom (Eigen's setRandom range); one shared SplitMix64 stream
inline uint64_t &random_state()
{
    static uint64_t s = 0x853c49e6748fea9bULL;
    return s;
}
inline double uniform_pm1()
{
    uint64_t &s = random_state();
    s += 0x9e3779b97f4a7c15ULL;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    z ^= z >> 31;
    return 2.0 * ((z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
}
inline void draw(double &x) { x = uniform_pm1(); }
inline void draw(complex<double> &x)
{
    const double re = uniform_pm1();
    x = complex<double>(re, uniform_pm1());
}
}  // namespace detail

// Column-major view of one (2l+1) x (2l+1) (or 2l+1 x 1) block.
template <class T>
class BlockView
{
public:
    BlockView(T *p, int r, int c) : p_(p), r_(r), c_(c) {}
    int rows() const { return r_; }
    int cols() const { return c_; }
    T &operator()(int i, int j) { return p_[i + (long)j * r_]; }
    const T &operator()(int i, int j) const { return p_[i + (long)j * r_]; }
    T &operator()(int i) { return p_[i]; }
    T *data() { return p_; }
    const T *data() const { return p_; }
    double norm() const
    {
        double s = 0.0;
        for (long i = 0; i < (long)r_ * c_; i++) s += detail::abs2(p_[i]);
        return std::sqrt(s);
    }

private:
    T *p_;
    int r_, c_;
};

template <class ScalarType>
class FFTSO3_matrix
{
public:
    int l_max = -1;
    std::vector<ScalarType> v;  // packed coefficients, off(l) + (m+l) + (n+l)(2l+1)

    FFTSO3_matrix() {}
    explicit FFTSO3_matrix(int l_max) { init(l_max); }

    static long offset(int l) { return (long)l * (4L * l * l - 1) / 3; }
    static long size_for(int l_max) { return offset(l_max + 1); }
    static long index(int l, int m, int n) { return offset(l) + (m + l) + (long)(n + l) * (2 * l + 1); }

    void init(int l_max_)
    {
        l_max = l_max_;
        v.assign(size_for(l_max), ScalarType(0));
    }

    BlockView<ScalarType> operator[](int l)
    {
        assert_index(l);
        return BlockView<ScalarType>(v.data() + offset(l), 2 * l + 1, 2 * l + 1);
    }
    const BlockView<ScalarType> operator[](int l) const
    {
        assert(l >= 0 && l <= l_max);
        return BlockView<ScalarType>(const_cast<ScalarType *>(v.data()) + offset(l), 2 * l + 1, 2 * l + 1);
    }
    ScalarType &operator()(int l, int m, int n)
    {
        assert_index(l, m, n);
        return v[index(l, m, n)];
    }
    const ScalarType &operator()(int l, int m, int n) const { return v[index(l, m, n)]; }

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

    // sum over l of the Frobenius norm of F[l] (src/fdcl_FFTSO3_matrix.cpp:154-163)
    double norm() const
    {
        double s = 0.0;
        for (int l = 0; l <= l_max; l++) s += (*this)[l].norm();
        return s;
    }

    FFTSO3_matrix<double> real() const
    {
        FFTSO3_matrix<double> R(l_max);
        for (long i = 0; i < size(); i++) R.v[i] = detail::real_part(v[i]);
        return R;
    }

    template <class U>
    auto operator+(const FFTSO3_matrix<U> &o) const -> FFTSO3_matrix<decltype(ScalarType() + U())>
    {
        assert(o.l_max == l_max);
        FFTSO3_matrix<decltype(ScalarType() + U())> R(l_max);
        for (long i = 0; i < size(); i++) R.v[i] = v[i] + o.v[i];
        return R;
    }
    template <class U>
    auto operator-(const FFTSO3_matrix<U> &o) const -> FFTSO3_matrix<decltype(ScalarType() - U())>
    {
        assert(o.l_max == l_max);
        FFTSO3_matrix<decltype(ScalarType() - U())> R(l_max);
        for (long i = 0; i < size(); i++) R.v[i] = v[i] - o.v[i];
        return R;
    }
    FFTSO3_matrix<complex<double>> operator*(const complex<double> &c) const
    {
        FFTSO3_matrix<complex<double>> R(l_max);
        for (long i = 0; i < size(); i++) R.v[i] = v[i] * c;
        return R;
    }
    FFTSO3_matrix<ScalarType> operator*(const double &c) const
    {
        FFTSO3_matrix<ScalarType> R(l_max);
        for (long i = 0; i < size(); i++) R.v[i] = v[i] * c;
        return R;
    }

    friend std::ostream &operator<<(std::ostream &os, const FFTSO3_matrix &M)
    {
        for (int l = 0; l <= M.l_max; l++) {
            os << "l=" << l << std::endl;
            const auto b = M[l];
            for (int i = 0; i < b.rows(); i++) {
                for (int j = 0; j < b.cols(); j++) os << b(i, j) << (j + 1 < b.cols() ? " " : "");
                os << std::endl;
            }
            os << std::endl;
        }
        return os;
    }

private:
    void assert_index(int l) const { assert(l >= 0 && l <= l_max); (void)l; }
    void assert_index(int l, int m, int n) const
    {
        assert(l >= 0 && l <= l_max && std::abs(m) <= l && std::abs(n) <= l);
        (void)l; (void)m; (void)n;
    }
};

typedef FFTSO3_matrix<double> FFTSO3_matrix_real;
typedef FFTSO3_matrix<complex<double>> FFTSO3_matrix_complex;

}  // namespace fdcl

#endif
