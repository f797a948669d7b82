// fdcl_Clebsch_Gordon.hpp -- Clebsch-Gordan coefficients of the complex and the real harmonics on
// SO(3) (host, FP64), with the accessors and the storage order of the reference's
// include/fdcl_Clebsch_Gordon.hpp:31-81.
//
// For one pair (l1, l2) the coefficients C^{l,m}_{l1,m1,l2,m2}, |l1-l2| <= l <= l1+l2, live in a
// (2l1+1)(2l2+1) x (2l1+1)(2l2+1) matrix: row (l1+m1)(2l2+1) + l2+m2 (the product state), column
// l^2 - (l2-l1)^2 + l+m (the coupled state), as src/fdcl_Clebsch_Gordon.cpp:18-26 lays them out,
// so code that reads the matrix members directly keeps working.
#pragma once

#include <complex>
#include <vector>

namespace fdcl
{

// dense column-major matrix: the part of Eigen::Matrix the Clebsch-Gordan members expose
template <typename T> class DenseMatrix
{
public:
    void resize(int r, int c)
    {
        rows_ = r;
        cols_ = c;
        v_.assign((size_t)r * c, T(0));
    }
    void setZero() { v_.assign(v_.size(), T(0)); }
    int rows() const { return rows_; }
    int cols() const { return cols_; }
    T &operator()(int i, int j) { return v_[(size_t)j * rows_ + i]; }
    const T &operator()(int i, int j) const { return v_[(size_t)j * rows_ + i]; }
    T *data() { return v_.data(); }
    const T *data() const { return v_.data(); }

private:
    int rows_ = 0, cols_ = 0;
    std::vector<T> v_;
};

// Clebsch-Gordan coefficients <l1 m1 l2 m2 | l m> of the complex harmonics (Condon-Shortley phase)
class Clebsch_Gordon_complex
{
public:
    Clebsch_Gordon_complex() {}
    Clebsch_Gordon_complex(int l1, int l2);
    virtual ~Clebsch_Gordon_complex() {}
    void init(int l1, int l2);  // sizes and zeroes the buffer for (l1, l2)
    int l1 = 0, l2 = 0;
    DenseMatrix<double> C;

    // C^{l,m}_{l1,m1,l2,m2}; the arguments must satisfy |l1-l2| <= l <= l1+l2, |m| <= l,
    // |m1| <= l1, |m2| <= l2 for the (l1, l2) last passed to compute() (std::out_of_range otherwise)
    double &operator()(int l, int m, int l1, int m1, int l2, int m2);
    void compute(int l1, int l2);  // all coefficients of (l1, l2); others are not kept

protected:
    int row(int l1, int m1, int l2, int m2) const { return (l1 + m1) * (2 * l2 + 1) + l2 + m2; }
    int col(int l, int m, int l1, int l2) const { return l * l - (l2 - l1) * (l2 - l1) + l + m; }
    void check_index(int l, int m, int l1, int m1, int l2, int m2) const;
};

// Coefficients c^{l,m}_{l1,m1,l2,m2} of the real harmonics U^l = conj(T^l) D^l (T^l)^T:
// U^{l1}_{m1 n1} U^{l2}_{m2 n2} = sum_{l,m,n} Re(c^{l,m}_{l1,m1,l2,m2} conj(c^{l,n}_{l1,n1,l2,n2})) U^l_{mn}
// (src/fdcl_FFTSO3.cpp:1685-1768 checks this identity)
class Clebsch_Gordon_real : public Clebsch_Gordon_complex
{
public:
    Clebsch_Gordon_real() {}
    Clebsch_Gordon_real(int l1, int l2);
    void init(int l1, int l2);
    DenseMatrix<std::complex<double>> c;

    std::complex<double> &operator()(int l, int m, int l1, int m1, int l2, int m2);
    void compute(int l1, int l2);
};

}  // namespace fdcl
