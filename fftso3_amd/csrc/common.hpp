// common.hpp -- device/host helpers shared by the SO(3) and S^2 transform kernels (gfx950).
//
// Index conventions (reference fdcl-gwu/FFTSO3):
//   SO(3) coefficient (l,m,n) lives at off(l) + (m+l) + (n+l)(2l+1), off(l) = l(4l^2-1)/3,
//   the column-major (2l+1)^2 block of FFTSO3_matrix (src/fdcl_FFTSO3_matrix.cpp:43-47).
//   S^2 coefficient (l,m) lives at l^2 + m + l (src/fdcl_FFTS2_matrix.cpp:43-47).
//
// Symmetry groups.  Wigner d obeys d(l,m,n) = (-1)^(m+n) d(l,-m,-n) = d(l,-n,-m)
// = (-1)^(m+n) d(l,n,m) (src/fdcl_FFTSO3.cpp:174-190), so the four (m,n) pairs
//   p0 = (m0,n0), p1 = (-m0,-n0), p2 = (-n0,-m0), p3 = (n0,m0)
// of a canonical (m0,n0) with m0 >= |n0| share one beta-row up to the sign s_p
// (s0 = s2 = 1, s1 = s3 = (-1)^(m0+n0)).  Canonical group g = m0^2 + (n0 + m0), g < B^2.
// For S^2 the group is m0 >= 0 with pairs (m0, -m0) and nP(l,-m) = (-1)^m nP(l,m)
// (src/fdcl_FFTS2.cpp:96-98).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define FFTSO3_HD __host__ __device__ __forceinline__

template <typename T> struct cplx_t;
template <> struct cplx_t<float> { typedef float2 type; };
template <> struct cplx_t<double> { typedef double2 type; };

template <typename T2> FFTSO3_HD T2 cmul(T2 a, T2 b)
{
    T2 r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <typename T2> FFTSO3_HD T2 cadd(T2 a, T2 b) { T2 r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename T2> FFTSO3_HD T2 csub(T2 a, T2 b) { T2 r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename T2> FFTSO3_HD T2 cmk(decltype(T2::x) x, decltype(T2::x) y) { T2 r; r.x = x; r.y = y; return r; }

FFTSO3_HD int so3_off(int l) { return l * (4 * l * l - 1) / 3; }
FFTSO3_HD int so3_idx(int l, int m, int n) { return so3_off(l) + (m + l) + (n + l) * (2 * l + 1); }
FFTSO3_HD long so3_ncoef(int B) { return (long)B * (4L * B * B - 1) / 3; }
FFTSO3_HD int s2_idx(int l, int m) { return l * l + m + l; }
FFTSO3_HD int iabs(int x) { return x < 0 ? -x : x; }
FFTSO3_HD int wrapN(int u, int N) { u %= N; return u < 0 ? u + N : u; }
FFTSO3_HD int parity_sign(int e) { return (e & 1) ? -1 : 1; }

// canonical group g -> (m0, n0), g = m0^2 + n0 + m0
FFTSO3_HD void so3_group(int g, int &m0, int &n0)
{
    int m = (int)sqrtf((float)g + 0.5f);
    while (m * m > g) m--;
    while ((m + 1) * (m + 1) <= g) m++;
    m0 = m;
    n0 = g - m * m - m;
}

// pair p of canonical (m0, n0)
FFTSO3_HD void so3_pair(int m0, int n0, int p, int &m, int &n)
{
    switch (p) {
    case 0: m = m0; n = n0; break;
    case 1: m = -m0; n = -n0; break;
    case 2: m = -n0; n = -m0; break;
    default: m = n0; n = m0; break;
    }
}

// storage slot (g, p) that holds (m, n); every (m,n) has exactly one slot
FFTSO3_HD void so3_slot(int m, int n, int &g, int &p)
{
    int am = iabs(m), an = iabs(n), m0, n0;
    if (m >= an) { m0 = m; n0 = n; p = 0; }
    else if (-m >= an) { m0 = -m; n0 = -n; p = 1; }
    else if (-n > am) { m0 = -n; n0 = -m; p = 2; }
    else { m0 = n; n0 = m; p = 3; }
    g = m0 * m0 + n0 + m0;
}

// pair p of group g is stored in its own slot (false for the duplicated pairs of orbits of size 1 or 2)
FFTSO3_HD bool so3_pair_valid(int m0, int n0, int p)
{
    if (p == 0) return true;
    if (m0 == 0) return false;                       // (0,0): orbit of size 1
    if (p == 1) return true;
    return !(n0 == m0 || n0 == -m0);                 // orbits of size 2 use p0, p1 only
}

FFTSO3_HD int so3_pair_sign(int m0, int n0, int p) { return (p & 1) ? parity_sign(m0 + n0) : 1; }

// Unitary T^l of src/fdcl_FFTSO3.cpp:1107-1139 (also src/fdcl_FFTS2.cpp:324-369):
//   m<0: T(m,m) = i/sqrt2, T(m,-m) = -i(-1)^m/sqrt2;  T(0,0) = 1;
//   m>0: T(m,-m) = 1/sqrt2, T(m,m) = (-1)^m/sqrt2.   Zero elsewhere.
template <typename T2> FFTSO3_HD T2 tmat(int m, int a)
{
    typedef decltype(T2::x) T;
    const T r2 = (T)0.70710678118654752440084436210485;
    if (m == 0) return cmk<T2>(a == 0 ? (T)1 : (T)0, (T)0);
    if (m < 0) {
        if (a == m) return cmk<T2>((T)0, r2);
        if (a == -m) return cmk<T2>((T)0, -(T)parity_sign(m) * r2);
        return cmk<T2>((T)0, (T)0);
    }
    if (a == -m) return cmk<T2>(r2, (T)0);
    if (a == m) return cmk<T2>((T)parity_sign(m) * r2, (T)0);
    return cmk<T2>((T)0, (T)0);
}
