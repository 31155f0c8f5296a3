// qcb_su3.hpp -- register-resident SU(3) algebra for gfx950 kernels.
//
// A 3x3 complex matrix is 18 reals in registers, row-major, (re,im)
// interleaved: m.v[(3*i+j)*2 + {0,1}], the order of the reference's Eigen
// RowMajor SU3 (src/su3/utils.h:12).  T is float or double; every function is
// written for fused multiply-adds on 64-lane wavefronts (no shared state).
#pragma once
#include <hip/hip_runtime.h>

namespace qcb {

template <class T>
struct M3 {
    T v[18];
    __device__ __forceinline__ T& re(int i, int j) { return v[(3 * i + j) * 2]; }
    __device__ __forceinline__ T& im(int i, int j) { return v[(3 * i + j) * 2 + 1]; }
    __device__ __forceinline__ T re(int i, int j) const { return v[(3 * i + j) * 2]; }
    __device__ __forceinline__ T im(int i, int j) const { return v[(3 * i + j) * 2 + 1]; }
};

template <class T>
__device__ __forceinline__ void m3_zero(M3<T>& a) {
#pragma unroll
    for (int k = 0; k < 18; ++k) a.v[k] = T(0);
}

template <class T>
__device__ __forceinline__ void m3_identity(M3<T>& a) {
    m3_zero(a);
    a.v[0] = a.v[8] = a.v[16] = T(1);
}

// c = a * b
template <class T>
__device__ __forceinline__ M3<T> mul_nn(const M3<T>& a, const M3<T>& b) {
    M3<T> c;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            T r = a.re(i, 0) * b.re(0, j);
            T s = a.re(i, 0) * b.im(0, j);
            r = fma(-a.im(i, 0), b.im(0, j), r);
            s = fma(a.im(i, 0), b.re(0, j), s);
#pragma unroll
            for (int k = 1; k < 3; ++k) {
                r = fma(a.re(i, k), b.re(k, j), r);
                r = fma(-a.im(i, k), b.im(k, j), r);
                s = fma(a.re(i, k), b.im(k, j), s);
                s = fma(a.im(i, k), b.re(k, j), s);
            }
            c.re(i, j) = r;
            c.im(i, j) = s;
        }
    return c;
}

// c = a * b^dag
template <class T>
__device__ __forceinline__ M3<T> mul_nd(const M3<T>& a, const M3<T>& b) {
    M3<T> c;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            // sum_k a(i,k) conj(b(j,k))
            T r = a.re(i, 0) * b.re(j, 0);
            T s = a.im(i, 0) * b.re(j, 0);
            r = fma(a.im(i, 0), b.im(j, 0), r);
            s = fma(-a.re(i, 0), b.im(j, 0), s);
#pragma unroll
            for (int k = 1; k < 3; ++k) {
                r = fma(a.re(i, k), b.re(j, k), r);
                r = fma(a.im(i, k), b.im(j, k), r);
                s = fma(a.im(i, k), b.re(j, k), s);
                s = fma(-a.re(i, k), b.im(j, k), s);
            }
            c.re(i, j) = r;
            c.im(i, j) = s;
        }
    return c;
}

// c = a^dag * b
template <class T>
__device__ __forceinline__ M3<T> mul_dn(const M3<T>& a, const M3<T>& b) {
    M3<T> c;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            // sum_k conj(a(k,i)) b(k,j)
            T r = a.re(0, i) * b.re(0, j);
            T s = a.re(0, i) * b.im(0, j);
            r = fma(a.im(0, i), b.im(0, j), r);
            s = fma(-a.im(0, i), b.re(0, j), s);
#pragma unroll
            for (int k = 1; k < 3; ++k) {
                r = fma(a.re(k, i), b.re(k, j), r);
                r = fma(a.im(k, i), b.im(k, j), r);
                s = fma(a.re(k, i), b.im(k, j), s);
                s = fma(-a.im(k, i), b.re(k, j), s);
            }
            c.re(i, j) = r;
            c.im(i, j) = s;
        }
    return c;
}

// acc += a * b^dag   (fused accumulate, saves a temporary)
template <class T>
__device__ __forceinline__ void acc_nd(M3<T>& acc, const M3<T>& a, const M3<T>& b) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            T r = acc.re(i, j), s = acc.im(i, j);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                r = fma(a.re(i, k), b.re(j, k), r);
                r = fma(a.im(i, k), b.im(j, k), r);
                s = fma(a.im(i, k), b.re(j, k), s);
                s = fma(-a.re(i, k), b.im(j, k), s);
            }
            acc.re(i, j) = r;
            acc.im(i, j) = s;
        }
}

// acc += a^dag * b
template <class T>
__device__ __forceinline__ void acc_dn(M3<T>& acc, const M3<T>& a, const M3<T>& b) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            T r = acc.re(i, j), s = acc.im(i, j);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                r = fma(a.re(k, i), b.re(k, j), r);
                r = fma(a.im(k, i), b.im(k, j), r);
                s = fma(a.re(k, i), b.im(k, j), s);
                s = fma(-a.im(k, i), b.re(k, j), s);
            }
            acc.re(i, j) = r;
            acc.im(i, j) = s;
        }
}

template <class T>
__device__ __forceinline__ T retrace(const M3<T>& a) {
    return a.v[0] + a.v[8] + a.v[16];
}

// Re Tr(a * b) without forming the product
template <class T>
__device__ __forceinline__ T retrace_nn(const M3<T>& a, const M3<T>& b) {
    T s = T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            s = fma(a.re(i, k), b.re(k, i), s);
            s = fma(-a.im(i, k), b.im(k, i), s);
        }
    return s;
}

// Re Tr(a * b^dag)
template <class T>
__device__ __forceinline__ T retrace_nd(const M3<T>& a, const M3<T>& b) {
    T s = T(0);
#pragma unroll
    for (int k = 0; k < 18; ++k) s = fma(a.v[k], b.v[k], s);
    return s;
}

// Traceless anti-Hermitian projection P_TA(A) = (A - A^dag)/2 - Tr(.)/3
// (reference src/su3/utils.cpp:130-134).
template <class T>
__device__ __forceinline__ void proj_lie(M3<T>& a) {
    M3<T> b;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            b.re(i, j) = T(0.5) * (a.re(i, j) - a.re(j, i));
            b.im(i, j) = T(0.5) * (a.im(i, j) + a.im(j, i));
        }
    T tr = (b.im(0, 0) + b.im(1, 1) + b.im(2, 2)) / T(3);
    // real parts of the diagonal of (A - A^dag)/2 are exactly zero
    T trr = (b.re(0, 0) + b.re(1, 1) + b.re(2, 2)) / T(3);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        b.re(i, i) -= trr;
        b.im(i, i) -= tr;
    }
    a = b;
}

// Taylor coefficients 1/n! built by successive division, as the reference does.
template <class T, int N>
struct InvFact {
    T c[N + 1];
    constexpr InvFact() : c() {
        double v = 1.0;
        c[0] = T(1);
        for (int n = 1; n <= N; ++n) {
            v = v / double(n);
            c[n] = T(v);
        }
    }
};

// exp(X) for X in su(3) by Luescher's Cayley-Hamilton Horner scheme
// (reference src/su3/utils.cpp:145-193): with t = -Re Tr X^2/2 and
// d = Re(i det X), q(n) = c_n - i d q2, q1 <- q0 - t q2, q2 <- q1 from n = N-1
// down to 0, and exp X = q0 + q1 X + q2 X^2.  N = 18 in double; float needs
// fewer terms for the same relative accuracy (||X|| <= 1 in the flow).
template <class T, int N>
__device__ __forceinline__ M3<T> exp_su3(const M3<T>& X) {
    M3<T> X2 = mul_nn(X, X);
    T t = T(-0.5) * retrace(X2);
    // Im det X by cofactors along row 0
    T c0r = X.re(1, 1) * X.re(2, 2) - X.im(1, 1) * X.im(2, 2) - (X.re(1, 2) * X.re(2, 1) - X.im(1, 2) * X.im(2, 1));
    T c0i = X.re(1, 1) * X.im(2, 2) + X.im(1, 1) * X.re(2, 2) - (X.re(1, 2) * X.im(2, 1) + X.im(1, 2) * X.re(2, 1));
    T c1r = X.re(1, 2) * X.re(2, 0) - X.im(1, 2) * X.im(2, 0) - (X.re(1, 0) * X.re(2, 2) - X.im(1, 0) * X.im(2, 2));
    T c1i = X.re(1, 2) * X.im(2, 0) + X.im(1, 2) * X.re(2, 0) - (X.re(1, 0) * X.im(2, 2) + X.im(1, 0) * X.re(2, 2));
    T c2r = X.re(1, 0) * X.re(2, 1) - X.im(1, 0) * X.im(2, 1) - (X.re(1, 1) * X.re(2, 0) - X.im(1, 1) * X.im(2, 0));
    T c2i = X.re(1, 0) * X.im(2, 1) + X.im(1, 0) * X.re(2, 1) - (X.re(1, 1) * X.im(2, 0) + X.im(1, 1) * X.re(2, 0));
    T det_im = X.re(0, 0) * c0i + X.im(0, 0) * c0r + X.re(0, 1) * c1i + X.im(0, 1) * c1r + X.re(0, 2) * c2i +
               X.im(0, 2) * c2r;
    T d = -det_im;
    constexpr InvFact<T, N> cf;
    T q0r = cf.c[N], q0i = T(0), q1r = T(0), q1i = T(0), q2r = T(0), q2i = T(0);
#pragma unroll
    for (int n = N - 1; n >= 0; --n) {
        T n0r = fma(d, q2i, cf.c[n]);
        T n0i = -d * q2r;
        T n1r = fma(-t, q2r, q0r);
        T n1i = fma(-t, q2i, q0i);
        q2r = q1r;
        q2i = q1i;
        q1r = n1r;
        q1i = n1i;
        q0r = n0r;
        q0i = n0i;
    }
    M3<T> e;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        T xr = X.v[2 * k], xi = X.v[2 * k + 1], yr = X2.v[2 * k], yi = X2.v[2 * k + 1];
        T r = q1r * xr - q1i * xi + (q2r * yr - q2i * yi);
        T s = q1r * xi + q1i * xr + (q2r * yi + q2i * yr);
        e.v[2 * k] = r;
        e.v[2 * k + 1] = s;
    }
    e.v[0] += q0r; e.v[1] += q0i;
    e.v[8] += q0r; e.v[9] += q0i;
    e.v[16] += q0r; e.v[17] += q0i;
    return e;
}

template <class T> struct ExpTerms;
template <> struct ExpTerms<double> { static constexpr int N = 18; };
template <> struct ExpTerms<float> { static constexpr int N = 12; };

// Gram-Schmidt on columns, c2 = conj(c0 x c1), all in double
// (reference src/gauge/GaugeField.cpp:36-58).
__device__ __forceinline__ void reunitarize_d(M3<double>& u) {
    double n0 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) n0 += u.re(i, 0) * u.re(i, 0) + u.im(i, 0) * u.im(i, 0);
    n0 = sqrt(n0);
#pragma unroll
    for (int i = 0; i < 3; ++i) { u.re(i, 0) /= n0; u.im(i, 0) /= n0; }
    double dr = 0.0, di = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        dr += u.re(i, 0) * u.re(i, 1) + u.im(i, 0) * u.im(i, 1);
        di += u.re(i, 0) * u.im(i, 1) - u.im(i, 0) * u.re(i, 1);
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        u.re(i, 1) -= u.re(i, 0) * dr - u.im(i, 0) * di;
        u.im(i, 1) -= u.re(i, 0) * di + u.im(i, 0) * dr;
    }
    double n1 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) n1 += u.re(i, 1) * u.re(i, 1) + u.im(i, 1) * u.im(i, 1);
    n1 = sqrt(n1);
#pragma unroll
    for (int i = 0; i < 3; ++i) { u.re(i, 1) /= n1; u.im(i, 1) /= n1; }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int a = (i + 1) % 3, b = (i + 2) % 3;
        double re = (u.re(a, 0) * u.re(b, 1) - u.im(a, 0) * u.im(b, 1)) - (u.re(b, 0) * u.re(a, 1) - u.im(b, 0) * u.im(a, 1));
        double im = (u.re(a, 0) * u.im(b, 1) + u.im(a, 0) * u.re(b, 1)) - (u.re(b, 0) * u.im(a, 1) + u.im(b, 0) * u.re(a, 1));
        u.re(i, 2) = re;
        u.im(i, 2) = -im;
    }
}

}  // namespace qcb
