// dual.hpp — forward-mode dual number for the per-cell Jacobian.
//
// Dual<N> carries a value and N partial derivatives in registers.  The
// propagation rules are the ones of the reference's DenseAd::Evaluation
// (ad/include/ad/Evaluation.hpp:389-470 product/quotient, :1100-1190 scalar
// mixes; ad/include/ad/Math.hpp:139-145 abs, :505-690 cos/acos/sqrt/exp/pow,
// :816-835 log) so that values and derivatives agree with the reference's AD
// path; in particular a Dual divided by a plain double multiplies by the
// reciprocal, as the reference does.
#pragma once

#include <hip/hip_runtime.h>

namespace nn
{

// Dual<0> carries no derivatives: it evaluates the AD code path for values
// only, with the same rounding as the AD evaluation (e.g. Dual / double is a
// multiplication by the reciprocal, not a division).
template <int N>
struct Dual
{
    static constexpr int size = N;
    double v;
    double d[N > 0 ? N : 1];

    __host__ __device__ Dual() : v(0.0)
    {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = 0.0;
    }
    __host__ __device__ Dual(double value) : v(value)
    {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = 0.0;
    }
    __host__ __device__ static Dual variable(double value, int index)
    {
        Dual r(value);
        r.d[index] = 1.0;
        return r;
    }
    __host__ __device__ void clearDerivatives()
    {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = 0.0;
    }

    __host__ __device__ Dual &operator+=(const Dual &o)
    {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] += o.d[i];
        v += o.v;
        return *this;
    }
    __host__ __device__ Dual &operator+=(double o)
    {
        v += o;
        return *this;
    }
    __host__ __device__ Dual &operator-=(const Dual &o)
    {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] -= o.d[i];
        v -= o.v;
        return *this;
    }
    __host__ __device__ Dual &operator-=(double o)
    {
        v -= o;
        return *this;
    }
    __host__ __device__ Dual &operator*=(const Dual &o)
    {
        const double u = v, w = o.v;
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = d[i] * w + o.d[i] * u;
        v = u * w;
        return *this;
    }
    __host__ __device__ Dual &operator*=(double s)
    {
        v *= s;
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] *= s;
        return *this;
    }
    __host__ __device__ Dual &operator/=(const Dual &o)
    {
        const double u = v, w = o.v, den = w * w;
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = (w * d[i] - u * o.d[i]) / den;
        v = u / w;
        return *this;
    }
    __host__ __device__ Dual &operator/=(double s)
    {
        const double r = 1.0 / s;
        v *= r;
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] *= r;
        return *this;
    }
    __host__ __device__ Dual operator-() const
    {
        Dual r;
        r.v = -v;
#pragma unroll
        for (int i = 0; i < N; ++i) r.d[i] = -d[i];
        return r;
    }
};

template <int N> __host__ __device__ inline Dual<N> operator+(Dual<N> a, const Dual<N> &b) { a += b; return a; }
template <int N> __host__ __device__ inline Dual<N> operator+(Dual<N> a, double b) { a += b; return a; }
template <int N> __host__ __device__ inline Dual<N> operator+(double a, Dual<N> b) { b += a; return b; }
template <int N> __host__ __device__ inline Dual<N> operator-(Dual<N> a, const Dual<N> &b) { a -= b; return a; }
template <int N> __host__ __device__ inline Dual<N> operator-(Dual<N> a, double b) { a -= b; return a; }
// scalar - Dual = -(Dual - scalar)
template <int N> __host__ __device__ inline Dual<N> operator-(double a, const Dual<N> &b) { return -(b - a); }
template <int N> __host__ __device__ inline Dual<N> operator*(Dual<N> a, const Dual<N> &b) { a *= b; return a; }
template <int N> __host__ __device__ inline Dual<N> operator*(Dual<N> a, double b) { a *= b; return a; }
template <int N> __host__ __device__ inline Dual<N> operator*(double a, Dual<N> b) { b *= a; return b; }
template <int N> __host__ __device__ inline Dual<N> operator/(Dual<N> a, const Dual<N> &b) { a /= b; return a; }
template <int N> __host__ __device__ inline Dual<N> operator/(Dual<N> a, double b) { a /= b; return a; }
// scalar / Dual: constant Dual divided by Dual.
template <int N> __host__ __device__ inline Dual<N> operator/(double a, const Dual<N> &b)
{
    Dual<N> r(a);
    r /= b;
    return r;
}

__host__ __device__ inline double sval(double x) { return x; }
template <int N> __host__ __device__ inline double sval(const Dual<N> &x) { return x.v; }

__host__ __device__ inline void clearDeriv(double &) {}
template <int N> __host__ __device__ inline void clearDeriv(Dual<N> &x) { x.clearDerivatives(); }

// ---- plain-double overloads (so that templated code finds one name set) ----
__host__ __device__ inline double abs(double x) { return ::fabs(x); }
__host__ __device__ inline double sqrt(double x) { return ::sqrt(x); }
__host__ __device__ inline double exp(double x) { return ::exp(x); }
__host__ __device__ inline double log(double x) { return ::log(x); }
__host__ __device__ inline double cos(double x) { return ::cos(x); }
__host__ __device__ inline double acos(double x) { return ::acos(x); }
__host__ __device__ inline double pow(double x, double e) { return ::pow(x, e); }

// ---- elementary functions (value first, then the chain-rule factor) ----
template <int N> __host__ __device__ inline Dual<N> abs(const Dual<N> &x) { return (x.v >= 0.0) ? x : -x; }

template <int N> __host__ __device__ inline Dual<N> sqrt(const Dual<N> &x)
{
    Dual<N> r;
    const double s = ::sqrt(x.v);
    r.v = s;
    const double f = 0.5 / s;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = f * x.d[i];
    return r;
}
template <int N> __host__ __device__ inline Dual<N> exp(const Dual<N> &x)
{
    Dual<N> r;
    const double e = ::exp(x.v);
    r.v = e;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = e * x.d[i];
    return r;
}
template <int N> __host__ __device__ inline Dual<N> log(const Dual<N> &x)
{
    Dual<N> r;
    r.v = ::log(x.v);
    const double f = 1.0 / x.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = f * x.d[i];
    return r;
}
template <int N> __host__ __device__ inline Dual<N> cos(const Dual<N> &x)
{
    Dual<N> r;
    r.v = ::cos(x.v);
    const double f = -::sin(x.v);
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = f * x.d[i];
    return r;
}
template <int N> __host__ __device__ inline Dual<N> acos(const Dual<N> &x)
{
    Dual<N> r;
    r.v = ::acos(x.v);
    const double f = -1.0 / ::sqrt(1.0 - x.v * x.v);
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = f * x.d[i];
    return r;
}
template <int N> __host__ __device__ inline Dual<N> pow(const Dual<N> &x, double e)
{
    Dual<N> r;
    r.v = ::pow(x.v, e);
    if (e == 0.0)
        return r;
    const double f = e * ::pow(x.v, e - 1.0);
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = f * x.d[i];
    return r;
}

} // namespace nn
