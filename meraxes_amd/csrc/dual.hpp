// dual.hpp - forward-mode dual numbers for the per-cell adjoints.
//
// Dual<K> carries a primal double and K directional derivatives.  Every
// operation computes the primal exactly as the same expression in double does
// (one IEEE operation per operation), so a template instantiated with Dual<K>
// reproduces the double instantiation's primal values bit for bit and adds the
// derivatives.  Branches of templated code test pval(x), the primal.
#pragma once
#include <hip/hip_runtime.h>

namespace mrx {

template <int K>
struct Dual {
  double v;
  double d[K];
  __device__ __forceinline__ Dual() : v(0.0)
  {
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] = 0.0;
  }
  __device__ __forceinline__ Dual(double x) : v(x)
  {
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] = 0.0;
  }
  // a variable: primal x, unit derivative along direction `dir`
  __device__ __forceinline__ static Dual var(double x, int dir)
  {
    Dual r(x);
    r.d[dir] = 1.0;
    return r;
  }
  __device__ __forceinline__ Dual& operator+=(const Dual& b);
  __device__ __forceinline__ Dual& operator-=(const Dual& b);
  __device__ __forceinline__ Dual& operator*=(const Dual& b);
  __device__ __forceinline__ Dual& operator/=(const Dual& b);
};

template <int K>
__device__ __forceinline__ double pval(const Dual<K>& x)
{
  return x.v;
}

template <int K>
__device__ __forceinline__ Dual<K> operator+(const Dual<K>& a, const Dual<K>& b)
{
  Dual<K> r(a.v + b.v);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a, const Dual<K>& b)
{
  Dual<K> r(a.v - b.v);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a)
{
  Dual<K> r(-a.v);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(const Dual<K>& a, const Dual<K>& b)
{
  Dual<K> r(a.v * b.v);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator/(const Dual<K>& a, const Dual<K>& b)
{
  Dual<K> r(a.v / b.v);
  const double ib = 1.0 / b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
  return r;
}
// mixed forms keep the double operand's position, so the primal rounds as the
// double expression does
template <int K>
__device__ __forceinline__ Dual<K> operator+(const Dual<K>& a, double b)
{
  Dual<K> r = a;
  r.v = a.v + b;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator+(double a, const Dual<K>& b)
{
  Dual<K> r = b;
  r.v = a + b.v;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a, double b)
{
  Dual<K> r = a;
  r.v = a.v - b;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(double a, const Dual<K>& b)
{
  Dual<K> r(a - b.v);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = -b.d[k];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(const Dual<K>& a, double b)
{
  Dual<K> r(a.v * b);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(double a, const Dual<K>& b)
{
  Dual<K> r(a * b.v);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a * b.d[k];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator/(const Dual<K>& a, double b)
{
  Dual<K> r(a.v / b);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] / b;
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator/(double a, const Dual<K>& b)
{
  Dual<K> r(a / b.v);
  const double g = -r.v / b.v;  // d(a/b)/db
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = g * b.d[k];
  return r;
}

template <int K>
__device__ __forceinline__ Dual<K>& Dual<K>::operator+=(const Dual<K>& b)
{
  *this = *this + b;
  return *this;
}
template <int K>
__device__ __forceinline__ Dual<K>& Dual<K>::operator-=(const Dual<K>& b)
{
  *this = *this - b;
  return *this;
}
template <int K>
__device__ __forceinline__ Dual<K>& Dual<K>::operator*=(const Dual<K>& b)
{
  *this = *this * b;
  return *this;
}
template <int K>
__device__ __forceinline__ Dual<K>& Dual<K>::operator/=(const Dual<K>& b)
{
  *this = *this / b;
  return *this;
}

template <int K>
__device__ __forceinline__ Dual<K> chain(const Dual<K>& a, double f, double dfda)
{
  Dual<K> r(f);
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = dfda * a.d[k];
  return r;
}
template <int K>
__device__ __forceinline__ Dual<K> log(const Dual<K>& a)
{
  return chain(a, ::log(a.v), 1.0 / a.v);
}
template <int K>
__device__ __forceinline__ Dual<K> exp(const Dual<K>& a)
{
  const double e = ::exp(a.v);
  return chain(a, e, e);
}
template <int K>
__device__ __forceinline__ Dual<K> pow(const Dual<K>& a, double p)
{
  const double f = ::pow(a.v, p);
  return chain(a, f, p * ::pow(a.v, p - 1.0));
}
template <int K>
__device__ __forceinline__ Dual<K> pow(const Dual<K>& a, int p)
{
  return pow(a, (double)p);
}
template <int K>
__device__ __forceinline__ Dual<K> fabs(const Dual<K>& a)
{
  return a.v < 0.0 ? -a : a;
}
// rounding to float: the primal is rounded, the derivatives pass through
template <int K>
__device__ __forceinline__ Dual<K> rnd_f(const Dual<K>& a)
{
  Dual<K> r = a;
  r.v = (double)(float)a.v;
  return r;
}
__device__ __forceinline__ double rnd_f(double a) { return (double)(float)a; }

}  // namespace mrx
