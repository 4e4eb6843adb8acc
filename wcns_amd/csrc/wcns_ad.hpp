// wcns_ad.hpp -- forward-mode dual numbers for the tangent and adjoint kernels.
//
// D1 carries a value and one directional derivative.  The value part of every
// operation is the plain double operation, in the same order, so a kernel
// instantiated on D1 computes bit-for-bit the values of its double instance
// (the build disables FP contraction) and, beside them, the exact derivative
// of that computation along the seeded direction.
#pragma once
#include <hip/hip_runtime.h>

namespace wcns {

// keep the double overloads visible beside the D1 ones below
using ::sqrt;
using ::fabs;
using ::pow;

struct D1 {
    double v, d;
    D1() = default;
    __host__ __device__ constexpr D1(double x) : v(x), d(0.0) {}
    __host__ __device__ constexpr D1(double x, double y) : v(x), d(y) {}
};

__device__ inline D1 operator+(D1 a, D1 b) { return D1(a.v + b.v, a.d + b.d); }
__device__ inline D1 operator-(D1 a, D1 b) { return D1(a.v - b.v, a.d - b.d); }
__device__ inline D1 operator-(D1 a) { return D1(-a.v, -a.d); }
__device__ inline D1 operator*(D1 a, D1 b) { return D1(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ inline D1 operator/(D1 a, D1 b)
{
    double q = a.v / b.v;
    return D1(q, (a.d - q * b.d) / b.v);
}
__device__ inline D1 operator+(D1 a, double b) { return D1(a.v + b, a.d); }
__device__ inline D1 operator+(double a, D1 b) { return D1(a + b.v, b.d); }
__device__ inline D1 operator-(D1 a, double b) { return D1(a.v - b, a.d); }
__device__ inline D1 operator-(double a, D1 b) { return D1(a - b.v, -b.d); }
__device__ inline D1 operator*(D1 a, double b) { return D1(a.v * b, a.d * b); }
__device__ inline D1 operator*(double a, D1 b) { return D1(a * b.v, a * b.d); }
__device__ inline D1 operator/(D1 a, double b) { return D1(a.v / b, a.d / b); }
__device__ inline D1 operator/(double a, D1 b)
{
    double q = a / b.v;
    return D1(q, -q * b.d / b.v);
}
__device__ inline bool operator<(D1 a, D1 b) { return a.v < b.v; }
__device__ inline bool operator<=(D1 a, D1 b) { return a.v <= b.v; }
__device__ inline bool operator>(D1 a, D1 b) { return a.v > b.v; }
__device__ inline bool operator>=(D1 a, D1 b) { return a.v >= b.v; }
__device__ inline bool operator<(D1 a, double b) { return a.v < b; }
__device__ inline bool operator<=(D1 a, double b) { return a.v <= b; }
__device__ inline bool operator>(D1 a, double b) { return a.v > b; }
__device__ inline bool operator>=(D1 a, double b) { return a.v >= b; }
__device__ inline D1 sqrt(D1 a)
{
    double s = ::sqrt(a.v);
    return D1(s, a.d / (2.0 * s));
}
__device__ inline D1 fabs(D1 a) { return a.v < 0.0 ? -a : a; }
// a^e for a constant exponent (the far-field boundary's isentropic relations)
__device__ inline D1 pow(D1 a, double e)
{
    double p = ::pow(a.v, e);
    return D1(p, e * (p / a.v) * a.d);
}

// value part of a double or a D1
__device__ inline double val(double a) { return a; }
__device__ inline double val(D1 a) { return a.v; }

// Colored seeding of a banded operator: along a pencil, row r of the residual
// depends on nodes r-5..r+5 only, so nodes whose index differs by a multiple
// of NCOLOR = 11 never meet in one row.  Seeding every node of one colour at
// once, the derivative found in row r belongs to the one seeded node within
// reach, node ad_node(r, colour).
constexpr int NCOLOR = 11;
__host__ __device__ inline int ad_mod(int l) { return ((l % NCOLOR) + NCOLOR) % NCOLOR; }
__host__ __device__ inline int ad_node(int r, int color)
{
    int off = ad_mod(color - r);
    if (off > NCOLOR / 2) off -= NCOLOR;
    return r + off;
}

}  // namespace wcns
