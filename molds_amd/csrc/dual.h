// Forward-mode dual numbers for the analytic pair-energy gradient: a value and
// its partial derivatives with respect to the three components of the
// internuclear vector R_B - R_A.  Every operation applies the chain rule
// exactly, so a function written once for double and for D3 yields its exact
// first derivatives (to rounding).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace nddo {

// the double overloads stay visible beside the D3 ones below
using ::exp;
using ::pow;
using ::sqrt;

struct D3 {
    double v, d[3];
    __host__ __device__ D3(double c = 0.0) : v(c), d{0.0, 0.0, 0.0} {}
    __host__ __device__ D3(double c, double d0, double d1, double d2) : v(c), d{d0, d1, d2} {}
};

__device__ inline double val(double x) { return x; }
__device__ inline double val(const D3& x) { return x.v; }

__device__ inline D3 operator+(const D3& a, const D3& b) { return D3(a.v + b.v, a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]); }
__device__ inline D3 operator+(const D3& a, double b) { return D3(a.v + b, a.d[0], a.d[1], a.d[2]); }
__device__ inline D3 operator+(double a, const D3& b) { return b + a; }
__device__ inline D3 operator-(const D3& a) { return D3(-a.v, -a.d[0], -a.d[1], -a.d[2]); }
__device__ inline D3 operator-(const D3& a, const D3& b) { return D3(a.v - b.v, a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]); }
__device__ inline D3 operator-(const D3& a, double b) { return D3(a.v - b, a.d[0], a.d[1], a.d[2]); }
__device__ inline D3 operator-(double a, const D3& b) { return D3(a - b.v, -b.d[0], -b.d[1], -b.d[2]); }
__device__ inline D3 operator*(const D3& a, const D3& b) {
    return D3(a.v * b.v, a.d[0] * b.v + a.v * b.d[0], a.d[1] * b.v + a.v * b.d[1], a.d[2] * b.v + a.v * b.d[2]);
}
__device__ inline D3 operator*(const D3& a, double b) { return D3(a.v * b, a.d[0] * b, a.d[1] * b, a.d[2] * b); }
__device__ inline D3 operator*(double a, const D3& b) { return b * a; }
__device__ inline D3 operator/(const D3& a, double b) { return a * (1.0 / b); }
__device__ inline D3 operator/(const D3& a, const D3& b) {
    const double q = a.v / b.v, ib = 1.0 / b.v;
    return D3(q, (a.d[0] - q * b.d[0]) * ib, (a.d[1] - q * b.d[1]) * ib, (a.d[2] - q * b.d[2]) * ib);
}
__device__ inline D3 operator/(double a, const D3& b) { return D3(a) / b; }
__device__ inline D3& operator+=(D3& a, const D3& b) { return a = a + b; }
__device__ inline D3& operator-=(D3& a, const D3& b) { return a = a - b; }
__device__ inline D3& operator*=(D3& a, const D3& b) { return a = a * b; }

__device__ inline D3 sqrt(const D3& a) {
    const double s = ::sqrt(a.v), h = 0.5 / s;
    return D3(s, a.d[0] * h, a.d[1] * h, a.d[2] * h);
}
__device__ inline D3 exp(const D3& a) {
    const double e = ::exp(a.v);
    return D3(e, a.d[0] * e, a.d[1] * e, a.d[2] * e);
}
__device__ inline D3 pow(const D3& a, double p) {
    const double y = ::pow(a.v, p), dy = p * ::pow(a.v, p - 1.0);
    return D3(y, a.d[0] * dy, a.d[1] * dy, a.d[2] * dy);
}

// c * r with the product rounded on its own (no fused multiply-add with a
// later operation), for double and D3.
__device__ inline double mul_rn(double c, double r) { return __dmul_rn(c, r); }
__device__ inline D3 mul_rn(double c, const D3& r) { return D3(__dmul_rn(c, r.v), c * r.d[0], c * r.d[1], c * r.d[2]); }

}  // namespace nddo
