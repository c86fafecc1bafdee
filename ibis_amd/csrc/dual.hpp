// dual.hpp -- forward-mode dual number for the tangent (JVP) residual.
//
// Same semantics as the reference's Ibis::Dual<double> (src/util/util/dual.h:
// 10-346), which its JFNK solver instantiates the whole residual on:
// comparisons, min, max and abs branch on the real part; min/max return the
// second argument on ties; floor is applied to the dual part as well and
// copysign has the reference's non-standard tangent.  Mixed operations with
// a plain double follow the reference's overloads operand for operand.
#ifndef IBIS_DUAL_HPP
#define IBIS_DUAL_HPP

#include <hip/hip_runtime.h>

namespace ibis {

// the double/float overloads stay visible next to the Dual ones below
using ::copysign;
using ::fabs;
using ::floor;
using ::pow;
using ::sqrt;

struct Dual {
    double r, d;
    __host__ __device__ Dual() : r(0.0), d(0.0) {}
    __host__ __device__ Dual(double re) : r(re), d(0.0) {}
    __host__ __device__ Dual(double re, double du) : r(re), d(du) {}
};

__host__ __device__ __forceinline__ bool operator<(const Dual& a, const Dual& b) { return a.r < b.r; }
__host__ __device__ __forceinline__ bool operator>(const Dual& a, const Dual& b) { return a.r > b.r; }
__host__ __device__ __forceinline__ bool operator<=(const Dual& a, const Dual& b) { return a.r <= b.r; }
__host__ __device__ __forceinline__ bool operator>=(const Dual& a, const Dual& b) { return a.r >= b.r; }
__host__ __device__ __forceinline__ bool operator==(const Dual& a, const Dual& b) { return a.r == b.r; }
__host__ __device__ __forceinline__ bool operator!=(const Dual& a, const Dual& b) { return a.r != b.r; }

__host__ __device__ __forceinline__ Dual operator+(const Dual& a, const Dual& b) { return Dual(a.r + b.r, a.d + b.d); }
__host__ __device__ __forceinline__ Dual operator+(const Dual& a, double b) { return Dual(a.r + b, a.d); }
__host__ __device__ __forceinline__ Dual operator+(double a, const Dual& b) { return Dual(a + b.r, b.d); }
__host__ __device__ __forceinline__ Dual operator-(const Dual& a, const Dual& b) { return Dual(a.r - b.r, a.d - b.d); }
__host__ __device__ __forceinline__ Dual operator-(const Dual& a, double b) { return Dual(a.r - b, a.d); }
__host__ __device__ __forceinline__ Dual operator-(double a, const Dual& b) { return Dual(a - b.r, -b.d); }
__host__ __device__ __forceinline__ Dual operator-(const Dual& a) { return Dual(-a.r, -a.d); }
__host__ __device__ __forceinline__ Dual operator*(const Dual& a, const Dual& b) {
    return Dual(a.r * b.r, a.r * b.d + a.d * b.r);
}
__host__ __device__ __forceinline__ Dual operator*(const Dual& a, double b) { return Dual(a.r * b, a.d * b); }
__host__ __device__ __forceinline__ Dual operator*(double a, const Dual& b) { return Dual(b.r * a, b.d * a); }
__host__ __device__ __forceinline__ Dual operator/(const Dual& n, const Dual& m) {
    return Dual(n.r / m.r, (n.d * m.r - n.r * m.d) / (m.r * m.r));
}
__host__ __device__ __forceinline__ Dual operator/(const Dual& n, double m) { return Dual(n.r / m, n.d / m); }
__host__ __device__ __forceinline__ Dual operator/(double n, const Dual& m) {
    return Dual(n / m.r, -m.d * n / (m.r * m.r));
}
__host__ __device__ __forceinline__ Dual& operator+=(Dual& a, const Dual& b) {
    a.r += b.r;
    a.d += b.d;
    return a;
}
__host__ __device__ __forceinline__ Dual& operator-=(Dual& a, const Dual& b) {
    a.r -= b.r;
    a.d -= b.d;
    return a;
}

__host__ __device__ __forceinline__ Dual sqrt(const Dual& x) {
    const double re = ::sqrt(x.r);
    return Dual(re, x.d / (2.0 * re));
}
__host__ __device__ __forceinline__ Dual fabs(const Dual& x) { return (x.r < 0.0) ? -x : x; }
__host__ __device__ __forceinline__ Dual pow(const Dual& b, const Dual& p) {
    const double re = ::pow(b.r, p.r);
    const double du = b.d * p.r * ::pow(b.r, p.r - 1.0) + p.d * re * ::log(b.r);
    return Dual(re, du);
}
__host__ __device__ __forceinline__ Dual floor(const Dual& x) { return Dual(::floor(x.r), ::floor(x.d)); }
__host__ __device__ __forceinline__ Dual copysign(const Dual& m, const Dual& s) {
    const double re = ::copysign(m.r, s.r);
    const double du = (re - m.r) < 1e-15 ? m.d : -m.d;
    return Dual(re, du);
}

// Kokkos::min / Kokkos::max on doubles return the first argument on ties;
// the Dual overloads return the second.
template <class S>
__host__ __device__ __forceinline__ S kmin(S a, S b) { return (b < a) ? b : a; }
template <class S>
__host__ __device__ __forceinline__ S kmax(S a, S b) { return (a < b) ? b : a; }
__host__ __device__ __forceinline__ Dual kmin(Dual a, Dual b) { return a.r < b.r ? a : b; }
__host__ __device__ __forceinline__ Dual kmax(Dual a, Dual b) { return a.r > b.r ? a : b; }
__host__ __device__ __forceinline__ Dual kmin(double a, Dual b) { return a < b.r ? Dual(a) : b; }
__host__ __device__ __forceinline__ Dual kmax(double a, Dual b) { return a > b.r ? Dual(a) : b; }
__host__ __device__ __forceinline__ Dual kmin(Dual a, double b) { return a.r < b ? a : Dual(b); }
__host__ __device__ __forceinline__ Dual kmax(Dual a, double b) { return a.r > b ? a : Dual(b); }

// Venkatakrishnan limiter: the running minimum of the face ratios
// num_f / den_f (den_f > 0) is kept as the fraction (bn, bd), and a face's
// ratio replaces it when num * bd < bn * den, the comparison of the ratios
// without dividing.  As with kmin above, a Dual candidate also replaces it on
// a tie.  One division, bn / bd, ends the cell's loop.
__host__ __device__ __forceinline__ bool venkat_take(double num, double den, double bn, double bd) {
    return num * bd < bn * den;
}
__host__ __device__ __forceinline__ bool venkat_take(const Dual& num, const Dual& den, const Dual& bn,
                                                     const Dual& bd) {
    return !(bn.r * den.r < num.r * bd.r);
}

__host__ __device__ __forceinline__ double real_part(double x) { return x; }
__host__ __device__ __forceinline__ double real_part(const Dual& x) { return x.r; }

// Type of the literal constants inside an S-valued expression: a Dual
// expression multiplies by plain doubles (as the reference's code does),
// float and double expressions use their own type.
template <class S>
struct lit {
    using type = S;
};
template <>
struct lit<Dual> {
    using type = double;
};

}  // namespace ibis
#endif
