// Core scalar/vector types and sphere geometry for the host side of the
// spherical Biot-Savart / Poisson sum.
//
// Semantics follow the reference's base layer:
//   scalar/index types and constants      src/LpmTypeDefs.h:11-35
//   XyzVector arithmetic                  src/LpmXyzVector.h:12-66
//   distance / sphere geometry            src/LpmXyzVector.cpp:17-107
// The floating-point operation order of every function below is the
// reference's (left-to-right sums, reciprocal-then-multiply in centroid,
// divide-by-norm in normalize), because mesh positions and areas must come
// out bit-identical to the reference's for the same seed and level.  This
// file is compiled without FMA contraction (-ffp-contract=off).
#pragma once

#include <cmath>
#include <cstdint>
#include <ostream>
#include <vector>

namespace Lpm {

typedef double scalar_type;
typedef long index_type;

static const scalar_type PI = 3.1415926535897932384626433832795027975;
static const scalar_type ZERO_TOL = 1.0e-13;

struct XyzVector {
    scalar_type x, y, z;
    XyzVector(scalar_type nx = 0.0, scalar_type ny = 0.0, scalar_type nz = 0.0) : x(nx), y(ny), z(nz) {}

    XyzVector& operator+=(const XyzVector& o) { x += o.x; y += o.y; z += o.z; return *this; }
    XyzVector& operator-=(const XyzVector& o) { x -= o.x; y -= o.y; z -= o.z; return *this; }

    scalar_type magnitudeSquared() const { return x * x + y * y + z * z; }
    scalar_type magnitude() const { return std::sqrt(x * x + y * y + z * z); }
    scalar_type dotProduct(const XyzVector& o) const { return x * o.x + y * o.y + z * o.z; }
    XyzVector crossProduct(const XyzVector& o) const {
        return XyzVector(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x);
    }
    void scale(scalar_type m) { x *= m; y *= m; z *= m; }
    XyzVector scalarMultiply(scalar_type c) const { return XyzVector(x * c, y * c, z * c); }
    void normalize() { const scalar_type n = magnitude(); x /= n; y /= n; z /= n; }
};

inline XyzVector operator+(const XyzVector& a, const XyzVector& b) { return XyzVector(a.x + b.x, a.y + b.y, a.z + b.z); }
inline XyzVector operator-(const XyzVector& a, const XyzVector& b) { return XyzVector(a.x - b.x, a.y - b.y, a.z - b.z); }
inline std::ostream& operator<<(std::ostream& os, const XyzVector& v) {
    return os << "(" << v.x << ", " << v.y << ", " << v.z << ")";
}

inline scalar_type square(scalar_type v) { return v * v; }
inline scalar_type cube(scalar_type v) { return v * v * v; }

// Longitude in [0, 2pi) with the reference's quadrant handling
// (src/LpmXyzVector.cpp:17-48).
inline scalar_type atan4(scalar_type y, scalar_type x) {
    if (x == 0.0) {
        if (y > 0.0) return 0.5 * PI;
        if (y < 0.0) return 1.5 * PI;
        return 0.0;
    }
    if (y == 0.0) return (x > 0.0) ? 0.0 : PI;
    const scalar_type t = std::atan2(std::abs(y), std::abs(x));
    if (x > 0.0 && y > 0.0) return t;
    if (x < 0.0 && y > 0.0) return PI - t;
    if (x < 0.0 && y < 0.0) return PI + t;
    return 2.0 * PI - t;
}
inline scalar_type longitude(const XyzVector& v) { return atan4(v.y, v.x); }
inline scalar_type latitude(const XyzVector& v) { return std::atan2(v.z, std::sqrt(v.x * v.x + v.y * v.y)); }

// Euclidean distance as sqrt of (b-a)^2 summed x, y, z (src/LpmXyzVector.cpp:66-68).
inline scalar_type distance(const XyzVector& a, const XyzVector& b) {
    return std::sqrt((b.x - a.x) * (b.x - a.x) + (b.y - a.y) * (b.y - a.y) + (b.z - a.z) * (b.z - a.z));
}

// Great-circle distance atan2(|a x b|, a.b) (src/LpmXyzVector.cpp:76-80).
inline scalar_type sphereDistance(const XyzVector& a, const XyzVector& b, scalar_type radius = 1.0) {
    const XyzVector c = a.crossProduct(b);
    return std::atan2(c.magnitude(), a.dotProduct(b)) * radius;
}

// L'Huilier spherical-triangle area (src/LpmXyzVector.cpp:82-90).
inline scalar_type sphereTriArea(const XyzVector& a, const XyzVector& b, const XyzVector& c, scalar_type radius = 1.0) {
    const scalar_type s1 = sphereDistance(a, b);
    const scalar_type s2 = sphereDistance(b, c);
    const scalar_type s3 = sphereDistance(c, a);
    const scalar_type hp = 0.5 * (s1 + s2 + s3);
    const scalar_type zz = std::tan(0.5 * hp) * std::tan(0.5 * (hp - s1)) * std::tan(0.5 * (hp - s2)) *
                           std::tan(0.5 * (hp - s3));
    return 4.0 * std::atan2(std::sqrt(zz), 1.0) * radius * radius;
}

// Normalized vertex mean: sum from zero, multiply by 1/n, divide by the norm,
// multiply by the radius (src/LpmXyzVector.cpp:92-100).
inline XyzVector sphereCentroid(const XyzVector* v, int n, scalar_type radius = 1.0) {
    XyzVector c(0.0, 0.0, 0.0);
    for (int i = 0; i < n; ++i) c += v[i];
    c.scale(1.0 / n);
    c.normalize();
    c.scale(radius);
    return c;
}

// Normalized chord midpoint (src/LpmXyzVector.cpp:102-107).
inline XyzVector sphereMidpoint(const XyzVector& a, const XyzVector& b, scalar_type radius = 1.0) {
    XyzVector m(0.5 * (a.x + b.x), 0.5 * (a.y + b.y), 0.5 * (a.z + b.z));
    m.normalize();
    m.scale(radius);
    return m;
}

}  // namespace Lpm
