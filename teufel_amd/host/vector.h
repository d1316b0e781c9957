// vector.h - 3-vector of doubles (interface of the reference's src/vector.h:
// public x, y, z; arithmetic operators; norm, abs2nd, normalize; dot, cross).
#pragma once

#include <cmath>

class Vector {
public:
    double x, y, z;
    Vector() : x(0.0), y(0.0), z(0.0) {}
    Vector(double x0, double y0, double z0) : x(x0), y(y0), z(z0) {}

    Vector operator+(const Vector& v) const { return Vector(x + v.x, y + v.y, z + v.z); }
    Vector operator-(const Vector& v) const { return Vector(x - v.x, y - v.y, z - v.z); }
    Vector operator-() const { return Vector(-x, -y, -z); }
    Vector operator*(double f) const { return Vector(x * f, y * f, z * f); }
    Vector operator/(double f) const { return Vector(x / f, y / f, z / f); }
    Vector& operator+=(const Vector& v) { x += v.x; y += v.y; z += v.z; return *this; }
    Vector& operator-=(const Vector& v) { x -= v.x; y -= v.y; z -= v.z; return *this; }
    Vector& operator*=(double f) { x *= f; y *= f; z *= f; return *this; }
    Vector& operator/=(double f) { x /= f; y /= f; z /= f; return *this; }

    double abs2nd() const { return x * x + y * y + z * z; }
    double norm() const { return std::sqrt(abs2nd()); }
    void normalize() { double n = norm(); x /= n; y /= n; z /= n; }
};

inline double dot(const Vector& a, const Vector& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vector cross(const Vector& a, const Vector& b)
{
    return Vector(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

#define VectorZero Vector(0.0, 0.0, 0.0)
