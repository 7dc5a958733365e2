// 3-vector value type with the reference's semantics (include/crpropa/Vector3.h).
// Component-wise arithmetic; getR() = sqrt((x*x + y*y) + z*z); dot and cross in
// the same evaluation order, so host-side normalisation (ParticleState::
// setDirection) produces the same bits as the reference.
#ifndef CRPROPA_AMD_VECTOR3_H
#define CRPROPA_AMD_VECTOR3_H

#include <cmath>
#include <limits>
#include <iostream>

namespace crpropa {

template <typename T>
class Vector3 {
public:
	T x, y, z;

	Vector3() : x(0), y(0), z(0) {}
	explicit Vector3(T t) : x(t), y(t), z(t) {}
	Vector3(T X, T Y, T Z) : x(X), y(Y), z(Z) {}
	template <typename U>
	Vector3(const Vector3<U> &v) : x(v.x), y(v.y), z(v.z) {}

	T getX() const { return x; }
	T getY() const { return y; }
	T getZ() const { return z; }
	void setX(T v) { x = v; }
	void setY(T v) { y = v; }
	void setZ(T v) { z = v; }
	void setXYZ(T X, T Y, T Z) { x = X; y = Y; z = Z; }

	T getR() const { return std::sqrt(x * x + y * y + z * z); }
	T getR2() const { return x * x + y * y + z * z; }
	Vector3 getUnitVector() const { return *this / getR(); }
	/* azimuth and zenith angles and their unit vectors (include/crpropa/Vector3.h:136-169 of the
	 * reference); the reference's build calls sin and cos of getUnitVectorPhi separately (its
	 * SphericalAdvectionShock samples show it), and the pointers keep them separate here */
	T getPhi() const {
		const T eps = std::numeric_limits<T>::min();
		if ((std::fabs(x) < eps) && (std::fabs(y) < eps))
			return 0.0;
		return std::atan2(y, x);
	}
	T getTheta() const {
		const T eps = std::numeric_limits<T>::min();
		if ((std::fabs(x) < eps) && (std::fabs(y) < eps) && (std::fabs(z) < eps))
			return 0.0;
		return std::atan2((T)std::sqrt(x * x + y * y), z);
	}
	Vector3 getUnitVectorTheta() const {
		const T theta = getTheta(), phi = getPhi();
		return Vector3(std::cos(theta) * std::cos(phi), std::cos(theta) * std::sin(phi), -std::sin(theta));
	}
	Vector3 getUnitVectorPhi() const {
		T (*volatile sin_)(T) = std::sin, (*volatile cos_)(T) = std::cos;
		return Vector3(-sin_(getPhi()), cos_(getPhi()), 0);
	}
	T dot(const Vector3 &v) const { return x * v.x + y * v.y + z * v.z; }
	Vector3 cross(const Vector3 &v) const { return Vector3(y * v.z - v.y * z, z * v.x - v.z * x, x * v.y - v.x * y); }
	T getDistanceTo(const Vector3 &p) const { return (*this - p).getR(); }
	T getAngleTo(const Vector3 &v) const {
		T c = dot(v) / v.getR() / getR();
		if (c >= 1.)
			return 0;
		if (c <= -1.)
			return M_PI;
		return std::acos(c);
	}
	bool isParallelTo(const Vector3 &v, T maxAngle) const { return getAngleTo(v) < maxAngle; }
	Vector3 floor() const { return Vector3(std::floor(x), std::floor(y), std::floor(z)); }

	bool operator==(const Vector3 &v) const { return x == v.x && y == v.y && z == v.z; }
	bool operator!=(const Vector3 &v) const { return !(*this == v); }
	Vector3 operator+(const Vector3 &v) const { return Vector3(x + v.x, y + v.y, z + v.z); }
	Vector3 operator-(const Vector3 &v) const { return Vector3(x - v.x, y - v.y, z - v.z); }
	Vector3 operator*(const Vector3 &v) const { return Vector3(x * v.x, y * v.y, z * v.z); }
	Vector3 operator/(const Vector3 &v) const { return Vector3(x / v.x, y / v.y, z / v.z); }
	Vector3 operator+(T f) const { return Vector3(x + f, y + f, z + f); }
	Vector3 operator-(T f) const { return Vector3(x - f, y - f, z - f); }
	Vector3 operator*(T f) const { return Vector3(x * f, y * f, z * f); }
	Vector3 operator/(T f) const { return Vector3(x / f, y / f, z / f); }
	Vector3 operator-() const { return Vector3(-x, -y, -z); }
	Vector3 &operator+=(const Vector3 &v) { x += v.x; y += v.y; z += v.z; return *this; }
	Vector3 &operator-=(const Vector3 &v) { x -= v.x; y -= v.y; z -= v.z; return *this; }
	Vector3 &operator*=(T f) { x *= f; y *= f; z *= f; return *this; }
	Vector3 &operator/=(T f) { x /= f; y /= f; z /= f; return *this; }
};

template <typename T>
inline Vector3<T> operator*(T f, const Vector3<T> &v) {
	return Vector3<T>(v.x * f, v.y * f, v.z * f);
}

template <typename T>
inline std::ostream &operator<<(std::ostream &out, const Vector3<T> &v) {
	return out << v.x << " " << v.y << " " << v.z;
}

typedef Vector3<double> Vector3d;
typedef Vector3<float> Vector3f;

} // namespace crpropa

#endif
