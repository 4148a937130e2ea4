// Float3 arithmetic shared by the host stages and the gfx950 kernels.
//
// Results must agree bit for bit with the reference baker built for x86-64
// (IEEE binary32 per operation, no FMA contraction), so every helper below
// evaluates in a fixed order: dot products left to right (x, y, z), cross
// products as (a.y*b.z - a.z*b.y, ...), normalisation as v * (1 / sqrt(|v|^2))
// returning zero for a zero vector (lighter_int.hpp:402-441).  All sources are
// compiled with -ffp-contract=off; HIP's default correctly rounded fp32
// division and square root keep device results equal to the host's.
#pragma once

#include <math.h>
#include <stdint.h>

// Host-only translation units (the CPU stages) are plain C++.
#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define LHD __host__ __device__ __forceinline__
#else
#define LHD inline __attribute__((always_inline))
#endif

namespace lgx {

constexpr float kSmall = 0.001f;        // threshold / segment shrink (lighter_int.hpp:261)
constexpr float kPenumbraSize = 2.0f;   // distance query radius, penumbra cap (lighter_int.hpp:964)
constexpr float kPenumbraStep = 1.0f;   // march step cap (lighter_int.hpp:966)
constexpr float kShadowOffset = 0.005f; // march origin offset along the normal (lighter_int.hpp:967)
constexpr float kPiF = 3.14159265358979323846f;

struct f3 {
	float x, y, z;
};

LHD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
LHD f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }
LHD void st3(float* p, f3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
LHD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
LHD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
LHD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
LHD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
LHD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
LHD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
LHD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
LHD f3& operator+=(f3& a, f3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
LHD bool operator==(f3 a, f3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
LHD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LHD float lensq(f3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
LHD float length(f3 a) { return sqrtf(lensq(a)); }
LHD f3 cross(f3 a, f3 b) { return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
LHD bool is_zero(f3 a) { return a.x == 0 && a.y == 0 && a.z == 0; }
LHD bool near_zero(f3 a) { return fabsf(a.x) < kSmall && fabsf(a.y) < kSmall && fabsf(a.z) < kSmall; }
LHD f3 normalized(f3 a)
{
	float l2 = lensq(a);
	if (l2 == 0)
		return f3{0, 0, 0};
	float inv = 1.0f / sqrtf(l2);
	return f3{a.x * inv, a.y * inv, a.z * inv};
}
// Ternary min/max with the reference's NaN behaviour (lighter_int.hpp:277-278).
LHD float tmin(float a, float b) { return a < b ? a : b; }
LHD float tmax(float a, float b) { return a > b ? a : b; }
LHD f3 vmin(f3 a, f3 b) { return f3{tmin(a.x, b.x), tmin(a.y, b.y), tmin(a.z, b.z)}; }
LHD f3 vmax(f3 a, f3 b) { return f3{tmax(a.x, b.x), tmax(a.y, b.y), tmax(a.z, b.z)}; }
// a * (1 - s) + b * s (lighter_int.hpp:285)
LHD f3 lerp3(f3 a, f3 b, float s) { return a * (1.0f - s) + b * s; }
LHD float lerpf(float a, float b, float s) { return a * (1.0f - s) + b * s; }

struct f2 {
	float x, y;
};
LHD f2 mk2(float x, float y) { return f2{x, y}; }
LHD f2 operator+(f2 a, f2 b) { return f2{a.x + b.x, a.y + b.y}; }
LHD f2 operator-(f2 a, f2 b) { return f2{a.x - b.x, a.y - b.y}; }
LHD f2 operator*(f2 a, f2 b) { return f2{a.x * b.x, a.y * b.y}; }
LHD f2 operator*(f2 a, float s) { return f2{a.x * s, a.y * s}; }
LHD f2 operator-(f2 a) { return f2{-a.x, -a.y}; }
LHD float dot2(f2 a, f2 b) { return a.x * b.x + a.y * b.y; }
LHD float cross2(f2 a, f2 b) { return (a.x * b.y) - (a.y * b.x); }
LHD f2 perp2(f2 a) { return f2{a.y, -a.x}; }
LHD float length2(f2 a) { return sqrtf(a.x * a.x + a.y * a.y); }
LHD f2 normalized2(f2 a)
{
	float l2 = a.x * a.x + a.y * a.y;
	if (l2 == 0)
		return f2{0, 0};
	float inv = 1.0f / sqrtf(l2);
	return f2{a.x * inv, a.y * inv};
}

struct f4 {
	float x, y, z, w;
};
LHD f4 operator+(f4 a, f4 b) { return f4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
LHD f4 operator-(f4 a, f4 b) { return f4{a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
LHD f4 operator*(f4 a, float s) { return f4{a.x * s, a.y * s, a.z * s, a.w * s}; }

} // namespace lgx
