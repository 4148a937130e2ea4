// Geometric primitives of the baking hot path, for host and gfx950 device.
//
// Two forms of each primitive exist:
//  * the plain form, taking triangle vertices, used by the host stages;
//  * a precomputed form used by the kernels, where every sub-expression that
//    depends only on the triangle is evaluated once on the host.  Those
//    sub-expressions are computed with exactly the operations the plain form
//    uses, so both forms return identical bits.
#pragma once

#include <float.h>

#include "vmath.h"

namespace lgx {

// ---------------------------------------------------------------------------
// Segment/triangle intersection (lighter_math.cpp:315-354).  Two-sided;
// returns the segment parameter r in [0,1] of the hit, or 2 for a miss.
// ---------------------------------------------------------------------------
LHD float seg_tri(f3 L1, f3 L2, f3 P1, f3 P2, f3 P3)
{
	f3 u = P2 - P1;
	f3 v = P3 - P1;
	f3 n = cross(u, v);
	if (near_zero(n))
		return 2.0f;
	f3 dir = L2 - L1;
	f3 w0 = L1 - P1;
	float a = -dot(n, w0);
	float b = dot(n, dir);
	if (fabsf(b) < kSmall)
		return 2.0f;
	float r = a / b;
	if (r < 0.0f || r > 1.0f)
		return 2.0f;
	f3 I = L1 + r * dir;
	float uu = dot(u, u);
	float uv = dot(u, v);
	float vv = dot(v, v);
	f3 w = I - P1;
	float wu = dot(w, u);
	float wv = dot(w, v);
	float D = uv * uv - uu * vv;
	float s = (uv * wv - vv * wu) / D;
	if (s < 0.0f || s > 1.0f)
		return 2.0f;
	float t = (uv * wu - uu * wv) / D;
	if (t < 0.0f || (s + t) > 1.0f)
		return 2.0f;
	return r;
}

// Precomputed triangle for segment tests: 16 floats (64 bytes).
struct alignas(16) RayTri {
	float p1x, p1y, p1z, uu;
	float ux, uy, uz, uv;
	float vx, vy, vz, vv;
	float nx, ny, nz, D;
};

LHD RayTri make_ray_tri(f3 P1, f3 P2, f3 P3)
{
	f3 u = P2 - P1, v = P3 - P1, n = cross(u, v);
	float uu = dot(u, u), uv = dot(u, v), vv = dot(v, v);
	float D = uv * uv - uu * vv;
	return RayTri{P1.x, P1.y, P1.z, uu, u.x, u.y, u.z, uv, v.x, v.y, v.z, vv, n.x, n.y, n.z, D};
}

// seg_tri for a triangle whose normal is not near zero (the only kind the
// scene keeps), with dir = L2 - L1 supplied by the caller.
LHD float seg_tri_pre(f3 L1, f3 dir, const RayTri& T)
{
	f3 P1 = mk3(T.p1x, T.p1y, T.p1z);
	f3 n = mk3(T.nx, T.ny, T.nz);
	f3 w0 = L1 - P1;
	float a = -dot(n, w0);
	float b = dot(n, dir);
	if (fabsf(b) < kSmall)
		return 2.0f;
	float r = a / b;
	if (r < 0.0f || r > 1.0f)
		return 2.0f;
	f3 I = L1 + r * dir;
	f3 w = I - P1;
	float wu = dot(w, mk3(T.ux, T.uy, T.uz));
	float wv = dot(w, mk3(T.vx, T.vy, T.vz));
	float s = (T.uv * wv - T.vv * wu) / T.D;
	if (s < 0.0f || s > 1.0f)
		return 2.0f;
	float t = (T.uv * wu - T.uu * wv) / T.D;
	if (t < 0.0f || (s + t) > 1.0f)
		return 2.0f;
	return r;
}

// ---------------------------------------------------------------------------
// Point/triangle distance (lighter_math.cpp:875-913): plane distance inside
// the triangle, edge distance beyond one edge, vertex distance in the corner
// regions delimited by the edge-tangent projections.
// ---------------------------------------------------------------------------
LHD float point_tri_dist(f3 pt, f3 t0, f3 t1, f3 t2)
{
	f3 nrm = normalized(cross(t1 - t0, t2 - t0));
	float pd = fabsf(dot(nrm, pt) - dot(nrm, t0));
	f3 tan0 = normalized(t1 - t0);
	f3 tan1 = normalized(t2 - t1);
	f3 tan2 = normalized(t0 - t2);
	float t0p = dot(tan0, pt), t0min = dot(tan0, t0), t0max = dot(tan0, t1);
	float t1p = dot(tan1, pt), t1min = dot(tan1, t1), t1max = dot(tan1, t2);
	float t2p = dot(tan2, pt), t2min = dot(tan2, t2), t2max = dot(tan2, t0);
	if (t0min >= t0p && t2max <= t2p) return length(pt - t0);
	if (t1min >= t1p && t0max <= t0p) return length(pt - t1);
	if (t2min >= t2p && t1max <= t1p) return length(pt - t2);
	f3 en0 = normalized(cross(t1 - t0, nrm));
	f3 en1 = normalized(cross(t2 - t1, nrm));
	f3 en2 = normalized(cross(t0 - t2, nrm));
	float ptd0 = dot(en0, pt) - dot(en0, t0);
	float ptd1 = dot(en1, pt) - dot(en1, t1);
	float ptd2 = dot(en2, pt) - dot(en2, t2);
	if (ptd0 >= 0 && t0p >= t0min && t0p <= t0max) return sqrtf(pd * pd + ptd0 * ptd0);
	if (ptd1 >= 0 && t1p >= t1min && t1p <= t1max) return sqrtf(pd * pd + ptd1 * ptd1);
	if (ptd2 >= 0 && t2p >= t2min && t2p <= t2max) return sqrtf(pd * pd + ptd2 * ptd2);
	return pd;
}

// Precomputed triangle for distance queries: 40 floats (160 bytes), read as
// ten float4.  Quads 0-3 serve every evaluation, quads 4-6 the corner cases,
// quads 7-9 the edge and interior cases.
struct alignas(16) DistTri {
	float nx, ny, nz, nd;         // plane normal, n . t0
	float a0x, a0y, a0z, t0min;   // edge tangent 0, tan0 . t0
	float a1x, a1y, a1z, t1min;   // edge tangent 1, tan1 . t1
	float a2x, a2y, a2z, t2min;   // edge tangent 2, tan2 . t2
	float t0max, t1max, t2max, v0x;
	float v0y, v0z, v1x, v1y;
	float v1z, v2x, v2y, v2z;
	float e0x, e0y, e0z, e0d;     // edge normal 0, en0 . t0
	float e1x, e1y, e1z, e1d;     // edge normal 1, en1 . t1
	float e2x, e2y, e2z, e2d;     // edge normal 2, en2 . t2
};

LHD DistTri make_dist_tri(f3 t0, f3 t1, f3 t2)
{
	f3 nrm = normalized(cross(t1 - t0, t2 - t0));
	f3 tan0 = normalized(t1 - t0), tan1 = normalized(t2 - t1), tan2 = normalized(t0 - t2);
	f3 en0 = normalized(cross(t1 - t0, nrm));
	f3 en1 = normalized(cross(t2 - t1, nrm));
	f3 en2 = normalized(cross(t0 - t2, nrm));
	DistTri d;
	d.nx = nrm.x; d.ny = nrm.y; d.nz = nrm.z; d.nd = dot(nrm, t0);
	d.a0x = tan0.x; d.a0y = tan0.y; d.a0z = tan0.z; d.t0min = dot(tan0, t0);
	d.a1x = tan1.x; d.a1y = tan1.y; d.a1z = tan1.z; d.t1min = dot(tan1, t1);
	d.a2x = tan2.x; d.a2y = tan2.y; d.a2z = tan2.z; d.t2min = dot(tan2, t2);
	d.t0max = dot(tan0, t1); d.t1max = dot(tan1, t2); d.t2max = dot(tan2, t0);
	d.v0x = t0.x; d.v0y = t0.y; d.v0z = t0.z;
	d.v1x = t1.x; d.v1y = t1.y; d.v1z = t1.z;
	d.v2x = t2.x; d.v2y = t2.y; d.v2z = t2.z;
	d.e0x = en0.x; d.e0y = en0.y; d.e0z = en0.z; d.e0d = dot(en0, t0);
	d.e1x = en1.x; d.e1y = en1.y; d.e1z = en1.z; d.e1d = dot(en1, t1);
	d.e2x = en2.x; d.e2y = en2.y; d.e2z = en2.z; d.e2d = dot(en2, t2);
	return d;
}

LHD float dist_tri_pre(f3 pt, const DistTri& T)
{
	float pd = fabsf(dot(mk3(T.nx, T.ny, T.nz), pt) - T.nd);
	float t0p = dot(mk3(T.a0x, T.a0y, T.a0z), pt);
	float t1p = dot(mk3(T.a1x, T.a1y, T.a1z), pt);
	float t2p = dot(mk3(T.a2x, T.a2y, T.a2z), pt);
	if (T.t0min >= t0p && T.t2max <= t2p) return length(pt - mk3(T.v0x, T.v0y, T.v0z));
	if (T.t1min >= t1p && T.t0max <= t0p) return length(pt - mk3(T.v1x, T.v1y, T.v1z));
	if (T.t2min >= t2p && T.t1max <= t1p) return length(pt - mk3(T.v2x, T.v2y, T.v2z));
	float ptd0 = dot(mk3(T.e0x, T.e0y, T.e0z), pt) - T.e0d;
	float ptd1 = dot(mk3(T.e1x, T.e1y, T.e1z), pt) - T.e1d;
	float ptd2 = dot(mk3(T.e2x, T.e2y, T.e2z), pt) - T.e2d;
	if (ptd0 >= 0 && t0p >= T.t0min && t0p <= T.t0max) return sqrtf(pd * pd + ptd0 * ptd0);
	if (ptd1 >= 0 && t1p >= T.t1min && t1p <= T.t1max) return sqrtf(pd * pd + ptd1 * ptd1);
	if (ptd2 >= 0 && t2p >= T.t2min && t2p <= T.t2max) return sqrtf(pd * pd + ptd2 * ptd2);
	return pd;
}

// ---------------------------------------------------------------------------
// Cosine-weighted golden-angle spiral (lighter_int.hpp:456-471).  sin_side
// and cos_side depend only on (i, n) and come from a host table built with
// the C library; the azimuth's cosine and sine are supplied by the caller.
// ---------------------------------------------------------------------------
constexpr float kGoldenAngle = 137.508f / 180.0f * kPiF;

LHD f3 spiral_dir(f3 dir, float cos_side, float sin_side, float cos_around, float sin_around)
{
	f3 diffvec = mk3(dir.y, -dir.z, dir.x);
	f3 up = normalized(cross(dir, diffvec));
	f3 rt = cross(dir, up);
	return (cos_around * sin_side) * rt + (sin_around * sin_side) * up + cos_side * dir;
}

// ---------------------------------------------------------------------------
// Sample placement (lighter.cpp:392-470, lighter_math.cpp:618-650, 915-1044),
// shared by the host restatement and the gfx950 placement kernel so that
// both evaluate the same operations.
// ---------------------------------------------------------------------------
LHD bool box_disjoint(f3 qmn, f3 qmx, f3 mn, f3 mx)
{
	return qmn.x > mx.x || qmx.x < mn.x || qmn.y > mx.y || qmx.y < mn.y || qmn.z > mx.z || qmx.z < mn.z;
}

// Segment query parameters (lighter_int.hpp:680-702).
struct RaySeg {
	f3 origin;
	float len;
	f3 inv_dir;
	f3 end;
	// the segment's bounding box, padded by 1e-3 (1 + largest |coordinate|):
	// a triangle the segment test accepts holds a point within rounding of the
	// segment, so a box disjoint from this one holds no such triangle
	f3 qmn, qmx;
	LHD void set(f3 r0, f3 r1)
	{
		origin = r0;
		end = r1;
		len = length(r1 - r0);
		f3 d = normalized(r1 - r0);
		inv_dir = mk3(d.x ? 1.0f / d.x : 0.0f, d.y ? 1.0f / d.y : 0.0f, d.z ? 1.0f / d.z : 0.0f);
		float m = tmax(tmax(tmax(fabsf(r0.x), fabsf(r0.y)), fabsf(r0.z)),
		               tmax(tmax(fabsf(r1.x), fabsf(r1.y)), fabsf(r1.z)));
		float pad = 1e-3f * (1.0f + m);
		f3 p3 = mk3(pad, pad, pad);
		qmn = vmin(r0, r1) - p3;
		qmx = vmax(r0, r1) + p3;
	}
};

// The conservative slab test (lighter_math.cpp:618-650): axes whose
// direction component is zero are skipped; no tmax >= 0 check.
LHD bool ray_box(const RaySeg& r, f3 mn, f3 mx)
{
	float lo = -FLT_MAX, hi = FLT_MAX;
	if (r.inv_dir.x != 0.0f) {
		float a = (mn.x - r.origin.x) * r.inv_dir.x, b = (mx.x - r.origin.x) * r.inv_dir.x;
		lo = tmax(lo, tmin(a, b));
		hi = tmin(hi, tmax(a, b));
	}
	if (r.inv_dir.y != 0.0f) {
		float a = (mn.y - r.origin.y) * r.inv_dir.y, b = (mx.y - r.origin.y) * r.inv_dir.y;
		lo = tmax(lo, tmin(a, b));
		hi = tmin(hi, tmax(a, b));
	}
	if (r.inv_dir.z != 0.0f) {
		float a = (mn.z - r.origin.z) * r.inv_dir.z, b = (mx.z - r.origin.z) * r.inv_dir.z;
		lo = tmax(lo, tmin(a, b));
		hi = tmin(hi, tmax(a, b));
	}
	return hi >= lo && r.len >= lo;
}

// PointProjOnTriangle (lighter_math.cpp:915-952): false when the point is
// clearly outside the triangle's prism (beyond SMALL_FLOAT).
LHD bool point_proj_on_tri(f3 pt, f3 t0, f3 t1, f3 t2)
{
	f3 nrm = normalized(cross(t1 - t0, t2 - t0));
	f3 tan0 = normalized(t1 - t0), tan1 = normalized(t2 - t1), tan2 = normalized(t0 - t2);
	float t0p = dot(tan0, pt), t0min = dot(tan0, t0), t0max = dot(tan0, t1);
	float t1p = dot(tan1, pt), t1min = dot(tan1, t1), t1max = dot(tan1, t2);
	float t2p = dot(tan2, pt), t2min = dot(tan2, t2), t2max = dot(tan2, t0);
	if (t0min - kSmall > t0p && t2max + kSmall < t2p) return false;
	if (t1min - kSmall > t1p && t0max + kSmall < t0p) return false;
	if (t2min - kSmall > t2p && t1max + kSmall < t1p) return false;
	f3 en0 = normalized(cross(t1 - t0, nrm));
	f3 en1 = normalized(cross(t2 - t1, nrm));
	f3 en2 = normalized(cross(t0 - t2, nrm));
	float ptd0 = dot(en0, pt) - dot(en0, t0);
	float ptd1 = dot(en1, pt) - dot(en1, t1);
	float ptd2 = dot(en2, pt) - dot(en2, t2);
	if (ptd0 > kSmall && t0p - kSmall > t0min && t0p + kSmall < t0max) return false;
	if (ptd1 > kSmall && t1p - kSmall > t1min && t1p + kSmall < t1max) return false;
	if (ptd2 > kSmall && t2p - kSmall > t2min && t2p + kSmall < t2max) return false;
	return true;
}

// OffsetSample's box pre-test: the distance the formula returns is that to a
// point of the triangle up to a few ulps of the coordinates, so a triangle
// whose box (bmn, bmx) misses the cube around the current P of half side
// dist + 1e-5 (1 + |P|) returns a distance >= dist and can be passed over.
LHD bool offset_box_skip(f3 P, float dist, f3 bmn, f3 bmx)
{
	const float m = dist + 1e-5f * (1.0f + fmaxf(fmaxf(fabsf(P.x), fabsf(P.y)), fabsf(P.z)));
	const f3 m3 = mk3(m, m, m);
	return box_disjoint(P - m3, P + m3, bmn, bmx);
}

// One triangle of OffsetSample (lighter_math.cpp:1000-1040) up to its
// visibility check: true when the sample at P, normal N, would move off the
// triangle (p1, p2, p3) to *pnew; the move is taken when the closest hit in
// the instance along *from -> *pnew is >= 0.9 or is this triangle.
LHD bool offset_candidate(f3 P, f3 N, float dist, f3 p1, f3 p2, f3 p3, f3* from, f3* pnew)
{
	float nd = point_tri_dist(P, p1, p2, p3);
	if (!(nd < dist))
		return false;
	f3 tpn = -normalized(cross(p3 - p1, p2 - p1));
	float tpd = dot(tpn, p1);
	float sig = dot(tpn, P) - tpd;
	if (!(sig <= kSmall))
		return false;
	f3 pex = P + normalized(tpn + N) * -sig * 1.41f;
	if (!point_proj_on_tri(pex, p1, p2, p3))
		return false;
	float pdot = dot(tpn, N);
	if (!(fabsf(pdot) < 0.95f))
		return false;
	f3 ptpn = normalized(tpn - N * pdot);
	float dfac = 1.0f - fabsf(pdot);
	*pnew = P + ptpn * (-sig / dfac + kSmall);
	*from = P + (N + ptpn) * kSmall;
	return true;
}

// Correction of a sample in front of overlapping faces (lighter.cpp:445-461):
// one step towards PEnd after a hit at q with normal hn; false ends the loop.
LHD bool correction_step(f3& P, f3 PEnd, f3 N, float& md, float q, f3 hn, float corr_min_dot)
{
	if (-dot(hn, N) < corr_min_dot)
		return false;
	if (q < kSmall)
		q = kSmall;
	P = P * (1.0f - q) + PEnd * q;
	md *= (1.0f - q);
	return true;
}

} // namespace lgx
