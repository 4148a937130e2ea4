// gfx950 kernels of the baking hot path and the DeviceScene that owns their
// device-resident data.  See device.h for the stage contracts.
//
// Query semantics (all order independent, hence equal to the reference's
// tree walks):
//  * distance(p): min(2, min over shadow triangles of PointTriangleDistance)
//    (lighter.cpp:150-188);
//  * closest(A,B): min over triangles of the segment hit parameter on A->B
//    shrunk by 0.001 at both ends, 2 when nothing is hit (lighter.cpp:254-289);
//  * occluded(A,B): some triangle hit with r < 1 on the same shrunk segment
//    (lighter.cpp:112-147).
#include <hip/hip_runtime.h>
#include <mutex>
#include <stdlib.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "bvh.h"
#include "device.h"
#include "geom.h"
#include "normals.h"

namespace lgx {

#define HIPCHECK(x)                                                                                          \
	do {                                                                                                     \
		hipError_t e_ = (x);                                                                                 \
		if (e_ != hipSuccess)                                                                                \
			throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " __FILE__ ":" + \
			                         std::to_string(__LINE__) + ": " #x);                                    \
	} while (0)

constexpr int kStackSize = 96; // >= 3 x (wide BVH depth) + 1, checked at build
constexpr int kBlock = 256;
constexpr int kNumCounters = 6;
constexpr uint32_t kSkipBits = 0xFFFFFFFFu; // weight of a texel the reference skips
constexpr uint32_t kSampleBits = 26;
constexpr uint32_t kSampleMask = (1u << kSampleBits) - 1;

struct View {
	const WideNode* __restrict__ nodes;
	const DistTri* __restrict__ dtri;
	const RayTri* __restrict__ rtri;
	int has_tris;
	// scene box grown by the distance-query radius plus a safety margin:
	// outside it every distance query returns exactly 2
	float xlo[3], xhi[3];
	// entry grid: cell (ix, iy, iz) of size 1 / ginv from glo; its entry node
	// is nodes[gbase + (iz * gn[1] + iy) * gn[0] + ix]; gbase = 0: no grid
	float glo[3], ginv;
	int gn[3];
	int32_t gbase;
	// the same cells again from gfine on, each a cut for the cell grown by
	// kGridFineReach only, for distance queries of radius <= 1 (0: none)
	int32_t gfine;
};

// The entry grid.  Every query of the kernels here that is local (a distance
// query of radius <= kGridReach, a segment no longer than kGridReach) can
// only be answered by triangles that meet the ball of that radius around
// its point (the segment's start).  Each cell of a uniform grid over the
// scene holds a virtual wide node (after the tree's nodes) whose children
// are a cut of the tree: at most four subtrees such that every triangle
// meeting the cell grown by kGridReach (plus a rounding margin) lies in one
// of them.  A second node per cell (from gfine on) is the same cut for the
// cell grown by kGridFineReach only, for distance queries of radius <= 1:
// a smaller region, a deeper cut.  A local query starts at its cell's node
// instead of the root, so it skips the levels above the cut and still
// meets every triangle that can change its answer; its result is the same
// minimum.
constexpr float kGridReach = kPenumbraSize;
constexpr float kGridFineReach = kPenumbraStep; // 1 m: the radius of many march queries
constexpr uint32_t kGridCells = 1u << 22; // cells at most (two nodes of 128 bytes each, 1 GB)
constexpr float kGridMinCell = 1.0f;      // metres: finer cells gain nothing against the 2 m reach

// entry node of a query at p (the root outside the grid); fine: the query's
// radius is at most kGridFineReach
__device__ __forceinline__ int32_t entry_node(const View& V, f3 p, bool fine = false)
{
	if (!V.gbase)
		return 0;
	const float fx = (p.x - V.glo[0]) * V.ginv, fy = (p.y - V.glo[1]) * V.ginv, fz = (p.z - V.glo[2]) * V.ginv;
	if (!(fx >= 0.0f && fy >= 0.0f && fz >= 0.0f && fx < (float)V.gn[0] && fy < (float)V.gn[1] &&
	      fz < (float)V.gn[2]))
		return 0;
	const int ix = min((int)fx, V.gn[0] - 1), iy = min((int)fy, V.gn[1] - 1), iz = min((int)fz, V.gn[2] - 1);
	return (fine && V.gfine ? V.gfine : V.gbase) + (iz * V.gn[1] + iy) * V.gn[0] + ix;
}

// ---------------------------------------------------------------------------
// transcendental helpers: evaluated in double and rounded once, which is the
// correctly rounded value except in vanishingly rare ties
// ---------------------------------------------------------------------------
__device__ __forceinline__ float pow_r(float x, float y)
{
	if (y == 1.0f)
		return x;
	return (float)pow((double)x, (double)y);
}
__device__ __forceinline__ float acos_r(float x) { return (float)acos((double)x); }

struct WNode {
	float4 lx, ly, lz, hx, hy, hz; // the four children's boxes, one plane per float4
	int4 c;                        // their references
};

__device__ __forceinline__ WNode load_node(const WideNode* nodes, int32_t i)
{
	const float4* q = reinterpret_cast<const float4*>(nodes + i);
	WNode n;
	n.lx = q[0];
	n.ly = q[1];
	n.lz = q[2];
	n.hx = q[3];
	n.hy = q[4];
	n.hz = q[5];
	n.c = reinterpret_cast<const int4*>(q)[6];
	return n;
}

// ascending order of (key, reference) pairs
__device__ __forceinline__ void cswap(float& ka, int32_t& ra, float& kb, int32_t& rb)
{
	const bool sw = kb < ka;
	const float k = sw ? kb : ka;
	kb = sw ? ka : kb;
	ka = k;
	const int32_t r = sw ? rb : ra;
	rb = sw ? ra : rb;
	ra = r;
}

__device__ __forceinline__ void sort4(float (&k)[4], int32_t (&r)[4])
{
	cswap(k[0], r[0], k[1], r[1]);
	cswap(k[2], r[2], k[3], r[3]);
	cswap(k[0], r[0], k[2], r[2]);
	cswap(k[1], r[1], k[3], r[3]);
	cswap(k[1], r[1], k[2], r[2]);
}

__device__ __forceinline__ float box_d2(f3 p, float lx, float ly, float lz, float hx, float hy, float hz)
{
	float dx = fmaxf(fmaxf(lx - p.x, p.x - hx), 0.0f);
	float dy = fmaxf(fmaxf(ly - p.y, p.y - hy), 0.0f);
	float dz = fmaxf(fmaxf(lz - p.z, p.z - hz), 0.0f);
	return dx * dx + dy * dy + dz * dz;
}

__device__ __forceinline__ DistTri load_dtri(const DistTri* t, uint32_t i)
{
	const float4* q = reinterpret_cast<const float4*>(t + i);
	DistTri r;
	float4* o = reinterpret_cast<float4*>(&r);
#pragma unroll
	for (int k = 0; k < 10; ++k)
		o[k] = q[k];
	return r;
}

__device__ __forceinline__ RayTri load_rtri(const RayTri* t, uint32_t i)
{
	const float4* q = reinterpret_cast<const float4*>(t + i);
	RayTri r;
	float4* o = reinterpret_cast<float4*>(&r);
#pragma unroll
	for (int k = 0; k < 4; ++k)
		o[k] = q[k];
	return r;
}

typedef float v4f __attribute__((ext_vector_type(4)));

// True when the triangle cannot lower the distance bound `best`: the box of
// its three vertices (DistTri quads 4-6, qa..qc) lies farther than best +
// margin.
// The distance the reference's formula returns is the distance to a point of
// the triangle (a vertex, a point of an edge line inside the edge's span, or
// the projection into the face), up to rounding in its region tests of a few
// ulps of the coordinates, so it is never below the box distance minus that;
// the margin (1e-5 of the coordinates' magnitude, hundreds of ulps) covers
// it, and a skipped triangle could only have returned a value >= best, which
// leaves the minimum unchanged.
__device__ __forceinline__ bool tri_box_beyond_q(f3 p, v4f qa, v4f qb, v4f qc, float best)
{
	// vertices (qa.w, qb.x, qb.y), (qb.z, qb.w, qc.x), (qc.y, qc.z, qc.w)
	float lx = fminf(fminf(qa.w, qb.z), qc.y), hx = fmaxf(fmaxf(qa.w, qb.z), qc.y);
	float ly = fminf(fminf(qb.x, qb.w), qc.z), hy = fmaxf(fmaxf(qb.x, qb.w), qc.z);
	float lz = fminf(fminf(qb.y, qc.x), qc.w), hz = fmaxf(fmaxf(qb.y, qc.x), qc.w);
	float lim = best + 1e-5f * (1.0f + fabsf(p.x) + fabsf(p.y) + fabsf(p.z));
	return box_d2(p, lx, ly, lz, hx, hy, hz) > lim * lim;
}

// Traversal stacks.  The traversals below are nearest-first: at an inner
// node they descend into the nearest child and push the other (at most
// three) children that are still within reach, so the stack holds at most
// three entries per level of the current path and 3 x the BVH's depth + 1
// bounds it.
// Entry = node reference + key (squared box distance, or the segment's entry
// parameter); a popped entry is skipped when its key exceeds the current
// bound.
struct PrivStack { // per-lane arrays (private segment)
	int32_t ref[kStackSize];
	float key[kStackSize];
	__device__ __forceinline__ void put(int i, int32_t r, float k)
	{
		ref[i] = r;
		key[i] = k;
	}
	__device__ __forceinline__ int32_t r(int i) const { return ref[i]; }
	__device__ __forceinline__ float k(int i) const { return key[i]; }
};

// A column of LDS per lane: references [level][lane] as int32, keys as FP16
// rounded toward zero.  Keys are non-negative, so the stored key never
// exceeds the true one and skipping stays conservative: at worst a node is
// visited that the exact key would have skipped, which changes no result.
__device__ __forceinline__ _Float16 f16_down(float x)
{
	_Float16 h = (_Float16)x;
	if ((float)h > x) {
		uint16_t b;
		memcpy(&b, &h, 2);
		b = (uint16_t)(b - 1); // x >= 0 and h > x imply h > 0
		memcpy(&h, &b, 2);
	}
	return h;
}

// The first kLdsLevels levels in LDS, deeper ones (rare: only a path that
// pushed at more than kLdsLevels of its levels) in a per-lane array.
constexpr int kLdsLevels = 12;

// LDS-qualified element types: through a generic pointer the stack would be
// accessed with FLAT instructions, which count against both the vector-memory
// and the LDS wait counters (every node fetch would then wait for the stack's
// stores and vice versa)
typedef __attribute__((address_space(3))) int32_t lds_int32_t;
typedef __attribute__((address_space(3))) _Float16 lds_half_t;

struct LdsStack {
	lds_int32_t* ref; // &refs[0][lane]
	lds_half_t* key;  // &keys[0][lane]
	int32_t oref[kStackSize - kLdsLevels];
	float okey[kStackSize - kLdsLevels];
	__device__ __forceinline__ void put(int i, int32_t r, float k)
	{
		if (i < kLdsLevels) {
			ref[i * kBlock] = r;
			key[i * kBlock] = f16_down(k);
		} else {
			oref[i - kLdsLevels] = r;
			okey[i - kLdsLevels] = k;
		}
	}
	__device__ __forceinline__ int32_t r(int i) const { return i < kLdsLevels ? ref[i * kBlock] : oref[i - kLdsLevels]; }
	__device__ __forceinline__ float k(int i) const
	{
		return i < kLdsLevels ? (float)key[i * kBlock] : okey[i - kLdsLevels];
	}
};

// The calling lane's LDS stack in the kernel's dynamic LDS (kBlock x depth
// int32 references, then kBlock x depth FP16 keys).
__device__ __forceinline__ LdsStack lds_stack_of_lane(unsigned char* lds, int depth)
{
	LdsStack st;
	st.ref = (lds_int32_t*)(reinterpret_cast<int32_t*>(lds) + threadIdx.x);
	st.key = (lds_half_t*)(reinterpret_cast<_Float16*>(lds + 4 * (size_t)kBlock * depth) + threadIdx.x);
	return st;
}

// Nearest-first traversal for the distance query: min(radius, distance to the
// nearest triangle closer than radius).
template <bool COUNT, class Stack>
__device__ float q_distance(const View& V, f3 p, float radius, unsigned long long* cnt, Stack& st)
{
	float best = radius;
	if (!V.has_tris)
		return best;
	int sp = 0;
	int32_t ref = radius <= kGridReach ? entry_node(V, p, radius <= kGridFineReach) : 0;
	unsigned nn = 0, nt = 0, ncull = 0;
	for (;;) {
		if (ref >= 0) {
			WNode n = load_node(V.nodes, ref);
			nn++;
			float d[4] = {box_d2(p, n.lx.x, n.ly.x, n.lz.x, n.hx.x, n.hy.x, n.hz.x),
			              box_d2(p, n.lx.y, n.ly.y, n.lz.y, n.hx.y, n.hy.y, n.hz.y),
			              box_d2(p, n.lx.z, n.ly.z, n.lz.z, n.hx.z, n.hy.z, n.hz.z),
			              box_d2(p, n.lx.w, n.ly.w, n.lz.w, n.hx.w, n.hy.w, n.hz.w)};
			int32_t r[4] = {n.c.x, n.c.y, n.c.z, n.c.w};
			sort4(d, r);
			float lim = best * best;
			if (d[0] <= lim) {
				// farthest first, so that the nearer ones pop first
				if (d[3] <= lim)
					st.put(sp++, r[3], d[3]);
				if (d[2] <= lim)
					st.put(sp++, r[2], d[2]);
				if (d[1] <= lim)
					st.put(sp++, r[1], d[1]);
				ref = r[0];
				continue;
			}
		} else {
			uint32_t f = leaf_first(ref), c = leaf_count(ref);
			for (uint32_t i = 0; i < c; ++i) {
				// the plane, tangents and vertices (quads 0-6) in one round
				// trip with the box pre-test's vertices; the edge normals
				// (quads 7-9) only where the vertex regions do not decide.
				// The empty asm takes the seven quads as register operands
				// here, so the compiler issues their loads together instead
				// of moving quads 0-3 behind the pre-test's branch (a second
				// round trip to memory for almost every triangle).
				const v4f* q = reinterpret_cast<const v4f*>(V.dtri + f + i);
				const v4f q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4], q5 = q[5], q6 = q[6];
				asm volatile("" ::"v"(q0), "v"(q1), "v"(q2), "v"(q3), "v"(q4), "v"(q5), "v"(q6));
				if (tri_box_beyond_q(p, q4, q5, q6, best)) {
					ncull++;
					continue;
				}
				DistTri T;
				v4f* o = reinterpret_cast<v4f*>(&T);
				o[0] = q0;
				o[1] = q1;
				o[2] = q2;
				o[3] = q3;
				o[4] = q4;
				o[5] = q5;
				o[6] = q6;
				o[7] = q[7];
				o[8] = q[8];
				o[9] = q[9];
				float d = dist_tri_pre(p, T);
				nt++;
				if (d < best)
					best = d;
			}
		}
		// pop the next entry still within reach
		bool found = false;
		while (sp > 0) {
			--sp;
			if (st.k(sp) <= best * best) {
				ref = st.r(sp);
				found = true;
				break;
			}
		}
		if (!found)
			break;
	}
	if (COUNT) {
		atomicAdd(cnt + 0, (unsigned long long)nn);
		atomicAdd(cnt + 1, (unsigned long long)nt);
		atomicAdd(cnt + 5, (unsigned long long)ncull);
		atomicAdd(cnt + 3, 1ull);
	}
	return best;
}

template <bool COUNT>
__device__ __forceinline__ float q_distance(const View& V, f3 p, float radius, unsigned long long* cnt)
{
	PrivStack st;
	return q_distance<COUNT>(V, p, radius, cnt, st);
}

struct Seg {
	f3 o, dir, inv;
};

__device__ __forceinline__ Seg make_seg(f3 A, f3 B)
{
	f3 dn = normalized(B - A);
	f3 mA = A + dn * kSmall;
	f3 mB = B - dn * kSmall;
	Seg s;
	s.o = mA;
	s.dir = mB - mA;
	float ex = fabsf(s.dir.x) > 1e-30f ? s.dir.x : copysignf(1e-30f, s.dir.x);
	float ey = fabsf(s.dir.y) > 1e-30f ? s.dir.y : copysignf(1e-30f, s.dir.y);
	float ez = fabsf(s.dir.z) > 1e-30f ? s.dir.z : copysignf(1e-30f, s.dir.z);
	s.inv = mk3(1.0f / ex, 1.0f / ey, 1.0f / ez);
	return s;
}

// entry parameter of the segment into a box, or +inf when [0, tlim] misses it
__device__ __forceinline__ float slab(const Seg& s, float lx, float ly, float lz, float hx, float hy, float hz,
                                      float tlim)
{
	float tx0 = (lx - s.o.x) * s.inv.x, tx1 = (hx - s.o.x) * s.inv.x;
	float ty0 = (ly - s.o.y) * s.inv.y, ty1 = (hy - s.o.y) * s.inv.y;
	float tz0 = (lz - s.o.z) * s.inv.z, tz1 = (hz - s.o.z) * s.inv.z;
	float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
	float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tlim));
	return tn <= tf ? tn : INFINITY;
}

// ANY = true: returns 0 on the first hit with r < 1 (occlusion), else 2.
// start >= 0: the walk starts at that node, which the caller took from the
// entry grid at A for a segment of at most kGridReach; -1: at A's cell for
// such a segment, else at the root.
template <bool ANY, bool COUNT, class Stack>
__device__ float q_segment(const View& V, f3 A, f3 B, unsigned long long* cnt, Stack& st, int32_t start = -1)
{
	float best = 2.0f;
	if (!V.has_tris)
		return best;
	Seg s = make_seg(A, B);
	int sp = 0;
	const f3 AB = B - A;
	int32_t ref = start >= 0 ? start : dot(AB, AB) <= kGridReach * kGridReach ? entry_node(V, A) : 0;
	unsigned nn = 0, nt = 0;
	for (;;) {
		if (ref >= 0) {
			WNode n = load_node(V.nodes, ref);
			nn++;
			float tl = fminf(best, 1.0f);
			float t[4] = {slab(s, n.lx.x, n.ly.x, n.lz.x, n.hx.x, n.hy.x, n.hz.x, tl),
			              slab(s, n.lx.y, n.ly.y, n.lz.y, n.hx.y, n.hy.y, n.hz.y, tl),
			              slab(s, n.lx.z, n.ly.z, n.lz.z, n.hx.z, n.hy.z, n.hz.z, tl),
			              slab(s, n.lx.w, n.ly.w, n.lz.w, n.hx.w, n.hy.w, n.hz.w, tl)};
			int32_t r[4] = {n.c.x, n.c.y, n.c.z, n.c.w};
			sort4(t, r);
			if (t[0] != INFINITY) {
				if (t[3] != INFINITY)
					st.put(sp++, r[3], t[3]);
				if (t[2] != INFINITY)
					st.put(sp++, r[2], t[2]);
				if (t[1] != INFINITY)
					st.put(sp++, r[1], t[1]);
				ref = r[0];
				continue;
			}
		} else {
			uint32_t f = leaf_first(ref), c = leaf_count(ref);
			nt += c;
			bool done = false;
			for (uint32_t i = 0; i < c; ++i) {
				RayTri T = load_rtri(V.rtri, f + i);
				float r = seg_tri_pre(s.o, s.dir, T);
				if (r < best)
					best = r;
				if (ANY && best < 1.0f) {
					done = true;
					break;
				}
			}
			if (done)
				break;
		}
		bool found = false;
		while (sp > 0) {
			--sp;
			if (st.k(sp) <= fminf(best, 1.0f)) {
				ref = st.r(sp);
				found = true;
				break;
			}
		}
		if (!found)
			break;
	}
	if (COUNT) {
		atomicAdd(cnt + 0, (unsigned long long)nn);
		atomicAdd(cnt + 1, (unsigned long long)nt);
		atomicAdd(cnt + 3, 1ull);
	}
	return best;
}

template <bool ANY, bool COUNT>
__device__ __forceinline__ float q_segment(const View& V, f3 A, f3 B, unsigned long long* cnt)
{
	PrivStack st;
	return q_segment<ANY, COUNT>(V, A, B, cnt, st);
}

// Exit parameter of the ray from + rd*t from the grown scene box, padded so
// that points beyond it are certainly outside; -inf if the ray misses it.
__device__ __forceinline__ float box_exit(const View& V, f3 from, f3 rd)
{
	float lo[3] = {V.xlo[0], V.xlo[1], V.xlo[2]}, hi[3] = {V.xhi[0], V.xhi[1], V.xhi[2]};
	float o[3] = {from.x, from.y, from.z}, d[3] = {rd.x, rd.y, rd.z};
	float tn = -INFINITY, tf = INFINITY;
#pragma unroll
	for (int a = 0; a < 3; ++a) {
		if (d[a] != 0.0f) {
			float t0 = (lo[a] - o[a]) / d[a], t1 = (hi[a] - o[a]) / d[a];
			tn = fmaxf(tn, fminf(t0, t1));
			tf = fminf(tf, fmaxf(t0, t1));
		} else if (o[a] < lo[a] || o[a] > hi[a]) {
			return -INFINITY;
		}
	}
	if (tn > tf)
		return -INFINITY;
	return tf * 1.0001f + 0.01f;
}

// CalcInvShadowFactor (lighter.cpp:190-207) as a resumable state machine, one
// distance query per step.  Two shortcuts leave the result unchanged:
//  * once the march has left the grown scene box every remaining step sees
//    h = 2 and, for k >= 0, cannot lower res (2 / min(t*k, 2) >= 1 >= res),
//    so the march stops there;
//  * the distance to the scene is 1-Lipschitz, so the next query's search
//    radius starts at min(2, h + step + margin) instead of 2: every
//    triangle closer than that is still found, hence the same minimum;
//  * for k >= 0 a step's distance h matters only below
//    need = max(1, res m (1 + 2^-20)), m = min(t k, 2): at or beyond it the
//    step is min(h, 1) = 1 and h / m >= res after rounding (the factor
//    1 + 2^-20 is far above the two roundings of res m and h / m), so res
//    keeps its value.  The query radius is therefore min(2, Lipschitz bound,
//    need); when it was set by need and nothing lies closer, the step is 1,
//    res is unchanged and the next Lipschitz bound is unknown (2).
// A march whose t no longer advances (t + step == t, where the reference
// would loop forever) ends with the current factor.
struct March {
	f3 from, rd;
	float t, maxt, k, res, t_out, radius;
	uint32_t steps;
};

__device__ __forceinline__ void march_init(March& M, const View& V, f3 from, f3 to, float k)
{
	M.from = from;
	M.rd = normalized(to - from);
	M.t = 0.001f;
	M.maxt = length(to - from);
	M.k = k;
	M.res = 1.0f;
	M.t_out = (k >= 0.0f) ? box_exit(V, from, M.rd) : INFINITY;
	M.radius = kPenumbraSize;
	M.steps = 0;
}

// margin of the next search radius: covers the rounding of the computed
// distances (a few ulps of the coordinates) and of the march position
__device__ __forceinline__ float lipschitz_margin(f3 p)
{
	return 1e-4f + 1e-5f * fmaxf(fmaxf(fabsf(p.x), fabsf(p.y)), fabsf(p.z));
}

// Advances one step; returns true when the march is over (result in M.res).
template <bool COUNT, class Stack>
__device__ __forceinline__ bool march_step(March& M, const View& V, unsigned long long* cnt, Stack& stack)
{
	if (!(M.t < M.maxt) || M.t > M.t_out)
		return true;
	f3 p = M.from + M.rd * M.t;
	const float m = tmin(M.t * M.k, kPenumbraSize);
	const float need = M.k >= 0.0f ? fmaxf(kPenumbraStep, M.res * m * (1.0f + 0x1p-20f)) : kPenumbraSize;
	const float R = fminf(M.radius, need);
	float h = q_distance<COUNT>(V, p, R, cnt, stack);
	M.steps++;
	if (need < M.radius && h >= R) {
		// nothing closer than need (>= 1): step 1, res unchanged
		float tn = M.t + kPenumbraStep;
		if (tn == M.t)
			return true;
		M.radius = kPenumbraSize;
		M.t = tn;
		return false;
	}
	if (h < 0.001f) {
		M.res = 0.0f;
		return true;
	}
	M.res = tmin(M.res, h / m);
	float st = tmin(h, kPenumbraStep);
	float tn = M.t + st;
	if (tn == M.t)
		return true;
	M.radius = fminf(h + st + lipschitz_margin(p), kPenumbraSize);
	M.t = tn;
	return false;
}

template <bool COUNT>
__device__ float shadow_march(const View& V, f3 from, f3 to, float k, unsigned long long* cnt, uint32_t* nsteps)
{
	March M;
	march_init(M, V, from, to, k);
	PrivStack st;
	while (!march_step<COUNT>(M, V, cnt, st)) {
	}
	if (COUNT) {
		atomicAdd(cnt + 2, (unsigned long long)M.steps);
		atomicAdd(cnt + 4, 1ull);
	}
	if (nsteps)
		*nsteps = M.steps;
	return M.res;
}

// ---------------------------------------------------------------------------
// Reference-shaped traversal counts (measurement only).  The byte count a
// roofline fraction is stated on must not depend on this build's own tree,
// so it is taken on a plain binary SAH tree (16 bins, at most 4 triangles
// per leaf) with the reference's query semantics: every march step of
// CalcInvShadowFactor (lighter.cpp:190-207) is a fresh distance query of
// radius MAX_PENUMBRA_SIZE (no exit test, no Lipschitz or need bound), the
// nearer child is visited first and the radius shrinks with every triangle
// distance found.  A node visit is one box tested (32 bytes of node in the
// reference layout), a triangle test one point-triangle distance (36 bytes
// of vertices).  The march's factor is compared with the product's, bit for
// bit, as a by-product.
// ---------------------------------------------------------------------------
struct CanonView {
	const BvhNode* __restrict__ nodes;
	const DistTri* __restrict__ dtri;
	float lo[3], hi[3]; // the root's box, padded as the builder pads node boxes
};

// The builder's outward padding of a box (bvh.cpp), for the root's box.
inline void canon_root_box(const BinaryBvh& b, float lo[3], float hi[3])
{
	const float l[3] = {b.lo.x, b.lo.y, b.lo.z}, h[3] = {b.hi.x, b.hi.y, b.hi.z};
	for (int a = 0; a < 3; ++a) {
		const float m = 1e-5f * std::max(fabsf(l[a]), fabsf(h[a])) + 1e-6f;
		lo[a] = l[a] - m;
		hi[a] = h[a] + m;
	}
}
constexpr int kCanonStack = 64; // >= the binary tree's depth, checked at build

__device__ float canon_distance(const CanonView& C, f3 p, unsigned long long& nvis, unsigned long long& ntri)
{
	float best = kPenumbraSize;
	nvis++; // the root's box
	if (box_d2(p, C.lo[0], C.lo[1], C.lo[2], C.hi[0], C.hi[1], C.hi[2]) > best * best)
		return best;
	int32_t st[kCanonStack];
	float sk[kCanonStack];
	int sp = 0;
	int32_t ref = 0;
	for (;;) {
		if (ref >= 0) {
			const BvhNode n = C.nodes[ref];
			float da = box_d2(p, n.lo0x, n.lo0y, n.lo0z, n.hi0x, n.hi0y, n.hi0z);
			float db = box_d2(p, n.lo1x, n.lo1y, n.lo1z, n.hi1x, n.hi1y, n.hi1z);
			int32_t a = n.c0, b = n.c1;
			nvis += 2;
			if (db < da) {
				float t = da;
				da = db;
				db = t;
				int32_t u = a;
				a = b;
				b = u;
			}
			const float lim = best * best;
			if (da <= lim) {
				if (db <= lim) {
					st[sp] = b;
					sk[sp] = db;
					sp++;
				}
				ref = a;
				continue;
			}
		} else {
			const uint32_t f = leaf_first(ref), c = leaf_count(ref);
			for (uint32_t i = 0; i < c; ++i) {
				const float d = dist_tri_pre(p, load_dtri(C.dtri, f + i));
				ntri++;
				if (d < best)
					best = d;
			}
		}
		bool found = false;
		while (sp > 0) {
			--sp;
			if (sk[sp] <= best * best) {
				ref = st[sp];
				found = true;
				break;
			}
		}
		if (!found)
			break;
	}
	return best;
}

// One lane per entry of the last direct-light run's march list.  out:
// node visits, triangle tests, distance queries, factors that differ from
// the product's (vis).
__global__ __launch_bounds__(kBlock) void k_march_canon(CanonView C, View V, const float* __restrict__ pos,
                                                        const float* __restrict__ nrm,
                                                        const DevLight* __restrict__ lights,
                                                        const uint32_t* __restrict__ items,
                                                        const uint32_t* __restrict__ mlist, uint32_t nmarch,
                                                        const float* __restrict__ vis, unsigned long long* out)
{
	const uint32_t idx = blockIdx.x * kBlock + threadIdx.x;
	if (idx >= nmarch)
		return;
	const uint32_t k = mlist[idx];
	const uint32_t code = items[k];
	const DevLight& L = lights[code >> kSampleBits];
	const uint32_t s = code & kSampleMask;
	const f3 SP = ld3(pos + 3 * s), SN = ld3(nrm + 3 * s);
	const f3 to = (L.type == 3) ? SP + ld3(L.dir) * L.range : ld3(L.pos);
	March M;
	march_init(M, V, SP + SN * kShadowOffset, to, L.radius);
	unsigned long long nv = 0, nt = 0, nq = 0;
	float res = 1.0f;
	for (float t = M.t; t < M.maxt;) {
		const float h = canon_distance(C, M.from + M.rd * t, nv, nt);
		nq++;
		if (h < 0.001f) {
			res = 0.0f;
			break;
		}
		res = tmin(res, h / tmin(t * M.k, kPenumbraSize));
		const float tn = t + tmin(h, kPenumbraStep);
		if (tn == t)
			break;
		t = tn;
	}
	atomicAdd(out + 0, nv);
	atomicAdd(out + 1, nt);
	atomicAdd(out + 2, nq);
	if (vis && __float_as_uint(res) != __float_as_uint(vis[k]))
		atomicAdd(out + 3, 1ull);
}

// ---------------------------------------------------------------------------
// direct light
// ---------------------------------------------------------------------------
// Shading gate of each (light, texel) item (Job_LMRender_*_Inner up to the
// shadow march, lighter.cpp:485-553, 586): items with a zero shading factor
// get their final weight here; the others are appended to the march list
// with the factor the shadow term multiplies.  A block takes kGateRounds x
// kBlock consecutive items, collects its marches in LDS, sorts them by
// (light, Morton code of the texel position) and appends them with one
// atomic (one per wave on a single counter serialised in L2).  Items come in
// atlas order, rows of texels; after the sort a wave's 64 marches start from
// a compact patch of space towards one light, so their traversals overlap
// (-9 % march time on the terrain workload).  Only the order in which
// marches are taken changes, not any result.
constexpr int kGateRounds = 8;
constexpr size_t kAOChunk = size_t(1) << 20; // texels per AO launch when progress is followed
constexpr uint32_t kMarchBatch = 64; // march-list positions a wave claims at a time
static_assert(kMarchBatch >= 64, "a refill of all 64 lanes must fit in one claim");

__device__ __forceinline__ bool gate_item(const float* __restrict__ pos, const float* __restrict__ nrm,
                                          const DevLight* __restrict__ lights, const uint32_t* __restrict__ items,
                                          uint32_t k, float* __restrict__ gate, float* __restrict__ wout,
                                          float* __restrict__ vout)
{
	bool march = false;
	{
		uint32_t code = items[k];
		const DevLight& L = lights[code >> kSampleBits];
		uint32_t s = code & kSampleMask;
		f3 SP = ld3(pos + 3 * s), SN = ld3(nrm + 3 * s);
		float w = __uint_as_float(kSkipBits), g = 0.0f;
		if (L.type == 3) {
			float fn = tmax(0.0f, dot(ld3(L.dir), SN));
			if (fn > 0) {
				march = true;
				g = fn;
			} else {
				w = fn; // +-0 times a finite non-negative shadow factor
			}
		} else if (L.type == 1 || L.type == 2) {
			f3 s2l = ld3(L.pos) - SP;
			float dist = length(s2l);
			if (dist)
				s2l = s2l / dist;
			float fd = pow_r(1 - tmin(1.0f, dist / L.range), L.power);
			float fn = tmax(0.0f, dot(s2l, SN));
			float gt;
			if (L.type == 2) {
				float ang = acos_r(tmin(1.0f, dot(s2l, -ld3(L.dir))));
				float fdir = tmax(0.0f, tmin(1.0f, (ang - L.spot_out_rad) / L.spot_diff));
				fdir = pow_r(fdir, L.spot_curve);
				gt = fd * fn * fdir;
			} else {
				gt = fd * fn;
			}
			if (!(gt <= 0)) {
				march = true;
				g = gt;
			}
		}
		if (march) {
			gate[k] = g;
		} else {
			wout[k] = w;
			if (vout)
				vout[k] = 1.0f;
		}
	}
	return march;
}

__global__ __launch_bounds__(kBlock) void k_direct_gate(const float* __restrict__ pos, const float* __restrict__ nrm,
                                                        const DevLight* __restrict__ lights,
                                                        const uint32_t* __restrict__ items, uint32_t nitems,
                                                        float* __restrict__ gate, float* __restrict__ wout,
                                                        float* __restrict__ vout, uint32_t* __restrict__ mlist,
                                                        uint32_t* __restrict__ mcount, const uint32_t* __restrict__ skey)
{
	__shared__ uint32_t list[kGateRounds * kBlock];
	__shared__ uint32_t keys[kGateRounds * kBlock];
	__shared__ uint32_t n, base, l0;
	const uint64_t k0 = (uint64_t)blockIdx.x * (kGateRounds * kBlock);
	if (threadIdx.x == 0) {
		n = 0;
		l0 = items[k0] >> kSampleBits;
	}
	__syncthreads();
	const int lane = threadIdx.x & 63;
	for (int r = 0; r < kGateRounds; ++r) {
		const uint64_t k = k0 + (uint64_t)r * kBlock + threadIdx.x;
		const bool march = k < nitems && gate_item(pos, nrm, lights, items, (uint32_t)k, gate, wout, vout);
		const uint64_t m = __ballot(march);
		uint32_t wbase = 0;
		if (lane == 0 && m)
			wbase = atomicAdd(&n, (uint32_t)__popcll(m)); // LDS atomic, one per wave
		wbase = (uint32_t)__shfl((int)wbase, 0);
		if (march) {
			const uint32_t slot = wbase + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
			const uint32_t code = items[k];
			const uint32_t dl = min((code >> kSampleBits) - l0, 2u); // keys stay below the padding 0xFFFFFFFF
			list[slot] = (uint32_t)k;
			keys[slot] = (dl << 30) | skey[code & kSampleMask];
		}
	}
	__syncthreads();
	// sort the block's marches by (light, Morton code of the texel)
	uint32_t N = 1;
	while (N < n)
		N <<= 1;
	for (uint32_t i = n + threadIdx.x; i < N; i += kBlock)
		keys[i] = 0xFFFFFFFFu;
	for (uint32_t size = 2; size <= N; size <<= 1)
		for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
			__syncthreads();
			for (uint32_t i = threadIdx.x; i < N / 2; i += kBlock) {
				const uint32_t lo = 2 * stride * (i / stride) + (i & (stride - 1)), hi = lo + stride;
				const bool up = (lo & size) == 0;
				const uint32_t a = keys[lo], b = keys[hi];
				if ((a > b) == up) {
					keys[lo] = b;
					keys[hi] = a;
					const uint32_t t = list[lo];
					list[lo] = list[hi];
					list[hi] = t;
				}
			}
		}
	__syncthreads();
	if (threadIdx.x == 0)
		base = n ? atomicAdd(mcount, n) : 0u;
	__syncthreads();
	for (uint32_t i = threadIdx.x; i < n; i += kBlock)
		mlist[base + i] = list[i];
}

// Persistent shadow-march kernel: every lane pulls a new march from the list
// as soon as its current one ends (one atomic per wave per refill), so the
// very uneven march lengths do not idle the wave.  Every wave leaves once the
// list is exhausted and its lanes' marches have ended.  ctl (optional, host
// memory): the claims of positions 0, 64 x kMarchBatch, .. post the
// position to ctl[1] (progress) and, when ctl[0] is set, raise the device
// flag next[1]; a wave that finds next[1] set at a claim leaves at once
// (abort), so every wave stops within one claim of the first such post.  (A variant that
// interleaves single traversal actions of all lanes in one loop measured
// 37% slower on the terrain workload.)
// The traversal stack's first levels live in LDS (dynamic, kBlock * depth * 6
// bytes: an int32 reference and an FP16 key per level and lane), depth =
// min(kLdsLevels, levels of the BVH plus margin).
template <bool COUNT>
__global__ __launch_bounds__(kBlock) void k_march(View V, const float* __restrict__ pos,
                                                  const float* __restrict__ nrm, const DevLight* __restrict__ lights,
                                                  const uint32_t* __restrict__ items,
                                                  const uint32_t* __restrict__ mlist, uint32_t nmarch,
                                                  const float* __restrict__ gate, float* __restrict__ wout,
                                                  float* __restrict__ vout, uint32_t* __restrict__ next,
                                                  unsigned long long* cnt, int depth, uint32_t* ctl)
{
	extern __shared__ __align__(16) unsigned char lds_stack[];
	LdsStack st = lds_stack_of_lane(lds_stack, depth);
	bool active = false, done = false;
	uint32_t k = 0;
	March M;
	// the wave's reservation [rbeg, rend) of march-list positions: refills
	// take from it and claim kMarchBatch more with one atomic when it runs
	// short, so a wave keeps working on neighbouring items (one texel region
	// and light) and the list counter sees one atomic per kMarchBatch marches
	const int lane = threadIdx.x & 63;
	uint32_t rbeg = 0, rend = 0;
	for (;;) {
		bool need = !active && !done;
		const uint64_t nm = __ballot(need);
		if (nm) {
			const uint32_t cnt_need = (uint32_t)__popcll(nm);
			const uint32_t rank = (uint32_t)__popcll(nm & ((1ull << lane) - 1ull));
			const uint32_t avail = rend - rbeg;
			uint32_t idx;
			if (avail >= cnt_need) {
				idx = rbeg + rank;
				rbeg += cnt_need;
			} else {
				uint32_t nb = 0, stop = 0;
				if (lane == 0) {
					// every claim reads the device abort flag next[1] (an L2
					// access in flight together with the claim's atomic);
					// one claim in 64 also makes the host round trip:
					// progress out, the host's abort flag in, passed on to
					// next[1] (a host round trip at every claim cost 3 %)
					if (ctl)
						stop = __hip_atomic_load(next + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
					nb = atomicAdd(next, kMarchBatch);
					if (ctl && ((nb / kMarchBatch) & 63u) == 0) {
						__hip_atomic_store(ctl + 1, nb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
						if (__hip_atomic_load(ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM))
							__hip_atomic_store(next + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
					}
				}
				nb = (uint32_t)__shfl((int)nb, 0);
				if (__shfl((int)stop, 0))
					break; // aborted: the wave leaves its marches unfinished
				idx = rank < avail ? rbeg + rank : nb + (rank - avail);
				rbeg = nb + (cnt_need - avail);
				rend = nb + kMarchBatch;
			}
			if (need) {
				if (idx < nmarch) {
					k = mlist[idx];
					uint32_t code = items[k];
					const DevLight& L = lights[code >> kSampleBits];
					uint32_t s = code & kSampleMask;
					f3 SP = ld3(pos + 3 * s), SN = ld3(nrm + 3 * s);
					f3 to = (L.type == 3) ? SP + ld3(L.dir) * L.range : ld3(L.pos);
					march_init(M, V, SP + SN * kShadowOffset, to, L.radius);
					active = true;
				} else {
					done = true;
				}
			}
		}
		if (__all(done))
			break;
		if (active && march_step<COUNT>(M, V, cnt, st)) {
			// weight = (shading factor) * shadow factor, in the reference's order
			wout[k] = gate[k] * M.res;
			if (vout)
				vout[k] = M.res;
			if (COUNT) {
				atomicAdd(cnt + 2, (unsigned long long)M.steps);
				atomicAdd(cnt + 4, 1ull);
			}
			active = false;
		}
	}
}

// lightmap[s] += color * w for the items of one light (no sample repeats
// within a light, so no two lanes touch the same texel).
__global__ __launch_bounds__(kBlock) void k_accumulate(const uint32_t* __restrict__ items,
                                                       const float* __restrict__ w, uint32_t begin, uint32_t end,
                                                       float cr, float cg, float cb, float* __restrict__ lm)
{
	uint32_t k = begin + blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= end)
		return;
	float wk = w[k];
	if (__float_as_uint(wk) == kSkipBits)
		return;
	uint32_t s = items[k] & kSampleMask;
	f3 c = mk3(cr, cg, cb);
	f3 v = ld3(lm + 3 * s) + c * wk;
	st3(lm + 3 * s, v);
}

// ---------------------------------------------------------------------------
// Reverse mode of the accumulation above: the lightmap gains color_l * w
// per item (lighter.cpp:505-513), so dL/dcolor_l[c] = sum over light l's
// items of w * lam[s][c], lam = dL/dlightmap per sample.  One launch, grid
// (kGradBlocks, lights); each block writes FP64 partial sums, reduced in a
// fixed order on the host.
// ---------------------------------------------------------------------------
constexpr int kGradBlocks = 256; // blocks per light

__global__ __launch_bounds__(kBlock) void k_color_grad(const uint32_t* __restrict__ items,
                                                       const float* __restrict__ w,
                                                       const uint32_t* __restrict__ light_off,
                                                       const double* __restrict__ lam, double* __restrict__ part)
{
	const uint32_t l = blockIdx.y;
	const uint32_t begin = light_off[l], end = light_off[l + 1];
	double a0 = 0, a1 = 0, a2 = 0;
	for (uint32_t k = begin + blockIdx.x * kBlock + threadIdx.x; k < end; k += kGradBlocks * kBlock) {
		float wk = w[k];
		if (__float_as_uint(wk) == kSkipBits)
			continue;
		uint32_t s = items[k] & kSampleMask;
		a0 += (double)wk * lam[3 * (uint64_t)s];
		a1 += (double)wk * lam[3 * (uint64_t)s + 1];
		a2 += (double)wk * lam[3 * (uint64_t)s + 2];
	}
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) {
		a0 += __shfl_xor(a0, o);
		a1 += __shfl_xor(a1, o);
		a2 += __shfl_xor(a2, o);
	}
	__shared__ double red[kBlock / 64][3];
	const int wave = threadIdx.x >> 6;
	if ((threadIdx.x & 63) == 0) {
		red[wave][0] = a0;
		red[wave][1] = a1;
		red[wave][2] = a2;
	}
	__syncthreads();
	if (threadIdx.x < 3) {
		double t = 0;
		for (int q = 0; q < kBlock / 64; ++q)
			t += red[q][threadIdx.x];
		part[3 * ((uint64_t)l * kGradBlocks + blockIdx.x) + threadIdx.x] = t;
	}
}

// ---------------------------------------------------------------------------
// ambient occlusion (lighter.cpp:799-841)
// ---------------------------------------------------------------------------
// a value that every lane of the wave holds, kept in a scalar register
__device__ __forceinline__ float wave_uniform(float x)
{
	return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

struct AOParams {
	int n;
	float distance, multiplier, falloff, effect;
	float color[3];
};

// One wave per texel, one lane per ray (rays beyond 64 in further rounds of
// the wave): the rays of a texel start together and stay coherent, and a
// small scene still fills the GPU.  The rays' terms are then added in ray
// order, the reference's summation order, so the sum is the same.
template <bool COUNT>
// at most 80 VGPRs (6 waves per SIMD instead of 5): -3 % on the terrain's
// 191 M rays; 8 waves spill and lose
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_ao(View V, const float* __restrict__ pos, const float* __restrict__ nrm,
                                               const float* __restrict__ randoff, const float2* __restrict__ side,
                                               AOParams P, uint32_t s0, uint32_t S, float* __restrict__ lm,
                                               unsigned long long* cnt, int depth, float* __restrict__ fac_out)
{
	extern __shared__ __align__(16) unsigned char lds_stack[];
	LdsStack st = lds_stack_of_lane(lds_stack, depth);
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	// The wave's texel is uniform, so its normal, ray origin and rotation
	// offset stay in scalar registers across the traversals: in vector
	// registers they pushed the kernel past its 80-register budget, and the
	// spills wrote about 1.5 KB per wave to scratch memory.
	const uint32_t s = __builtin_amdgcn_readfirstlane(s0 + blockIdx.x * (kBlock / 64) + wave); // texels [s0, S)
	if (s >= S)
		return; // whole waves: no barrier below
	const f3 SP = ld3(pos + 3 * s), SNv = ld3(nrm + 3 * s);
	const f3 SN = mk3(wave_uniform(SNv.x), wave_uniform(SNv.y), wave_uniform(SNv.z));
	const f3 ov = SP + SN * (kSmall * 2);
	const f3 origin = mk3(wave_uniform(ov.x), wave_uniform(ov.y), wave_uniform(ov.z));
	const float ro = wave_uniform(randoff[s]);
	// The rays (at most P.distance <= kGridReach long) start at the origin's
	// entry-grid cell; its node is fetched here, before the directions are
	// computed, so that the fetch overlaps that arithmetic (taken at the
	// first walk's start instead, it was one exposed miss per texel and
	// cost more than the levels it skips).
	const int32_t start = P.distance <= kGridReach ? entry_node(V, origin) : 0;
	const uint32_t touch = reinterpret_cast<const uint32_t*>(V.nodes + start)[lane & 31];
	float ao = 0;
	for (int i0 = 0; i0 < P.n; i0 += 64) {
		const int i = i0 + lane;
		float term = 0.0f;
		if (i < P.n) {
			float angle = (i + ro) * kGoldenAngle;
			double sa, ca;
			sincos((double)angle, &sa, &ca);
			float2 sd = side[i];
			f3 d = spiral_dir(SN, sd.x, sd.y, (float)ca, (float)sa) * P.distance;
			asm volatile("" ::"v"(touch));
			float hit = q_segment<false, COUNT>(V, origin, origin + d, cnt, st, start);
			if (hit < 1.0f)
				term = 1.0f - hit; // a miss adds nothing (ao + 0 = ao)
		}
		// every lane adds the 64 terms in ray order (the same sum on every lane)
		const int m = P.n - i0 < 64 ? P.n - i0 : 64;
		for (int k = 0; k < m; ++k)
			ao += __shfl(term, k);
	}
	if (lane != 0)
		return;
	ao /= P.n;
	ao = tmin(ao * P.multiplier, 1.0f);
	if (P.falloff)
		ao = pow_r(ao, P.falloff);
	if (fac_out)
		fac_out[s] = ao;
	f3 out = ld3(lm + 3 * s);
	f3 col = mk3(P.color[0], P.color[1], P.color[2]);
	if (P.effect >= 0) {
		out = out * (1 - ao * (1 - P.effect)) + col * ao;
	} else {
		out = lerp3(out, lerp3(col, col * out, -P.effect), ao);
	}
	st3(lm + 3 * s, out);
}

// Closest hit of DistanceTestBBT (lighter.cpp:254-289) on the binary tree
// with the canonical semantics: nearest child first by entry parameter, the
// segment's far end shrinking to the closest hit found.  Counts as
// canon_distance does.
// ANY = true: any hit, the walk ending at the first hit with r < 1 (the
// reference's VisibilityTest, lighter.cpp:112-147).
template <bool ANY>
__device__ float canon_segment(const CanonView& C, const RayTri* __restrict__ rtri, f3 A, f3 B,
                               unsigned long long& nvis, unsigned long long& ntri)
{
	float best = 2.0f;
	const Seg s = make_seg(A, B);
	nvis++; // the root's box
	if (slab(s, C.lo[0], C.lo[1], C.lo[2], C.hi[0], C.hi[1], C.hi[2], 1.0f) == INFINITY)
		return best;
	int32_t st[kCanonStack];
	float sk[kCanonStack];
	int sp = 0;
	int32_t ref = 0;
	for (;;) {
		if (ref >= 0) {
			const BvhNode n = C.nodes[ref];
			const float tl = fminf(best, 1.0f);
			float ta = slab(s, n.lo0x, n.lo0y, n.lo0z, n.hi0x, n.hi0y, n.hi0z, tl);
			float tb = slab(s, n.lo1x, n.lo1y, n.lo1z, n.hi1x, n.hi1y, n.hi1z, tl);
			int32_t a = n.c0, b = n.c1;
			nvis += 2;
			if (tb < ta) {
				float t = ta;
				ta = tb;
				tb = t;
				int32_t u = a;
				a = b;
				b = u;
			}
			if (ta != INFINITY) {
				if (tb != INFINITY) {
					st[sp] = b;
					sk[sp] = tb;
					sp++;
				}
				ref = a;
				continue;
			}
		} else {
			const uint32_t f = leaf_first(ref), c = leaf_count(ref);
			for (uint32_t i = 0; i < c; ++i) {
				const float r = seg_tri_pre(s.o, s.dir, load_rtri(rtri, f + i));
				ntri++;
				if (r < best)
					best = r;
				if (ANY && best < 1.0f)
					return best;
			}
		}
		bool found = false;
		while (sp > 0) {
			--sp;
			if (sk[sp] <= fminf(best, 1.0f)) {
				ref = st[sp];
				found = true;
				break;
			}
		}
		if (!found)
			break;
	}
	return best;
}

// The AO factor of every texel (Job_AORender_Inner, lighter.cpp:812-831)
// with the rays traced by canon_segment: one wave per texel, one lane per
// ray, the terms added in ray order, as k_ao does.  out: node visits,
// triangle tests, rays.
__global__ __launch_bounds__(kBlock) void k_ao_canon(CanonView C, const RayTri* __restrict__ rtri,
                                                     const float* __restrict__ pos, const float* __restrict__ nrm,
                                                     const float* __restrict__ randoff,
                                                     const float2* __restrict__ side, AOParams P, uint32_t S,
                                                     float* __restrict__ fac_out, unsigned long long* out)
{
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const uint32_t s = blockIdx.x * (kBlock / 64) + wave;
	if (s >= S)
		return;
	f3 SP = ld3(pos + 3 * s), SN = ld3(nrm + 3 * s);
	f3 origin = SP + SN * (kSmall * 2);
	float ro = randoff[s];
	float ao = 0;
	unsigned long long nv = 0, nt = 0, nr = 0;
	for (int i0 = 0; i0 < P.n; i0 += 64) {
		const int i = i0 + lane;
		float term = 0.0f;
		if (i < P.n) {
			float angle = (i + ro) * kGoldenAngle;
			double sa, ca;
			sincos((double)angle, &sa, &ca);
			float2 sd = side[i];
			f3 d = spiral_dir(SN, sd.x, sd.y, (float)ca, (float)sa) * P.distance;
			float hit = canon_segment<false>(C, rtri, origin, origin + d, nv, nt);
			nr++;
			if (hit < 1.0f)
				term = 1.0f - hit;
		}
		const int m = P.n - i0 < 64 ? P.n - i0 : 64;
		for (int k = 0; k < m; ++k)
			ao += __shfl(term, k);
	}
	atomicAdd(out + 0, nv);
	atomicAdd(out + 1, nt);
	atomicAdd(out + 2, nr);
	if (lane != 0)
		return;
	ao /= P.n;
	ao = tmin(ao * P.multiplier, 1.0f);
	if (P.falloff)
		ao = pow_r(ao, P.falloff);
	fac_out[s] = ao;
}

// The occlusion test of every candidate pair of the last link search
// (Int_RDGenLinks, lighter.cpp:751) with the canonical any-hit walk; out:
// node visits, triangle tests, pairs whose link bit differs from the
// product's.
__global__ __launch_bounds__(kBlock) void k_link_canon(CanonView C, const RayTri* __restrict__ rtri,
                                                       const float* __restrict__ pos,
                                                       const uint2* __restrict__ cand, uint64_t ncand,
                                                       const uint8_t* __restrict__ linked_by_product,
                                                       unsigned long long* out)
{
	const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= ncand)
		return;
	const uint2 c = cand[k];
	unsigned long long nv = 0, nt = 0;
	const bool occluded = canon_segment<true>(C, rtri, ld3(pos + 3 * c.x), ld3(pos + 3 * c.y), nv, nt) < 1.0f;
	const bool linked = linked_by_product[k] != 0;
	atomicAdd(out + 0, nv);
	atomicAdd(out + 1, nt);
	if (occluded != linked)
		atomicAdd(out + 2, 1ull);
}

// ---------------------------------------------------------------------------
// point-query kernels (tests, timing)
// ---------------------------------------------------------------------------
__global__ void k_q_distance(View V, const float* p, uint32_t n, float* out)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		out[i] = q_distance<false>(V, ld3(p + 3 * i), kPenumbraSize, nullptr);
}

__global__ void k_q_shadow(View V, const float* a, const float* b, const float* k, uint32_t n, float* out,
                           uint32_t* steps)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) {
		uint32_t st = 0;
		out[i] = shadow_march<false>(V, ld3(a + 3 * i), ld3(b + 3 * i), k[i], nullptr, &st);
		steps[i] = st;
	}
}

__global__ void k_q_closest(View V, const float* a, const float* b, uint32_t n, float* out)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		out[i] = q_segment<false, false>(V, ld3(a + 3 * i), ld3(b + 3 * i), nullptr);
}

__global__ void k_q_occluded(View V, const float* a, const float* b, uint32_t n, uint8_t* out)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		out[i] = q_segment<true, false>(V, ld3(a + 3 * i), ld3(b + 3 * i), nullptr) < 1.0f ? 1 : 0;
}

// ---------------------------------------------------------------------------
// radiosity links (lighter.cpp:728-759)
// ---------------------------------------------------------------------------
// Form factor of the pair (i < j) exactly as the reference evaluates it, or
// -1 when the pair fails the geometric thresholds (NaN inputs fail neither
// comparison, as in the reference, and give a NaN factor).
__device__ __forceinline__ float pair_factor(f3 ap, f3 an, f3 bp, f3 bn)
{
	f3 d = bp - ap;
	float da = dot(an, d);
	float db = dot(bn, -d);
	if (da <= kSmall || db <= kSmall)
		return -1.0f;
	float l2 = lensq(d);
	float f = da * db / (l2 * l2 * kPiF);
	if (f < kSmall)
		return -1.0f;
	return f;
}

constexpr int kPairChunk = 256;

// Pass 0 counts, pass 1 writes: candidates of row i in columns
// [c*kPairChunk, +kPairChunk) restricted to j > i, in ascending j.
template <bool WRITE>
__global__ __launch_bounds__(kBlock) void k_pairs(const float* __restrict__ pos, const float* __restrict__ nrm,
                                                  uint32_t N, uint32_t nchunks, uint32_t* __restrict__ counts,
                                                  const uint64_t* __restrict__ offsets, uint2* __restrict__ cand)
{
	uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	uint32_t c = blockIdx.y;
	if (i >= N)
		return;
	uint32_t j0 = c * kPairChunk, j1 = min(N, j0 + kPairChunk);
	if (j0 <= i)
		j0 = i + 1;
	uint32_t n = 0;
	uint64_t base = WRITE ? offsets[(uint64_t)i * nchunks + c] : 0;
	if (j0 < j1) {
		f3 ap = ld3(pos + 3 * i), an = ld3(nrm + 3 * i);
		for (uint32_t j = j0; j < j1; ++j) {
			float f = pair_factor(ap, an, ld3(pos + 3 * j), ld3(nrm + 3 * j));
			if (!(f < 0.0f)) { // -1: failed; NaN passes, as in the reference's comparisons
				if (WRITE)
					cand[base + n] = make_uint2(i, j);
				n++;
			}
		}
	}
	if (!WRITE)
		counts[(uint64_t)i * nchunks + c] = n;
}

// ---------------------------------------------------------------------------
// Link candidates over a uniform grid of the patches.  A pair passes the
// factor thresholds only when f = dA dB / (pi l^4) >= kSmall, and dA, dB <=
// |n| l, so l^2 <= |nA| |nB| / (pi kSmall): 17.8 m for unit normals.  With
// cells at least that wide every pair that can pass lies in the same or a
// neighbouring cell; those pairs are tested with the arithmetic of the
// all-pairs search (pair_factor, the lower index as A), so both searches
// find the same candidate set, in another order.
// ---------------------------------------------------------------------------
struct LinkGrid {
	float lo[3];
	float inv;     // 1 / cell width
	uint32_t n[3]; // cells per axis (product below 2^24)
	int reach;     // cells searched on each side: the cells are 1 / reach of the linking distance wide
};

__device__ __forceinline__ uint32_t grid_cell(const LinkGrid& G, f3 p)
{
	const float q[3] = {(p.x - G.lo[0]) * G.inv, (p.y - G.lo[1]) * G.inv, (p.z - G.lo[2]) * G.inv};
	uint32_t c[3];
#pragma unroll
	for (int a = 0; a < 3; ++a)
		c[a] = min(G.n[a] - 1, (uint32_t)floorf(q[a])); // p >= lo: q >= 0
	return (c[2] * G.n[1] + c[1]) * G.n[0] + c[0];
}

__global__ void k_grid_count(const float* __restrict__ pos, uint32_t N, LinkGrid G, uint32_t* __restrict__ cell,
                             uint32_t* __restrict__ ccount)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= N)
		return;
	const uint32_t k = grid_cell(G, ld3(pos + 3 * i));
	cell[i] = k;
	atomicAdd(ccount + k, 1u);
}

__global__ void k_grid_fill(uint32_t N, const uint32_t* __restrict__ cell, const uint64_t* __restrict__ cstart,
                            uint32_t* __restrict__ cfill, uint32_t* __restrict__ order)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= N)
		return;
	const uint32_t k = cell[i];
	order[cstart[k] + atomicAdd(cfill + k, 1u)] = i;
}

// Pass 0 counts, pass 1 writes the candidates (i, j > i) of row i from the
// (2 reach + 1)^3 cells around i's cell; thread t takes the t-th patch in
// cell order.
template <bool WRITE>
__global__ __launch_bounds__(kBlock) void k_pairs_grid(const float* __restrict__ pos, const float* __restrict__ nrm,
                                                       uint32_t N, LinkGrid G, const uint32_t* __restrict__ cell,
                                                       const uint64_t* __restrict__ cstart,
                                                       const uint32_t* __restrict__ order,
                                                       uint32_t* __restrict__ counts,
                                                       const uint64_t* __restrict__ offs, uint2* __restrict__ cand)
{
	const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= N)
		return;
	const uint32_t i = order[t]; // neighbouring lanes share cells: they walk the same lists
	const uint32_t k = cell[i];
	const int cx = (int)(k % G.n[0]), cy = (int)((k / G.n[0]) % G.n[1]), cz = (int)(k / (G.n[0] * G.n[1]));
	const f3 ap = ld3(pos + 3 * i), an = ld3(nrm + 3 * i);
	uint32_t n = 0;
	const uint64_t base = WRITE ? offs[i] : 0;
	const int q = G.reach;
	for (int z = max(cz - q, 0); z <= min(cz + q, (int)G.n[2] - 1); ++z)
		for (int y = max(cy - q, 0); y <= min(cy + q, (int)G.n[1] - 1); ++y)
			for (int x = max(cx - q, 0); x <= min(cx + q, (int)G.n[0] - 1); ++x) {
				const uint32_t cc = ((uint32_t)z * G.n[1] + (uint32_t)y) * G.n[0] + (uint32_t)x;
				const uint64_t e = cstart[cc + 1];
				uint64_t q = cstart[cc];
				// four patches per round: their loads are in flight together
				for (; q + 4 <= e; q += 4) {
					uint32_t j[4];
					f3 bp[4], bn[4];
#pragma unroll
					for (int u = 0; u < 4; ++u)
						j[u] = order[q + u];
#pragma unroll
					for (int u = 0; u < 4; ++u) {
						bp[u] = ld3(pos + 3 * j[u]);
						bn[u] = ld3(nrm + 3 * j[u]);
					}
#pragma unroll
					for (int u = 0; u < 4; ++u)
						if (j[u] > i && !(pair_factor(ap, an, bp[u], bn[u]) < 0.0f)) {
							if (WRITE)
								cand[base + n] = make_uint2(i, j[u]);
							n++;
						}
				}
				for (; q < e; ++q) {
					const uint32_t j = order[q];
					if (j <= i)
						continue;
					if (!(pair_factor(ap, an, ld3(pos + 3 * j), ld3(nrm + 3 * j)) < 0.0f)) {
						if (WRITE)
							cand[base + n] = make_uint2(i, j);
						n++;
					}
				}
			}
	if (!WRITE)
		counts[i] = n;
}

// The occlusion test of each candidate: the pair links when the segment
// between the patches hits a shadow caster (VisibilityTest returns the
// any-hit query's hit, lighter.cpp:137-147, and Int_RDGenLinks keeps the
// pair when it is true, lighter.cpp:751-752).  Each link (i, j > i) also
// counts toward row j's links below the diagonal; the candidates of one row
// are contiguous, so the lanes of a wave add to different rows.
__global__ __launch_bounds__(kBlock) void k_link_test(View V, const float* __restrict__ pos,
                                                      const uint2* __restrict__ cand, uint64_t ncand,
                                                      uint8_t* __restrict__ linked, uint32_t* __restrict__ lower,
                                                      int depth)
{
	extern __shared__ __align__(16) unsigned char lds_stack[];
	LdsStack st = lds_stack_of_lane(lds_stack, depth);
	const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= ncand)
		return;
	const uint2 c = cand[k];
	const bool l = q_segment<true, false>(V, ld3(pos + 3 * c.x), ld3(pos + 3 * c.y), nullptr, st) < 1.0f;
	linked[k] = l ? 1 : 0;
	if (l)
		atomicAdd(lower + c.y, 1u);
}

// Links of row i above the diagonal: its own linked candidates, one wave
// per row.  With colx null, their count goes to upper[i]; otherwise they
// are written after the row's lower part, in candidate order.
__global__ __launch_bounds__(kBlock) void k_link_upper(uint32_t N, const uint64_t* __restrict__ coff,
                                                       const uint2* __restrict__ cand,
                                                       const uint8_t* __restrict__ linked,
                                                       uint32_t* __restrict__ upper,
                                                       const uint64_t* __restrict__ rowptr,
                                                       const uint32_t* __restrict__ lower, uint32_t* __restrict__ colx)
{
	const uint32_t i = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	if (i >= N)
		return; // the whole wave
	const uint64_t e = coff[i + 1];
	uint64_t at = colx ? rowptr[i] + lower[i] : 0;
	for (uint64_t k0 = coff[i]; k0 < e; k0 += 64) {
		const uint64_t k = k0 + lane;
		const bool l = k < e && linked[k];
		const uint64_t m = __ballot(l);
		if (colx && l)
			colx[at + __popcll(m & ((1ull << lane) - 1ull))] = cand[k].y;
		at += __popcll(m);
	}
	if (!colx && lane == 0)
		upper[i] = (uint32_t)at;
}

// Links of row j below the diagonal: every linked candidate (i, j) puts i
// into row j's lower part (any order; the row sort orders it)
__global__ void k_link_lower(const uint2* __restrict__ cand, const uint8_t* __restrict__ linked, uint64_t ncand,
                             const uint64_t* __restrict__ rowptr, uint32_t* __restrict__ lfill,
                             uint32_t* __restrict__ colx)
{
	const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= ncand || !linked[k])
		return;
	const uint2 c = cand[k];
	colx[rowptr[c.y] + atomicAdd(lfill + c.y, 1u)] = c.x;
}

__global__ void k_row_sum(uint32_t N, const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                          uint32_t* __restrict__ out)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < N)
		out[i] = a[i] + b[i];
}

// factor of every link (columns ascending within a row), one wave per row;
// the reference evaluates each pair with the lower index as A
__global__ __launch_bounds__(kBlock) void k_link_factors(uint32_t N, const uint64_t* __restrict__ rowptr,
                                                         const uint32_t* __restrict__ col,
                                                         const float* __restrict__ pos,
                                                         const float* __restrict__ nrm, float* __restrict__ val)
{
	const uint32_t i = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	if (i >= N)
		return;
	const f3 ip = ld3(pos + 3 * i), in_ = ld3(nrm + 3 * i);
	const uint64_t e = rowptr[i + 1];
	for (uint64_t o = rowptr[i] + lane; o < e; o += 64) {
		const uint32_t j = col[o];
		const f3 jp = ld3(pos + 3 * j), jn = ld3(nrm + 3 * j);
		val[o] = (i < j) ? pair_factor(ip, in_, jp, jn) : pair_factor(jp, jn, ip, in_);
	}
}

// first candidate of every row from the all-pairs search's per-(row, chunk)
// offsets
__global__ void k_row_first(uint32_t N, uint32_t nchunks, const uint64_t* __restrict__ offs, uint64_t total,
                            uint64_t* __restrict__ coff)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i <= N)
		coff[i] = i < N ? offs[(uint64_t)i * nchunks] : total;
}

__global__ void k_u64_to_u32(const uint64_t* __restrict__ in, uint64_t n, uint32_t* __restrict__ out)
{
	const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		out[i] = (uint32_t)in[i];
}

// One bounce of Int_RDBounce (lighter.cpp:767-797) as a gather over the
// symmetric links: in_i = sum_j ((out_j * diffuse_j) * area_j) * f_ij.  The
// factor-free part x_j = (out_j * diffuse_j) * area_j is formed once per
// patch (the same products in the same order as the reference's per-link
// expression); one wave per row then sums x_j * f_ij over the row's links,
// lane l taking links l, l + 64, .., and the lanes' partial sums meet in a
// fixed butterfly (deterministic, and rows of thousands of links read
// their columns and factors in coalesced runs).
__global__ void k_bounce_prep(uint32_t N, const float* __restrict__ out, const float* __restrict__ diffuse,
                              const float* __restrict__ area, float4* __restrict__ x)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= N)
		return;
	const f3 v = (ld3(out + 3 * i) * ld3(diffuse + 3 * i)) * area[i];
	x[i] = make_float4(v.x, v.y, v.z, 0.0f);
}

__global__ __launch_bounds__(kBlock) void k_bounce_rows(uint32_t N, const uint64_t* __restrict__ rowptr,
                                                        const uint32_t* __restrict__ col,
                                                        const float* __restrict__ val, const float4* __restrict__ x,
                                                        float* __restrict__ in, float* __restrict__ total)
{
	const uint32_t i = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	if (i >= N)
		return; // the whole wave
	float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
	const uint64_t e = rowptr[i + 1];
	for (uint64_t k = rowptr[i] + lane; k < e; k += 64) {
		const float4 v = x[col[k]];
		const float f = val[k];
		a0 += v.x * f;
		a1 += v.y * f;
		a2 += v.z * f;
	}
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) {
		a0 += __shfl_xor(a0, o);
		a1 += __shfl_xor(a1, o);
		a2 += __shfl_xor(a2, o);
	}
	if (lane == 0) {
		const f3 acc = mk3(a0, a1, a2);
		st3(in + 3 * i, acc);
		st3(total + 3 * i, ld3(total + 3 * i) + acc);
	}
}

// ---------------------------------------------------------------------------
// exclusive scan of u32 counts into u64 offsets (three-phase)
// ---------------------------------------------------------------------------
__global__ void k_scan_blocks(const uint32_t* __restrict__ in, uint64_t n, uint64_t* __restrict__ out,
                              uint64_t* __restrict__ block_sums)
{
	__shared__ uint64_t sh[kBlock];
	constexpr int kPer = 16;
	uint64_t base = (uint64_t)blockIdx.x * kBlock * kPer + (uint64_t)threadIdx.x * kPer;
	uint64_t local[kPer];
	uint64_t s = 0;
	for (int k = 0; k < kPer; ++k) {
		uint64_t idx = base + k;
		local[k] = s;
		s += (idx < n) ? in[idx] : 0;
	}
	sh[threadIdx.x] = s;
	__syncthreads();
	for (int off = 1; off < kBlock; off <<= 1) {
		uint64_t v = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
		__syncthreads();
		sh[threadIdx.x] += v;
		__syncthreads();
	}
	uint64_t excl = sh[threadIdx.x] - s;
	for (int k = 0; k < kPer; ++k) {
		uint64_t idx = base + k;
		if (idx < n)
			out[idx] = excl + local[k];
	}
	if (threadIdx.x == kBlock - 1)
		block_sums[blockIdx.x] = sh[threadIdx.x];
}

__global__ void k_scan_sums(uint64_t* sums, uint32_t nb, uint64_t* total)
{
	if (threadIdx.x == 0 && blockIdx.x == 0) {
		uint64_t s = 0;
		for (uint32_t b = 0; b < nb; ++b) {
			uint64_t v = sums[b];
			sums[b] = s;
			s += v;
		}
		*total = s;
	}
}

__global__ void k_scan_add(uint64_t* out, uint64_t n, const uint64_t* sums)
{
	constexpr int kPer = 16;
	uint64_t base = (uint64_t)blockIdx.x * kBlock * kPer + (uint64_t)threadIdx.x * kPer;
	uint64_t add = sums[blockIdx.x];
	for (int k = 0; k < kPer; ++k)
		if (base + k < n)
			out[base + k] += add;
}

// ===========================================================================
// host side
// ===========================================================================
static inline unsigned grid_for(uint64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }
static inline unsigned gridn_waves(uint64_t n) { return (unsigned)((n + kBlock / 64 - 1) / (kBlock / 64)); }

// The HIP runtime draws from (or reseeds) the C library's random generator
// while it initialises (the stream, memory and code-object set-up behind the
// process's first device scene).  A bake replays that generator's sequence
// for the AO rotation offsets, one draw per texel as the reference's randf()
// makes them (lighter.cpp:819, lighter_int.hpp:271), so the runtime is
// initialised once per process, on the first GPU request, with a private
// generator state switched in; the caller's sequence is left where it was.
namespace {
struct PrivateRandState {
	char buf[128];
	char* saved;
	PrivateRandState() { saved = initstate(0x4C54u, buf, sizeof(buf)); }
	~PrivateRandState() { setstate(saved); }
};
} // namespace

__global__ void k_warm(uint32_t* p)
{
	p[threadIdx.x] = threadIdx.x;
}

// A stream, device and pinned memory and one kernel launch: the runtime's
// one-time set-up, best effort (errors surface again in the calls that need
// the runtime).
static void init_runtime_once()
{
	hipStream_t s = nullptr;
	if (hipStreamCreate(&s) != hipSuccess)
		return;
	void* d = nullptr;
	if (hipMalloc(&d, 256) == hipSuccess) {
		k_warm<<<1, 64, 0, s>>>((uint32_t*)d);
		(void)hipStreamSynchronize(s);
		(void)hipFree(d);
	}
	void* h = nullptr;
	if (hipHostMalloc(&h, 256, hipHostMallocDefault) == hipSuccess)
		(void)hipHostFree(h);
	(void)hipStreamDestroy(s);
}

void require_gpu()
{
	static std::once_flag once;
	std::call_once(once, [] {
		PrivateRandState private_rand;
		int n = 0;
		if (hipGetDeviceCount(&n) == hipSuccess && n > 0)
			init_runtime_once();
		(void)hipGetLastError();
	});
	int n = 0;
	hipError_t e = hipGetDeviceCount(&n);
	if (e != hipSuccess || n == 0)
		throw std::runtime_error("liblighter_amd: no HIP device available (the bake stages run only on the GPU)");
}

std::string device_name()
{
	hipDeviceProp_t p;
	if (hipGetDeviceProperties(&p, 0) != hipSuccess)
		return "unknown";
	return std::string(p.name) + " " + p.gcnArchName;
}

template <class T> static T* dalloc(size_t n)
{
	void* p = nullptr;
	if (n == 0)
		n = 1;
	HIPCHECK(hipMalloc(&p, n * sizeof(T)));
	return (T*)p;
}

template <class T> static void dfree(T*& p)
{
	if (p)
		(void)hipFree((void*)p);
	p = nullptr;
}

// a device array owned for one scope
template <class T> struct DevArray {
	T* p = nullptr;
	explicit DevArray(size_t n) : p(dalloc<T>(n)) {}
	DevArray(const DevArray&) = delete;
	DevArray& operator=(const DevArray&) = delete;
	~DevArray() { dfree(p); }
};

// ---------------------------------------------------------------------------
// Scratch of the allocation-free direct-light pass (run_direct_noalloc): the
// buffers run_direct allocates (per-item shading factor, the march list, the
// march count and list cursor, the traversal counters, the lights and their
// item ranges) carved out of one file-scope device array.  Each GPU holds
// its own copy for the life of the process; every call on a GPU uses that
// copy, so calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocGateOff = 0;                                          // float per item
constexpr size_t kNoAllocListOff = kNoAllocGateOff + 4 * kNoAllocMaxItems;      // uint32 per item
constexpr size_t kNoAllocHeadOff = kNoAllocListOff + 4 * kNoAllocMaxItems;      // count, cursor, bad-item flag
constexpr size_t kNoAllocCntOff = kNoAllocHeadOff + 256;                        // traversal counters
constexpr size_t kNoAllocLightOff = kNoAllocCntOff + 256;                       // 64 lights
constexpr size_t kNoAllocOffOff = kNoAllocLightOff + 64 * sizeof(DevLight);     // 65 item offsets
constexpr size_t kNoAllocBytes = kNoAllocOffOff + 512;
static_assert(kNumCounters * 8 <= 256, "counters fit their slot");
static_assert(kNoAllocLightOff % 256 == 0 && kNoAllocOffOff % 256 == 0, "aligned slots");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Device address of `scratch` on the current GPU, looked up once per device
// (the array's name is only a host-side stand-in for it).
static void* noalloc_scratch()
{
	static std::mutex mu;
	static void* addr[kMaxDevices] = {};
	int dev = 0;
	HIPCHECK(hipGetDevice(&dev));
	if (dev < 0 || dev >= kMaxDevices)
		throw std::runtime_error("device ordinal beyond the supported " + std::to_string(kMaxDevices));
	std::lock_guard<std::mutex> g(mu);
	if (!addr[dev]) {
		PrivateRandState private_rand; // the lookup may set up the device's code object
		void* p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p)
			throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
		addr[dev] = p;
	}
	return addr[dev];
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t v)
{
	for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
		p[i] = v;
}

// Sets *bad when an item names a light other than its segment's or a sample
// beyond the device's samples (the kernels index samples by the items).
__global__ void k_check_items(const uint32_t* __restrict__ items, uint32_t nitems, const uint32_t* __restrict__ off,
                              uint32_t nlights, uint32_t nsamples, uint32_t* bad)
{
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nitems)
		return;
	const uint32_t code = items[i], l = code >> kSampleBits;
	if (l >= nlights || i < off[l] || i >= off[l + 1] || (code & kSampleMask) >= nsamples)
		atomicOr(bad, 1u);
}

DeviceScene::DeviceScene()
{
	require_gpu();
	hipStream_t s;
	HIPCHECK(hipStreamCreate(&s));
	stream_ = s;
	d_counters_ = dalloc<unsigned long long>(kNumCounters);
	d_mcount_ = dalloc<uint32_t>(3);
	d_scratch_ = noalloc_scratch();
	// the march's abort flag and progress word: pinned, coherent host memory
	// the kernel reads and writes while it runs
	HIPCHECK(hipHostMalloc((void**)&ctl_host_, 2 * sizeof(uint32_t), hipHostMallocMapped | hipHostMallocCoherent));
	ctl_host_[0] = ctl_host_[1] = 0u;
	HIPCHECK(hipHostGetDevicePointer((void**)&ctl_dev_, ctl_host_, 0));
}

// Blocks of the persistent march kernel: as many as can be resident at once.
// LDS stack levels of the traversal kernels (march, AO, link visibility): a
// leaf at depth d has at most d entries below it on the stack; at most
// kLdsLevels levels (18 KB per block) live in LDS
int DeviceScene::march_stack_depth() const { return std::min(kLdsLevels, 3 * bvh_depth_ + 1); }

size_t DeviceScene::march_lds_bytes() const { return (size_t)kBlock * march_stack_depth() * 6; }

unsigned DeviceScene::persistent_blocks()
{
	if (!persist_blocks_) {
		int dev = 0, per_cu = 0;
		hipDeviceProp_t p;
		HIPCHECK(hipGetDevice(&dev));
		HIPCHECK(hipGetDeviceProperties(&p, dev));
		HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_march<false>, kBlock, march_lds_bytes()));
		persist_blocks_ = (unsigned)std::max(1, per_cu) * (unsigned)p.multiProcessorCount;
	}
	return persist_blocks_;
}

DeviceScene::~DeviceScene()
{
	free_all();
	dfree(d_counters_);
	dfree(d_mcount_);
	if (stream_)
		(void)hipStreamDestroy((hipStream_t)stream_);
	if (ctl_host_)
		(void)hipHostFree(ctl_host_);
}

void DeviceScene::free_all()
{
	nm_clear();
	dfree(d_cand_);
	dfree(d_clinked_);
	ncand_ = 0;
	dfree(d_nodes_);
	dfree(d_dtri_);
	dfree(d_rtri_);
	dfree(d_pos_);
	dfree(d_nrm_);
	dfree(d_skey_);
	dfree(d_lm_);
	dfree(d_w_);
	dfree(d_vis_);
	dfree(d_items_);
	dfree(d_lights_);
	dfree(d_gate_);
	dfree(d_mlist_);
	dfree(rad.pos);
	dfree(rad.nrm);
	dfree(rad.rowptr);
	dfree(rad.rowptr64);
	dfree(rad.col);
	dfree(rad.val);
	dfree(rad.diffuse);
	dfree(rad.area);
	dfree(rad.total);
	dfree(rad.out);
	dfree(rad.in);
	dfree(rad.x);
}

static View make_view(void* nodes, void* dtri, void* rtri, size_t ntri, const float* lo, const float* hi,
                      const GridSpec& G, int32_t gbase, int32_t gfine)
{
	View v;
	v.gfine = gfine;
	for (int a = 0; a < 3; ++a) {
		v.glo[a] = G.lo[a];
		v.gn[a] = G.n[a];
	}
	v.ginv = 1.0f / G.g;
	v.gbase = gbase;
	v.nodes = (const WideNode*)nodes;
	v.dtri = (const DistTri*)dtri;
	v.rtri = (const RayTri*)rtri;
	v.has_tris = ntri > 0;
	for (int a = 0; a < 3; ++a) {
		float m = kPenumbraSize * 1.001f + 0.01f + 1e-4f * fmaxf(fabsf(lo[a]), fabsf(hi[a]));
		v.xlo[a] = lo[a] - m;
		v.xhi[a] = hi[a] + m;
	}
	return v;
}

// The triangles' precomputed forms in leaf order (make_dist_tri and
// make_ray_tri, the host's functions, on the device).
__global__ __launch_bounds__(kBlock) void k_tri_forms(const float* __restrict__ tris,
                                                      const uint32_t* __restrict__ order, size_t n,
                                                      DistTri* __restrict__ dt, RayTri* __restrict__ rt)
{
	const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
	if (i >= n)
		return;
	const float* p = tris + 9 * (size_t)order[i];
	const f3 a = ld3(p), c = ld3(p + 3), e = ld3(p + 6);
	dt[i] = make_dist_tri(a, c, e);
	rt[i] = make_ray_tri(a, c, e);
}

__device__ __forceinline__ bool box_meets(const float* lo, const float* hi, float lx, float ly, float lz, float hx,
                                          float hy, float hz)
{
	return lx <= hi[0] && hx >= lo[0] && ly <= hi[1] && hy >= lo[1] && lz <= hi[2] && hz >= lo[2];
}

// One lane per grid cell: its virtual node.  Starting from the root's
// children that meet the cell's region (the cell grown by pad), the largest
// inner node of the set is replaced by its children that meet the region
// while the set stays within four nodes; a node whose children would not
// fit stays.  Children that miss the region hold no triangle meeting it (a
// node's box holds its subtree's triangles; boxes with a NaN bound hold only
// triangles whose distances are NaN, which lower no minimum).
__global__ __launch_bounds__(256) void k_entry_cells(const WideNode* __restrict__ nodes, GridSpec G,
                                                     WideNode* __restrict__ cells)
{
	const uint32_t c = blockIdx.x * 256 + threadIdx.x;
	const uint32_t total = (uint32_t)G.n[0] * G.n[1] * G.n[2];
	if (c >= total)
		return;
	const int ix = (int)(c % (uint32_t)G.n[0]), iy = (int)((c / (uint32_t)G.n[0]) % (uint32_t)G.n[1]),
	          iz = (int)(c / ((uint32_t)G.n[0] * G.n[1]));
	const int ic[3] = {ix, iy, iz};
	float rlo[3], rhi[3];
	for (int a = 0; a < 3; ++a) {
		rlo[a] = G.lo[a] + (float)ic[a] * G.g - G.pad;
		rhi[a] = G.lo[a] + (float)(ic[a] + 1) * G.g + G.pad;
	}
	float blo[4][3], bhi[4][3];
	int32_t ref[4];
	bool open[4];
	int k = 0;
	{
		const WideNode& r = nodes[0];
		for (int j = 0; j < 4; ++j)
			if (box_meets(rlo, rhi, r.lox[j], r.loy[j], r.loz[j], r.hix[j], r.hiy[j], r.hiz[j])) {
				blo[k][0] = r.lox[j]; blo[k][1] = r.loy[j]; blo[k][2] = r.loz[j];
				bhi[k][0] = r.hix[j]; bhi[k][1] = r.hiy[j]; bhi[k][2] = r.hiz[j];
				ref[k] = r.child[j];
				open[k] = r.child[j] >= 0;
				++k;
			}
	}
	for (;;) {
		int e = -1;
		float area = -1.0f;
		for (int i = 0; i < k; ++i) {
			if (!open[i])
				continue;
			const float dx = bhi[i][0] - blo[i][0], dy = bhi[i][1] - blo[i][1], dz = bhi[i][2] - blo[i][2];
			const float s2 = dx * dy + dy * dz + dz * dx;
			if (e < 0 || s2 > area) {
				e = i;
				area = s2;
			}
		}
		if (e < 0)
			break;
		const WideNode& n = nodes[ref[e]];
		int take[4], m = 0;
		for (int j = 0; j < 4; ++j)
			if (box_meets(rlo, rhi, n.lox[j], n.loy[j], n.loz[j], n.hix[j], n.hiy[j], n.hiz[j]))
				take[m++] = j;
		if (k - 1 + m > 4) {
			open[e] = false;
			continue;
		}
		// entry e is replaced by the last entry, then the children appended
		--k;
		for (int a = 0; a < 3; ++a) {
			blo[e][a] = blo[k][a];
			bhi[e][a] = bhi[k][a];
		}
		ref[e] = ref[k];
		open[e] = open[k];
		for (int t = 0; t < m; ++t) {
			const int j = take[t];
			blo[k][0] = n.lox[j]; blo[k][1] = n.loy[j]; blo[k][2] = n.loz[j];
			bhi[k][0] = n.hix[j]; bhi[k][1] = n.hiy[j]; bhi[k][2] = n.hiz[j];
			ref[k] = n.child[j];
			open[k] = n.child[j] >= 0;
			++k;
		}
	}
	WideNode w;
	for (int j = 0; j < 4; ++j) {
		const bool use = j < k;
		w.lox[j] = use ? blo[j][0] : kFarBox;
		w.loy[j] = use ? blo[j][1] : kFarBox;
		w.loz[j] = use ? blo[j][2] : kFarBox;
		w.hix[j] = use ? bhi[j][0] : kFarBox;
		w.hiy[j] = use ? bhi[j][1] : kFarBox;
		w.hiz[j] = use ? bhi[j][2] : kFarBox;
		w.child[j] = use ? ref[j] : leaf_ref(0, 1);
		w.pad[j] = 0;
	}
	cells[c] = w;
}

// Sizes the entry grid over the scene box grown by kGridReach: cubic cells
// of at least kGridMinCell, at most kGridCells of them; false when the box
// is not finite.
static bool grid_spec(const float lo[3], const float hi[3], GridSpec& G)
{
	double e[3], vol = 1.0, mx = 0.0;
	for (int a = 0; a < 3; ++a) {
		if (!std::isfinite(lo[a]) || !std::isfinite(hi[a]) || !(hi[a] >= lo[a]))
			return false;
		G.lo[a] = (float)((double)lo[a] - kGridReach);
		e[a] = (double)hi[a] - lo[a] + 2.0 * kGridReach;
		vol *= e[a];
		mx = std::max(mx, std::max(fabs((double)lo[a]), fabs((double)hi[a])));
	}
	double g = std::max((double)kGridMinCell, std::cbrt(vol / kGridCells));
	for (int it = 0; it < 64; ++it) {
		double cells = 1.0;
		for (int a = 0; a < 3; ++a)
			cells *= std::max(1.0, std::ceil(e[a] / g));
		if (cells <= kGridCells)
			break;
		g *= 1.05;
	}
	for (int a = 0; a < 3; ++a)
		G.n[a] = (int)std::max(1.0, std::ceil(e[a] / g));
	if ((double)G.n[0] * G.n[1] * G.n[2] > kGridCells)
		return false;
	G.g = (float)g;
	// rounding of the cell index and of the region's bounds: a query point
	// lies within 1e-3 of a cell plus 1e-5 of the coordinates' magnitude of
	// its cell; distances within the same relative bound of the radius
	G.pad = kGridReach * (1.0f + 1e-4f) + 1e-3f * G.g + 1e-5f * (float)mx + 1e-3f;
	return true;
}

void DeviceScene::set_triangles(const float* tris, size_t ntri)
{
	dfree(d_nodes_);
	dfree(d_dtri_);
	dfree(d_rtri_);
	ntri_ = ntri;
	times.bvh_ms = 0;
	if (ntri == 0) {
		nnodes_ = 0;
		return;
	}
	if (ntri >= (1u << 27))
		throw std::runtime_error("too many triangles for the leaf encoding");
	hipStream_t s = (hipStream_t)stream_;
	// triangles up, the tree built on the device (the host build's tree,
	// bvh_gpu.hip), the triangle forms made there in leaf order
	DevArray<float> d_tris(9 * ntri);
	DevArray<uint32_t> d_order(ntri);
	HIPCHECK(hipMemcpyAsync(d_tris.p, tris, 36 * ntri, hipMemcpyHostToDevice, s));
	double kms = 0;
	BvhBuild b = build_bvh_gpu(d_tris.p, ntri, d_order.p, stream_, &kms);
	times.bvh_ms = kms;
	if (3 * b.depth + 1 > kStackSize)
		throw std::runtime_error("BVH deeper than the traversal stack");
	bvh_depth_ = b.depth;
	persist_blocks_ = 0; // the march kernel's LDS stack follows the depth
	nnodes_ = b.nodes.size();
	lo_[0] = b.lo.x; lo_[1] = b.lo.y; lo_[2] = b.lo.z;
	hi_[0] = b.hi.x; hi_[1] = b.hi.y; hi_[2] = b.hi.z;
	GridSpec G{};
	const bool grid = grid_spec(lo_, hi_, G) && nnodes_ + (size_t)G.n[0] * G.n[1] * G.n[2] < (size_t)INT32_MAX;
	const size_t ncells = grid ? (size_t)G.n[0] * G.n[1] * G.n[2] : 0;
	const bool fine = grid && nnodes_ + 2 * ncells < (size_t)INT32_MAX;
	d_nodes_ = dalloc<WideNode>(nnodes_ + (fine ? 2 : 1) * ncells);
	d_dtri_ = dalloc<DistTri>(ntri);
	d_rtri_ = dalloc<RayTri>(ntri);
	HIPCHECK(hipMemcpyAsync(d_nodes_, b.nodes.data(), sizeof(WideNode) * nnodes_, hipMemcpyHostToDevice, s));
	k_tri_forms<<<grid_for(ntri), kBlock, 0, s>>>(d_tris.p, d_order.p, ntri, (DistTri*)d_dtri_, (RayTri*)d_rtri_);
	HIPCHECK(hipGetLastError());
	grid_ = G;
	grid_base_ = 0;
	if (grid) {
		k_entry_cells<<<(unsigned)((ncells + 255) / 256), 256, 0, s>>>((const WideNode*)d_nodes_, G,
		                                                               (WideNode*)d_nodes_ + nnodes_);
		HIPCHECK(hipGetLastError());
		grid_base_ = (int32_t)nnodes_;
	}
	grid_fine_ = 0;
	if (fine) {
		GridSpec F = G;
		F.pad = G.pad - kGridReach * (1.0f + 1e-4f) + kGridFineReach * (1.0f + 1e-4f);
		k_entry_cells<<<(unsigned)((ncells + 255) / 256), 256, 0, s>>>((const WideNode*)d_nodes_, F,
		                                                               (WideNode*)d_nodes_ + nnodes_ + ncells);
		HIPCHECK(hipGetLastError());
		grid_fine_ = (int32_t)(nnodes_ + ncells);
	}
	HIPCHECK(hipStreamSynchronize(s));
}

void DeviceScene::set_samples(const float* pos, const float* nrm, size_t n)
{
	if (n > kSampleMask)
		throw std::runtime_error("too many samples for the work-item encoding");
	dfree(d_pos_);
	dfree(d_nrm_);
	dfree(d_lm_);
	dfree(d_skey_);
	nsamples_ = n;
	{
		// 30-bit Morton code of each texel position over the samples' box
		std::vector<uint32_t> key(n);
		float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
		for (size_t i = 0; i < n; ++i)
			for (int a = 0; a < 3; ++a) {
				mn[a] = fminf(mn[a], pos[3 * i + a]);
				mx[a] = fmaxf(mx[a], pos[3 * i + a]);
			}
		const float ext = fmaxf(fmaxf(mx[0] - mn[0], mx[1] - mn[1]), mx[2] - mn[2]);
		const float q = ext > 0.0f ? 1023.0f / ext : 0.0f;
		auto spread = [](uint32_t v) {
			v &= 0x3ffu;
			v = (v | (v << 16)) & 0x030000ffu;
			v = (v | (v << 8)) & 0x0300f00fu;
			v = (v | (v << 4)) & 0x030c30c3u;
			v = (v | (v << 2)) & 0x09249249u;
			return v;
		};
		for (size_t i = 0; i < n; ++i) {
			uint32_t c = 0;
			for (int a = 0; a < 3; ++a) {
				float f = (pos[3 * i + a] - mn[a]) * q;
				c |= spread(f > 0.0f ? (uint32_t)fminf(f, 1023.0f) : 0u) << a;
			}
			key[i] = c;
		}
		d_skey_ = dalloc<uint32_t>(n);
		if (n)
			HIPCHECK(hipMemcpy(d_skey_, key.data(), 4 * n, hipMemcpyHostToDevice));
	}
	d_pos_ = dalloc<float>(3 * n);
	d_nrm_ = dalloc<float>(3 * n);
	d_lm_ = dalloc<float>(3 * n);
	hipStream_t s = (hipStream_t)stream_;
	if (n) {
		HIPCHECK(hipMemcpyAsync(d_pos_, pos, 12 * n, hipMemcpyHostToDevice, s));
		HIPCHECK(hipMemcpyAsync(d_nrm_, nrm, 12 * n, hipMemcpyHostToDevice, s));
		HIPCHECK(hipMemsetAsync(d_lm_, 0, 12 * n, s));
	}
	HIPCHECK(hipStreamSynchronize(s));
}

void DeviceScene::set_lightmap(const float* rgb)
{
	if (nsamples_)
		HIPCHECK(hipMemcpy(d_lm_, rgb, 12 * nsamples_, hipMemcpyHostToDevice));
}

void DeviceScene::get_lightmap(float* rgb) const
{
	if (nsamples_)
		HIPCHECK(hipMemcpy(rgb, d_lm_, 12 * nsamples_, hipMemcpyDeviceToHost));
}

struct EventTimer {
	hipEvent_t a, b;
	hipStream_t s;
	explicit EventTimer(hipStream_t st) : s(st)
	{
		HIPCHECK(hipEventCreate(&a));
		HIPCHECK(hipEventCreate(&b));
		HIPCHECK(hipEventRecord(a, s));
	}
	double stop()
	{
		HIPCHECK(hipEventRecord(b, s));
		HIPCHECK(hipEventSynchronize(b));
		float ms = 0;
		HIPCHECK(hipEventElapsedTime(&ms, a, b));
		return ms;
	}
	~EventTimer()
	{
		(void)hipEventDestroy(a);
		(void)hipEventDestroy(b);
	}
};

void DeviceScene::set_direct_work(const std::vector<DevLight>& lights, const std::vector<uint32_t>& items,
                                  const std::vector<uint32_t>& light_off, bool keep_weights)
{
	hipStream_t s = (hipStream_t)stream_;
	if (lights.size() > 64)
		throw std::runtime_error("direct_light: at most 64 lights per batch");
	if (light_off.size() != lights.size() + 1 || light_off.back() != items.size() || light_off[0] != 0)
		throw std::runtime_error("direct_light: light offsets do not match the work items");
	// every item must name its segment's light and a sample on the device
	for (size_t l = 0; l < lights.size(); ++l) {
		if (light_off[l + 1] < light_off[l])
			throw std::runtime_error("direct_light: light offsets must not decrease");
		for (uint32_t k = light_off[l]; k < light_off[l + 1]; ++k)
			if ((items[k] >> kSampleBits) != l || (items[k] & kSampleMask) >= nsamples_)
				throw std::runtime_error("direct_light: work item " + std::to_string(k) +
				                         " names another light or a sample beyond the device's samples");
	}
	dfree(d_w_);
	dfree(d_vis_);
	dfree(d_items_);
	dfree(d_lights_);
	dfree(d_gate_);
	dfree(d_mlist_);
	nitems_ = items.size();
	d_w_ = dalloc<float>(nitems_);
	if (keep_weights)
		d_vis_ = dalloc<float>(nitems_);
	d_gate_ = dalloc<float>(nitems_);
	d_mlist_ = dalloc<uint32_t>(nitems_);
	d_items_ = dalloc<uint32_t>(nitems_);
	d_lights_ = dalloc<DevLight>(lights.size());
	HIPCHECK(hipMemcpyAsync(d_items_, items.data(), 4 * nitems_, hipMemcpyHostToDevice, s));
	HIPCHECK(hipMemcpyAsync(d_lights_, lights.data(), sizeof(DevLight) * lights.size(), hipMemcpyHostToDevice, s));
	HIPCHECK(hipStreamSynchronize(s));
	h_lights_ = lights;
	h_light_off_ = light_off;
}

// While the march kernel runs: progress to keep_going (the fraction of the
// march list claimed) every millisecond, and the abort flag raised when it
// returns false.  True when the march was aborted (also when the device's
// abort request was set before the launch: the kernel then leaves at once).
bool DeviceScene::wait_march(uint32_t nmarch, const std::function<bool(float)>& keep_going)
{
	if (abort_requested_)
		return true;
	if (!keep_going)
		return false;
	hipStream_t s = (hipStream_t)stream_;
	volatile uint32_t* ctl = ctl_host_;
	bool stop = false;
	for (;;) {
		const hipError_t e = hipStreamQuery(s);
		if (e == hipSuccess)
			break;
		if (e != hipErrorNotReady)
			HIPCHECK(e);
		if (!stop) {
			const float f = std::min(1.0f, (float)ctl[1] / (float)nmarch);
			if (!keep_going(f)) {
				stop = true;
				ctl[0] = 1u;
			}
		}
		std::this_thread::sleep_for(std::chrono::milliseconds(1));
	}
	return stop;
}

bool DeviceScene::run_direct(const std::function<bool(float)>& keep_going)
{
	hipStream_t s = (hipStream_t)stream_;
	View V = make_view(d_nodes_, d_dtri_, d_rtri_, ntri_, lo_, hi_, grid_, use_grid_ ? grid_base_ : 0, use_grid_ ? grid_fine_ : 0);
	bool aborted = false;
	ctl_host_[0] = 0u; // read by the kernel after the launch below
	ctl_host_[1] = 0u;
	HIPCHECK(hipMemsetAsync(d_counters_, 0, kNumCounters * sizeof(unsigned long long), s));
	EventTimer all(s);
	if (nitems_) {
		HIPCHECK(hipMemsetAsync(d_mcount_, 0, 2 * sizeof(uint32_t), s));
		// the device abort flag starts raised under request_abort: every wave
		// then leaves at its first claim, before any march
		HIPCHECK(hipMemsetAsync(d_mcount_ + 2, abort_requested_ ? 1 : 0, sizeof(uint32_t), s));
		k_direct_gate<<<(unsigned)((nitems_ + kGateRounds * kBlock - 1) / (kGateRounds * kBlock)), kBlock, 0, s>>>(
		    d_pos_, d_nrm_, d_lights_, d_items_, (uint32_t)nitems_, d_gate_, d_w_, d_vis_, d_mlist_, d_mcount_, d_skey_);
		HIPCHECK(hipGetLastError());
		uint32_t nmarch = 0;
		HIPCHECK(hipMemcpyAsync(&nmarch, d_mcount_, 4, hipMemcpyDeviceToHost, s));
		HIPCHECK(hipStreamSynchronize(s));
		nmarch_ = nmarch;
		EventTimer tm(s);
		if (nmarch) {
			unsigned blocks = persistent_blocks();
			const size_t lds = march_lds_bytes();
			// the abort and progress path only for runs that can be stopped
			// (it costs up to 2 % of the march on the city block)
			uint32_t* ctl = (keep_going || abort_requested_) ? ctl_dev_ : nullptr;
			if (count_queries)
				k_march<true><<<blocks, kBlock, lds, s>>>(V, d_pos_, d_nrm_, d_lights_, d_items_, d_mlist_, nmarch,
				                                          d_gate_, d_w_, d_vis_, d_mcount_ + 1, d_counters_,
				                                          march_stack_depth(), ctl);
			else
				k_march<false><<<blocks, kBlock, lds, s>>>(V, d_pos_, d_nrm_, d_lights_, d_items_, d_mlist_, nmarch,
				                                           d_gate_, d_w_, d_vis_, d_mcount_ + 1, d_counters_,
				                                           march_stack_depth(), ctl);
			HIPCHECK(hipGetLastError());
			aborted = wait_march(nmarch, keep_going);
		}
		times.direct_march_ms = tm.stop();
	}
	if (aborted) { // the lightmap keeps its colours
		times.direct_ms = all.stop();
		read_counters();
		return false;
	}
	for (size_t l = 0; l < h_lights_.size(); ++l) {
		uint32_t b = h_light_off_[l], e = h_light_off_[l + 1];
		if (e > b) {
			k_accumulate<<<grid_for(e - b), kBlock, 0, s>>>(d_items_, d_w_, b, e, h_lights_[l].color[0],
			                                                h_lights_[l].color[1], h_lights_[l].color[2], d_lm_);
			HIPCHECK(hipGetLastError());
		}
	}
	times.direct_ms = all.stop();
	read_counters();
	return true;
}

// [p, p + bytes) must lie inside one device allocation of the current GPU.
static void check_device_range(const void* p, size_t bytes, const char* what)
{
	hipDeviceptr_t base = nullptr;
	size_t size = 0;
	const hipError_t e = hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p);
	if (e != hipSuccess)
		(void)hipGetLastError(); // the failed lookup must not surface as a later launch's error
	if (e != hipSuccess || !base || (const char*)p + bytes > (const char*)base + size)
		throw std::runtime_error(std::string("direct_light_noalloc: ") + what +
		                         " is not a device buffer of the current GPU large enough for the work items");
}

void DeviceScene::run_direct_noalloc(const DevLight* lights, int nlights, const uint32_t* d_items, size_t nitems,
                                     const uint32_t* light_off, float* d_w, float* d_vis)
{
	if (nitems > kNoAllocMaxItems)
		throw std::runtime_error("direct_light_noalloc: " + std::to_string(nitems) +
		                         " work items exceed the scratch's maximum of " + std::to_string(kNoAllocMaxItems));
	if (nlights < 0 || nlights > 64)
		throw std::runtime_error("direct_light_noalloc: 0..64 lights per call");
	if (light_off[0] != 0 || light_off[nlights] != nitems)
		throw std::runtime_error("direct_light_noalloc: light offsets do not match the work items");
	for (int l = 0; l < nlights; ++l)
		if (light_off[l + 1] < light_off[l])
			throw std::runtime_error("direct_light_noalloc: light offsets must not decrease");
	if (nitems == 0)
		return;
	if (!nsamples_)
		throw std::runtime_error("direct_light_noalloc: the device holds no samples");
	check_device_range(d_items, 4 * nitems, "the work-item array");
	check_device_range(d_w, 4 * nitems, "the weight array");
	if (d_vis)
		check_device_range(d_vis, 4 * nitems, "the shadow-factor array");
	hipStream_t s = (hipStream_t)stream_;
	char* sc = (char*)d_scratch_;
	float* gate = (float*)(sc + kNoAllocGateOff);
	uint32_t* mlist = (uint32_t*)(sc + kNoAllocListOff);
	uint32_t* head = (uint32_t*)(sc + kNoAllocHeadOff); // [0] march count, [1] list cursor, [2] bad item
	unsigned long long* cnt = (unsigned long long*)(sc + kNoAllocCntOff);
	DevLight* dl = (DevLight*)(sc + kNoAllocLightOff);
	uint32_t* doff = (uint32_t*)(sc + kNoAllocOffOff);
	View V = make_view(d_nodes_, d_dtri_, d_rtri_, ntri_, lo_, hi_, grid_, use_grid_ ? grid_base_ : 0, use_grid_ ? grid_fine_ : 0);
	if (nlights)
		HIPCHECK(hipMemcpyAsync(dl, lights, sizeof(DevLight) * nlights, hipMemcpyHostToDevice, s));
	HIPCHECK(hipMemcpyAsync(doff, light_off, 4 * (size_t)(nlights + 1), hipMemcpyHostToDevice, s));
	// counters and flags start at zero (the array keeps the last call's contents)
	k_fill_u32<<<1, 64, 0, s>>>(head, 4, 0u);
	k_fill_u32<<<1, 64, 0, s>>>((uint32_t*)cnt, 2 * kNumCounters, 0u);
	k_check_items<<<grid_for(nitems), kBlock, 0, s>>>(d_items, (uint32_t)nitems, doff, (uint32_t)nlights,
	                                                  (uint32_t)nsamples_, head + 2);
	HIPCHECK(hipGetLastError());
	uint32_t h[3] = {0, 0, 0};
	HIPCHECK(hipMemcpyAsync(h, head, sizeof(h), hipMemcpyDeviceToHost, s));
	HIPCHECK(hipStreamSynchronize(s));
	if (h[2])
		throw std::runtime_error("direct_light_noalloc: a work item names another light or a sample beyond the "
		                         "device's samples");
	k_direct_gate<<<(unsigned)((nitems + kGateRounds * kBlock - 1) / (kGateRounds * kBlock)), kBlock, 0, s>>>(
	    d_pos_, d_nrm_, dl, d_items, (uint32_t)nitems, gate, d_w, d_vis, mlist, head, d_skey_);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipMemcpyAsync(h, head, 4, hipMemcpyDeviceToHost, s));
	HIPCHECK(hipStreamSynchronize(s));
	const uint32_t nmarch = h[0];
	if (nmarch) {
		k_march<false><<<persistent_blocks(), kBlock, march_lds_bytes(), s>>>(
		    V, d_pos_, d_nrm_, dl, d_items, mlist, nmarch, gate, d_w, d_vis, head + 1, cnt, march_stack_depth(),
		    nullptr);
		HIPCHECK(hipGetLastError());
	}
	for (int l = 0; l < nlights; ++l) {
		const uint32_t b = light_off[l], e = light_off[l + 1];
		if (e > b) {
			k_accumulate<<<grid_for(e - b), kBlock, 0, s>>>(d_items, d_w, b, e, lights[l].color[0], lights[l].color[1],
			                                                lights[l].color[2], d_lm_);
			HIPCHECK(hipGetLastError());
		}
	}
	HIPCHECK(hipStreamSynchronize(s));
}

void DeviceScene::reference_counts(const float* tris, size_t ntri, uint64_t out[5])
{
	for (int i = 0; i < 5; ++i)
		out[i] = 0;
	if (ntri != ntri_)
		throw std::runtime_error("reference_counts: the triangles must be those of the device scene");
	if (!d_vis_)
		throw std::runtime_error("reference_counts: run the direct light with kept weights first");
	if (nmarch_ == 0 || ntri == 0)
		return;
	BinaryBvh b = build_bvh_binary(tris, ntri, 16, 4);
	if (b.depth > kCanonStack)
		throw std::runtime_error("reference_counts: binary tree deeper than the counting stack");
	std::vector<DistTri> dt(ntri);
	for (size_t i = 0; i < ntri; ++i) {
		const float* p = tris + 9 * (size_t)b.order[i];
		dt[i] = make_dist_tri(ld3(p), ld3(p + 3), ld3(p + 6));
	}
	BvhNode* d_nodes = dalloc<BvhNode>(b.nodes.size());
	DistTri* d_dt = dalloc<DistTri>(ntri);
	unsigned long long* d_out = dalloc<unsigned long long>(4);
	struct Free {
		BvhNode*& a;
		DistTri*& b;
		unsigned long long*& c;
		~Free()
		{
			dfree(a);
			dfree(b);
			dfree(c);
		}
	} fr{d_nodes, d_dt, d_out};
	hipStream_t s = (hipStream_t)stream_;
	HIPCHECK(hipMemcpyAsync(d_nodes, b.nodes.data(), b.nodes.size() * sizeof(BvhNode), hipMemcpyHostToDevice, s));
	HIPCHECK(hipMemcpyAsync(d_dt, dt.data(), ntri * sizeof(DistTri), hipMemcpyHostToDevice, s));
	HIPCHECK(hipMemsetAsync(d_out, 0, 4 * sizeof(unsigned long long), s));
	CanonView C;
	C.nodes = d_nodes;
	C.dtri = d_dt;
	canon_root_box(b, C.lo, C.hi);
	View V = make_view(d_nodes_, d_dtri_, d_rtri_, ntri_, lo_, hi_, grid_, use_grid_ ? grid_base_ : 0, use_grid_ ? grid_fine_ : 0);
	k_march_canon<<<grid_for(nmarch_), kBlock, 0, s>>>(C, V, d_pos_, d_nrm_, d_lights_, d_items_, d_mlist_, nmarch_,
	                                                   d_vis_, d_out);
	HIPCHECK(hipGetLastError());
	unsigned long long h[4];
	HIPCHECK(hipMemcpyAsync(h, d_out, sizeof(h), hipMemcpyDeviceToHost, s));
	HIPCHECK(hipStreamSynchronize(s));
	for (int i = 0; i < 4; ++i)
		out[i] = h[i];
	out[4] = nmarch_;
}

void DeviceScene::reference_counts_ao(const float* tris, size_t ntri, const float* randoff, int num_samples,
                                      float distance, float multiplier, float falloff, uint64_t out[5])
{
	for (int i = 0; i < 5; ++i)
		out[i] = 0;
	if (ntri != ntri_)
		throw std::runtime_error("reference_counts_ao: the triangles must be those of the device scene");
	if (nsamples_ == 0 || num_samples <= 0 || ntri == 0)
		return;
	BinaryBvh b = build_bvh_binary(tris, ntri, 16, 4);
	if (b.depth > kCanonStack)
		throw std::runtime_error("reference_counts_ao: binary tree deeper than the counting stack");
	std::vector<RayTri> rt(ntri);
	for (size_t i = 0; i < ntri; ++i) {
		const float* p = tris + 9 * (size_t)b.order[i];
		rt[i] = make_ray_tri(ld3(p), ld3(p + 3), ld3(p + 6));
	}
	std::vector<float2> side(num_samples);
	for (int i = 0; i < num_samples; ++i) {
		float q = (i + 0.5f) / num_samples;
		float cs = sqrtf(q);
		side[i] = make_float2(cs, sinf(acosf(cs)));
	}
	hipStream_t s = (hipStream_t)stream_;
	BvhNode* d_nodes = dalloc<BvhNode>(b.nodes.size());
	RayTri* d_rt = dalloc<RayTri>(ntri);
	float2* d_side = dalloc<float2>(num_samples);
	float* d_ro = dalloc<float>(nsamples_);
	float* d_fp = dalloc<float>(nsamples_);
	float* d_fc = dalloc<float>(nsamples_);
	float* d_lm_keep = dalloc<float>(3 * nsamples_);
	unsigned long long* d_out = dalloc<unsigned long long>(4);
	struct Free {
		std::vector<void*> p;
		~Free()
		{
			for (void* q : p)
				if (q)
					(void)hipFree(q);
		}
	} fr{{d_nodes, d_rt, d_side, d_ro, d_fp, d_fc, d_lm_keep, d_out}};
	HIPCHECK(hipMemcpyAsync(d_nodes, b.nodes.data(), b.nodes.size() * sizeof(BvhNode), hipMemcpyHostToDevice, s));
	HIPCHECK(hipMemcpyAsync(d_rt, rt.data(), ntri * sizeof(RayTri), hipMemcpyHostToDevice, s));
	HIPCHECK(hipMemcpyAsync(d_side, side.data(), sizeof(float2) * num_samples, hipMemcpyHostToDevice, s));
	HIPCHECK(hipMemcpyAsync(d_ro, randoff, 4 * nsamples_, hipMemcpyHostToDevice, s));
	HIPCHECK(hipMemsetAsync(d_out, 0, 4 * sizeof(unsigned long long), s));
	// the product's factors, with the lightmap put back afterwards
	HIPCHECK(hipMemcpyAsync(d_lm_keep, d_lm_, 12 * nsamples_, hipMemcpyDeviceToDevice, s));
	const float black[3] = {0, 0, 0};
	const bool counting = count_queries;
	count_queries = false;
	const StageTimes keep_times = times;
	ambient_occlusion(randoff, num_samples, distance, multiplier, falloff, 0.0f, black, d_fp);
	count_queries = counting;
	times = keep_times;
	HIPCHECK(hipMemcpyAsync(d_lm_, d_lm_keep, 12 * nsamples_, hipMemcpyDeviceToDevice, s));
	AOParams P;
	P.n = num_samples;
	P.distance = distance;
	P.multiplier = multiplier;
	P.falloff = falloff;
	P.effect = 0.0f;
	P.color[0] = P.color[1] = P.color[2] = 0.0f;
	CanonView C;
	C.nodes = d_nodes;
	C.dtri = nullptr;
	canon_root_box(b, C.lo, C.hi);
	const unsigned grid = (unsigned)((nsamples_ + kBlock / 64 - 1) / (kBlock / 64));
	k_ao_canon<<<grid, kBlock, 0, s>>>(C, d_rt, d_pos_, d_nrm_, d_ro, d_side, P, (uint32_t)nsamples_, d_fc, d_out);
	HIPCHECK(hipGetLastError());
	unsigned long long h[4];
	HIPCHECK(hipMemcpyAsync(h, d_out, sizeof(h), hipMemcpyDeviceToHost, s));
	std::vector<float> fp(nsamples_), fc(nsamples_);
	HIPCHECK(hipMemcpyAsync(fp.data(), d_fp, 4 * nsamples_, hipMemcpyDeviceToHost, s));
	HIPCHECK(hipMemcpyAsync(fc.data(), d_fc, 4 * nsamples_, hipMemcpyDeviceToHost, s));
	HIPCHECK(hipStreamSynchronize(s));
	uint64_t differ = 0;
	for (size_t i = 0; i < nsamples_; ++i)
		differ += memcmp(&fp[i], &fc[i], 4) != 0;
	out[0] = h[0];
	out[1] = h[1];
	out[2] = h[2];
	out[3] = differ;
	out[4] = nsamples_;
}

void DeviceScene::reference_counts_links(const float* tris, size_t ntri, uint64_t out[5])
{
	for (int i = 0; i < 5; ++i)
		out[i] = 0;
	if (ntri != ntri_)
		throw std::runtime_error("reference_counts_links: the triangles must be those of the device scene");
	if (!d_cand_)
		throw std::runtime_error("reference_counts_links: run the link search with counting on first");
	if (ncand_ == 0 || ntri == 0)
		return;
	BinaryBvh b = build_bvh_binary(tris, ntri, 16, 4);
	if (b.depth > kCanonStack)
		throw std::runtime_error("reference_counts_links: binary tree deeper than the counting stack");
	std::vector<RayTri> rt(ntri);
	for (size_t i = 0; i < ntri; ++i) {
		const float* p = tris + 9 * (size_t)b.order[i];
		rt[i] = make_ray_tri(ld3(p), ld3(p + 3), ld3(p + 6));
	}
	hipStream_t s = (hipStream_t)stream_;
	BvhNode* d_nodes = dalloc<BvhNode>(b.nodes.size());
	RayTri* d_rt = dalloc<RayTri>(ntri);
	unsigned long long* d_out = dalloc<unsigned long long>(3);
	struct Free {
		std::vector<void*> p;
		~Free()
		{
			for (void* q : p)
				if (q)
					(void)hipFree(q);
		}
	} fr{{d_nodes, d_rt, d_out}};
	HIPCHECK(hipMemcpyAsync(d_nodes, b.nodes.data(), b.nodes.size() * sizeof(BvhNode), hipMemcpyHostToDevice, s));
	HIPCHECK(hipMemcpyAsync(d_rt, rt.data(), ntri * sizeof(RayTri), hipMemcpyHostToDevice, s));
	HIPCHECK(hipMemsetAsync(d_out, 0, 3 * sizeof(unsigned long long), s));
	CanonView C;
	C.nodes = d_nodes;
	C.dtri = nullptr;
	canon_root_box(b, C.lo, C.hi);
	k_link_canon<<<grid_for(ncand_), kBlock, 0, s>>>(C, d_rt, rad.pos, (const uint2*)d_cand_, ncand_,
	                                                 (const uint8_t*)d_clinked_, d_out);
	HIPCHECK(hipGetLastError());
	unsigned long long h[3];
	HIPCHECK(hipMemcpyAsync(h, d_out, sizeof(h), hipMemcpyDeviceToHost, s));
	HIPCHECK(hipStreamSynchronize(s));
	out[0] = h[0];
	out[1] = h[1];
	out[2] = ncand_;
	out[3] = h[2];
	out[4] = rad.nnz / 2;
}

void DeviceScene::read_counters()
{
	if (!count_queries)
		return;
	unsigned long long c[kNumCounters];
	HIPCHECK(hipMemcpy(c, d_counters_, sizeof(c), hipMemcpyDeviceToHost));
	counters.nodes = c[0];
	counters.tris = c[1];
	counters.steps = c[2];
	counters.queries = c[3];
	counters.marches = c[4];
	counters.pretests = c[5];
}

bool DeviceScene::direct_light(const std::vector<DevLight>& lights, const std::vector<uint32_t>& items,
                               const std::vector<uint32_t>& light_off, bool keep_weights,
                               const std::function<bool(float)>& keep_going)
{
	set_direct_work(lights, items, light_off, keep_weights);
	return run_direct(keep_going);
}

// ---------------------------------------------------------------------------
// Normal-map data on the GPU (lighter.cpp:505-513, 959-1030): one lane per
// sample walks the kept lights in order twice, as the host restatement
// walks its contribution list: first the sum of the scaled sample normal
// and the contributions and their count, then, against the mean, the
// min-dot product (normals.h serves both sides).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_nm_data(const float* __restrict__ pos, const float* __restrict__ nrm,
                                                    uint32_t n, const uint32_t* __restrict__ inst_off, int ninst,
                                                    const DevLight* __restrict__ lights, int nl,
                                                    const int32_t* __restrict__ first,
                                                    const float* const* __restrict__ visp, float ab,
                                                    float* __restrict__ out)
{
	const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
	if (s >= n)
		return;
	int lo = 0, hi = ninst - 1; // the sample's instance: last with inst_off <= s
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (inst_off[mid] <= s)
			lo = mid;
		else
			hi = mid - 1;
	}
	const int inst = lo;
	const uint32_t within = s - inst_off[inst];
	const f3 SP = ld3(pos + 3 * (size_t)s), SN = ld3(nrm + 3 * (size_t)s);
	f3 sum = SN * ab;
	int count = 1;
	for (int l = 0; l < nl; ++l) {
		const int32_t f = first[(size_t)l * ninst + inst];
		f3 c;
		if (f >= 0 && nm_contrib(lights[l], SP, SN, visp[l][(uint32_t)f + within], &c)) {
			sum += c;
			++count;
		}
	}
	const f3 mean = sum / (float)count;
	float md = 1.0f;
	for (int l = 0; l < nl; ++l) {
		const int32_t f = first[(size_t)l * ninst + inst];
		f3 c;
		if (f >= 0 && nm_contrib(lights[l], SP, SN, visp[l][(uint32_t)f + within], &c))
			md *= nm_mindot_factor(mean, c);
	}
	const f3 N = normalized(mean);
	out[4 * (size_t)s + 0] = N.x;
	out[4 * (size_t)s + 1] = N.y;
	out[4 * (size_t)s + 2] = N.z;
	out[4 * (size_t)s + 3] = md;
}

void DeviceScene::nm_clear()
{
	for (float*& p : nm_vis_)
		dfree(p);
	nm_vis_.clear();
	nm_lights_.clear();
	nm_first_.clear();
	nm_light_batch_.clear();
	nm_batch_items_.clear();
	nm_ninst_ = 0;
}

void DeviceScene::nm_keep_batch(const std::vector<DevLight>& lights, const std::vector<int32_t>& first_item,
                                size_t ninst)
{
	if (!d_vis_)
		throw std::runtime_error("nm_keep_batch: the direct-light run kept no shadow factors");
	if (first_item.size() != lights.size() * ninst || (nm_ninst_ && nm_ninst_ != ninst))
		throw std::runtime_error("nm_keep_batch: item table does not match the lights and instances");
	for (int32_t f : first_item)
		if (f >= 0 && (size_t)f >= nitems_)
			throw std::runtime_error("nm_keep_batch: item offset beyond the run's items");
	nm_ninst_ = ninst;
	float* v = dalloc<float>(nitems_);
	hipStream_t s = (hipStream_t)stream_;
	HIPCHECK(hipMemcpyAsync(v, d_vis_, 4 * std::max<size_t>(nitems_, 1), hipMemcpyDeviceToDevice, s));
	HIPCHECK(hipStreamSynchronize(s));
	nm_vis_.push_back(v);
	nm_batch_items_.push_back(nitems_);
	for (size_t l = 0; l < lights.size(); ++l) {
		nm_lights_.push_back(lights[l]);
		nm_light_batch_.push_back((uint32_t)(nm_vis_.size() - 1));
	}
	nm_first_.insert(nm_first_.end(), first_item.begin(), first_item.end());
}

void DeviceScene::nm_compute(const std::vector<uint32_t>& inst_off, float ambient_bright, float* out4)
{
	const size_t ni = inst_off.size() - 1;
	if (inst_off.size() < 2 || ni != nm_ninst_ || inst_off.back() != nsamples_)
		throw std::runtime_error("nm_compute: instance offsets do not match the kept batches or the samples");
	// every kept item must lie inside its batch: first + instance size <= items
	const int nl = (int)nm_lights_.size();
	for (int l = 0; l < nl; ++l)
		for (size_t i = 0; i < ni; ++i) {
			const int32_t f = nm_first_[(size_t)l * ni + i];
			if (f >= 0 && (size_t)f + (inst_off[i + 1] - inst_off[i]) > nm_batch_items_[nm_light_batch_[l]])
				throw std::runtime_error("nm_compute: an instance's items run past its batch");
		}
	hipStream_t s = (hipStream_t)stream_;
	std::vector<const float*> visp(nl);
	for (int l = 0; l < nl; ++l)
		visp[l] = nm_vis_[nm_light_batch_[l]];
	DevArray<uint32_t> d_off(inst_off.size());
	DevArray<DevLight> d_l(std::max(nl, 1));
	DevArray<int32_t> d_first(std::max<size_t>(nm_first_.size(), 1));
	DevArray<const float*> d_visp(std::max(nl, 1));
	DevArray<float> d_out(4 * std::max<size_t>(nsamples_, 1));
	HIPCHECK(hipMemcpyAsync(d_off.p, inst_off.data(), 4 * inst_off.size(), hipMemcpyHostToDevice, s));
	if (nl) {
		HIPCHECK(hipMemcpyAsync(d_l.p, nm_lights_.data(), sizeof(DevLight) * nl, hipMemcpyHostToDevice, s));
		HIPCHECK(hipMemcpyAsync(d_first.p, nm_first_.data(), 4 * nm_first_.size(), hipMemcpyHostToDevice, s));
		HIPCHECK(hipMemcpyAsync(d_visp.p, visp.data(), sizeof(float*) * nl, hipMemcpyHostToDevice, s));
	}
	if (nsamples_) {
		k_nm_data<<<grid_for(nsamples_), kBlock, 0, s>>>(d_pos_, d_nrm_, (uint32_t)nsamples_, d_off.p, (int)ni, d_l.p,
		                                                 nl, d_first.p, d_visp.p, ambient_bright, d_out.p);
		HIPCHECK(hipGetLastError());
		HIPCHECK(hipMemcpyAsync(out4, d_out.p, 16 * nsamples_, hipMemcpyDeviceToHost, s));
	}
	HIPCHECK(hipStreamSynchronize(s));
}

void DeviceScene::direct_grad(const double* lam3, double* dcolor3)
{
	hipStream_t s = (hipStream_t)stream_;
	const size_t L = h_lights_.size();
	if (!d_w_ && nitems_)
		throw std::runtime_error("direct_grad: no direct-light work on the device");
	double* dlam = dalloc<double>(3 * std::max<size_t>(nsamples_, 1));
	double* dpart = dalloc<double>(3 * kGradBlocks * std::max<size_t>(L, 1));
	HIPCHECK(hipMemcpyAsync(dlam, lam3, 24 * nsamples_, hipMemcpyHostToDevice, s));
	HIPCHECK(hipMemsetAsync(dpart, 0, 24 * kGradBlocks * std::max<size_t>(L, 1), s));
	uint32_t* doff = dalloc<uint32_t>(L + 1);
	HIPCHECK(hipMemcpyAsync(doff, h_light_off_.data(), 4 * (L + 1), hipMemcpyHostToDevice, s));
	EventTimer tm(s);
	if (L) {
		k_color_grad<<<dim3(kGradBlocks, (unsigned)L), kBlock, 0, s>>>(d_items_, d_w_, doff, dlam, dpart);
		HIPCHECK(hipGetLastError());
	}
	times.grad_ms = tm.stop();
	std::vector<double> hp(3 * kGradBlocks * L);
	if (L)
		HIPCHECK(hipMemcpy(hp.data(), dpart, 8 * hp.size(), hipMemcpyDeviceToHost));
	dfree(dlam);
	dfree(dpart);
	dfree(doff);
	for (size_t l = 0; l < L; ++l)
		for (int c = 0; c < 3; ++c) {
			double t = 0;
			for (int b = 0; b < kGradBlocks; ++b)
				t += hp[3 * (kGradBlocks * l + b) + c];
			dcolor3[3 * l + c] = t;
		}
}

void DeviceScene::get_weights(float* out, size_t n) const
{
	if (n > nitems_)
		throw std::runtime_error("get_weights: more than available");
	if (n)
		HIPCHECK(hipMemcpy(out, d_w_, 4 * n, hipMemcpyDeviceToHost));
}

void DeviceScene::get_vis(float* out, size_t n) const
{
	if (!d_vis_ || n > nitems_)
		throw std::runtime_error("get_vis: shadow factors were not kept");
	if (n)
		HIPCHECK(hipMemcpy(out, d_vis_, 4 * n, hipMemcpyDeviceToHost));
}

void DeviceScene::ambient_occlusion(const float* randoff, int num_samples, float distance, float multiplier,
                                    float falloff, float effect, const float* color, float* fac_out,
                                    const std::function<bool(float)>& between)
{
	hipStream_t s = (hipStream_t)stream_;
	if (nsamples_ == 0 || num_samples <= 0)
		return;
	View V = make_view(d_nodes_, d_dtri_, d_rtri_, ntri_, lo_, hi_, grid_, use_grid_ ? grid_base_ : 0, use_grid_ ? grid_fine_ : 0);
	// (i + 0.5) / n, its square root and sin(acos(.)) depend only on (i, n):
	// evaluated here with the C library the reference uses
	std::vector<float2> side(num_samples);
	for (int i = 0; i < num_samples; ++i) {
		float q = (i + 0.5f) / num_samples;
		float cs = sqrtf(q);
		side[i] = make_float2(cs, sinf(acosf(cs)));
	}
	float2* d_side = dalloc<float2>(num_samples);
	float* d_ro = dalloc<float>(nsamples_);
	HIPCHECK(hipMemcpyAsync(d_side, side.data(), sizeof(float2) * num_samples, hipMemcpyHostToDevice, s));
	HIPCHECK(hipMemcpyAsync(d_ro, randoff, 4 * nsamples_, hipMemcpyHostToDevice, s));
	HIPCHECK(hipMemsetAsync(d_counters_, 0, kNumCounters * sizeof(unsigned long long), s));
	AOParams P;
	P.n = num_samples;
	P.distance = distance;
	P.multiplier = multiplier;
	P.falloff = falloff;
	P.effect = effect;
	P.color[0] = color[0];
	P.color[1] = color[1];
	P.color[2] = color[2];
	// one launch, or, when the caller follows progress, chunks of texels
	// with the caller asked between them whether to go on
	const size_t chunk = between ? kAOChunk : nsamples_;
	EventTimer tm(s);
	for (size_t b = 0; b < nsamples_; b += chunk) {
		const size_t e = std::min(nsamples_, b + chunk);
		const unsigned ao_grid = (unsigned)((e - b + kBlock / 64 - 1) / (kBlock / 64)); // a wave per texel
		if (count_queries)
			k_ao<true><<<ao_grid, kBlock, march_lds_bytes(), s>>>(V, d_pos_, d_nrm_, d_ro, d_side, P, (uint32_t)b,
			                                                      (uint32_t)e, d_lm_, d_counters_,
			                                                      march_stack_depth(), fac_out);
		else
			k_ao<false><<<ao_grid, kBlock, march_lds_bytes(), s>>>(V, d_pos_, d_nrm_, d_ro, d_side, P, (uint32_t)b,
			                                                       (uint32_t)e, d_lm_, d_counters_,
			                                                       march_stack_depth(), fac_out);
		HIPCHECK(hipGetLastError());
		if (between) {
			HIPCHECK(hipStreamSynchronize(s));
			if (!between((float)e / (float)nsamples_))
				break;
		}
	}
	times.ao_ms = tm.stop();
	read_counters();
	dfree(d_side);
	dfree(d_ro);
}

// ---------------------------------------------------------------------------
void DeviceScene::query_distance(const float* pts, size_t n, float* out)
{
	View V = make_view(d_nodes_, d_dtri_, d_rtri_, ntri_, lo_, hi_, grid_, use_grid_ ? grid_base_ : 0, use_grid_ ? grid_fine_ : 0);
	float *dp = dalloc<float>(3 * n), *dout = dalloc<float>(n);
	HIPCHECK(hipMemcpy(dp, pts, 12 * n, hipMemcpyHostToDevice));
	k_q_distance<<<grid_for(n), kBlock, 0, (hipStream_t)stream_>>>(V, dp, (uint32_t)n, dout);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipStreamSynchronize((hipStream_t)stream_));
	HIPCHECK(hipMemcpy(out, dout, 4 * n, hipMemcpyDeviceToHost));
	dfree(dp);
	dfree(dout);
}

void DeviceScene::query_shadow(const float* from, const float* to, const float* k, size_t n, float* out,
                               uint32_t* steps)
{
	View V = make_view(d_nodes_, d_dtri_, d_rtri_, ntri_, lo_, hi_, grid_, use_grid_ ? grid_base_ : 0, use_grid_ ? grid_fine_ : 0);
	float *da = dalloc<float>(3 * n), *db = dalloc<float>(3 * n), *dk = dalloc<float>(n), *dout = dalloc<float>(n);
	uint32_t* dst = dalloc<uint32_t>(n);
	HIPCHECK(hipMemcpy(da, from, 12 * n, hipMemcpyHostToDevice));
	HIPCHECK(hipMemcpy(db, to, 12 * n, hipMemcpyHostToDevice));
	HIPCHECK(hipMemcpy(dk, k, 4 * n, hipMemcpyHostToDevice));
	k_q_shadow<<<grid_for(n), kBlock, 0, (hipStream_t)stream_>>>(V, da, db, dk, (uint32_t)n, dout, dst);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipStreamSynchronize((hipStream_t)stream_));
	HIPCHECK(hipMemcpy(out, dout, 4 * n, hipMemcpyDeviceToHost));
	if (steps)
		HIPCHECK(hipMemcpy(steps, dst, 4 * n, hipMemcpyDeviceToHost));
	dfree(da);
	dfree(db);
	dfree(dk);
	dfree(dout);
	dfree(dst);
}

void DeviceScene::query_closest(const float* a, const float* b, size_t n, float* out)
{
	View V = make_view(d_nodes_, d_dtri_, d_rtri_, ntri_, lo_, hi_, grid_, use_grid_ ? grid_base_ : 0, use_grid_ ? grid_fine_ : 0);
	float *da = dalloc<float>(3 * n), *db = dalloc<float>(3 * n), *dout = dalloc<float>(n);
	HIPCHECK(hipMemcpy(da, a, 12 * n, hipMemcpyHostToDevice));
	HIPCHECK(hipMemcpy(db, b, 12 * n, hipMemcpyHostToDevice));
	k_q_closest<<<grid_for(n), kBlock, 0, (hipStream_t)stream_>>>(V, da, db, (uint32_t)n, dout);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipStreamSynchronize((hipStream_t)stream_));
	HIPCHECK(hipMemcpy(out, dout, 4 * n, hipMemcpyDeviceToHost));
	dfree(da);
	dfree(db);
	dfree(dout);
}

void DeviceScene::query_occluded(const float* a, const float* b, size_t n, uint8_t* out)
{
	View V = make_view(d_nodes_, d_dtri_, d_rtri_, ntri_, lo_, hi_, grid_, use_grid_ ? grid_base_ : 0, use_grid_ ? grid_fine_ : 0);
	float *da = dalloc<float>(3 * n), *db = dalloc<float>(3 * n);
	uint8_t* dout = dalloc<uint8_t>(n);
	HIPCHECK(hipMemcpy(da, a, 12 * n, hipMemcpyHostToDevice));
	HIPCHECK(hipMemcpy(db, b, 12 * n, hipMemcpyHostToDevice));
	k_q_occluded<<<grid_for(n), kBlock, 0, (hipStream_t)stream_>>>(V, da, db, (uint32_t)n, dout);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipStreamSynchronize((hipStream_t)stream_));
	HIPCHECK(hipMemcpy(out, dout, n, hipMemcpyDeviceToHost));
	dfree(da);
	dfree(db);
	dfree(dout);
}

// ---------------------------------------------------------------------------
// radiosity
// ---------------------------------------------------------------------------
static void scan_u32_to_u64(const uint32_t* d_in, uint64_t n, uint64_t* d_out, uint64_t* h_total, hipStream_t s)
{
	constexpr uint64_t per_block = (uint64_t)kBlock * 16;
	uint32_t nb = (uint32_t)((n + per_block - 1) / per_block);
	if (nb == 0) {
		*h_total = 0;
		return;
	}
	uint64_t* sums = dalloc<uint64_t>(nb + 1);
	k_scan_blocks<<<nb, kBlock, 0, s>>>(d_in, n, d_out, sums);
	k_scan_sums<<<1, 1, 0, s>>>(sums, nb, sums + nb);
	k_scan_add<<<nb, kBlock, 0, s>>>(d_out, n, sums);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipMemcpyAsync(h_total, sums + nb, 8, hipMemcpyDeviceToHost, s));
	HIPCHECK(hipStreamSynchronize(s));
	dfree(sums);
}

void DeviceScene::rad_set_patches(const float* pos, const float* nrm, size_t n)
{
	dfree(rad.pos);
	dfree(rad.nrm);
	npatches = n;
	rad.pos = dalloc<float>(3 * n);
	rad.nrm = dalloc<float>(3 * n);
	if (n) {
		HIPCHECK(hipMemcpy(rad.pos, pos, 12 * n, hipMemcpyHostToDevice));
		HIPCHECK(hipMemcpy(rad.nrm, nrm, 12 * n, hipMemcpyHostToDevice));
	}
	// bounds and the longest normal, for the grid of the link search
	links_lo_[0] = links_lo_[1] = links_lo_[2] = INFINITY;
	links_hi_[0] = links_hi_[1] = links_hi_[2] = -INFINITY;
	links_nmax2_ = 0.0;
	links_finite_ = true;
	for (size_t i = 0; i < n; ++i) {
		double l2 = 0.0;
		for (int a = 0; a < 3; ++a) {
			const float p = pos[3 * i + a], q = nrm[3 * i + a];
			if (!std::isfinite(p) || !std::isfinite(q))
				links_finite_ = false;
			links_lo_[a] = std::min(links_lo_[a], p);
			links_hi_[a] = std::max(links_hi_[a], p);
			l2 += (double)q * q;
		}
		links_nmax2_ = std::max(links_nmax2_, l2);
	}
}

// All pairs (i, j > i) that pass the factor thresholds, in chunks of
// kPairChunk columns per row (the reference's loop, lighter.cpp:728-749).
void* DeviceScene::link_candidates_all(uint64_t& ncand, uint64_t* coff)
{
	hipStream_t s = (hipStream_t)stream_;
	const uint32_t N = (uint32_t)npatches;
	if (npatches > kMaxAllPairsPatches)
		throw std::runtime_error("radiosity: the all-pairs link search takes at most " +
		                         std::to_string(kMaxAllPairsPatches) + " patches (" + std::to_string(npatches) +
		                         (links_finite_ ? " given)" : " given; positions or normals are not finite)"));
	const uint32_t nchunks = (N + kPairChunk - 1) / kPairChunk;
	const uint64_t ncounts = (uint64_t)N * nchunks;
	DevArray<uint32_t> counts(ncounts);
	DevArray<uint64_t> offs(ncounts);
	const dim3 g(grid_for(N), nchunks);
	k_pairs<false><<<g, kBlock, 0, s>>>(rad.pos, rad.nrm, N, nchunks, counts.p, nullptr, nullptr);
	HIPCHECK(hipGetLastError());
	scan_u32_to_u64(counts.p, ncounts, offs.p, &ncand, s);
	uint2* cand = dalloc<uint2>(ncand);
	k_pairs<true><<<g, kBlock, 0, s>>>(rad.pos, rad.nrm, N, nchunks, nullptr, offs.p, cand);
	k_row_first<<<grid_for((uint64_t)N + 1), kBlock, 0, s>>>(N, nchunks, offs.p, ncand, coff);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipStreamSynchronize(s));
	return cand;
}

// The same candidate set from the 27 grid cells around each patch.
void* DeviceScene::link_candidates_grid(uint64_t& ncand, uint64_t* coff)
{
	hipStream_t s = (hipStream_t)stream_;
	const uint32_t N = (uint32_t)npatches;
	ncand = 0;
	// the longest linking distance (1 per mille margin over the float
	// rounding of the factor); cells of half that distance (1e-4 above),
	// searched two cells to each side: every pair within the distance is
	// at most two cells apart per axis, and the searched box is 2.5 rather
	// than 3 distances wide (about 40 % fewer pair tests than whole cells)
	const double r2 = links_nmax2_ / (M_PI * (double)kSmall) * 1.001;
	if (!(r2 > 0.0)) { // zero normals: no pair passes
		HIPCHECK(hipMemsetAsync(coff, 0, 8 * ((size_t)N + 1), s));
		HIPCHECK(hipStreamSynchronize(s));
		return dalloc<uint2>(0);
	}
	double h = std::sqrt(r2) / 2 * 1.0001;
	LinkGrid G;
	G.reach = 2;
	double na[3], cells = 0;
	for (;;) { // sized in double: a large extent over a short distance must not overflow the counts
		cells = 1;
		for (int a = 0; a < 3; ++a) {
			na[a] = std::floor(((double)links_hi_[a] - links_lo_[a]) / h) + 1;
			cells *= na[a];
		}
		if (cells <= (double)(1u << 24))
			break;
		h *= 1.26; // wider cells keep every linking pair within reach
	}
	for (int a = 0; a < 3; ++a) {
		G.lo[a] = links_lo_[a];
		G.n[a] = (uint32_t)na[a];
	}
	G.inv = (float)(1.0 / h);
	const uint32_t nc = G.n[0] * G.n[1] * G.n[2];
	DevArray<uint32_t> cell(N), ccount(nc + 1), order(N), counts((uint64_t)N + 1);
	DevArray<uint64_t> cstart(nc + 1);
	HIPCHECK(hipMemsetAsync(ccount.p, 0, 4 * ((size_t)nc + 1), s));
	k_grid_count<<<grid_for(N), kBlock, 0, s>>>(rad.pos, N, G, cell.p, ccount.p);
	HIPCHECK(hipGetLastError());
	if (links_mode == 0 && npatches <= kMaxAllPairsPatches) {
		// the pairs the grid tests (both orders, each patch against the
		// cells within reach) against the all-pairs loop's N^2 / 2: where
		// the grid does not prune (a room within one linking distance), the
		// loop is faster
		std::vector<uint32_t> hist(nc);
		HIPCHECK(hipMemcpyAsync(hist.data(), ccount.p, 4 * (size_t)nc, hipMemcpyDeviceToHost, s));
		HIPCHECK(hipStreamSynchronize(s));
		double tests = 0;
		for (uint32_t z = 0; z < G.n[2]; ++z)
			for (uint32_t y = 0; y < G.n[1]; ++y)
				for (uint32_t x = 0; x < G.n[0]; ++x) {
					const uint32_t c = (z * G.n[1] + y) * G.n[0] + x;
					if (!hist[c])
						continue;
					double nb = 0;
					const uint32_t q = (uint32_t)G.reach;
					for (uint32_t zz = z > q ? z - q : 0; zz <= std::min(z + q, G.n[2] - 1); ++zz)
						for (uint32_t yy = y > q ? y - q : 0; yy <= std::min(y + q, G.n[1] - 1); ++yy)
							for (uint32_t xx = x > q ? x - q : 0; xx <= std::min(x + q, G.n[0] - 1); ++xx)
								nb += hist[(zz * G.n[1] + yy) * G.n[0] + xx];
					tests += (double)hist[c] * nb;
				}
		if (tests > 0.25 * (double)N * (double)N)
			return link_candidates_all(ncand, coff);
	}
	uint64_t total = 0;
	scan_u32_to_u64(ccount.p, (uint64_t)nc + 1, cstart.p, &total, s);
	HIPCHECK(hipMemsetAsync(ccount.p, 0, 4 * ((size_t)nc + 1), s));
	k_grid_fill<<<grid_for(N), kBlock, 0, s>>>(N, cell.p, cstart.p, ccount.p, order.p);
	HIPCHECK(hipMemsetAsync(counts.p + N, 0, 4, s));
	k_pairs_grid<false><<<grid_for(N), kBlock, 0, s>>>(rad.pos, rad.nrm, N, G, cell.p, cstart.p, order.p, counts.p,
	                                                   nullptr, nullptr);
	HIPCHECK(hipGetLastError());
	scan_u32_to_u64(counts.p, (uint64_t)N + 1, coff, &ncand, s);
	uint2* cand = dalloc<uint2>(ncand);
	k_pairs_grid<true><<<grid_for(N), kBlock, 0, s>>>(rad.pos, rad.nrm, N, G, cell.p, cstart.p, order.p, nullptr,
	                                                  coff, cand);
	HIPCHECK(hipGetLastError());
	HIPCHECK(hipStreamSynchronize(s));
	return cand;
}

size_t DeviceScene::rad_links()
{
	hipStream_t s = (hipStream_t)stream_;
	const uint32_t N = (uint32_t)npatches;
	View V = make_view(d_nodes_, d_dtri_, d_rtri_, ntri_, lo_, hi_, grid_, use_grid_ ? grid_base_ : 0, use_grid_ ? grid_fine_ : 0);
	dfree(rad.rowptr);
	dfree(rad.rowptr64);
	dfree(rad.col);
	dfree(rad.val);
	dfree(d_cand_);
	dfree(d_clinked_);
	ncand_ = 0;
	rad.nnz = 0;
	rad.rowptr64 = dalloc<uint64_t>((uint64_t)N + 1);
	EventTimer tm(s);
	uint64_t ncand = 0;
	uint2* cand = nullptr;
	DevArray<uint64_t> coff((uint64_t)N + 1); // first candidate of each row
	if (N > 1) {
		cand = (uint2*)((links_mode == 1 || !links_finite_) ? link_candidates_all(ncand, coff.p)
		                                                    : link_candidates_grid(ncand, coff.p));
	} else {
		HIPCHECK(hipMemsetAsync(coff.p, 0, 8 * ((size_t)N + 1), s));
	}
	// occlusion test of every candidate; links per row below the diagonal
	// (atomics from the other end) and above it (the row's own candidates)
	DevArray<uint32_t> lower((uint64_t)N + 1), upper((uint64_t)N + 1), cnt((uint64_t)N + 1);
	HIPCHECK(hipMemsetAsync(lower.p, 0, 4 * ((size_t)N + 1), s));
	HIPCHECK(hipMemsetAsync(cnt.p, 0, 4 * ((size_t)N + 1), s));
	uint8_t* linked = dalloc<uint8_t>(ncand);
	if (ncand) {
		k_link_test<<<grid_for(ncand), kBlock, march_lds_bytes(), s>>>(V, rad.pos, cand, ncand, linked, lower.p,
		                                                               march_stack_depth());
		HIPCHECK(hipGetLastError());
	}
	if (N) {
		k_link_upper<<<gridn_waves(N), kBlock, 0, s>>>(N, coff.p, cand, linked, upper.p, nullptr, nullptr, nullptr);
		k_row_sum<<<grid_for(N), kBlock, 0, s>>>(N, upper.p, lower.p, cnt.p);
		HIPCHECK(hipGetLastError());
	}
	// symmetric CSR, columns ascending within a row; 32-bit row offsets too
	// while they fit (the solver forms take those)
	uint64_t nnz = 0;
	scan_u32_to_u64(cnt.p, (uint64_t)N + 1, rad.rowptr64, &nnz, s);
	if (nnz < (1ull << 32)) {
		rad.rowptr = dalloc<uint32_t>((uint64_t)N + 1);
		k_u64_to_u32<<<grid_for((uint64_t)N + 1), kBlock, 0, s>>>(rad.rowptr64, (uint64_t)N + 1, rad.rowptr);
		HIPCHECK(hipGetLastError());
	}
	HIPCHECK(hipStreamSynchronize(s));
	rad.nnz = nnz;
	rad.col = dalloc<uint32_t>(nnz);
	rad.val = dalloc<float>(nnz);
	if (nnz) {
		DevArray<uint32_t> colx(nnz);
		k_link_upper<<<gridn_waves(N), kBlock, 0, s>>>(N, coff.p, cand, linked, nullptr, rad.rowptr64, lower.p,
		                                                colx.p);
		HIPCHECK(hipMemsetAsync(upper.p, 0, 4 * ((size_t)N + 1), s)); // the lower parts' fill counters
		k_link_lower<<<grid_for(ncand), kBlock, 0, s>>>(cand, linked, ncand, rad.rowptr64, upper.p, colx.p);
		HIPCHECK(hipGetLastError());
		int bits = 1;
		while (bits < 32 && (1ull << bits) < N)
			++bits;
		sort_row_columns(colx.p, rad.col, nnz, N, rad.rowptr64, bits, s);
		k_link_factors<<<gridn_waves(N), kBlock, 0, s>>>(N, rad.rowptr64, rad.col, rad.pos, rad.nrm, rad.val);
		HIPCHECK(hipGetLastError());
		HIPCHECK(hipStreamSynchronize(s)); // before colx goes
	}
	HIPCHECK(hipStreamSynchronize(s));
	if (count_queries) { // kept for reference_counts_links
		d_cand_ = (void*)cand;
		d_clinked_ = (void*)linked;
		ncand_ = ncand;
	} else {
		dfree(cand);
		dfree(linked);
	}
	times.links_ms = tm.stop();
	nlinks_ = nnz / 2;
	return nlinks_;
}

void DeviceScene::rad_get_csr(std::vector<uint32_t>& rowptr, std::vector<uint32_t>& col,
                              std::vector<float>& val) const
{
	if (!rad.rowptr)
		throw std::runtime_error("radiosity: the link CSR holds " + std::to_string(rad.nnz) +
		                         " entries; reading it takes fewer than 2^32");
	rowptr.resize(npatches + 1);
	col.resize(rad.nnz);
	val.resize(rad.nnz);
	HIPCHECK(hipMemcpy(rowptr.data(), rad.rowptr, 4 * (npatches + 1), hipMemcpyDeviceToHost));
	if (rad.nnz) {
		HIPCHECK(hipMemcpy(col.data(), rad.col, 4 * rad.nnz, hipMemcpyDeviceToHost));
		HIPCHECK(hipMemcpy(val.data(), rad.val, 4 * rad.nnz, hipMemcpyDeviceToHost));
	}
}

void DeviceScene::rad_get_links(uint32_t* rowmap, uint32_t* other, float* factor) const
{
	std::vector<uint32_t> rp, col;
	std::vector<float> val;
	rad_get_csr(rp, col, val);
	uint32_t o = 0;
	for (size_t i = 0; i < npatches; ++i) {
		rowmap[2 * i] = o;
		uint32_t c = 0;
		for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
			if (col[k] > i) {
				other[o] = col[k];
				factor[o] = val[k];
				o++;
				c++;
			}
		}
		rowmap[2 * i + 1] = c;
	}
}

void DeviceScene::rad_set_colors(const float* diffuse3, const float* emit3, const float* area)
{
	size_t n = npatches;
	dfree(rad.diffuse);
	dfree(rad.area);
	dfree(rad.total);
	dfree(rad.out);
	dfree(rad.in);
	dfree(rad.x);
	rad.diffuse = dalloc<float>(3 * n);
	rad.area = dalloc<float>(n);
	rad.total = dalloc<float>(3 * n);
	rad.out = dalloc<float>(3 * n);
	rad.in = dalloc<float>(3 * n);
	rad.x = dalloc<float>(4 * n);
	if (n) {
		HIPCHECK(hipMemcpy(rad.diffuse, diffuse3, 12 * n, hipMemcpyHostToDevice));
		HIPCHECK(hipMemcpy(rad.area, area, 4 * n, hipMemcpyHostToDevice));
		HIPCHECK(hipMemcpy(rad.total, emit3, 12 * n, hipMemcpyHostToDevice));
		HIPCHECK(hipMemcpy(rad.out, emit3, 12 * n, hipMemcpyHostToDevice));
	}
}

void DeviceScene::rad_bounce(int count)
{
	hipStream_t s = (hipStream_t)stream_;
	uint32_t N = (uint32_t)npatches;
	EventTimer tm(s);
	for (int b = 0; b < count && N; ++b) {
		k_bounce_prep<<<grid_for(N), kBlock, 0, s>>>(N, rad.out, rad.diffuse, rad.area, (float4*)rad.x);
		k_bounce_rows<<<gridn_waves(N), kBlock, 0, s>>>(N, rad.rowptr64, rad.col, rad.val, (const float4*)rad.x,
		                                                 rad.in, rad.total);
		HIPCHECK(hipGetLastError());
		std::swap(rad.out, rad.in);
	}
	times.bounce_ms = tm.stop();
}

void DeviceScene::rad_get_total(float* out3) const
{
	if (npatches)
		HIPCHECK(hipMemcpy(out3, rad.total, 12 * npatches, hipMemcpyDeviceToHost));
}

} // namespace lgx
