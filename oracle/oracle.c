/* CPU oracle for the lightmap-baking hot path — TEST INFRASTRUCTURE ONLY.
 * See oracle.h for scope.  Each function names the reference lines it
 * restates; the arithmetic keeps the reference's operation order because the
 * tests compare bit patterns. */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_SMALL 0.001f      /* SMALL_FLOAT, lighter_int.hpp:261 */
#define ORC_PEN_SIZE 2.0f     /* MAX_PENUMBRA_SIZE, lighter_int.hpp:964 */
#define ORC_PEN_STEP 1.0f     /* MAX_PENUMBRA_STEP, lighter_int.hpp:966 */
#define ORC_SHADOW_OFS 0.005f /* SAMPLE_SHADOW_OFFSET, lighter_int.hpp:967 */
#define ORC_PI_F ((float)M_PI)

typedef struct { float x, y, z; } v3;

static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 ld(const float* p) { return mk(p[0], p[1], p[2]); }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 scl(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 sclf(float s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float lensq(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline float len(v3 a) { return sqrtf(lensq(a)); }
static inline v3 cross(v3 a, v3 b)
{
	return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* Vec3::Normalized, lighter_int.hpp:404-415 */
static inline v3 nrmz(v3 a)
{
	float l2 = lensq(a);
	if (l2 == 0)
		return mk(0, 0, 0);
	float inv = 1.0f / sqrtf(l2);
	return mk(a.x * inv, a.y * inv, a.z * inv);
}
/* TMIN / TMAX, lighter_int.hpp:277-278 (NaN behaviour of the ternary kept) */
static inline float tmin(float a, float b) { return a < b ? a : b; }
static inline float tmax(float a, float b) { return a > b ? a : b; }
static inline int near_zero(v3 a)
{
	return fabsf(a.x) < ORC_SMALL && fabsf(a.y) < ORC_SMALL && fabsf(a.z) < ORC_SMALL;
}

/* IntersectLineSegmentTriangle, lighter_math.cpp:315-354 */
static float seg_tri(v3 L1, v3 L2, v3 P1, v3 P2, v3 P3)
{
	v3 u = sub(P2, P1), v = sub(P3, P1);
	v3 n = cross(u, v);
	if (near_zero(n))
		return 2.0f;
	v3 dir = sub(L2, L1);
	v3 w0 = sub(L1, P1);
	float a = -dot(n, w0);
	float b = dot(n, dir);
	if (fabsf(b) < ORC_SMALL)
		return 2.0f;
	float r = a / b;
	if (r < 0.0f || r > 1.0f)
		return 2.0f;
	v3 I = add(L1, sclf(r, dir));
	float uu = dot(u, u), uv = dot(u, v), vv = dot(v, v);
	v3 w = sub(I, P1);
	float wu = dot(w, u), wv = dot(w, v);
	float D = uv * uv - uu * vv;
	float s = (uv * wv - vv * wu) / D;
	if (s < 0.0f || s > 1.0f)
		return 2.0f;
	float t = (uv * wu - uu * wv) / D;
	if (t < 0.0f || (s + t) > 1.0f)
		return 2.0f;
	return r;
}

/* PointTriangleDistance, lighter_math.cpp:875-913 */
static float pt_tri(v3 pt, v3 t0, v3 t1, v3 t2)
{
	v3 nrm = nrmz(cross(sub(t1, t0), sub(t2, t0)));
	float pd = fabsf(dot(nrm, pt) - dot(nrm, t0));
	v3 tan0 = nrmz(sub(t1, t0)), tan1 = nrmz(sub(t2, t1)), tan2 = nrmz(sub(t0, t2));
	float t0p = dot(tan0, pt), t0min = dot(tan0, t0), t0max = dot(tan0, t1);
	float t1p = dot(tan1, pt), t1min = dot(tan1, t1), t1max = dot(tan1, t2);
	float t2p = dot(tan2, pt), t2min = dot(tan2, t2), t2max = dot(tan2, t0);
	if (t0min >= t0p && t2max <= t2p) return len(sub(pt, t0));
	if (t1min >= t1p && t0max <= t0p) return len(sub(pt, t1));
	if (t2min >= t2p && t1max <= t1p) return len(sub(pt, t2));
	v3 en0 = nrmz(cross(sub(t1, t0), nrm));
	v3 en1 = nrmz(cross(sub(t2, t1), nrm));
	v3 en2 = nrmz(cross(sub(t0, t2), nrm));
	float ptd0 = dot(en0, pt) - dot(en0, t0);
	float ptd1 = dot(en1, pt) - dot(en1, t1);
	float ptd2 = dot(en2, pt) - dot(en2, t2);
	/* V2(pd, ptd).Length() = sqrtf(pd*pd + ptd*ptd), lighter_int.hpp:324 */
	if (ptd0 >= 0 && t0p >= t0min && t0p <= t0max) return sqrtf(pd * pd + ptd0 * ptd0);
	if (ptd1 >= 0 && t1p >= t1min && t1p <= t1max) return sqrtf(pd * pd + ptd1 * ptd1);
	if (ptd2 >= 0 && t2p >= t2min && t2p <= t2max) return sqrtf(pd * pd + ptd2 * ptd2);
	return pd;
}

float orc_seg_tri(const float* l1, const float* l2, const float* p1, const float* p2, const float* p3)
{
	return seg_tri(ld(l1), ld(l2), ld(p1), ld(p2), ld(p3));
}

float orc_point_tri_dist(const float* p, const float* t0, const float* t1, const float* t2)
{
	return pt_tri(ld(p), ld(t0), ld(t1), ld(t2));
}

/* Triangle::CheckIsUseful, lighter_int.hpp:604-608 */
int orc_tri_useful(const float* t)
{
	v3 P1 = ld(t), P2 = ld(t + 3), P3 = ld(t + 6);
	return !near_zero(cross(sub(P2, P1), sub(P3, P1)));
}

/* Vec3::CreateSpiralDirVector, lighter_int.hpp:456-471 */
static v3 spiral(v3 dir, float randoff, int i, int n)
{
	const float golden = 137.508f / 180.0f * ORC_PI_F; /* DEG2RAD(137.508f) */
	float q = (i + 0.5f) / n;
	float cos_side = sqrtf(q);
	float sin_side = sinf(acosf(cos_side));
	float angle = (i + randoff) * golden;
	float cos_around = cosf(angle);
	float sin_around = sinf(angle);
	v3 diffvec = mk(dir.y, -dir.z, dir.x);
	v3 up = nrmz(cross(dir, diffvec));
	v3 rt = cross(dir, up);
	v3 a = sclf(cos_around * sin_side, rt);
	v3 b = sclf(sin_around * sin_side, up);
	v3 c = sclf(cos_side, dir);
	return add(add(a, b), c);
}

void orc_spiral(const float* dir, float randoff, int i, int n, float* out)
{
	v3 r = spiral(ld(dir), randoff, i, n);
	out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* ltr_Scene::Distance / SceneDistanceQuery, lighter.cpp:150-188, with the
 * per-instance DistanceBBQuery, lighter_math.cpp:954-989: the minimum of
 * PointTriangleDistance over all shadow triangles, capped at 2.0. */
static float scene_distance(const float* tris, size_t ntri, v3 p)
{
	float dist = ORC_PEN_SIZE;
	for (size_t i = 0; i < ntri; ++i) {
		const float* t = tris + 9 * i;
		float d = pt_tri(p, ld(t), ld(t + 3), ld(t + 6));
		if (d < dist)
			dist = d;
	}
	return dist;
}

float orc_scene_distance(const float* tris, size_t ntri, const float* p)
{
	return scene_distance(tris, ntri, ld(p));
}

/* ltr_Scene::CalcInvShadowFactor, lighter.cpp:190-207 */
static float shadow_factor(const float* tris, size_t ntri, v3 from, v3 to, float k, uint32_t* steps)
{
	v3 rd = nrmz(sub(to, from));
	float mint = 0.001f;
	float maxt = len(sub(to, from));
	float res = 1.0f;
	uint32_t n = 0;
	for (float t = mint; t < maxt;) {
		float h = scene_distance(tris, ntri, add(from, scl(rd, t)));
		n++;
		if (h < 0.001f) {
			res = 0.0f;
			break;
		}
		res = tmin(res, h / tmin(t * k, ORC_PEN_SIZE));
		h = tmin(h, ORC_PEN_STEP);
		t += h;
	}
	if (steps)
		*steps = n;
	return res;
}

float orc_shadow_factor(const float* tris, size_t ntri, const float* from, const float* to, float k,
                        uint32_t* steps_out)
{
	return shadow_factor(tris, ntri, ld(from), ld(to), k, steps_out);
}

/* ltr_Scene::DistanceTestBBT, lighter.cpp:254-289 (segment shrunk by
 * SMALL_FLOAT at both ends; minimum r over all shadow triangles, 2 = miss) */
static float closest_hit(const float* tris, size_t ntri, v3 A, v3 B)
{
	v3 dn = nrmz(sub(B, A));
	v3 mA = add(A, scl(dn, ORC_SMALL));
	v3 mB = sub(B, scl(dn, ORC_SMALL));
	float closest = 2.0f;
	for (size_t i = 0; i < ntri; ++i) {
		const float* t = tris + 9 * i;
		float d = seg_tri(mA, mB, ld(t), ld(t + 3), ld(t + 6));
		if (d < closest)
			closest = d;
	}
	return closest;
}

float orc_closest_hit(const float* tris, size_t ntri, const float* a, const float* b)
{
	return closest_hit(tris, ntri, ld(a), ld(b));
}

/* ltr_Scene::VisibilityTest, lighter.cpp:112-147: true when any triangle is
 * hit with r < 1 on the shrunk segment (i.e. the pair is occluded). */
static int occluded(const float* tris, size_t ntri, v3 A, v3 B)
{
	v3 dn = nrmz(sub(B, A));
	v3 mA = add(A, scl(dn, ORC_SMALL));
	v3 mB = sub(B, scl(dn, ORC_SMALL));
	for (size_t i = 0; i < ntri; ++i) {
		const float* t = tris + 9 * i;
		if (seg_tri(mA, mB, ld(t), ld(t + 3), ld(t + 6)) < 1.0f)
			return 1;
	}
	return 0;
}

int orc_occluded(const float* tris, size_t ntri, const float* a, const float* b)
{
	return occluded(tris, ntri, ld(a), ld(b));
}

/* Job_LMRender_{Point,Spot,Direct}_Inner, lighter.cpp:485-600, with the spot
 * angle set-up of Int_LMRender, lighter.cpp:639-645. */
size_t orc_direct_light(const float* tris, size_t ntri, size_t S, const float* pos, const float* nrm,
                        const orc_light* L, float* lightmap, float* weight_out)
{
	v3 lpos = ld(L->pos), ldir = ld(L->dir), col = ld(L->color);
	float out_rad = L->spot_out_deg / 180.0f * ORC_PI_F;
	float in_rad = L->spot_in_deg / 180.0f * ORC_PI_F;
	if (in_rad == out_rad)
		in_rad -= ORC_SMALL;
	float angle_diff = in_rad - out_rad;
	size_t marches = 0;
	for (size_t i = 0; i < S; ++i) {
		v3 SP = ld(pos + 3 * i), SN = ld(nrm + 3 * i);
		float w;
		if (L->type == 3) {
			float f_ndotl = tmax(0.0f, dot(ldir, SN));
			float f_vis = shadow_factor(tris, ntri, add(SP, scl(SN, ORC_SHADOW_OFS)),
			                            add(SP, scl(ldir, L->range)), L->radius, NULL);
			if (f_ndotl > 0)
				marches++;
			w = f_ndotl * f_vis;
		} else if (L->type == 1 || L->type == 2) {
			v3 s2l = sub(lpos, SP);
			float dist = len(s2l);
			if (dist)
				s2l = mk(s2l.x / dist, s2l.y / dist, s2l.z / dist);
			float f_dist = powf(1 - tmin(1.0f, dist / L->range), L->power);
			float f_ndotl = tmax(0.0f, dot(s2l, SN));
			float f_dir = 1.0f;
			float gate;
			if (L->type == 2) {
				float angle = acosf(tmin(1.0f, dot(s2l, neg(ldir))));
				f_dir = tmax(0.0f, tmin(1.0f, (angle - out_rad) / angle_diff));
				f_dir = powf(f_dir, L->spot_curve);
				gate = f_dist * f_ndotl * f_dir;
			} else {
				gate = f_dist * f_ndotl;
			}
			if (gate <= 0) {
				if (weight_out)
					weight_out[i] = -1.0f;
				continue;
			}
			marches++;
			float f_vis = shadow_factor(tris, ntri, add(SP, scl(SN, ORC_SHADOW_OFS)), lpos, L->radius, NULL);
			w = (L->type == 2) ? f_dist * f_ndotl * f_dir * f_vis : f_dist * f_ndotl * f_vis;
		} else {
			if (weight_out)
				weight_out[i] = -1.0f;
			continue;
		}
		if (weight_out)
			weight_out[i] = w;
		v3 lm = ld(lightmap + 3 * i);
		lm = add(lm, scl(col, w));
		lightmap[3 * i + 0] = lm.x;
		lightmap[3 * i + 1] = lm.y;
		lightmap[3 * i + 2] = lm.z;
	}
	return marches;
}

/* Job_AORender_Inner, lighter.cpp:799-841 */
void orc_ao(const float* tris, size_t ntri, size_t S, const float* pos, const float* nrm,
            const float* randoff, float ao_distance, float ao_multiplier, float ao_falloff,
            float ao_effect, const float* ao_color, int num_samples, float* lightmap,
            float* ao_factor_out)
{
	v3 aocol = ld(ao_color);
	for (size_t i = 0; i < S; ++i) {
		v3 SP = ld(pos + 3 * i), SN = ld(nrm + 3 * i);
		v3 out = ld(lightmap + 3 * i);
		v3 origin = add(SP, scl(SN, ORC_SMALL * 2));
		float ao = 0;
		for (int s = 0; s < num_samples; ++s) {
			v3 rdir = scl(spiral(SN, randoff[i], s, num_samples), ao_distance);
			float hit = closest_hit(tris, ntri, origin, add(origin, rdir));
			if (hit < 1.0f)
				ao += 1.0f - hit;
		}
		ao /= num_samples;
		ao = tmin(ao * ao_multiplier, 1.0f);
		if (ao_falloff)
			ao = powf(ao, ao_falloff);
		if (ao_factor_out)
			ao_factor_out[i] = ao;
		if (ao_effect >= 0) {
			out = add(scl(out, 1 - ao * (1 - ao_effect)), scl(aocol, ao));
		} else {
			/* TLERP(a, b, s) = a * (1 - s) + b * s, lighter_int.hpp:285 */
			float e = -ao_effect;
			v3 inner = add(scl(aocol, 1.0f - e), scl(mul(aocol, out), e));
			out = add(scl(out, 1.0f - ao), scl(inner, ao));
		}
		lightmap[3 * i + 0] = out.x;
		lightmap[3 * i + 1] = out.y;
		lightmap[3 * i + 2] = out.z;
	}
}

/* Int_RDGenLinks pair loop, lighter.cpp:728-759 */
size_t orc_rad_links(const float* tris, size_t ntri, size_t N, const float* pos, const float* nrm,
                     uint32_t* rowmap, uint32_t* other, float* factor, size_t cap)
{
	size_t nl = 0;
	for (size_t i = 0; i < N; ++i) {
		rowmap[2 * i] = (uint32_t)nl;
		rowmap[2 * i + 1] = 0;
		v3 ASP = ld(pos + 3 * i), ASN = ld(nrm + 3 * i);
		for (size_t j = i + 1; j < N; ++j) {
			v3 BSP = ld(pos + 3 * j), BSN = ld(nrm + 3 * j);
			v3 d = sub(BSP, ASP);
			float dA = dot(ASN, d);
			float dB = dot(BSN, neg(d));
			if (dA <= ORC_SMALL || dB <= ORC_SMALL)
				continue;
			float l2 = lensq(d);
			float f = dA * dB / (l2 * l2 * ORC_PI_F);
			if (f < ORC_SMALL)
				continue;
			if (!occluded(tris, ntri, ASP, BSP))
				continue;
			if (nl < cap) {
				other[nl] = (uint32_t)j;
				factor[nl] = f;
			}
			nl++;
			rowmap[2 * i + 1]++;
		}
	}
	return nl;
}

/* Int_RDBounce, lighter.cpp:767-797 */
void orc_rad_bounce_f32(size_t N, const uint32_t* rowmap, const uint32_t* other, const float* factor,
                        float* c)
{
	for (size_t i = 0; i < N; ++i) {
		size_t b = rowmap[2 * i], e = b + rowmap[2 * i + 1];
		for (size_t s = b; s < e; ++s) {
			size_t j = other[s];
			float f = factor[s];
			float* A = c + 13 * i;
			float* B = c + 13 * j;
			v3 Aout = scl(scl(mul(ld(A + 6), ld(A + 0)), A[12]), f);
			v3 Bout = scl(scl(mul(ld(B + 6), ld(B + 0)), B[12]), f);
			for (int k = 0; k < 3; ++k) {
				A[3 + k] += (&Bout.x)[k];
				B[3 + k] += (&Aout.x)[k];
				A[9 + k] += (&Bout.x)[k];
				B[9 + k] += (&Aout.x)[k];
			}
		}
	}
	for (size_t i = 0; i < N; ++i) {
		float* A = c + 13 * i;
		memcpy(A + 6, A + 9, 12);
		A[9] = A[10] = A[11] = 0;
	}
}

/* Closed form of the bounce loop + commit (lighter.cpp:1104-1123):
 * total = sum_{k=0..b} K^k E with K = F diag(area) diag(diffuse_c), F the
 * symmetric link matrix, evaluated in FP64. */
void orc_rad_series_f64(size_t N, const uint32_t* rowmap, const uint32_t* other, const float* factor,
                        const double* area, const double* diffuse3, const double* E, int R, int bounces,
                        double* total)
{
	double* x = (double*)malloc(sizeof(double) * N * (size_t)R);
	double* y = (double*)malloc(sizeof(double) * N * (size_t)R);
	memcpy(x, E, sizeof(double) * N * (size_t)R);
	memcpy(total, E, sizeof(double) * N * (size_t)R);
	for (int b = 0; b < bounces; ++b) {
		memset(y, 0, sizeof(double) * N * (size_t)R);
		for (size_t i = 0; i < N; ++i) {
			size_t s0 = rowmap[2 * i], s1 = s0 + rowmap[2 * i + 1];
			for (size_t s = s0; s < s1; ++s) {
				size_t j = other[s];
				double f = factor[s];
				for (int r = 0; r < R; ++r) {
					int ch = r % 3;
					y[i * R + r] += f * area[j] * diffuse3[3 * j + ch] * x[j * R + r];
					y[j * R + r] += f * area[i] * diffuse3[3 * i + ch] * x[i * R + r];
				}
			}
		}
		for (size_t k = 0; k < N * (size_t)R; ++k)
			total[k] += y[k];
		double* t = x; x = y; y = t;
	}
	free(x);
	free(y);
}
