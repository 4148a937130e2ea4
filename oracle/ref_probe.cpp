// Probe driver for the reference baker (archo5/lighter), compiled together with
// the reference's own sources straight from the reference checkout (see
// oracle/Makefile).  It is test infrastructure only: it lets the tests run the
// reference's bake one stage at a time and read the state between stages
// (samples, per-stage lightmaps, radiosity links), and it exposes the
// reference's scene queries for point-wise parity checks.
//
// Stage order follows Job_MainProc (lighter.cpp:1046-1145); every stage body
// is the reference's own member function, this file only sequences them.
#include "lighter_int.hpp"

#include <algorithm>
#include <vector>

// Defined with C++ linkage in lighter_math.cpp:315.
float IntersectLineSegmentTriangle(const Vec3& L1, const Vec3& L2, const Vec3& P1, const Vec3& P2, const Vec3& P3);

extern "C" {

// PreXForm, ColInfo, instance tree, Samples (lighter.cpp:1050-1073).
int rp_prepare(ltr_Scene* S, int nthreads)
{
	if (S->m_meshInstances.empty())
		return 0;
	S->m_worker.Init(nthreads < 1 ? 1 : nthreads);
	S->m_worker.DoWork(S, NULL, 0, S->m_meshInstances.size(), ltr_Scene::Job_PreXForm);
	S->m_worker.DoWork(S, NULL, 0, S->m_meshInstances.size(), ltr_Scene::Job_ColInfo);
	std::vector<AABB3> bbs;
	for (size_t i = 0; i < S->m_meshInstances.size(); ++i) {
		AABBTree::Node& N = S->m_meshInstances[i]->m_triTree.m_bbTree.m_nodes[0];
		AABB3 bb = {N.bbmin, N.bbmax};
		bbs.push_back(bb);
	}
	S->m_instTree.SetAABBs(&bbs[0], bbs.size());
	S->m_worker.DoWork(S, NULL, 0, S->m_meshInstances.size(), ltr_Scene::Job_Samples);
	return 1;
}

// PreXForm, ColInfo and the instance tree only: the collision structures the
// scene queries need, without sample placement.
int rp_prepare_collision(ltr_Scene* S, int nthreads)
{
	if (S->m_meshInstances.empty())
		return 0;
	S->m_worker.Init(nthreads < 1 ? 1 : nthreads);
	S->m_worker.DoWork(S, NULL, 0, S->m_meshInstances.size(), ltr_Scene::Job_PreXForm);
	S->m_worker.DoWork(S, NULL, 0, S->m_meshInstances.size(), ltr_Scene::Job_ColInfo);
	std::vector<AABB3> bbs;
	for (size_t i = 0; i < S->m_meshInstances.size(); ++i) {
		AABBTree::Node& N = S->m_meshInstances[i]->m_triTree.m_bbTree.m_nodes[0];
		AABB3 bb = {N.bbmin, N.bbmax};
		bbs.push_back(bb);
	}
	S->m_instTree.SetAABBs(&bbs[0], bbs.size());
	return 1;
}

// CalcInvShadowFactor over a batch on nthreads threads (contiguous chunks),
// for timing the reference's shadow march on the host cores.
struct ShadowJob {
	ltr_Scene* S;
	const float *from, *to, *k;
	float* out;
	size_t b, e;
};
static void* shadow_job(void* p)
{
	ShadowJob* j = (ShadowJob*)p;
	for (size_t i = j->b; i < j->e; ++i)
		j->out[i] = j->S->CalcInvShadowFactor(V3P(j->from + 3 * i), V3P(j->to + 3 * i), j->k[i]);
	return NULL;
}
void rp_shadow_factor_batch_mt(ltr_Scene* S, size_t n, const float* from, const float* to, const float* k,
                               float* out, int nthreads)
{
	if (nthreads < 1)
		nthreads = 1;
	std::vector<pthread_t> th(nthreads);
	std::vector<ShadowJob> jobs(nthreads);
	for (int t = 0; t < nthreads; ++t) {
		ShadowJob j = {S, from, to, k, out, n * t / nthreads, n * (t + 1) / nthreads};
		jobs[t] = j;
		pthread_create(&th[t], NULL, shadow_job, &jobs[t]);
	}
	for (int t = 0; t < nthreads; ++t)
		pthread_join(th[t], NULL);
}

// The AO factor of Job_AORender_Inner (lighter.cpp:812-828) for a batch of
// texels given as position, normal and spiral offset, on nthreads threads
// (contiguous chunks): the per-texel ray loop over the reference's own
// CreateSpiralDirVector and DistanceTestBBT, up to the clamp by the
// multiplier (falloff and the colour blend are left to the caller).
struct AoJob {
	ltr_Scene* S;
	const float *pos, *nrm, *randoff;
	int nsamples;
	float distance, multiplier;
	float* out;
	size_t b, e;
};
static void* ao_job(void* p)
{
	AoJob* j = (AoJob*)p;
	for (size_t i = j->b; i < j->e; ++i) {
		Vec3 SP = V3P(j->pos + 3 * i), SN = V3P(j->nrm + 3 * i);
		Vec3 origin = SP + SN * (SMALL_FLOAT * 2);
		float f = 0;
		for (int s = 0; s < j->nsamples; ++s) {
			Vec3 dir = Vec3::CreateSpiralDirVector(SN, j->randoff[i], s, j->nsamples) * j->distance;
			float hit = j->S->DistanceTestBBT(origin, origin + dir);
			if (hit < 1.0f)
				f += 1.0f - hit;
		}
		f /= j->nsamples;
		j->out[i] = TMIN(f * j->multiplier, 1.0f);
	}
	return NULL;
}
void rp_ao_factor_batch_mt(ltr_Scene* S, size_t n, const float* pos, const float* nrm, const float* randoff,
                           int nsamples, float distance, float multiplier, float* out, int nthreads)
{
	if (nthreads < 1)
		nthreads = 1;
	std::vector<pthread_t> th(nthreads);
	std::vector<AoJob> jobs(nthreads);
	for (int t = 0; t < nthreads; ++t) {
		AoJob j = {S, pos, nrm, randoff, nsamples, distance, multiplier, out, n * t / nthreads, n * (t + 1) / nthreads};
		jobs[t] = j;
		pthread_create(&th[t], NULL, ao_job, &jobs[t]);
	}
	for (int t = 0; t < nthreads; ++t)
		pthread_join(th[t], NULL);
}

// Instance ids a light is rendered into, in render order (lighter.cpp:1080-1093).
int rp_light_instances(ltr_Scene* S, int light, int32_t* out, int cap)
{
	std::vector<int32_t> ids;
	ids.push_back(0);
	S->m_lights[light].QueryMeshInsts(S->m_instTree, ids);
	std::sort(ids.begin(), ids.end());
	int n = 0;
	int32_t prev = -1;
	for (size_t i = 0; i < ids.size(); ++i) {
		if (i > 0 && ids[i] == prev)
			continue;
		prev = ids[i];
		if (n < cap)
			out[n] = prev;
		n++;
	}
	return n;
}

// Direct light for lights [l0, l1) (lighter.cpp:1075-1096).
void rp_direct(ltr_Scene* S, int l0, int l1)
{
	std::vector<int32_t> ids(S->m_meshInstances.size() + 1);
	for (int l = l0; l < l1; ++l) {
		int n = rp_light_instances(S, l, &ids[0], (int)ids.size());
		for (int k = 0; k < n; ++k)
			S->Int_LMRender(S->m_lights[l], S->m_meshInstances[ids[k]]);
	}
}

void rp_rad_links(ltr_Scene* S) { S->Int_RDGenLinks(); }

void rp_rad_bounce(ltr_Scene* S, int count)
{
	for (int b = 0; b < count; ++b)
		S->Int_RDBounce();
}

// "committing radiosity" (lighter.cpp:1112-1123).
void rp_rad_commit(ltr_Scene* S)
{
	size_t sp = 0;
	for (size_t m = 0; m < S->m_meshInstances.size(); ++m) {
		ltr_MeshInstance* mi = S->m_meshInstances[m];
		for (size_t i = 0; i < mi->m_lightmap.size(); ++i)
			mi->m_lightmap[i] = S->m_radSampleColors[sp++].totalLight;
	}
}

void rp_ao(ltr_Scene* S)
{
	for (size_t m = 0; m < S->m_meshInstances.size(); ++m) {
		ltr_MeshInstance* mi = S->m_meshInstances[m];
		S->m_worker.DoWork(S, mi, 0, mi->m_lightmap.size(), ltr_Scene::Job_AORender);
	}
}

void rp_finalize(ltr_Scene* S) { S->Int_Finalize(); }

// ---- state access --------------------------------------------------------
int rp_num_instances(ltr_Scene* S) { return (int)S->m_meshInstances.size(); }

void rp_instance_info(ltr_Scene* S, int i, int32_t* out4)
{
	ltr_MeshInstance* mi = S->m_meshInstances[i];
	out4[0] = (int32_t)mi->lm_width;
	out4[1] = (int32_t)mi->lm_height;
	out4[2] = (int32_t)mi->m_samples_pos.size();
	out4[3] = (int32_t)mi->m_triTree.m_tris.size();
}

void rp_get_samples(ltr_Scene* S, int i, float* pos, float* nrm, uint32_t* loc, float* radinfo)
{
	ltr_MeshInstance* mi = S->m_meshInstances[i];
	for (size_t k = 0; k < mi->m_samples_pos.size(); ++k) {
		memcpy(pos + 3 * k, &mi->m_samples_pos[k], 12);
		memcpy(nrm + 3 * k, &mi->m_samples_nrm[k], 12);
		if (loc && k < mi->m_samples_loc.size())
			loc[k] = mi->m_samples_loc[k];
		if (radinfo && k < mi->m_samples_radinfo.size())
			memcpy(radinfo + 4 * k, &mi->m_samples_radinfo[k], 16);
	}
}

void rp_get_tris(ltr_Scene* S, int i, float* out9)
{
	ltr_MeshInstance* mi = S->m_meshInstances[i];
	for (size_t k = 0; k < mi->m_triTree.m_tris.size(); ++k)
		memcpy(out9 + 9 * k, &mi->m_triTree.m_tris[k], 36);
}

int rp_instance_shadow(ltr_Scene* S, int i) { return S->m_meshInstances[i]->m_shadow ? 1 : 0; }

void rp_get_lightmap(ltr_Scene* S, int i, float* rgb)
{
	ltr_MeshInstance* mi = S->m_meshInstances[i];
	for (size_t k = 0; k < mi->m_lightmap.size(); ++k)
		memcpy(rgb + 3 * k, &mi->m_lightmap[k], 12);
}

void rp_set_lightmap(ltr_Scene* S, int i, const float* rgb)
{
	ltr_MeshInstance* mi = S->m_meshInstances[i];
	for (size_t k = 0; k < mi->m_lightmap.size(); ++k)
		memcpy(&mi->m_lightmap[k], rgb + 3 * k, 12);
}

size_t rp_num_links(ltr_Scene* S) { return S->m_radLinks.size(); }
size_t rp_num_patches(ltr_Scene* S) { return S->m_radSampleGeoms.size(); }

void rp_get_links(ltr_Scene* S, uint32_t* rowmap, uint32_t* other, float* factor)
{
	for (size_t k = 0; k < S->m_radLinkMap.size(); ++k)
		rowmap[k] = S->m_radLinkMap[k];
	for (size_t k = 0; k < S->m_radLinks.size(); ++k) {
		other[k] = S->m_radLinks[k].other;
		factor[k] = S->m_radLinks[k].factor;
	}
}

// 13 floats per patch: diffuse, total, out, in, area.
void rp_get_rad_colors(ltr_Scene* S, float* out13)
{
	for (size_t k = 0; k < S->m_radSampleColors.size(); ++k)
		memcpy(out13 + 13 * k, &S->m_radSampleColors[k], 52);
}

void rp_get_rad_geoms(ltr_Scene* S, float* out6)
{
	for (size_t k = 0; k < S->m_radSampleGeoms.size(); ++k)
		memcpy(out6 + 6 * k, &S->m_radSampleGeoms[k], 24);
}

// ---- point queries (lighter.cpp:138-207, 280-289) --------------------------
float rp_distance(ltr_Scene* S, const float* p) { return S->Distance(V3P(p)); }

float rp_shadow_factor(ltr_Scene* S, const float* from, const float* to, float k)
{
	return S->CalcInvShadowFactor(V3P(from), V3P(to), k);
}

float rp_closest_hit(ltr_Scene* S, const float* a, const float* b)
{
	return S->DistanceTestBBT(V3P(a), V3P(b));
}

int rp_visibility(ltr_Scene* S, const float* a, const float* b)
{
	return S->VisibilityTest(V3P(a), V3P(b)) ? 1 : 0;
}

// Batched forms of the queries above, for timing and bulk parity.
void rp_shadow_factor_batch(ltr_Scene* S, size_t n, const float* from, const float* to, float k, float* out)
{
	for (size_t i = 0; i < n; ++i)
		out[i] = S->CalcInvShadowFactor(V3P(from + 3 * i), V3P(to + 3 * i), k);
}

void rp_closest_hit_batch(ltr_Scene* S, size_t n, const float* a, const float* b, float* out)
{
	for (size_t i = 0; i < n; ++i)
		out[i] = S->DistanceTestBBT(V3P(a + 3 * i), V3P(b + 3 * i));
}

void rp_visibility_batch(ltr_Scene* S, size_t n, const float* a, const float* b, uint8_t* out)
{
	for (size_t i = 0; i < n; ++i)
		out[i] = S->VisibilityTest(V3P(a + 3 * i), V3P(b + 3 * i)) ? 1 : 0;
}

// Free-standing primitive (lighter_math.cpp:315-354).
float rp_seg_tri(const float* l1, const float* l2, const float* p1, const float* p2, const float* p3)
{
	return IntersectLineSegmentTriangle(V3P(l1), V3P(l2), V3P(p1), V3P(p2), V3P(p3));
}

// The spiral used by AO (lighter_int.hpp:457-471).
void rp_spiral(const float* dir, float randoff, int i, int n, float* out)
{
	Vec3 v = Vec3::CreateSpiralDirVector(V3P(dir), randoff, i, n);
	out[0] = v.x; out[1] = v.y; out[2] = v.z;
}

// Triangle::CheckIsUseful (lighter_int.hpp:604-608) on one triangle, 9 floats.
int rp_tri_useful(const float* t9)
{
	Triangle T = {V3P(t9), V3P(t9 + 3), V3P(t9 + 6)};
	return T.CheckIsUseful() ? 1 : 0;
}

// Point-triangle distance as the reference's distance query sees it: a
// one-triangle TriTree queried with the initial radius MAX_PENUMBRA_SIZE.
float rp_point_tri_dist_capped(const float* p, const float* t0, const float* t1, const float* t2)
{
	Triangle T = {V3P(t0), V3P(t1), V3P(t2)};
	TriTree tt;
	tt.SetTris(&T, 1);
	return tt.GetDistance(V3P(p), 1e30f);
}

} // extern "C"
