// The bake pipeline behind ltr_Start: host stages before and after the GPU
// hot path, and the GPU stages themselves.
//
// Stage order, status strings and completion updates follow Job_MainProc
// (lighter.cpp:1046-1145).  The host stages restate the reference's
// instance transform, collision set-up, sample placement and lightmap
// finalisation (lighter.cpp:292-482, 849-1044) so that the samples the
// kernels see and the images the caller gets back are the reference's.
#include "pipeline.h"
#include "normals.h"
#include "raster_gpu.h"
#include "texel.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <functional>
#include <memory>
#include <stdexcept>
#include <thread>

namespace lgx {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
static void parallel_for(int nthreads, size_t n, const std::function<void(size_t)>& fn)
{
	if (nthreads <= 1 || n <= 1) {
		for (size_t i = 0; i < n; ++i)
			fn(i);
		return;
	}
	std::atomic<size_t> next{0};
	std::vector<std::thread> th;
	int nt = (int)std::min<size_t>((size_t)nthreads, n);
	for (int t = 0; t < nt; ++t)
		th.emplace_back([&]() {
			for (;;) {
				size_t i = next.fetch_add(1);
				if (i >= n)
					break;
				fn(i);
			}
		});
	for (auto& t : th)
		t.join();
}

static void set_stage(ltr_Scene* S, const char* st)
{
	S->stage.store(st);
	S->completion.store(0.0f);
}

static void check_abort(ltr_Scene* S)
{
	if (S->abort.load())
		throw AbortBake();
}

// ---------------------------------------------------------------------------
// triangles of a mesh part
// ---------------------------------------------------------------------------
// The reference forms a part's triangles with `tri < index_count` in steps
// of 3 over the mesh's index array from the part's first index, vertex
// indices relative to the part's first vertex (lighter.cpp:41-47, 308-317,
// 367-373).  A part whose index count is not a multiple of 3 (MeshAddPart
// refuses only counts 1 and 2, lighter.cpp:1237) therefore completes its
// last triangle with the indices that follow it in the mesh, the next
// part's.  fn(a, b, c) receives mesh vertex indices.  Where such a read
// would leave the mesh's index array, or an index the mesh's vertex arrays,
// the reference reads past the end of its vectors (undefined); that
// triangle is skipped.
template <class F> static void part_triangles(const ltr_Mesh* mesh, const MeshPart& mp, F&& fn)
{
	const size_t ni = mesh->indices.size(), nv = mesh->vpos.size();
	for (uint32_t t = 0; t < mp.icount; t += 3) {
		const size_t first = (size_t)mp.ioff + t;
		if (first + 3 > ni)
			break;
		const size_t a = (size_t)mp.voff + mesh->indices[first], b = (size_t)mp.voff + mesh->indices[first + 1],
		             c = (size_t)mp.voff + mesh->indices[first + 2];
		if (a >= nv || b >= nv || c >= nv)
			continue;
		fn((uint32_t)a, (uint32_t)b, (uint32_t)c);
	}
}

// ---------------------------------------------------------------------------
// transforming spatial data (lighter.cpp:292-341)
// ---------------------------------------------------------------------------
// World-space vertices of an instance: positions by its matrix, normals by
// the inverse transpose, renormalised.
static void world_vertices(Instance* mi)
{
	const ltr_Mesh* mesh = mi->mesh;
	const Mat4 to_world = mi->matrix, for_normals = normal_matrix(mi->matrix);
	mi->vpos.clear();
	mi->vnrm.clear();
	for (f3 p : mesh->vpos)
		mi->vpos.push_back(mat4_transform(to_world, p, 1.0f));
	for (f3 n : mesh->vnrm)
		mi->vnrm.push_back(normalized(mat4_transform(for_normals, n, 0.0f)));
}

// Atlas size from the size callback, given the instance's world-space
// surface area (triangles summed in part and index order); the config's
// default size when there is no callback or it declines.
static void atlas_size(ltr_Scene* S, Instance* mi)
{
	const ltr_Mesh* mesh = mi->mesh;
	float area = 0.0f;
	for (const MeshPart& mp : mesh->parts)
		part_triangles(mesh, mp, [&](uint32_t a, uint32_t b, uint32_t c) {
			area += triangle_area(mi->vpos[a], mi->vpos[b], mi->vpos[c]);
		});
	ltr_Config& cfg = S->config;
	u32 wh[2] = {cfg.default_width, cfg.default_height};
	const bool sized = cfg.size_fn && cfg.size_fn(&cfg, mesh->ident.c_str(), mesh->ident.size(), mi->ident.c_str(),
	                                              mi->ident.size(), area, mi->importance, wh);
	mi->lm_w = sized ? wh[0] : cfg.default_width;
	mi->lm_h = sized ? wh[1] : cfg.default_height;
}

// Per-vertex texture coordinates: the first set as given, the second in
// texel units of the atlas (texel centres at integers).
static void atlas_coords(Instance* mi)
{
	const ltr_Mesh* mesh = mi->mesh;
	const f2 texels = mk2((float)mi->lm_w, (float)mi->lm_h), centre = mk2(0.5f, 0.5f);
	mi->vtex.assign(mesh->vtex1.begin(), mesh->vtex1.end());
	mi->ltex.resize(mesh->vtex2.size());
	for (size_t i = 0; i < mesh->vtex2.size(); ++i)
		mi->ltex[i] = mesh->vtex2[i] * texels - centre;
}

static void pre_xform(ltr_Scene* S, Instance* mi)
{
	if (mi->samplecont)
		return;
	world_vertices(mi);
	atlas_size(S, mi);
	atlas_coords(mi);
}

// ---------------------------------------------------------------------------
// generating data structures (lighter.cpp:349-384, 1060-1069)
// ---------------------------------------------------------------------------
// The instance's collision set: triangles of its shadow-casting parts that
// are not degenerate (Triangle::CheckIsUseful, lighter_int.hpp:604-608).
static void col_info(ltr_Scene* S, Instance* mi)
{
	(void)S;
	if (mi->samplecont) {
		mi->tris.set(nullptr, 0);
		return;
	}
	std::vector<Tri> tris;
	const ltr_Mesh* mesh = mi->mesh;
	for (const MeshPart& mp : mesh->parts) {
		if (!mp.shadow)
			continue;
		part_triangles(mesh, mp, [&](uint32_t a, uint32_t b, uint32_t c) {
			const Tri T{mi->vpos[a], mi->vpos[b], mi->vpos[c]};
			if (tri_useful(T))
				tris.push_back(T);
		});
	}
	if (!tris.empty())
		mi->split.set(tris.data(), tris.size());
	mi->tris.set(tris.data(), tris.size());
}

// Scene-level closest hit through the split trees (lighter.cpp:210-251).
static float scene_split_hit(ltr_Scene* S, f3 A, f3 B, f3* outnormal)
{
	f3 dn = normalized(B - A);
	f3 mA = A + dn * kSmall;
	f3 mB = B - dn * kSmall;
	RaySeg r;
	r.set(mA, mB);
	float closest = 2.0f;
	auto fn = [&](const int32_t* ids, int32_t n) {
		for (int32_t i = 0; i < n; ++i) {
			Instance* mi = S->insts[ids[i]];
			if (!mi->shadow)
				continue;
			// the tree hands over every instance of a node the segment
			// enters; one whose own box (that of the triangles its split
			// tree holds) misses the segment's padded box cannot be hit
			const BoxTree::Node& root = mi->tris.tree().root();
			if (box_disjoint(r.qmn, r.qmx, root.mn, root.mx))
				continue;
			f3 nrm = mk3(0, 0, 0);
			float d = mi->split.intersect(r.origin, r.end, &nrm);
			if (d < closest) {
				closest = d;
				if (outnormal)
					*outnormal = nrm;
			}
		}
		return true;
	};
	S->inst_tree.ray_query(r, fn);
	return closest;
}

// ---------------------------------------------------------------------------
// generating samples (lighter.cpp:32-66, 392-476)
// ---------------------------------------------------------------------------
// The instance's triangles as the rasteriser draws them, in part and index
// order: texel-space corners, world positions and normals, and
// (uv0, part, texel area) per corner.
// The drawing-order triangles of one instance, each given to emit (a
// RasterTri); returns their count.
template <class Emit>
static size_t raster_tris_each(Instance* mi, Emit&& emit)
{
	ltr_Mesh* mesh = mi->mesh;
	size_t n = 0;
	for (uint32_t part = 0; part < mesh->parts.size(); ++part) {
		part_triangles(mesh, mesh->parts[part], [&](uint32_t a, uint32_t b, uint32_t c) {
			const float area = sample_area(mi->ltex[a], mi->ltex[b], mi->ltex[c], mi->vpos[a], mi->vpos[b],
			                               mi->vpos[c]);
			const f4 ia{mi->vtex[a].x, mi->vtex[a].y, (float)part, area};
			const f4 ib{mi->vtex[b].x, mi->vtex[b].y, (float)part, area};
			const f4 ic{mi->vtex[c].x, mi->vtex[c].y, (float)part, area};
			emit(n++, raster_setup(mi->ltex[a], mi->ltex[b], mi->ltex[c], mi->vpos[a], mi->vpos[b], mi->vpos[c],
			                       mi->vnrm[a], mi->vnrm[b], mi->vnrm[c], ia, ib, ic));
		});
	}
	return n;
}

static size_t count_raster_tris(Instance* mi)
{
	size_t n = 0;
	for (const MeshPart& mp : mi->mesh->parts)
		part_triangles(mi->mesh, mp, [&](uint32_t, uint32_t, uint32_t) { ++n; });
	return n;
}

static void raster_tris(Instance* mi, std::vector<RasterTri>& out)
{
	raster_tris_each(mi, [&](size_t, const RasterTri& t) { out.push_back(t); });
}

// Rasterised texels of one instance with a nonzero normal, in texel order:
// position, normalised normal, (uv0, part, area) and the texel index.
struct RasterSamples {
	std::vector<f3> pos, nrm;
	std::vector<f4> info;
	std::vector<uint32_t> loc;
	std::vector<float> dist; // square root of the world area per texel
};

// The host rasteriser: both passes of raster.h's rule in drawing order over
// the instance's atlas, then the texels with a nonzero normal in texel order.
static void raster_samples_host(Instance* mi, RasterSamples& out)
{
	std::vector<RasterTri> tris;
	raster_tris(mi, tris);
	const int32_t w = (int32_t)mi->lm_w, h = (int32_t)mi->lm_h;
	const size_t n = (size_t)mi->lm_w * mi->lm_h;
	std::vector<f3> pos(n, mk3(0, 0, 0)), nrm(n, mk3(0, 0, 0));
	std::vector<f4> inf(n, f4{0, 0, 0, 0});
	for (const float margin : {kRasterMargin0, 0.0f})
		for (const RasterTri& t : tris) {
			PixelBox b;
			if (!raster_box(t, margin, w, h, &b))
				continue;
			for (int32_t y = b.y0; y <= b.y1; ++y)
				for (int32_t x = b.x0; x <= b.x1; ++x)
					if (raster_covers(t, margin, x, y)) {
						const size_t o = (size_t)x + (size_t)w * (size_t)y;
						raster_values(t, x, y, &pos[o], &nrm[o], &inf[o]);
					}
		}
	for (size_t i = 0; i < n; ++i) {
		if (is_zero(nrm[i]))
			continue;
		out.pos.push_back(pos[i]);
		out.nrm.push_back(normalized(nrm[i]));
		out.info.push_back(inf[i]);
		out.loc.push_back((uint32_t)i);
		out.dist.push_back(sqrtf(inf[i].w));
	}
}

// Every instance rasterised in one set of GPU launches (raster.hip); the
// same samples as raster_samples_host's.
static void raster_samples_gpu(ltr_Scene* S, std::vector<RasterSamples>& raw)
{
	std::vector<size_t> ids;
	for (size_t i = 0; i < S->insts.size(); ++i)
		if (!S->insts[i]->samplecont)
			ids.push_back(i);
	// every instance's triangles in one array, each instance filling its own
	// range in parallel
	std::vector<size_t> first(ids.size() + 1, 0);
	std::vector<size_t> cnt(ids.size());
	parallel_for(S->nthreads, ids.size(), [&](size_t q) { cnt[q] = count_raster_tris(S->insts[ids[q]]); });
	for (size_t q = 0; q < ids.size(); ++q)
		first[q + 1] = first[q] + cnt[q];
	// left uninitialised: every element is written below
	std::unique_ptr<RasterTri[]> tris(new RasterTri[first.back() + 1]);
	std::unique_ptr<uint32_t[]> tri_atlas(new uint32_t[first.back() + 1]);
	std::vector<RasterAtlasSize> sizes;
	for (size_t q = 0; q < ids.size(); ++q)
		sizes.push_back(RasterAtlasSize{S->insts[ids[q]]->lm_w, S->insts[ids[q]]->lm_h});
	std::atomic<bool> miscount{false};
	parallel_for(S->nthreads, ids.size(), [&](size_t q) {
		RasterTri* dst = tris.get() + first[q];
		const size_t n = raster_tris_each(S->insts[ids[q]], [&](size_t k, const RasterTri& t) {
			if (k < cnt[q])
				dst[k] = t;
		});
		if (n != cnt[q])
			miscount.store(true);
		std::fill(tri_atlas.get() + first[q], tri_atlas.get() + first[q + 1], (uint32_t)q);
	});
	if (miscount.load())
		throw std::runtime_error("rasterisation: an instance's triangle count changed between two passes");
	if (ids.empty())
		return;
	require_gpu();
	RasterOut out;
	rasterize_gpu(tris.get(), tri_atlas.get(), first.back(), sizes, out, &S->raster_times);
	parallel_for(S->nthreads, ids.size(), [&](size_t q) {
		const uint64_t f = out.first[q], c = out.count[q];
		RasterSamples& r = raw[ids[q]];
		r.pos.resize(c);
		r.nrm.resize(c);
		r.info.resize(c);
		r.loc.assign(out.loc.get() + f, out.loc.get() + f + c);
		r.dist.resize(c);
		for (size_t i = 0; i < c; ++i) {
			const size_t k = f + i;
			r.pos[i] = ld3(&out.pos[3 * k]);
			r.nrm[i] = ld3(&out.nrm[3 * k]);
			r.info[i] = f4{out.inf[4 * k], out.inf[4 * k + 1], out.inf[4 * k + 2], out.inf[4 * k + 3]};
			r.dist[i] = sqrtf(r.info[i].w);
		}
	});
}

// The host restatement of the placement of one instance's samples
// (lighter.cpp:436-466); the GPU placement (place.hip) gives the same bits.
static void place_host(ltr_Scene* S, RasterSamples& r)
{
	const ltr_Config& cfg = S->config;
	float corr_min_dot = cosf(cfg.max_correct_angle / 180.0f * kPiF);
	std::vector<Box> roots(S->insts.size());
	for (size_t k = 0; k < S->insts.size(); ++k) {
		const BoxTree::Node& rt = S->insts[k]->tris.tree().root();
		roots[k] = Box{rt.mn, rt.mx};
	}
	for (size_t i = 0; i < r.pos.size(); ++i) {
		f3 P = r.pos[i];
		const f3 N = r.nrm[i];
		// move the sample out of concave edges; an instance whose tree root is
		// disjoint from the query cube contributes nothing (its query stops at
		// the root), so that test is made here first, on the current P
		const float dist = r.dist[i];
		for (size_t k = 0; k < S->insts.size(); ++k) {
			const f3 d3 = mk3(dist, dist, dist);
			if (box_disjoint(P - d3, P + d3, roots[k].mn, roots[k].mx))
				continue;
			S->insts[k]->tris.offset_sample(P, N, dist);
		}
		// move the sample in front of overlapping faces
		if (cfg.max_correct_dist) {
			int left = 100;
			float md = cfg.max_correct_dist;
			f3 PEnd = P + N * md;
			while (md > kSmall && left-- > 0) {
				f3 hn = mk3(0, 0, 0);
				float q = scene_split_hit(S, P, PEnd, &hn);
				if (!correction_step(P, PEnd, N, md, q, hn, corr_min_dot))
					break;
			}
		}
		r.pos[i] = P;
	}
}

// Every instance's rasterised samples placed in one GPU launch.
static void place_gpu(ltr_Scene* S, std::vector<RasterSamples>& raw)
{
	size_t total = 0;
	for (const RasterSamples& r : raw)
		total += r.pos.size();
	S->place_times = PlaceTimes{};
	if (total == 0)
		return;
	require_gpu();
	std::vector<PlaceSamples> runs;
	for (RasterSamples& r : raw)
		runs.push_back(PlaceSamples{r.pos.data(), r.nrm.data(), r.dist.data(), r.pos.size()});
	std::vector<PlaceInstance> pi;
	for (Instance* mi : S->insts)
		pi.push_back(PlaceInstance{&mi->tris, &mi->split, mi->shadow});
	const ltr_Config& cfg = S->config;
	float corr_min_dot = cosf(cfg.max_correct_angle / 180.0f * kPiF);
	place_samples_gpu(pi, S->inst_tree, runs, cfg.max_correct_dist, corr_min_dot, S->nthreads, &S->place_times);
}

static void commit_samples(ltr_Scene* S, Instance* mi, RasterSamples& r)
{
	mi->s_pos = std::move(r.pos);
	mi->s_nrm = std::move(r.nrm);
	mi->s_loc = std::move(r.loc);
	mi->s_info = std::move(r.info);
	mi->lightmap.assign(mi->s_pos.size(), ld3(S->config.ambient_color));
}

// generating samples (lighter.cpp:392-476): rasterisation and compaction and
// placement on the GPU (or, when selected, their host restatements)
static void gen_samples(ltr_Scene* S)
{
	const int nt = S->nthreads;
	std::vector<RasterSamples> raw(S->insts.size());
	for (Instance* mi : S->insts)
		if (mi->samplecont)
			for (auto& c : mi->lightmap)
				c = ld3(S->config.ambient_color);
	S->raster_times = RasterTimes{};
	using clk = std::chrono::steady_clock;
	auto t0 = clk::now();
	if (S->host_raster)
		parallel_for(nt, S->insts.size(), [&](size_t i) {
			if (!S->insts[i]->samplecont)
				raster_samples_host(S->insts[i], raw[i]);
		});
	else
		raster_samples_gpu(S, raw);
	auto t1 = clk::now();
	check_abort(S);
	if (S->host_place)
		parallel_for(nt, S->insts.size(), [&](size_t i) { place_host(S, raw[i]); });
	else
		place_gpu(S, raw);
	auto t2 = clk::now();
	parallel_for(nt, S->insts.size(), [&](size_t i) {
		if (!S->insts[i]->samplecont)
			commit_samples(S, S->insts[i], raw[i]);
	});
	auto t3 = clk::now();
	S->prep_ms[4] = std::chrono::duration<double, std::milli>(t1 - t0).count();
	S->prep_ms[5] = std::chrono::duration<double, std::milli>(t2 - t1).count();
	S->prep_ms[6] = std::chrono::duration<double, std::milli>(t3 - t2).count();
}

// Per-bake state back to that of a scene that has not been baked: the next
// preparation runs every host stage again (the reference's ltr_Start re-runs
// them too, lighter.cpp:1046-1073), so lightmaps restart from the ambient
// colour and no normal-map contribution carries over.
void reset_bake(ltr_Scene* S)
{
	S->dev.reset();
	S->prepared = false;
	for (Instance* mi : S->insts)
		mi->contrib.clear();
}

void prepare(ltr_Scene* S)
{
	if (S->prepared)
		return;
	int nt = S->nthreads;
	using clk = std::chrono::steady_clock;
	auto ms_since = [](clk::time_point t0) {
		return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
	};
	auto t0 = clk::now();
	set_stage(S, "transforming spatial data");
	parallel_for(nt, S->insts.size(), [&](size_t i) { pre_xform(S, S->insts[i]); });
	S->prep_ms[0] = ms_since(t0);
	check_abort(S);

	t0 = clk::now();
	set_stage(S, "generating data structures");
	parallel_for(nt, S->insts.size(), [&](size_t i) { col_info(S, S->insts[i]); });
	S->prep_ms[1] = ms_since(t0);
	S->completion.store(0.99f);
	t0 = clk::now();
	std::vector<Box> bbs;
	for (Instance* mi : S->insts) {
		const BoxTree::Node& r = mi->tris.tree().root();
		bbs.push_back(Box{r.mn, r.mx});
	}
	S->inst_tree.build(bbs.data(), bbs.size());
	S->prep_ms[2] = ms_since(t0);
	check_abort(S);

	t0 = clk::now();
	set_stage(S, "generating samples");
	gen_samples(S);
	S->prep_ms[3] = ms_since(t0);
	check_abort(S);

	S->inst_off.assign(S->insts.size() + 1, 0);
	for (size_t i = 0; i < S->insts.size(); ++i)
		S->inst_off[i + 1] = S->inst_off[i] + S->insts[i]->lightmap.size();
	S->total_samples = S->inst_off.back();
	S->prepared = true;
}

// ---------------------------------------------------------------------------
// GPU scene
// ---------------------------------------------------------------------------
void ensure_device(ltr_Scene* S)
{
	if (S->dev)
		return;
	std::unique_ptr<DeviceScene> d(new DeviceScene());
	std::vector<float> tris;
	for (Instance* mi : S->insts) {
		if (!mi->shadow)
			continue;
		for (const Tri& t : mi->tris.tris()) {
			const float v[9] = {t.p1.x, t.p1.y, t.p1.z, t.p2.x, t.p2.y, t.p2.z, t.p3.x, t.p3.y, t.p3.z};
			tris.insert(tris.end(), v, v + 9);
		}
	}
	d->set_triangles(tris.data(), tris.size() / 9);
	std::vector<float> pos(3 * S->total_samples), nrm(3 * S->total_samples), lm(3 * S->total_samples);
	size_t k = 0;
	for (Instance* mi : S->insts) {
		for (size_t i = 0; i < mi->lightmap.size(); ++i, ++k) {
			st3(&pos[3 * k], mi->s_pos[i]);
			st3(&nrm[3 * k], mi->s_nrm[i]);
			st3(&lm[3 * k], mi->lightmap[i]);
		}
	}
	d->set_samples(pos.data(), nrm.data(), S->total_samples);
	d->set_lightmap(lm.data());
	S->dev = std::move(d);
}

void download_lightmaps(ltr_Scene* S)
{
	std::vector<float> lm(3 * S->total_samples);
	S->dev->get_lightmap(lm.data());
	size_t k = 0;
	for (Instance* mi : S->insts)
		for (size_t i = 0; i < mi->lightmap.size(); ++i, ++k)
			mi->lightmap[i] = ld3(&lm[3 * k]);
}

void upload_lightmaps(ltr_Scene* S)
{
	std::vector<float> lm(3 * S->total_samples);
	size_t k = 0;
	for (Instance* mi : S->insts)
		for (size_t i = 0; i < mi->lightmap.size(); ++i, ++k)
			st3(&lm[3 * k], mi->lightmap[i]);
	S->dev->set_lightmap(lm.data());
}

// Instances a light renders into, in render order (lighter.cpp:80-98,
// 1080-1093): the sample container plus the instance-tree query, sorted and
// de-duplicated.
std::vector<int32_t> light_instances(ltr_Scene* S, const Light& L)
{
	std::vector<int32_t> ids;
	ids.push_back(0);
	auto add = [&](const int32_t* p, int32_t n) { ids.insert(ids.end(), p, p + n); };
	if (L.type == LTR_LT_POINT || L.type == LTR_LT_SPOT) {
		f3 r3 = mk3(L.range, L.range, L.range);
		S->inst_tree.query(L.pos - r3, L.pos + r3, add);
	} else if (L.type == LTR_LT_DIRECT) {
		S->inst_tree.all(add);
	}
	std::sort(ids.begin(), ids.end());
	ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
	return ids;
}

DevLight to_dev_light(const Light& L)
{
	DevLight d{};
	d.type = (int32_t)L.type;
	st3(d.pos, L.pos);
	st3(d.dir, L.dir);
	st3(d.color, L.color);
	d.range = L.range;
	d.power = L.power;
	d.radius = L.radius;
	// Int_LMRender spot set-up (lighter.cpp:641-644)
	float out_rad = L.spot_out / 180.0f * kPiF, in_rad = L.spot_in / 180.0f * kPiF;
	if (in_rad == out_rad)
		in_rad -= kSmall;
	d.spot_out_rad = out_rad;
	d.spot_diff = in_rad - out_rad;
	d.spot_curve = L.spot_curve;
	return d;
}

static float brightness(f3 c) { return (c.x + c.y + c.z) * (1.0f / 3.0f); }

// Normal-map light contributions of one light (lighter.cpp:505-513,
// 543-551, 590-598, 656-664), from the shadow factors the GPU computed:
// the host restatement of the GPU pass (nm_contrib, normals.h).
static void collect_contribs(ltr_Scene* S, const Light& L, const std::vector<int32_t>& ids, const float* vis)
{
	const DevLight D = to_dev_light(L);
	size_t k = 0;
	for (int32_t id : ids) {
		Instance* mi = S->insts[id];
		for (size_t i = 0; i < mi->lightmap.size(); ++i, ++k) {
			f3 c;
			if (!mi->samplecont && nm_contrib(D, mi->s_pos[i], mi->s_nrm[i], vis[k], &c))
				mi->contrib.push_back(LightContrib{c, (int32_t)i});
		}
	}
}

// ---------------------------------------------------------------------------
// rendering lightmaps (lighter.cpp:1075-1096)
// ---------------------------------------------------------------------------
void stage_direct(ltr_Scene* S)
{
	ensure_device(S);
	const bool want_vis = S->config.generate_normalmap_data != 0;
	S->dev->nm_clear();
	S->nm_data.clear();
	S->times.direct_ms = S->times.direct_march_ms = 0;
	for (size_t l0 = 0; l0 < S->lights.size(); l0 += 64) {
		size_t l1 = std::min(S->lights.size(), l0 + 64);
		std::vector<DevLight> dl;
		std::vector<uint32_t> off;
		std::vector<std::vector<int32_t>> ids_of;
		off.push_back(0);
		for (size_t l = l0; l < l1; ++l) {
			const Light& L = S->lights[l];
			dl.push_back(to_dev_light(L));
			std::vector<int32_t> ids = light_instances(S, L);
			size_t n = off.back();
			for (int32_t id : ids)
				n += S->inst_off[id + 1] - S->inst_off[id];
			if (n >= ((size_t)1 << 32))
				throw std::runtime_error("direct light: more than 2^32 work items in one batch of lights");
			off.push_back((uint32_t)n);
			ids_of.push_back(std::move(ids));
		}
		// (light << 26) | sample for every sample of every instance the
		// light reaches, in light order and instance order; one light per
		// task
		std::vector<uint32_t> items(off.back());
		parallel_for(S->nthreads, l1 - l0, [&](size_t q) {
			uint32_t* dst = items.data() + off[q];
			for (int32_t id : ids_of[q])
				for (size_t s = S->inst_off[id]; s < S->inst_off[id + 1]; ++s)
					*dst++ = (uint32_t)(q << 26) | (uint32_t)s;
		});
		// progress and abort while the march kernel runs
		const float n_lights = (float)S->lights.size();
		const bool done = S->dev->direct_light(dl, items, off, want_vis, [&](float f) {
			S->completion.store(((float)l0 + f * (float)(l1 - l0)) / n_lights);
			return !S->abort.load();
		});
		if (!done)
			throw AbortBake{};
		S->times.direct_ms += S->dev->times.direct_ms;
		S->times.direct_march_ms += S->dev->times.direct_march_ms;
		if (want_vis && S->host_finalize) {
			std::vector<float> vis(items.size());
			S->dev->get_vis(vis.data(), vis.size());
			for (size_t l = l0; l < l1; ++l)
				collect_contribs(S, S->lights[l], ids_of[l - l0], vis.data() + off[l - l0]);
		} else if (want_vis) {
			// the GPU export's normal-map pass reads these shadow factors:
			// per light, each instance's first item (or -1)
			const size_t ni = S->insts.size();
			std::vector<int32_t> first((l1 - l0) * ni, -1);
			for (size_t l = l0; l < l1; ++l) {
				uint32_t at = off[l - l0];
				for (int32_t id : ids_of[l - l0]) {
					first[(l - l0) * ni + id] = (int32_t)at;
					at += (uint32_t)(S->inst_off[id + 1] - S->inst_off[id]);
				}
			}
			S->dev->nm_keep_batch(dl, first, ni);
		}
		S->completion.store((float)l1 / (float)S->lights.size());
		check_abort(S);
	}
}

// ---------------------------------------------------------------------------
// radiosity (lighter.cpp:667-765, 1098-1124)
// ---------------------------------------------------------------------------
void stage_radiosity(ltr_Scene* S)
{
	ensure_device(S);
	size_t N = S->total_samples;
	if (N >= ((size_t)1 << 32))
		throw std::runtime_error("radiosity: patch indices are 32-bit (" + std::to_string(N) + " patches)");
	download_lightmaps(S);
	std::vector<float> pos(3 * N), nrm(3 * N), diff(3 * N), emit(3 * N), area(N);
	const ltr_Config& cfg = S->config;
	// each instance's patches at its samples' offsets; in parallel unless
	// the caller's sample_fn is called, which the reference calls from one
	// thread (its thread safety is not promised)
	auto patches_of = [&](size_t m) {
		Instance* mi = S->insts[m];
		size_t k = S->inst_off[m];
		for (size_t i = 0; i < mi->lightmap.size(); ++i, ++k) {
			f3 P = mi->s_pos[i];
			f3 Nn = normalized(mi->s_nrm[i]);
			st3(&pos[3 * k], P);
			st3(&nrm[3 * k], Nn);
			f3 d = mk3(1, 1, 1);
			f3 e = mi->lightmap[i];
			float a = mi->samplecont ? 0 : mi->s_info[i].w;
			if (mi->samplecont) {
				d = mk3(0, 0, 0);
			} else if (cfg.sample_fn) {
				uint32_t L = mi->s_loc[i];
				int lx = (int)(L % mi->lm_w), ly = (int)(L / mi->lm_w);
				ltr_SampleRequest req;
				memset(&req, 0, sizeof(req));
				st3(req.position, P);
				st3(req.normal, Nn);
				req.tex0u = mi->s_info[i].x;
				req.tex0v = mi->s_info[i].y;
				req.tex1u = (lx + 0.5f) / mi->lm_w;
				req.tex1v = (ly + 0.5f) / mi->lm_h;
				req.part_id = (uint32_t)mi->s_info[i].z;
				req.mesh_ident = mi->mesh->ident.c_str();
				req.mesh_ident_size = mi->mesh->ident.size();
				req.inst_ident = mi->ident.c_str();
				req.inst_ident_size = mi->ident.size();
				LTR_VEC3_SET(req.out_diffuse_color, 1, 1, 1);
				LTR_VEC3_SET(req.out_emissive_color, 0, 0, 0);
				if (cfg.sample_fn(&S->config, &req)) {
					d = ld3(req.out_diffuse_color);
					e += ld3(req.out_emissive_color);
				}
			}
			st3(&diff[3 * k], d);
			st3(&emit[3 * k], e);
			area[k] = a;
		}
	};
	parallel_for(cfg.sample_fn ? 1 : S->nthreads, S->insts.size(), patches_of);
	check_abort(S);
	DeviceScene& D = *S->dev;
	D.rad_set_patches(pos.data(), nrm.data(), N);
	D.rad_links();
	S->times.links_ms = D.times.links_ms;
	check_abort(S);

	set_stage(S, "bouncing light");
	D.rad_set_colors(diff.data(), emit.data(), area.data());
	S->times.bounce_ms = 0;
	for (int b = 0; b < cfg.bounce_count; ++b) {
		D.rad_bounce(1);
		S->times.bounce_ms += D.times.bounce_ms;
		S->completion.store((b + 1) / (float)cfg.bounce_count);
	}
	check_abort(S);

	set_stage(S, "committing radiosity");
	std::vector<float> total(3 * N);
	D.rad_get_total(total.data());
	parallel_for(S->nthreads, S->insts.size(), [&](size_t m) {
		Instance* mi = S->insts[m];
		for (size_t i = 0, k = S->inst_off[m]; i < mi->lightmap.size(); ++i, ++k)
			mi->lightmap[i] = ld3(&total[3 * k]);
	});
	S->completion.store(1.0f);
	upload_lightmaps(S);
}

// ---------------------------------------------------------------------------
// rendering ambient occlusion (lighter.cpp:799-841, 1126-1136)
// ---------------------------------------------------------------------------
void stage_ao(ltr_Scene* S)
{
	ensure_device(S);
	const ltr_Config& cfg = S->config;
	// One rotation offset per texel, drawn from the C library generator in
	// instance order and sample order, as a single-threaded reference bake
	// draws them (lighter.cpp:819, lighter_int.hpp:271).
	std::vector<float> ro(S->total_samples);
	for (size_t i = 0; i < ro.size(); ++i)
		ro[i] = (float)rand() / (float)RAND_MAX;
	// chunks of texels: progress after each, and an abort stops between them
	S->dev->ambient_occlusion(ro.data(), cfg.ao_num_samples, cfg.ao_distance, cfg.ao_multiplier, cfg.ao_falloff,
	                          cfg.ao_effect, cfg.ao_color_rgb, nullptr, [S](float done) {
		                          S->completion.store(done);
		                          return !S->abort.load();
	                          });
	S->times.ao_ms = S->dev->times.ao_ms;
	check_abort(S);
	S->completion.store(1.0f);
}

// ---------------------------------------------------------------------------
// exporting lightmaps (lighter.cpp:849-1044, lighter_math.cpp:357-424)
// ---------------------------------------------------------------------------
// The image of one instance on the host (lighter.cpp:866-958), texel by
// texel through the functions the GPU export runs (texel.h): the samples'
// colours in their texels, three rounds of extrapolation, the separable
// blur, the 2x downsample.  The exported size goes to out_w/out_h; lm_w/lm_h
// keep the atlas size the samples' texel indices refer to.  Instances run in
// parallel.
static float* image_host(ltr_Scene* S, size_t mid)
{
	Instance* mi = S->insts[mid];
	const ltr_Config& cfg = S->config;
	uint32_t w = mi->lm_w, h = mi->lm_h;
	const size_t npx = (size_t)w * h;
	for (uint32_t loc : mi->s_loc)
		if (loc >= npx)
			throw std::runtime_error("export: a sample's texel index lies outside its instance's atlas");
	std::vector<float> img(3 * npx, 0.0f);
	std::vector<uint8_t> flag(npx, 0);
	for (size_t i = 0; i < mi->s_loc.size(); ++i) {
		st3(&img[3 * (size_t)mi->s_loc[i]], mi->lightmap[i]);
		flag[mi->s_loc[i]] = 1;
	}
	// each round reads the flags it started with and sets the new ones apart
	for (int round = 0; round < 3; ++round) {
		std::vector<uint8_t> grown(flag);
		const AtlasView view{img.data(), flag.data(), w, h};
		for (uint32_t y = 0; y < h; ++y)
			for (uint32_t x = 0; x < w; ++x) {
				const size_t t = x + (size_t)y * w;
				f3 c;
				if (!flag[t] && extrapolate_texel(view, x, y, &c)) {
					st3(&img[3 * t], c);
					grown[t] = 1;
				}
			}
		flag.swap(grown);
	}
	if (cfg.blur_size) {
		const int ext = (int)ceilf(cfg.blur_size);
		const std::vector<float> taps = blur_taps(ext, cfg.blur_size);
		std::vector<float> across(3 * npx);
		// rows of w texels into columns, then rows of h texels back
		for (uint32_t y = 0; y < h; ++y)
			for (uint32_t x = 0; x < w; ++x)
				st3(&across[3 * (y + (size_t)h * x)], blur_texel(&img[3 * (size_t)y * w], w, (int)x, taps.data(), ext));
		for (uint32_t x = 0; x < w; ++x)
			for (uint32_t y = 0; y < h; ++y)
				st3(&img[3 * (x + (size_t)w * y)], blur_texel(&across[3 * (size_t)x * h], h, (int)y, taps.data(), ext));
	}
	if (cfg.ds2x) {
		const uint32_t hw = std::max(w / 2, 1u), hh = std::max(h / 2, 1u);
		std::vector<float> half(3 * (size_t)hw * hh);
		for (uint32_t y = 0; y < hh; ++y)
			for (uint32_t x = 0; x < hw; ++x)
				st3(&half[3 * (x + (size_t)y * hw)], half_texel(img.data(), w, h, x, y));
		img.swap(half);
		w = hw;
		h = hh;
	}
	mi->out_w = w;
	mi->out_h = h;
	float* rgb = new float[img.size()];
	memcpy(rgb, img.data(), sizeof(float) * img.size());
	return rgb;
}

// Normal-map data (lighter.cpp:959-1030) and the output record of one
// instance around its finished image.
static ltr_WorkOutput output_of(ltr_Scene* S, size_t mid, float* rgb)
{
	Instance* mi = S->insts[mid];
	const ltr_Config& cfg = S->config;
	float* normals = nullptr;
	if (cfg.generate_normalmap_data && !S->nm_data.empty()) {
		// from the GPU pass: normal and min-dot term per sample
		const size_t nn = (size_t)mi->out_w * mi->out_h, base = S->inst_off[mid];
		normals = new float[nn * 4]();
		for (size_t i = 0; i < mi->s_loc.size(); ++i) {
			const size_t loc = (size_t)mi->s_loc[i];
			if (loc >= nn)
				continue; // after ds2x the full-resolution texel index is out of range
			memcpy(normals + loc * 4, &S->nm_data[4 * (base + i)], 4 * sizeof(float));
		}
	} else if (cfg.generate_normalmap_data) {
		struct Sum {
			f3 normal;
			float mindot;
			int32_t count;
		};
		std::vector<Sum> cs(mi->s_loc.size(), Sum{mk3(0, 0, 0), 1, 1});
		float ab = brightness(ld3(cfg.ambient_color));
		for (size_t i = 0; i < mi->s_nrm.size(); ++i)
			cs[i].normal = mi->s_nrm[i] * ab;
		for (const LightContrib& c : mi->contrib) {
			cs[c.sid].normal += c.normal;
			cs[c.sid].count++;
		}
		for (Sum& c : cs)
			c.normal = c.normal / (float)c.count;
		for (const LightContrib& c : mi->contrib) {
			Sum& t = cs[c.sid];
			t.mindot *= nm_mindot_factor(t.normal, c.normal);
		}
		size_t nn = (size_t)mi->out_w * mi->out_h;
		normals = new float[nn * 4]();
		for (size_t i = 0; i < mi->s_loc.size(); ++i) {
			size_t loc = (size_t)mi->s_loc[i];
			if (loc >= nn)
				continue; // after ds2x the full-resolution texel index is out of range
			f3 N = normalized(cs[i].normal);
			normals[loc * 4 + 0] = N.x;
			normals[loc * 4 + 1] = N.y;
			normals[loc * 4 + 2] = N.z;
			normals[loc * 4 + 3] = cs[i].mindot;
		}
	}
	ltr_WorkOutput wo;
	wo.uid = (u32)mid;
	wo.mesh_ident = mi->mesh->ident.c_str();
	wo.mesh_ident_size = mi->mesh->ident.size();
	wo.inst_ident = mi->ident.c_str();
	wo.inst_ident_size = mi->ident.size();
	wo.lightmap_rgb = rgb;
	wo.normals_xyzf = normals;
	wo.width = mi->out_w;
	wo.height = mi->out_h;
	return wo;
}

// The images of every instance on the GPU (finalize.hip), identical to
// image_host's.
static void images_gpu(ltr_Scene* S, std::vector<float*>& rgb)
{
	const ltr_Config& cfg = S->config;
	std::vector<FinImage> imgs;
	std::vector<size_t> ids;
	for (size_t mid = 0; mid < S->insts.size(); ++mid) {
		Instance* mi = S->insts[mid];
		if (mi->samplecont)
			continue;
		uint32_t fw = mi->lm_w, fh = mi->lm_h;
		if (cfg.ds2x) {
			fw = std::max(mi->lm_w / 2, 1u);
			fh = std::max(mi->lm_h / 2, 1u);
		}
		rgb[mid] = new float[(size_t)fw * fh * 3];
		imgs.push_back(FinImage{mi->lm_w, mi->lm_h, mi->s_loc.size(), (const float*)mi->lightmap.data(),
		                        mi->s_loc.data(), rgb[mid], fw, fh});
		ids.push_back(mid);
	}
	S->fin_times = FinTimes{};
	if (imgs.empty())
		return;
	require_gpu();
	std::vector<float> kernel;
	if (cfg.blur_size)
		kernel = blur_taps((int)ceilf(cfg.blur_size), cfg.blur_size);
	// the colours stay on the device when a GPU stage left them there: the
	// images' samples follow the sample container's in the device order
	const float* d_rgb = nullptr;
	if (S->dev && S->dev->num_samples() == S->total_samples && S->inst_off.size() > 1)
		d_rgb = S->dev->lightmap_device() + 3 * S->inst_off[1];
	finalize_gpu(imgs, d_rgb, cfg.blur_size, kernel, cfg.ds2x != 0, S->nthreads, &S->fin_times);
	for (size_t q = 0; q < ids.size(); ++q) {
		S->insts[ids[q]]->out_w = imgs[q].final_w;
		S->insts[ids[q]]->out_h = imgs[q].final_h;
	}
}

void finalize(ltr_Scene* S)
{
	// an export replaces the output list; the images of an earlier one stay
	// allocated (the caller may hold them) until the scene is destroyed
	S->retired.insert(S->retired.end(), S->outputs.begin(), S->outputs.end());
	S->outputs.clear();
	const size_t n = S->insts.size();
	std::vector<ltr_WorkOutput> outs(n);
	for (size_t mid = 0; mid < n; ++mid) {
		Instance* mi = S->insts[mid];
		if (mi->samplecont)
			for (size_t i = 0; i < mi->lightmap.size(); ++i)
				st3(S->samples[i].out_color, mi->lightmap[i]);
	}
	// normal-map data of the GPU pass, when the direct stage kept its
	// shadow factors on the device (else the host's contributions serve)
	S->nm_data.clear();
	if (S->config.generate_normalmap_data && !S->host_finalize && S->dev && S->dev->nm_lights() > 0) {
		std::vector<uint32_t> off(S->inst_off.begin(), S->inst_off.end());
		S->nm_data.resize(4 * S->total_samples);
		S->dev->nm_compute(off, brightness(ld3(S->config.ambient_color)), S->nm_data.data());
	}
	std::vector<float*> rgb(n, nullptr);
	if (S->host_finalize) {
		S->fin_times = FinTimes{};
		parallel_for(S->nthreads, n, [&](size_t mid) {
			if (!S->insts[mid]->samplecont)
				rgb[mid] = image_host(S, mid);
		});
	} else {
		try {
			images_gpu(S, rgb);
		} catch (...) {
			for (float* p : rgb)
				delete[] p;
			throw;
		}
	}
	parallel_for(S->nthreads, n, [&](size_t mid) {
		if (!S->insts[mid]->samplecont)
			outs[mid] = output_of(S, mid, rgb[mid]);
	});
	// outputs in instance order (uid = instance index), as the reference's
	for (size_t mid = 0; mid < n; ++mid)
		if (!S->insts[mid]->samplecont)
			S->outputs.push_back(outs[mid]);
}

// ---------------------------------------------------------------------------
void bake(ltr_Scene* S)
{
	try {
		if (!S->insts.empty()) {
			prepare(S);
			bool gpu = !S->lights.empty() || S->config.bounce_count || S->config.ao_distance;
			if (gpu)
				ensure_device(S);
			if (!S->lights.empty()) {
				set_stage(S, "rendering lightmaps");
				stage_direct(S);
			}
			if (S->config.bounce_count) {
				set_stage(S, "calculating radiosity");
				stage_radiosity(S);
			}
			if (S->config.ao_distance) {
				set_stage(S, "rendering ambient occlusion");
				stage_ao(S);
			}
			if (gpu)
				download_lightmaps(S);
			set_stage(S, "exporting lightmaps");
			finalize(S);
		}
		S->dev.reset();
		S->completion.store(1.0f);
		S->stage.store(nullptr);
	} catch (const AbortBake&) {
		S->dev.reset();
		S->completion.store(1.0f);
		S->stage.store(nullptr);
	} catch (const std::exception& e) {
		S->error = e.what();
		fprintf(stderr, "lighter_amd: bake failed: %s\n", e.what());
		S->dev.reset();
		S->failed.store(true);
		S->completion.store(0.0f);
		S->stage.store(nullptr);
	}
}

} // namespace lgx
