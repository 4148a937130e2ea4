// The C entry points: lighter's ltr_* API (include/lighter.h) and the lgx_*
// extensions (include/lighter_gpu.h).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "../../include/lighter_gpu.h"
#include "bvh.h"
#include "pipeline.h"
#include "radiosity.h"

using namespace lgx;

// ---------------------------------------------------------------------------
// ABI layout of the public structures (x86-64), identical to the reference
// header's so existing callers link unchanged.
// ---------------------------------------------------------------------------
static_assert(sizeof(ltr_MeshPartInfo) == 72, "ltr_MeshPartInfo");
static_assert(sizeof(ltr_MeshInstanceInfo) == 88, "ltr_MeshInstanceInfo");
static_assert(sizeof(ltr_LightInfo) == 80, "ltr_LightInfo");
static_assert(sizeof(ltr_SampleInfo) == 40, "ltr_SampleInfo");
static_assert(sizeof(ltr_SampleRequest) == 104, "ltr_SampleRequest");
static_assert(sizeof(ltr_Config) == 144, "ltr_Config");
static_assert(sizeof(ltr_WorkOutputInfo) == 16, "ltr_WorkOutputInfo");
static_assert(sizeof(ltr_WorkOutput) == 64, "ltr_WorkOutput");
static_assert(sizeof(ltr_WorkStatus) == 16, "ltr_WorkStatus");
static_assert(offsetof(ltr_Config, size_fn) == 8 && offsetof(ltr_Config, max_num_threads) == 16, "ltr_Config");
static_assert(offsetof(ltr_Config, bounce_count) == 80 && offsetof(ltr_Config, sample_fn) == 88, "ltr_Config");
static_assert(offsetof(ltr_Config, ao_distance) == 96 && offsetof(ltr_Config, ao_num_samples) == 128, "ltr_Config");
static_assert(offsetof(ltr_Config, blur_size) == 132 && offsetof(ltr_Config, generate_normalmap_data) == 140,
              "ltr_Config");
static_assert(offsetof(ltr_WorkOutput, lightmap_rgb) == 40 && offsetof(ltr_WorkOutput, width) == 56,
              "ltr_WorkOutput");

static thread_local std::string g_error;

template <class F> static int guarded(F&& f)
{
	try {
		f();
		return 1;
	} catch (const std::exception& e) {
		g_error = e.what();
		return 0;
	} catch (const AbortBake&) {
		g_error = "aborted";
		return 0;
	}
}

static int online_cpus()
{
	long n = sysconf(_SC_NPROCESSORS_ONLN);
	return n > 0 ? (int)n : 1;
}

// ---------------------------------------------------------------------------
// ltr_* (lighter.h)
// ---------------------------------------------------------------------------
extern "C" {

// lighter.cpp:7-28
LTRBOOL ltr_DefaultSizeFunc(ltr_Config* config, const char*, size_t, const char*, size_t, float computed_surface_area,
                            float inst_importance, u32 out_size[2])
{
	float size = config->global_size_factor * sqrtf(computed_surface_area) * inst_importance;
	if (size < 1)
		size = 1;
	u32 s = ltr_NextPowerOfTwo((u32)size);
	if (s > config->max_lightmap_size)
		return 0;
	out_size[0] = s;
	out_size[1] = s;
	return 1;
}

// lighter.cpp:1348-1357
u32 ltr_NextPowerOfTwo(u32 x)
{
	x--;
	for (int sh = 1; sh < 32; sh <<= 1)
		x |= x >> sh;
	return x + 1;
}

// lighter.cpp:1171-1208
void ltr_GetConfig(ltr_Config* cfg, ltr_Scene* opt_scene)
{
	if (opt_scene) {
		memcpy(cfg, &opt_scene->config, sizeof(*cfg));
		return;
	}
	memset(cfg, 0, sizeof(*cfg));
	cfg->userdata = NULL;
	cfg->size_fn = ltr_DefaultSizeFunc;
	cfg->max_num_threads = 0x7fff;
	cfg->max_tree_memory = 128 * 1024 * 1024;
	cfg->max_lightmap_size = 1024;
	cfg->default_width = 64;
	cfg->default_height = 64;
	cfg->global_size_factor = 4;
	cfg->max_correct_dist = 0.1f;
	cfg->max_correct_angle = 60;
	LTR_VEC3_SET(cfg->clear_color, 0, 0, 0);
	LTR_VEC3_SET(cfg->ambient_color, 0, 0, 0);
	cfg->bounce_count = 0;
	cfg->sample_fn = NULL;
	cfg->ao_distance = 0;
	cfg->ao_multiplier = 1.2f;
	cfg->ao_falloff = 1;
	cfg->ao_effect = 0;
	cfg->ao_divergence = 0;
	LTR_VEC3_SET(cfg->ao_color_rgb, 0, 0, 0);
	cfg->ao_num_samples = 17;
	cfg->blur_size = 0.5f;
	cfg->ds2x = 0;
	cfg->generate_normalmap_data = 0;
}

LTRCODE ltr_SetConfig(ltr_Scene* scene, ltr_Config* cfg)
{
	memcpy(&scene->config, cfg, sizeof(*cfg));
	return 1;
}

// lighter_int.hpp:971-981
ltr_Scene* ltr_CreateScene()
{
	ltr_Scene* S = new ltr_Scene();
	ltr_GetConfig(&S->config, NULL);
	Instance* c = new Instance();
	c->samplecont = true;
	S->insts.push_back(c);
	return S;
}

static void join_bake(ltr_Scene* S)
{
	if (S->worker.joinable())
		S->worker.join();
}

void ltr_DestroyScene(ltr_Scene* scene)
{
	if (!scene)
		return;
	scene->abort.store(true);
	join_bake(scene);
	scene->dev.reset();
	for (Instance* mi : scene->insts)
		delete mi;
	for (ltr_Mesh* m : scene->meshes)
		delete m;
	for (auto* list : {&scene->outputs, &scene->retired})
		for (ltr_WorkOutput& o : *list) {
			delete[] o.lightmap_rgb;
			delete[] o.normals_xyzf;
		}
	delete scene;
}

// lighter.cpp:1147-1152: returns at once; the bake runs on its own thread.
void ltr_Start(ltr_Scene* scene)
{
	if (scene->stage.load() != nullptr && scene->worker.joinable())
		return; // a bake is already running
	join_bake(scene);
	// every bake starts from the scene as added: samples, lightmaps and
	// device state of an earlier bake are dropped (its output images stay
	// allocated until ltr_DestroyScene)
	reset_bake(scene);
	scene->nthreads = std::max(std::min(scene->config.max_num_threads, online_cpus()), 1);
	scene->abort.store(false);
	scene->failed.store(false);
	scene->stage.store("transforming spatial data");
	scene->completion.store(0.0f);
	scene->worker = std::thread([scene]() { bake(scene); });
}

// The reference leaves this empty (lighter.cpp:1154-1157); here the bake
// stops at the next stage boundary and reports "finished" with no outputs.
void ltr_Abort(ltr_Scene* scene) { scene->abort.store(true); }

// lighter.cpp:1159-1164; after a failure the stage reads "failed".
LTRBOOL ltr_GetStatus(ltr_Scene* scene, ltr_WorkStatus* wsout)
{
	const char* st = scene->stage.load();
	wsout->completion = scene->completion.load();
	wsout->stage = st ? st : (scene->failed.load() ? "failed" : "finished");
	return st != nullptr;
}

void ltr_Sleep(int ms)
{
	if (ms >= 1000) {
		sleep(ms / 1000);
		ms %= 1000;
	}
	if (ms > 0)
		usleep(ms * 1000);
}

// lighter.cpp:1217-1224
ltr_Mesh* ltr_CreateMesh(ltr_Scene* scene, const char* ident, size_t ident_size)
{
	ltr_Mesh* m = new ltr_Mesh();
	m->scene = scene;
	if (ident)
		m->ident.assign(ident, ident_size);
	scene->meshes.push_back(m);
	return m;
}

// lighter.cpp:1226-1264
LTRBOOL ltr_MeshAddPart(ltr_Mesh* mesh, ltr_MeshPartInfo* mp)
{
	if (mp->index_count < 3 && mp->index_count % 3 != 0)
		return 0;
	MeshPart part{mp->vertex_count, (uint32_t)mesh->vpos.size(), mp->index_count, (uint32_t)mesh->indices.size(),
	              mp->shadow};
	size_t base = mesh->vpos.size();
	mesh->vpos.resize(base + mp->vertex_count);
	mesh->vnrm.resize(base + mp->vertex_count);
	mesh->vtex1.resize(base + mp->vertex_count);
	mesh->vtex2.resize(base + mp->vertex_count);
	for (u32 i = 0; i < mp->vertex_count; ++i) {
		memcpy(&mesh->vpos[base + i], (const char*)mp->positions_f3 + (size_t)mp->stride_positions * i, 12);
		memcpy(&mesh->vnrm[base + i], (const char*)mp->normals_f3 + (size_t)mp->stride_normals * i, 12);
		memcpy(&mesh->vtex1[base + i], (const char*)mp->texcoords1_f2 + (size_t)mp->stride_texcoords1 * i, 8);
		memcpy(&mesh->vtex2[base + i], (const char*)mp->texcoords2_f2 + (size_t)mp->stride_texcoords2 * i, 8);
	}
	mesh->indices.insert(mesh->indices.end(), mp->indices, mp->indices + mp->index_count);
	mesh->parts.push_back(part);
	return 1;
}

// lighter.cpp:1266-1280
LTRBOOL ltr_MeshAddInstance(ltr_Mesh* mesh, ltr_MeshInstanceInfo* mii)
{
	Instance* mi = new Instance();
	mi->mesh = mesh;
	mi->importance = mii->importance;
	mi->shadow = !!mii->shadow;
	mi->samplecont = false;
	if (mii->ident)
		mi->ident.assign(mii->ident, mii->ident_size);
	memcpy(&mi->matrix.m[0][0], mii->matrix, sizeof(Mat4));
	mi->lm_w = 128;
	mi->lm_h = 128;
	mesh->scene->insts.push_back(mi);
	return 1;
}

// lighter.cpp:1282-1317.  The reference draws one value from the C library
// generator per light here; the call is kept so that the sequence the AO
// stage draws from stays the same.
void ltr_LightAdd(ltr_Scene* scene, ltr_LightInfo* li)
{
	Light L;
	L.type = li->type;
	L.pos = ld3(li->position);
	L.dir = normalized(ld3(li->direction));
	L.up = normalized(ld3(li->up_direction));
	L.color = ld3(li->color_rgb);
	L.range = li->range;
	L.power = li->power;
	L.radius = li->light_radius;
	L.samples = std::max(1, li->shadow_sample_count);
	L.spot_out = li->spot_angle_out;
	L.spot_in = li->spot_angle_in;
	L.spot_curve = li->spot_curve;
	(void)rand();
	scene->lights.push_back(L);
}

// lighter.cpp:1319-1329
void ltr_SampleAdd(ltr_Scene* scene, ltr_SampleInfo* si)
{
	ltr_SampleInfo s = *si;
	s.out_color[0] = s.out_color[1] = s.out_color[2] = 0;
	Instance* c = scene->insts[0];
	c->s_pos.push_back(ld3(s.position));
	c->s_nrm.push_back(ld3(s.normal));
	c->lightmap.push_back(mk3(0, 0, 0));
	scene->samples.push_back(s);
}

// lighter.cpp:1332-1345
void ltr_GetWorkOutputInfo(ltr_Scene* scene, ltr_WorkOutputInfo* woutinfo)
{
	woutinfo->lightmap_count = (u32)scene->outputs.size();
	woutinfo->sample_count = (u32)scene->samples.size();
	woutinfo->samples = scene->samples.empty() ? NULL : scene->samples.data();
}

LTRBOOL ltr_GetWorkOutput(ltr_Scene* scene, u32 which, ltr_WorkOutput* wout)
{
	if (which >= scene->outputs.size())
		return 0;
	*wout = scene->outputs[which];
	return 1;
}

// ---------------------------------------------------------------------------
// lgx_* staged bake
// ---------------------------------------------------------------------------
const char* lgx_LastError(void) { return g_error.c_str(); }

int lgx_HasDevice(void)
{
	return guarded([] { require_gpu(); });
}

static void staged_begin(ltr_Scene* S)
{
	if (S->worker.joinable() && S->stage.load() != nullptr)
		throw std::runtime_error("a bake started with ltr_Start is running");
	S->nthreads = std::max(std::min(S->config.max_num_threads, online_cpus()), 1);
}

int lgx_Prepare(ltr_Scene* S)
{
	return guarded([&] {
		staged_begin(S);
		S->abort.store(false);
		reset_bake(S);
		prepare(S);
		S->stage.store(nullptr);
	});
}

int lgx_StageDirect(ltr_Scene* S)
{
	return guarded([&] {
		staged_begin(S);
		prepare(S);
		if (!S->lights.empty())
			stage_direct(S);
		S->stage.store(nullptr);
	});
}

int lgx_StageRadiosity(ltr_Scene* S)
{
	return guarded([&] {
		staged_begin(S);
		prepare(S);
		stage_radiosity(S);
		S->stage.store(nullptr);
	});
}

int lgx_StageAO(ltr_Scene* S)
{
	return guarded([&] {
		staged_begin(S);
		prepare(S);
		stage_ao(S);
		S->stage.store(nullptr);
	});
}

int lgx_StageFinalize(ltr_Scene* S)
{
	return guarded([&] {
		staged_begin(S);
		prepare(S);
		if (S->dev)
			download_lightmaps(S);
		finalize(S);
		S->stage.store(nullptr);
	});
}

int lgx_SetHostStages(ltr_Scene* S, int mask)
{
	return guarded([&] {
		if (mask & ~(LGX_HOST_PLACEMENT | LGX_HOST_FINALIZE | LGX_HOST_RASTER))
			throw std::runtime_error("lgx_SetHostStages: unknown bits in mask");
		if (S->worker.joinable() && S->stage.load() != nullptr)
			throw std::runtime_error("a bake started with ltr_Start is running");
		S->host_place = (mask & LGX_HOST_PLACEMENT) != 0;
		S->host_finalize = (mask & LGX_HOST_FINALIZE) != 0;
		S->host_raster = (mask & LGX_HOST_RASTER) != 0;
	});
}

int lgx_RasterTimes(ltr_Scene* S, double out[3])
{
	const RasterTimes& t = S->raster_times;
	out[0] = t.kernel_ms;
	out[1] = (double)t.texels;
	out[2] = (double)t.samples;
	return 1;
}

int lgx_FinalizeTimes(ltr_Scene* S, double out[5])
{
	const FinTimes& t = S->fin_times;
	out[0] = t.pack_ms;
	out[1] = t.upload_ms;
	out[2] = t.kernels_ms;
	out[3] = t.download_ms;
	out[4] = (double)t.texels;
	return 1;
}

int lgx_PlacementTimes(ltr_Scene* S, double out[5])
{
	const PlaceTimes& t = S->place_times;
	out[0] = t.flatten_ms;
	out[1] = t.upload_ms;
	out[2] = t.kernel_ms;
	out[3] = t.download_ms;
	out[4] = (double)t.samples;
	return 1;
}

int lgx_SplitTreeCompare(const float* tris9, size_t n, double out[3])
{
	return guarded([&] {
		std::vector<Tri> t(n);
		for (size_t i = 0; i < n; ++i)
			t[i] = Tri{mk3(tris9[9 * i], tris9[9 * i + 1], tris9[9 * i + 2]),
			           mk3(tris9[9 * i + 3], tris9[9 * i + 4], tris9[9 * i + 5]),
			           mk3(tris9[9 * i + 6], tris9[9 * i + 7], tris9[9 * i + 8])};
		out[0] = split_tree_compare(t.data(), n, &out[1], &out[2]) ? 1.0 : 0.0;
	});
}

int lgx_PrepareTimes(ltr_Scene* S, double out[7])
{
	for (int i = 0; i < 7; ++i)
		out[i] = S->prep_ms[i];
	return 1;
}

int lgx_InstanceCount(ltr_Scene* S) { return (int)S->insts.size(); }

int lgx_InstanceInfo(ltr_Scene* S, int inst, int32_t out[4])
{
	if (inst < 0 || (size_t)inst >= S->insts.size())
		return 0;
	Instance* mi = S->insts[inst];
	out[0] = (int32_t)mi->lm_w;
	out[1] = (int32_t)mi->lm_h;
	out[2] = (int32_t)mi->lightmap.size();
	out[3] = (int32_t)mi->tris.tris().size();
	return 1;
}

int lgx_GetSamples(ltr_Scene* S, int inst, float* pos, float* nrm, uint32_t* loc, float* radinfo)
{
	if (inst < 0 || (size_t)inst >= S->insts.size())
		return 0;
	Instance* mi = S->insts[inst];
	for (size_t i = 0; i < mi->s_pos.size(); ++i) {
		if (pos)
			st3(pos + 3 * i, mi->s_pos[i]);
		if (nrm)
			st3(nrm + 3 * i, mi->s_nrm[i]);
		if (loc && i < mi->s_loc.size())
			loc[i] = mi->s_loc[i];
		if (radinfo && i < mi->s_info.size())
			memcpy(radinfo + 4 * i, &mi->s_info[i], 16);
	}
	return 1;
}

int lgx_GetLightmap(ltr_Scene* S, int inst, float* rgb)
{
	return guarded([&] {
		if (inst < 0 || (size_t)inst >= S->insts.size())
			throw std::runtime_error("instance out of range");
		if (S->dev)
			download_lightmaps(S);
		Instance* mi = S->insts[inst];
		for (size_t i = 0; i < mi->lightmap.size(); ++i)
			st3(rgb + 3 * i, mi->lightmap[i]);
	});
}

int lgx_SetLightmap(ltr_Scene* S, int inst, const float* rgb)
{
	return guarded([&] {
		if (inst < 0 || (size_t)inst >= S->insts.size())
			throw std::runtime_error("instance out of range");
		if (S->dev)
			download_lightmaps(S);
		Instance* mi = S->insts[inst];
		for (size_t i = 0; i < mi->lightmap.size(); ++i)
			mi->lightmap[i] = ld3(rgb + 3 * i);
		if (S->dev)
			upload_lightmaps(S);
	});
}

int lgx_GetTriangles(ltr_Scene* S, int inst, float* tris9)
{
	if (inst < 0 || (size_t)inst >= S->insts.size())
		return 0;
	const auto& T = S->insts[inst]->tris.tris();
	for (size_t i = 0; i < T.size(); ++i) {
		st3(tris9 + 9 * i, T[i].p1);
		st3(tris9 + 9 * i + 3, T[i].p2);
		st3(tris9 + 9 * i + 6, T[i].p3);
	}
	return 1;
}

int lgx_LightInstances(ltr_Scene* S, int light, int32_t* out, int cap)
{
	if (light < 0 || (size_t)light >= S->lights.size() || !S->prepared)
		return -1;
	std::vector<int32_t> ids = light_instances(S, S->lights[light]);
	for (int i = 0; i < cap && (size_t)i < ids.size(); ++i)
		out[i] = ids[i];
	return (int)ids.size();
}

size_t lgx_RadLinkCount(ltr_Scene* S) { return S->dev ? S->dev->rad_num_links() : 0; }

int lgx_RadGetLinks(ltr_Scene* S, uint32_t* rowmap, uint32_t* other, float* factor)
{
	return guarded([&] {
		if (!S->dev)
			throw std::runtime_error("no radiosity state");
		S->dev->rad_get_links(rowmap, other, factor);
	});
}

int lgx_StageTimes(ltr_Scene* S, double out[5])
{
	out[0] = S->times.direct_ms;
	out[1] = S->times.direct_march_ms;
	out[2] = S->times.ao_ms;
	out[3] = S->times.links_ms;
	out[4] = S->times.bounce_ms;
	return 1;
}

// ---------------------------------------------------------------------------
// lgx_* device-level
// ---------------------------------------------------------------------------
struct lgx_Device {
	DeviceScene dev;
	std::unique_ptr<RadSolver> rad;
};

lgx_Device* lgx_DeviceCreate(const float* tris9, size_t ntri)
{
	lgx_Device* d = nullptr;
	if (!guarded([&] {
		    d = new lgx_Device();
		    d->dev.set_triangles(tris9, ntri);
	    })) {
		delete d;
		return nullptr;
	}
	return d;
}

void lgx_DeviceDestroy(lgx_Device* d) { delete d; }

int lgx_DeviceInfo(lgx_Device* d, int64_t out[3])
{
	out[0] = (int64_t)d->dev.num_triangles();
	out[1] = (int64_t)d->dev.bvh_nodes();
	out[2] = (int64_t)d->dev.bvh_depth();
	return 1;
}

int lgx_SetEntryGrid(lgx_Device* d, int on, int64_t* cells)
{
	return guarded([&] {
		d->dev.set_entry_grid(on != 0);
		if (cells)
			*cells = (int64_t)d->dev.grid_cells();
	});
}

int lgx_QueryDistance(lgx_Device* d, const float* p, size_t n, float* out)
{
	return guarded([&] { d->dev.query_distance(p, n, out); });
}

int lgx_QueryShadow(lgx_Device* d, const float* from, const float* to, const float* k, size_t n, float* out,
                    uint32_t* steps)
{
	return guarded([&] { d->dev.query_shadow(from, to, k, n, out, steps); });
}

int lgx_QueryClosest(lgx_Device* d, const float* a, const float* b, size_t n, float* out)
{
	return guarded([&] { d->dev.query_closest(a, b, n, out); });
}

int lgx_QueryOccluded(lgx_Device* d, const float* a, const float* b, size_t n, uint8_t* out)
{
	return guarded([&] { d->dev.query_occluded(a, b, n, out); });
}

int lgx_SetSamples(lgx_Device* d, const float* pos, const float* nrm, size_t n, const float* lm)
{
	return guarded([&] {
		d->dev.set_samples(pos, nrm, n);
		if (lm)
			d->dev.set_lightmap(lm);
	});
}

int lgx_ReadLightmap(lgx_Device* d, float* lm)
{
	return guarded([&] { d->dev.get_lightmap(lm); });
}

int lgx_DirectLight(lgx_Device* d, const lgx_Light* lights, int nlights, const uint32_t* items, size_t nitems,
                    const uint32_t* light_off, int keep_vis)
{
	return lgx_DirectLightSetup(d, lights, nlights, items, nitems, light_off, keep_vis) && lgx_DirectLightRun(d);
}

int lgx_DirectLightRun(lgx_Device* d)
{
	return guarded([&] {
		if (!d->dev.run_direct())
			throw std::runtime_error("direct light aborted");
	});
}

int lgx_BvhCompare(const float* tris, size_t ntri, double out[6])
{
	return guarded([&] {
		require_gpu();
		compare_bvh_builds(tris, ntri, out);
	});
}

int lgx_DeviceAbort(lgx_Device* d, int on)
{
	return guarded([&] { d->dev.request_abort(on != 0); });
}

int lgx_DirectLightSetup(lgx_Device* d, const lgx_Light* lights, int nlights, const uint32_t* items, size_t nitems,
                         const uint32_t* light_off, int keep_vis)
{
	return guarded([&] {
		if (nlights < 0 || nlights > 64)
			throw std::runtime_error("lgx_DirectLight: 0..64 lights per call");
		std::vector<DevLight> dl;
		for (int l = 0; l < nlights; ++l) {
			Light L;
			L.type = (uint32_t)lights[l].type;
			L.pos = ld3(lights[l].pos);
			L.dir = ld3(lights[l].dir);
			L.color = ld3(lights[l].color);
			L.range = lights[l].range;
			L.power = lights[l].power;
			L.radius = lights[l].radius;
			L.spot_out = lights[l].spot_angle_out;
			L.spot_in = lights[l].spot_angle_in;
			L.spot_curve = lights[l].spot_curve;
			dl.push_back(to_dev_light(L));
		}
		std::vector<uint32_t> it(items, items + nitems), off(light_off, light_off + nlights + 1);
		if (off.back() != nitems)
			throw std::runtime_error("lgx_DirectLight: light_off does not end at nitems");
		for (uint32_t v : it)
			if ((v >> 26) >= (uint32_t)nlights || (v & ((1u << 26) - 1)) >= d->dev.num_samples())
				throw std::runtime_error("lgx_DirectLight: work item out of range");
		d->dev.set_direct_work(dl, it, off, keep_vis != 0);
	});
}

int lgx_DirectLightNoAlloc(lgx_Device* d, const lgx_Light* lights, int nlights, const uint32_t* d_items,
                           size_t nitems, const uint32_t* light_off, float* d_weights, float* d_vis)
{
	return guarded([&] {
		if (nlights < 0 || nlights > 64)
			throw std::runtime_error("lgx_DirectLightNoAlloc: 0..64 lights per call");
		if (!light_off || (nitems && (!d_items || !d_weights)) || (nlights && !lights))
			throw std::runtime_error("lgx_DirectLightNoAlloc: null array");
		DevLight dl[64];
		for (int l = 0; l < nlights; ++l) {
			Light L;
			L.type = (uint32_t)lights[l].type;
			L.pos = ld3(lights[l].pos);
			L.dir = ld3(lights[l].dir);
			L.color = ld3(lights[l].color);
			L.range = lights[l].range;
			L.power = lights[l].power;
			L.radius = lights[l].radius;
			L.spot_out = lights[l].spot_angle_out;
			L.spot_in = lights[l].spot_angle_in;
			L.spot_curve = lights[l].spot_curve;
			dl[l] = to_dev_light(L);
		}
		d->dev.run_direct_noalloc(dl, nlights, d_items, nitems, light_off, d_weights, d_vis);
	});
}

int lgx_ReadWeights(lgx_Device* d, float* w, float* vis, size_t n)
{
	return guarded([&] {
		if (w)
			d->dev.get_weights(w, n);
		if (vis)
			d->dev.get_vis(vis, n);
	});
}

int lgx_AmbientOcclusion(lgx_Device* d, const float* randoff, int num_samples, float distance, float multiplier,
                         float falloff, float effect, const float* color3)
{
	return guarded(
	    [&] { d->dev.ambient_occlusion(randoff, num_samples, distance, multiplier, falloff, effect, color3); });
}

int lgx_RadLinks(lgx_Device* d, const float* pos, const float* nrm, size_t n, size_t* nlinks)
{
	return guarded([&] {
		if (n >= ((size_t)1 << 32))
			throw std::runtime_error("lgx_RadLinks: patch indices are 32-bit");
		d->rad.reset(); // its dense matrix belongs to the old links
		d->dev.rad_set_patches(pos, nrm, n);
		size_t nl = d->dev.rad_links();
		if (nlinks)
			*nlinks = nl;
	});
}

int lgx_RadReadLinks(lgx_Device* d, uint32_t* rowmap, uint32_t* other, float* factor)
{
	return guarded([&] { d->dev.rad_get_links(rowmap, other, factor); });
}

int lgx_RadBounce(lgx_Device* d, const float* diffuse3, const float* emit3, const float* area, int bounces,
                  float* total3)
{
	return guarded([&] {
		d->dev.rad_set_colors(diffuse3, emit3, area);
		d->dev.rad_bounce(bounces);
		d->dev.rad_get_total(total3);
	});
}

int lgx_RadSolve(lgx_Device* d, const double* E, uint32_t R, const float* diffuse3, const float* area, int form,
                 int bounces, int refine, double* total, double times_ms[4])
{
	return guarded([&] {
		if (form < 0 || form >= kNumForms)
			throw std::runtime_error("lgx_RadSolve: unknown form");
		if (bounces < 0 || refine < 0)
			throw std::runtime_error("lgx_RadSolve: negative count");
		if (times_ms)
			times_ms[0] = times_ms[1] = times_ms[2] = times_ms[3] = 0.0;
		if (d->dev.npatches == 0)
			return; // no patches: nothing to solve
		if (!d->rad || d->rad->N != d->dev.npatches)
			d->rad.reset(new RadSolver(d->dev));
		d->rad->set_problem(E, R, diffuse3, area);
		d->rad->solve(form, bounces, refine, total);
		if (times_ms) {
			times_ms[0] = d->rad->times.build_ms;
			times_ms[1] = d->rad->times.series_ms;
			times_ms[2] = d->rad->times.per_bounce_ms;
			times_ms[3] = d->rad->times.refine_ms;
		}
	});
}

int lgx_RadSolveGrad(lgx_Device* d, const double* E, uint32_t R, const float* diffuse3, const float* area,
                     int form, int bounces, const double* g, double* total, double* lam, double* dalb,
                     double times_ms[4])
{
	return guarded([&] {
		if (form < 0 || form >= kNumForms)
			throw std::runtime_error("lgx_RadSolveGrad: unknown form");
		if (bounces < 0)
			throw std::runtime_error("lgx_RadSolveGrad: negative bounce count");
		if (!g || !total || !lam || !dalb)
			throw std::runtime_error("lgx_RadSolveGrad: null output");
		if (times_ms)
			times_ms[0] = times_ms[1] = times_ms[2] = times_ms[3] = 0.0;
		if (d->dev.npatches == 0)
			return; // no patches: nothing to differentiate
		if (!d->rad || d->rad->N != d->dev.npatches)
			d->rad.reset(new RadSolver(d->dev));
		d->rad->set_problem(E, R, diffuse3, area);
		d->rad->solve_grad(form, bounces, g, total, lam, dalb);
		if (times_ms) {
			times_ms[0] = d->rad->times.build_ms;
			times_ms[1] = d->rad->times.series_ms;
			times_ms[2] = d->rad->times.per_bounce_ms;
			times_ms[3] = d->rad->times.refine_ms;
		}
	});
}

int lgx_RadTiles(lgx_Device* d, uint64_t out[2])
{
	return guarded([&] {
		if (!out)
			throw std::runtime_error("lgx_RadTiles: null output");
		out[0] = out[1] = 0;
		if (d->rad && d->rad->ntiles_built()) {
			out[0] = d->rad->ntiles();
			out[1] = (uint64_t)(d->rad->Np / 256) * (d->rad->Np / 32);
		}
	});
}

int lgx_DirectLightGrad(lgx_Device* d, const double* lam3, double* dcolor3)
{
	return guarded([&] {
		if (!lam3 || !dcolor3)
			throw std::runtime_error("lgx_DirectLightGrad: null argument");
		d->dev.direct_grad(lam3, dcolor3);
	});
}

int lgx_SetLinkSearch(lgx_Device* d, int mode)
{
	return guarded([&] {
		if (mode < 0 || mode > 2)
			throw std::runtime_error("lgx_SetLinkSearch: mode 0 (automatic), 1 (all pairs) or 2 (grid)");
		d->dev.links_mode = mode;
	});
}

int lgx_SetCounting(lgx_Device* d, int on)
{
	d->dev.count_queries = on != 0;
	return 1;
}

int lgx_ReferenceCounts(lgx_Device* d, const float* tris, size_t ntri, uint64_t out[5])
{
	return guarded([&] { d->dev.reference_counts(tris, ntri, out); });
}

int lgx_ReferenceCountsAO(lgx_Device* d, const float* tris, size_t ntri, const float* randoff, int num_samples,
                          float distance, float multiplier, float falloff, uint64_t out[5])
{
	return guarded([&] {
		d->dev.reference_counts_ao(tris, ntri, randoff, num_samples, distance, multiplier, falloff, out);
	});
}

int lgx_ReferenceCountsLinks(lgx_Device* d, const float* tris, size_t ntri, uint64_t out[5])
{
	return guarded([&] { d->dev.reference_counts_links(tris, ntri, out); });
}

int lgx_ReadCounters(lgx_Device* d, uint64_t out[6])
{
	out[0] = d->dev.counters.nodes;
	out[1] = d->dev.counters.tris;
	out[2] = d->dev.counters.steps;
	out[3] = d->dev.counters.queries;
	out[4] = d->dev.counters.marches;
	out[5] = d->dev.counters.pretests;
	return 1;
}

int lgx_DeviceTimes(lgx_Device* d, double out[6])
{
	const StageTimes& t = d->dev.times;
	out[0] = t.direct_ms;
	out[1] = t.direct_march_ms;
	out[2] = t.ao_ms;
	out[3] = t.links_ms;
	out[4] = t.bounce_ms;
	out[5] = t.grad_ms;
	return 1;
}

} // extern "C"
