// Internal scene model behind the ltr_* API.
#pragma once

#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lighter.h"
#include "device.h"
#include "host_geom.h"
#include "finalize.h"
#include "place.h"
#include "raster_gpu.h"

namespace lgx {

struct MeshPart {
	uint32_t vcount, voff, icount, ioff;
	int shadow;
};

struct LightContrib {
	f3 normal;
	int32_t sid;
};

struct Instance {
	ltr_Mesh* mesh = nullptr;
	std::string ident;
	float importance = 0;
	Mat4 matrix{};
	// atlas size from the size callback (full resolution, as the samples'
	// texel indices use it); out_w/out_h: the exported image's size (halved
	// by ds2x), set by each export
	uint32_t lm_w = 0, lm_h = 0;
	uint32_t out_w = 0, out_h = 0;
	bool shadow = false;
	bool samplecont = false;
	// world-space vertex data
	std::vector<f3> vpos, vnrm;
	std::vector<f2> vtex, ltex;
	// collision data
	TriSet tris;
	SplitTree split;
	// samples
	std::vector<f3> s_pos, s_nrm;
	std::vector<uint32_t> s_loc;
	std::vector<f4> s_info; // uv0.x, uv0.y, part, world area per texel
	std::vector<f3> lightmap;
	std::vector<LightContrib> contrib;
};

struct Light {
	uint32_t type;
	f3 pos, dir, up, color;
	float range, power, radius;
	int samples;
	float spot_out, spot_in, spot_curve;
};

} // namespace lgx

struct ltr_Mesh {
	ltr_Scene* scene = nullptr;
	std::string ident;
	std::vector<lgx::f3> vpos, vnrm;
	std::vector<lgx::f2> vtex1, vtex2;
	std::vector<uint32_t> indices;
	std::vector<lgx::MeshPart> parts;
};

struct ltr_Scene {
	ltr_Config config;
	std::vector<ltr_Mesh*> meshes;
	std::vector<lgx::Instance*> insts; // insts[0] holds the externally added samples
	std::vector<lgx::Light> lights;
	std::vector<ltr_SampleInfo> samples;
	lgx::BoxTree inst_tree;
	std::vector<ltr_WorkOutput> outputs;
	// images of earlier exports: the caller may still hold their pointers,
	// so they stay allocated until ltr_DestroyScene
	std::vector<ltr_WorkOutput> retired;

	// bake state
	std::atomic<const char*> stage{"not started"};
	std::atomic<float> completion{0.0f};
	std::atomic<bool> abort{false};
	std::atomic<bool> failed{false};
	std::thread worker;
	std::string error;
	int nthreads = 1;

	// GPU state, alive from the first GPU stage to the end of the bake
	std::unique_ptr<lgx::DeviceScene> dev;
	std::vector<size_t> inst_off; // first global sample index of each instance
	size_t total_samples = 0;
	bool prepared = false;
	std::vector<float> vis; // per work item shadow factor, when normal-map data is requested
	lgx::StageTimes times;
	// rasterisation, sample placement and lightmap export: on the GPU unless
	// the host restatement is selected
	bool host_raster = false, host_place = false, host_finalize = false;
	lgx::RasterTimes raster_times;
	std::vector<float> nm_data; // normal-map data of the GPU pass: normal xyz + min-dot per sample
	lgx::PlaceTimes place_times;
	// wall ms of the preparation steps: transforms, collision set-up (the
	// instances' split and triangle trees), the instance tree, samples; and
	// within the samples: rasterisation, placement, the samples' commit
	double prep_ms[7] = {0, 0, 0, 0, 0, 0, 0};
	lgx::FinTimes fin_times;
};
