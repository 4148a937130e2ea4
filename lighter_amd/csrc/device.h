// Host interface of the gfx950 hot path: device-resident scene (BVH,
// triangles, samples, lightmap) and the stage launchers used by the bake
// pipeline and by the lgx_* entry points.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <functional>
#include <string>
#include <vector>

namespace lgx {

// Light as stored after ltr_LightAdd (direction normalised), as the kernels
// read it.
// Entry grid of the traversals (kernels.hip): cubic cells of size g from
// lo, n cells per axis; pad grows a cell to the region its entry node covers.
struct GridSpec {
	float lo[3], g;
	int n[3];
	float pad;
};

struct DevLight {
	int32_t type;
	float pos[3];
	float dir[3];
	float color[3];
	float range, power, radius;
	float spot_out_rad, spot_diff; // outer angle and (inner - outer), radians
	float spot_curve;
};

// Work items the allocation-free direct-light pass takes at most: its
// scratch array holds 8 bytes per item (32 MiB) plus a few KiB.
constexpr size_t kNoAllocMaxItems = (size_t)1 << 22;
// GPUs per process whose scratch address is kept.
constexpr int kMaxDevices = 64;
// Patches the all-pairs link search takes at most: its per-row chunk counts
// and offsets take 12 bytes per 256 pairs (12 GB here).
constexpr size_t kMaxAllPairsPatches = 524288;

// Sorts the columns within every row of a CSR structure (links_sort.hip):
// in -> out, nnz entries, rows given by rowptr (nrows + 1 entries, device),
// column values below 2^bits.
void sort_row_columns(const uint32_t* in, uint32_t* out, uint64_t nnz, uint32_t nrows, const uint64_t* rowptr,
                      int bits, void* stream);

// Per-stage timing in milliseconds measured with HIP events on the stream
// the kernels run on.
struct StageTimes {
	double direct_ms = 0, direct_march_ms = 0, ao_ms = 0, links_ms = 0, bounce_ms = 0, grad_ms = 0;
	double bvh_ms = 0; // the device's BVH build (kernels and the level read-backs)
};

// Optional traversal counters (when enabled): node visits and triangle
// tests, for the algorithmic-bytes figure.
struct QueryCounters {
	unsigned long long nodes = 0, tris = 0, steps = 0, queries = 0, marches = 0;
	unsigned long long pretests = 0; // triangles the box pre-test culled (distance queries)
};

class DeviceScene {
public:
	DeviceScene();
	~DeviceScene();
	DeviceScene(const DeviceScene&) = delete;
	DeviceScene& operator=(const DeviceScene&) = delete;

	// Shadow-casting triangles, 9 floats each (world space).
	void set_triangles(const float* tris, size_t ntri);
	// Samples (points + normals) and the lightmap they accumulate into.
	void set_samples(const float* pos, const float* nrm, size_t n);
	void set_lightmap(const float* rgb);
	void get_lightmap(float* rgb) const;
	// the device lightmap (3 floats per sample, in sample order); valid
	// once the stage that wrote it has completed
	const float* lightmap_device() const { return d_lm_; }
	size_t num_samples() const { return nsamples_; }
	size_t num_triangles() const { return ntri_; }
	int bvh_depth() const { return bvh_depth_; }
	int march_stack_depth() const;
	size_t march_lds_bytes() const;
	size_t bvh_nodes() const { return nnodes_; }
	// on = false: every query starts at the root (the entry grid's answers
	// are the same; this is for comparing them)
	void set_entry_grid(bool on) { use_grid_ = on; }
	size_t grid_cells() const { return grid_base_ ? (size_t)grid_.n[0] * grid_.n[1] * grid_.n[2] : 0; }

	// --- direct light (lighter.cpp:485-600, 1075-1096) -------------------
	// items: (light << 26) | sample, grouped by light in ascending light
	// order; light_off has lights.size()+1 entries.  The march kernel writes
	// one weight per item (bit pattern 0xFFFFFFFF where the reference skips
	// the texel); the accumulation adds color * weight per light in order.
	// keep_going (optional) is called about every millisecond while the
	// march kernel runs, with the fraction of marches started; when it
	// returns false the kernel is stopped from the inside (every wave leaves
	// at its next claim of work).  Returns false when the run was aborted,
	// by keep_going or by request_abort: the lightmap then keeps its colours
	// and the weights are incomplete.
	bool direct_light(const std::vector<DevLight>& lights, const std::vector<uint32_t>& items,
	                  const std::vector<uint32_t>& light_off, bool keep_weights,
	                  const std::function<bool(float)>& keep_going = {});
	// the same in two steps: upload the work once, run it (repeatedly)
	void set_direct_work(const std::vector<DevLight>& lights, const std::vector<uint32_t>& items,
	                     const std::vector<uint32_t>& light_off, bool keep_weights);
	bool run_direct(const std::function<bool(float)>& keep_going = {});
	// While set, every direct-light run stops inside its march kernel at the
	// first claim of work and returns false.
	void request_abort(bool on) { abort_requested_ = on; }
	// The same pass without device allocations, on the caller's device
	// buffers (items, weights, optional shadow factors) with host lights and
	// light offsets; the kernels' scratch is a file-scope device array of
	// kNoAllocMaxItems items, shared by every call on one GPU (calls that
	// overlap on one GPU race).  Synchronous.
	void run_direct_noalloc(const DevLight* lights, int nlights, const uint32_t* d_items, size_t nitems,
	                        const uint32_t* light_off, float* d_w, float* d_vis);
	const float* weights_device() const { return d_w_; }
	void get_weights(float* out, size_t n) const;
	// reverse mode of the last direct-light run: dcolor3[l] = sum over light
	// l's items of weight * lam3[sample] (lam3: 3 per device sample)
	void direct_grad(const double* lam3, double* dcolor3);
	// shadow factor per item (1 where no march ran); kept when keep_weights
	void get_vis(float* out, size_t n) const;
	size_t num_items() const { return nitems_; }

	// --- normal-map data (lighter.cpp:505-513, 959-1030) ----------------
	// Keeps the last direct-light run's shadow factors (run with kept
	// weights) with its lights; first_item[l * ninst + i] is the item of
	// instance i's first sample in light l's items, or -1.
	void nm_keep_batch(const std::vector<DevLight>& lights, const std::vector<int32_t>& first_item, size_t ninst);
	size_t nm_lights() const { return nm_lights_.size(); }
	void nm_clear();
	// Per sample: the normal (normalised mean of the sample normal scaled by
	// ambient_bright and the kept contributions, in light order) and the
	// min-dot term, 4 floats into out4; inst_off: first sample of each
	// instance plus the total.
	void nm_compute(const std::vector<uint32_t>& inst_off, float ambient_bright, float* out4);

	// --- ambient occlusion (lighter.cpp:799-841) -----------------------
	// fac_out (device, one float per sample, optional) receives each
	// texel's AO factor before the colour blend.  With between, the texels
	// go in chunks and between(fraction done) is called after each; it
	// returns false to stop (the remaining texels keep their colour).
	void ambient_occlusion(const float* randoff, int num_samples, float distance, float multiplier, float falloff,
	                       float effect, const float* color, float* fac_out = nullptr,
	                       const std::function<bool(float)>& between = {});
	// The AO factors again with the rays traced on a plain binary SAH tree
	// with the canonical closest-hit semantics (measurement only; the
	// lightmap is left as it was): out[0] node visits, out[1] triangle
	// tests, out[2] rays, out[3] texels whose factor differs from the
	// product's, out[4] texels.
	void reference_counts_ao(const float* tris, size_t ntri, const float* randoff, int num_samples, float distance,
	                         float multiplier, float falloff, uint64_t out[5]);
	// The occlusion test of every candidate pair of the last link search
	// (run with counting on, which keeps the pairs) with the canonical
	// any-hit walk on the same binary tree: out[0] node visits, out[1]
	// triangle tests, out[2] pairs tested, out[3] pairs whose link differs
	// from the product's, out[4] links.
	void reference_counts_links(const float* tris, size_t ntri, uint64_t out[5]);

	// --- radiosity (lighter.cpp:667-797) ---------------------------------
	// Patches = samples.  rad_links builds the symmetric link structure of
	// the pairs that pass the form-factor thresholds and whose segment hits
	// a shadow caster (the reference's VisibilityTest returns the hit).
	// Candidates come from a uniform grid of cells as wide as the longest
	// distance a pair can link over, or from all pairs as the reference
	// enumerates them (at most kMaxAllPairsPatches patches); both give the
	// same links.  links_mode 0: the grid, unless it would test more than
	// N^2 / 4 pairs (then all pairs); 1: all pairs; 2: the grid.  Patches
	// with non-finite positions or normals always take all pairs.
	void rad_set_patches(const float* pos, const float* nrm, size_t n);
	size_t rad_links();
	int links_mode = 0;
	size_t rad_num_links() const { return nlinks_; }
	// Reference link list: rowmap {offset,count} per patch, other, factor.
	void rad_get_links(uint32_t* rowmap, uint32_t* other, float* factor) const;
	// Symmetric CSR (both directions).
	void rad_get_csr(std::vector<uint32_t>& rowptr, std::vector<uint32_t>& col, std::vector<float>& val) const;
	// Reference bounce state per patch and bounce loop; E = emitted light.
	void rad_set_colors(const float* diffuse3, const float* emit3, const float* area);
	void rad_bounce(int count);
	void rad_get_total(float* out3) const;

	// --- point queries for tests and timing --------------------------
	void query_distance(const float* pts, size_t n, float* out);
	void query_shadow(const float* from, const float* to, const float* k, size_t n, float* out, uint32_t* steps);
	void query_closest(const float* a, const float* b, size_t n, float* out);
	void query_occluded(const float* a, const float* b, size_t n, uint8_t* out);

	// Traversal counts of the last direct-light run's marches on a plain
	// binary SAH tree with the reference's query semantics (measurement
	// only; tris must be this scene's triangles): out[0] node visits,
	// out[1] triangle tests, out[2] distance queries, out[3] shadow factors
	// that differ from the product's, out[4] marches.
	void reference_counts(const float* tris, size_t ntri, uint64_t out[5]);

	StageTimes times;
	QueryCounters counters;
	bool count_queries = false;

	// device pointers for the radiosity solvers
	struct Rad {
		float* pos = nullptr;
		float* nrm = nullptr;
		uint64_t* rowptr64 = nullptr; // N + 1 row offsets of the symmetric CSR
		uint32_t* rowptr = nullptr;   // the same in 32 bits, while nnz < 2^32 (the solver forms)
		uint32_t* col = nullptr;
		float* val = nullptr;
		size_t nnz = 0;
		float* diffuse = nullptr;
		float* area = nullptr;
		float* total = nullptr;
		float* out = nullptr;
		float* in = nullptr;
		float* x = nullptr; // (out * diffuse) * area per patch, 4 floats (the bounce's gather operand)
	};
	Rad rad;
	size_t npatches = 0;
	void* stream() const { return stream_; }

private:
	void free_all();
	void read_counters();
	void* stream_ = nullptr;
	// BVH and triangles
	void* d_nodes_ = nullptr; // the tree's nodes, then the entry grid's cells
	GridSpec grid_{};
	int32_t grid_base_ = 0; // index of the first cell node (0: no grid)
	int32_t grid_fine_ = 0; // the fine cuts' first cell node (0: none)
	bool use_grid_ = true;
	void* d_dtri_ = nullptr;
	void* d_rtri_ = nullptr;
	size_t ntri_ = 0, nnodes_ = 0;
	int bvh_depth_ = 0;
	float lo_[3] = {0, 0, 0}, hi_[3] = {0, 0, 0};
	// samples
	float* d_pos_ = nullptr;
	uint32_t* d_skey_ = nullptr; // 30-bit Morton code of each sample position (march ordering)
	float* d_nrm_ = nullptr;
	float* d_lm_ = nullptr;
	void* d_cand_ = nullptr;    // candidate pairs (uint2) of the last link search, kept in counting mode
	void* d_clinked_ = nullptr; // and whether each one linked (uint8)
	uint64_t ncand_ = 0;
	// link search: patch bounds, the largest squared normal length, all finite
	float links_lo_[3] = {0, 0, 0}, links_hi_[3] = {0, 0, 0};
	double links_nmax2_ = 0.0;
	bool links_finite_ = true;
	// candidate pairs (uint2, i < j, device memory), grouped by row i;
	// coff (N + 1, device) receives the first candidate of each row
	void* link_candidates_all(uint64_t& ncand, uint64_t* coff);
	void* link_candidates_grid(uint64_t& ncand, uint64_t* coff); // the same set from the grid
	size_t nsamples_ = 0;
	// direct light
	float* d_w_ = nullptr;
	float* d_vis_ = nullptr;
	uint32_t* d_items_ = nullptr;
	DevLight* d_lights_ = nullptr;
	float* d_gate_ = nullptr;     // shading factor per item (marched items)
	uint32_t* d_mlist_ = nullptr; // compacted march list
	uint32_t* d_mcount_ = nullptr; // [0] march count, [1] persistent-kernel cursor, [2] abort flag
	uint32_t nmarch_ = 0;
	unsigned persistent_blocks();
	unsigned persist_blocks_ = 0;
	std::vector<DevLight> h_lights_;
	std::vector<uint32_t> h_light_off_;
	size_t nitems_ = 0;
	size_t nlinks_ = 0;
	unsigned long long* d_counters_ = nullptr;
	void* d_scratch_ = nullptr; // this GPU's allocation-free scratch array
	// march control in pinned, coherent host memory: [0] abort flag (host
	// writes, kernel reads), [1] march-list position last claimed (kernel
	// writes, host reads); ctl_dev_ is its device address
	std::vector<DevLight> nm_lights_;      // every kept light, in order
	std::vector<int32_t> nm_first_;        // per kept light: first item per instance
	std::vector<float*> nm_vis_;           // per kept batch: its shadow factors (device)
	std::vector<uint32_t> nm_light_batch_; // batch of each kept light
	std::vector<size_t> nm_batch_items_;   // items of each kept batch
	size_t nm_ninst_ = 0;
	uint32_t* ctl_host_ = nullptr;
	uint32_t* ctl_dev_ = nullptr;
	bool abort_requested_ = false;
	bool wait_march(uint32_t nmarch, const std::function<bool(float)>& keep_going);
};

// Throws std::runtime_error unless a gfx950-class GPU is usable.
void require_gpu();
std::string device_name();

} // namespace lgx
