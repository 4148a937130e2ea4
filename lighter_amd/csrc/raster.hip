// Lightmap rasterisation and sample compaction on the GPU; see raster.h for
// the per-pixel rule and raster_gpu.h for the interface.
//
// Every instance's atlas sits in one texel range (each atlas starting on a
// 256-texel boundary, so that a block of k_raster_count / k_raster_emit
// belongs to one atlas).  Three launches:
//  * k_raster_own: one wave per (pass, triangle); the lanes step through
//    the triangle's pixel box and raise the pixel's owner to the draw key
//    pass * ntri + triangle + 1 by atomicMax.  Draw keys grow in drawing
//    order, so the owner is the last draw that covers the pixel: the one
//    whose values the drawing order leaves there.
//  * k_raster_count: one lane per texel; the owner's values at the texel,
//    and a sample where the interpolated normal is not zero; per-block
//    counts.
//  * k_raster_emit: the same per texel, written at the block's offset (an
//    exclusive scan of the counts on the host) plus the texel's rank in its
//    block: samples in atlas order and texel order, as the host compaction
//    gives them.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "raster.h"
#include "raster_gpu.h"

namespace lgx {

#define RAS_CHECK(x)                                                                                         \
	do {                                                                                                     \
		hipError_t e_ = (x);                                                                                 \
		if (e_ != hipSuccess)                                                                                \
			throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " __FILE__ ":" + \
			                         std::to_string(__LINE__) + ": " #x);                                    \
	} while (0)

namespace {

constexpr int kRasBlock = 256;

struct RasAtlas {
	uint64_t px_off; // first texel (a multiple of kRasBlock)
	uint32_t w, h;
};

template <class T> struct DevArr {
	T* p = nullptr;
	explicit DevArr(size_t n) { RAS_CHECK(hipMalloc((void**)&p, sizeof(T) * (n ? n : 1))); }
	DevArr(const DevArr&) = delete;
	DevArr& operator=(const DevArr&) = delete;
	~DevArr()
	{
		if (p)
			(void)hipFree(p);
	}
};

__global__ __launch_bounds__(kRasBlock) void k_raster_own(const RasterTri* __restrict__ tris,
                                                          const uint32_t* __restrict__ tri_atlas, uint32_t ntri,
                                                          const RasAtlas* __restrict__ atlas,
                                                          uint32_t* __restrict__ owner)
{
	const uint64_t draw = (uint64_t)blockIdx.x * (kRasBlock / 64) + (threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	if (draw >= 2ull * ntri)
		return;
	const uint32_t pass = (uint32_t)(draw / ntri), t = (uint32_t)(draw % ntri);
	const float margin = pass == 0 ? kRasterMargin0 : 0.0f;
	const RasterTri tri = tris[t];
	const RasAtlas A = atlas[tri_atlas[t]];
	PixelBox b;
	if (!raster_box(tri, margin, (int32_t)A.w, (int32_t)A.h, &b))
		return;
	const uint32_t nx = (uint32_t)(b.x1 - b.x0 + 1);
	const uint64_t n = (uint64_t)nx * (uint32_t)(b.y1 - b.y0 + 1);
	const uint32_t key = (uint32_t)draw + 1;
	for (uint64_t k = lane; k < n; k += 64) {
		const int32_t x = b.x0 + (int32_t)(k % nx), y = b.y0 + (int32_t)(k / nx);
		if (raster_covers(tri, margin, x, y))
			atomicMax(owner + A.px_off + (uint32_t)x + (uint64_t)(uint32_t)y * A.w, key);
	}
}

// the values the owner leaves at texel p of atlas A; false: no sample there
__device__ __forceinline__ bool texel_sample(const RasterTri* __restrict__ tris, uint32_t ntri, uint32_t own,
                                             const RasAtlas& A, uint64_t p, f3* pos, f3* nrm, f4* inf)
{
	if (own == 0 || p >= (uint64_t)A.w * A.h)
		return false;
	const uint32_t t = (own - 1) % ntri;
	raster_values(tris[t], (int32_t)(p % A.w), (int32_t)(p / A.w), pos, nrm, inf);
	return !is_zero(*nrm);
}

__device__ __forceinline__ uint32_t block_atlas(const uint32_t* __restrict__ first_block, uint32_t n, uint32_t b)
{
	uint32_t lo = 0, hi = n - 1;
	while (lo < hi) {
		const uint32_t mid = (lo + hi + 1) >> 1;
		if (first_block[mid] <= b)
			lo = mid;
		else
			hi = mid - 1;
	}
	return lo;
}

__global__ __launch_bounds__(kRasBlock) void k_raster_count(const RasterTri* __restrict__ tris, uint32_t ntri,
                                                            const RasAtlas* __restrict__ atlas,
                                                            const uint32_t* __restrict__ first_block,
                                                            uint32_t natlas, const uint32_t* __restrict__ owner,
                                                            uint32_t* __restrict__ bcount)
{
	const RasAtlas A = atlas[block_atlas(first_block, natlas, blockIdx.x)];
	const uint64_t g = (uint64_t)blockIdx.x * kRasBlock + threadIdx.x;
	f3 pos, nrm;
	f4 inf;
	const bool s = texel_sample(tris, ntri, owner[g], A, g - A.px_off, &pos, &nrm, &inf);
	const int n = __syncthreads_count(s);
	if (threadIdx.x == 0)
		bcount[blockIdx.x] = (uint32_t)n;
}

__global__ __launch_bounds__(kRasBlock) void k_raster_emit(const RasterTri* __restrict__ tris, uint32_t ntri,
                                                           const RasAtlas* __restrict__ atlas,
                                                           const uint32_t* __restrict__ first_block,
                                                           uint32_t natlas, const uint32_t* __restrict__ owner,
                                                           const uint64_t* __restrict__ boff,
                                                           float* __restrict__ out_pos, float* __restrict__ out_nrm,
                                                           float* __restrict__ out_inf, uint32_t* __restrict__ out_loc)
{
	__shared__ uint32_t wave_n[kRasBlock / 64];
	const RasAtlas A = atlas[block_atlas(first_block, natlas, blockIdx.x)];
	const uint64_t g = (uint64_t)blockIdx.x * kRasBlock + threadIdx.x;
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	f3 pos, nrm;
	f4 inf;
	const bool s = texel_sample(tris, ntri, owner[g], A, g - A.px_off, &pos, &nrm, &inf);
	const uint64_t m = __ballot(s);
	if (lane == 0)
		wave_n[wave] = (uint32_t)__popcll(m);
	__syncthreads();
	if (!s)
		return;
	uint64_t at = boff[blockIdx.x] + (uint64_t)__popcll(m & ((1ull << lane) - 1));
	for (int v = 0; v < wave; ++v)
		at += wave_n[v];
	st3(out_pos + 3 * at, pos);
	st3(out_nrm + 3 * at, normalized(nrm));
	out_inf[4 * at + 0] = inf.x;
	out_inf[4 * at + 1] = inf.y;
	out_inf[4 * at + 2] = inf.z;
	out_inf[4 * at + 3] = inf.w;
	out_loc[at] = (uint32_t)(g - A.px_off);
}

} // namespace

void rasterize_gpu(const RasterTri* tris, const uint32_t* tri_atlas, size_t ntris,
                   const std::vector<RasterAtlasSize>& sizes, RasterOut& out, RasterTimes* times)
{
	hipEvent_t e0, e1;
	const size_t na = sizes.size();
	out = RasterOut{};
	out.first.assign(na, 0);
	out.count.assign(na, 0);
	if (na == 0)
		return;
	if (ntris >= 0x7fffffffu)
		throw std::runtime_error("rasterize_gpu: too many triangles");
	// atlases back to back, each from a block boundary
	std::vector<RasAtlas> atlas(na);
	std::vector<uint32_t> first_block(na);
	uint64_t px = 0;
	for (size_t a = 0; a < na; ++a) {
		atlas[a] = RasAtlas{px, sizes[a].w, sizes[a].h};
		first_block[a] = (uint32_t)(px / kRasBlock);
		px += ((uint64_t)sizes[a].w * sizes[a].h + kRasBlock - 1) / kRasBlock * kRasBlock;
	}
	for (size_t t = 0; t < ntris; ++t)
		if (tri_atlas[t] >= na)
			throw std::runtime_error("rasterize_gpu: triangle of an unknown atlas");
	const uint64_t nblocks = px / kRasBlock;
	if (nblocks == 0)
		return;
	if (nblocks >= 0xffffffffull)
		throw std::runtime_error("rasterize_gpu: atlases too large");
	const uint32_t ntri = (uint32_t)ntris;
	hipStream_t s;
	RAS_CHECK(hipStreamCreate(&s));
	struct StreamGuard {
		hipStream_t s;
		~StreamGuard() { (void)hipStreamDestroy(s); }
	} sg{s};
	RAS_CHECK(hipEventCreate(&e0));
	RAS_CHECK(hipEventCreate(&e1));
	struct EvGuard {
		hipEvent_t a, b;
		~EvGuard()
		{
			(void)hipEventDestroy(a);
			(void)hipEventDestroy(b);
		}
	} eg{e0, e1};
	DevArr<RasterTri> d_tris(ntris);
	DevArr<uint32_t> d_tatlas(ntris), d_first(na), d_owner(px), d_count(nblocks);
	DevArr<RasAtlas> d_atlas(na);
	DevArr<uint64_t> d_boff(nblocks);
	if (ntri) {
		RAS_CHECK(hipMemcpyAsync(d_tris.p, tris, sizeof(RasterTri) * ntri, hipMemcpyHostToDevice, s));
		RAS_CHECK(hipMemcpyAsync(d_tatlas.p, tri_atlas, 4 * (size_t)ntri, hipMemcpyHostToDevice, s));
	}
	RAS_CHECK(hipMemcpyAsync(d_atlas.p, atlas.data(), sizeof(RasAtlas) * na, hipMemcpyHostToDevice, s));
	RAS_CHECK(hipMemcpyAsync(d_first.p, first_block.data(), 4 * na, hipMemcpyHostToDevice, s));
	RAS_CHECK(hipMemsetAsync(d_owner.p, 0, 4 * px, s));
	RAS_CHECK(hipEventRecord(e0, s));
	if (ntri) {
		const uint64_t waves = 2ull * ntri;
		k_raster_own<<<(unsigned)((waves + 3) / 4), kRasBlock, 0, s>>>(d_tris.p, d_tatlas.p, ntri, d_atlas.p,
		                                                             d_owner.p);
	}
	k_raster_count<<<(unsigned)nblocks, kRasBlock, 0, s>>>(d_tris.p, ntri ? ntri : 1, d_atlas.p, d_first.p,
	                                                      (uint32_t)na, d_owner.p, d_count.p);
	RAS_CHECK(hipGetLastError());
	std::vector<uint32_t> cnt(nblocks);
	RAS_CHECK(hipMemcpyAsync(cnt.data(), d_count.p, 4 * nblocks, hipMemcpyDeviceToHost, s));
	RAS_CHECK(hipStreamSynchronize(s));
	std::vector<uint64_t> boff(nblocks);
	uint64_t total = 0;
	for (uint64_t b = 0; b < nblocks; ++b) {
		boff[b] = total;
		total += cnt[b];
	}
	// per atlas: its first sample and count
	for (size_t a = 0; a < na; ++a) {
		const uint64_t b0 = first_block[a], b1 = a + 1 < na ? first_block[a + 1] : nblocks;
		const uint64_t first = b0 < nblocks ? boff[b0] : total, end = b1 < nblocks ? boff[b1] : total;
		out.first[a] = first;
		out.count[a] = end - first;
	}
	out.total = total;
	DevArr<float> d_pos(3 * total), d_nrm(3 * total), d_inf(4 * total);
	DevArr<uint32_t> d_loc(total);
	RAS_CHECK(hipMemcpyAsync(d_boff.p, boff.data(), 8 * nblocks, hipMemcpyHostToDevice, s));
	k_raster_emit<<<(unsigned)nblocks, kRasBlock, 0, s>>>(d_tris.p, ntri ? ntri : 1, d_atlas.p, d_first.p,
	                                                     (uint32_t)na, d_owner.p, d_boff.p, d_pos.p, d_nrm.p,
	                                                     d_inf.p, d_loc.p);
	RAS_CHECK(hipGetLastError());
	RAS_CHECK(hipEventRecord(e1, s));
	// one copy per array (host buffers left uninitialised: every element is copied)
	out.pos.reset(new float[3 * total + 1]);
	out.nrm.reset(new float[3 * total + 1]);
	out.inf.reset(new float[4 * total + 1]);
	out.loc.reset(new uint32_t[total + 1]);
	if (total) {
		RAS_CHECK(hipMemcpyAsync(out.pos.get(), d_pos.p, 12 * total, hipMemcpyDeviceToHost, s));
		RAS_CHECK(hipMemcpyAsync(out.nrm.get(), d_nrm.p, 12 * total, hipMemcpyDeviceToHost, s));
		RAS_CHECK(hipMemcpyAsync(out.inf.get(), d_inf.p, 16 * total, hipMemcpyDeviceToHost, s));
		RAS_CHECK(hipMemcpyAsync(out.loc.get(), d_loc.p, 4 * total, hipMemcpyDeviceToHost, s));
	}
	RAS_CHECK(hipStreamSynchronize(s));
	if (times) {
		float ms = 0;
		RAS_CHECK(hipEventElapsedTime(&ms, e0, e1));
		times->kernel_ms = ms; // includes the count read-back between the launches
		times->texels = px;
		times->samples = total;
	}
}

} // namespace lgx
