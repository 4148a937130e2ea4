// Lightmap export kernels for gfx950; see finalize.h.
//
// All instances' atlases sit back to back in one device image (texel order
// x + y w within an instance).  Every kernel is one lane per output texel;
// a block belongs to one instance, found by a binary search of the
// instances' first blocks (the same for every lane of the block).
//  * k_fin_scatter: sample colours into their texels, texel flags set;
//  * k_fin_extrap (three rounds): extrapolate_texel at every unset texel;
//    flags of the round's start are read and the round's new flags written
//    to a second array, so a round reads only texels it does not write, as
//    lighter.cpp:879-940 does in scan order;
//  * k_fin_conv (twice): blur_texel along the rows, written transposed
//    (lighter_math.cpp:367-398): the second pass transposes back;
//  * k_fin_ds: half_texel, the 2x2 mean with wrap-around
//    (lighter_math.cpp:400-424).
// The per-texel functions (texel.h) are the host export's too.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <stdexcept>
#include <string>
#include <thread>

#include "finalize.h"
#include "texel.h"
#include "vmath.h"

namespace lgx {

#define FIN_CHECK(x)                                                                                         \
	do {                                                                                                     \
		hipError_t e_ = (x);                                                                                 \
		if (e_ != hipSuccess)                                                                                \
			throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " __FILE__ ":" + \
			                         std::to_string(__LINE__) + ": " #x);                                    \
	} while (0)

constexpr int kFinBlock = 256;

struct FinInst {
	uint64_t px_off; // first texel of the atlas in the full-size image
	uint64_t ds_off; // first texel in the downsampled image
	uint32_t w, h, hw, hh;
};

__device__ int fin_inst(const uint32_t* __restrict__ first_block, int n, uint32_t b)
{
	int lo = 0, hi = n - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (first_block[mid] <= b)
			lo = mid;
		else
			hi = mid - 1;
	}
	return lo;
}

__global__ __launch_bounds__(kFinBlock) void k_fin_scatter(const uint32_t* __restrict__ map,
                                                           const float* __restrict__ lm, uint64_t n,
                                                           float* __restrict__ rgb, uint8_t* __restrict__ set)
{
	const uint64_t i = (uint64_t)blockIdx.x * kFinBlock + threadIdx.x;
	if (i >= n)
		return;
	const uint64_t t = map[i];
	st3(rgb + 3 * t, ld3(lm + 3 * i));
	set[t] = 1;
}

__global__ __launch_bounds__(kFinBlock) void k_fin_extrap(const FinInst* __restrict__ inst,
                                                          const uint32_t* __restrict__ fb, int ninst,
                                                          float* __restrict__ rgb, const uint8_t* __restrict__ set0,
                                                          uint8_t* __restrict__ set1)
{
	const int k = fin_inst(fb, ninst, blockIdx.x);
	const FinInst I = inst[k];
	const uint64_t p = (uint64_t)(blockIdx.x - fb[k]) * kFinBlock + threadIdx.x;
	if (p >= (uint64_t)I.w * I.h || set0[I.px_off + p])
		return;
	const AtlasView view{rgb + 3 * I.px_off, set0 + I.px_off, I.w, I.h};
	f3 c;
	if (extrapolate_texel(view, (uint32_t)(p % I.w), (uint32_t)(p / I.w), &c)) {
		st3(rgb + 3 * (I.px_off + p), c);
		set1[I.px_off + p] = 1;
	}
}

// One blur pass over every atlas: rows of length W (w on the first pass, h
// on the second), H rows, each result written transposed.
__global__ __launch_bounds__(kFinBlock) void k_fin_conv(const FinInst* __restrict__ inst,
                                                        const uint32_t* __restrict__ fb, int ninst,
                                                        const float* __restrict__ src, float* __restrict__ dst,
                                                        const float* __restrict__ kern, int ext, int second)
{
	const int k = fin_inst(fb, ninst, blockIdx.x);
	const FinInst I = inst[k];
	const uint32_t W = second ? I.h : I.w, H = second ? I.w : I.h;
	const uint64_t p = (uint64_t)(blockIdx.x - fb[k]) * kFinBlock + threadIdx.x;
	if (p >= (uint64_t)W * H)
		return;
	const uint32_t x = (uint32_t)(p % W);
	const uint64_t y = p / W;
	st3(dst + 3 * (I.px_off + y + (uint64_t)H * x), blur_texel(src + 3 * (I.px_off + y * W), W, (int)x, kern, ext));
}

__global__ __launch_bounds__(kFinBlock) void k_fin_ds(const FinInst* __restrict__ inst,
                                                      const uint32_t* __restrict__ dfb, int ninst,
                                                      const float* __restrict__ src, float* __restrict__ dst)
{
	const int k = fin_inst(dfb, ninst, blockIdx.x);
	const FinInst I = inst[k];
	const uint64_t p = (uint64_t)(blockIdx.x - dfb[k]) * kFinBlock + threadIdx.x;
	if (p >= (uint64_t)I.hw * I.hh)
		return;
	const uint32_t x = (uint32_t)(p % I.hw), y = (uint32_t)(p / I.hw);
	st3(dst + 3 * (I.ds_off + x + (uint64_t)y * I.hw), half_texel(src + 3 * I.px_off, I.w, I.h, x, y));
}

// ===========================================================================
// host side
// ===========================================================================
namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

struct Pinned {
	char* p = nullptr;
	explicit Pinned(size_t bytes) { FIN_CHECK(hipHostMalloc((void**)&p, bytes ? bytes : 1, hipHostMallocDefault)); }
	Pinned(const Pinned&) = delete;
	Pinned& operator=(const Pinned&) = delete;
	~Pinned()
	{
		if (p)
			(void)hipHostFree(p);
	}
};

struct Dev {
	char* p = nullptr;
	explicit Dev(size_t bytes) { FIN_CHECK(hipMalloc((void**)&p, bytes ? bytes : 1)); }
	Dev(const Dev&) = delete;
	Dev& operator=(const Dev&) = delete;
	~Dev()
	{
		if (p)
			(void)hipFree(p);
	}
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

template <class F> void parallel_items(int nthreads, size_t n, const F& fn)
{
	std::atomic<size_t> next{0};
	auto work = [&]() {
		for (size_t i; (i = next.fetch_add(1)) < n;)
			fn(i);
	};
	const int nt = std::max(1, std::min<int>(nthreads, (int)std::min<size_t>(n, 1024)));
	std::vector<std::thread> th;
	for (int t = 1; t < nt; ++t)
		th.emplace_back(work);
	work();
	for (auto& t : th)
		t.join();
}

} // namespace

void finalize_gpu(const std::vector<FinImage>& imgs, const float* d_rgb, float blur_size,
                  const std::vector<float>& kernel, bool ds2x, int nthreads, FinTimes* times)
{
	FinTimes tm;
	const int n = (int)imgs.size();
	if (n == 0) {
		if (times)
			*times = tm;
		return;
	}
	auto t0 = Clock::now();
	std::vector<FinInst> inst(n);
	std::vector<uint32_t> fb(n), dfb(n);
	std::vector<size_t> s_off(n);
	uint64_t px = 0, dpx = 0, blocks = 0, dblocks = 0, ns = 0;
	for (int k = 0; k < n; ++k) {
		const FinImage& im = imgs[k];
		const uint64_t np = (uint64_t)im.w * im.h;
		const uint64_t nd = (uint64_t)im.final_w * im.final_h;
		inst[k] = FinInst{px, dpx, im.w, im.h, ds2x ? im.final_w : 0u, ds2x ? im.final_h : 0u};
		fb[k] = (uint32_t)blocks;
		dfb[k] = (uint32_t)dblocks;
		s_off[k] = ns;
		px += np;
		dpx += ds2x ? nd : 0;
		blocks += (np + kFinBlock - 1) / kFinBlock;
		dblocks += ds2x ? (nd + kFinBlock - 1) / kFinBlock : 0;
		ns += im.nsamples;
	}
	if (px > 0xFFFFFFFFull || blocks > 0x7FFFFFFFull || dblocks > 0x7FFFFFFFull)
		throw std::runtime_error("finalize: more than 2^32 texels in all atlases");
	tm.texels = px;
	const int ext = blur_size != 0.0f ? (int)((kernel.size() - 1) / 2) : 0;
	if (blur_size != 0.0f && kernel.size() != (size_t)(2 * ext + 1))
		throw std::runtime_error("finalize: the blur kernel must have an odd number of taps");

	// pinned block: instance table, first-block tables, taps, sample texel
	// indices and colours
	size_t o = 0;
	auto take = [&](size_t bytes) {
		size_t at = o;
		o = align256(o + bytes);
		return at;
	};
	const size_t o_inst = take(n * sizeof(FinInst)), o_fb = take(n * 4), o_dfb = take(n * 4),
	             o_kern = take(std::max<size_t>(kernel.size(), 1) * 4), o_map = take(ns * 4),
	             o_lm = take(d_rgb ? 0 : ns * 12);
	const size_t up_bytes = o;
	const uint64_t out_px = ds2x ? dpx : px;
	Pinned host(std::max<size_t>(up_bytes, out_px * 12));
	memcpy(host.p + o_inst, inst.data(), n * sizeof(FinInst));
	memcpy(host.p + o_fb, fb.data(), n * 4);
	memcpy(host.p + o_dfb, dfb.data(), n * 4);
	if (!kernel.empty())
		memcpy(host.p + o_kern, kernel.data(), kernel.size() * 4);
	std::atomic<bool> bad_loc{false};
	parallel_items(nthreads, (size_t)n, [&](size_t k) {
		const FinImage& im = imgs[k];
		const uint64_t np = (uint64_t)im.w * im.h;
		uint32_t* map = (uint32_t*)(host.p + o_map) + s_off[k];
		for (size_t i = 0; i < im.nsamples; ++i) {
			if (im.loc[i] >= np) {
				bad_loc.store(true);
				map[i] = 0;
				continue;
			}
			map[i] = (uint32_t)(inst[k].px_off + im.loc[i]);
		}
		if (!d_rgb)
			memcpy((float*)(host.p + o_lm) + 3 * s_off[k], im.rgb, im.nsamples * 12);
	});
	if (bad_loc.load())
		throw std::runtime_error("finalize: a sample's texel index lies outside its instance's atlas");
	tm.pack_ms = ms_since(t0);

	t0 = Clock::now();
	hipStream_t s;
	FIN_CHECK(hipStreamCreate(&s));
	struct StreamGuard {
		hipStream_t s;
		~StreamGuard() { (void)hipStreamDestroy(s); }
	} sg{s};
	Dev meta(up_bytes), rgb(px * 12), tmp(std::max<uint64_t>(px, dpx) * 12), set0(px), set1(px);
	FIN_CHECK(hipMemcpyAsync(meta.p, host.p, up_bytes, hipMemcpyHostToDevice, s));
	FIN_CHECK(hipStreamSynchronize(s));
	tm.upload_ms = ms_since(t0);

	const FinInst* d_inst = (const FinInst*)(meta.p + o_inst);
	const uint32_t* d_fb = (const uint32_t*)(meta.p + o_fb);
	const uint32_t* d_dfb = (const uint32_t*)(meta.p + o_dfb);
	const float* d_kern = (const float*)(meta.p + o_kern);
	hipEvent_t e0, e1;
	FIN_CHECK(hipEventCreate(&e0));
	FIN_CHECK(hipEventCreate(&e1));
	FIN_CHECK(hipEventRecord(e0, s));
	FIN_CHECK(hipMemsetAsync(rgb.p, 0, px * 12, s));
	FIN_CHECK(hipMemsetAsync(set0.p, 0, px, s));
	if (ns)
		k_fin_scatter<<<(unsigned)((ns + kFinBlock - 1) / kFinBlock), kFinBlock, 0, s>>>(
		    (const uint32_t*)(meta.p + o_map), d_rgb ? d_rgb : (const float*)(meta.p + o_lm), ns, (float*)rgb.p,
		    (uint8_t*)set0.p);
	uint8_t *sa = (uint8_t*)set0.p, *sb = (uint8_t*)set1.p;
	for (int its = 1; its <= 3; ++its) {
		FIN_CHECK(hipMemcpyAsync(sb, sa, px, hipMemcpyDeviceToDevice, s));
		k_fin_extrap<<<(unsigned)blocks, kFinBlock, 0, s>>>(d_inst, d_fb, n, (float*)rgb.p, sa, sb);
		std::swap(sa, sb);
	}
	if (blur_size != 0.0f) {
		k_fin_conv<<<(unsigned)blocks, kFinBlock, 0, s>>>(d_inst, d_fb, n, (const float*)rgb.p, (float*)tmp.p, d_kern,
		                                                  ext, 0);
		k_fin_conv<<<(unsigned)blocks, kFinBlock, 0, s>>>(d_inst, d_fb, n, (const float*)tmp.p, (float*)rgb.p, d_kern,
		                                                  ext, 1);
	}
	const char* final_img = rgb.p;
	if (ds2x) {
		if (dblocks)
			k_fin_ds<<<(unsigned)dblocks, kFinBlock, 0, s>>>(d_inst, d_dfb, n, (const float*)rgb.p, (float*)tmp.p);
		final_img = tmp.p;
	}
	hipError_t le = hipGetLastError();
	FIN_CHECK(hipEventRecord(e1, s));
	hipError_t se = hipStreamSynchronize(s);
	float kms = 0;
	if (le == hipSuccess && se == hipSuccess)
		(void)hipEventElapsedTime(&kms, e0, e1);
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	FIN_CHECK(le);
	FIN_CHECK(se);
	tm.kernels_ms = kms;

	t0 = Clock::now();
	FIN_CHECK(hipMemcpyAsync(host.p, final_img, out_px * 12, hipMemcpyDeviceToHost, s));
	FIN_CHECK(hipStreamSynchronize(s));
	parallel_items(nthreads, (size_t)n, [&](size_t k) {
		const FinImage& im = imgs[k];
		const uint64_t off = ds2x ? inst[k].ds_off : inst[k].px_off;
		memcpy(im.out, (const float*)host.p + 3 * off, (size_t)im.final_w * im.final_h * 12);
	});
	tm.download_ms = ms_since(t0);
	if (times)
		*times = tm;
}

} // namespace lgx
