// The binned-SAH BVH build of bvh.cpp on the GPU, level by level: the same
// binary tree, node for node and leaf for leaf, then the host's collapse.
//
// The host build's decisions depend only on sets, never on the order of the
// triangles inside a range: centroid and box bounds are minima and maxima,
// the bins are counts and boxes, the SAH sweep reads the bins in a fixed
// order, the partition is stable and the fallback split is by position.  A
// level holds every node made by the previous one (ranges of the leaf-order
// array `idx`, sorted by their first position); per level:
//  * k_lvl_bounds: one lane per position (its node found by a binary
//    search of the level's first positions) adds its triangle's box and
//    centroid to the node's bounds, as order-preserving integer keys with
//    atomicMin / atomicMax (exact, so the result does not depend on the
//    order of the atomics); a wave that lies in one node reduces first;
//  * k_lvl_bins: the centroid's bin on every axis with an extent, count
//    and box per bin (in LDS first when a block lies in one node);
//  * k_lvl_sah: one lane per node, the host's SAH sweep on the decoded bins
//    (empty bins read as the reset box), giving the split or the fallback;
//  * k_lvl_flags, a scan of the flags, k_lvl_scatter: the stable partition.
// The host keeps the node list (creation order = the binary node numbering
// of its own build is irrelevant: the collapse follows the references),
// reads each level's decisions and bounds, and makes the next level.
#include <hip/hip_runtime.h>
#include <float.h>
#include <string.h>

#include <chrono>
#include <stdexcept>
#include <string>
#include <vector>

#include "bvh.h"

namespace lgx {

#define BV_CHECK(x)                                                                                          \
	do {                                                                                                     \
		hipError_t e_ = (x);                                                                                 \
		if (e_ != hipSuccess)                                                                                \
			throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " __FILE__ ":" + \
			                         std::to_string(__LINE__) + ": " #x);                                    \
	} while (0)

namespace {

constexpr int kBB = 256;                  // threads per block
constexpr int kNB = kSahBins;             // bins per axis
constexpr int kBinSlots = 3 * kNB;        // (axis, bin) per node
constexpr uint32_t kModeLeaf = 0, kModeSah = 1, kModeIndex = 2;

// order-preserving key of a float (finite values): unsigned comparison of
// the keys orders the floats
__device__ __forceinline__ uint32_t fkey(float f)
{
	const uint32_t u = __float_as_uint(f);
	return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ __forceinline__ float fval(uint32_t k)
{
	const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
	float f;
	memcpy(&f, &u, 4);
	return f;
}

template <class T> struct DArr {
	T* p = nullptr;
	size_t n = 0;
	DArr() = default;
	explicit DArr(size_t cnt) { alloc(cnt); }
	void alloc(size_t cnt)
	{
		if (p)
			(void)hipFree(p);
		p = nullptr;
		n = cnt;
		BV_CHECK(hipMalloc((void**)&p, sizeof(T) * (cnt ? cnt : 1)));
	}
	DArr(const DArr&) = delete;
	DArr& operator=(const DArr&) = delete;
	~DArr()
	{
		if (p)
			(void)hipFree(p);
	}
};

// a ? b : a, b < a: the host build's std::min / std::max, whose +-0 ties do
// not matter here (padded boxes and bin indices take equal values)
__device__ __forceinline__ float mn(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float mx(float a, float b) { return a < b ? b : a; }

// triangle boxes, centroids, the identity order and the scene bounds
__global__ __launch_bounds__(kBB) void k_tri_prep(const float* __restrict__ tris, uint32_t n,
                                                  float* __restrict__ box, float* __restrict__ cen,
                                                  uint32_t* __restrict__ idx, uint32_t* __restrict__ all)
{
	const uint32_t t = blockIdx.x * kBB + threadIdx.x;
	uint32_t k[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
	if (t < n) {
		float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
		for (int v = 0; v < 3; ++v)
			for (int a = 0; a < 3; ++a) {
				const float p = tris[9 * (size_t)t + 3 * v + a];
				lo[a] = mn(lo[a], p);
				hi[a] = mx(hi[a], p);
			}
		for (int a = 0; a < 3; ++a) {
			box[6 * (size_t)t + a] = lo[a];
			box[6 * (size_t)t + 3 + a] = hi[a];
			cen[3 * (size_t)t + a] = 0.5f * (lo[a] + hi[a]);
			k[a] = fkey(lo[a]);
			k[3 + a] = fkey(hi[a]);
		}
		idx[t] = t;
	}
	for (int o = 32; o > 0; o >>= 1)
		for (int c = 0; c < 6; ++c) {
			const uint32_t other = (uint32_t)__shfl_xor((int)k[c], o);
			k[c] = c < 3 ? min(k[c], other) : max(k[c], other);
		}
	if ((threadIdx.x & 63) == 0)
		for (int c = 0; c < 6; ++c)
			c < 3 ? atomicMin(all + c, k[c]) : atomicMax(all + c, k[c]);
}

// the level slot holding position i (first positions sorted), or -1
__device__ __forceinline__ int slot_of(const uint32_t* __restrict__ lb, const uint32_t* __restrict__ le, int m,
                                       uint32_t i)
{
	int lo = 0, hi = m - 1;
	if (m == 0 || i < lb[0])
		return -1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (lb[mid] <= i)
			lo = mid;
		else
			hi = mid - 1;
	}
	return i < le[lo] ? lo : -1;
}

// per slot: 12 keys (centroid lo x/y/z, hi x/y/z, box lo x/y/z, hi x/y/z)
__global__ __launch_bounds__(kBB) void k_lvl_bounds(uint32_t n, const uint32_t* __restrict__ lb,
                                                    const uint32_t* __restrict__ le, int m,
                                                    const uint32_t* __restrict__ idx, const float* __restrict__ box,
                                                    const float* __restrict__ cen, uint32_t* __restrict__ acc)
{
	const uint32_t i = blockIdx.x * kBB + threadIdx.x;
	const int lane = threadIdx.x & 63;
	const int s = i < n ? slot_of(lb, le, m, i) : -1;
	uint32_t k[12];
	if (s >= 0) {
		const uint32_t t = idx[i];
		for (int a = 0; a < 3; ++a) {
			k[a] = k[3 + a] = fkey(cen[3 * (size_t)t + a]);
			k[6 + a] = fkey(box[6 * (size_t)t + a]);
			k[9 + a] = fkey(box[6 * (size_t)t + 3 + a]);
		}
	}
	// a block inside one slot reduces in LDS and makes 12 atomics; else a
	// wave inside one slot reduces by shuffles; else every lane makes its own
	__shared__ int bs[2];
	__shared__ uint32_t red[kBB / 64][12];
	if (threadIdx.x == 0)
		bs[0] = s;
	if (threadIdx.x == kBB - 1)
		bs[1] = s;
	__syncthreads();
	const bool block_uni = bs[0] >= 0 && bs[0] == bs[1];
	const int s0 = __shfl(s, 0);
	const bool wave_uni = __all(s == s0) && s0 >= 0;
	if (wave_uni) {
		for (int o = 32; o > 0; o >>= 1)
			for (int c = 0; c < 12; ++c) {
				const uint32_t other = (uint32_t)__shfl_xor((int)k[c], o);
				k[c] = ((c / 3) % 2 == 0) ? min(k[c], other) : max(k[c], other);
			}
	}
	if (block_uni) { // every wave is then inside the slot too
		if (lane == 0)
			for (int c = 0; c < 12; ++c)
				red[threadIdx.x >> 6][c] = k[c];
		__syncthreads();
		if (threadIdx.x < 12) {
			const int c = threadIdx.x;
			uint32_t v = red[0][c];
			for (int w = 1; w < kBB / 64; ++w)
				v = ((c / 3) % 2 == 0) ? min(v, red[w][c]) : max(v, red[w][c]);
			((c / 3) % 2 == 0) ? atomicMin(acc + 12 * s0 + c, v) : atomicMax(acc + 12 * s0 + c, v);
		}
	} else if (wave_uni) {
		if (lane == 0)
			for (int c = 0; c < 12; ++c)
				((c / 3) % 2 == 0) ? atomicMin(acc + 12 * s0 + c, k[c]) : atomicMax(acc + 12 * s0 + c, k[c]);
	} else if (s >= 0) {
		for (int c = 0; c < 12; ++c)
			((c / 3) % 2 == 0) ? atomicMin(acc + 12 * s + c, k[c]) : atomicMax(acc + 12 * s + c, k[c]);
	}
}

// per slot and (axis, bin): count, then box keys lo x/y/z, hi x/y/z
__device__ __forceinline__ bool bin_setup(const uint32_t* __restrict__ acc, int s, int a, float* lo, float* sc)
{
	*lo = fval(acc[12 * s + a]);
	const float ext = fval(acc[12 * s + 3 + a]) - *lo;
	if (!(ext > 0))
		return false;
	*sc = (float)kNB / ext;
	return true;
}

__device__ __forceinline__ int bin_of(float c, float lo, float sc)
{
	const int k = (int)((c - lo) * sc);
	return k < kNB - 1 ? k : kNB - 1;
}

__global__ __launch_bounds__(kBB) void k_lvl_bins(uint32_t n, const uint32_t* __restrict__ lb,
                                                  const uint32_t* __restrict__ le, int m,
                                                  const uint32_t* __restrict__ idx, const float* __restrict__ box,
                                                  const float* __restrict__ cen, const uint32_t* __restrict__ acc,
                                                  uint32_t* __restrict__ bins)
{
	__shared__ uint32_t sb[kBinSlots * 7];
	__shared__ int uni;
	const uint32_t i = blockIdx.x * kBB + threadIdx.x;
	const int s = i < n ? slot_of(lb, le, m, i) : -1;
	const bool inner = s >= 0 && le[s] - lb[s] > (uint32_t)kLeafMax;
	if (threadIdx.x == 0) {
		const uint32_t last = min(n - 1, blockIdx.x * kBB + kBB - 1);
		const int s0 = slot_of(lb, le, m, blockIdx.x * kBB), s1 = slot_of(lb, le, m, last);
		uni = (s0 >= 0 && s0 == s1 && last == blockIdx.x * kBB + kBB - 1) ? s0 : -1;
	}
	__syncthreads();
	const int u = uni;
	if (u >= 0) {
		for (int q = threadIdx.x; q < kBinSlots * 7; q += kBB)
			sb[q] = (q % 7) == 0 ? 0u : ((q % 7) <= 3 ? 0xffffffffu : 0u);
		__syncthreads();
	}
	if (inner) {
		const uint32_t t = idx[i];
		const size_t row = u >= 0 ? 0 : (size_t)s * kBinSlots;
		for (int a = 0; a < 3; ++a) {
			float lo, sc;
			if (!bin_setup(acc, s, a, &lo, &sc))
				continue;
			const int k = bin_of(cen[3 * (size_t)t + a], lo, sc);
			uint32_t* w = (u >= 0 ? sb : bins) + 7 * (row + (size_t)a * kNB + k);
			atomicAdd(w, 1u);
			for (int c = 0; c < 3; ++c) {
				atomicMin(w + 1 + c, fkey(box[6 * (size_t)t + c]));
				atomicMax(w + 4 + c, fkey(box[6 * (size_t)t + 3 + c]));
			}
		}
	}
	if (u >= 0) {
		__syncthreads();
		for (int q = threadIdx.x; q < kBinSlots; q += kBB) {
			const uint32_t* w = sb + 7 * q;
			if (!w[0])
				continue;
			uint32_t* g = bins + 7 * ((size_t)u * kBinSlots + q);
			atomicAdd(g, w[0]);
			for (int c = 0; c < 3; ++c) {
				atomicMin(g + 1 + c, w[1 + c]);
				atomicMax(g + 4 + c, w[4 + c]);
			}
		}
	}
}

// initial accumulators: per slot 12 keys (minima all ones, maxima zero) and
// per (slot, axis, bin) a count of 0, box minima all ones, maxima zero
__global__ __launch_bounds__(kBB) void k_lvl_init(int m, uint32_t* __restrict__ acc, uint32_t* __restrict__ bins)
{
	const uint64_t q = (uint64_t)blockIdx.x * kBB + threadIdx.x;
	if (q < 12ull * m) {
		const int c = (int)(q % 12);
		acc[q] = ((c / 3) % 2 == 0) ? 0xffffffffu : 0u;
	}
	if (q < 7ull * kBinSlots * m) {
		const int w = (int)(q % 7);
		bins[q] = w == 0 ? 0u : (w <= 3 ? 0xffffffffu : 0u);
	}
}

struct Box3 {
	float lo[3], hi[3];
	__device__ void reset()
	{
		for (int a = 0; a < 3; ++a) {
			lo[a] = FLT_MAX;
			hi[a] = -FLT_MAX;
		}
	}
	__device__ void grow(const Box3& b)
	{
		for (int a = 0; a < 3; ++a) {
			lo[a] = mn(lo[a], b.lo[a]);
			hi[a] = mx(hi[a], b.hi[a]);
		}
	}
	__device__ float area() const
	{
		float d[3];
		for (int a = 0; a < 3; ++a)
			d[a] = mx(0.0f, hi[a] - lo[a]);
		return 2.0f * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
	}
};

// per slot: {mode, axis, split bin, left count}; the sweep is the host
// build's (bvh.cpp, Builder::split)
__global__ __launch_bounds__(kBB) void k_lvl_sah(int m, const uint32_t* __restrict__ lb,
                                                 const uint32_t* __restrict__ le, const int32_t* __restrict__ ld,
                                                 const uint32_t* __restrict__ acc, const uint32_t* __restrict__ bins,
                                                 uint4* __restrict__ dec)
{
	const int s = blockIdx.x * kBB + threadIdx.x;
	if (s >= m)
		return;
	const uint32_t n = le[s] - lb[s];
	if (n <= (uint32_t)kLeafMax) {
		dec[s] = make_uint4(kModeLeaf, 0, 0, 0);
		return;
	}
	float best_cost = FLT_MAX;
	int best_axis = -1, best_split = -1;
	for (int a = 0; a < 3; ++a) {
		float lo, sc;
		if (!bin_setup(acc, s, a, &lo, &sc))
			continue;
		const uint32_t* B = bins + 7 * ((size_t)s * kBinSlots + (size_t)a * kNB);
		auto bin_box = [&](int k) {
			Box3 b;
			const uint32_t* w = B + 7 * k;
			if (!w[0]) {
				b.reset();
			} else {
				for (int c = 0; c < 3; ++c) {
					b.lo[c] = fval(w[1 + c]);
					b.hi[c] = fval(w[4 + c]);
				}
			}
			return b;
		};
		float ra[kNB];
		uint32_t rc[kNB];
		Box3 acc3;
		acc3.reset();
		uint32_t c = 0;
		for (int k = kNB - 1; k > 0; --k) {
			acc3.grow(bin_box(k));
			c += B[7 * k];
			ra[k] = acc3.area();
			rc[k] = c;
		}
		acc3.reset();
		c = 0;
		for (int k = 0; k < kNB - 1; ++k) {
			acc3.grow(bin_box(k));
			c += B[7 * k];
			if (c == 0 || rc[k + 1] == 0)
				continue;
			const float cost = acc3.area() * c + ra[k + 1] * rc[k + 1];
			if (cost < best_cost) {
				best_cost = cost;
				best_axis = a;
				best_split = k;
			}
		}
	}
	uint32_t mode = kModeIndex, left = n / 2;
	if (ld[s] < kSahMaxDepth && best_axis >= 0) {
		const uint32_t* B = bins + 7 * ((size_t)s * kBinSlots + (size_t)best_axis * kNB);
		uint32_t c = 0;
		for (int k = 0; k <= best_split; ++k)
			c += B[7 * k];
		if (c != 0 && c != n) {
			mode = kModeSah;
			left = c;
		}
	}
	dec[s] = make_uint4(mode, (uint32_t)(best_axis < 0 ? 0 : best_axis), (uint32_t)(best_split < 0 ? 0 : best_split),
	                    left);
}

__device__ __forceinline__ bool goes_left(uint32_t i, int s, const uint32_t* __restrict__ lb,
                                          const uint32_t* __restrict__ le, const uint4* __restrict__ dec,
                                          const uint32_t* __restrict__ acc, const uint32_t* __restrict__ idx,
                                          const float* __restrict__ cen)
{
	const uint4 d = dec[s];
	if (d.x == kModeIndex)
		return i - lb[s] < (le[s] - lb[s]) / 2;
	float lo, sc;
	bin_setup(acc, s, (int)d.y, &lo, &sc);
	return bin_of(cen[3 * (size_t)idx[i] + d.y], lo, sc) <= (int)d.z;
}

__global__ __launch_bounds__(kBB) void k_lvl_flags(uint32_t n, const uint32_t* __restrict__ lb,
                                                   const uint32_t* __restrict__ le, int m,
                                                   const uint4* __restrict__ dec, const uint32_t* __restrict__ acc,
                                                   const uint32_t* __restrict__ idx, const float* __restrict__ cen,
                                                   uint32_t* __restrict__ flag)
{
	const uint32_t i = blockIdx.x * kBB + threadIdx.x;
	if (i >= n)
		return;
	const int s = slot_of(lb, le, m, i);
	flag[i] = (s >= 0 && dec[s].x != kModeLeaf && goes_left(i, s, lb, le, dec, acc, idx, cen)) ? 1u : 0u;
}

// exclusive scan of n flags in blocks of kBB: block sums, then their scan
// (one block, serial chunks), then the block offsets added
__global__ __launch_bounds__(kBB) void k_scan_local(uint32_t n, const uint32_t* __restrict__ in,
                                                    uint32_t* __restrict__ out, uint32_t* __restrict__ bsum)
{
	__shared__ uint32_t sh[kBB];
	const uint32_t i = blockIdx.x * kBB + threadIdx.x;
	const uint32_t v = i < n ? in[i] : 0u;
	sh[threadIdx.x] = v;
	__syncthreads();
	for (int o = 1; o < kBB; o <<= 1) {
		const uint32_t add = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0u;
		__syncthreads();
		sh[threadIdx.x] += add;
		__syncthreads();
	}
	if (i < n)
		out[i] = sh[threadIdx.x] - v;
	if (threadIdx.x == kBB - 1)
		bsum[blockIdx.x] = sh[kBB - 1];
}

__global__ __launch_bounds__(kBB) void k_scan_sums(uint32_t nb, uint32_t* __restrict__ bsum, uint32_t* __restrict__ total)
{
	__shared__ uint32_t part[kBB];
	const uint32_t per = (nb + kBB - 1) / kBB, b0 = threadIdx.x * per, b1 = min(nb, b0 + per);
	uint32_t s = 0;
	for (uint32_t b = b0; b < b1; ++b)
		s += bsum[b];
	part[threadIdx.x] = s;
	__syncthreads();
	if (threadIdx.x == 0) {
		uint32_t run = 0;
		for (int t = 0; t < kBB; ++t) {
			const uint32_t x = part[t];
			part[t] = run;
			run += x;
		}
		*total = run;
	}
	__syncthreads();
	uint32_t run = part[threadIdx.x];
	for (uint32_t b = b0; b < b1; ++b) {
		const uint32_t x = bsum[b];
		bsum[b] = run;
		run += x;
	}
}

__global__ __launch_bounds__(kBB) void k_scan_add(uint32_t n, uint32_t* __restrict__ out,
                                                  const uint32_t* __restrict__ bsum)
{
	const uint32_t i = blockIdx.x * kBB + threadIdx.x;
	if (i < n)
		out[i] += bsum[blockIdx.x];
}

__global__ __launch_bounds__(kBB) void k_lvl_scatter(uint32_t n, const uint32_t* __restrict__ lb,
                                                     const uint32_t* __restrict__ le, int m,
                                                     const uint4* __restrict__ dec, const uint32_t* __restrict__ flag,
                                                     const uint32_t* __restrict__ pre, const uint32_t* __restrict__ idx,
                                                     uint32_t* __restrict__ idx2, uint32_t* __restrict__ bad)
{
	const uint32_t i = blockIdx.x * kBB + threadIdx.x;
	if (i >= n)
		return;
	const int s = slot_of(lb, le, m, i);
	if (s < 0 || dec[s].x == kModeLeaf) {
		idx2[i] = idx[i];
		return;
	}
	const uint32_t b = lb[s], e = le[s], left = dec[s].w;
	const uint32_t l = pre[i] - pre[b];
	if (i == b) { // the flags' count must equal the bins' (the same predicate)
		const uint32_t cnt = (e < n ? pre[e] : pre[n - 1] + flag[n - 1]) - pre[b];
		if (cnt != left)
			atomicAdd(bad, 1u);
	}
	const uint32_t dst = flag[i] ? b + l : b + left + ((i - b) - l);
	if (dst < b || dst >= e) { // only if flags and bins disagreed (reported above)
		atomicAdd(bad, 1u);
		return;
	}
	idx2[dst] = idx[i];
}

// per slot: its node's box into the node table, and 1 for an inner node
__global__ __launch_bounds__(kBB) void k_lvl_nodes(int m, const int32_t* __restrict__ lid,
                                                   const uint32_t* __restrict__ acc, const uint4* __restrict__ dec,
                                                   float* __restrict__ nbox, uint32_t* __restrict__ inner)
{
	const int s = blockIdx.x * kBB + threadIdx.x;
	if (s >= m)
		return;
	const int32_t id = lid[s];
	for (int c = 0; c < 6; ++c)
		nbox[6 * (size_t)id + c] = fval(acc[12 * (size_t)s + 6 + c]);
	inner[s] = dec[s].x != kModeLeaf ? 1u : 0u;
}

// per inner slot of rank r (exclusive scan of the inner flags): children
// base + 2 r and base + 2 r + 1 in the node table and at 2 r, 2 r + 1 of the
// next level (which stays sorted by first position)
__global__ __launch_bounds__(kBB) void k_lvl_children(int m, uint32_t base, const int32_t* __restrict__ lid,
                                                      const uint32_t* __restrict__ lb,
                                                      const uint32_t* __restrict__ le,
                                                      const int32_t* __restrict__ ld, const uint4* __restrict__ dec,
                                                      const uint32_t* __restrict__ rank, int2* __restrict__ nchild,
                                                      uint32_t* __restrict__ nlb, uint32_t* __restrict__ nle,
                                                      int32_t* __restrict__ nid, uint32_t* __restrict__ nb2,
                                                      uint32_t* __restrict__ ne2, int32_t* __restrict__ nd2)
{
	const int s = blockIdx.x * kBB + threadIdx.x;
	if (s >= m || dec[s].x == kModeLeaf)
		return;
	const uint32_t r = rank[s], c = base + 2 * r, mid = lb[s] + dec[s].w;
	nchild[lid[s]] = make_int2((int)c, (int)c + 1);
	nlb[c] = lb[s];
	nle[c] = mid;
	nlb[c + 1] = mid;
	nle[c + 1] = le[s];
	nid[2 * r] = (int32_t)c;
	nid[2 * r + 1] = (int32_t)c + 1;
	nb2[2 * r] = lb[s];
	ne2[2 * r] = mid;
	nb2[2 * r + 1] = mid;
	ne2[2 * r + 1] = le[s];
	nd2[2 * r] = nd2[2 * r + 1] = ld[s] + 1;
}

} // namespace

BvhBuild build_bvh_gpu(const float* d_tris, size_t ntri, uint32_t* d_order, void* stream, double* kernel_ms)
{
	hipStream_t s = (hipStream_t)stream;
	BvhBuild out;
	if (kernel_ms)
		*kernel_ms = 0;
	if (ntri == 0) {
		out.lo = mk3(0, 0, 0);
		out.hi = mk3(0, 0, 0);
		return out;
	}
	if (ntri >= (1u << 27))
		throw std::runtime_error("too many triangles for the leaf encoding");
	const uint32_t n = (uint32_t)ntri;
	const unsigned grid = (unsigned)((ntri + kBB - 1) / kBB);
	hipEvent_t e0, e1;
	BV_CHECK(hipEventCreate(&e0));
	BV_CHECK(hipEventCreate(&e1));
	struct EvGuard {
		hipEvent_t a, b;
		~EvGuard()
		{
			(void)hipEventDestroy(a);
			(void)hipEventDestroy(b);
		}
	} eg{e0, e1};
	DArr<float> box(6 * ntri), cen(3 * ntri);
	DArr<uint32_t> idx2(ntri), all(6), flag(ntri), pre(ntri), bsum(grid), total(1), bad(1);
	uint32_t* idx = d_order;
	BV_CHECK(hipEventRecord(e0, s));
	const uint32_t init_all[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
	BV_CHECK(hipMemcpyAsync(all.p, init_all, sizeof(init_all), hipMemcpyHostToDevice, s));
	BV_CHECK(hipMemsetAsync(bad.p, 0, 4, s));
	k_tri_prep<<<grid, kBB, 0, s>>>(d_tris, n, box.p, cen.p, idx, all.p);
	BV_CHECK(hipGetLastError());
	uint32_t allk[6];
	BV_CHECK(hipMemcpyAsync(allk, all.p, sizeof(allk), hipMemcpyDeviceToHost, s));
	BV_CHECK(hipStreamSynchronize(s));
	float lo[3], hi[3];
	for (int a = 0; a < 3; ++a) {
		lo[a] = fval(allk[a]);
		hi[a] = fval(allk[3 + a]);
	}
	out.lo = mk3(lo[0], lo[1], lo[2]);
	out.hi = mk3(hi[0], hi[1], hi[2]);
	std::vector<BvhNode> bin;
	if (ntri <= (size_t)kLeafMax) {
		// as the host build: one leaf and an inverted second child
		const float ilo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, ihi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
		bin.push_back(bvh_node(lo, hi, leaf_ref(0, n), ilo, ihi, leaf_ref(0, 1)));
		// the inverted box stays unpadded, as the host build leaves it
		bin[0].lo1x = bin[0].lo1y = bin[0].lo1z = FLT_MAX;
		bin[0].hi1x = bin[0].hi1y = bin[0].hi1z = -FLT_MAX;
	} else {
		// the node table and the levels live on the device; per level the
		// host reads only the number of inner nodes
		const size_t cap = 2 * ntri; // nodes of a binary tree over ntri leaves' ranges
		DArr<uint32_t> t_lb(cap), t_le(cap);
		DArr<int2> t_child(cap);
		DArr<float> t_box(6 * cap);
		DArr<int32_t> l_id(ntri), l_id2(ntri), l_d(ntri), l_d2(ntri);
		DArr<uint32_t> l_b(ntri), l_b2(ntri), l_e(ntri), l_e2(ntri), l_inner(ntri), l_rank(ntri);
		DArr<uint32_t> d_acc(12 * ntri), d_bins;
		DArr<uint4> d_dec(ntri);
		size_t bins_cap = 0;
		BV_CHECK(hipMemsetAsync(t_child.p, 0xff, sizeof(int2) * cap, s)); // -1: no children
		const int32_t zero = 0;
		BV_CHECK(hipMemcpyAsync(l_id.p, &zero, 4, hipMemcpyHostToDevice, s));
		BV_CHECK(hipMemcpyAsync(l_d.p, &zero, 4, hipMemcpyHostToDevice, s));
		BV_CHECK(hipMemsetAsync(l_b.p, 0, 4, s));
		BV_CHECK(hipMemcpyAsync(l_e.p, &n, 4, hipMemcpyHostToDevice, s));
		BV_CHECK(hipMemsetAsync(t_lb.p, 0, 4, s));
		BV_CHECK(hipMemcpyAsync(t_le.p, &n, 4, hipMemcpyHostToDevice, s));
		uint32_t nnodes = 1;
		int m = 1;
		while (m > 0) {
			if ((size_t)m > bins_cap) {
				bins_cap = std::min<size_t>(ntri, std::max<size_t>((size_t)m, 2 * bins_cap));
				d_bins.alloc((size_t)7 * kBinSlots * bins_cap);
			}
			const unsigned mg = (unsigned)((m + kBB - 1) / kBB);
			k_lvl_init<<<(unsigned)((7ull * kBinSlots * m + kBB - 1) / kBB), kBB, 0, s>>>(m, d_acc.p, d_bins.p);
			k_lvl_bounds<<<grid, kBB, 0, s>>>(n, l_b.p, l_e.p, m, idx, box.p, cen.p, d_acc.p);
			k_lvl_bins<<<grid, kBB, 0, s>>>(n, l_b.p, l_e.p, m, idx, box.p, cen.p, d_acc.p, d_bins.p);
			k_lvl_sah<<<mg, kBB, 0, s>>>(m, l_b.p, l_e.p, l_d.p, d_acc.p, d_bins.p, d_dec.p);
			k_lvl_flags<<<grid, kBB, 0, s>>>(n, l_b.p, l_e.p, m, d_dec.p, d_acc.p, idx, cen.p, flag.p);
			k_scan_local<<<grid, kBB, 0, s>>>(n, flag.p, pre.p, bsum.p);
			k_scan_sums<<<1, kBB, 0, s>>>(grid, bsum.p, total.p);
			k_scan_add<<<grid, kBB, 0, s>>>(n, pre.p, bsum.p);
			k_lvl_scatter<<<grid, kBB, 0, s>>>(n, l_b.p, l_e.p, m, d_dec.p, flag.p, pre.p, idx, idx2.p, bad.p);
			BV_CHECK(hipGetLastError());
			BV_CHECK(hipMemcpyAsync(idx, idx2.p, 4 * ntri, hipMemcpyDeviceToDevice, s));
			// the level's boxes, inner flags, their ranks, the next level
			k_lvl_nodes<<<mg, kBB, 0, s>>>(m, l_id.p, d_acc.p, d_dec.p, t_box.p, l_inner.p);
			k_scan_local<<<mg, kBB, 0, s>>>((uint32_t)m, l_inner.p, l_rank.p, bsum.p);
			k_scan_sums<<<1, kBB, 0, s>>>(mg, bsum.p, total.p);
			k_scan_add<<<mg, kBB, 0, s>>>((uint32_t)m, l_rank.p, bsum.p);
			k_lvl_children<<<mg, kBB, 0, s>>>(m, nnodes, l_id.p, l_b.p, l_e.p, l_d.p, d_dec.p, l_rank.p, t_child.p,
			                                  t_lb.p, t_le.p, l_id2.p, l_b2.p, l_e2.p, l_d2.p);
			BV_CHECK(hipGetLastError());
			uint32_t ninner = 0;
			BV_CHECK(hipMemcpyAsync(&ninner, total.p, 4, hipMemcpyDeviceToHost, s));
			BV_CHECK(hipStreamSynchronize(s));
			if ((size_t)nnodes + 2 * (size_t)ninner > cap)
				throw std::runtime_error("build_bvh_gpu: node table overflow");
			nnodes += 2 * ninner;
			m = (int)(2 * ninner);
			std::swap(l_id.p, l_id2.p);
			std::swap(l_b.p, l_b2.p);
			std::swap(l_e.p, l_e2.p);
			std::swap(l_d.p, l_d2.p);
		}
		uint32_t nbad = 0;
		BV_CHECK(hipMemcpyAsync(&nbad, bad.p, 4, hipMemcpyDeviceToHost, s));
		std::vector<uint32_t> hb(nnodes), he(nnodes);
		std::vector<int2> hc(nnodes);
		std::vector<float> hbox(6 * (size_t)nnodes);
		BV_CHECK(hipMemcpyAsync(hb.data(), t_lb.p, 4 * (size_t)nnodes, hipMemcpyDeviceToHost, s));
		BV_CHECK(hipMemcpyAsync(he.data(), t_le.p, 4 * (size_t)nnodes, hipMemcpyDeviceToHost, s));
		BV_CHECK(hipMemcpyAsync(hc.data(), t_child.p, sizeof(int2) * nnodes, hipMemcpyDeviceToHost, s));
		BV_CHECK(hipMemcpyAsync(hbox.data(), t_box.p, 4 * hbox.size(), hipMemcpyDeviceToHost, s));
		BV_CHECK(hipStreamSynchronize(s));
		if (nbad)
			throw std::runtime_error("build_bvh_gpu: a partition's flags disagree with its bins");
		// binary nodes: the inner nodes in creation order (the root first)
		std::vector<int32_t> inner_id(nnodes, -1);
		int32_t ni = 0;
		for (uint32_t q = 0; q < nnodes; ++q)
			if (hc[q].x >= 0)
				inner_id[q] = ni++;
		bin.resize(ni);
		for (uint32_t q = 0; q < nnodes; ++q) {
			if (inner_id[q] < 0)
				continue;
			const int32_t ch[2] = {hc[q].x, hc[q].y};
			int32_t ref[2];
			for (int c = 0; c < 2; ++c)
				ref[c] = hc[ch[c]].x >= 0 ? inner_id[ch[c]] : leaf_ref(hb[ch[c]], he[ch[c]] - hb[ch[c]]);
			const float* L = &hbox[6 * (size_t)ch[0]];
			const float* R = &hbox[6 * (size_t)ch[1]];
			bin[inner_id[q]] = bvh_node(L, L + 3, ref[0], R, R + 3, ref[1]);
		}
	}
	BV_CHECK(hipEventRecord(e1, s));
	BV_CHECK(hipEventSynchronize(e1));
	if (kernel_ms) {
		float ms = 0;
		BV_CHECK(hipEventElapsedTime(&ms, e0, e1));
		*kernel_ms = ms;
	}
	collapse_bvh(bin, out);
	out.order.resize(ntri);
	BV_CHECK(hipMemcpy(out.order.data(), idx, 4 * ntri, hipMemcpyDeviceToHost));
	return out;
}

void compare_bvh_builds(const float* tris, size_t ntri, double out[6])
{
	using Clock = std::chrono::steady_clock;
	auto t0 = Clock::now();
	const BvhBuild h = build_bvh(tris, ntri);
	const double host_ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
	hipStream_t s = nullptr;
	BV_CHECK(hipStreamCreate(&s));
	struct StreamGuard {
		hipStream_t s;
		~StreamGuard() { (void)hipStreamDestroy(s); }
	} sg{s};
	DArr<float> t(9 * ntri);
	DArr<uint32_t> o(ntri);
	if (ntri)
		BV_CHECK(hipMemcpy(t.p, tris, 36 * ntri, hipMemcpyHostToDevice));
	t0 = Clock::now();
	double kms = 0;
	const BvhBuild d = build_bvh_gpu(t.p, ntri, o.p, s, &kms);
	const double gpu_ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
	out[0] = h.nodes.size() == d.nodes.size() &&
	         (h.nodes.empty() || !memcmp(h.nodes.data(), d.nodes.data(), sizeof(WideNode) * h.nodes.size()));
	out[1] = h.order == d.order && h.depth == d.depth;
	out[2] = host_ms;
	out[3] = gpu_ms;
	out[4] = kms;
	out[5] = (double)d.nodes.size();
}

} // namespace lgx
