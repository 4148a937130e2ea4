// Sample placement kernel for gfx950; see place.h.
//
// One lane per rasterised texel.  A lane runs the whole placement of its
// sample: OffsetSample against every instance whose tree root the search
// cube meets (lighter.cpp:444-446), then up to 100 correction steps, each a
// closest hit with normal through the instance tree and the instances'
// split trees (lighter.cpp:448-465, 210-251).  The sample moves while a
// tree is walked, so the walks keep the host's orders exactly:
//  * BoxTree: node n, then its item list, then children n + 1 before ch
//    (a stack, ch pushed first);
//  * SplitTree: the near side first; the far side only when the segment
//    reaches it and the near side found no hit below 1, which a stack of
//    node indices and "result 2" markers reproduces (the value of the last
//    subtree searched is the tree's result, and a hit below 1 ends the walk).
// Stacks are per lane and sized from the trees' depths, checked on the host
// before the launch.  This is a preprocessing step that runs once per bake;
// its cost is dominated by divergent tree walks, not by memory traffic.
#include <hip/hip_runtime.h>

#include <chrono>
#include <stdexcept>
#include <string>

#include "place.h"

namespace lgx {

#define PLACE_CHECK(x)                                                                                       \
	do {                                                                                                     \
		hipError_t e_ = (x);                                                                                 \
		if (e_ != hipSuccess)                                                                                \
			throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " __FILE__ ":" + \
			                         std::to_string(__LINE__) + ": " #x);                                    \
	} while (0)

constexpr int kPlaceBlock = 256;
// Stack entries a walk needs equal the tree's depth (nodes on the longest
// root-to-leaf path): BoxTrees stop splitting at depth 16 and SplitTrees at
// 32 (lighter_math.cpp:654, 430-435), so 17 and 33 suffice; both are
// checked per tree before the launch.
constexpr int kBoxStack = 24;
constexpr int kSplitStack = 40;

struct PlaceTri {
	f3 p1, p2, p3;
	f3 mn, mx; // the box the tree was built from
};

struct PlaceInst {
	int32_t node_off, item_off, tri_off, split_off;
	int32_t stri_off, nsplit, shadow, pad;
	f3 mn, mx; // root box of the instance's BoxTree
};

struct PlaceView {
	const BoxTree::Node* __restrict__ nodes;
	const int32_t* __restrict__ items;
	const PlaceTri* __restrict__ tris;
	const FlatSplitNode* __restrict__ snodes;
	const FlatSplitTri* __restrict__ stris;
	const PlaceInst* __restrict__ inst;
	const BoxTree::Node* __restrict__ inodes; // instance tree
	const int32_t* __restrict__ iitems;
	int32_t ninst;
};

// TriSet::ray_dist (lighter_math.cpp:835-871): closest hit on from -> to in
// one instance, with the index of the triangle that gave it.
__host__ __device__ float pl_ray_dist(const PlaceView& V, const PlaceInst& I, f3 from, f3 to, int32_t* tid)
{
	RaySeg r;
	r.set(from, to);
	float closest = 2.0f;
	int32_t hit = -1;
	int32_t st[kBoxStack];
	int sp = 0;
	st[sp++] = 0;
	while (sp) {
		const int32_t n = st[--sp];
		const BoxTree::Node N = V.nodes[I.node_off + n];
		if (box_disjoint(r.qmn, r.qmx, N.mn, N.mx) || !ray_box(r, N.mn, N.mx))
			continue;
		if (N.ido != -1) {
			const int32_t* it = V.items + I.item_off + N.ido;
			const int32_t c = it[0];
			for (int32_t i = 0; i < c; ++i) {
				const int32_t id = it[1 + i];
				const PlaceTri& T = V.tris[I.tri_off + id];
				float d = seg_tri(r.origin, r.end, T.p1, T.p2, T.p3);
				if (d < closest) {
					closest = d;
					hit = id;
				}
			}
		}
		if (N.ch != -1) {
			st[sp++] = N.ch;
			st[sp++] = n + 1;
		}
	}
	if (hit != -1)
		*tid = hit;
	return closest;
}

// TriSet::offset_sample (lighter_math.cpp:991-1044).
__host__ __device__ void pl_offset(const PlaceView& V, const PlaceInst& I, f3& P, f3 N, float dist)
{
	const f3 d3 = mk3(dist, dist, dist);
	const f3 qmn = P - d3, qmx = P + d3;
	int32_t st[kBoxStack];
	int sp = 0;
	st[sp++] = 0;
	while (sp) {
		const int32_t n = st[--sp];
		const BoxTree::Node B = V.nodes[I.node_off + n];
		if (box_disjoint(qmn, qmx, B.mn, B.mx))
			continue;
		if (B.ido != -1) {
			const int32_t* it = V.items + I.item_off + B.ido;
			const int32_t c = it[0];
			for (int32_t i = 0; i < c; ++i) {
				const int32_t id = it[1 + i];
				const PlaceTri& T = V.tris[I.tri_off + id];
				if (offset_box_skip(P, dist, T.mn, T.mx))
					continue;
				f3 from, pnew;
				if (!offset_candidate(P, N, dist, T.p1, T.p2, T.p3, &from, &pnew))
					continue;
				int32_t tid = -1;
				if (pl_ray_dist(V, I, from, pnew, &tid) >= 0.9f || tid == id)
					P = pnew;
			}
		}
		if (B.ch != -1) {
			st[sp++] = B.ch;
			st[sp++] = n + 1;
		}
	}
}

// SplitTree::intersect (lighter_math.cpp:490-533).
__host__ __device__ float pl_split(const PlaceView& V, const PlaceInst& I, f3 from, f3 to, f3* outnormal)
{
	if (I.nsplit == 0)
		return 2.0f;
	const FlatSplitNode* nodes = V.snodes + I.split_off;
	const FlatSplitTri* tris = V.stris + I.stri_off;
	const f3 dir = to - from;
	int32_t st[kSplitStack];
	int sp = 0;
	st[sp++] = 0;
	float last = 2.0f;
	while (sp) {
		const int32_t e = st[--sp];
		if (e < 0) { // the far side is not searched: this subtree gives 2
			last = 2.0f;
			continue;
		}
		const FlatSplitNode n = nodes[e];
		if (n.front != -1) {
			const f3 pn = mk3(n.nx, n.ny, n.nz);
			const float df = dot(from, pn) - n.d;
			const float dt = dot(to, pn) - n.d;
			if (df < 0) {
				st[sp++] = dt > 0 ? n.front : -1;
				st[sp++] = n.back;
			} else {
				st[sp++] = dt < 0 ? n.back : -1;
				st[sp++] = n.front;
			}
			continue;
		}
		float closest = 2.0f;
		int32_t ct = -1;
		for (int32_t k = 0; k < n.leaf_cnt; ++k) {
			const float hit = seg_tri_pre(from, dir, tris[n.leaf_off + k].rt);
			if (hit < closest) {
				closest = hit;
				ct = k;
			}
		}
		if (ct != -1) {
			const FlatSplitTri& t = tris[n.leaf_off + ct];
			*outnormal = mk3(t.nx, t.ny, t.nz);
		}
		last = closest;
		if (last < 1.0f)
			return last;
	}
	return last;
}

// Scene-level closest hit with normal (lighter.cpp:210-251): the segment
// shrunk by SMALL_FLOAT at both ends, through the instance tree to the
// split trees of the shadow-casting instances.
__host__ __device__ float pl_scene_hit(const PlaceView& V, f3 A, f3 B, f3* outnormal)
{
	const f3 dn = normalized(B - A);
	const f3 mA = A + dn * kSmall;
	const f3 mB = B - dn * kSmall;
	RaySeg r;
	r.set(mA, mB);
	float closest = 2.0f;
	int32_t st[kBoxStack];
	int sp = 0;
	st[sp++] = 0;
	while (sp) {
		const int32_t n = st[--sp];
		const BoxTree::Node N = V.inodes[n];
		if (box_disjoint(r.qmn, r.qmx, N.mn, N.mx) || !ray_box(r, N.mn, N.mx))
			continue;
		if (N.ido != -1) {
			const int32_t* it = V.iitems + N.ido;
			const int32_t c = it[0];
			for (int32_t i = 0; i < c; ++i) {
				const PlaceInst I = V.inst[it[1 + i]];
				if (!I.shadow)
					continue;
				// an instance whose own box misses the segment's padded box
				// cannot be hit
				if (box_disjoint(r.qmn, r.qmx, I.mn, I.mx))
					continue;
				f3 nrm = mk3(0, 0, 0);
				const float d = pl_split(V, I, r.origin, r.end, &nrm);
				if (d < closest) {
					closest = d;
					*outnormal = nrm;
				}
			}
		}
		if (N.ch != -1) {
			st[sp++] = N.ch;
			st[sp++] = n + 1;
		}
	}
	return closest;
}

// The placement of one sample (lighter.cpp:441-465).
__host__ __device__ f3 pl_place(const PlaceView& V, f3 P, f3 N, float d, float max_corr, float corr_min_dot)
{
	// move the sample out of concave edges; an instance whose tree root the
	// cube around the current P misses contributes nothing
	const f3 d3 = mk3(d, d, d);
	for (int32_t k = 0; k < V.ninst; ++k) {
		const PlaceInst I = V.inst[k];
		if (box_disjoint(P - d3, P + d3, I.mn, I.mx))
			continue;
		pl_offset(V, I, P, N, d);
	}
	// move the sample in front of overlapping faces
	if (max_corr != 0.0f) {
		int left = 100;
		float md = max_corr;
		const f3 PEnd = P + N * md;
		while (md > kSmall && left-- > 0) {
			f3 hn = mk3(0, 0, 0);
			const float q = pl_scene_hit(V, P, PEnd, &hn);
			if (!correction_step(P, PEnd, N, md, q, hn, corr_min_dot))
				break;
		}
	}
	return P;
}

__global__ __launch_bounds__(kPlaceBlock) void k_place(PlaceView V, float* __restrict__ pos,
                                                       const float* __restrict__ nrm,
                                                       const float* __restrict__ dist, uint32_t n, float max_corr,
                                                       float corr_min_dot)
{
	const uint32_t i = blockIdx.x * kPlaceBlock + threadIdx.x;
	if (i >= n)
		return;
	const f3 P = pl_place(V, ld3(pos + 3 * (size_t)i), ld3(nrm + 3 * (size_t)i), dist[i], max_corr, corr_min_dot);
	st3(pos + 3 * (size_t)i, P);
}

// ===========================================================================
// host side
// ===========================================================================
namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0) { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

template <class T> struct DevBuf {
	T* p = nullptr;
	DevBuf() = default;
	DevBuf(const DevBuf&) = delete;
	DevBuf& operator=(const DevBuf&) = delete;
	~DevBuf()
	{
		if (p)
			(void)hipFree(p);
	}
	void upload(const T* src, size_t n)
	{
		PLACE_CHECK(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
		if (n)
			PLACE_CHECK(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
	}
	void upload(const std::vector<T>& v) { upload(v.data(), v.size()); }
};

void check_depth(int depth, int cap, const char* what)
{
	if (depth > cap)
		throw std::runtime_error(std::string("sample placement: a ") + what + " of depth " + std::to_string(depth) +
		                         " exceeds the kernel's stack of " + std::to_string(cap));
}

// The per-instance trees and triangles in shared arrays, instance ranges
// in PlaceInst.
struct PlaceData {
	std::vector<PlaceInst> inst;
	std::vector<BoxTree::Node> nodes;
	std::vector<int32_t> items;
	std::vector<PlaceTri> tris;
	std::vector<FlatSplitNode> snodes;
	std::vector<FlatSplitTri> stris;
};

void flatten(const std::vector<PlaceInstance>& insts, const BoxTree& inst_tree, PlaceData& D)
{
	// flatten: per-instance ranges into shared arrays, sizes first
	D.inst.resize(insts.size());
	std::vector<PlaceInst>& inst = D.inst;
	size_t nn = 0, ni = 0, nt = 0;
	for (size_t k = 0; k < insts.size(); ++k) {
		const BoxTree& bt = insts[k].tris->tree();
		check_depth(bt.depth(), kBoxStack, "triangle tree");
		PlaceInst& I = inst[k];
		I.node_off = (int32_t)nn;
		I.item_off = (int32_t)ni;
		I.tri_off = (int32_t)nt;
		I.shadow = insts[k].shadow ? 1 : 0;
		I.pad = 0;
		I.mn = bt.root().mn;
		I.mx = bt.root().mx;
		nn += bt.nodes().size();
		ni += bt.items().size();
		nt += insts[k].tris->tris().size();
	}
	if (nn > 0x7FFFFFFF || ni > 0x7FFFFFFF || nt > 0x7FFFFFFF)
		throw std::runtime_error("sample placement: collision data beyond 2^31 entries");
	D.nodes.resize(nn);
	D.items.resize(ni);
	D.tris.resize(nt);
	std::vector<BoxTree::Node>& nodes = D.nodes;
	std::vector<int32_t>& items = D.items;
	std::vector<PlaceTri>& tris = D.tris;
	std::vector<FlatSplitNode>& snodes = D.snodes;
	std::vector<FlatSplitTri>& stris = D.stris;
	for (size_t k = 0; k < insts.size(); ++k) {
		const TriSet& ts = *insts[k].tris;
		const BoxTree& bt = ts.tree();
		PlaceInst& I = inst[k];
		std::copy(bt.nodes().begin(), bt.nodes().end(), nodes.begin() + I.node_off);
		std::copy(bt.items().begin(), bt.items().end(), items.begin() + I.item_off);
		const std::vector<Tri>& tv = ts.tris();
		const std::vector<Box>& bv = ts.boxes();
		for (size_t j = 0; j < tv.size(); ++j)
			tris[I.tri_off + j] = PlaceTri{tv[j].p1, tv[j].p2, tv[j].p3, bv[j].mn, bv[j].mx};
		I.split_off = (int32_t)snodes.size();
		I.stri_off = (int32_t)stris.size();
		const int sd = insts[k].split->flatten(snodes, stris);
		check_depth(sd, kSplitStack, "split tree");
		I.nsplit = (int32_t)(snodes.size() - I.split_off);
	}
	if (snodes.size() > 0x7FFFFFFF || stris.size() > 0x7FFFFFFF)
		throw std::runtime_error("sample placement: split trees beyond 2^31 entries");
	check_depth(inst_tree.depth(), kBoxStack, "instance tree");
}

} // namespace

void place_samples_gpu(const std::vector<PlaceInstance>& insts, const BoxTree& inst_tree, float* pos,
                       const float* nrm, const float* dist, size_t n, float max_corr, float corr_min_dot,
                       PlaceTimes* times)
{
	PlaceTimes tm;
	tm.samples = n;
	if (n == 0) {
		if (times)
			*times = tm;
		return;
	}
	if (n > 0xFFFFFFFFull)
		throw std::runtime_error("sample placement: more than 2^32 samples");
	auto t0 = Clock::now();
	PlaceData D;
	flatten(insts, inst_tree, D);
	tm.flatten_ms = ms_since(t0);

	t0 = Clock::now();
	DevBuf<BoxTree::Node> d_nodes, d_inodes;
	DevBuf<int32_t> d_items, d_iitems;
	DevBuf<PlaceTri> d_tris;
	DevBuf<FlatSplitNode> d_snodes;
	DevBuf<FlatSplitTri> d_stris;
	DevBuf<PlaceInst> d_inst;
	DevBuf<float> d_pos, d_nrm, d_dist;
	d_nodes.upload(D.nodes);
	d_items.upload(D.items);
	d_tris.upload(D.tris);
	d_snodes.upload(D.snodes);
	d_stris.upload(D.stris);
	d_inst.upload(D.inst);
	d_inodes.upload(inst_tree.nodes());
	d_iitems.upload(inst_tree.items());
	d_pos.upload(pos, 3 * n);
	d_nrm.upload(nrm, 3 * n);
	d_dist.upload(dist, n);
	tm.upload_ms = ms_since(t0);

	PlaceView V;
	V.nodes = d_nodes.p;
	V.items = d_items.p;
	V.tris = d_tris.p;
	V.snodes = d_snodes.p;
	V.stris = d_stris.p;
	V.inst = d_inst.p;
	V.inodes = d_inodes.p;
	V.iitems = d_iitems.p;
	V.ninst = (int32_t)insts.size();
	hipStream_t s;
	PLACE_CHECK(hipStreamCreate(&s));
	hipEvent_t e0, e1;
	PLACE_CHECK(hipEventCreate(&e0));
	PLACE_CHECK(hipEventCreate(&e1));
	const unsigned grid = (unsigned)((n + kPlaceBlock - 1) / kPlaceBlock);
	PLACE_CHECK(hipEventRecord(e0, s));
	k_place<<<grid, kPlaceBlock, 0, s>>>(V, d_pos.p, d_nrm.p, d_dist.p, (uint32_t)n, max_corr, corr_min_dot);
	hipError_t le = hipGetLastError();
	PLACE_CHECK(hipEventRecord(e1, s));
	hipError_t se = hipStreamSynchronize(s);
	float kms = 0;
	if (le == hipSuccess && se == hipSuccess)
		(void)hipEventElapsedTime(&kms, e0, e1);
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	(void)hipStreamDestroy(s);
	PLACE_CHECK(le);
	PLACE_CHECK(se);
	tm.kernel_ms = kms;
	t0 = Clock::now();
	PLACE_CHECK(hipMemcpy(pos, d_pos.p, 3 * n * sizeof(float), hipMemcpyDeviceToHost));
	tm.download_ms = ms_since(t0);
	if (times)
		*times = tm;
}

} // namespace lgx
