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
// The BoxTrees are stored in pre-order (node n's first child is n + 1, its
// second child ch follows the first child's subtree), so their walks need
// no stack: every node carries the node that follows its subtree, and a
// walk goes to n + 1 after a node with children and to that node after a
// leaf or a culled box, which visits the nodes in the stack walk's order.
// The split-tree walk keeps a per-lane stack, sized from the trees' depths
// and checked on the host before the launch.  This is a preprocessing step
// that runs once per bake; its cost is dominated by divergent tree walks.
#include <hip/hip_runtime.h>

#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <stdexcept>
#include <string>
#include <thread>

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
// Stack entries the split-tree walk needs equal the tree's depth (nodes on
// the longest root-to-leaf path): SplitTrees stop splitting at depth 32
// (lighter_math.cpp:430-435), so 33 suffice; checked per tree before the
// launch.
constexpr int kSplitStack = 40;

// A BoxTree node for the stackless walk: link = (the node after this
// node's subtree) << 1 | (the node has children); kWalkEnd after the root's
// subtree.
struct PlaceNode {
	f3 mn, mx;
	int32_t ido;
	uint32_t link;
};
constexpr uint32_t kWalkEnd = 0x3FFFFFFFu;

__host__ __device__ __forceinline__ uint32_t walk_next(uint32_t n, uint32_t link)
{
	return (link & 1u) ? n + 1 : link >> 1;
}

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
	const PlaceNode* __restrict__ nodes;
	const int32_t* __restrict__ items;
	const PlaceTri* __restrict__ tris;
	const FlatSplitNode* __restrict__ snodes;
	const FlatSplitTri* __restrict__ stris;
	const PlaceInst* __restrict__ inst;
	const PlaceNode* __restrict__ inodes; // instance tree
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
	for (uint32_t n = 0; n != kWalkEnd;) {
		const PlaceNode N = V.nodes[I.node_off + n];
		if (box_disjoint(r.qmn, r.qmx, N.mn, N.mx) || !ray_box(r, N.mn, N.mx)) {
			n = N.link >> 1;
			continue;
		}
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
		n = walk_next(n, N.link);
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
	for (uint32_t n = 0; n != kWalkEnd;) {
		const PlaceNode B = V.nodes[I.node_off + n];
		if (box_disjoint(qmn, qmx, B.mn, B.mx)) {
			n = B.link >> 1;
			continue;
		}
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
		n = walk_next(n, B.link);
	}
}

// SplitTree::intersect (lighter_math.cpp:490-533).  The stack's entries
// are node indices or -1 (the far side of a split, not searched); it lives
// in private memory (a 16-bit copy in LDS measured slower).
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
	for (uint32_t n = 0; n != kWalkEnd;) {
		const PlaceNode N = V.inodes[n];
		if (box_disjoint(r.qmn, r.qmx, N.mn, N.mx) || !ray_box(r, N.mn, N.mx)) {
			n = N.link >> 1;
			continue;
		}
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
		n = walk_next(n, N.link);
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

// 8 waves per SIMD (64 VGPRs): the divergent walks hide their latency with
// more waves.  C4 city block with the BoxTree walks on stacks: 358 ms at
// the compiler's 3 waves, 235 ms at 5, 222 ms at 6, 196 ms at 8; with the
// stackless walks 165 ms at 8 and 176 ms at 6.
__global__ __launch_bounds__(kPlaceBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_place(
    PlaceView V, float* __restrict__ pos, const float* __restrict__ nrm, const float* __restrict__ dist, uint32_t n,
    float max_corr, float corr_min_dot)
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

// Pinned host block (hipHostMalloc: the upload runs at DMA speed).
struct PinnedBlock {
	char* p = nullptr;
	explicit PinnedBlock(size_t bytes) { PLACE_CHECK(hipHostMalloc((void**)&p, bytes ? bytes : 1, hipHostMallocDefault)); }
	PinnedBlock(const PinnedBlock&) = delete;
	PinnedBlock& operator=(const PinnedBlock&) = delete;
	~PinnedBlock()
	{
		if (p)
			(void)hipHostFree(p);
	}
};

struct DevBlock {
	char* p = nullptr;
	explicit DevBlock(size_t bytes) { PLACE_CHECK(hipMalloc((void**)&p, bytes ? bytes : 1)); }
	DevBlock(const DevBlock&) = delete;
	DevBlock& operator=(const DevBlock&) = delete;
	~DevBlock()
	{
		if (p)
			(void)hipFree(p);
	}
};

void check_depth(int depth, int cap, const char* what)
{
	if (depth > cap)
		throw std::runtime_error(std::string("sample placement: a ") + what + " of depth " + std::to_string(depth) +
		                         " exceeds the kernel's stack of " + std::to_string(cap));
}

// All placement inputs in one block (one pinned host copy, one device copy):
// the instance records, the instances' BoxTree nodes, item lists and
// triangles, their split trees, the instance tree and the samples.  Byte
// offsets of each array, aligned to 256.
struct PlaceLayout {
	size_t inst, nodes, items, tris, snodes, stris, inodes, iitems, pos, nrm, dist, bytes;
};

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// A BoxTree's nodes for the stackless walk.  Parents come before their
// children, so one pass in index order gives every node the node after its
// subtree: the first child's is the second child, the second child's is
// its parent's.
void walk_nodes(const std::vector<BoxTree::Node>& in, PlaceNode* out)
{
	const size_t n = in.size();
	if (n == 0)
		return;
	std::vector<uint32_t> after(n, kWalkEnd);
	for (size_t k = 0; k < n; ++k) {
		const int32_t ch = in[k].ch;
		if (ch == -1)
			continue;
		if ((size_t)ch <= k + 1 || (size_t)ch >= n)
			throw std::runtime_error("sample placement: a tree node's children are not in pre-order");
		after[k + 1] = (uint32_t)ch;
		after[ch] = after[k];
	}
	for (size_t k = 0; k < n; ++k)
		out[k] = PlaceNode{in[k].mn, in[k].mx, in[k].ido, (after[k] << 1) | (in[k].ch != -1 ? 1u : 0u)};
}

// Sizes, offsets and the instance records (serial; sizes only).
PlaceLayout plan(const std::vector<PlaceInstance>& insts, const BoxTree& inst_tree, size_t n,
                 std::vector<PlaceInst>& inst)
{
	inst.assign(insts.size(), PlaceInst{});
	size_t nn = 0, ni = 0, nt = 0, sn = 0, stt = 0;
	for (size_t k = 0; k < insts.size(); ++k) {
		const BoxTree& bt = insts[k].tris->tree();
		PlaceInst& I = inst[k];
		I.node_off = (int32_t)nn;
		I.item_off = (int32_t)ni;
		I.tri_off = (int32_t)nt;
		I.split_off = (int32_t)sn;
		I.stri_off = (int32_t)stt;
		I.nsplit = (int32_t)insts[k].split->num_nodes();
		I.shadow = insts[k].shadow ? 1 : 0;
		I.pad = 0;
		I.mn = bt.root().mn;
		I.mx = bt.root().mx;
		nn += bt.nodes().size();
		ni += bt.items().size();
		nt += insts[k].tris->tris().size();
		sn += insts[k].split->num_nodes();
		stt += insts[k].split->num_leaf_tris();
		if (nn >= kWalkEnd || ni > 0x7FFFFFFF || nt > 0x7FFFFFFF || sn > 0x7FFFFFFF || stt > 0x7FFFFFFF)
			throw std::runtime_error("sample placement: collision data beyond 2^31 entries");
	}
	PlaceLayout L;
	size_t o = 0;
	auto take = [&](size_t bytes) {
		size_t at = o;
		o = align256(o + bytes);
		return at;
	};
	L.inst = take(inst.size() * sizeof(PlaceInst));
	L.nodes = take(nn * sizeof(PlaceNode));
	L.items = take(ni * sizeof(int32_t));
	L.tris = take(nt * sizeof(PlaceTri));
	L.snodes = take(sn * sizeof(FlatSplitNode));
	L.stris = take(stt * sizeof(FlatSplitTri));
	L.inodes = take(inst_tree.nodes().size() * sizeof(PlaceNode));
	L.iitems = take(inst_tree.items().size() * sizeof(int32_t));
	L.pos = take(3 * n * sizeof(float));
	L.nrm = take(3 * n * sizeof(float));
	L.dist = take(n * sizeof(float));
	L.bytes = o;
	return L;
}

// Writes the block at base (instances in parallel on nthreads threads);
// throws when a tree is deeper than the kernel's stacks.
void fill(const std::vector<PlaceInstance>& insts, const BoxTree& inst_tree, const std::vector<PlaceInst>& inst,
          const PlaceLayout& L, char* base, int nthreads)
{
	memcpy(base + L.inst, inst.data(), inst.size() * sizeof(PlaceInst));
	walk_nodes(inst_tree.nodes(), (PlaceNode*)(base + L.inodes));
	std::copy(inst_tree.items().begin(), inst_tree.items().end(), (int32_t*)(base + L.iitems));
	std::atomic<size_t> next{0};
	std::atomic<int> split_depth{0};
	auto work = [&]() {
		for (size_t k; (k = next.fetch_add(1)) < insts.size();) {
			const TriSet& ts = *insts[k].tris;
			const BoxTree& bt = ts.tree();
			const PlaceInst& I = inst[k];
			walk_nodes(bt.nodes(), (PlaceNode*)(base + L.nodes) + I.node_off);
			std::copy(bt.items().begin(), bt.items().end(), (int32_t*)(base + L.items) + I.item_off);
			const std::vector<Tri>& tv = ts.tris();
			const std::vector<Box>& bv = ts.boxes();
			PlaceTri* tris = (PlaceTri*)(base + L.tris) + I.tri_off;
			for (size_t j = 0; j < tv.size(); ++j)
				tris[j] = PlaceTri{tv[j].p1, tv[j].p2, tv[j].p3, bv[j].mn, bv[j].mx};
			const int sd = insts[k].split->flatten((FlatSplitNode*)(base + L.snodes) + I.split_off,
			                                       (FlatSplitTri*)(base + L.stris) + I.stri_off);
			for (int cur = split_depth.load(); sd > cur && !split_depth.compare_exchange_weak(cur, sd);) {
			}
		}
	};
	const int nt = std::max(1, std::min<int>(nthreads, (int)insts.size()));
	std::vector<std::thread> th;
	for (int t = 1; t < nt; ++t)
		th.emplace_back(work);
	work();
	for (auto& t : th)
		t.join();
	check_depth(split_depth.load(), kSplitStack, "split tree");
}

// A view of the block at base (host or device).
PlaceView view_of(const char* base, const PlaceLayout& L, int32_t ninst)
{
	PlaceView V;
	V.nodes = (const PlaceNode*)(base + L.nodes);
	V.items = (const int32_t*)(base + L.items);
	V.tris = (const PlaceTri*)(base + L.tris);
	V.snodes = (const FlatSplitNode*)(base + L.snodes);
	V.stris = (const FlatSplitTri*)(base + L.stris);
	V.inst = (const PlaceInst*)(base + L.inst);
	V.inodes = (const PlaceNode*)(base + L.inodes);
	V.iitems = (const int32_t*)(base + L.iitems);
	V.ninst = ninst;
	return V;
}

} // namespace

// f(k) for k in [0, n) on nthreads threads
template <class F>
static void run_parallel(int nthreads, size_t n, F&& f)
{
	const int nt = (int)std::max<size_t>(1, std::min<size_t>((size_t)std::max(1, nthreads), n));
	std::atomic<size_t> next{0};
	auto work = [&]() {
		for (size_t k; (k = next.fetch_add(1)) < n;)
			f(k);
	};
	std::vector<std::thread> th;
	for (int t = 1; t < nt; ++t)
		th.emplace_back(work);
	work();
	for (auto& t : th)
		t.join();
}

void place_samples_gpu(const std::vector<PlaceInstance>& insts, const BoxTree& inst_tree,
                       const std::vector<PlaceSamples>& runs, float max_corr, float corr_min_dot, int nthreads,
                       PlaceTimes* times)
{
	PlaceTimes tm;
	std::vector<size_t> first(runs.size() + 1, 0);
	for (size_t r = 0; r < runs.size(); ++r)
		first[r + 1] = first[r] + runs[r].n;
	const size_t n = first.back();
	tm.samples = n;
	if (n == 0) {
		if (times)
			*times = tm;
		return;
	}
	if (n > 0xFFFFFFFFull)
		throw std::runtime_error("sample placement: more than 2^32 samples");
	auto t0 = Clock::now();
	std::vector<PlaceInst> inst;
	const PlaceLayout L = plan(insts, inst_tree, n, inst);
	PinnedBlock host(L.bytes);
	fill(insts, inst_tree, inst, L, host.p, nthreads);
	float* hpos = (float*)(host.p + L.pos);
	float* hnrm = (float*)(host.p + L.nrm);
	float* hdist = (float*)(host.p + L.dist);
	run_parallel(nthreads, runs.size(), [&](size_t r) {
		const PlaceSamples& R = runs[r];
		for (size_t i = 0; i < R.n; ++i) {
			const size_t k = first[r] + i;
			st3(hpos + 3 * k, R.pos[i]);
			st3(hnrm + 3 * k, R.nrm[i]);
			hdist[k] = R.dist[i];
		}
	});
	tm.flatten_ms = ms_since(t0);

	t0 = Clock::now();
	DevBlock dev(L.bytes);
	PLACE_CHECK(hipMemcpy(dev.p, host.p, L.bytes, hipMemcpyHostToDevice));
	tm.upload_ms = ms_since(t0);

	const PlaceView V = view_of(dev.p, L, (int32_t)insts.size());
	float* d_pos = (float*)(dev.p + L.pos);
	hipStream_t s;
	PLACE_CHECK(hipStreamCreate(&s));
	hipEvent_t e0, e1;
	PLACE_CHECK(hipEventCreate(&e0));
	PLACE_CHECK(hipEventCreate(&e1));
	const unsigned grid = (unsigned)((n + kPlaceBlock - 1) / kPlaceBlock);
	PLACE_CHECK(hipEventRecord(e0, s));
	k_place<<<grid, kPlaceBlock, 0, s>>>(V, d_pos, (const float*)(dev.p + L.nrm), (const float*)(dev.p + L.dist),
	                                     (uint32_t)n, max_corr, corr_min_dot);
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
	PLACE_CHECK(hipMemcpy(hpos, d_pos, 3 * n * sizeof(float), hipMemcpyDeviceToHost));
	run_parallel(nthreads, runs.size(), [&](size_t r) {
		const PlaceSamples& R = runs[r];
		for (size_t i = 0; i < R.n; ++i)
			R.pos[i] = ld3(hpos + 3 * (first[r] + i));
	});
	tm.download_ms = ms_since(t0);
	if (times)
		*times = tm;
}

} // namespace lgx
