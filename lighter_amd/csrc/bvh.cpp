// Binned-SAH BVH builder for the gfx950 kernels; see bvh.h.
#include "bvh.h"

#include <float.h>
#include <math.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <stdexcept>
#include <thread>

namespace lgx {

namespace {

struct Aabb {
	float lo[3], hi[3];
	void reset()
	{
		for (int a = 0; a < 3; ++a) {
			lo[a] = FLT_MAX;
			hi[a] = -FLT_MAX;
		}
	}
	void grow(const Aabb& b)
	{
		for (int a = 0; a < 3; ++a) {
			lo[a] = std::min(lo[a], b.lo[a]);
			hi[a] = std::max(hi[a], b.hi[a]);
		}
	}
	void grow(const float* p)
	{
		for (int a = 0; a < 3; ++a) {
			lo[a] = std::min(lo[a], p[a]);
			hi[a] = std::max(hi[a], p[a]);
		}
	}
	float area() const
	{
		float d[3];
		for (int a = 0; a < 3; ++a)
			d[a] = std::max(0.0f, hi[a] - lo[a]);
		return 2.0f * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
	}
};

constexpr int kBins = kSahBins; // bins of the build the kernels traverse (at most kMaxBins)
constexpr int kMaxBins = 32;
static_assert(kBins <= kMaxBins, "bins");

struct Builder {
	const float* tris;
	std::vector<Aabb> box;
	std::vector<float> cen; // 3 per triangle
	std::vector<uint32_t> idx;
	std::vector<BvhNode> nodes;
	int max_depth = 0;
	int bins = kBins;          // SAH bins per axis
	int leaf_max = kLeafMax;   // triangles per leaf at most

	// Subtrees below the top levels are built as independent tasks (their
	// triangle ranges are disjoint) and spliced back in depth-first order,
	// which gives the same node array as building everything in one pass.
	struct Child {
		int kind; // 0: final reference (leaf), 1: top node, 2: task
		int32_t ref;
	};
	struct TopNode {
		Aabb lb, rb;
		Child c[2];
	};
	struct Task {
		uint32_t b, e;
		int depth;
		std::vector<BvhNode> nodes;
		int32_t ref = 0;
		int maxd = 0;
	};
	std::vector<TopNode> top;
	std::vector<Task> tasks;

	std::mutex mu; // guards top, tasks and max_depth while the top levels build

	// A pass over idx[b, e) on up to par threads (bounds and bins merge
	// exactly: minima, maxima and counts).
	static constexpr uint32_t kParMin = 1u << 16;

	template <class T, class Fn, class Merge>
	T reduce(uint32_t b, uint32_t e, int par, const T& init, Fn fn, Merge merge) const
	{
		const uint32_t n = e - b;
		const int nt = (n >= kParMin && par > 1) ? par : 1;
		if (nt == 1) {
			T r = init;
			fn(b, e, r);
			return r;
		}
		std::vector<T> part(nt, init);
		std::vector<std::thread> th;
		for (int t = 0; t < nt; ++t) {
			const uint32_t cb = b + (uint32_t)((uint64_t)n * t / nt), ce = b + (uint32_t)((uint64_t)n * (t + 1) / nt);
			th.emplace_back([&, t, cb, ce] { fn(cb, ce, part[t]); });
		}
		for (auto& x : th)
			x.join();
		T r = init;
		for (const T& q : part)
			merge(r, q);
		return r;
	}

	struct Bins {
		Aabb bb[3][kMaxBins];
		uint32_t cnt[3][kMaxBins];
	};

	// Pads a box outward so that float rounding in the kernels' box tests
	// never rejects a point of a contained triangle.
	static Aabb padded(const Aabb& b)
	{
		Aabb o = b;
		for (int a = 0; a < 3; ++a) {
			float m = 1e-5f * std::max(fabsf(b.lo[a]), fabsf(b.hi[a])) + 1e-6f;
			o.lo[a] = b.lo[a] - m;
			o.hi[a] = b.hi[a] + m;
		}
		return o;
	}

	void set_child(BvhNode& n, int which, const Aabb& b, int32_t ref) { set_child_of(n, which, b, ref); }
	static void set_child_of(BvhNode& n, int which, const Aabb& b, int32_t ref)
	{
		Aabb p = padded(b);
		if (which == 0) {
			n.lo0x = p.lo[0]; n.lo0y = p.lo[1]; n.lo0z = p.lo[2];
			n.hi0x = p.hi[0]; n.hi0y = p.hi[1]; n.hi0z = p.hi[2];
			n.c0 = ref;
		} else {
			n.lo1x = p.lo[0]; n.lo1y = p.lo[1]; n.lo1z = p.lo[2];
			n.hi1x = p.hi[0]; n.hi1y = p.hi[1]; n.hi1z = p.hi[2];
			n.c1 = ref;
		}
	}

	Aabb bounds(uint32_t b, uint32_t e, int par = 1) const
	{
		Aabb init;
		init.reset();
		return reduce(
		    b, e, par, init,
		    [&](uint32_t cb, uint32_t ce, Aabb& r) {
			    for (uint32_t i = cb; i < ce; ++i)
				    r.grow(box[idx[i]]);
		    },
		    [](Aabb& r, const Aabb& q) { r.grow(q); });
	}

	// Partitions idx[b, e) (more than kLeafMax triangles) and returns the
	// split position.
	uint32_t split(uint32_t b, uint32_t e, int depth, int par = 1)
	{
		uint32_t n = e - b;
		Aabb init;
		init.reset();
		const Aabb cb = reduce(
		    b, e, par, init,
		    [&](uint32_t rb, uint32_t re, Aabb& r) {
			    for (uint32_t i = rb; i < re; ++i)
				    r.grow(&cen[3 * idx[i]]);
		    },
		    [](Aabb& r, const Aabb& q) { r.grow(q); });
		// bins of every axis with a centroid extent, in one pass
		Bins b0;
		for (int a = 0; a < 3; ++a)
			for (int k = 0; k < kMaxBins; ++k) {
				b0.bb[a][k].reset();
				b0.cnt[a][k] = 0;
			}
		bool use[3];
		float sc[3];
		for (int a = 0; a < 3; ++a) {
			float ext = cb.hi[a] - cb.lo[a];
			use[a] = ext > 0;
			sc[a] = use[a] ? bins / ext : 0.0f;
		}
		const Bins hist = reduce(
		    b, e, par, b0,
		    [&](uint32_t rb, uint32_t re, Bins& r) {
			    for (uint32_t i = rb; i < re; ++i) {
				    const uint32_t t = idx[i];
				    for (int a = 0; a < 3; ++a) {
					    if (!use[a])
						    continue;
					    int k = std::min(bins - 1, (int)((cen[3 * t + a] - cb.lo[a]) * sc[a]));
					    r.cnt[a][k]++;
					    r.bb[a][k].grow(box[t]);
				    }
			    }
		    },
		    [](Bins& r, const Bins& q) {
			    for (int a = 0; a < 3; ++a)
				    for (int k = 0; k < kMaxBins; ++k) {
					    r.bb[a][k].grow(q.bb[a][k]);
					    r.cnt[a][k] += q.cnt[a][k];
				    }
		    });
		// binned SAH over the axes with a centroid extent
		float best_cost = FLT_MAX;
		int best_axis = -1, best_split = -1;
		for (int a = 0; a < 3; ++a) {
			if (!use[a])
				continue;
			const Aabb* bb = hist.bb[a];
			const uint32_t* cnt = hist.cnt[a];
			float ra[kMaxBins];
			uint32_t rc[kMaxBins];
			Aabb acc;
			acc.reset();
			uint32_t c = 0;
			for (int k = bins - 1; k > 0; --k) {
				acc.grow(bb[k]);
				c += cnt[k];
				ra[k] = acc.area();
				rc[k] = c;
			}
			acc.reset();
			c = 0;
			for (int k = 0; k < bins - 1; ++k) {
				acc.grow(bb[k]);
				c += cnt[k];
				if (c == 0 || rc[k + 1] == 0)
					continue;
				float cost = acc.area() * c + ra[k + 1] * rc[k + 1];
				if (cost < best_cost) {
					best_cost = cost;
					best_axis = a;
					best_split = k;
				}
			}
		}
		// Order-independent rules throughout (bins, stable partition, the
		// index split), so the GPU build (bvh_gpu.hip) gives the same tree.
		uint32_t mid;
		if (depth >= kSahMaxDepth || best_axis < 0) {
			// bound the depth; coincident centroids: split by position
			mid = b + n / 2;
		} else {
			float ext = cb.hi[best_axis] - cb.lo[best_axis];
			float sc = bins / ext;
			auto it = std::stable_partition(idx.begin() + b, idx.begin() + e, [&](uint32_t t) {
				int k = std::min(bins - 1, (int)((cen[3 * t + best_axis] - cb.lo[best_axis]) * sc));
				return k <= best_split;
			});
			mid = (uint32_t)(it - idx.begin());
			if (mid == b || mid == e)
				mid = b + n / 2;
		}
		return mid;
	}

	// Returns a child reference for triangles idx[b, e), appending the
	// subtree's nodes to out (indices relative to out).
	int32_t build(uint32_t b, uint32_t e, int depth, std::vector<BvhNode>& out, int& maxd)
	{
		maxd = std::max(maxd, depth);
		uint32_t n = e - b;
		if (n <= (uint32_t)leaf_max)
			return leaf_ref(b, n);
		const uint32_t mid = split(b, e, depth);
		int32_t self = (int32_t)out.size();
		out.push_back(BvhNode{});
		Aabb lb = bounds(b, mid), rb = bounds(mid, e);
		int32_t l = build(b, mid, depth + 1, out, maxd);
		int32_t r = build(mid, e, depth + 1, out, maxd);
		set_child(out[self], 0, lb, l);
		set_child(out[self], 1, rb, r);
		return self;
	}

	// The top levels: subtrees of at most grain triangles become tasks.  par
	// threads share the work: the passes over a range, then the two halves
	// (the left one on a new thread).  Only the tree structure, not the
	// order in which top nodes and tasks are recorded, decides the output.
	Child build_top(uint32_t b, uint32_t e, int depth, uint32_t grain, int par)
	{
		uint32_t n = e - b;
		{
			std::lock_guard<std::mutex> g(mu);
			max_depth = std::max(max_depth, depth);
			if (n > (uint32_t)leaf_max && n <= grain) {
				tasks.push_back(Task{b, e, depth, {}, 0, 0});
				return Child{2, (int32_t)tasks.size() - 1};
			}
		}
		if (n <= (uint32_t)leaf_max)
			return Child{0, leaf_ref(b, n)};
		const uint32_t mid = split(b, e, depth, par);
		TopNode node;
		node.lb = bounds(b, mid, par);
		node.rb = bounds(mid, e, par);
		Child l, r;
		if (par > 1) {
			const int pl = par / 2;
			std::thread th([&] { l = build_top(b, mid, depth + 1, grain, pl); });
			r = build_top(mid, e, depth + 1, grain, par - pl);
			th.join();
		} else {
			l = build_top(b, mid, depth + 1, grain, 1);
			r = build_top(mid, e, depth + 1, grain, 1);
		}
		node.c[0] = l;
		node.c[1] = r;
		std::lock_guard<std::mutex> g(mu);
		top.push_back(node);
		return Child{1, (int32_t)top.size() - 1};
	}

	// Depth-first splice of the top nodes and the tasks' node blocks.
	int32_t emit(Child c)
	{
		if (c.kind == 0)
			return c.ref;
		if (c.kind == 2) {
			const Task& t = tasks[c.ref];
			const int32_t off = (int32_t)nodes.size();
			for (BvhNode nd : t.nodes) {
				if (nd.c0 >= 0)
					nd.c0 += off;
				if (nd.c1 >= 0)
					nd.c1 += off;
				nodes.push_back(nd);
			}
			return t.ref >= 0 ? t.ref + off : t.ref;
		}
		const TopNode& T = top[c.ref];
		const int32_t self = (int32_t)nodes.size();
		nodes.push_back(BvhNode{});
		const int32_t l = emit(T.c[0]);
		const int32_t r = emit(T.c[1]);
		set_child(nodes[self], 0, T.lb, l);
		set_child(nodes[self], 1, T.rb, r);
		return self;
	}

	void build_all(uint32_t ntri, int nthreads)
	{
		const uint32_t grain = std::max<uint32_t>(4096, ntri / (uint32_t)(8 * std::max(nthreads, 1)));
		const Child root = build_top(0, ntri, 0, grain, nthreads);
		std::atomic<size_t> next{0};
		auto work = [&]() {
			for (size_t k; (k = next.fetch_add(1)) < tasks.size();) {
				Task& t = tasks[k];
				t.ref = build(t.b, t.e, t.depth, t.nodes, t.maxd);
			}
		};
		std::vector<std::thread> th;
		for (int i = 1; i < nthreads && (size_t)i < tasks.size(); ++i)
			th.emplace_back(work);
		work();
		for (auto& x : th)
			x.join();
		for (const Task& t : tasks)
			max_depth = std::max(max_depth, t.maxd);
		size_t total = top.size();
		for (const Task& t : tasks)
			total += t.nodes.size();
		nodes.reserve(total);
		emit(root); // the root is an inner node (more than leaf_max triangles): index 0
	}
};

} // namespace

namespace {

// Collapse binary node b into a wide node: start from its two children and
// open the inner candidate of largest surface area until four candidates
// (or only leaves) remain; then collapse the inner candidates in turn.
struct Collapser {
	const std::vector<BvhNode>& bin;
	std::vector<WideNode>& out;
	int max_depth = 0;

	struct Cand {
		float lo[3], hi[3];
		int32_t ref;
		float area() const
		{
			float d[3];
			for (int a = 0; a < 3; ++a)
				d[a] = std::max(0.0f, hi[a] - lo[a]);
			return d[0] * d[1] + d[1] * d[2] + d[2] * d[0];
		}
	};

	static Cand child(const BvhNode& n, int k)
	{
		Cand c;
		if (k == 0) {
			c.lo[0] = n.lo0x, c.lo[1] = n.lo0y, c.lo[2] = n.lo0z;
			c.hi[0] = n.hi0x, c.hi[1] = n.hi0y, c.hi[2] = n.hi0z;
			c.ref = n.c0;
		} else {
			c.lo[0] = n.lo1x, c.lo[1] = n.lo1y, c.lo[2] = n.lo1z;
			c.hi[0] = n.hi1x, c.hi[1] = n.hi1y, c.hi[2] = n.hi1z;
			c.ref = n.c1;
		}
		return c;
	}

	int32_t collapse(int32_t b, int depth)
	{
		max_depth = std::max(max_depth, depth + 1);
		Cand c[4];
		int n = 0;
		c[n++] = child(bin[b], 0);
		c[n++] = child(bin[b], 1);
		while (n < 4) {
			int pick = -1;
			float best = -1.0f;
			for (int i = 0; i < n; ++i)
				if (c[i].ref >= 0 && c[i].area() > best) {
					best = c[i].area();
					pick = i;
				}
			if (pick < 0)
				break;
			const BvhNode& m = bin[c[pick].ref];
			c[pick] = child(m, 0);
			c[n++] = child(m, 1);
		}
		const int32_t self = (int32_t)out.size();
		out.push_back(WideNode{});
		int32_t refs[4];
		for (int i = 0; i < n; ++i)
			refs[i] = c[i].ref >= 0 ? collapse(c[i].ref, depth + 1) : c[i].ref;
		WideNode& w = out[self];
		for (int i = 0; i < 4; ++i) {
			// an unused slot, or a box that holds nothing (lo > hi), gets the
			// far degenerate box and a harmless real leaf
			const bool use = i < n && c[i].lo[0] <= c[i].hi[0] && c[i].lo[1] <= c[i].hi[1] && c[i].lo[2] <= c[i].hi[2];
			w.lox[i] = use ? c[i].lo[0] : kFarBox;
			w.loy[i] = use ? c[i].lo[1] : kFarBox;
			w.loz[i] = use ? c[i].lo[2] : kFarBox;
			w.hix[i] = use ? c[i].hi[0] : kFarBox;
			w.hiy[i] = use ? c[i].hi[1] : kFarBox;
			w.hiz[i] = use ? c[i].hi[2] : kFarBox;
			w.child[i] = use ? refs[i] : leaf_ref(0, 1);
		}
		return self;
	}
};

} // namespace

// Boxes, centroids and the identity order of the input triangles; returns
// the scene bounds.
static Aabb prepare(Builder& B, const float* tris, size_t ntri)
{
	B.tris = tris;
	B.box.resize(ntri);
	B.cen.resize(3 * ntri);
	B.idx.resize(ntri);
	Aabb all;
	all.reset();
	for (size_t t = 0; t < ntri; ++t) {
		Aabb& b = B.box[t];
		b.reset();
		for (int v = 0; v < 3; ++v)
			b.grow(tris + 9 * t + 3 * v);
		for (int a = 0; a < 3; ++a)
			B.cen[3 * t + a] = 0.5f * (b.lo[a] + b.hi[a]);
		B.idx[t] = (uint32_t)t;
		all.grow(b);
	}
	return all;
}

// The root (node 0) is always an inner node; a small scene gets a second
// child whose inverted box is never entered.
static void build_binary(Builder& B, const Aabb& all, size_t ntri)
{
	if (ntri <= (size_t)B.leaf_max) {
		B.nodes.push_back(BvhNode{});
		B.set_child(B.nodes[0], 0, all, leaf_ref(0, (uint32_t)ntri));
		B.nodes[0].lo1x = B.nodes[0].lo1y = B.nodes[0].lo1z = FLT_MAX;
		B.nodes[0].hi1x = B.nodes[0].hi1y = B.nodes[0].hi1z = -FLT_MAX;
		B.nodes[0].c1 = leaf_ref(0, 1);
	} else {
		const unsigned hw = std::thread::hardware_concurrency();
		B.build_all((uint32_t)ntri, (int)std::min(16u, std::max(1u, hw)));
	}
}

void collapse_bvh(const std::vector<BvhNode>& bin, BvhBuild& out)
{
	out.nodes.clear();
	Collapser W{bin, out.nodes};
	out.nodes.reserve(bin.size() / 2 + 1);
	W.collapse(0, 0);
	out.depth = W.max_depth;
}

BvhNode bvh_node(const float lo0[3], const float hi0[3], int32_t c0, const float lo1[3], const float hi1[3],
                 int32_t c1)
{
	BvhNode n{};
	Aabb a, b;
	for (int k = 0; k < 3; ++k) {
		a.lo[k] = lo0[k];
		a.hi[k] = hi0[k];
		b.lo[k] = lo1[k];
		b.hi[k] = hi1[k];
	}
	Builder::set_child_of(n, 0, a, c0);
	Builder::set_child_of(n, 1, b, c1);
	return n;
}

BvhBuild build_bvh(const float* tris, size_t ntri)
{
	Builder B;
	const Aabb all = prepare(B, tris, ntri);
	BvhBuild out;
	if (ntri == 0) {
		out.nodes.clear();
		out.lo = mk3(0, 0, 0);
		out.hi = mk3(0, 0, 0);
		return out;
	}
	build_binary(B, all, ntri);
	collapse_bvh(B.nodes, out);
	out.order = std::move(B.idx);
	out.lo = mk3(all.lo[0], all.lo[1], all.lo[2]);
	out.hi = mk3(all.hi[0], all.hi[1], all.hi[2]);
	return out;
}

BinaryBvh build_bvh_binary(const float* tris, size_t ntri, int bins, int leaf_max)
{
	if (bins < 2 || bins > kMaxBins || leaf_max < 1 || leaf_max > 16)
		throw std::runtime_error("build_bvh_binary: bins must be 2..32 and leaves 1..16 triangles");
	Builder B;
	B.bins = bins;
	B.leaf_max = leaf_max;
	const Aabb all = prepare(B, tris, ntri);
	BinaryBvh out;
	out.lo = mk3(all.lo[0], all.lo[1], all.lo[2]);
	out.hi = mk3(all.hi[0], all.hi[1], all.hi[2]);
	if (ntri == 0)
		return out;
	build_binary(B, all, ntri);
	out.nodes = std::move(B.nodes);
	out.order = std::move(B.idx);
	out.depth = B.max_depth + 1;
	return out;
}

} // namespace lgx
