// Host geometry for the CPU-side bake stages; see host_geom.h.
#include "host_geom.h"

#include <float.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>

namespace lgx {

// ---------------------------------------------------------------------------
// 4x4 inverse by cofactor expansion (lighter_math.cpp:19-147).  Element
// (i, j) of the inverse is the (j, i) cofactor: the 3x3 minor without row j
// and column i, expanded along its first column, each term written as
// (column element) * (upper row factor) * (lower row factor) and summed left
// to right.  This fixed term order keeps the result identical to the
// reference's.
// ---------------------------------------------------------------------------
bool mat4_invert(const Mat4& M, Mat4& out)
{
	const float* a = &M.m[0][0];
	float inv[16];
	for (int i = 0; i < 4; ++i) {
		for (int j = 0; j < 4; ++j) {
			int rows[3], cols[3], nr = 0, nc = 0;
			for (int k = 0; k < 4; ++k) {
				if (k != j) rows[nr++] = k;
				if (k != i) cols[nc++] = k;
			}
			float sign = ((i + j) & 1) ? -1.0f : 1.0f;
			float acc = 0.0f;
			bool first = true;
			for (int k = 0; k < 3; ++k) {
				int ra = rows[k == 0 ? 1 : 0], rb = rows[k == 2 ? 1 : 2];
				float e = a[4 * rows[k] + cols[0]];
				float s = (k == 1) ? -sign : sign;
				float tp = e * a[4 * ra + cols[1]] * a[4 * rb + cols[2]];
				float tn = e * a[4 * ra + cols[2]] * a[4 * rb + cols[1]];
				if (first) {
					acc = s > 0 ? tp : (-e) * a[4 * ra + cols[1]] * a[4 * rb + cols[2]];
					first = false;
				} else {
					acc = s > 0 ? acc + tp : acc - tp;
				}
				acc = s > 0 ? acc - tn : acc + tn;
			}
			inv[4 * i + j] = acc;
		}
	}
	float det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
	if (det == 0)
		return false;
	det = 1.0f / det;
	float* o = &out.m[0][0];
	for (int i = 0; i < 16; ++i)
		o[i] = inv[i] * det;
	return true;
}

f3 mat4_transform(const Mat4& M, f3 v, float w)
{
	const float(*m)[4] = M.m;
	return mk3(v.x * m[0][0] + v.y * m[1][0] + v.z * m[2][0] + m[3][0] * w,
	           v.x * m[0][1] + v.y * m[1][1] + v.z * m[2][1] + m[3][1] * w,
	           v.x * m[0][2] + v.y * m[1][2] + v.z * m[2][2] + m[3][2] * w);
}

Mat4 normal_matrix(const Mat4& M)
{
	Mat4 o;
	for (int r = 0; r < 3; ++r)
		for (int c = 0; c < 3; ++c)
			o.m[r][c] = M.m[r][c];
	o.m[0][3] = o.m[1][3] = o.m[2][3] = 0;
	o.m[3][0] = o.m[3][1] = o.m[3][2] = 0;
	o.m[3][3] = 1;
	Mat4 inv;
	if (mat4_invert(o, inv))
		o = inv;
	for (int r = 0; r < 4; ++r)
		for (int c = r + 1; c < 4; ++c)
			std::swap(o.m[r][c], o.m[c][r]);
	return o;
}

static float heron(float a, float b, float c)
{
	float p = (a + b + c) * 0.5f;
	float pre = p * (p - a) * (p - b) * (p - c);
	if (pre < 0)
		return 0;
	return sqrtf(pre);
}

float triangle_area(f3 p1, f3 p2, f3 p3) { return heron(length(p2 - p1), length(p3 - p2), length(p1 - p3)); }
float triangle_area(f2 p1, f2 p2, f2 p3) { return heron(length2(p2 - p1), length2(p3 - p2), length2(p1 - p3)); }

float sample_area(f2 t1, f2 t2, f2 t3, f3 p1, f3 p2, f3 p3)
{
	float lm = triangle_area(t1, t2, t3);
	float co = triangle_area(p1, p2, p3);
	return lm > 0 ? co / lm : 0;
}

// ---------------------------------------------------------------------------
// BoxTree construction (lighter_math.cpp:674-781): median split of the
// "splittable" items (volume * 3 < node volume) along the axis of their
// largest extent; the others stay on the node.  Leaves at <= 4 items, depth
// 16, or fewer than 4 splittable items.
// ---------------------------------------------------------------------------
static constexpr size_t kMinSplit = 4;
static constexpr int kMaxDepth = 16;
static constexpr float kSizeFactor = 3.0f;

void BoxTree::build(const Box* boxes, size_t count)
{
	nodes_.clear();
	items_.clear();
	nodes_.push_back(Node{mk3(FLT_MAX, FLT_MAX, FLT_MAX), mk3(-FLT_MAX, -FLT_MAX, -FLT_MAX), -1, -1});
	if (count == 0)
		return;
	std::vector<int32_t> idx;
	for (size_t i = 0; i < count; ++i)
		if (boxes[i].valid())
			idx.push_back((int32_t)i);
	make_node(0, boxes, idx.data(), idx.size(), 0);
}

void BoxTree::make_node(int32_t node, const Box* boxes, int32_t* idx, size_t count, int depth)
{
	f3 bmn = mk3(FLT_MAX, FLT_MAX, FLT_MAX), bmx = mk3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
	for (size_t i = 0; i < count; ++i) {
		bmn = vmin(bmn, boxes[idx[i]].mn);
		bmx = vmax(bmx, boxes[idx[i]].mx);
	}
	nodes_[node].mn = bmn;
	nodes_[node].mx = bmx;
	float vol = Box{bmn, bmx}.volume();

	bool leaf = true;
	if (count > kMinSplit && depth < kMaxDepth) {
		nodes_[node].ido = -1;
		nodes_[node].ch = -1;
		int splittable = 0;
		f3 smn = mk3(FLT_MAX, FLT_MAX, FLT_MAX), smx = mk3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
		for (size_t i = 0; i < count; ++i) {
			const Box& b = boxes[idx[i]];
			if (b.volume() * kSizeFactor < vol) {
				splittable++;
				smn = vmin(smn, b.mn);
				smx = vmax(smx, b.mx);
			}
		}
		f3 ssz = smx - smn;
		f3 axis = mk3(0, 0, 1);
		if (ssz.x > ssz.y && ssz.x > ssz.z)
			axis = mk3(1, 0, 0);
		else if (ssz.y > ssz.x && ssz.y > ssz.z)
			axis = mk3(0, 1, 0);
		if (splittable >= (int)kMinSplit) {
			leaf = false;
			std::vector<int32_t> keep, sub;
			for (size_t i = 0; i < count; ++i) {
				if (boxes[idx[i]].volume() * kSizeFactor < vol)
					sub.push_back(idx[i]);
				else
					keep.push_back(idx[i]);
			}
			if (!keep.empty()) {
				nodes_[node].ido = (int32_t)items_.size();
				items_.push_back((int32_t)keep.size());
				items_.insert(items_.end(), keep.begin(), keep.end());
			}
			std::sort(sub.begin(), sub.end(), [&](int32_t a, int32_t b) {
				return dot(axis, boxes[a].center()) < dot(axis, boxes[b].center());
			});
			size_t mid = sub.size() / 2;
			nodes_.push_back(Node{});
			make_node((int32_t)nodes_.size() - 1, boxes, sub.data(), mid, depth + 1);
			nodes_[node].ch = (int32_t)nodes_.size();
			nodes_.push_back(Node{});
			make_node((int32_t)nodes_.size() - 1, boxes, sub.data() + mid, sub.size() - mid, depth + 1);
		}
	}
	if (leaf) {
		nodes_[node].ido = (int32_t)items_.size();
		nodes_[node].ch = -1;
		items_.push_back((int32_t)count);
		items_.insert(items_.end(), idx, idx + count);
	}
}

int BoxTree::depth() const
{
	int depth = 0;
	std::vector<std::pair<int32_t, int>> st{{0, 1}};
	while (!st.empty()) {
		auto [k, d] = st.back();
		st.pop_back();
		depth = std::max(depth, d);
		if (nodes_[k].ch != -1) {
			st.push_back({nodes_[k].ch, d + 1});
			st.push_back({k + 1, d + 1});
		}
	}
	return depth;
}

// ---------------------------------------------------------------------------
void TriSet::set(const Tri* tris, size_t count)
{
	tris_.clear();
	boxes_.clear();
	for (size_t i = 0; i < count; ++i) {
		if (tri_useful(tris[i])) {
			boxes_.push_back(tri_box(tris[i]));
			tris_.push_back(tris[i]);
		}
	}
	tree_.build(boxes_.data(), boxes_.size());
}

float TriSet::ray_dist(f3 from, f3 to, int32_t* outtid) const
{
	RaySeg r;
	r.set(from, to);
	float closest = 2.0f;
	int32_t hit = -1;
	auto fn = [&](const int32_t* ids, int32_t n) {
		for (int32_t i = 0; i < n; ++i) {
			const Tri& T = tris_[ids[i]];
			float d = seg_tri(r.origin, r.end, T.p1, T.p2, T.p3);
			if (d < closest) {
				closest = d;
				hit = ids[i];
			}
		}
		return true;
	};
	tree_.ray_query(r, fn);
	if (hit != -1 && outtid)
		*outtid = hit;
	return closest;
}

// OffsetSample (lighter_math.cpp:991-1044): the triangles of every node the
// cube around the starting P overlaps, in tree order, each possibly moving P.
void TriSet::offset_sample(f3& P, f3 N, float dist) const
{
	auto fn = [&](const int32_t* ids, int32_t n) {
		for (int32_t i = 0; i < n; ++i) {
			const Box& b = boxes_[ids[i]];
			if (offset_box_skip(P, dist, b.mn, b.mx))
				continue;
			const Tri& T = tris_[ids[i]];
			f3 from, pnew;
			if (!offset_candidate(P, N, dist, T.p1, T.p2, T.p3, &from, &pnew))
				continue;
			int32_t tid = -1;
			if (ray_dist(from, pnew, &tid) >= 0.9f || tid == ids[i])
				P = pnew;
		}
	};
	f3 d3 = mk3(dist, dist, dist);
	tree_.query(P - d3, P + d3, fn);
}

// ---------------------------------------------------------------------------
// SplitTree (lighter_math.cpp:428-614)
// ---------------------------------------------------------------------------
static constexpr size_t kSplitMaxTris = 16;
static constexpr int kSplitMaxDepth = 32;

void SplitTree::set(const Tri* tris, size_t count, bool brute_pairs)
{
	brute_pairs_ = brute_pairs;
	nodes_.clear();
	nodes_.emplace_back();
	nodes_[0].tris.assign(tris, tris + count);
	if (count)
		split(0, 0);
	// a triangle whose normal is near zero never reports a hit, so the leaf
	// lists keep only the others, in order, with their operands precomputed
	for (Node& n : nodes_) {
		if (n.front != -1)
			continue;
		for (size_t i = 0; i < n.tris.size(); ++i) {
			const Tri& t = n.tris[i];
			if (!tri_useful(t))
				continue;
			n.rt.push_back(make_ray_tri(t.p1, t.p2, t.p3));
			n.rti.push_back((int32_t)i);
		}
	}
}

namespace {
struct PointKey {
	uint32_t x, y, z;
	bool operator==(const PointKey& o) const { return x == o.x && y == o.y && z == o.z; }
};
struct PointKeyHash {
	size_t operator()(const PointKey& k) const
	{
		uint64_t h = k.x * 0x9E3779B97F4A7C15ull;
		h ^= (k.y + 0x632BE59BD9B4E019ull) * 0xC2B2AE3D27D4EB4Full;
		h ^= (k.z + 0x165667B19E3779F9ull) * 0x85EBCA77C2B2AE63ull;
		return (size_t)(h ^ (h >> 29));
	}
};
inline uint32_t canon_bits(float f)
{
	if (f == 0.0f)
		f = 0.0f; // +0 and -0 compare equal
	uint32_t u;
	memcpy(&u, &f, 4);
	return u;
}
} // namespace

// pts[j] - pts[i] of the pair i < j whose float squared length is the
// largest, the first such pair in (i, j) order (lighter_math.cpp:451-460
// scans every pair and keeps a pair only when it is strictly longer).  With
// brute set, or any coordinate not finite, every pair is scanned.  Otherwise
// the scan is restricted to the points that can be in such a pair: L, the
// float squared length of a real pair (the longest among the extreme points
// along 7 directions), is at most the maximum M; a point whose distance to
// the farthest corner of the points' box is below sqrt(L) (with a margin
// for the float rounding of the difference and the sum of squares) is in no
// pair reaching M.  Every pair reaching M therefore joins two surviving
// points, and scanning those in their original order finds the same first
// pair, with the same float difference.
static f3 farthest_pair_dir(const std::vector<f3>& pts, bool brute)
{
	auto scan = [&](const std::vector<uint32_t>* idx) {
		const size_t n = idx ? idx->size() : pts.size();
		f3 dir = mk3(0, 0, 0);
		float best = 0;
		for (size_t a = 0; a < n; ++a) {
			const f3 pi = pts[idx ? (*idx)[a] : a];
			for (size_t b = a + 1; b < n; ++b) {
				const f3 nd = pts[idx ? (*idx)[b] : b] - pi;
				const float l = lensq(nd);
				if (l > best) {
					dir = nd;
					best = l;
				}
			}
		}
		return dir;
	};
	bool finite = true;
	for (const f3& p : pts)
		finite = finite && std::isfinite(p.x) && std::isfinite(p.y) && std::isfinite(p.z);
	if (brute || !finite || pts.size() < 64)
		return scan(nullptr);
	double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
	// extreme points along x, y, z and the four diagonals
	static const int dirs[7][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {-1, 1, 1}};
	size_t emin[7], emax[7];
	double vmin[7], vmax[7];
	for (int k = 0; k < 7; ++k) {
		emin[k] = emax[k] = 0;
		vmin[k] = INFINITY;
		vmax[k] = -INFINITY;
	}
	for (size_t i = 0; i < pts.size(); ++i) {
		const double c[3] = {pts[i].x, pts[i].y, pts[i].z};
		for (int a = 0; a < 3; ++a) {
			lo[a] = std::min(lo[a], c[a]);
			hi[a] = std::max(hi[a], c[a]);
		}
		for (int k = 0; k < 7; ++k) {
			const double v = dirs[k][0] * c[0] + dirs[k][1] * c[1] + dirs[k][2] * c[2];
			if (v < vmin[k]) {
				vmin[k] = v;
				emin[k] = i;
			}
			if (v > vmax[k]) {
				vmax[k] = v;
				emax[k] = i;
			}
		}
	}
	std::vector<size_t> ext;
	for (int k = 0; k < 7; ++k) {
		ext.push_back(emin[k]);
		ext.push_back(emax[k]);
	}
	float L = 0;
	for (size_t a = 0; a < ext.size(); ++a)
		for (size_t b = 0; b < ext.size(); ++b)
			if (ext[a] < ext[b])
				L = std::max(L, lensq(pts[ext[b]] - pts[ext[a]]));
	// the rounding of each difference (1 + 2^-24) and of the squares and
	// sums (four more roundings) stays far inside this margin
	const double margin = 1.0 + 1e-5;
	std::vector<uint32_t> keep;
	for (size_t i = 0; i < pts.size(); ++i) {
		const double c[3] = {pts[i].x, pts[i].y, pts[i].z};
		double ub = 0;
		for (int a = 0; a < 3; ++a) {
			const double e = std::max(hi[a] - c[a], c[a] - lo[a]);
			ub += e * e;
		}
		if (ub * margin >= (double)L)
			keep.push_back((uint32_t)i);
	}
	return scan(&keep);
}

bool SplitTree::pick_plane(Node& node)
{
	// unique vertices in first-seen order, equality as float == (+0 and -0
	// alike, NaN never equal): an open-addressing table of indices into
	// pts, reused by the thread's later nodes
	std::vector<f3> pts;
	pts.reserve(3 * node.tris.size());
	static thread_local std::vector<uint32_t> slots;
	size_t cap = 16;
	while (cap < 6 * node.tris.size())
		cap <<= 1;
	slots.assign(cap, 0u);
	const size_t mask = cap - 1;
	PointKeyHash hash;
	auto add = [&](f3 p) {
		if (p.x != p.x || p.y != p.y || p.z != p.z) {
			pts.push_back(p);
			return;
		}
		const PointKey k{canon_bits(p.x), canon_bits(p.y), canon_bits(p.z)};
		for (size_t h = hash(k) & mask;; h = (h + 1) & mask) {
			const uint32_t e = slots[h];
			if (e == 0) {
				pts.push_back(p);
				slots[h] = (uint32_t)pts.size();
				return;
			}
			const f3 q = pts[e - 1];
			if (PointKey{canon_bits(q.x), canon_bits(q.y), canon_bits(q.z)} == k)
				return;
		}
	};
	for (const Tri& t : node.tris) {
		add(t.p1);
		add(t.p2);
		add(t.p3);
	}
	f3 center = mk3(0, 0, 0);
	for (size_t i = 0; i < pts.size(); ++i)
		center += pts[i];
	const f3 dir = farthest_pair_dir(pts, brute_pairs_);
	if (!pts.empty())
		center = center / (float)pts.size();
	node.N = normalized(dir);
	node.D = dot(node.N, center);
	int na = 0, nb = 0;
	for (const Tri& t : node.tris) {
		float d1 = dot(t.p1, node.N) - node.D;
		float d2 = dot(t.p2, node.N) - node.D;
		float d3 = dot(t.p3, node.N) - node.D;
		if (d1 * d2 > -kSmall && d2 * d3 > -kSmall) {
			if (d1 > 0)
				na++;
			else
				nb++;
		} else {
			na += (d1 > 0) + (d2 > 0) + (d3 > 0);
			nb += (d1 <= 0) + (d2 <= 0) + (d3 <= 0);
		}
	}
	if (na != 0 && nb != 0 && (float)(na + nb) < (float)node.tris.size() * 1.7f) {
		float q = (float)na / (float)nb;
		if (q < 1)
			q = 1 / q;
		return q < 3.0f;
	}
	return false;
}

void SplitTree::split(int32_t node, int depth)
{
	if (!(nodes_[node].tris.size() > kSplitMaxTris && depth < kSplitMaxDepth))
		return;
	if (!pick_plane(nodes_[node]))
		return;
	Node front, back;
	const Node& n = nodes_[node];
	for (const Tri& t : n.tris) {
		float p1 = dot(n.N, t.p1) - n.D, p2 = dot(n.N, t.p2) - n.D, p3 = dot(n.N, t.p3) - n.D;
		if (p1 * p2 >= -kSmall && p1 * p3 >= -kSmall) {
			(p1 + p2 + p3 > 0 ? front : back).tris.push_back(t);
		} else {
			front.tris.push_back(t);
			back.tris.push_back(t);
		}
	}
	if (front.tris.size() + back.tris.size() > n.tris.size() * 3 / 2)
		return;
	int32_t fi = (int32_t)nodes_.size();
	nodes_.push_back(std::move(front));
	nodes_.push_back(std::move(back));
	nodes_[node].front = fi;
	nodes_[node].back = fi + 1;
	std::vector<Tri>().swap(nodes_[node].tris);
	split(fi, depth + 1);
	split(fi + 1, depth + 1);
}

float SplitTree::node_intersect(int32_t node, f3 from, f3 to, f3* outnormal) const
{
	const Node& n = nodes_[node];
	if (n.front != -1) {
		float df = dot(from, n.N) - n.D;
		float dt = dot(to, n.N) - n.D;
		if (df < 0) {
			float hit = node_intersect(n.back, from, to, outnormal);
			if (hit < 1.0f)
				return hit;
			return dt > 0 ? node_intersect(n.front, from, to, outnormal) : 2.0f;
		}
		float hit = node_intersect(n.front, from, to, outnormal);
		if (hit < 1.0f)
			return hit;
		return dt < 0 ? node_intersect(n.back, from, to, outnormal) : 2.0f;
	}
	float closest = 2.0f;
	const Tri* ct = nullptr;
	const f3 dir = to - from;
	for (size_t k = 0; k < n.rt.size(); ++k) {
		float hit = seg_tri_pre(from, dir, n.rt[k]); // = seg_tri(from, to, ...)
		if (hit < closest) {
			closest = hit;
			ct = &n.tris[n.rti[k]];
		}
	}
	if (outnormal && ct)
		*outnormal = normalized(cross(ct->p3 - ct->p1, ct->p2 - ct->p1));
	return closest;
}

size_t SplitTree::num_leaf_tris() const
{
	size_t n = 0;
	for (const Node& nd : nodes_)
		if (nd.front == -1)
			n += nd.rt.size();
	return n;
}

int SplitTree::flatten(FlatSplitNode* out, FlatSplitTri* tris) const
{
	int depth = 0;
	size_t nt = 0;
	std::vector<std::pair<int32_t, int>> st;
	if (!nodes_.empty())
		st.push_back({0, 1});
	while (!st.empty()) {
		auto [k, d] = st.back();
		st.pop_back();
		depth = std::max(depth, d);
		const Node& n = nodes_[k];
		FlatSplitNode& f = out[k];
		f.nx = n.N.x; f.ny = n.N.y; f.nz = n.N.z; f.d = n.D;
		f.front = n.front;
		f.back = n.back;
		f.leaf_off = f.leaf_cnt = 0;
		if (n.front != -1) {
			st.push_back({n.front, d + 1});
			st.push_back({n.back, d + 1});
			continue;
		}
		f.leaf_off = (int32_t)nt;
		f.leaf_cnt = (int32_t)n.rt.size();
		for (size_t q = 0; q < n.rt.size(); ++q) {
			const Tri& t = n.tris[n.rti[q]];
			// the normal node_intersect reports for this triangle
			f3 hn = normalized(cross(t.p3 - t.p1, t.p2 - t.p1));
			tris[nt++] = FlatSplitTri{n.rt[q], hn.x, hn.y, hn.z, 0.0f};
		}
	}
	return depth;
}

float SplitTree::intersect(f3 from, f3 to, f3* outnormal) const
{
	if (nodes_.empty())
		return 2.0f;
	return node_intersect(0, from, to, outnormal);
}

// The split tree of a triangle set built with the pruned and with the full
// scan of vertex pairs; true when both give the same flat tree (nodes and
// leaf triangles, byte for byte).
bool split_tree_compare(const Tri* tris, size_t count, double* ms_pruned, double* ms_full)
{
	using clk = std::chrono::steady_clock;
	SplitTree a, b;
	auto t0 = clk::now();
	a.set(tris, count, false);
	auto t1 = clk::now();
	b.set(tris, count, true);
	auto t2 = clk::now();
	if (ms_pruned)
		*ms_pruned = std::chrono::duration<double, std::milli>(t1 - t0).count();
	if (ms_full)
		*ms_full = std::chrono::duration<double, std::milli>(t2 - t1).count();
	if (a.num_nodes() != b.num_nodes() || a.num_leaf_tris() != b.num_leaf_tris())
		return false;
	std::vector<FlatSplitNode> na(a.num_nodes()), nb(b.num_nodes());
	std::vector<FlatSplitTri> ta(a.num_leaf_tris()), tb(b.num_leaf_tris());
	const int da = a.flatten(na.data(), ta.data()), db = b.flatten(nb.data(), tb.data());
	return da == db && memcmp(na.data(), nb.data(), na.size() * sizeof(FlatSplitNode)) == 0 &&
	       memcmp(ta.data(), tb.data(), ta.size() * sizeof(FlatSplitTri)) == 0;
}

} // namespace lgx
