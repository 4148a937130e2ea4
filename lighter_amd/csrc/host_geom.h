// Host-side geometry for the bake stages that run on the CPU before and
// after the GPU hot path: instance transforms, lightmap rasterisation, the
// sample placement fixes, and the two search structures those fixes query.
//
// The sample positions every kernel consumes come out of these stages, and
// they depend on visit order (OffsetSample moves the sample while the tree is
// being walked), so the structures here reproduce the reference's own tree
// layouts and traversal orders:
//  * BoxTree   - lighter_int.hpp:706-821, lighter_math.cpp:653-781
//  * SplitTree - lighter_int.hpp:623-677, lighter_math.cpp:428-614
// The hot-path kernels use their own BVH (bvh.cpp); these trees serve the
// sample placement, on the host or flattened for the placement kernel
// (place.hip), which walks them in the same orders.
#pragma once

#include <stdint.h>

#include <vector>

#include "geom.h"

namespace lgx {

// ---------------------------------------------------------------------------
// matrices (lighter_int.hpp:490-556, lighter_math.cpp:19-147)
// ---------------------------------------------------------------------------
struct Mat4 {
	float m[4][4];
};

// Inverse by cofactor expansion; false (and out untouched) when singular.
bool mat4_invert(const Mat4& a, Mat4& out);
// Transform with row vectors: v.x*m[0] + v.y*m[1] + v.z*m[2] + m[3]*w.
f3 mat4_transform(const Mat4& m, f3 v, float w);
// Inverse-transpose of the upper 3x3 (lighter_int.hpp:534-545).
Mat4 normal_matrix(const Mat4& m);

// Heron's formula (lighter_math.cpp:150-173).
float triangle_area(f3 p1, f3 p2, f3 p3);
float triangle_area(f2 p1, f2 p2, f2 p3);
// World area per lightmap texel of a triangle (lighter_math.cpp:175-181).
float sample_area(f2 t1, f2 t2, f2 t3, f3 p1, f3 p2, f3 p3);

struct Tri {
	f3 p1, p2, p3;
};

inline bool tri_useful(const Tri& t) { return !near_zero(cross(t.p2 - t.p1, t.p3 - t.p1)); }
inline f3 tri_normal(const Tri& t) { return normalized(cross(t.p3 - t.p1, t.p2 - t.p1)); }

struct Box {
	f3 mn, mx;
	bool valid() const { return mn.x <= mx.x && mn.y <= mx.y && mn.z <= mx.z; }
	f3 center() const { return (mn + mx) * 0.5f; }
	float volume() const { return (mx.x - mn.x) * (mx.y - mn.y) * (mx.z - mn.z); }
};

inline Box tri_box(const Tri& t)
{
	return Box{mk3(tmin(t.p1.x, tmin(t.p2.x, t.p3.x)), tmin(t.p1.y, tmin(t.p2.y, t.p3.y)),
	               tmin(t.p1.z, tmin(t.p2.z, t.p3.z))),
	           mk3(tmax(t.p1.x, tmax(t.p2.x, t.p3.x)), tmax(t.p1.y, tmax(t.p2.y, t.p3.y)),
	               tmax(t.p1.z, tmax(t.p2.z, t.p3.z)))};
}

// ---------------------------------------------------------------------------
// BoxTree: binary tree in depth-first order; node n's children are n+1 and
// ch; item lists (count-prefixed) hang off leaves and, for items too large
// to split, off inner nodes.
// ---------------------------------------------------------------------------
class BoxTree {
public:
	struct Node {
		f3 mn, mx;
		int32_t ch;
		int32_t ido;
	};
	void build(const Box* boxes, size_t count);
	const Node& root() const { return nodes_[0]; }
	const std::vector<Node>& nodes() const { return nodes_; }
	const std::vector<int32_t>& items() const { return items_; }
	// nodes on the longest root-to-leaf path (the root alone is 1)
	int depth() const;

	// Items of every node whose box overlaps [qmn, qmx]; fn(ids, count).
	template <class F> void query(f3 qmn, f3 qmx, F& fn) const
	{
		int32_t stack[128];
		int sp = 0;
		stack[sp++] = 0;
		while (sp) {
			int32_t n = stack[--sp];
			const Node& N = nodes_[n];
			if (box_disjoint(qmn, qmx, N.mn, N.mx))
				continue;
			if (N.ido != -1)
				fn(&items_[N.ido + 1], items_[N.ido]);
			if (N.ch != -1) {
				stack[sp++] = N.ch;
				stack[sp++] = n + 1;
			}
		}
	}
	// Items of every node the segment's slab test accepts, until fn returns
	// false.  The slab test skips the axes where the direction is zero, so an
	// axis-aligned segment accepts every box in a whole slab of space; the
	// callers look for triangles the segment meets, so boxes disjoint from the
	// segment's padded bounding box are passed over as well (same visit order
	// for the rest, same answers).
	template <class F> void ray_query(const RaySeg& r, F& fn) const
	{
		int32_t stack[128];
		int sp = 0;
		stack[sp++] = 0;
		while (sp) {
			int32_t n = stack[--sp];
			const Node& N = nodes_[n];
			if (box_disjoint(r.qmn, r.qmx, N.mn, N.mx) || !ray_box(r, N.mn, N.mx))
				continue;
			if (N.ido != -1 && !fn(&items_[N.ido + 1], items_[N.ido]))
				return;
			if (N.ch != -1) {
				stack[sp++] = N.ch;
				stack[sp++] = n + 1;
			}
		}
	}
	template <class F> void all(F& fn) const
	{
		for (size_t i = 0; i < items_.size(); i += 1 + items_[i])
			fn(&items_[i + 1], items_[i]);
	}

private:
	void make_node(int32_t node, const Box* boxes, int32_t* idx, size_t count, int depth);
	std::vector<Node> nodes_;
	std::vector<int32_t> items_;
};

// Per-instance triangle set with its BoxTree (lighter_int.hpp:823-833).
class TriSet {
public:
	void set(const Tri* tris, size_t count);
	const std::vector<Tri>& tris() const { return tris_; }
	const BoxTree& tree() const { return tree_; }
	const std::vector<Box>& boxes() const { return boxes_; }
	// Closest segment hit; hit index through outtid (lighter_math.cpp:835-871).
	float ray_dist(f3 from, f3 to, int32_t* outtid) const;
	// Push a sample out of concave edges (lighter_math.cpp:991-1044).
	void offset_sample(f3& P, f3 N, float dist) const;

private:
	BoxTree tree_;
	std::vector<Tri> tris_;
	std::vector<Box> boxes_; // per triangle, as the tree was built from
};

// ---------------------------------------------------------------------------
// SplitTree: the plane-split tree used to move samples in front of
// overlapping faces.  Triangles that straddle a plane go to both sides.
// ---------------------------------------------------------------------------
// Flat form of a SplitTree for the placement kernel: inner nodes hold the
// plane and the child indices, leaves (front == -1) a range of leaf
// triangles with the segment-test operands and the normal a hit reports;
// node and triangle indices are local to the tree.
struct FlatSplitNode {
	float nx, ny, nz, d;
	int32_t front, back, leaf_off, leaf_cnt;
};
struct FlatSplitTri {
	RayTri rt;
	float nx, ny, nz, pad;
};

class SplitTree {
public:
	// brute_pairs: pick every plane with the full scan of vertex pairs (the
	// reference's loop) instead of the pruned scan; both give the same tree
	void set(const Tri* tris, size_t count, bool brute_pairs = false);
	bool empty() const { return nodes_.empty(); }
	size_t num_nodes() const { return nodes_.size(); }
	size_t num_leaf_tris() const;
	// Writes the flat form (num_nodes() nodes, num_leaf_tris() triangles;
	// indices local to this tree); returns the depth (nodes on the longest
	// root-to-leaf path, 0 for an empty tree).
	int flatten(FlatSplitNode* nodes, FlatSplitTri* tris) const;
	// Closest hit in the first subtree along the segment that has one
	// (lighter_math.cpp:490-533); 2 when nothing is hit.
	float intersect(f3 from, f3 to, f3* outnormal) const;

private:
	struct Node {
		f3 N = {0, 0, 0}; // the plane (leaves that were never split keep zeros)
		float D = 0;
		int32_t front = -1, back = -1;
		std::vector<Tri> tris;
		// leaves: the segment-test operands of the triangles with a usable
		// normal, in list order, and their positions in tris
		std::vector<RayTri> rt;
		std::vector<int32_t> rti;
	};
	void split(int32_t node, int depth);
	bool pick_plane(Node& n);
	bool brute_pairs_ = false; // scan every vertex pair when picking a plane (tests)
	float node_intersect(int32_t node, f3 from, f3 to, f3* outnormal) const;
	std::vector<Node> nodes_;
};

// Builds the split tree of a triangle set with the pruned and with the full
// scan of vertex pairs; true when the two flat trees are identical.
bool split_tree_compare(const Tri* tris, size_t count, double* ms_pruned, double* ms_full);

} // namespace lgx
