// Bounding volume hierarchy used by the gfx950 kernels.
//
// Binary BVH built on the host with binned SAH, then collapsed into 4-wide
// nodes of 128 bytes flattened depth-first: the four children's boxes as
// six float4 planes (lo x, lo y, lo z, hi x, hi y, hi z over the children)
// and their references, so one node fetch (seven dwordx4 loads) lets a lane
// order and cull four children, and a query visits about half as many
// nodes as in the binary tree.  A child reference >= 0 is an inner node;
// < 0 encodes a leaf as ~((first_triangle << 4) | (count - 1)),
// count <= 16.  Unused child slots hold a degenerate box far outside any
// scene (never entered by either box test).  Triangles are stored in leaf
// order in two precomputed forms (DistTri for distance queries, RayTri for
// segment queries).
//
// Queries are order independent (minimum distance, minimum hit parameter,
// any hit), so this structure gives the same answers as the reference's
// per-instance trees; node boxes are padded outward by a small relative
// margin so that rounding in the box tests never culls a triangle the exact
// test would accept.
#pragma once

#include <stdint.h>

#include <vector>

#include "geom.h"

namespace lgx {

// binary node of the build (both children's boxes and references)
struct alignas(16) BvhNode {
	float lo0x, lo0y, lo0z, hi0x;
	float hi0y, hi0z, lo1x, lo1y;
	float lo1z, hi1x, hi1y, hi1z;
	int32_t c0, c1, pad0, pad1;
};
static_assert(sizeof(BvhNode) == 64, "node size");

struct alignas(16) WideNode {
	float lox[4], loy[4], loz[4];
	float hix[4], hiy[4], hiz[4];
	int32_t child[4];
	int32_t pad[4];
};
static_assert(sizeof(WideNode) == 128, "wide node size");
constexpr float kFarBox = 1e37f; // corner of the unused slots' degenerate box

constexpr int kLeafMax = 2; // measured on the terrain workload: 2 beats 1, 4 and 8
constexpr int kSahBins = 32;     // SAH bins per axis of the build the kernels traverse
constexpr int kSahMaxDepth = 40; // from this binary depth on, subtrees split by position (depth bound)

LHD int32_t leaf_ref(uint32_t first, uint32_t count) { return ~(int32_t)((first << 4) | (count - 1)); }
LHD uint32_t leaf_first(int32_t ref) { return ((uint32_t)~ref) >> 4; }
LHD uint32_t leaf_count(int32_t ref) { return (((uint32_t)~ref) & 15u) + 1u; }

struct BvhBuild {
	std::vector<WideNode> nodes; // 4-wide, root at 0
	std::vector<uint32_t> order; // leaf-order position -> input triangle index
	f3 lo, hi;                   // scene bounds (unpadded)
	int depth = 0;               // levels of wide nodes on the longest path
};

// tris: 9 floats per triangle.  Binned SAH (kSahBins per axis, the split of
// least SAH cost over the axes with a centroid extent), stable partitions,
// a split by position where all centroids coincide or below kSahMaxDepth,
// leaves of at most kLeafMax triangles; then collapsed into wide nodes.
BvhBuild build_bvh(const float* tris, size_t ntri);

// The same build with the binary tree and the leaf order made on the GPU
// (bvh_gpu.hip) from the triangles on the device (d_tris, 9 floats each;
// stream: a hipStream_t): every node, box and leaf of the host build,
// collapsed on the host.  d_order (ntri entries, device) receives the leaf
// order; kernel_ms the device build's time (HIP events).
BvhBuild build_bvh_gpu(const float* d_tris, size_t ntri, uint32_t* d_order, void* stream, double* kernel_ms);

// Both builds of one triangle soup (host array): out[0] = 1 when the wide
// node arrays are identical byte for byte, out[1] = 1 when the leaf orders
// and depths are, out[2] host ms, out[3] GPU ms (wall), out[4] GPU kernels
// and read-backs ms, out[5] wide nodes.
void compare_bvh_builds(const float* tris, size_t ntri, double out[6]);

// The collapse of a binary tree (root at 0) into 4-wide nodes; sets nodes
// and depth of out.
void collapse_bvh(const std::vector<BvhNode>& bin, BvhBuild& out);
// A binary node from its children's (unpadded) boxes and references, the
// boxes padded as the build pads them.
BvhNode bvh_node(const float lo0[3], const float hi0[3], int32_t c0, const float lo1[3], const float hi1[3],
                 int32_t c1);

// A plain binary binned-SAH tree (bins per axis, at most leaf_max triangles
// per leaf), not collapsed: the reference tree shape the traversal byte
// counts are stated on.  depth counts node levels, leaves included.
struct BinaryBvh {
	std::vector<BvhNode> nodes; // root at 0
	std::vector<uint32_t> order;
	f3 lo, hi;
	int depth = 0;
};
BinaryBvh build_bvh_binary(const float* tris, size_t ntri, int bins, int leaf_max);

} // namespace lgx
