// Bounding volume hierarchy used by the gfx950 kernels.
//
// Binary BVH built on the host with binned SAH, flattened depth-first into
// 64-byte nodes.  Each node stores the boxes of both children, so one node
// fetch (four dwordx4 loads) lets a lane order and cull both children.  A
// child reference >= 0 is an inner node; < 0 encodes a leaf as
// ~((first_triangle << 4) | (count - 1)), count <= 16.  Triangles are stored
// in leaf order in two precomputed forms (DistTri for distance queries,
// RayTri for segment queries).
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

struct alignas(16) BvhNode {
	float lo0x, lo0y, lo0z, hi0x;
	float hi0y, hi0z, lo1x, lo1y;
	float lo1z, hi1x, hi1y, hi1z;
	int32_t c0, c1, pad0, pad1;
};
static_assert(sizeof(BvhNode) == 64, "node size");

constexpr int kLeafMax = 2; // measured on the terrain workload: 2 beats 1, 4 and 8

LHD int32_t leaf_ref(uint32_t first, uint32_t count) { return ~(int32_t)((first << 4) | (count - 1)); }
LHD uint32_t leaf_first(int32_t ref) { return ((uint32_t)~ref) >> 4; }
LHD uint32_t leaf_count(int32_t ref) { return (((uint32_t)~ref) & 15u) + 1u; }

struct BvhBuild {
	std::vector<BvhNode> nodes;
	std::vector<uint32_t> order; // leaf-order position -> input triangle index
	f3 lo, hi;                   // scene bounds (unpadded)
	int depth = 0;
};

// tris: 9 floats per triangle.
BvhBuild build_bvh(const float* tris, size_t ntri);

} // namespace lgx
