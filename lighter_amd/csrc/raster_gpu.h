// Lightmap rasterisation of every instance in one set of GPU launches
// (raster.hip): the per-pixel rule of raster.h, then the texels whose
// interpolated normal is not zero compacted into samples in atlas order and
// texel order, as the host rasteriser gives them.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <memory>
#include <vector>

#include "raster.h"

namespace lgx {

struct RasterAtlasSize {
	uint32_t w, h;
};

// Every atlas's samples, atlas after atlas: position (3 floats), normalised
// normal (3), (uv0, part, area) (4) and the texel index x + y w; atlas a's
// samples are first[a] .. first[a] + count[a] - 1.
struct RasterOut {
	std::vector<uint64_t> first, count;
	std::unique_ptr<float[]> pos, nrm, inf;
	std::unique_ptr<uint32_t[]> loc;
	uint64_t total = 0;
};

struct RasterTimes {
	double kernel_ms = 0; // the three launches (HIP events)
	uint64_t texels = 0, samples = 0;
};

// tris: every atlas's triangles in drawing order (part and index order),
// the atlases one after another; tri_atlas: the atlas of each triangle.
// Throws std::runtime_error on HIP errors and malformed input.
void rasterize_gpu(const RasterTri* tris, const uint32_t* tri_atlas, size_t ntris,
                   const std::vector<RasterAtlasSize>& sizes, RasterOut& out, RasterTimes* times);

} // namespace lgx
