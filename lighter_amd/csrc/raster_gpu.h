// Lightmap rasterisation of every instance in one set of GPU launches
// (raster.hip): the per-pixel rule of raster.h, then the texels whose
// interpolated normal is not zero compacted into samples in atlas order and
// texel order, as the host rasteriser gives them.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "raster.h"

namespace lgx {

struct RasterAtlasSize {
	uint32_t w, h;
};

// One atlas's samples: position, normalised normal, (uv0, part, area) and
// the texel index x + y w.
struct RasterOut {
	uint64_t first = 0, count = 0;
	std::vector<float> pos, nrm, inf;
	std::vector<uint32_t> loc;
};

struct RasterTimes {
	double kernel_ms = 0; // the three launches (HIP events)
	uint64_t texels = 0, samples = 0;
};

// tris: every atlas's triangles in drawing order (part and index order),
// the atlases one after another; tri_atlas: the atlas of each triangle.
// Throws std::runtime_error on HIP errors and malformed input.
void rasterize_gpu(const std::vector<RasterTri>& tris, const std::vector<uint32_t>& tri_atlas,
                   const std::vector<RasterAtlasSize>& sizes, std::vector<RasterOut>& out, RasterTimes* times);

} // namespace lgx
