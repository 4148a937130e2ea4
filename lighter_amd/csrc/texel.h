// Per-texel operations of the lightmap export (lighter.cpp:866-958,
// lighter_math.cpp:357-424), shared by the GPU export kernels (finalize.hip)
// and the host export (pipeline.cpp).  Each function produces one output
// texel from its inputs, so both exports run the same arithmetic in the same
// order and give the same bits (-ffp-contract=off on both sides).
#pragma once

#include <math.h>
#include <stdint.h>

#include <vector>

#include "vmath.h"

namespace lgx {

// An atlas image in texel order (x + y w), 3 floats per texel, and the
// flags of the texels that hold a colour.
struct AtlasView {
	const float* rgb;
	const uint8_t* set;
	uint32_t w, h;
};

LHD bool atlas_set(const AtlasView& a, uint32_t x, uint32_t y)
{
	return x < a.w && y < a.h && a.set[x + (uint64_t)y * a.w]; // x, y wrap past zero to huge values
}

LHD f3 atlas_rgb(const AtlasView& a, uint32_t x, uint32_t y) { return ld3(a.rgb + 3 * (x + (uint64_t)y * a.w)); }

// One round of edge extrapolation at an unset texel.  The neighbours are
// visited in the order -x, +x, -y, +y; a set neighbour contributes its
// colour, continued one step along the line from the texel beyond it when
// that one is set as well (lerp(far, near, 2) = 2 near - far).  The texel
// becomes the mean of the contributions, clamped at zero; false (and *out
// untouched) when no neighbour is set.  Reads only texels flagged set, which
// a round does not write, so the texels of a round can go in any order.
LHD bool extrapolate_texel(const AtlasView& a, uint32_t x, uint32_t y, f3* out)
{
	const int step[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
	f3 acc = mk3(0, 0, 0);
	int n = 0;
	for (int k = 0; k < 4; ++k) {
		const uint32_t nx = x + (uint32_t)step[k][0], ny = y + (uint32_t)step[k][1];
		if (!atlas_set(a, nx, ny))
			continue;
		const uint32_t fx = nx + (uint32_t)step[k][0], fy = ny + (uint32_t)step[k][1];
		const f3 near = atlas_rgb(a, nx, ny);
		acc += atlas_set(a, fx, fy) ? lerp3(atlas_rgb(a, fx, fy), near, 2.0f) : near;
		++n;
	}
	if (n == 0)
		return false;
	const f3 m = acc / (float)n;
	*out = mk3(tmax(m.x, 0.0f), tmax(m.y, 0.0f), tmax(m.z, 0.0f));
	return true;
}

// One texel of a Gaussian pass along a row of `len` texels (3 floats each):
// the taps -ext .. ext in order, the row's end texels repeated beyond its
// ends.  The pass writes its result transposed, so two passes blur both
// axes and restore the layout.
LHD f3 blur_texel(const float* row, uint32_t len, int x, const float* taps, int ext)
{
	float r = 0.0f, g = 0.0f, b = 0.0f;
	for (int i = -ext; i <= ext; ++i) {
		int at = x + i;
		at = at < 0 ? 0 : (at >= (int)len ? (int)len - 1 : at);
		const float w = taps[i + ext];
		r += row[3 * at + 0] * w;
		g += row[3 * at + 1] * w;
		b += row[3 * at + 2] * w;
	}
	return mk3(r, g, b);
}

// One texel of the half-size image: the mean of the 2 x 2 block at (2x, 2y)
// of an sw x sh image, indices wrapping at the edges (odd sizes).
LHD f3 half_texel(const float* img, uint32_t sw, uint32_t sh, uint32_t x, uint32_t y)
{
	const uint32_t xa = (2 * x) % sw, xb = (2 * x + 1) % sw, ya = (2 * y) % sh, yb = (2 * y + 1) % sh;
	const f3 s = ld3(img + 3 * (xa + (uint64_t)ya * sw)) + ld3(img + 3 * (xb + (uint64_t)ya * sw)) +
	             ld3(img + 3 * (xa + (uint64_t)yb * sw)) + ld3(img + 3 * (xb + (uint64_t)yb * sw));
	return s * 0.25f;
}

// The blur's taps: the Gaussian of deviation `radius` at -ext .. ext,
// normalised by its sum (accumulated in tap order).
inline std::vector<float> blur_taps(int ext, float radius)
{
	std::vector<float> taps(2 * (size_t)ext + 1);
	const float peak = 1.0f / sqrtf(2.0f * kPiF * radius * radius);
	float total = 0.0f;
	for (size_t k = 0; k < taps.size(); ++k) {
		const float u = (float)((int)k - ext) / radius;
		taps[k] = expf(-0.5f * (u * u)) * peak;
		total += taps[k];
	}
	for (float& t : taps)
		t /= total;
	return taps;
}

} // namespace lgx
