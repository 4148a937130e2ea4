// Normal-map data (lighter.cpp:505-513, 543-551, 590-598 for the per-light
// contributions, 959-1030 for their combination per texel), shared by the
// host restatement (pipeline.cpp) and the GPU pass (kernels.hip), so that
// both give the same bits.  pow and acos are evaluated in double and
// rounded once on both sides (the GPU's direct light does the same).
#pragma once

#include <math.h>

#include "device.h"
#include "vmath.h"

namespace lgx {

LHD float pow_once(float x, float y) { return y == 1.0f ? x : (float)pow((double)x, (double)y); }
LHD float acos_once(float x) { return (float)acos((double)x); }

// The contribution of light L at a texel (position SP, normal SN) with
// shadow factor vis: the light's direction scaled by its weight at the
// texel and the mean of its colour.  False when the weight is not positive
// (no contribution).
LHD bool nm_contrib(const DevLight& L, f3 SP, f3 SN, float vis, f3* out)
{
	const float bright = (L.color[0] + L.color[1] + L.color[2]) * (1.0f / 3.0f);
	if (L.type == 3) { // directional
		const f3 dir = ld3(L.dir);
		const float fn = tmax(0.0f, dot(dir, SN));
		const float factor = vis * fn * bright;
		if (!(factor > 0))
			return false;
		*out = dir * factor;
		return true;
	}
	f3 s2l = ld3(L.pos) - SP;
	const float dist = length(s2l);
	if (dist)
		s2l = s2l / dist;
	const float fd = pow_once(1 - tmin(1.0f, dist / L.range), L.power);
	const float fn = tmax(0.0f, dot(s2l, SN));
	float factor;
	if (L.type == 2) { // spot
		const float ang = acos_once(tmin(1.0f, dot(s2l, -ld3(L.dir))));
		float fdir = tmax(0.0f, tmin(1.0f, (ang - L.spot_out_rad) / L.spot_diff));
		fdir = pow_once(fdir, L.spot_curve);
		if (fd * fn * fdir <= 0)
			return false;
		factor = fd * fdir * vis * fn * bright;
	} else {
		if (fd * fn <= 0)
			return false;
		factor = fd * vis * fn * bright;
	}
	if (!(factor > 0))
		return false;
	*out = s2l * factor;
	return true;
}

// One contribution's effect on the texel's min-dot term, given the texel's
// mean normal (the sum of the scaled sample normal and the contributions,
// divided by their count): the cosine between the two directions, blended
// towards 1 as the contribution is weak against the mean.
LHD float nm_mindot_factor(f3 mean, f3 c)
{
	float d = tmax(dot(normalized(mean), normalized(c)), 0.0f);
	return lerpf(1.0f, d, tmin(length(c) / length(mean), 1.0f));
}

} // namespace lgx
