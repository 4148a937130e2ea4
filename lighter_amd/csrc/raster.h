// Lightmap rasterisation of one triangle (lighter.cpp:32-66, 392-420,
// lighter_math.cpp:245-302), per pixel, shared by the GPU rasteriser
// (raster.hip) and the host one (pipeline.cpp).
//
// A triangle in texel coordinates covers pixel (x, y) of its box when the
// point (x, y) lies within `margin` of the inside of all three edge lines,
// for either winding; a covered pixel takes the triangle's attributes at the
// barycentric coordinates of (x, y) clamped to [0, 1] each (not projected
// onto the triangle).  An instance is drawn twice, every triangle with
// margin 0.5 + kSmall and then every triangle with margin 0, in part and
// index order, later triangles overwriting earlier ones: a pixel's final
// value comes from the last (pass, triangle) that covers it.
#pragma once

#include <stdint.h>

#include "vmath.h"

namespace lgx {

// Margin of the first of the two passes.
constexpr float kRasterMargin0 = 0.5f + kSmall;

struct RasterTri {
	f2 o;             // first vertex
	f2 eu, ev;        // edges to the second and third vertex
	float area2;      // cross(eu, ev); 0: the triangle draws nothing
	f2 edge_n[3];     // unit normals of the edges 1-2, 2-3 and 3-1 (the last pointing the other way)
	float edge_d[3];  // edge_n[k] . (a vertex of edge k)
	float lo_x, hi_x, lo_y, hi_y; // extent of the three vertices
	f3 pos, pos_u, pos_v;         // attribute at the first vertex and its changes along eu, ev
	f3 nrm, nrm_u, nrm_v;
	f4 inf, inf_u, inf_v;
};

LHD RasterTri raster_setup(f2 p1, f2 p2, f2 p3, f3 pa, f3 pb, f3 pc, f3 na, f3 nb, f3 nc, f4 ia, f4 ib, f4 ic)
{
	RasterTri t;
	t.o = p1;
	t.eu = p2 - p1;
	t.ev = p3 - p1;
	t.area2 = cross2(t.eu, t.ev);
	t.edge_n[0] = normalized2(perp2(t.eu));
	t.edge_n[1] = normalized2(perp2(p3 - p2));
	t.edge_n[2] = -normalized2(perp2(t.ev));
	t.edge_d[0] = dot2(t.edge_n[0], p1);
	t.edge_d[1] = dot2(t.edge_n[1], p2);
	t.edge_d[2] = dot2(t.edge_n[2], p3);
	t.lo_x = tmin(p1.x, tmin(p2.x, p3.x));
	t.hi_x = tmax(p1.x, tmax(p2.x, p3.x));
	t.lo_y = tmin(p1.y, tmin(p2.y, p3.y));
	t.hi_y = tmax(p1.y, tmax(p2.y, p3.y));
	t.pos = pa;
	t.pos_u = pb - pa;
	t.pos_v = pc - pa;
	t.nrm = na;
	t.nrm_u = nb - na;
	t.nrm_v = nc - na;
	t.inf = ia;
	t.inf_u = ib - ia;
	t.inf_v = ic - ia;
	return t;
}

// Pixel box [x0, x1] x [y0, y1] of the triangle grown by `margin`, clipped
// to a w x h image (bounds truncated toward zero); false when it misses the
// image or the triangle is degenerate.
struct PixelBox {
	int32_t x0, x1, y0, y1;
};

LHD bool raster_box(const RasterTri& t, float margin, int32_t w, int32_t h, PixelBox* b)
{
	int32_t x1 = (int32_t)(t.hi_x + margin), x0 = (int32_t)(t.lo_x - margin);
	int32_t y1 = (int32_t)(t.hi_y + margin), y0 = (int32_t)(t.lo_y - margin);
	if (x1 < 0 || x0 >= w || y1 < 0 || y0 >= h || t.area2 == 0)
		return false;
	b->x0 = x0 < 0 ? 0 : x0;
	b->x1 = x1 >= w ? w - 1 : x1;
	b->y0 = y0 < 0 ? 0 : y0;
	b->y1 = y1 >= h ? h - 1 : y1;
	return true;
}

LHD bool raster_covers(const RasterTri& t, float margin, int32_t x, int32_t y)
{
	const f2 p = mk2((float)x, (float)y);
	const float s0 = dot2(p, t.edge_n[0]), s1 = dot2(p, t.edge_n[1]), s2 = dot2(p, t.edge_n[2]);
	const bool one_side = s0 <= t.edge_d[0] + margin && s1 <= t.edge_d[1] + margin && s2 <= t.edge_d[2] + margin;
	const bool other_side = s0 + margin >= t.edge_d[0] && s1 + margin >= t.edge_d[1] && s2 + margin >= t.edge_d[2];
	return one_side || other_side;
}

// position, normal (not normalised) and (uv0, part, area) at pixel (x, y)
LHD void raster_values(const RasterTri& t, int32_t x, int32_t y, f3* pos, f3* nrm, f4* inf)
{
	const f2 q = mk2((float)x - t.o.x, (float)y - t.o.y);
	float u = cross2(q, t.ev) / t.area2;
	float v = cross2(t.eu, q) / t.area2;
	u = tmax(tmin(u, 1.0f), 0.0f);
	v = tmax(tmin(v, 1.0f), 0.0f);
	*pos = t.pos + t.pos_u * u + t.pos_v * v;
	*nrm = t.nrm + t.nrm_u * u + t.nrm_v * v;
	*inf = t.inf + t.inf_u * u + t.inf_v * v;
}

} // namespace lgx
