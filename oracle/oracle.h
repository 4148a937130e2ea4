/* CPU oracle for the lightmap-baking hot path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference baker's per-texel shadow march,
 * ambient-occlusion rays and radiosity gathering (archo5/lighter), used by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the
 * checker.  The product (lighter_amd) never links, loads or calls it.
 *
 * Every scene query is brute force over the full shadow-casting triangle list
 * (no acceleration structure), so its results are the exact minimum / any-hit
 * over all triangles, which is what the reference's tree traversals compute.
 *
 * All arithmetic is IEEE binary32 evaluated in source order with no FMA
 * contraction (built with -ffp-contract=off), matching the reference built
 * for x86-64 without -march flags.  Transcendentals (powf, acosf, sinf, cosf)
 * come from the same C library the reference calls.
 *
 * Pinning (tests/test_oracle.py, tests/test_golden.py): each function is
 * checked against the reference compiled from its own sources (oracle/_ref)
 * on seeded inputs, against golden fixtures of the reference's own test
 * scenarios (tests/golden), and against closed forms / brute-force FP64
 * geometry.
 */
#ifndef LIGHTER_ORACLE_H
#define LIGHTER_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Light as stored after ltr_LightAdd (direction already normalised). */
typedef struct orc_light {
	int32_t type; /* 1 point, 2 spot, 3 directional */
	float pos[3];
	float dir[3];
	float color[3];
	float range;
	float power;
	float radius;
	float spot_out_deg;
	float spot_in_deg;
	float spot_curve;
} orc_light;

/* --- primitives --------------------------------------------------------- */
float orc_seg_tri(const float* l1, const float* l2, const float* p1, const float* p2, const float* p3);
float orc_point_tri_dist(const float* p, const float* t0, const float* t1, const float* t2);
int orc_tri_useful(const float* t9);
void orc_spiral(const float* dir, float randoff, int i, int n, float* out3);

/* --- scene queries over a triangle soup (9 floats per triangle) --------- */
float orc_scene_distance(const float* tris, size_t ntri, const float* p);
float orc_shadow_factor(const float* tris, size_t ntri, const float* from, const float* to, float k,
                        uint32_t* steps_out);
float orc_closest_hit(const float* tris, size_t ntri, const float* a, const float* b);
int orc_occluded(const float* tris, size_t ntri, const float* a, const float* b);

/* --- stages ------------------------------------------------------------- */
/* One light over samples [0,S).  lightmap (S*3) is accumulated in place.
 * weight_out (S, optional) receives the scalar multiplying the light colour,
 * or -1 where the reference skips the texel.  Returns the number of shadow
 * marches with a nonzero shading factor (work unit U1). */
size_t orc_direct_light(const float* tris, size_t ntri, size_t S, const float* pos, const float* nrm,
                        const orc_light* L, float* lightmap, float* weight_out);

/* Ambient occlusion over samples [0,S) with one rotation offset per texel. */
void orc_ao(const float* tris, size_t ntri, size_t S, const float* pos, const float* nrm,
            const float* randoff, float ao_distance, float ao_multiplier, float ao_falloff,
            float ao_effect, const float* ao_color, int num_samples, float* lightmap,
            float* ao_factor_out);

/* Radiosity links over N patches (pos, nrm as in the patch geometry).
 * rowmap receives 2N entries {offset, count}; returns the link count, and
 * writes at most cap links. */
size_t orc_rad_links(const float* tris, size_t ntri, size_t N, const float* pos, const float* nrm,
                     uint32_t* rowmap, uint32_t* other, float* factor, size_t cap);

/* One bounce in the reference's order; colors13 holds per patch
 * {diffuse[3], total[3], out[3], in[3], area}. */
void orc_rad_bounce_f32(size_t N, const uint32_t* rowmap, const uint32_t* other, const float* factor,
                        float* colors13);

/* Truncated series total = sum_{k=0..bounces} K^k E in FP64, K = F diag(area)
 * diag(diffuse[c]) per column; column r of E/total uses channel r % 3. */
void orc_rad_series_f64(size_t N, const uint32_t* rowmap, const uint32_t* other, const float* factor,
                        const double* area, const double* diffuse3, const double* E, int R, int bounces,
                        double* total);

#ifdef __cplusplus
}
#endif
#endif
