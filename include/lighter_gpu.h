/*
 * lighter_gpu.h — additional C entry points of liblighter_amd.so.
 *
 * The ltr_* API (lighter.h) is the drop-in boundary and runs a whole bake.
 * The lgx_* functions below have no counterpart in the reference: they expose
 * the same hot-path stages one at a time, for callers that drive the stages
 * themselves (tests, benchmarks, tools) and for the radiosity solver forms
 * and adjoint that the reference does not have.  Each comment names the
 * reference code whose behaviour the stage reproduces.
 *
 * Conventions: plain pointers and sizes, host memory in and out; functions
 * return 1 on success and 0 on failure, with the message available from
 * lgx_LastError() (per calling thread).  All GPU work runs on one HIP device
 * (the current one).
 */
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "lighter.h"

#ifdef __cplusplus
extern "C" {
#endif

LTRAPI const char* lgx_LastError(void);
/* 1 when a HIP device is usable, else 0 (and lgx_LastError says why). */
LTRAPI int lgx_HasDevice(void);

/* ---- staged bake of an ltr_Scene (synchronous, calling thread) -------- */
/* Starts a staged bake: drops the samples, lightmaps and device state of an
 * earlier bake (as ltr_Start does), then runs the stages
 * "transforming spatial data" .. "generating samples"
 * (lighter.cpp:1050-1073): transforms, collision set-up and rasterisation
 * on the host threads; the sample placement (OffsetSample and the
 * correction in front of overlapping faces, lighter.cpp:436-466) on the GPU
 * unless lgx_SetHostStages selected the host. Scenes with no texel to
 * place need no GPU. */
LTRAPI int lgx_Prepare(ltr_Scene* scene);
/* No counterpart in the reference (it runs these stages on its worker
 * threads). Three host stages of the reference also run on the GPU here:
 * the lightmap rasterisation and sample compaction (lighter.cpp:32-66,
 * 392-420), the sample placement (lighter.cpp:436-466) and the lightmap
 * export's image work (extrapolation, blur, downsample;
 * lighter.cpp:849-958). They use the GPU by default, also under ltr_Start;
 * a bit set in mask selects the host restatement instead, which the GPU
 * forms are tested against. Takes effect at the next preparation or
 * export. */
#define LGX_HOST_PLACEMENT 1
#define LGX_HOST_FINALIZE 2
#define LGX_HOST_RASTER 4
LTRAPI int lgx_SetHostStages(ltr_Scene* scene, int mask);
/* No counterpart in the reference. Last GPU rasterisation: out[0] kernels
 * ms (HIP events, the count read-back between them included), out[1] atlas
 * texels, out[2] samples (zeros after a host rasterisation). */
LTRAPI int lgx_RasterTimes(ltr_Scene* scene, double out[3]);
/* No counterpart in the reference. Last GPU placement: out[0] flatten ms,
 * out[1] upload ms, out[2] kernel ms (HIP events), out[3] download ms,
 * out[4] samples placed (zeros after a host placement). */
LTRAPI int lgx_PlacementTimes(ltr_Scene* scene, double out[5]);
/* No counterpart in the reference. Wall ms of the last preparation's steps:
 * out[0] transforms, out[1] collision set-up (each instance's split tree
 * and triangle tree), out[2] the instance tree, out[3] sample generation,
 * and within it out[4] rasterisation, out[5] placement, out[6] the
 * samples' commit to their instances. */
LTRAPI int lgx_PrepareTimes(ltr_Scene* scene, double out[7]);
/* No counterpart in the reference (tests). The split tree (BSP) of n
 * triangles (9 floats each; lighter_math.cpp:428-488) built with the pruned
 * farthest-pair search and with the reference's full pair scan: out[0] 1
 * when the two trees are identical, out[1] and out[2] their build ms. */
LTRAPI int lgx_SplitTreeCompare(const float* tris9, size_t n, double out[3]);
/* No counterpart in the reference. Last GPU export: out[0] pack ms, out[1]
 * upload ms, out[2] kernels ms (HIP events), out[3] download and unpack ms,
 * out[4] atlas texels (zeros after a host export). */
LTRAPI int lgx_FinalizeTimes(ltr_Scene* scene, double out[5]);
/* "rendering lightmaps" (lighter.cpp:1075-1096) on the GPU. */
LTRAPI int lgx_StageDirect(ltr_Scene* scene);
/* "calculating radiosity" .. "committing radiosity" (lighter.cpp:1098-1124). */
LTRAPI int lgx_StageRadiosity(ltr_Scene* scene);
/* "rendering ambient occlusion" (lighter.cpp:1126-1136). */
LTRAPI int lgx_StageAO(ltr_Scene* scene);
/* "exporting lightmaps" (lighter.cpp:1138-1140, 849-1044): fills the work
 * outputs; the images (extrapolation, blur, downsample) and the normal-map
 * data on the GPU unless lgx_SetHostStages selected the host.
 * Each export replaces the output list; the images of an earlier export
 * stay allocated until ltr_DestroyScene. */
LTRAPI int lgx_StageFinalize(ltr_Scene* scene);

LTRAPI int lgx_InstanceCount(ltr_Scene* scene);
/* out: lightmap width, height, sample count, collision triangle count. */
LTRAPI int lgx_InstanceInfo(ltr_Scene* scene, int inst, int32_t out[4]);
/* Sample arrays of one instance (lighter_int.hpp:897-901); NULL skips. */
LTRAPI int lgx_GetSamples(ltr_Scene* scene, int inst, float* pos3, float* nrm3, uint32_t* loc, float* radinfo4);
LTRAPI int lgx_GetLightmap(ltr_Scene* scene, int inst, float* rgb3);
LTRAPI int lgx_SetLightmap(ltr_Scene* scene, int inst, const float* rgb3);
/* Collision triangles of one instance, 9 floats each. */
LTRAPI int lgx_GetTriangles(ltr_Scene* scene, int inst, float* tris9);
/* Instances light `light` renders into (lighter.cpp:1080-1093). */
LTRAPI int lgx_LightInstances(ltr_Scene* scene, int light, int32_t* out, int cap);
/* Radiosity links after lgx_StageRadiosity: rowmap {offset, count} per
 * patch, then (other, factor) per link, as Int_RDGenLinks lists them. */
LTRAPI size_t lgx_RadLinkCount(ltr_Scene* scene);
LTRAPI int lgx_RadGetLinks(ltr_Scene* scene, uint32_t* rowmap, uint32_t* other, float* factor);
/* Stage times in ms: direct, direct march, AO, links, bounces. */
LTRAPI int lgx_StageTimes(ltr_Scene* scene, double out[5]);

/* ---- device-level hot path over raw arrays -------------------------- */
typedef struct lgx_Device lgx_Device;

/* Light parameters as the reference stores them after ltr_LightAdd
 * (lighter.cpp:1282-1298): dir normalised by the caller. */
typedef struct lgx_Light {
	int32_t type;
	float pos[3];
	float dir[3];
	float color[3];
	float range;
	float power;
	float radius;
	float spot_angle_out; /* degrees */
	float spot_angle_in;  /* degrees */
	float spot_curve;
} lgx_Light;

/* Uploads shadow-casting triangles (9 floats each) and builds the BVH. */
LTRAPI lgx_Device* lgx_DeviceCreate(const float* tris9, size_t ntri);
LTRAPI void lgx_DeviceDestroy(lgx_Device* dev);
/* out: triangles, BVH nodes, BVH depth. */
LTRAPI int lgx_DeviceInfo(lgx_Device* dev, int64_t out[3]);
/* The entry grid (no counterpart in the reference, whose queries start at
 * the root of each BBT, lighter_math.cpp:653-781): a uniform grid over the
 * scene whose cells each hold a node listing at most four subtrees that
 * contain every triangle within 2 m of the cell, and one for 1 m; distance
 * queries (radius <= 2, <= 1) and segments of at most 2 m start at their
 * cell's node.  Results
 * are the same with it on (the default) or off (on = 0: every query starts
 * at the root).  cells (may be null) receives the grid's cell count, 0
 * when the scene has none (no triangles, or a non-finite scene box). */
LTRAPI int lgx_SetEntryGrid(lgx_Device* dev, int on, int64_t* cells);

/* Scene queries, one result per query:
 *  distance: min(2, nearest shadow-triangle distance)   (lighter.cpp:183-188)
 *  shadow:   CalcInvShadowFactor(from, to, k) + step count (lighter.cpp:190-207)
 *  closest:  DistanceTestBBT(a, b)                       (lighter.cpp:280-289)
 *  occluded: VisibilityTest(a, b)                        (lighter.cpp:138-147) */
LTRAPI int lgx_QueryDistance(lgx_Device* dev, const float* p3, size_t n, float* out);
LTRAPI int lgx_QueryShadow(lgx_Device* dev, const float* from3, const float* to3, const float* k, size_t n,
                           float* out, uint32_t* steps);
LTRAPI int lgx_QueryClosest(lgx_Device* dev, const float* a3, const float* b3, size_t n, float* out);
LTRAPI int lgx_QueryOccluded(lgx_Device* dev, const float* a3, const float* b3, size_t n, uint8_t* out);

/* Samples and the lightmap they accumulate into (lightmap NULL = zeros). */
LTRAPI int lgx_SetSamples(lgx_Device* dev, const float* pos3, const float* nrm3, size_t n, const float* lightmap3);
LTRAPI int lgx_ReadLightmap(lgx_Device* dev, float* lightmap3);
/* Direct light (Job_LMRender_*_Inner, lighter.cpp:485-600) for work items
 * (light << 26) | sample grouped by light; light_off has nlights+1 entries.
 * At most 64 lights per call; contributions are added in light order. */
LTRAPI int lgx_DirectLight(lgx_Device* dev, const lgx_Light* lights, int nlights, const uint32_t* items,
                           size_t nitems, const uint32_t* light_off, int keep_vis);
/* The same in two steps: upload the work items once, then run them (each run
 * adds the lights' contributions to the lightmap again). */
LTRAPI int lgx_DirectLightSetup(lgx_Device* dev, const lgx_Light* lights, int nlights, const uint32_t* items,
                                size_t nitems, const uint32_t* light_off, int keep_vis);
LTRAPI int lgx_DirectLightRun(lgx_Device* dev);
/* No counterpart in the reference (its AABB tree builds on the host,
 * lighter_math.cpp:653-801; the device scene's BVH, bvh.h, is this
 * library's).  Builds the device scene's BVH of a triangle soup (9 floats
 * each) twice, with the host build and with the GPU build that
 * lgx_DeviceCreate uses: out[0] = 1 when the 4-wide node arrays are
 * identical byte for byte, out[1] = 1 when the leaf orders and depths are,
 * out[2] host build ms, out[3] GPU build ms (wall, collapse included),
 * out[4] GPU kernels and level read-backs ms (HIP events), out[5] nodes. */
LTRAPI int lgx_BvhCompare(const float* tris, size_t ntri, double out[6]);
/* Abort of the direct-light march from inside the kernel (ltr_Abort, which
 * is a no-op in the reference, lighter.cpp:1154-1157, stops a bake's march
 * the same way through the scene's worker).  While on is set, every
 * lgx_DirectLight / lgx_DirectLightRun on this device stops in its march
 * kernel at the first claim of work, leaves the lightmap as it was and
 * returns 0 with the error "direct light aborted"; on = 0 clears it. */
LTRAPI int lgx_DeviceAbort(lgx_Device* dev, int on);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The pass of lgx_DirectLight (Job_LMRender_*_Inner,
 * lighter.cpp:485-600) with no device allocation: the work items (d_items,
 * (light << 26) | sample grouped by light), the per-item weights
 * (d_weights) and shadow factors (d_vis, or NULL) are device buffers of the
 * caller on the device's GPU; lights and light_off (nlights + 1 entries) are
 * host arrays.  The lights' contributions are added to the device's
 * lightmap (lgx_SetSamples) in light order, as lgx_DirectLight adds them,
 * and the weights and factors are those lgx_ReadWeights would return.  The
 * kernels' scratch (per-item shading factors, the march list, counters) is
 * a file-scope device array for 4,194,304 work items (32 MiB + 4 KiB), one
 * per GPU for the life of the process: more items are refused, and calls
 * that overlap on one GPU race on it.  Synchronous. */
LTRAPI int lgx_DirectLightNoAlloc(lgx_Device* dev, const lgx_Light* lights, int nlights, const uint32_t* d_items,
                                  size_t nitems, const uint32_t* light_off, float* d_weights, float* d_vis);
/* Per item: weight multiplying the light colour (bits 0xFFFFFFFF where the
 * reference skips the texel) and, if kept, the shadow factor. */
LTRAPI int lgx_ReadWeights(lgx_Device* dev, float* weights, float* vis, size_t n);
/* Ambient occlusion (Job_AORender_Inner, lighter.cpp:799-841) with one
 * rotation offset per sample. */
LTRAPI int lgx_AmbientOcclusion(lgx_Device* dev, const float* randoff, int num_samples, float distance,
                                float multiplier, float falloff, float effect, const float* color3);
/* Radiosity links among N patches (Int_RDGenLinks pair loop,
 * lighter.cpp:728-759); returns the link count through nlinks.  The pairs
 * are found on a uniform grid of cells as wide as the longest distance a
 * pair can link over (f >= 0.001 needs l^2 <= |nA| |nB| / (0.001 pi)), so
 * any patch count whose links fit 2^32 CSR entries is taken; with
 * non-finite positions or normals the search falls back to all pairs. */
LTRAPI int lgx_RadLinks(lgx_Device* dev, const float* pos3, const float* nrm3, size_t n, size_t* nlinks);
/* No counterpart in the reference (tests). How the next lgx_RadLinks
 * finds its candidate pairs: mode 0 (default) the grid, unless it would
 * test more than N^2 / 4 pairs (a scene within one cell width), then all
 * pairs; 1 all pairs, as the reference's loop does (at most 524,288
 * patches); 2 the grid.  Every mode gives the same links. */
LTRAPI int lgx_SetLinkSearch(lgx_Device* dev, int mode);
LTRAPI int lgx_RadReadLinks(lgx_Device* dev, uint32_t* rowmap, uint32_t* other, float* factor);
/* bounce_count rounds of Int_RDBounce (lighter.cpp:767-797) from emitted
 * light E; total3 receives the accumulated light per patch. */
LTRAPI int lgx_RadBounce(lgx_Device* dev, const float* diffuse3, const float* emit3, const float* area, int bounces,
                         float* total3);
/* Radiosity for R right-hand sides (1..192) over the links of the last
 * lgx_RadLinks: total = sum_{k=0..bounces} K^k E with
 * K = F diag(area) diag(diffuse_{c % 3}) for column c, the closed form of
 * bounce_count rounds of Int_RDBounce (lighter.cpp:767-797, 1104-1123).
 * E and total are N x R row-major FP64.  form: 0 sparse FP64 gather,
 * 1 sparse FP32 gather, 2 dense FP32, 3 BF16 on MFMA and 4 FP16 on MFMA over
 * the 256 x 32 tiles of the matrix that hold a link (patches in Morton
 * order), 5 FP32 on the FP32 matrix cores over the same tiles, 6 FP16 and
 * 7 BF16 with the error-compensated product (hi + lo planes of matrix and
 * operand, three MFMA products) on every bounce; every form accumulates in
 * FP32 or FP64.  The 16-bit forms (3, 4, 6, 7) keep the largest factors out
 * of the tiles (the top binades of the factors' exponents, at most 12 below
 * the largest and at most max(nnz / 4096, 256) factors: the near-coincident
 * pairs where two surfaces meet) and apply them in FP32 to the FP32
 * operand.  refine: refinement sweeps of forms 3 and 4 with an FP64
 * residual from the error-compensated product (an error for other forms).
 * times_ms: matrix build, series, per bounce, refinement. */
LTRAPI int lgx_RadSolve(lgx_Device* dev, const double* E, uint32_t R, const float* diffuse3, const float* area,
                        int form, int bounces, int refine, double* total, double times_ms[4]);
/* Reverse mode of lgx_RadSolve (no counterpart in the reference, which has
 * no derivatives), in any of its forms (forward and reverse products both
 * the form's own): for the upstream gradient g = dL/dtotal (N x R FP64),
 * total again, lam = dL/dE (which is also dL/d emission) and
 * dalb = dL/d diffuse per column (column c holds the derivative through the
 * albedo of channel c % 3; sum the columns of one channel).  The adjoint
 * sweep applies K^T = diag(diffuse) diag(area) F with the same kernels.
 * times_ms: dense build, forward series, reverse per bounce, reverse sweep. */
LTRAPI int lgx_RadSolveGrad(lgx_Device* dev, const double* E, uint32_t R, const float* diffuse3,
                            const float* area, int form, int bounces, const double* g, double* total, double* lam,
                            double* dalb, double times_ms[4]);
/* Tile store of the 16-bit forms after an lgx_RadSolve / lgx_RadSolveGrad in
 * form 3 or 4 (no counterpart in the reference): out[0] = stored 256 x 32
 * tiles (those holding a link, patches in Morton order), out[1] = all tiles
 * of the padded matrix.  Zeros before such a solve. */
LTRAPI int lgx_RadTiles(lgx_Device* dev, uint64_t out[2]);
/* Reverse mode of the last lgx_DirectLight / lgx_DirectLightRun (no
 * counterpart in the reference): the lightmap gains colour * weight per work
 * item (Job_LMRender_*_Inner, lighter.cpp:485-589), so for lam3 = dL/dlightmap
 * (3 FP64 per device sample) dcolor3[3 l + c] = sum over light l's items of
 * weight * lam3[3 s + c]. */
LTRAPI int lgx_DirectLightGrad(lgx_Device* dev, const double* lam3, double* dcolor3);
/* Traversal counters (node visits, full triangle tests, march steps,
 * queries, marches, triangles culled by the distance query's box pre-test)
 * collected by the next stage when enabled; with counting on, lgx_RadLinks
 * also keeps its candidate pairs for lgx_ReferenceCountsLinks. */
LTRAPI int lgx_SetCounting(lgx_Device* dev, int on);
LTRAPI int lgx_ReadCounters(lgx_Device* dev, uint64_t out[6]);
/* No counterpart in the reference (measurement only). The last
 * lgx_DirectLightRun's marches (run with kept weights) again, with the
 * reference's query semantics (lighter.cpp:150-207: a distance query of
 * radius 2 per step) on a plain binary SAH tree of 16 bins and at most 4
 * triangles per leaf, counting out[0] node visits (boxes tested), out[1]
 * triangle tests, out[2] distance queries, out[3] shadow factors that
 * differ from the product's bits, out[4] marches. tris: the device's
 * triangles, as given to lgx_DeviceCreate. */
LTRAPI int lgx_ReferenceCounts(lgx_Device* dev, const float* tris, size_t ntri, uint64_t out[5]);
/* No counterpart in the reference (measurement only). The AO factors of
 * every sample (lighter.cpp:812-831; the lightmap is left as it was) with
 * the rays traced with the canonical closest-hit semantics (nearest child
 * first, the segment's end shrinking to the closest hit) on the same plain
 * binary SAH tree: out[0] node visits, out[1] triangle tests, out[2] rays,
 * out[3] samples whose factor differs from the product's bits, out[4]
 * samples. */
LTRAPI int lgx_ReferenceCountsAO(lgx_Device* dev, const float* tris, size_t ntri, const float* randoff,
                                 int num_samples, float distance, float multiplier, float falloff, uint64_t out[5]);
/* No counterpart in the reference (measurement only). The occlusion test of
 * every candidate pair of the last lgx_RadLinks (run with lgx_SetCounting
 * on, which keeps the pairs) again with the canonical any-hit walk (nearest
 * child first, ending at the first hit with r < 1) on the same plain binary
 * SAH tree: out[0] node visits, out[1] triangle tests, out[2] pairs tested,
 * out[3] pairs whose link differs from the product's, out[4] links. */
LTRAPI int lgx_ReferenceCountsLinks(lgx_Device* dev, const float* tris, size_t ntri, uint64_t out[5]);
/* Last stage times in ms: direct, direct march, AO, links, bounces,
 * direct-light gradient. */
LTRAPI int lgx_DeviceTimes(lgx_Device* dev, double out[6]);

#ifdef __cplusplus
}
#endif
