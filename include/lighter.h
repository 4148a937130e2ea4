/*
 * lighter.h — public C API of the lightmap baker, MI355X build.
 *
 * This header declares the same entry points, structures and constants as the
 * reference baker's public header (archo5/lighter, lighter.h:1-264), with the
 * same memory layout, so that applications written against the reference
 * link against liblighter_amd.so unchanged.  The layout is fixed by the ABI;
 * the sizes and offsets below are asserted in lighter_amd/csrc/api.cpp.
 *
 *   struct               size   (x86-64)
 *   ltr_MeshPartInfo       72
 *   ltr_MeshInstanceInfo   88
 *   ltr_LightInfo          80
 *   ltr_SampleInfo         40
 *   ltr_SampleRequest     104
 *   ltr_Config            144
 *   ltr_WorkOutputInfo     16
 *   ltr_WorkOutput         64
 *   ltr_WorkStatus         16
 *
 * As in the reference, ltr_Config refers to itself inside its own callback
 * types, so the header is consumed as C++ (the linkage is C).
 */
#pragma once

#include <stdlib.h>
#include <inttypes.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LTRAPI __attribute__((visibility("default")))

#define LTRBOOL int
#define LTRCODE int

typedef int32_t i32;
typedef uint32_t u32;

typedef float ltr_VEC3[3];
typedef float ltr_VEC4[4];
typedef ltr_VEC4 ltr_MAT4[4];

#define LTR_VEC3_SET(v, x, y, z) do { (v)[0] = (x); (v)[1] = (y); (v)[2] = (z); } while (0)

#define LTRC_SUCCESS 0
#define LTRC_ATEND 1

#define LTR_LT_POINT 1
#define LTR_LT_SPOT 2
#define LTR_LT_DIRECT 3

/* One part of a mesh: strided vertex streams plus a triangle list.  Copied at
 * ltr_MeshAddPart time.  Parts with shadow == 0 receive light but cast no
 * shadow. */
typedef struct ltr_MeshPartInfo {
	const void* positions_f3;
	const void* normals_f3;
	const void* texcoords1_f2;
	const void* texcoords2_f2; /* lightmap UVs */
	u32 stride_positions;
	u32 stride_normals;
	u32 stride_texcoords1;
	u32 stride_texcoords2;
	const u32* indices;
	u32 vertex_count;
	u32 index_count;
	int shadow;
} ltr_MeshPartInfo;

/* A placement of a mesh; matrix uses row vectors (translation in row 3). */
typedef struct ltr_MeshInstanceInfo {
	ltr_MAT4 matrix;
	float importance;
	int shadow;
	const char* ident;
	size_t ident_size;
} ltr_MeshInstanceInfo;

typedef struct ltr_LightInfo {
	u32 type; /* LTR_LT_* */
	ltr_VEC3 position;
	ltr_VEC3 direction; /* spot axis, or direction towards a directional light */
	ltr_VEC3 up_direction;
	ltr_VEC3 color_rgb;
	float range;        /* falloff range; for directional lights the march length */
	float power;        /* distance falloff exponent */
	float light_radius; /* penumbra constant of the soft-shadow march */
	int shadow_sample_count;
	float spot_angle_out; /* degrees */
	float spot_angle_in;  /* degrees */
	float spot_curve;
} ltr_LightInfo;

/* An externally supplied sample point; out_color is filled by the bake. */
typedef struct ltr_SampleInfo {
	u32 id;
	ltr_VEC3 position;
	ltr_VEC3 normal;
	ltr_VEC3 out_color;
} ltr_SampleInfo;

/* Per-patch material query issued during radiosity set-up. */
typedef struct ltr_SampleRequest {
	ltr_VEC3 position;
	ltr_VEC3 normal;
	float tex0u, tex0v;
	float tex1u, tex1v;
	uint32_t part_id;

	const char* mesh_ident;
	size_t mesh_ident_size;
	const char* inst_ident;
	size_t inst_ident_size;

	ltr_VEC3 out_diffuse_color;
	ltr_VEC3 out_emissive_color;
} ltr_SampleRequest;

typedef struct ltr_Config {
	void* userdata;
	LTRBOOL (*size_fn)(ltr_Config* config, const char* mesh_ident, size_t mesh_ident_size,
	                   const char* inst_ident, size_t inst_ident_size, float computed_surface_area,
	                   float inst_importance, u32 out_size[2]);
	int max_num_threads;
	size_t max_tree_memory;
	u32 max_lightmap_size;
	u32 default_width;
	u32 default_height;
	float global_size_factor;
	float max_correct_dist;
	float max_correct_angle;
	ltr_VEC3 clear_color;
	ltr_VEC3 ambient_color;
	/* radiosity */
	int bounce_count;
	LTRBOOL (*sample_fn)(ltr_Config* config, ltr_SampleRequest* req);
	/* ambient occlusion: per texel, ao = sum over rays of (1 - hit fraction)
	 * averaged and scaled by ao_multiplier (clamped to 1), raised to
	 * ao_falloff; ao_effect >= 0: C * (1 - ao * (1 - ao_effect)) + ao_color * ao,
	 * ao_effect < 0: lerp(C, lerp(ao_color, ao_color * C, -ao_effect), ao) */
	float ao_distance;
	float ao_multiplier;
	float ao_falloff;
	float ao_effect;
	float ao_divergence;
	ltr_VEC3 ao_color_rgb;
	int ao_num_samples;
	/* separable Gaussian blur radius in texels (0 = off) */
	float blur_size;
	/* halve the output resolution */
	int ds2x;
	/* also produce per-texel dominant light direction + focus */
	int generate_normalmap_data;
} ltr_Config;

LTRAPI LTRBOOL ltr_DefaultSizeFunc(ltr_Config* config, const char* mesh_ident, size_t mesh_ident_size,
                                   const char* inst_ident, size_t inst_ident_size, float computed_surface_area,
                                   float inst_importance, u32 out_size[2]);

typedef struct ltr_WorkOutputInfo {
	u32 lightmap_count;
	u32 sample_count;
	ltr_SampleInfo* samples;
} ltr_WorkOutputInfo;

typedef struct ltr_WorkOutput {
	u32 uid;
	const char* mesh_ident;
	size_t mesh_ident_size;
	const char* inst_ident;
	size_t inst_ident_size;
	float* lightmap_rgb; /* width*height*3, row-major, owned by the scene */
	float* normals_xyzf; /* width*height*4 or NULL, owned by the scene */
	u32 width;
	u32 height;
} ltr_WorkOutput;

typedef struct ltr_Mesh ltr_Mesh;
typedef struct ltr_Scene ltr_Scene;

typedef struct ltr_WorkStatus {
	float completion;
	const char* stage;
} ltr_WorkStatus;

/* scene */
LTRAPI ltr_Scene* ltr_CreateScene();
LTRAPI void ltr_DestroyScene(ltr_Scene* scene);
LTRAPI void ltr_Start(ltr_Scene* scene);
LTRAPI void ltr_Abort(ltr_Scene* scene);
LTRAPI LTRBOOL ltr_GetStatus(ltr_Scene* scene, ltr_WorkStatus* wsout);
LTRAPI void ltr_Sleep(int ms);
LTRAPI void ltr_GetConfig(ltr_Config* cfg, ltr_Scene* opt_scene);
LTRAPI LTRCODE ltr_SetConfig(ltr_Scene* scene, ltr_Config* cfg);

/* input */
LTRAPI ltr_Mesh* ltr_CreateMesh(ltr_Scene* scene, const char* ident, size_t ident_size);
LTRAPI LTRBOOL ltr_MeshAddPart(ltr_Mesh* mesh, ltr_MeshPartInfo* mpinfo);
LTRAPI LTRBOOL ltr_MeshAddInstance(ltr_Mesh* mesh, ltr_MeshInstanceInfo* mii);
LTRAPI void ltr_LightAdd(ltr_Scene* scene, ltr_LightInfo* li);
LTRAPI void ltr_SampleAdd(ltr_Scene* scene, ltr_SampleInfo* si);

/* output */
LTRAPI void ltr_GetWorkOutputInfo(ltr_Scene* scene, ltr_WorkOutputInfo* woutinfo);
LTRAPI LTRBOOL ltr_GetWorkOutput(ltr_Scene* scene, u32 which, ltr_WorkOutput* wout);

/* misc */
LTRAPI u32 ltr_NextPowerOfTwo(u32 x);

#ifdef __cplusplus
}
#endif
