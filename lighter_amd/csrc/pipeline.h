// Bake pipeline entry points shared by the ltr_* API (api.cpp) and the
// staged lgx_* entry points.
#pragma once

#include <stdexcept>
#include <vector>

#include "scene.h"

namespace lgx {

// Thrown between stages when ltr_Abort was called.
struct AbortBake {};

void bake(ltr_Scene* S);                  // the whole pipeline (ltr_Start's thread)
void reset_bake(ltr_Scene* S);            // forget the last bake's samples, lightmaps and device state
void prepare(ltr_Scene* S);               // host stages up to sample generation
void ensure_device(ltr_Scene* S);         // upload triangles, samples, lightmaps
void download_lightmaps(ltr_Scene* S);
void upload_lightmaps(ltr_Scene* S);
void stage_direct(ltr_Scene* S);          // "rendering lightmaps"
void stage_radiosity(ltr_Scene* S);       // "calculating radiosity" .. "committing radiosity"
void stage_ao(ltr_Scene* S);              // "rendering ambient occlusion"
void finalize(ltr_Scene* S);              // "exporting lightmaps"
std::vector<int32_t> light_instances(ltr_Scene* S, const Light& L);
DevLight to_dev_light(const Light& L);

} // namespace lgx
