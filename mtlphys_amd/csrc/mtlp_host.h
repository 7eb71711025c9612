// mtlp_host.h -- host-side helpers shared by engine.cpp and scenes.cpp.
#pragma once

#include "mtlp.h"

namespace mtlp {

void default_params(mtlp_params* p);
mtlp_status scene_params(uint32_t scene, uint32_t n, mtlp_params* out);
mtlp_status generate_scene(uint32_t scene, uint32_t n, uint64_t seed, const mtlp_params& p,
                           float* x4, float* v4);

}  // namespace mtlp
