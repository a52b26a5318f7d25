// rd_envmap.h — host preprocessing of environment lights (product side).
#pragma once
#include <vector>
#include "rd_types.h"
#include "../../include/radiantium_hip.h"

namespace rd {

// The sampling tables of one InfiniteAreaLight (Lights/InfiniteAreaLight.cs:21-54):
// luminance x sin(theta) of the radiance map sampled at texel corners, as the
// rows of a Distribution2D (Distribution2D.cs:14-31) whose rows and marginal
// are Distribution1D (Distribution1D.cs:19-50).
struct EnvTables {
    int w = 1, h = 1;
    std::vector<float> cond_func;  // [h][w]
    std::vector<float> cond_cdf;   // [h][w+1]
    std::vector<float> marg_func;  // [h]  (row integrals)
    std::vector<float> marg_cdf;   // [h+1]
    float marg_int = 0.0f;
};

void build_env_tables(const rd_texture_desc& tex, EnvTables* out);

}  // namespace rd
