// rd_envmap.cpp — environment-light sampling tables, built on the host at
// scene creation and read by the device's env_sample / env_pdf_dir.
#include "rd_envmap.h"

#include <cmath>

namespace rd {
namespace {

// TextureSampler.Sample (TextureSampler.cs:28-72) on a host texture
// description: Clamp / Repeat wrap, Nearest / Linear filter; images are
// column-major (x*h + y)*3 like ColorBuffer.  Same float operations as the
// device's tex_sample.
f3 host_tex_sample(const rd_texture_desc& t, float su, float sv) {
    if (t.width == 0) return mk(t.value[0], t.value[1], t.value[2]);
    auto read = [&](int x, int y) {
        const float* p = t.data + ((size_t)x * t.height + y) * 3;
        return mk(p[0], p[1], p[2]);
    };
    float u, v;
    if (t.wrap == 0) { u = nclamp(su, 0.0f, 1.0f); v = nclamp(sv, 0.0f, 1.0f); }
    else { u = nclamp(su - std::floor(su), 0.0f, 1.0f); v = nclamp(sv - std::floor(sv), 0.0f, 1.0f); }
    float fu = u * (t.width - 1), fv = v * (t.height - 1);
    int pu = (int)fu, pv = (int)fv;
    if (t.filter == 0) return read(pu, pv);
    int dpu = (fu > pu + 0.5f) ? 1 : -1;
    int dpv = (fv > pv + 0.5f) ? 1 : -1;
    int apu = iclamp(pu + dpu, 0, t.width - 1);
    int apv = iclamp(pv + dpv, 0, t.height - 1);
    float du = nmin(std::fabs(fu - pu - 0.5f), 1.0f);
    float dv = nmin(std::fabs(fv - pv - 0.5f), 1.0f);
    f3 a = read(pu, pv), b = read(apu, pv), c = read(pu, apv), e = read(apu, apv);
    return (a * (1 - du) + b * du) * (1 - dv) + (c * (1 - du) + e * du) * dv;
}

// Distribution1D ctor (Distribution1D.cs:19-50): cdf[i] = cdf[i-1] + f[i-1]/n,
// normalised by the integral (uniform when it is zero).  Returns the integral.
float build_1d(const float* f, int n, float* cdf) {
    cdf[0] = 0;
    for (int i = 1; i < n + 1; ++i) cdf[i] = cdf[i - 1] + f[i - 1] / n;
    float integ = cdf[n];
    if (integ == 0) {
        for (int i = 1; i < n + 1; i++) cdf[i] = i / (float)n;
    } else {
        for (int i = 1; i < n + 1; i++) cdf[i] /= integ;
    }
    return integ;
}

}  // namespace

void build_env_tables(const rd_texture_desc& t, EnvTables* out) {
    const int w = t.width > 0 ? t.width : 1;   // ConstColorTexture2D is 1x1
    const int h = t.width > 0 ? t.height : 1;
    out->w = w;
    out->h = h;
    out->cond_func.assign((size_t)w * h, 0.0f);
    out->cond_cdf.assign((size_t)(w + 1) * h, 0.0f);
    out->marg_func.assign(h, 0.0f);
    out->marg_cdf.assign(h + 1, 0.0f);
    for (int y = 0; y < h; y++) {
        float sin_theta = std::sin((y + 0.5f) / h * kPi);
        float v = y / (float)h;
        for (int x = 0; x < w; x++) {
            float u = x / (float)w;
            f3 c = host_tex_sample(t, u, v);
            // Color3F.GetLuminance (Color3F.cs:49)
            out->cond_func[(size_t)y * w + x] = (c.x * 0.212671f + c.y * 0.715160f + c.z * 0.072169f) * sin_theta;
        }
        out->marg_func[y] = build_1d(&out->cond_func[(size_t)y * w], w, &out->cond_cdf[(size_t)y * (w + 1)]);
    }
    out->marg_int = build_1d(out->marg_func.data(), h, out->marg_cdf.data());
}

}  // namespace rd
