/* Oracle (test infrastructure only): the post-process pass that turns a
 * frame's energy into the HDR accumulator and the RGBA8 final colour.
 *
 * Follows post_process.hlsl:25-72 and linear_to_srgb (common.hlsl:251-259):
 *   - NaN or Inf in any channel of the energy -> magenta (1, 0, 1)  (:34-38)
 *   - accumulate and view mode NONE: accum = accum*(1-w) + (e,1)*w, with
 *     w = 1/sample_count; otherwise accum = (e, 1)                  (:40-51)
 *   - final = accum; Reinhard x/(1+x) for view mode NONE only       (:53-60)
 *   - linear -> sRGB for NONE, MATERIAL_BASE_COLOR (6) and
 *     MATERIAL_EMISSIVE (9) only                                     (:62-69)
 *   - the R8G8B8A8_UNORM store of final (renderer.cpp:492): NaN -> 0,
 *     clamp to [0, 1], round to nearest.
 * The energy passed in is the frame's per-pixel mean over its samples (the
 * reference renders one sample per frame). */
#include "oracle.h"

#include <math.h>

static float clamp01(float x)
{
    /* HLSL clamp = min(max(x, 0), 1) with IEEE maxNum/minNum: NaN -> 0 */
    return fminf(fmaxf(x, 0.0f), 1.0f);
}

static float linear_to_srgb(float c)
{
    const float x = clamp01(c);
    /* step(x, 0.0031308) is 1 when 0.0031308 >= x: the linear segment */
    return (x <= 0.0031308f) ? x * 12.92f : 1.055f * powf(x, 1.0f / 2.4f) - 0.055f;
}

static uint32_t unorm8(float x)
{
    return (uint32_t)fminf(fmaxf(x * 255.0f + 0.5f, 0.0f), 255.0f);
}

void orc_post_process(const float* energy_rgba, uint32_t n, uint32_t sample_count, uint32_t accumulate,
                      uint32_t view_mode, float* accum_rgba, uint32_t* final_rgba8, float* final_rgba)
{
    const int tracer = view_mode == 0;
    const int srgb = tracer || view_mode == 6 || view_mode == 9;
    const float w = 1.0f / (float)(sample_count ? sample_count : 1);
    for (uint32_t i = 0; i < n; ++i) {
        float e[3] = {energy_rgba[4 * i + 0], energy_rgba[4 * i + 1], energy_rgba[4 * i + 2]};
        if (isnan(e[0]) || isnan(e[1]) || isnan(e[2]) || isinf(e[0]) || isinf(e[1]) || isinf(e[2])) {
            e[0] = 1.0f; e[1] = 0.0f; e[2] = 1.0f;
        }
        float* acc = accum_rgba + 4 * i;
        if (accumulate && tracer) {
            for (int k = 0; k < 3; ++k) acc[k] = acc[k] * (1.0f - w) + e[k] * w;
            acc[3] = acc[3] * (1.0f - w) + 1.0f * w;
        } else {
            for (int k = 0; k < 3; ++k) acc[k] = e[k];
            acc[3] = 1.0f;
        }
        float f[4] = {acc[0], acc[1], acc[2], acc[3]};
        for (int k = 0; k < 3; ++k) {
            if (tracer) f[k] = f[k] / (1.0f + f[k]);
            if (srgb) f[k] = linear_to_srgb(f[k]);
        }
        if (final_rgba)
            for (int k = 0; k < 4; ++k) final_rgba[4 * i + k] = f[k];
        final_rgba8[i] = unorm8(f[0]) | unorm8(f[1]) << 8 | unorm8(f[2]) << 16 | unorm8(f[3]) << 24;
    }
}
