/*
 * Oracle: wavefront render loop.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

int orc_render(const orc_scene* s, const orc_shading* sh, const orc_view* v, uint32_t width, uint32_t height,
               const orc_render_params* p, float* energy_rgb, int nthreads)
{
    (void)s; (void)sh; (void)v; (void)width; (void)height; (void)p; (void)energy_rgb; (void)nthreads;
    return ORC_EINVAL;
}
