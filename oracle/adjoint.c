/*
 * Oracle: FP64 re-weighting of recorded paths and the FP64 path-replay
 * adjoint (gradients of an image loss with respect to the texels of one
 * base-color texture and to per-instance emission strengths).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Method: path replay backpropagation (Vicini, Speierer, Jakob, "Path
 * Replay Backpropagation: Differentiating Light Paths using Constant Memory
 * and Linear Time", SIGGRAPH 2021).  Every contribution of a path is linear
 * in the albedo of each diffuse vertex it passes and in the emission
 * strength it ends on, so with L_rem = the part of the path's estimate
 * collected at or after vertex d (its NEE term included),
 *   d estimate / d albedo_d = L_rem / albedo_d   (per channel),
 * and the estimate is replayed front to back, subtracting each collected
 * contribution, so only the path's total is stored.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define INV_PI_F 0.31830988f

static void texel_rgb(const orc_shading* sh, int32_t vtex, uint32_t texel, int32_t tex, const double* texels, double out[3])
{
    if (vtex == tex && texels) {
        for (int k = 0; k < 3; ++k) out[k] = texels[4 * (size_t)texel + k];
    } else {
        const float* t = sh->tex_rgba + sh->tex_off[vtex];
        for (int k = 0; k < 3; ++k) out[k] = (double)t[4 * (size_t)texel + k];
    }
}

/* albedo of a diffuse vertex: base_color x bilinear texture, in FP64 */
static void albedo_f64(const orc_shading* sh, const orc_vertex* v, int32_t tex, const double* texels, double a[3])
{
    const float* bc = sh->mi[v->inst].base_color;
    if (v->tex < 0) {
        for (int k = 0; k < 3; ++k) a[k] = (double)bc[k];
        return;
    }
    double c[4][3];
    for (int j = 0; j < 4; ++j) texel_rgb(sh, v->tex, v->texel[j], tex, texels, c[j]);
    double fx = v->fx, fy = v->fy;
    for (int k = 0; k < 3; ++k) {
        double top = c[0][k] + fx * (c[1][k] - c[0][k]);
        double bot = c[2][k] + fx * (c[3][k] - c[2][k]);
        a[k] = (double)bc[k] * (top + fy * (bot - top));
    }
}

static void emission_f64(const double* factor3, const double* strength, uint32_t inst, double le[3])
{
    for (int k = 0; k < 3; ++k) le[k] = factor3[3 * (size_t)inst + k] * strength[inst];
}

/* the FP64 estimate of one path */
static void eval_path(const orc_shading* sh, const orc_render_params* p, const orc_vertex* rec, int32_t tex,
                      const double* texels, const double* factor3, const double* strength, double e[3])
{
    double thr[3] = { 1.0, 1.0, 1.0 };
    e[0] = e[1] = e[2] = 0.0;
    const double inv_pi = (double)INV_PI_F;
    for (uint32_t d = 0; d <= p->max_bounces; ++d) {
        const orc_vertex* v = &rec[d];
        if (v->kind == ORC_V_END) break;
        if (v->kind == ORC_V_MISS) {
            if (sh->env_tex >= 0) {
                const float* c = sh->tex_rgba + sh->tex_off[sh->env_tex] + 4 * (size_t)v->texel[0];
                for (int k = 0; k < 3; ++k) e[k] += thr[k] * (double)sh->env_strength * (double)c[k];
            } else {
                for (int k = 0; k < 3; ++k) e[k] += thr[k] * (double)sh->env[k];
            }
            break;
        }
        if (v->kind == ORC_V_EMIT) {
            if (v->emit_counted) {
                double le[3];
                emission_f64(factor3, strength, v->inst, le);
                for (int k = 0; k < 3; ++k) e[k] += thr[k] * le[k] * (double)v->em_tex[k];
            }
            break;
        }
        double a[3];
        albedo_f64(sh, v, tex, texels, a);
        if (v->nee) {
            double le[3];
            emission_f64(factor3, strength, v->light_inst, le);
            for (int k = 0; k < 3; ++k) e[k] += ((thr[k] * a[k]) * inv_pi) * le[k] * (double)v->nee_tex[k] * (double)v->nee_wgt;
        }
        if (d + 1 > p->max_bounces || rec[d + 1].kind == ORC_V_END) break;
        for (int k = 0; k < 3; ++k) {
            thr[k] *= ((double)v->noL * (a[k] * inv_pi)) * (double)v->inv_pdf;
            if (v->rr) thr[k] /= (double)p->rr_survival;
        }
    }
}

/* Only the albedo use of `tex` is a variable here: a texture that also gives
 * emission or the environment is refused (its other terms would be held
 * at the recorded values). */
static int tex_is_albedo_only(const orc_shading* sh, int32_t tex)
{
    if (tex < 0) return 1;
    if (sh->env_tex == tex) return 0;
    for (uint32_t i = 0; i < sh->n_inst; ++i)
        if (sh->mi[i].emissive_tex == tex) return 0;
    return 1;
}

int orc_eval_f64(const orc_shading* sh, const orc_render_params* p, uint32_t width, uint32_t height,
                 const orc_vertex* rec, int32_t tex, const double* texels, const double* factor3,
                 const double* strength, double* image)
{
    if (!sh || !p || !rec || !factor3 || !strength || !image) return ORC_EINVAL;
    if (!tex_is_albedo_only(sh, tex)) return ORC_EINVAL;
    const uint32_t npix = width * height, stride = p->max_bounces + 1;
    memset(image, 0, sizeof(double) * 3 * (size_t)npix);
    for (uint32_t s = 0; s < p->spp; ++s)
        for (uint32_t px = 0; px < npix; ++px) {
            double e[3];
            eval_path(sh, p, rec + ((size_t)s * npix + px) * stride, tex, texels, factor3, strength, e);
            for (int k = 0; k < 3; ++k) image[3 * (size_t)px + k] += e[k];
        }
    for (size_t i = 0; i < 3 * (size_t)npix; ++i) image[i] /= (double)p->spp;
    return ORC_OK;
}

int orc_adjoint_f64(const orc_shading* sh, const orc_render_params* p, uint32_t width, uint32_t height,
                    const orc_vertex* rec, int32_t tex, const double* texels, const double* factor3,
                    const double* strength, const double* dl_dimage, double* d_texels, double* d_strength)
{
    if (!sh || !p || !rec || !factor3 || !strength || !dl_dimage || !d_strength) return ORC_EINVAL;
    if (!tex_is_albedo_only(sh, tex)) return ORC_EINVAL;
    const uint32_t npix = width * height, stride = p->max_bounces + 1;
    const double inv_pi = (double)INV_PI_F;
    for (uint32_t s = 0; s < p->spp; ++s)
        for (uint32_t px = 0; px < npix; ++px) {
            const orc_vertex* r = rec + ((size_t)s * npix + px) * stride;
            double lrem[3], thr[3] = { 1.0, 1.0, 1.0 }, delta[3];
            eval_path(sh, p, r, tex, texels, factor3, strength, lrem);
            for (int k = 0; k < 3; ++k) delta[k] = dl_dimage[3 * (size_t)px + k] / (double)p->spp;
            for (uint32_t d = 0; d <= p->max_bounces; ++d) {
                const orc_vertex* v = &r[d];
                if (v->kind == ORC_V_END || v->kind == ORC_V_MISS) break;
                if (v->kind == ORC_V_EMIT) {
                    if (v->emit_counted) {
                        double g = 0.0;
                        for (int k = 0; k < 3; ++k) g += delta[k] * thr[k] * factor3[3 * (size_t)v->inst + k] * (double)v->em_tex[k];
                        d_strength[v->inst] += g;
                    }
                    break;
                }
                double a[3];
                albedo_f64(sh, v, tex, texels, a);
                if (v->tex == tex && tex >= 0 && d_texels) {
                    const float* bc = sh->mi[v->inst].base_color;
                    const double fx = v->fx, fy = v->fy;
                    const double w[4] = { (1.0 - fx) * (1.0 - fy), fx * (1.0 - fy), (1.0 - fx) * fy, fx * fy };
                    for (int k = 0; k < 3; ++k) {
                        if (!(a[k] > 0.0)) continue;
                        const double g = delta[k] * lrem[k] / a[k] * (double)bc[k];
                        for (int j = 0; j < 4; ++j) d_texels[4 * (size_t)v->texel[j] + k] += g * w[j];
                    }
                }
                if (v->nee) {
                    double le[3], g = 0.0;
                    emission_f64(factor3, strength, v->light_inst, le);
                    for (int k = 0; k < 3; ++k) {
                        const double unit = ((thr[k] * a[k]) * inv_pi) * (double)v->nee_wgt * (double)v->nee_tex[k];
                        lrem[k] -= unit * le[k];
                        g += delta[k] * unit * factor3[3 * (size_t)v->light_inst + k];
                    }
                    d_strength[v->light_inst] += g;
                }
                if (d + 1 > p->max_bounces || r[d + 1].kind == ORC_V_END) break;
                for (int k = 0; k < 3; ++k) {
                    thr[k] *= ((double)v->noL * (a[k] * inv_pi)) * (double)v->inv_pdf;
                    if (v->rr) thr[k] /= (double)p->rr_survival;
                }
            }
        }
    return ORC_OK;
}
