// Procedural scene inputs (see include/wf_scene.h).
#include "../../include/wf_scene.h"
#include "host_math.h"

#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

// PCG32, XSH-RR output (O'Neill 2014, pcg32_random_r / pcg32_srandom_r)
constexpr uint64_t kPcgMult = 6364136223846793005ull;

uint32_t pcg_next(wfs_pcg32* r)
{
    const uint64_t old = r->state;
    r->state = old * kPcgMult + r->inc;
    const uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
}

void pcg_seed(wfs_pcg32* r, uint64_t initstate, uint64_t initseq)
{
    r->state = 0u;
    r->inc = (initseq << 1u) | 1u;
    pcg_next(r);
    r->state += initstate;
    pcg_next(r);
}

float pcg_float(wfs_pcg32* r) { return (float)(pcg_next(r) >> 8) * (1.0f / 16777216.0f); }
float uniform(wfs_pcg32* r, float lo, float hi) { return lo + (hi - lo) * pcg_float(r); }

void set_vertex(wf_vertex& v, const float p[3], const float n[3], float u, float w)
{
    std::memset(&v, 0, sizeof(v));
    std::memcpy(v.position, p, 12);
    std::memcpy(v.normal, n, 12);
    v.tangent[0] = 1.0f; v.tangent[3] = 1.0f;
    v.uv[0] = u; v.uv[1] = w;
}

void face_normal(const float a[3], const float b[3], const float c[3], float n[3])
{
    const float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    float x[3];
    wf::cross3(e1, e2, x);
    const float l2 = wf::dot3(x, x);
    if (l2 > 0.0f) wf::normalize3(x, n);
    else { n[0] = 0.0f; n[1] = 1.0f; n[2] = 0.0f; }
}

struct ValueNoise2D {
    uint32_t period;
    std::vector<float> lattice;
    ValueNoise2D(wfs_pcg32* r, uint32_t p) : period(p), lattice((size_t)p * p)
    {
        for (auto& v : lattice) v = pcg_float(r);
    }
    // periodic, smoothstep-bilinear; x, y in lattice units
    float eval(double x, double y) const
    {
        const double fx = std::floor(x), fy = std::floor(y);
        const double tx = x - fx, ty = y - fy;
        const int64_t ix = (int64_t)fx, iy = (int64_t)fy;
        auto at = [&](int64_t i, int64_t j) {
            const uint32_t a = (uint32_t)(((i % period) + period) % period);
            const uint32_t b = (uint32_t)(((j % period) + period) % period);
            return (double)lattice[(size_t)b * period + a];
        };
        const double sx = tx * tx * (3.0 - 2.0 * tx), sy = ty * ty * (3.0 - 2.0 * ty);
        const double v0 = at(ix, iy) + (at(ix + 1, iy) - at(ix, iy)) * sx;
        const double v1 = at(ix, iy + 1) + (at(ix + 1, iy + 1) - at(ix, iy + 1)) * sx;
        return (float)(v0 + (v1 - v0) * sy);
    }
};

} // namespace

extern "C" {

WF_API void wfs_pcg32_seed(wfs_pcg32* rng, uint64_t initstate, uint64_t initseq) { pcg_seed(rng, initstate, initseq); }
WF_API uint32_t wfs_pcg32_next(wfs_pcg32* rng) { return pcg_next(rng); }
WF_API float wfs_pcg32_float(wfs_pcg32* rng) { return pcg_float(rng); }

WF_API wf_status wfs_random_soup(uint64_t seed, uint32_t n, float extent, float off_min, float off_max,
                                 wf_vertex* vertices, uint32_t* indices)
{
    if (!vertices || !indices || n == 0) return WF_ERROR_INVALID_ARGUMENT;
    wfs_pcg32 r;
    pcg_seed(&r, seed, 0x5eed0001ull);
    for (uint32_t t = 0; t < n; ++t) {
        const float c[3] = {uniform(&r, -extent, extent), uniform(&r, -extent, extent), uniform(&r, -extent, extent)};
        float p[3][3];
        for (int v = 0; v < 3; ++v)
            for (int k = 0; k < 3; ++k) {
                const float sign = (pcg_next(&r) & 1u) ? 1.0f : -1.0f;
                p[v][k] = c[k] + sign * uniform(&r, off_min, off_max);
            }
        float nrm[3];
        face_normal(p[0], p[1], p[2], nrm);
        set_vertex(vertices[3 * t + 0], p[0], nrm, 0.0f, 0.0f);
        set_vertex(vertices[3 * t + 1], p[1], nrm, 1.0f, 0.0f);
        set_vertex(vertices[3 * t + 2], p[2], nrm, 0.0f, 1.0f);
        indices[3 * t + 0] = 3 * t + 0;
        indices[3 * t + 1] = 3 * t + 1;
        indices[3 * t + 2] = 3 * t + 2;
    }
    return WF_OK;
}

WF_API wf_status wfs_heightfield(uint64_t seed, uint32_t nx, uint32_t nz, float amplitude,
                                 wf_vertex* vertices, uint32_t* indices)
{
    if (!vertices || !indices || nx == 0 || nz == 0) return WF_ERROR_INVALID_ARGUMENT;
    wfs_pcg32 r;
    pcg_seed(&r, seed, 0x5eed0002ull);
    ValueNoise2D noise(&r, 8);
    const uint32_t vx = nx + 1;
    for (uint32_t j = 0; j <= nz; ++j)
        for (uint32_t i = 0; i <= nx; ++i) {
            const float u = (float)i / (float)nx, w = (float)j / (float)nz;
            const float x = -1.0f + 2.0f * u, z = -1.0f + 2.0f * w;
            const float y = amplitude * (2.0f * noise.eval(u * 8.0, w * 8.0) - 1.0f);
            const float p[3] = {x, y, z}, up[3] = {0.0f, 1.0f, 0.0f};
            set_vertex(vertices[(size_t)j * vx + i], p, up, u, w);
        }
    // flat normals would need unshared vertices; the shared-vertex normal is
    // the average of the cell's two face normals at the vertex's lower-left cell
    for (uint32_t j = 0; j <= nz; ++j)
        for (uint32_t i = 0; i <= nx; ++i) {
            const uint32_t i0 = i < nx ? i : nx - 1, j0 = j < nz ? j : nz - 1;
            const float* a = vertices[(size_t)j0 * vx + i0].position;
            const float* b = vertices[(size_t)(j0 + 1) * vx + i0].position;
            const float* c = vertices[(size_t)j0 * vx + i0 + 1].position;
            float nrm[3];
            face_normal(a, b, c, nrm);
            std::memcpy(vertices[(size_t)j * vx + i].normal, nrm, 12);
        }
    size_t k = 0;
    for (uint32_t j = 0; j < nz; ++j)
        for (uint32_t i = 0; i < nx; ++i) {
            const uint32_t v00 = j * vx + i, v10 = j * vx + i + 1, v01 = (j + 1) * vx + i, v11 = (j + 1) * vx + i + 1;
            // (v00, v01, v10) and (v10, v01, v11): (v1-v0)x(v2-v0) points +y
            indices[k++] = v00; indices[k++] = v01; indices[k++] = v10;
            indices[k++] = v10; indices[k++] = v01; indices[k++] = v11;
        }
    return WF_OK;
}

WF_API wf_status wfs_icosphere(uint32_t subdiv, wf_vertex* vertices, uint32_t* indices,
                               uint32_t* out_vertex_count, uint32_t* out_index_count)
{
    if (subdiv > 9) return WF_ERROR_INVALID_ARGUMENT;
    const uint64_t ntri = 20ull << (2 * subdiv);
    if (out_vertex_count) *out_vertex_count = (uint32_t)(3 * ntri);
    if (out_index_count) *out_index_count = (uint32_t)(3 * ntri);
    if (!vertices || !indices) return WF_OK;
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    double base[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                          {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    static const int faces[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                                     {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                                     {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    struct V { double x, y, z; };
    auto unit = [](V a) { const double l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); return V{a.x / l, a.y / l, a.z / l}; };
    std::vector<V> tris;
    tris.reserve(3 * ntri);
    for (auto& f : faces)
        for (int k = 0; k < 3; ++k) tris.push_back(unit(V{base[f[k]][0], base[f[k]][1], base[f[k]][2]}));
    for (uint32_t s = 0; s < subdiv; ++s) {
        std::vector<V> nt;
        nt.reserve(tris.size() * 4);
        for (size_t i = 0; i < tris.size(); i += 3) {
            const V a = tris[i], b = tris[i + 1], c = tris[i + 2];
            const V ab = unit(V{a.x + b.x, a.y + b.y, a.z + b.z});
            const V bc = unit(V{b.x + c.x, b.y + c.y, b.z + c.z});
            const V ca = unit(V{c.x + a.x, c.y + a.y, c.z + a.z});
            const V q[12] = {a, ab, ca, ab, b, bc, ca, bc, c, ab, bc, ca};
            nt.insert(nt.end(), q, q + 12);
        }
        tris.swap(nt);
    }
    const double kPi = 3.14159265358979323846;
    for (size_t tri = 0; tri < tris.size() / 3; ++tri) {
        const V a = tris[3 * tri], b = tris[3 * tri + 1], c = tris[3 * tri + 2];
        // orient so that (v1-v0)x(v2-v0) points away from the centre
        const double e1[3] = {b.x - a.x, b.y - a.y, b.z - a.z}, e2[3] = {c.x - a.x, c.y - a.y, c.z - a.z};
        const double nx = e1[1] * e2[2] - e1[2] * e2[1], ny = e1[2] * e2[0] - e1[0] * e2[2], nz = e1[0] * e2[1] - e1[1] * e2[0];
        const bool flip = nx * (a.x + b.x + c.x) + ny * (a.y + b.y + c.y) + nz * (a.z + b.z + c.z) < 0.0;
        const V q[3] = {a, flip ? c : b, flip ? b : c};
        for (int k = 0; k < 3; ++k) {
            const V& p = q[k];
            const size_t i = 3 * tri + k;
            const float pos[3] = {(float)p.x, (float)p.y, (float)p.z};
            const float u = (float)(0.5 + std::atan2(p.z, p.x) / (2.0 * kPi));
            const float w = (float)(0.5 - std::asin(p.y) / kPi);
            set_vertex(vertices[i], pos, pos, u, w);
            indices[i] = (uint32_t)i;
        }
    }
    return WF_OK;
}

WF_API wf_status wfs_box_room(wf_vertex* vertices, uint32_t* indices)
{
    if (!vertices || !indices) return WF_ERROR_INVALID_ARGUMENT;
    // each wall: inward normal n, two in-plane axes (a, b) with a x b = -n so
    // that (v1-v0)x(v2-v0) = +n for the (p00, p10, p01) winding below
    struct Wall { float n[3]; float a[3]; float b[3]; };
    const Wall walls[6] = {
        {{0, 1, 0}, {1, 0, 0}, {0, 0, 1}},    // floor  y=-1, normal +y
        {{0, -1, 0}, {0, 0, 1}, {1, 0, 0}},   // ceiling y=+1, normal -y
        {{1, 0, 0}, {0, 0, 1}, {0, 1, 0}},    // wall x=-1, normal +x
        {{-1, 0, 0}, {0, 1, 0}, {0, 0, 1}},   // wall x=+1, normal -x
        {{0, 0, 1}, {0, 1, 0}, {1, 0, 0}},    // wall z=-1, normal +z
        {{0, 0, -1}, {1, 0, 0}, {0, 1, 0}},   // wall z=+1, normal -z
    };
    uint32_t vi = 0, ii = 0;
    for (const Wall& w : walls) {
        for (int half = 0; half < 2; ++half) {
            const float a0 = half ? 0.0f : -1.0f, a1 = half ? 1.0f : 0.0f;
            float p[4][3];
            const float as[4] = {a0, a1, a0, a1}, bs[4] = {-1.0f, -1.0f, 1.0f, 1.0f};
            for (int c = 0; c < 4; ++c)
                for (int k = 0; k < 3; ++k) p[c][k] = -w.n[k] + as[c] * w.a[k] + bs[c] * w.b[k];
            // winding check: (p1-p0)x(p2-p0) must point along n
            float nrm[3];
            face_normal(p[0], p[1], p[2], nrm);
            const bool flip = wf::dot3(nrm, w.n) < 0.0f;
            for (int c = 0; c < 4; ++c) set_vertex(vertices[vi + c], p[c], w.n, 0.5f * (as[c] + 1.0f), 0.5f * (bs[c] + 1.0f));
            if (!flip) {
                indices[ii++] = vi + 0; indices[ii++] = vi + 1; indices[ii++] = vi + 2;
                indices[ii++] = vi + 2; indices[ii++] = vi + 1; indices[ii++] = vi + 3;
            } else {
                indices[ii++] = vi + 0; indices[ii++] = vi + 2; indices[ii++] = vi + 1;
                indices[ii++] = vi + 1; indices[ii++] = vi + 2; indices[ii++] = vi + 3;
            }
            vi += 4;
        }
    }
    return WF_OK;
}

WF_API wf_status wfs_quad(wf_vertex* vertices, uint32_t* indices)
{
    if (!vertices || !indices) return WF_ERROR_INVALID_ARGUMENT;
    const float down[3] = {0.0f, -1.0f, 0.0f};
    const float p[4][3] = {{-0.5f, 0.0f, -0.5f}, {0.5f, 0.0f, -0.5f}, {-0.5f, 0.0f, 0.5f}, {0.5f, 0.0f, 0.5f}};
    const float uv[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
    for (int c = 0; c < 4; ++c) set_vertex(vertices[c], p[c], down, uv[c][0], uv[c][1]);
    // (p1-p0)x(p2-p0) = (1,0,0)x(0,0,1) = (0,-1,0): front faces down
    const uint32_t idx[6] = {0, 1, 2, 2, 1, 3};
    std::memcpy(indices, idx, sizeof(idx));
    return WF_OK;
}

WF_API void wfs_random_transform(wfs_pcg32* rng, float scale_min, float scale_max, float t_extent, float out[16])
{
    const double kPi = 3.14159265358979323846;
    const double u1 = pcg_float(rng), u2 = pcg_float(rng), u3 = pcg_float(rng);
    // Shoemake, "Uniform random rotations", Graphics Gems III (1992)
    const double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
    const double qx = a * std::sin(2 * kPi * u2), qy = a * std::cos(2 * kPi * u2);
    const double qz = b * std::sin(2 * kPi * u3), qw = b * std::cos(2 * kPi * u3);
    const double R[3][3] = {{1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)},
                            {2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)},
                            {2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)}};
    const double s = scale_min + (scale_max - scale_min) * (double)pcg_float(rng);
    const double t[3] = {t_extent * (2.0 * pcg_float(rng) - 1.0), t_extent * (2.0 * pcg_float(rng) - 1.0),
                         t_extent * (2.0 * pcg_float(rng) - 1.0)};
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            double v;
            if (c < 3) v = (r < 3) ? R[r][c] * s : 0.0;
            else v = (r < 3) ? t[r] : 1.0;
            out[c * 4 + r] = (float)v;
        }
}

WF_API wf_status wfs_sphere_rays(uint64_t seed, const float bmin[3], const float bmax[3], uint32_t n, wf_ray* out)
{
    if (!bmin || !bmax || !out) return WF_ERROR_INVALID_ARGUMENT;
    const double kPi = 3.14159265358979323846;
    wfs_pcg32 r;
    pcg_seed(&r, seed, 0x5eed0003ull);
    for (uint32_t i = 0; i < n; ++i) {
        wf_ray& ray = out[i];
        for (int k = 0; k < 3; ++k) ray.origin[k] = bmin[k] + (bmax[k] - bmin[k]) * pcg_float(&r);
        const double z = 1.0 - 2.0 * (double)pcg_float(&r);
        const double phi = 2.0 * kPi * (double)pcg_float(&r);
        const double s = std::sqrt(std::fmax(0.0, 1.0 - z * z));
        ray.direction[0] = (float)(s * std::cos(phi));
        ray.direction[1] = (float)(s * std::sin(phi));
        ray.direction[2] = (float)z;
        ray.tmin = 1e-8f;
        ray.tmax = FLT_MAX;
    }
    return WF_OK;
}

WF_API wf_status wfs_value_noise_texture(uint64_t seed, uint32_t w, uint32_t h, uint32_t cells, float lo, float hi, float* out_rgb)
{
    if (!out_rgb || w == 0 || h == 0 || cells == 0) return WF_ERROR_INVALID_ARGUMENT;
    wfs_pcg32 r;
    pcg_seed(&r, seed, 0x5eed0004ull);
    ValueNoise2D ch[3] = {ValueNoise2D(&r, cells), ValueNoise2D(&r, cells), ValueNoise2D(&r, cells)};
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) {
                const double fx = ((double)x + 0.5) / w * cells, fy = ((double)y + 0.5) / h * cells;
                out_rgb[((size_t)y * w + x) * 3 + c] = lo + (hi - lo) * ch[c].eval(fx, fy);
            }
    return WF_OK;
}

} // extern "C"
