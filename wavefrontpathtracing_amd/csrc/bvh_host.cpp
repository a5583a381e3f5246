// Host BVH builders: binned-SAH BLAS (bvh_builder.cpp) and agglomerative
// TLAS (tlas_builder.cpp), producing the reference's node order and bytes.
//
// The BLAS build differs from the reference in mechanics only:
//  * per-triangle bounds and centroids are precomputed once (SoA) instead of
//    being re-derived from the three vertices in every pass; min/max are exact,
//    so node and bin bounds are the same values;
//  * large subtrees are built on worker threads into a scratch tree that is
//    renumbered afterwards in the reference's depth-first allocation order
//    (children of a node get two consecutive slots when the node is split,
//    slot 1 is left unused, bvh_builder.cpp:44-45,146-154).
#include "wf_host.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <future>
#include <vector>

namespace wf {

namespace {

// Bin of a centroid, min(B-1, int((c - lo) * scale)) (bvh_builder.cpp:194-195).
// A centroid extent so small that `scale` overflows makes the product inf or
// NaN, where the reference's float->int conversion is undefined; the bin then
// saturates (inf -> B-1, NaN -> 0).  Finite products give the reference's bin.
inline int32_t bin_of(float c, float lo, float scale, uint32_t bins)
{
    const float f = (c - lo) * scale;
    if (!(f >= 0.0f)) return 0;
    if (f >= (float)bins) return (int32_t)bins - 1;
    const int32_t b = (int32_t)f;
    return b < (int32_t)bins - 1 ? b : (int32_t)bins - 1;
}

struct TmpNode {
    float mn[3], mx[3];
    uint32_t first, count;   // triangle range (leaf) / original range (interior)
    int32_t left, right;     // scratch indices of children, -1 for a leaf
};

struct BlasBuilder {
    uint32_t n = 0;
    uint32_t bins = 8;
    bool single_prim = false;
    std::vector<float> tmin[3], tmax[3], cen[3];   // per triangle, SoA
    std::vector<uint32_t> idx;
    std::vector<TmpNode> tmp;
    std::atomic<uint32_t> tmp_at{0};

    uint32_t alloc_tmp() { return tmp_at.fetch_add(1, std::memory_order_relaxed); }

    // calc_node_min_max (bvh_builder.cpp:71-93)
    void node_bounds(TmpNode& nd, float cmn[3], float cmx[3]) const
    {
        for (int k = 0; k < 3; ++k) { nd.mn[k] = FLT_MAX; nd.mx[k] = -FLT_MAX; cmn[k] = FLT_MAX; cmx[k] = -FLT_MAX; }
        for (uint32_t i = nd.first, e = nd.first + nd.count; i < e; ++i) {
            const uint32_t t = idx[i];
            for (int k = 0; k < 3; ++k) {
                nd.mn[k] = gmin(nd.mn[k], tmin[k][t]);
                nd.mx[k] = gmax(nd.mx[k], tmax[k][t]);
                cmn[k] = gmin(cmn[k], cen[k][t]);
                cmx[k] = gmax(cmx[k], cen[k][t]);
            }
        }
    }

    // find_best_split_plane (bvh_builder.cpp:170-250)
    float best_split(const TmpNode& nd, const float cmn[3], const float cmx[3], uint32_t& axis_out, uint32_t& split_out) const
    {
        float best = FLT_MAX;
        const uint32_t B = bins;
        std::vector<float> bmn(3 * B), bmx(3 * B), la(B), ra(B);
        std::vector<uint32_t> bc(B);
        for (uint32_t axis = 0; axis < 3; ++axis) {
            const float lo = cmn[axis], hi = cmx[axis];
            if (lo == hi) continue;
            std::fill(bc.begin(), bc.end(), 0u);
            std::fill(bmn.begin(), bmn.end(), FLT_MAX);
            std::fill(bmx.begin(), bmx.end(), -FLT_MAX);
            const float scale = (float)B / (hi - lo);
            const float* c = cen[axis].data();
            for (uint32_t i = nd.first, e = nd.first + nd.count; i < e; ++i) {
                const uint32_t t = idx[i];
                const int32_t b = bin_of(c[t], lo, scale, B);
                bc[b]++;
                for (int k = 0; k < 3; ++k) {
                    bmn[3 * b + k] = gmin(bmn[3 * b + k], tmin[k][t]);
                    bmx[3 * b + k] = gmax(bmx[3 * b + k], tmax[k][t]);
                }
            }
            float lmn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, lmx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            float rmn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, rmx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            uint32_t ls = 0, rs = 0;
            for (uint32_t i = 0; i + 1 < B; ++i) {
                ls += bc[i];
                const uint32_t j = B - 1 - i;
                rs += bc[j];
                for (int k = 0; k < 3; ++k) {
                    lmn[k] = gmin(lmn[k], bmn[3 * i + k]); lmx[k] = gmax(lmx[k], bmx[3 * i + k]);
                    rmn[k] = gmin(rmn[k], bmn[3 * j + k]); rmx[k] = gmax(rmx[k], bmx[3 * j + k]);
                }
                la[i] = (float)ls * half_area(lmx[0] - lmn[0], lmx[1] - lmn[1], lmx[2] - lmn[2]);
                ra[B - 2 - i] = (float)rs * half_area(rmx[0] - rmn[0], rmx[1] - rmn[1], rmx[2] - rmn[2]);
            }
            for (uint32_t i = 0; i + 1 < B; ++i) {
                const float cost = la[i] + ra[i];
                if (cost < best) { best = cost; axis_out = axis; split_out = i + 1; }
            }
        }
        return best;
    }

    // subdivide_node (bvh_builder.cpp:100-168) for the subtree rooted at `ni`
    void build_subtree(uint32_t ni, int depth)
    {
        std::vector<std::pair<uint32_t, int>> work;   // (node, depth)
        work.emplace_back(ni, depth);
        std::vector<std::future<void>> forks;
        while (!work.empty()) {
            const uint32_t cur = work.back().first;
            const int d = work.back().second;
            work.pop_back();
            TmpNode& nd = tmp[cur];
            float cmn[3], cmx[3];
            node_bounds(nd, cmn, cmx);
            uint32_t axis = 0, split = 0;
            const float cost = best_split(nd, cmn, cmx, axis, split);
            if (single_prim) {
                if (nd.count == 1) continue;
            } else {
                const float parent = (float)nd.count * half_area(nd.mx[0] - nd.mn[0], nd.mx[1] - nd.mn[1], nd.mx[2] - nd.mn[2]);
                if (cost >= parent) continue;
            }
            if (cost == FLT_MAX) continue;   // no axis with centroid extent: stays a leaf
            int32_t i = (int32_t)nd.first;
            int32_t j = i + (int32_t)nd.count - 1;
            const float scale = (float)bins / (cmx[axis] - cmn[axis]);
            const float lo = cmn[axis];
            const float* c = cen[axis].data();
            while (i <= j) {
                const int32_t b = bin_of(c[idx[i]], lo, scale, bins);
                if ((uint32_t)b < split) ++i;
                else { std::swap(idx[i], idx[j]); --j; }
            }
            const uint32_t nl = (uint32_t)i - nd.first;
            if (nl == 0 || nl == nd.count) continue;
            const uint32_t l = alloc_tmp(), r = alloc_tmp();
            tmp[l] = TmpNode{{0, 0, 0}, {0, 0, 0}, nd.first, nl, -1, -1};
            tmp[r] = TmpNode{{0, 0, 0}, {0, 0, 0}, (uint32_t)i, nd.count - nl, -1, -1};
            nd.left = (int32_t)l;
            nd.right = (int32_t)r;
            // large right subtrees go to a worker; subtrees are disjoint in idx
            if (d < 6 && tmp[r].count >= 65536) {
                forks.emplace_back(std::async(std::launch::async, [this, r, d] { build_subtree(r, d + 1); }));
                work.emplace_back(l, d + 1);
            } else {
                work.emplace_back(r, d + 1);
                work.emplace_back(l, d + 1);
            }
        }
        for (auto& f : forks) f.get();
    }
};

} // namespace

wf_status build_blas(const float* pos9, uint32_t n, const wf_build_options& opt, BlasData& out)
{
    if (!pos9 || n == 0) return WF_ERROR_INVALID_ARGUMENT;
    if (opt.interval_count < 2 || opt.interval_count > 1024) return WF_ERROR_INVALID_ARGUMENT;
    if (n > (1u << 30)) return WF_ERROR_CAPACITY;
    for (uint64_t i = 0; i < 9ull * n; ++i)
        if (!std::isfinite(pos9[i])) return WF_ERROR_INVALID_ARGUMENT;

    BlasBuilder b;
    b.n = n;
    b.bins = opt.interval_count;
    b.single_prim = opt.subdivide_single_prim != 0;
    for (int k = 0; k < 3; ++k) { b.tmin[k].resize(n); b.tmax[k].resize(n); b.cen[k].resize(n); }
    b.idx.resize(n);
    for (uint32_t t = 0; t < n; ++t) {
        const float* p = pos9 + 9ull * t;
        b.idx[t] = t;
        for (int k = 0; k < 3; ++k) {
            // get_triangle_min_max (bvh_builder.cpp:257-264)
            b.tmin[k][t] = gmin(gmin(p[k], p[3 + k]), p[6 + k]);
            b.tmax[k][t] = gmax(gmax(p[k], p[3 + k]), p[6 + k]);
            // get_triangle_centroid (bvh_builder.cpp:252-255)
            b.cen[k][t] = ((p[k] + p[3 + k]) + p[6 + k]) * 0.3333f;
        }
    }
    b.tmp.resize(2ull * n);
    const uint32_t root = b.alloc_tmp();
    b.tmp[root] = TmpNode{{0, 0, 0}, {0, 0, 0}, 0, n, -1, -1};
    b.build_subtree(root, 0);

    // renumber in the reference's allocation order
    out.nodes.assign(2ull * n, wf_bvh_node{});
    uint32_t counter = 2;
    std::vector<std::pair<uint32_t, uint32_t>> st;   // (scratch, final)
    st.emplace_back(root, 0);
    while (!st.empty()) {
        auto [ti, fi] = st.back();
        st.pop_back();
        const TmpNode& t = b.tmp[ti];
        wf_bvh_node& o = out.nodes[fi];
        std::memcpy(o.aabb_min, t.mn, 12);
        std::memcpy(o.aabb_max, t.mx, 12);
        if (t.left >= 0) {
            const uint32_t L = counter, R = counter + 1;
            counter += 2;
            o.left_first = L;
            o.prim_count = 0;
            st.emplace_back((uint32_t)t.right, R);
            st.emplace_back((uint32_t)t.left, L);
        } else {
            o.left_first = t.first;
            o.prim_count = t.count;
        }
    }
    out.node_count = counter;
    out.nodes.resize(counter);
    out.idx = std::move(b.idx);
    out.tris.resize(n);
    std::memcpy(out.tris.data(), pos9, sizeof(wf_bvh_triangle) * (size_t)n);
    uint32_t max_leaf = 0;
    for (const auto& nd : out.nodes) max_leaf = std::max(max_leaf, nd.prim_count);
    out.max_leaf = max_leaf;
    return WF_OK;
}

void blas_blob(const BlasData& d, std::vector<uint8_t>& blob)
{
    const uint32_t n = (uint32_t)d.tris.size();
    wf_bvh_header h;
    h.nodes_offset = sizeof(wf_bvh_header);
    h.triangles_offset = h.nodes_offset + (uint32_t)(sizeof(wf_bvh_node) * d.node_count);
    h.indices_offset = h.triangles_offset + (uint32_t)(sizeof(wf_bvh_triangle) * n);
    blob.resize((size_t)h.indices_offset + 4ull * n);
    std::memcpy(blob.data(), &h, sizeof(h));
    std::memcpy(blob.data() + h.nodes_offset, d.nodes.data(), sizeof(wf_bvh_node) * d.node_count);
    std::memcpy(blob.data() + h.triangles_offset, d.tris.data(), sizeof(wf_bvh_triangle) * n);
    std::memcpy(blob.data() + h.indices_offset, d.idx.data(), 4ull * n);
}

// find_best_match (tlas_builder.cpp:86-110): the first candidate of least
// metric.  WF_TLAS_METRIC_REFERENCE is the reference's "max of mins / max of
// maxes"; WF_TLAS_METRIC_UNION the half area of the union box.
static uint32_t best_match(const std::vector<wf_tlas_node>& nodes, uint32_t a, const uint32_t* ids, uint32_t count, uint32_t metric)
{
    float smallest = FLT_MAX;
    uint32_t best = ~0u;
    const wf_tlas_node& A = nodes[ids[a]];
    for (uint32_t b = 0; b < count; ++b) {
        if (b == a) continue;
        const wf_tlas_node& Bn = nodes[ids[b]];
        float area;
        if (metric == WF_TLAS_METRIC_UNION) {
            float lo[3], hi[3];
            for (int k = 0; k < 3; ++k) {
                lo[k] = gmin(A.aabb_min[k], Bn.aabb_min[k]);
                hi[k] = gmax(A.aabb_max[k], Bn.aabb_max[k]);
            }
            area = half_area(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]);
        } else {
            float hi[3], lo[3];
            for (int k = 0; k < 3; ++k) {
                hi[k] = gmax(A.aabb_min[k], Bn.aabb_min[k]);
                lo[k] = gmax(A.aabb_max[k], Bn.aabb_max[k]);
            }
            area = half_area(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]);
        }
        if (area < smallest) { smallest = area; best = b; }
    }
    return best;
}

// tlas_builder.cpp:6-68.  With the union metric, an A that was the last list
// entry follows its move into B's slot (the reference keeps the stale index).
wf_status build_tlas(const wf_bvh_instance* inst, uint32_t m, TlasData& out, uint32_t metric)
{
    if (!inst || m == 0 || metric > WF_TLAS_METRIC_UNION) return WF_ERROR_INVALID_ARGUMENT;
    if (m > kMaxTlasInstances) return WF_ERROR_CAPACITY;
    out.nodes.assign(2ull * m, wf_tlas_node{});
    std::vector<uint32_t> ids(m);
    uint32_t at = 1;
    for (uint32_t i = 0; i < m; ++i) {
        ids[i] = at;
        wf_tlas_node& nd = out.nodes[at++];
        std::memcpy(nd.aabb_min, inst[i].aabb_min, 12);
        std::memcpy(nd.aabb_max, inst[i].aabb_max, 12);
        nd.instance_idx = i;
        nd.left_right = 0;
    }
    uint32_t count = m;
    uint32_t A = 0;
    uint32_t B = best_match(out.nodes, A, ids.data(), count, metric);
    uint64_t guard = 0;
    const uint64_t guard_max = 4ull * m * m + 16;
    while (count > 1) {
        if (B == ~0u || ++guard > guard_max) return WF_ERROR_INVALID_ARGUMENT;
        const uint32_t C = best_match(out.nodes, B, ids.data(), count, metric);
        if (C != A) { A = B; B = C; continue; }
        const uint32_t ia = ids[A], ib = ids[B];
        wf_tlas_node& nn = out.nodes[at];
        nn.left_right = ia + (ib << 16);
        for (int k = 0; k < 3; ++k) {
            nn.aabb_min[k] = gmin(out.nodes[ia].aabb_min[k], out.nodes[ib].aabb_min[k]);
            nn.aabb_max[k] = gmax(out.nodes[ia].aabb_max[k], out.nodes[ib].aabb_max[k]);
        }
        ids[A] = at++;
        ids[B] = ids[count - 1];
        if (metric == WF_TLAS_METRIC_UNION && A == count - 1) A = B;
        --count;
        B = best_match(out.nodes, A, ids.data(), count, metric);
    }
    out.nodes[0] = out.nodes[ids[A]];
    out.node_count = at;
    out.nodes.resize(at);
    out.instances.assign(inst, inst + m);
    return WF_OK;
}

void tlas_blob(const TlasData& d, std::vector<uint8_t>& blob)
{
    wf_tlas_header h;
    h.nodes_offset = sizeof(wf_tlas_header);
    h.instances_offset = h.nodes_offset + (uint32_t)(sizeof(wf_tlas_node) * d.node_count);
    blob.resize((size_t)h.instances_offset + sizeof(wf_bvh_instance) * d.instances.size());
    std::memcpy(blob.data(), &h, sizeof(h));
    std::memcpy(blob.data() + h.nodes_offset, d.nodes.data(), sizeof(wf_tlas_node) * d.node_count);
    std::memcpy(blob.data() + h.instances_offset, d.instances.data(), sizeof(wf_bvh_instance) * d.instances.size());
}

void make_bvh_instance(const float* transform, const float* bmin, const float* bmax, uint32_t bvh_index, wf_bvh_instance& out)
{
    std::memset(&out, 0, sizeof(out));
    mat4_inverse(transform, out.world_to_local);
    out.bvh_index = bvh_index;
    for (uint32_t c = 0; c < 8; ++c) {
        const float p[4] = {(c & 1) ? bmax[0] : bmin[0], (c & 2) ? bmax[1] : bmin[1], (c & 4) ? bmax[2] : bmin[2], 1.0f};
        float w[4];
        mat4_mul_vec4(transform, p, w);
        for (int k = 0; k < 3; ++k) {
            out.aabb_min[k] = gmin(out.aabb_min[k], w[k]);
            out.aabb_max[k] = gmax(out.aabb_max[k], w[k]);
        }
    }
}

} // namespace wf

// ---- C-ABI wrappers --------------------------------------------------------

extern "C" WF_API wf_status wf_bvh_build(const float* tri_positions, uint32_t n, const wf_build_options* options,
                                         void* out_blob, uint64_t capacity, uint64_t* out_size, uint32_t* out_node_count)
{
    if (!tri_positions || !out_size) return WF_ERROR_INVALID_ARGUMENT;
    wf_build_options opt = options ? *options : wf_build_options{8, 0};
    wf::BlasData d;
    wf_status s = wf::build_blas(tri_positions, n, opt, d);
    if (s != WF_OK) return s;
    std::vector<uint8_t> blob;
    wf::blas_blob(d, blob);
    *out_size = blob.size();
    if (out_node_count) *out_node_count = d.node_count;
    if (!out_blob) return WF_OK;
    if (capacity < blob.size()) return WF_ERROR_CAPACITY;
    std::memcpy(out_blob, blob.data(), blob.size());
    return WF_OK;
}

extern "C" WF_API wf_status wf_tlas_build(const wf_bvh_instance* instances, uint32_t m,
                                          void* out_blob, uint64_t capacity, uint64_t* out_size, uint32_t* out_node_count)
{
    return wf_tlas_build_metric(instances, m, WF_TLAS_METRIC_REFERENCE, out_blob, capacity, out_size, out_node_count);
}

extern "C" WF_API wf_status wf_tlas_build_metric(const wf_bvh_instance* instances, uint32_t m, uint32_t metric,
                                                 void* out_blob, uint64_t capacity, uint64_t* out_size, uint32_t* out_node_count)
{
    if (!instances || !out_size) return WF_ERROR_INVALID_ARGUMENT;
    wf::TlasData d;
    wf_status s = wf::build_tlas(instances, m, d, metric);
    if (s != WF_OK) return s;
    std::vector<uint8_t> blob;
    wf::tlas_blob(d, blob);
    *out_size = blob.size();
    if (out_node_count) *out_node_count = d.node_count;
    if (!out_blob) return WF_OK;
    if (capacity < blob.size()) return WF_ERROR_CAPACITY;
    std::memcpy(out_blob, blob.data(), blob.size());
    return WF_OK;
}

extern "C" WF_API wf_status wf_make_bvh_instance(const float transform[16], const float blas_min[3], const float blas_max[3],
                                                 uint32_t bvh_index, wf_bvh_instance* out)
{
    if (!transform || !blas_min || !blas_max || !out) return WF_ERROR_INVALID_ARGUMENT;
    wf::make_bvh_instance(transform, blas_min, blas_max, bvh_index, *out);
    return WF_OK;
}
