/*
 * vgi_bvh.cpp — scene upload and the host BVH builder (binned SAH).
 *
 * Replaces violet's driver-built BLAS/TLAS (render_scene.rs:1170-1210 BLAS
 * per geometry group, :1557-1712 TLAS over instances).  Instances are baked
 * to world space by the caller, so one BVH over all triangles serves every
 * ray.  Build: a binary tree by 3-axis binned SAH (32 bins; leaves hold at
 * most 15 triangles; from depth 36 the split is the object median, so the
 * depth stays bounded by kBvhMaxDepth), then collapsed into a 4-wide tree:
 * each 4-wide node takes its binary node's two children and repeatedly
 * opens the inner child with the largest surface area until it has four
 * children (or only leaves are left).  Each 4-wide node stores its
 * children's padded boxes on an 8-bit grid spanning them (vgi_bvh.h).
 */
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <new>
#include <vector>

#include "vgi_bvh.h"
#include "vgi_internal.h"

using namespace vgi;

namespace {

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Box& b) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    void grow(const float* p) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    double area() const {
        double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0 && dy >= 0 && dz >= 0)) return 0.0;
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

// Binary build node: both child boxes (padded) and links.
struct Node2 {
    Box box[2];
    int32_t link[2];
};

struct Prim {
    Box box;
    float c[3];
    uint32_t id;
};

constexpr int kBins = 32;
// Largest vertex coordinate magnitude a scene may have: keeps every node's
// grid step a float and the traversal's plane distances finite (the step
// is below 2^34, so the stored step * 2^24 times |1/d| stays below 2^125
// and q * step * |1/d| below 2^109 for |1/d| <= 1e20).
constexpr float kMaxCoord = 1099511627776.0f;  // 2^40
constexpr int kMaxLeaf = 15;
constexpr int kMedianDepth = 36;

struct Builder {
    std::vector<Prim> prims;
    std::vector<Node2> nodes;
    std::vector<uint32_t> order;
    int max_depth = 0;

    static void pad(Box& b) {
        float s = 1.0f;
        for (int a = 0; a < 3; ++a) s = std::max(s, std::max(std::fabs(b.lo[a]), std::fabs(b.hi[a])));
        float p = s * (1.0f / 65536.0f);
        for (int a = 0; a < 3; ++a) {
            b.lo[a] -= p;
            b.hi[a] += p;
        }
    }

    int32_t make_leaf(int begin, int end) {
        uint32_t first = (uint32_t)order.size();
        for (int i = begin; i < end; ++i) order.push_back(prims[i].id);
        return ~(int32_t)((first << 4) | (uint32_t)(end - begin));
    }

    int32_t build(int begin, int end, int depth, Box& out) {
        max_depth = std::max(max_depth, depth);
        Box box, cbox;
        for (int i = begin; i < end; ++i) {
            box.grow(prims[i].box);
            cbox.grow(prims[i].c);
        }
        out = box;
        int n = end - begin;
        if (n <= 2 && n <= kMaxLeaf) return make_leaf(begin, end);
        int mid = -1;
        if (depth < kMedianDepth) {
            // binned SAH over 3 axes
            double best_cost = INFINITY;
            int best_axis = -1, best_split = -1;
            for (int a = 0; a < 3; ++a) {
                float ext = cbox.hi[a] - cbox.lo[a];
                if (!(ext > 0.0f)) continue;
                Box bb[kBins];
                int cnt[kBins] = {0};
                float k = (float)kBins / ext;
                for (int i = begin; i < end; ++i) {
                    int b = (int)((prims[i].c[a] - cbox.lo[a]) * k);
                    b = std::min(std::max(b, 0), kBins - 1);
                    cnt[b]++;
                    bb[b].grow(prims[i].box);
                }
                double right_area[kBins];
                int right_cnt[kBins];
                Box acc;
                int c = 0;
                for (int b = kBins - 1; b > 0; --b) {
                    acc.grow(bb[b]);
                    c += cnt[b];
                    right_area[b] = acc.area();
                    right_cnt[b] = c;
                }
                Box lacc;
                int lc = 0;
                for (int b = 1; b < kBins; ++b) {
                    lacc.grow(bb[b - 1]);
                    lc += cnt[b - 1];
                    if (lc == 0 || right_cnt[b] == 0) continue;
                    double cost = lacc.area() * lc + right_area[b] * right_cnt[b];
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = a;
                        best_split = b;
                    }
                }
            }
            double parent = box.area();
            double leaf_cost = (double)n;
            double split_cost = parent > 0 ? 0.125 + best_cost / parent : INFINITY;
            if (best_axis >= 0 && !(n <= 4 && split_cost >= leaf_cost)) {
                float ext = cbox.hi[best_axis] - cbox.lo[best_axis];
                float k = (float)kBins / ext;
                auto it = std::partition(prims.begin() + begin, prims.begin() + end, [&](const Prim& p) {
                    int b = (int)((p.c[best_axis] - cbox.lo[best_axis]) * k);
                    b = std::min(std::max(b, 0), kBins - 1);
                    return b < best_split;
                });
                mid = (int)(it - prims.begin());
                if (mid == begin || mid == end) mid = -1;
            } else if (n <= kMaxLeaf && best_axis >= 0) {
                return make_leaf(begin, end);
            } else if (n <= kMaxLeaf && best_axis < 0) {
                return make_leaf(begin, end);
            }
        }
        if (mid < 0) {
            if (n <= kMaxLeaf && depth >= kMedianDepth) return make_leaf(begin, end);
            // object median along the widest centroid axis
            int a = 0;
            float best = -1;
            for (int k = 0; k < 3; ++k) {
                float e = cbox.hi[k] - cbox.lo[k];
                if (e > best) {
                    best = e;
                    a = k;
                }
            }
            mid = begin + n / 2;
            std::nth_element(prims.begin() + begin, prims.begin() + mid, prims.begin() + end,
                             [a](const Prim& x, const Prim& y) { return x.c[a] < y.c[a] || (x.c[a] == y.c[a] && x.id < y.id); });
        }
        int32_t idx = (int32_t)nodes.size();
        nodes.push_back(Node2{});
        Box lb, rb;
        int32_t l = build(begin, mid, depth + 1, lb);
        int32_t r = build(mid, end, depth + 1, rb);
        pad(lb);
        pad(rb);
        Node2& nd = nodes[idx];
        nd.box[0] = lb;
        nd.box[1] = rb;
        nd.link[0] = l;
        nd.link[1] = r;
        return idx;
    }
};

// Collapse of the binary tree into 4-wide nodes (depth-first, so a parent
// precedes its children in memory).
struct Collapser {
    const std::vector<Node2>& n2;
    std::vector<BvhNode> out;
    int max_level = 0;

    int32_t emit(int32_t bin, int level) {
        max_level = std::max(max_level, level);
        Box box[4];
        int32_t link[4];
        int cnt = 2;
        for (int k = 0; k < 2; ++k) {
            box[k] = n2[bin].box[k];
            link[k] = n2[bin].link[k];
        }
        while (cnt < 4) {
            int best = -1;
            double best_area = -1.0;
            for (int k = 0; k < cnt; ++k)
                if (link[k] >= 0 && box[k].area() > best_area) {
                    best_area = box[k].area();
                    best = k;
                }
            if (best < 0) break;
            const Node2& c = n2[link[best]];
            box[best] = c.box[0];
            link[best] = c.link[0];
            box[cnt] = c.box[1];
            link[cnt] = c.link[1];
            ++cnt;
        }
        const int32_t idx = (int32_t)out.size();
        out.push_back(BvhNode{});
        int32_t child[4];
        for (int k = 0; k < 4; ++k) child[k] = k < cnt && link[k] >= 0 ? emit(link[k], level + 1) : (k < cnt ? link[k] : ~0);
        BvhNode& nd = out[idx];
        quantise(nd, box, child, cnt);
        return idx;
    }

    // Smallest power of two >= x (x >= 0), at least 2^-60.  Positions are
    // bounded by kMaxCoord, so x < 2^34 and the power of two is a float.
    static double pow2_at_least(double x) {
        int e = 0;
        const double m = std::frexp(x, &e);  // x = m * 2^e, m in [0.5, 1)
        if (!(x > 0.0)) e = -60;
        else if (m == 0.5) e -= 1;
        return std::ldexp(1.0, std::max(-60, e));
    }

    // The node's grid: per axis, the origin is the lowest used child plane
    // and 255 steps of the (power-of-two) scale reach the highest.  A child's
    // lo rounds down and its hi up, so the quantised box contains the padded
    // one exactly (computed in double: org, lo and hi are floats and the
    // scale a power of two, so (lo - org) / scale is exact).
    static void quantise(BvhNode& nd, const Box* box, const int32_t* child, int cnt) {
        bool used[4];
        for (int k = 0; k < 4; ++k) used[k] = k < cnt && child[k] != ~0;
        for (int a = 0; a < 3; ++a) {
            float lo = INFINITY, hi = -INFINITY;
            for (int k = 0; k < 4; ++k)
                if (used[k]) {
                    lo = std::min(lo, box[k].lo[a]);
                    hi = std::max(hi, box[k].hi[a]);
                }
            if (!(lo <= hi)) lo = hi = 0.0f;  // no used child (cannot happen for a built tree)
            const double s = pow2_at_least(((double)hi - (double)lo) / 255.0);
            nd.org[a] = lo;
            nd.scale[a] = (float)std::ldexp(s, 24);  // see BvhNode::scale
            uint32_t qlo = 0, qhi = 0;
            for (int k = 0; k < 4; ++k) {
                uint32_t l = 255u, h = 0u;  // unused: an empty interval
                if (used[k]) {
                    const double fl = std::floor(((double)box[k].lo[a] - (double)lo) / s);
                    const double ce = std::ceil(((double)box[k].hi[a] - (double)lo) / s);
                    l = (uint32_t)std::min(255.0, std::max(0.0, fl));
                    h = (uint32_t)std::min(255.0, std::max(0.0, ce));
                }
                qlo |= l << (8 * k);
                qhi |= h << (8 * k);
            }
            nd.q[2 * a] = qlo;
            nd.q[2 * a + 1] = qhi;
        }
        for (int k = 0; k < 4; ++k) nd.link[k] = k < cnt ? child[k] : ~0;
    }
};

}  // namespace

struct vgi_scene {
    DevBuf nodes, tris, shade, positions, normals, materials, texcoords, tangents, textures, texels, texel_lut;
    SceneView view{};
    uint32_t num_vertices = 0, num_triangles = 0, depth = 0;
};

namespace vgi {
const SceneView& scene_view(const vgi_scene* s) { return s->view; }
}  // namespace vgi

template <class T>
static vgi_status upload(DevBuf& b, const std::vector<T>& v) {
    vgi_status s = b.alloc(v.size() * sizeof(T));
    if (s != VGI_OK) return s;
    if (!v.empty()) VGI_TRY_HIP(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return VGI_OK;
}

namespace vgi {

// Texel table, texture descriptors and the byte -> float lookup of the
// sampling code (render_scene.rs:344-346,768-833 for the two built-in
// textures, which the device code treats as constants).
static vgi_status build_textures(const vgi_scene_desc* desc, HostScene& hs) {
    VGI_REQUIRE(desc->num_textures == 0 || desc->textures, "vgi_scene_create: null texture array");
    VGI_REQUIRE(desc->num_textures <= 0xFFFFu - VGI_TEXTURE_FIRST, "vgi_scene_create: %u textures exceed the 16-bit indices",
                desc->num_textures);
    const uint32_t limit = VGI_TEXTURE_FIRST + desc->num_textures;
    for (uint32_t m = 0; m < desc->num_materials; ++m) {
        const vgi_material& mt = desc->materials[m];
        const uint32_t bc = mt.color_metalrough_texture & 0xFFFFu, mr = mt.color_metalrough_texture >> 16,
                       nm = mt.normal_texture & 0xFFFFu;
        VGI_REQUIRE(bc < limit && mr < limit && nm < limit,
                    "vgi_scene_create: material %u names texture %u but the scene has %u", m,
                    bc >= limit ? bc : mr >= limit ? mr : nm, desc->num_textures);
    }
    hs.textures.assign(desc->num_textures, TexDesc{});
    uint64_t total = 0;
    for (uint32_t t = 0; t < desc->num_textures; ++t) {
        const vgi_texture& x = desc->textures[t];
        VGI_REQUIRE(x.texels && x.width >= 1 && x.height >= 1 && x.width <= 32768u && x.height <= 32768u,
                    "vgi_scene_create: texture %u: bad size %ux%u or no texels", t, x.width, x.height);
        VGI_REQUIRE(x.levels >= 1 && x.levels <= kMaxTextureLevels, "vgi_scene_create: texture %u: %u mip levels", t,
                    x.levels);
        TexDesc& d = hs.textures[t];
        d.width = x.width;
        d.height = x.height;
        d.levels = x.levels;
        d.srgb = x.srgb ? 1u : 0u;
        for (uint32_t l = 0; l < x.levels; ++l) {
            d.level_offset[l] = (uint32_t)total;
            total += (uint64_t)std::max(1u, x.width >> l) * std::max(1u, x.height >> l);
            VGI_REQUIRE(total < 0xFFFFFFFFull, "vgi_scene_create: textures exceed 2^32 texels");
        }
    }
    hs.texels.resize((size_t)total);
    for (uint32_t t = 0; t < desc->num_textures; ++t) {
        const TexDesc& d = hs.textures[t];
        const uint64_t end = t + 1 < desc->num_textures ? hs.textures[t + 1].level_offset[0] : total;
        memcpy(hs.texels.data() + d.level_offset[0], desc->textures[t].texels, (size_t)(end - d.level_offset[0]) * 4u);
    }
    hs.texel_lut.resize(kTexelLutEntries);
    for (uint32_t i = 0; i < 256; ++i) {
        const double c = i / 255.0;
        hs.texel_lut[i] = (float)c;
        hs.texel_lut[256 + i] = (float)(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
    }
    return VGI_OK;
}

vgi_status build_scene_host(const vgi_scene_desc* desc, HostScene& hs) {
    VGI_REQUIRE(desc, "vgi_scene_create: null argument");
    VGI_REQUIRE(desc->positions && desc->indices && desc->num_vertices > 0 && desc->num_triangles > 0,
                "vgi_scene_create: empty scene");
    VGI_REQUIRE(desc->num_triangles < (1u << 27), "vgi_scene_create: too many triangles (%u)", desc->num_triangles);
    VGI_REQUIRE(desc->materials && desc->num_materials > 0, "vgi_scene_create: no materials");
    const uint32_t nt = desc->num_triangles, nv = desc->num_vertices;
    for (size_t i = 0; i < 3 * (size_t)nv; ++i)
        VGI_REQUIRE(std::fabs(desc->positions[i]) <= kMaxCoord,
                    "vgi_scene_create: vertex coordinate %zu is not finite or beyond 2^40", i);
    for (uint32_t i = 0; i < 3 * nt; ++i)
        VGI_REQUIRE(desc->indices[i] < nv, "vgi_scene_create: index %u out of range at %u", desc->indices[i], i);
    if (desc->triangle_material)
        for (uint32_t i = 0; i < nt; ++i)
            VGI_REQUIRE(desc->triangle_material[i] < desc->num_materials,
                        "vgi_scene_create: material index out of range at triangle %u", i);

    Builder b;
    b.prims.resize(nt);
    for (uint32_t t = 0; t < nt; ++t) {
        Prim& p = b.prims[t];
        for (int k = 0; k < 3; ++k) p.box.grow(desc->positions + 3 * desc->indices[3 * t + k]);
        for (int a = 0; a < 3; ++a) p.c[a] = 0.5f * (p.box.lo[a] + p.box.hi[a]);
        p.id = t;
    }
    b.nodes.reserve(2 * (size_t)nt / 4 + 4);
    b.order.reserve(nt);
    Box root_box;
    int32_t root = b.build(0, (int)nt, 0, root_box);
    if (root < 0) {
        // tiny scene: wrap the single leaf in a node whose other child is an
        // empty leaf with a box at +inf on every axis, where no slab interval
        // can reach it (an inverted box would not do: the slab test orders
        // the planes)
        Node2 nd{};
        nd.box[0] = root_box;
        Builder::pad(nd.box[0]);
        for (int a = 0; a < 3; ++a) nd.box[1].lo[a] = nd.box[1].hi[a] = INFINITY;
        nd.link[0] = root;
        nd.link[1] = ~0;
        b.nodes.push_back(nd);
        root = 0;
    }
    if (b.max_depth + 1 > kBvhMaxDepth)
        return set_error(VGI_E_UNSUPPORTED, "vgi_scene_create: BVH depth %d exceeds %d", b.max_depth + 1, kBvhMaxDepth);
    Collapser c4{b.nodes, {}, 0};
    c4.out.reserve(b.nodes.size() / 2 + 1);
    root = c4.emit(root, 0);

    std::vector<BvhTri>& tris = hs.tris;
    std::vector<TriShade>& shade = hs.shade;
    tris.assign(nt, BvhTri{});
    shade.assign(nt, TriShade{});
    for (uint32_t k = 0; k < nt; ++k) {
        uint32_t t = b.order[k];
        const float* p0 = desc->positions + 3 * desc->indices[3 * t + 0];
        const float* p1 = desc->positions + 3 * desc->indices[3 * t + 1];
        const float* p2 = desc->positions + 3 * desc->indices[3 * t + 2];
        tris[k].v0 = float4_{p0[0], p0[1], p0[2], asfloat(t)};
        tris[k].e1 = float4_{p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2], 0.0f};
        tris[k].e2 = float4_{p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2], 0.0f};
    }
    for (uint32_t t = 0; t < nt; ++t)
        shade[t] = TriShade{desc->indices[3 * t], desc->indices[3 * t + 1], desc->indices[3 * t + 2],
                            desc->triangle_material ? desc->triangle_material[t] : 0u};
    std::vector<float>& pos = hs.positions;
    pos.assign(desc->positions, desc->positions + 3 * (size_t)nv);
    std::vector<float>& nrm = hs.normals;
    if (desc->normals) {
        nrm.assign(desc->normals, desc->normals + 3 * (size_t)nv);
    } else {
        // flat shading: per-vertex normal slot unused; use the geometric normal
        nrm.assign(3 * (size_t)nv, 0.0f);
    }
    std::vector<Material>& mats = hs.materials;
    mats.resize(desc->num_materials);
    memcpy(mats.data(), desc->materials, sizeof(Material) * desc->num_materials);
    if (desc->texcoords) hs.texcoords.assign(desc->texcoords, desc->texcoords + 2 * (size_t)nv);
    if (desc->tangents) hs.tangents.assign(desc->tangents, desc->tangents + 4 * (size_t)nv);
    VGI_CHECK(build_textures(desc, hs));
    hs.nodes = std::move(c4.out);
    hs.root = root;
    hs.levels = (uint32_t)c4.max_level + 1;  // levels of 4-wide nodes
    VGI_REQUIRE(3u * hs.levels <= (uint32_t)kStackMaxEntries,
                "vgi_scene_create: BVH of %u levels exceeds the traversal stack", hs.levels);
    // traversal reads both arrays through 32-bit buffer offsets
    if ((uint64_t)hs.nodes.size() * sizeof(BvhNode) > 0x7FFFFFFFull || (uint64_t)nt * sizeof(BvhTri) > 0x7FFFFFFFull)
        return set_error(VGI_E_UNSUPPORTED, "vgi_scene_create: %u triangles exceed the 2 GiB node/triangle arrays", nt);
    return VGI_OK;
}

}  // namespace vgi

extern "C" vgi_status vgi_scene_create(const vgi_scene_desc* desc, vgi_scene** out) {
    VGI_REQUIRE(desc && out, "vgi_scene_create: null argument");
    *out = nullptr;
    HostScene hs;
    VGI_CHECK(build_scene_host(desc, hs));
    const uint32_t nt = desc->num_triangles, nv = desc->num_vertices;
    vgi_scene* sc = new (std::nothrow) vgi_scene();
    if (!sc) return set_error(VGI_E_NOMEM, "vgi_scene_create: out of host memory");
    vgi_status st;
    if ((st = upload(sc->nodes, hs.nodes)) != VGI_OK || (st = upload(sc->tris, hs.tris)) != VGI_OK ||
        (st = upload(sc->shade, hs.shade)) != VGI_OK || (st = upload(sc->positions, hs.positions)) != VGI_OK ||
        (st = upload(sc->normals, hs.normals)) != VGI_OK || (st = upload(sc->materials, hs.materials)) != VGI_OK ||
        (st = upload(sc->texcoords, hs.texcoords)) != VGI_OK || (st = upload(sc->tangents, hs.tangents)) != VGI_OK ||
        (st = upload(sc->textures, hs.textures)) != VGI_OK || (st = upload(sc->texels, hs.texels)) != VGI_OK ||
        (st = upload(sc->texel_lut, hs.texel_lut)) != VGI_OK) {
        delete sc;
        return st;
    }
    sc->num_vertices = nv;
    sc->num_triangles = nt;
    sc->depth = hs.levels;
    sc->view.nodes = sc->nodes.as<BvhNode>();
    sc->view.tris = sc->tris.as<BvhTri>();
    sc->view.shade = sc->shade.as<TriShade>();
    sc->view.positions = sc->positions.as<float>();
    sc->view.normals = desc->normals ? sc->normals.as<float>() : nullptr;
    sc->view.materials = sc->materials.as<Material>();
    sc->view.texcoords = desc->texcoords ? sc->texcoords.as<float>() : nullptr;
    sc->view.tangents = desc->tangents ? sc->tangents.as<float>() : nullptr;
    sc->view.textures = sc->textures.as<TexDesc>();
    sc->view.texels = sc->texels.as<uint32_t>();
    sc->view.texel_lut = sc->texel_lut.as<float>();
    sc->view.num_nodes = (uint32_t)hs.nodes.size();
    sc->view.num_tris = nt;
    sc->view.root = hs.root;
    sc->view.stack_depth = 3u * sc->depth;  // at most 3 far children pushed per level
    sc->view.counters = nullptr;
    *out = sc;
    return VGI_OK;
}

extern "C" vgi_status vgi_scene_destroy(vgi_scene* s) {
    delete s;
    return VGI_OK;
}

extern "C" vgi_status vgi_scene_info(const vgi_scene* s, vgi_scene_stats* out) {
    VGI_REQUIRE(s && out, "vgi_scene_info: null argument");
    out->num_triangles = s->num_triangles;
    out->num_vertices = s->num_vertices;
    out->num_nodes = s->view.num_nodes;
    out->depth = s->depth;
    out->device_bytes = (uint64_t)(s->nodes.bytes + s->tris.bytes + s->shade.bytes + s->positions.bytes +
                                   s->normals.bytes + s->materials.bytes + s->texcoords.bytes + s->tangents.bytes +
                                   s->textures.bytes + s->texels.bytes + s->texel_lut.bytes);
    return VGI_OK;
}
