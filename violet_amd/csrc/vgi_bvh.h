/*
 * vgi_bvh.h — software ray tracing on gfx950 (CDNA4 has no ray-tracing
 * hardware): device layout of the scene BVH and the traversal routines that
 * replace violet's VK_KHR_acceleration_structure + TraceRay
 * (restir/raytrace.inc.hlsl:31-56,67-82; render_scene.rs:1170-1712).
 *
 * Layout (built on the host by vgi_bvh.cpp, uploaded once per scene):
 *   nodes : 4-wide BVH, one 64-byte node (half an L2 line, four 16-byte
 *           loads) holds its four children's boxes quantised to 8 bits per
 *           plane on a per-node grid, plane = org + q * step with step a
 *           power of two per axis (stored as step * 2^24, the factor that
 *           turns the traversal's binary16 grid lines q * 2^-24 back into
 *           q), and four child links.  A child link >= 0
 *           is an inner node index; < 0 encodes a leaf ~((first << 4) | n).
 *           A child's quantised box contains its padded box (lo rounded
 *           down, hi rounded up).  An unused child has lo = 255, hi = 0 on
 *           every axis (an empty interval) and the empty leaf link ~0.
 *   tris  : leaf-ordered triangles as {v0, e1 = v1-v0, e2 = v2-v0} (3 x
 *           float4, 48 bytes); v0.w carries the scene triangle id.
 * Child boxes are padded by a small relative margin so the slab test is
 * conservative: a triangle the intersection routine accepts is never pruned
 * by its ancestors' boxes, and the closest hit (smallest t, ties to the
 * lowest triangle id) equals a brute-force loop over all triangles.
 *
 * Hit semantics follow violet: opaque, two-sided, t in [tmin, tmax];
 * shadow rays accept the first hit (RAY_FLAG_ACCEPT_FIRST_HIT_AND_END_SEARCH).
 */
#pragma once
#include <vector>

#include "../../include/vgi.h"
#include "vgi_common.h"

namespace vgi {

struct BvhNode {
    float org[3];    // grid origin (at or below every used child's padded box)
    float scale[3];  // grid step per axis times 2^24, a power of two (255 steps span the node)
    uint32_t q[6];   // planes lo.x, hi.x, lo.y, hi.y, lo.z, hi.z: byte k = child k
    int32_t link[4];
};
static_assert(sizeof(BvhNode) == 64, "BvhNode layout");

struct BvhTri {
    float4_ v0;  // w = triangle id (bits)
    float4_ e1;
    float4_ e2;
};
static_assert(sizeof(BvhTri) == 48, "BvhTri layout");

// Per-triangle shading data, indexed by scene triangle id.
struct TriShade {
    uint32_t i0, i1, i2;  // vertex indices
    uint32_t material;
};

// MaterialParams (scene_bindings.hlsl:24-51): the layout of vgi_material.
struct Material {
    float base_color[4];
    float metallic;
    float roughness;
    uint32_t color_metalrough_texture;  // base colour | metallic-roughness << 16
    uint32_t normal_texture;            // low 16 bits
};
static_assert(sizeof(Material) == sizeof(vgi_material), "Material layout");

// One material texture on the device: RGBA8 texels (one u32 each, R in the
// low byte) of every mip level, level l at texels + level_offset[l].
constexpr uint32_t kMaxTextureLevels = 16;
struct TexDesc {
    uint32_t width, height, levels, srgb;
    uint32_t level_offset[kMaxTextureLevels];
};
static_assert(sizeof(TexDesc) == 80, "TexDesc layout");
// Texel byte -> float: [0, 256) UNORM (c / 255), [256, 512) the sRGB
// transfer function's inverse, both evaluated in double on the host.
constexpr uint32_t kTexelLutEntries = 512;

struct SceneView {
    const BvhNode* nodes;
    const BvhTri* tris;
    const TriShade* shade;
    const float* positions;  // 3 floats per vertex
    const float* normals;    // 3 floats per vertex
    const Material* materials;
    const float* texcoords;  // 2 floats per vertex, or null
    const float* tangents;   // 4 floats per vertex, or null
    const TexDesc* textures; // the scene's textures (material index - 2)
    const uint32_t* texels;
    const float* texel_lut;  // kTexelLutEntries floats
    uint32_t num_nodes;
    uint32_t num_tris;
    int32_t root;           // link of the root (node 0)
    uint32_t stack_depth;   // most traversal-stack entries a ray can hold (3 per node level)
    unsigned long long* counters;  // traversal statistics slots (instrumented kernels only)
};

constexpr float kRayTMin = 0.0005f;  // raytrace.inc.hlsl:39,70
constexpr float kRayTMax = 1000.0f;  // raytrace.inc.hlsl:40,71
constexpr int kBvhMaxDepth = 64;     // binary build depth bound (the 4-wide tree is shallower)

// Traversal stack: the bottom kLdsStackEntries entries of each lane live in
// LDS, deeper ones in private (scratch) memory.  A one-wave workgroup with a
// 4 KiB stack leaves LDS for more resident waves than the registers allow,
// and the scratch part is touched only by rays whose stack grows past 16.
constexpr uint32_t kLdsStackEntries = 16;
constexpr uint32_t kStackMaxEntries = 3 * kBvhMaxDepth;  // 3 far children per 4-wide level

// Register budget of the per-pixel ray-tracing kernels, as waves per SIMD:
// the traversal hides node-fetch latency with resident waves, and 4 KiB of
// LDS stack per wave allows more of them than ~80 registers would.  (The
// cache-ray pass runs too few waves to gain from it and keeps its registers.)
constexpr int kTraceWaves = 8;

// The host side of vgi_scene_create: validation, the binned-SAH build, the
// 4-wide collapse and the device arrays, before any upload (so host tools
// can build and check a scene without a GPU).
struct HostScene {
    std::vector<BvhNode> nodes;
    std::vector<BvhTri> tris;
    std::vector<TriShade> shade;
    std::vector<float> positions, normals, texcoords, tangents;
    std::vector<Material> materials;
    std::vector<TexDesc> textures;
    std::vector<uint32_t> texels;
    std::vector<float> texel_lut;
    int32_t root = 0;
    uint32_t levels = 0;  // levels of 4-wide nodes
};
vgi_status build_scene_host(const vgi_scene_desc* desc, HostScene& out);

// Dynamic LDS bytes of a 64-lane traversal stack for this scene.
inline size_t traversal_lds_bytes(const SceneView& s) {
    const uint32_t e = s.stack_depth < kLdsStackEntries ? s.stack_depth : kLdsStackEntries;
    return (size_t)e * 64u * sizeof(int32_t);
}

struct Hit {
    float t;
    float u, v;  // barycentric weights of vertex 1 and vertex 2
    uint32_t tri;  // scene triangle id; 0xFFFFFFFF = miss
};

// Moller-Trumbore, two-sided.  Fixed evaluation order (bit-identical on the
// host and the device).
VGI_HD bool intersect_tri(f3 o, f3 d, f3 v0, f3 e1, f3 e2, float tmin, float& t, float& u, float& v) {
    f3 pvec = cross3(d, e2);
    float det = dot3(e1, pvec);
    if (det == 0.0f) return false;
    float inv_det = 1.0f / det;
    f3 tvec = o - v0;
    u = dot3(tvec, pvec) * inv_det;
    if (!(u >= 0.0f && u <= 1.0f)) return false;
    f3 qvec = cross3(tvec, e1);
    v = dot3(d, qvec) * inv_det;
    if (!(v >= 0.0f && u + v <= 1.0f)) return false;
    t = dot3(e2, qvec) * inv_det;
    return t >= tmin;
}

}  // namespace vgi
