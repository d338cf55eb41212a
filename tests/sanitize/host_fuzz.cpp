// Host-code fuzz driver for the sanitizer builds (AddressSanitizer and
// UndefinedBehaviorSanitizer): the glTF importer (vgi_gltf.cpp) on every
// document named on the command line, and the scene builder (the host half of
// vgi_scene_create: validation, binned SAH, 4-wide collapse, quantised nodes)
// on each imported model and on seeded triangle soups, some of them invalid.
// Nothing here touches a GPU.  Exit status 0 unless an invariant fails; a
// sanitizer report aborts the process.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>

#include "vgi_bvh.h"
#include "vgi_internal.h"

using namespace vgi;

namespace {

int failures = 0;

void check(bool ok, const char* what, const char* where) {
    if (!ok) {
        std::fprintf(stderr, "host_fuzz: %s (%s)\n", what, where);
        ++failures;
    }
}

// Every triangle sits in exactly one leaf, every inner link is in range.
void check_scene(const HostScene& hs, uint32_t nt, const char* where) {
    std::vector<uint32_t> seen(nt, 0);
    std::vector<int32_t> stack{hs.root};
    size_t visited = 0;
    while (!stack.empty() && visited <= hs.nodes.size()) {
        const int32_t n = stack.back();
        stack.pop_back();
        check(n >= 0 && (size_t)n < hs.nodes.size(), "node link out of range", where);
        if (n < 0 || (size_t)n >= hs.nodes.size()) return;
        ++visited;
        for (int k = 0; k < 4; ++k) {
            const int32_t l = hs.nodes[n].link[k];
            if (l >= 0) {
                stack.push_back(l);
            } else if (l != ~0) {
                const uint32_t code = ~(uint32_t)l, first = code >> 4, count = code & 15u;
                check((size_t)first + count <= hs.tris.size(), "leaf range out of bounds", where);
                for (uint32_t t = first; t < first + count && t < hs.tris.size(); ++t) {
                    uint32_t id;
                    std::memcpy(&id, &hs.tris[t].v0.w, 4);
                    check(id < nt, "leaf triangle id out of range", where);
                    if (id < nt) ++seen[id];
                }
            }
        }
    }
    check(visited == hs.nodes.size(), "node reached twice or unreachable", where);
    for (uint32_t t = 0; t < nt; ++t) check(seen[t] == 1, "triangle not in exactly one leaf", where);
}

uint32_t rng_state = 12345u;
uint32_t next() {
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 17;
    rng_state ^= rng_state << 5;
    return rng_state;
}
float unit() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }

}  // namespace

int main(int argc, char** argv) {
    int docs = 0, imported = 0, scenes = 0, built = 0;
    for (int a = 1; a < argc; ++a) {
        ++docs;
        vgi_gltf_model* m = nullptr;
        if (vgi_gltf_import(argv[a], nullptr, &m) != VGI_OK) continue;
        ++imported;
        vgi_scene_desc desc{};
        check(vgi_gltf_model_desc(m, &desc) == VGI_OK, "model desc", argv[a]);
        HostScene hs;
        ++scenes;
        if (build_scene_host(&desc, hs) == VGI_OK) {
            ++built;
            check_scene(hs, desc.num_triangles, argv[a]);
        }
        vgi_gltf_destroy(m);
    }
    // triangle soups: clustered, degenerate, huge, non-finite, bad indices
    for (int s = 0; s < 60; ++s) {
        const uint32_t nt = 1 + next() % (s < 50 ? 3000u : 20u), nv = 3 * nt;
        std::vector<float> pos(3 * (size_t)nv), nrm(3 * (size_t)nv, 0.0f);
        std::vector<uint32_t> idx(3 * (size_t)nt), tm(nt, 0);
        const int kind = s % 6;
        for (uint32_t v = 0; v < nv; ++v)
            for (int k = 0; k < 3; ++k) {
                float x = (unit() - 0.5f) * 100.0f;
                if (kind == 1) x = 1.0f;                                        // every vertex equal
                if (kind == 2) x = (unit() + 1.0f) * 1e13f;                     // beyond 2^40
                if (kind == 3 && v == nv / 2) x = std::numeric_limits<float>::quiet_NaN();
                if (kind == 4) x = (float)(next() % 4) * 1e-30f;               // tiny, mostly coplanar
                pos[3 * v + k] = x;
            }
        for (uint32_t i = 0; i < 3 * nt; ++i) idx[i] = (kind == 5 && i == 0) ? nv : next() % nv;
        vgi_material mat{};
        mat.base_color[0] = mat.base_color[1] = mat.base_color[2] = mat.base_color[3] = 1.0f;
        mat.roughness = 0.5f;
        vgi_scene_desc desc{};
        desc.positions = pos.data();
        desc.normals = (s & 1) ? nrm.data() : nullptr;
        desc.num_vertices = nv;
        desc.indices = idx.data();
        desc.num_triangles = nt;
        desc.triangle_material = tm.data();
        desc.materials = &mat;
        desc.num_materials = 1;
        HostScene hs;
        ++scenes;
        const vgi_status st = build_scene_host(&desc, hs);
        const bool must_fail = kind == 2 || kind == 3 || kind == 5;
        check((st == VGI_OK) != must_fail, must_fail ? "invalid soup accepted" : "valid soup rejected", "soup");
        if (st == VGI_OK) {
            ++built;
            check_scene(hs, nt, "soup");
        }
    }
    std::printf("host_fuzz: %d documents (%d imported), %d scenes (%d built), %d failures\n", docs, imported, scenes,
                built, failures);
    return failures ? 1 : 0;
}
