/*
 * vgi_gltf.cpp — glTF 2.0 scene import for the scene/BVH builder.
 *
 * Follows violet's importer (crates/violet/src/model.rs:389-720,
 * import_gltf_uncached) for what reaches the ray tracer:
 *   - every glTF mesh is a geometry group of its primitives; a primitive is
 *     kept only in TRIANGLES mode with a material, indices and POSITION
 *     (model.rs:424-460); indices are read as u16, so u32 indices are
 *     truncated the way `i as u16` truncates them (model.rs:437-447);
 *   - the default scene (the document's `scene` property; none is an
 *     error, model.rs:600-602) is walked depth first from each root node,
 *     with the node's matrix (or T * R * S) applied after its parent's
 *     (model.rs:563-598, :172-180); every node with a mesh is an instance;
 *   - the root transform turns glTF's Y-up into violet's Z-up: rotation of
 *     +90 degrees about X times the configured uniform scale (:590-594);
 *   - materials keep the PBR factors (base colour, metallic, roughness;
 *     :700-709) and their base-colour, metallic-roughness and normal
 *     textures (:640-702; texture coordinate set 0 only);
 *   - TEXCOORD_0 and TANGENT are read; a primitive with texture coordinates
 *     and normals but no TANGENT gets generated tangents (violet runs
 *     MikkTSpace, :507-531; here per-vertex sums of the triangles' dP/du
 *     made orthogonal to the normal, E. Lengyel 2001, without MikkTSpace's
 *     vertex splitting);
 *   - every image a material uses is decoded once (PNG, baseline JPEG;
 *     vgi_image.h),
 *     numbered in image order, marked with the uses it has: an image used
 *     only as a metallic-roughness map has its B and G moved to R and G
 *     (:811-846), one used only as a base colour is sRGB (:936-941); its mip
 *     chain is violet's Lanczos3 halving (:853-870).  An image in another
 *     format (progressive JPEG, ...) is counted in stats.images_unsupported
 *     and its slots keep violet's defaults (white, flat normal).
 * Instances are baked to world space, the form vgi_scene_create takes: a
 * position goes through the instance's affine transform, a normal through
 * the inverse transpose of its 3x3 part without normalisation (the hit
 * shader interpolates the local normal, transforms it and then normalises,
 * geometry_rchit.inc.hlsl:95-100, which the baked normal reproduces because
 * the transform is linear), and a mirroring transform (negative
 * determinant) reverses the winding so that the baked geometric normal
 * keeps the direction the hit shader derives (:108-112).  A kept primitive
 * without NORMAL gets one vertex per triangle corner carrying the
 * geometric normal.  Texture coordinates are copied and tangents go
 * through the normal transform with their sign w multiplied by the sign of
 * the determinant (the bitangent the hit shader forms in local space and
 * then transforms).
 *
 * Sources: a .gltf JSON file with its buffers as relative files or base64
 * data URIs, or a .glb container (JSON chunk + BIN chunk).
 */
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "vgi_image.h"
#include "vgi_internal.h"

using namespace vgi;

namespace {

// ---------------------------------------------------------------------------
// JSON (RFC 8259) document tree
// ---------------------------------------------------------------------------
struct JVal {
    enum Kind { kNull, kBool, kNum, kStr, kArr, kObj } kind = kNull;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;

    const JVal* get(const char* key) const {
        if (kind != kObj) return nullptr;
        for (auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    bool is_num() const { return kind == kNum; }
};

struct JsonParser {
    const char* p;
    const char* end;
    std::string err;
    int depth = 0;

    void ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool fail(const char* what) {
        if (err.empty()) err = what;
        return false;
    }
    static void put_utf8(std::string& s, uint32_t cp) {
        if (cp < 0x80) {
            s += (char)cp;
        } else if (cp < 0x800) {
            s += (char)(0xC0 | (cp >> 6));
            s += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            s += (char)(0xE0 | (cp >> 12));
            s += (char)(0x80 | ((cp >> 6) & 0x3F));
            s += (char)(0x80 | (cp & 0x3F));
        } else {
            s += (char)(0xF0 | (cp >> 18));
            s += (char)(0x80 | ((cp >> 12) & 0x3F));
            s += (char)(0x80 | ((cp >> 6) & 0x3F));
            s += (char)(0x80 | (cp & 0x3F));
        }
    }
    bool hex4(uint32_t& v) {
        if (end - p < 4) return fail("truncated \\u escape");
        v = 0;
        for (int k = 0; k < 4; ++k) {
            char c = *p++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else return fail("bad \\u escape");
        }
        return true;
    }
    bool string(std::string& out) {
        ++p;  // opening quote
        while (p < end && *p != '"') {
            char c = *p++;
            if ((unsigned char)c < 0x20) return fail("control character in string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p >= end) return fail("truncated escape");
            char e = *p++;
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    uint32_t cp;
                    if (!hex4(cp)) return false;
                    if (cp >= 0xD800 && cp < 0xDC00) {  // surrogate pair
                        uint32_t lo;
                        if (end - p < 2 || p[0] != '\\' || p[1] != 'u') return fail("unpaired surrogate");
                        p += 2;
                        if (!hex4(lo) || lo < 0xDC00 || lo >= 0xE000) return fail("bad surrogate pair");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    put_utf8(out, cp);
                    break;
                }
                default: return fail("bad escape");
            }
        }
        if (p >= end) return fail("unterminated string");
        ++p;
        return true;
    }
    bool number(double& v) {
        const char* s = p;
        if (p < end && *p == '-') ++p;
        if (p >= end || !(*p >= '0' && *p <= '9')) return fail("bad number");
        while (p < end && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' || *p == '+' || *p == '-'))
            ++p;
        std::string tok(s, p);
        char* q = nullptr;
        v = strtod(tok.c_str(), &q);
        if (!q || *q != '\0') return fail("bad number");
        return true;
    }
    bool literal(const char* w) {
        size_t n = strlen(w);
        if ((size_t)(end - p) < n || strncmp(p, w, n) != 0) return fail("bad literal");
        p += n;
        return true;
    }
    bool value(JVal& v) {
        if (++depth > 256) return fail("nesting too deep");
        ws();
        if (p >= end) return fail("unexpected end");
        bool ok = true;
        char c = *p;
        if (c == '{') {
            v.kind = JVal::kObj;
            ++p;
            ws();
            if (p < end && *p == '}') {
                ++p;
            } else {
                for (;;) {
                    ws();
                    if (p >= end || *p != '"') {
                        ok = fail("expected key");
                        break;
                    }
                    std::pair<std::string, JVal> kv;
                    if (!string(kv.first)) {
                        ok = false;
                        break;
                    }
                    ws();
                    if (p >= end || *p != ':') {
                        ok = fail("expected ':'");
                        break;
                    }
                    ++p;
                    if (!value(kv.second)) {
                        ok = false;
                        break;
                    }
                    v.obj.push_back(std::move(kv));
                    ws();
                    if (p < end && *p == ',') {
                        ++p;
                        continue;
                    }
                    if (p < end && *p == '}') {
                        ++p;
                        break;
                    }
                    ok = fail("expected ',' or '}'");
                    break;
                }
            }
        } else if (c == '[') {
            v.kind = JVal::kArr;
            ++p;
            ws();
            if (p < end && *p == ']') {
                ++p;
            } else {
                for (;;) {
                    JVal e;
                    if (!value(e)) {
                        ok = false;
                        break;
                    }
                    v.arr.push_back(std::move(e));
                    ws();
                    if (p < end && *p == ',') {
                        ++p;
                        continue;
                    }
                    if (p < end && *p == ']') {
                        ++p;
                        break;
                    }
                    ok = fail("expected ',' or ']'");
                    break;
                }
            }
        } else if (c == '"') {
            v.kind = JVal::kStr;
            ok = string(v.str);
        } else if (c == 't') {
            v.kind = JVal::kBool;
            v.b = true;
            ok = literal("true");
        } else if (c == 'f') {
            v.kind = JVal::kBool;
            ok = literal("false");
        } else if (c == 'n') {
            ok = literal("null");
        } else {
            v.kind = JVal::kNum;
            ok = number(v.num);
        }
        --depth;
        return ok;
    }
};

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
bool read_file(const std::string& path, std::vector<uint8_t>& out) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    bool ok = fseek(f, 0, SEEK_END) == 0;
    long n = ok ? ftell(f) : -1;
    ok = ok && n >= 0 && fseek(f, 0, SEEK_SET) == 0;
    if (ok) {
        out.resize((size_t)n);
        ok = n == 0 || fread(out.data(), 1, (size_t)n, f) == (size_t)n;
    }
    fclose(f);
    return ok;
}

bool base64_decode(const char* s, size_t n, std::vector<uint8_t>& out) {
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+') return 62;
        if (c == '/') return 63;
        return -1;
    };
    uint32_t acc = 0;
    int bits = 0;
    for (size_t i = 0; i < n; ++i) {
        char c = s[i];
        if (c == '=') break;
        int v = val(c);
        if (v < 0) return false;
        acc = (acc << 6) | (uint32_t)v;
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back((uint8_t)((acc >> bits) & 0xFF));
        }
    }
    return true;
}

bool as_u32(const JVal* v, uint32_t& out) {
    if (!v || !v->is_num() || v->num < 0 || v->num > 4294967295.0 || v->num != std::floor(v->num)) return false;
    out = (uint32_t)v->num;
    return true;
}

// Column-major 4x4, element (row r, column c) at m[c * 4 + r] (glam's layout).
struct Mat4 {
    float m[16];
    static Mat4 identity() {
        Mat4 a;
        for (int k = 0; k < 16; ++k) a.m[k] = (k % 5 == 0) ? 1.0f : 0.0f;
        return a;
    }
    Mat4 operator*(const Mat4& b) const {
        Mat4 r;
        for (int c = 0; c < 4; ++c)
            for (int row = 0; row < 4; ++row) {
                float s = 0.0f;
                for (int k = 0; k < 4; ++k) s += m[k * 4 + row] * b.m[c * 4 + k];
                r.m[c * 4 + row] = s;
            }
        return r;
    }
};

// T * R * S of a glTF node (rotation quaternion x, y, z, w).
Mat4 trs(const float t[3], const float q[4], const float s[3]) {
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z, wx = w * x, wy = w * y,
                wz = w * z;
    Mat4 a = Mat4::identity();
    // columns of the rotation, each scaled by the matching scale component
    a.m[0] = (1.0f - 2.0f * (yy + zz)) * s[0];
    a.m[1] = (2.0f * (xy + wz)) * s[0];
    a.m[2] = (2.0f * (xz - wy)) * s[0];
    a.m[4] = (2.0f * (xy - wz)) * s[1];
    a.m[5] = (1.0f - 2.0f * (xx + zz)) * s[1];
    a.m[6] = (2.0f * (yz + wx)) * s[1];
    a.m[8] = (2.0f * (xz + wy)) * s[2];
    a.m[9] = (2.0f * (yz - wx)) * s[2];
    a.m[10] = (1.0f - 2.0f * (xx + yy)) * s[2];
    a.m[12] = t[0];
    a.m[13] = t[1];
    a.m[14] = t[2];
    return a;
}

}  // namespace

// ---------------------------------------------------------------------------
// the model
// ---------------------------------------------------------------------------
struct vgi_gltf_model {
    std::vector<float> positions, normals, texcoords, tangents;
    bool has_texcoords = false, has_tangents = false;
    std::vector<uint32_t> indices, tri_material;
    std::vector<vgi_material> materials;
    std::vector<std::vector<uint8_t>> texels;  // per texture: every mip level back to back
    std::vector<vgi_texture> textures;
    vgi_gltf_stats stats{};
};

namespace {

constexpr uint32_t kMaxInstances = 1u << 20;
constexpr uint32_t kMaxTriangles = 0x7FFFFFFFu / 48u;
constexpr uint64_t kMaxVertices = 3ull * kMaxTriangles;  // baked vertices (indices are u32)

struct Importer {
    JVal doc;
    std::vector<std::vector<uint8_t>> buffers;
    std::string dir;
    const uint8_t* glb_bin = nullptr;
    size_t glb_bin_len = 0;
    std::vector<uint8_t> file;

    struct Accessor {
        const uint8_t* base = nullptr;
        size_t count = 0, stride = 0;
        uint32_t comp_type = 0, ncomp = 0;
        bool normalized = false;
    };
    struct Prim {
        uint32_t material = 0;
        std::vector<uint32_t> idx;  // u16 values
        std::vector<float> pos, nrm, uv, tan;
        bool has_normals = false;
    };
    // material texture references, as image indices (-1: none)
    struct MatImages {
        int64_t base_color = -1, metal_rough = -1, normal = -1;
    };
    std::vector<MatImages> mat_images;
    enum : uint32_t { kUseBaseColor = 1, kUseMetalRough = 2, kUseNormal = 4 };
    std::vector<std::vector<Prim>> meshes;

    static uint32_t comp_size(uint32_t t) {
        switch (t) {
            case 5120: case 5121: return 1;
            case 5122: case 5123: return 2;
            case 5125: case 5126: return 4;
            default: return 0;
        }
    }
    static uint32_t type_comps(const std::string& t) {
        if (t == "SCALAR") return 1;
        if (t == "VEC2") return 2;
        if (t == "VEC3") return 3;
        if (t == "VEC4") return 4;
        if (t == "MAT4") return 16;
        return 0;
    }

    vgi_status load(const char* path) {
        std::string ps(path);
        size_t slash = ps.find_last_of('/');
        dir = slash == std::string::npos ? std::string() : ps.substr(0, slash + 1);
        if (!read_file(ps, file)) return set_error(VGI_E_INVALID, "vgi_gltf_import: cannot read '%s'", path);
        const char* json = (const char*)file.data();
        size_t json_len = file.size();
        if (file.size() >= 12 && memcmp(file.data(), "glTF", 4) == 0) {
            uint32_t hdr[3];
            memcpy(hdr, file.data(), 12);
            if (hdr[1] != 2) return set_error(VGI_E_UNSUPPORTED, "vgi_gltf_import: GLB version %u (2 expected)", hdr[1]);
            if (hdr[2] > file.size()) return set_error(VGI_E_INVALID, "vgi_gltf_import: GLB length exceeds the file");
            size_t off = 12;
            json = nullptr;
            while (off + 8 <= hdr[2]) {
                uint32_t ch[2];
                memcpy(ch, file.data() + off, 8);
                off += 8;
                if (ch[0] > hdr[2] - off) return set_error(VGI_E_INVALID, "vgi_gltf_import: GLB chunk exceeds the file");
                if (ch[1] == 0x4E4F534Au && !json) {  // "JSON"
                    json = (const char*)file.data() + off;
                    json_len = ch[0];
                } else if (ch[1] == 0x004E4942u && !glb_bin) {  // "BIN\0"
                    glb_bin = file.data() + off;
                    glb_bin_len = ch[0];
                }
                off += (ch[0] + 3u) & ~3u;
            }
            if (!json) return set_error(VGI_E_INVALID, "vgi_gltf_import: GLB without a JSON chunk");
        }
        JsonParser jp{json, json + json_len, {}};
        if (!jp.value(doc)) return set_error(VGI_E_INVALID, "vgi_gltf_import: JSON: %s", jp.err.c_str());
        jp.ws();
        if (jp.p != jp.end) return set_error(VGI_E_INVALID, "vgi_gltf_import: JSON: trailing characters");
        if (doc.kind != JVal::kObj) return set_error(VGI_E_INVALID, "vgi_gltf_import: JSON root is not an object");
        const JVal* asset = doc.get("asset");
        const JVal* ver = asset ? asset->get("version") : nullptr;
        if (!ver || ver->kind != JVal::kStr || ver->str.compare(0, 2, "2.") != 0)
            return set_error(VGI_E_UNSUPPORTED, "vgi_gltf_import: asset.version 2.x required");
        return load_buffers();
    }

    // The bytes a uri names: a base64 data URI or a file relative to the
    // document (percent-decoded; absolute paths and schemes are refused).
    vgi_status uri_bytes(const std::string& u, std::vector<uint8_t>& data, const char* what, uint32_t k) {
        if (u.compare(0, 5, "data:") == 0) {
            size_t comma = u.find(',');
            if (comma == std::string::npos || u.rfind(";base64", comma) == std::string::npos)
                return set_error(VGI_E_UNSUPPORTED, "vgi_gltf_import: %s %u: only base64 data URIs", what, k);
            if (!base64_decode(u.data() + comma + 1, u.size() - comma - 1, data))
                return set_error(VGI_E_INVALID, "vgi_gltf_import: %s %u: bad base64", what, k);
            return VGI_OK;
        }
        if (u.find("://") != std::string::npos || (!u.empty() && u[0] == '/'))
            return set_error(VGI_E_UNSUPPORTED, "vgi_gltf_import: %s %u: only relative file URIs", what, k);
        std::string path;
        for (size_t i = 0; i < u.size(); ++i) {  // percent-decoding of the relative path
            if (u[i] == '%' && i + 2 < u.size()) {
                char hx[3] = {u[i + 1], u[i + 2], 0};
                path += (char)strtol(hx, nullptr, 16);
                i += 2;
            } else {
                path += u[i];
            }
        }
        if (!read_file(dir + path, data))
            return set_error(VGI_E_INVALID, "vgi_gltf_import: cannot read %s '%s'", what, (dir + path).c_str());
        return VGI_OK;
    }

    vgi_status load_buffers() {
        const JVal* bs = doc.get("buffers");
        if (!bs) return VGI_OK;
        if (bs->kind != JVal::kArr) return set_error(VGI_E_INVALID, "vgi_gltf_import: buffers is not an array");
        for (size_t k = 0; k < bs->arr.size(); ++k) {
            const JVal& b = bs->arr[k];
            uint32_t len = 0;
            if (!as_u32(b.get("byteLength"), len)) return set_error(VGI_E_INVALID, "vgi_gltf_import: buffer %zu: byteLength", k);
            std::vector<uint8_t> data;
            const JVal* uri = b.get("uri");
            if (!uri) {
                if (k != 0 || !glb_bin) return set_error(VGI_E_INVALID, "vgi_gltf_import: buffer %zu has no uri", k);
                data.assign(glb_bin, glb_bin + glb_bin_len);
            } else {
                if (uri->kind != JVal::kStr) return set_error(VGI_E_INVALID, "vgi_gltf_import: buffer %zu: uri", k);
                vgi_status s = uri_bytes(uri->str, data, "buffer", (uint32_t)k);
                if (s != VGI_OK) return s;
            }
            if (data.size() < len) return set_error(VGI_E_INVALID, "vgi_gltf_import: buffer %zu shorter than byteLength", k);
            data.resize(len);
            buffers.push_back(std::move(data));
        }
        return VGI_OK;
    }

    vgi_status accessor(uint32_t index, Accessor& a) {
        const JVal* accs = doc.get("accessors");
        if (!accs || accs->kind != JVal::kArr || index >= accs->arr.size())
            return set_error(VGI_E_INVALID, "vgi_gltf_import: accessor %u does not exist", index);
        const JVal& ac = accs->arr[index];
        uint32_t count = 0, ct = 0, bv = 0, off = 0;
        const JVal* type = ac.get("type");
        if (!as_u32(ac.get("count"), count) || !as_u32(ac.get("componentType"), ct) || !type ||
            type->kind != JVal::kStr)
            return set_error(VGI_E_INVALID, "vgi_gltf_import: accessor %u: count/componentType/type", index);
        if (ac.get("sparse")) return set_error(VGI_E_UNSUPPORTED, "vgi_gltf_import: accessor %u is sparse", index);
        if (!as_u32(ac.get("bufferView"), bv))
            return set_error(VGI_E_UNSUPPORTED, "vgi_gltf_import: accessor %u has no bufferView", index);
        if (ac.get("byteOffset") && !as_u32(ac.get("byteOffset"), off))
            return set_error(VGI_E_INVALID, "vgi_gltf_import: accessor %u: byteOffset", index);
        a.comp_type = ct;
        a.ncomp = type_comps(type->str);
        const uint32_t cs = comp_size(ct);
        if (!cs || !a.ncomp) return set_error(VGI_E_INVALID, "vgi_gltf_import: accessor %u: bad type", index);
        const JVal* nz = ac.get("normalized");
        a.normalized = nz && nz->kind == JVal::kBool && nz->b;
        const JVal* views = doc.get("bufferViews");
        if (!views || views->kind != JVal::kArr || bv >= views->arr.size())
            return set_error(VGI_E_INVALID, "vgi_gltf_import: accessor %u: bufferView %u does not exist", index, bv);
        const JVal& v = views->arr[bv];
        uint32_t buf = 0, voff = 0, vlen = 0, stride = 0;
        if (!as_u32(v.get("buffer"), buf) || !as_u32(v.get("byteLength"), vlen) || buf >= buffers.size())
            return set_error(VGI_E_INVALID, "vgi_gltf_import: bufferView %u: buffer/byteLength", bv);
        if (v.get("byteOffset") && !as_u32(v.get("byteOffset"), voff))
            return set_error(VGI_E_INVALID, "vgi_gltf_import: bufferView %u: byteOffset", bv);
        if (v.get("byteStride") && !as_u32(v.get("byteStride"), stride))
            return set_error(VGI_E_INVALID, "vgi_gltf_import: bufferView %u: byteStride", bv);
        const size_t elem = (size_t)cs * a.ncomp;
        a.stride = stride ? stride : elem;
        a.count = count;
        if ((uint64_t)voff + vlen > buffers[buf].size())
            return set_error(VGI_E_INVALID, "vgi_gltf_import: bufferView %u exceeds its buffer", bv);
        if (count && (uint64_t)off + (uint64_t)a.stride * (count - 1) + elem > vlen)
            return set_error(VGI_E_INVALID, "vgi_gltf_import: accessor %u exceeds its bufferView", index);
        a.base = buffers[buf].data() + voff + off;
        return VGI_OK;
    }

    // Component c of element i as a float (normalised integers mapped as glTF specifies).
    static float read_f(const Accessor& a, size_t i, uint32_t c) {
        const uint8_t* q = a.base + i * a.stride;
        switch (a.comp_type) {
            case 5126: { float v; memcpy(&v, q + 4 * c, 4); return v; }
            case 5121: { float v = q[c]; return a.normalized ? v / 255.0f : v; }
            case 5120: { float v = (int8_t)q[c]; return a.normalized ? std::fmax(v / 127.0f, -1.0f) : v; }
            case 5123: { uint16_t u; memcpy(&u, q + 2 * c, 2); return a.normalized ? u / 65535.0f : (float)u; }
            case 5122: { int16_t s; memcpy(&s, q + 2 * c, 2); return a.normalized ? std::fmax(s / 32767.0f, -1.0f) : (float)s; }
            default: { uint32_t u; memcpy(&u, q + 4 * c, 4); return (float)u; }
        }
    }
    static uint32_t read_u(const Accessor& a, size_t i) {
        const uint8_t* q = a.base + i * a.stride;
        switch (a.comp_type) {
            case 5121: return q[0];
            case 5123: { uint16_t u; memcpy(&u, q, 2); return u; }
            default: { uint32_t u; memcpy(&u, q, 4); return u; }
        }
    }

    vgi_status load_meshes(vgi_gltf_stats& st) {
        const JVal* ms = doc.get("meshes");
        if (!ms) return VGI_OK;
        if (ms->kind != JVal::kArr) return set_error(VGI_E_INVALID, "vgi_gltf_import: meshes is not an array");
        const JVal* mats = doc.get("materials");
        const size_t nmat = mats && mats->kind == JVal::kArr ? mats->arr.size() : 0;
        for (size_t mi = 0; mi < ms->arr.size(); ++mi) {
            const JVal* prims = ms->arr[mi].get("primitives");
            if (!prims || prims->kind != JVal::kArr)
                return set_error(VGI_E_INVALID, "vgi_gltf_import: mesh %zu has no primitives", mi);
            std::vector<Prim> group;
            for (size_t pi = 0; pi < prims->arr.size(); ++pi) {
                const JVal& pr = prims->arr[pi];
                uint32_t mode = 4, mat = 0, iacc = 0, pacc = 0, nacc = 0;
                if (pr.get("mode") && !as_u32(pr.get("mode"), mode))
                    return set_error(VGI_E_INVALID, "vgi_gltf_import: mesh %zu primitive %zu: mode", mi, pi);
                const JVal* attrs = pr.get("attributes");
                const bool has_mat = as_u32(pr.get("material"), mat);
                const bool has_idx = as_u32(pr.get("indices"), iacc);
                const bool has_pos = attrs && as_u32(attrs->get("POSITION"), pacc);
                if (mode != 4 || !has_mat || !has_idx || !has_pos) {  // model.rs:426-460: skipped with a warning
                    ++st.primitives_skipped;
                    continue;
                }
                if (mat >= nmat)
                    return set_error(VGI_E_INVALID, "vgi_gltf_import: mesh %zu primitive %zu: material %u", mi, pi, mat);
                Accessor ia, pa, na;
                vgi_status s = accessor(iacc, ia);
                if (s == VGI_OK) s = accessor(pacc, pa);
                if (s != VGI_OK) return s;
                if (ia.ncomp != 1 || !(ia.comp_type == 5121 || ia.comp_type == 5123 || ia.comp_type == 5125) ||
                    ia.count % 3 != 0)
                    return set_error(VGI_E_INVALID, "vgi_gltf_import: mesh %zu primitive %zu: indices", mi, pi);
                if (pa.ncomp != 3 || pa.comp_type != 5126)
                    return set_error(VGI_E_INVALID, "vgi_gltf_import: mesh %zu primitive %zu: POSITION", mi, pi);
                Prim p;
                p.material = mat;
                p.idx.resize(ia.count);
                for (size_t k = 0; k < ia.count; ++k) {
                    const uint32_t u = read_u(ia, k);
                    if (u > 0xFFFFu) ++st.truncated_indices;
                    p.idx[k] = u & 0xFFFFu;  // read as u16 (model.rs:437-447)
                    if (p.idx[k] >= pa.count)
                        return set_error(VGI_E_INVALID, "vgi_gltf_import: mesh %zu primitive %zu: index %u >= %zu vertices",
                                         mi, pi, p.idx[k], pa.count);
                }
                p.pos.resize(3 * pa.count);
                for (size_t k = 0; k < pa.count; ++k)
                    for (uint32_t c = 0; c < 3; ++c) p.pos[3 * k + c] = read_f(pa, k, c);
                if (as_u32(attrs->get("NORMAL"), nacc)) {
                    if ((s = accessor(nacc, na)) != VGI_OK) return s;
                    if (na.ncomp != 3 || na.count != pa.count)
                        return set_error(VGI_E_INVALID, "vgi_gltf_import: mesh %zu primitive %zu: NORMAL", mi, pi);
                    p.nrm.resize(3 * na.count);
                    for (size_t k = 0; k < na.count; ++k)
                        for (uint32_t c = 0; c < 3; ++c) p.nrm[3 * k + c] = read_f(na, k, c);
                    p.has_normals = true;
                }
                uint32_t tacc = 0, gacc = 0;
                Accessor ta, ga;
                if (as_u32(attrs->get("TEXCOORD_0"), tacc)) {
                    if ((s = accessor(tacc, ta)) != VGI_OK) return s;
                    if (ta.ncomp != 2 || ta.count != pa.count ||
                        !(ta.comp_type == 5126 || (ta.normalized && (ta.comp_type == 5121 || ta.comp_type == 5123))))
                        return set_error(VGI_E_INVALID, "vgi_gltf_import: mesh %zu primitive %zu: TEXCOORD_0", mi, pi);
                    p.uv.resize(2 * ta.count);
                    for (size_t k = 0; k < ta.count; ++k)
                        for (uint32_t c = 0; c < 2; ++c) p.uv[2 * k + c] = read_f(ta, k, c);
                }
                if (as_u32(attrs->get("TANGENT"), gacc)) {
                    if ((s = accessor(gacc, ga)) != VGI_OK) return s;
                    if (ga.ncomp != 4 || ga.count != pa.count || ga.comp_type != 5126)
                        return set_error(VGI_E_INVALID, "vgi_gltf_import: mesh %zu primitive %zu: TANGENT", mi, pi);
                    p.tan.resize(4 * ga.count);
                    for (size_t k = 0; k < ga.count; ++k)
                        for (uint32_t c = 0; c < 4; ++c) p.tan[4 * k + c] = read_f(ga, k, c);
                } else if (!p.uv.empty() && p.has_normals) {
                    generate_tangents(p);
                }
                ++st.primitives;
                group.push_back(std::move(p));
            }
            meshes.push_back(std::move(group));
        }
        st.meshes = (uint32_t)meshes.size();
        return VGI_OK;
    }

    vgi_status load_materials(vgi_gltf_model& m) {
        const JVal* mats = doc.get("materials");
        if (!mats) return VGI_OK;
        if (mats->kind != JVal::kArr) return set_error(VGI_E_INVALID, "vgi_gltf_import: materials is not an array");
        for (const JVal& mt : mats->arr) {
            vgi_material o{};
            o.base_color[0] = o.base_color[1] = o.base_color[2] = o.base_color[3] = 1.0f;
            o.metallic = 1.0f;  // glTF defaults of the PBR factors
            o.roughness = 1.0f;
            bool textured = mt.get("normalTexture") || mt.get("occlusionTexture") || mt.get("emissiveTexture");
            MatImages mi;
            vgi_status s = texture_image(mt.get("normalTexture"), mi.normal);
            if (const JVal* pbr = mt.get("pbrMetallicRoughness")) {
                const JVal* bc = pbr->get("baseColorFactor");
                if (bc && bc->kind == JVal::kArr && bc->arr.size() == 4)
                    for (int c = 0; c < 4; ++c)
                        if (bc->arr[c].is_num()) o.base_color[c] = (float)bc->arr[c].num;
                const JVal* mf = pbr->get("metallicFactor");
                if (mf && mf->is_num()) o.metallic = (float)mf->num;
                const JVal* rf = pbr->get("roughnessFactor");
                if (rf && rf->is_num()) o.roughness = (float)rf->num;
                textured = textured || pbr->get("baseColorTexture") || pbr->get("metallicRoughnessTexture");
                if (s == VGI_OK) s = texture_image(pbr->get("baseColorTexture"), mi.base_color);
                if (s == VGI_OK) s = texture_image(pbr->get("metallicRoughnessTexture"), mi.metal_rough);
            }
            if (s != VGI_OK) return s;
            if (textured) ++m.stats.textured_materials;
            o.color_metalrough_texture = 0;
            o.normal_texture = VGI_TEXTURE_NORMAL_UP;
            m.materials.push_back(o);
            mat_images.push_back(mi);
        }
        m.stats.materials = (uint32_t)m.materials.size();
        return VGI_OK;
    }

    // textureInfo -> the image index of its texture's source (-1 if absent).
    vgi_status texture_image(const JVal* info, int64_t& image) {
        image = -1;
        if (!info) return VGI_OK;
        uint32_t t = 0, src = 0;
        const JVal* texs = doc.get("textures");
        if (!as_u32(info->get("index"), t) || !texs || texs->kind != JVal::kArr || t >= texs->arr.size())
            return set_error(VGI_E_INVALID, "vgi_gltf_import: material texture index");
        const JVal* imgs = doc.get("images");
        if (!as_u32(texs->arr[t].get("source"), src) || !imgs || imgs->kind != JVal::kArr || src >= imgs->arr.size())
            return set_error(VGI_E_INVALID, "vgi_gltf_import: texture %u has no image source", t);
        image = src;
        return VGI_OK;
    }

    // The encoded bytes of image k (a data URI, a relative file or a bufferView).
    vgi_status image_bytes(uint32_t k, std::vector<uint8_t>& out) {
        const JVal& im = doc.get("images")->arr[k];
        uint32_t bv = 0;
        if (const JVal* uri = im.get("uri")) {
            if (uri->kind != JVal::kStr) return set_error(VGI_E_INVALID, "vgi_gltf_import: image %u: uri", k);
            return uri_bytes(uri->str, out, "image", k);
        }
        if (!as_u32(im.get("bufferView"), bv)) return set_error(VGI_E_INVALID, "vgi_gltf_import: image %u has no source", k);
        const JVal* views = doc.get("bufferViews");
        uint32_t buf = 0, off = 0, len = 0;
        if (!views || views->kind != JVal::kArr || bv >= views->arr.size() || !as_u32(views->arr[bv].get("buffer"), buf) ||
            !as_u32(views->arr[bv].get("byteLength"), len) || buf >= buffers.size() ||
            (views->arr[bv].get("byteOffset") && !as_u32(views->arr[bv].get("byteOffset"), off)) ||
            (uint64_t)off + len > buffers[buf].size())
            return set_error(VGI_E_INVALID, "vgi_gltf_import: image %u: bufferView %u", k, bv);
        out.assign(buffers[buf].begin() + off, buffers[buf].begin() + off + len);
        return VGI_OK;
    }

    // Decodes every image the materials use and sets their texture indices
    // (model.rs:743-790: used images renumbered in image order).
    vgi_status load_images(vgi_gltf_model& m) {
        const JVal* imgs = doc.get("images");
        const size_t nimg = imgs && imgs->kind == JVal::kArr ? imgs->arr.size() : 0;
        std::vector<uint32_t> use(nimg, 0);
        for (const MatImages& mi : mat_images) {
            if (mi.base_color >= 0) use[mi.base_color] |= kUseBaseColor;
            if (mi.metal_rough >= 0) use[mi.metal_rough] |= kUseMetalRough;
            if (mi.normal >= 0) use[mi.normal] |= kUseNormal;
        }
        std::vector<int64_t> tex_of(nimg, -1);  // texture index, -1: unused or undecodable
        for (uint32_t k = 0; k < nimg; ++k) {
            if (!use[k]) continue;
            ++m.stats.images;
            std::vector<uint8_t> bytes;
            vgi_status s = image_bytes(k, bytes);
            if (s != VGI_OK) return s;
            Rgba8Image img;
            std::string err;
            if (is_png(bytes.data(), bytes.size())) {
                if (!decode_png(bytes.data(), bytes.size(), img, err))
                    return set_error(VGI_E_INVALID, "vgi_gltf_import: image %u: %s", k, err.c_str());
            } else if (is_jpeg(bytes.data(), bytes.size())) {
                const int r = decode_jpeg(bytes.data(), bytes.size(), img, err);
                if (r < 0) return set_error(VGI_E_INVALID, "vgi_gltf_import: image %u: %s", k, err.c_str());
                if (r == 0) {  // a JPEG form this decoder does not read (progressive, ...)
                    ++m.stats.images_unsupported;
                    continue;
                }
            } else {
                ++m.stats.images_unsupported;
                continue;
            }
            if (img.width > 32768u || img.height > 32768u)
                return set_error(VGI_E_UNSUPPORTED, "vgi_gltf_import: image %u is %ux%u (at most 32768 a side)", k,
                                 img.width, img.height);
            if (use[k] == kUseMetalRough)  // R: metallic (glTF B), G: roughness (glTF G)
                for (size_t i = 0; i < img.texels.size(); i += 4) std::swap(img.texels[i], img.texels[i + 2]);
            const std::vector<Rgba8Image> mips = mip_chain(img);
            std::vector<uint8_t> all;
            for (const Rgba8Image& l : mips) all.insert(all.end(), l.texels.begin(), l.texels.end());
            tex_of[k] = (int64_t)m.textures.size();
            m.texels.push_back(std::move(all));
            vgi_texture t{};
            t.width = img.width;
            t.height = img.height;
            t.levels = (uint32_t)mips.size();
            t.srgb = use[k] == kUseBaseColor ? 1u : 0u;
            m.textures.push_back(t);
        }
        for (size_t t = 0; t < m.textures.size(); ++t) m.textures[t].texels = m.texels[t].data();
        auto slot = [&](int64_t img, uint32_t dflt) {
            return img >= 0 && tex_of[img] >= 0 ? VGI_TEXTURE_FIRST + (uint32_t)tex_of[img] : dflt;
        };
        for (size_t i = 0; i < mat_images.size(); ++i) {
            const MatImages& mi = mat_images[i];
            m.materials[i].color_metalrough_texture =
                slot(mi.base_color, VGI_TEXTURE_WHITE) | (slot(mi.metal_rough, VGI_TEXTURE_WHITE) << 16);
            m.materials[i].normal_texture = slot(mi.normal, VGI_TEXTURE_NORMAL_UP);
        }
        m.stats.textures = (uint32_t)m.textures.size();
        return VGI_OK;
    }

    // Tangents from the texture coordinates (E. Lengyel, "Computing Tangent
    // Space Basis Vectors for an Arbitrary Mesh", 2001): per vertex, the sum
    // of its triangles' dP/du and dP/dv; the tangent is dP/du made
    // orthogonal to the normal and normalised, w the sign of the bitangent
    // against dP/dv.  A vertex without a usable direction gets any unit
    // vector orthogonal to its normal.
    static void generate_tangents(Prim& p) {
        const size_t nv = p.pos.size() / 3;
        std::vector<double> T(3 * nv, 0.0), B(3 * nv, 0.0);
        for (size_t t = 0; t + 2 < p.idx.size(); t += 3) {
            const uint32_t i[3] = {p.idx[t], p.idx[t + 1], p.idx[t + 2]};
            double e1[3], e2[3];
            for (int c = 0; c < 3; ++c) {
                e1[c] = (double)p.pos[3 * i[1] + c] - p.pos[3 * i[0] + c];
                e2[c] = (double)p.pos[3 * i[2] + c] - p.pos[3 * i[0] + c];
            }
            const double d1u = (double)p.uv[2 * i[1]] - p.uv[2 * i[0]], d1v = (double)p.uv[2 * i[1] + 1] - p.uv[2 * i[0] + 1];
            const double d2u = (double)p.uv[2 * i[2]] - p.uv[2 * i[0]], d2v = (double)p.uv[2 * i[2] + 1] - p.uv[2 * i[0] + 1];
            const double r = d1u * d2v - d2u * d1v;
            if (!(std::fabs(r) > 1e-20)) continue;
            const double inv = 1.0 / r;
            for (int k = 0; k < 3; ++k)
                for (int c = 0; c < 3; ++c) {
                    T[3 * i[k] + c] += (e1[c] * d2v - e2[c] * d1v) * inv;
                    B[3 * i[k] + c] += (e2[c] * d1u - e1[c] * d2u) * inv;
                }
        }
        p.tan.assign(4 * nv, 0.0f);
        for (size_t v = 0; v < nv; ++v) {
            double n[3] = {p.nrm[3 * v], p.nrm[3 * v + 1], p.nrm[3 * v + 2]};
            const double nl = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            if (nl > 0)
                for (double& c : n) c /= nl;
            const double dn = n[0] * T[3 * v] + n[1] * T[3 * v + 1] + n[2] * T[3 * v + 2];
            double t[3] = {T[3 * v] - n[0] * dn, T[3 * v + 1] - n[1] * dn, T[3 * v + 2] - n[2] * dn};
            double tl = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
            if (!(tl > 1e-12)) {  // any unit vector orthogonal to n
                const double a[3] = {std::fabs(n[0]) < 0.9 ? 1.0 : 0.0, std::fabs(n[0]) < 0.9 ? 0.0 : 1.0, 0.0};
                t[0] = n[1] * a[2] - n[2] * a[1];
                t[1] = n[2] * a[0] - n[0] * a[2];
                t[2] = n[0] * a[1] - n[1] * a[0];
                tl = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
            }
            const double cx = n[1] * t[2] - n[2] * t[1], cy = n[2] * t[0] - n[0] * t[2], cz = n[0] * t[1] - n[1] * t[0];
            const double hand = cx * B[3 * v] + cy * B[3 * v + 1] + cz * B[3 * v + 2];
            for (int c = 0; c < 3; ++c) p.tan[4 * v + c] = (float)(t[c] / tl);
            p.tan[4 * v + 3] = hand < 0.0 ? -1.0f : 1.0f;
        }
    }

    vgi_status node_local(const JVal& n, Mat4& out) {
        if (const JVal* mx = n.get("matrix")) {
            if (mx->kind != JVal::kArr || mx->arr.size() != 16)
                return set_error(VGI_E_INVALID, "vgi_gltf_import: node matrix needs 16 numbers");
            for (int k = 0; k < 16; ++k) out.m[k] = (float)mx->arr[k].num;
            return VGI_OK;
        }
        float t[3] = {0, 0, 0}, q[4] = {0, 0, 0, 1}, s[3] = {1, 1, 1};
        auto vec = [](const JVal* v, float* dst, size_t n) {
            if (v && v->kind == JVal::kArr && v->arr.size() == n)
                for (size_t k = 0; k < n; ++k) dst[k] = (float)v->arr[k].num;
        };
        vec(n.get("translation"), t, 3);
        vec(n.get("rotation"), q, 4);
        vec(n.get("scale"), s, 3);
        out = trs(t, q, s);
        return VGI_OK;
    }

    void bake(vgi_gltf_model& m, const Prim& p, const Mat4& x) {
        const float* a = x.m;
        // inverse transpose of the 3x3 part (cofactor matrix / determinant)
        const float a00 = a[0], a10 = a[1], a20 = a[2], a01 = a[4], a11 = a[5], a21 = a[6], a02 = a[8], a12 = a[9],
                    a22 = a[10];
        const float c00 = a11 * a22 - a12 * a21, c01 = a12 * a20 - a10 * a22, c02 = a10 * a21 - a11 * a20;
        const float c10 = a02 * a21 - a01 * a22, c11 = a00 * a22 - a02 * a20, c12 = a01 * a20 - a00 * a21;
        const float c20 = a01 * a12 - a02 * a11, c21 = a02 * a10 - a00 * a12, c22 = a00 * a11 - a01 * a10;
        const float det = a00 * c00 + a01 * c01 + a02 * c02;
        const float id = det != 0.0f ? 1.0f / det : 0.0f;
        // (A^-1)^T = cofactor(A) / det(A): row r of the normal transform is cofactor row r
        const float nt[9] = {c00 * id, c01 * id, c02 * id, c10 * id, c11 * id, c12 * id, c20 * id, c21 * id, c22 * id};
        auto xp = [&](const float* v, float* o) {
            for (int r = 0; r < 3; ++r) o[r] = ((a[r] * v[0] + a[4 + r] * v[1]) + a[8 + r] * v[2]) + a[12 + r];
        };
        auto xn = [&](const float* v, float* o) {
            for (int r = 0; r < 3; ++r) o[r] = (nt[3 * r] * v[0] + nt[3 * r + 1] * v[1]) + nt[3 * r + 2] * v[2];
        };
        const bool mirror = det < 0.0f;
        const size_t ntri = p.idx.size() / 3;
        auto put_uv_tan = [&](uint32_t v) {  // texture coordinate and tangent of source vertex v
            const float uv[2] = {p.uv.empty() ? 0.0f : p.uv[2 * v], p.uv.empty() ? 0.0f : p.uv[2 * v + 1]};
            m.texcoords.insert(m.texcoords.end(), uv, uv + 2);
            float t[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            if (!p.tan.empty()) {
                xn(&p.tan[4 * v], t);
                t[3] = mirror ? -p.tan[4 * v + 3] : p.tan[4 * v + 3];
            }
            m.tangents.insert(m.tangents.end(), t, t + 4);
        };
        m.has_texcoords = m.has_texcoords || !p.uv.empty();
        m.has_tangents = m.has_tangents || !p.tan.empty();
        if (p.has_normals) {
            const uint32_t base = (uint32_t)(m.positions.size() / 3);
            const size_t nv = p.pos.size() / 3;
            for (size_t k = 0; k < nv; ++k) {
                float o[3], n[3];
                xp(&p.pos[3 * k], o);
                xn(&p.nrm[3 * k], n);
                m.positions.insert(m.positions.end(), o, o + 3);
                m.normals.insert(m.normals.end(), n, n + 3);
                put_uv_tan((uint32_t)k);
            }
            for (size_t t = 0; t < ntri; ++t) {
                uint32_t i0 = p.idx[3 * t], i1 = p.idx[3 * t + 1], i2 = p.idx[3 * t + 2];
                if (mirror) std::swap(i1, i2);
                m.indices.push_back(base + i0);
                m.indices.push_back(base + i1);
                m.indices.push_back(base + i2);
                m.tri_material.push_back(p.material);
            }
            m.stats.vertices += (uint32_t)nv;
        } else {
            for (size_t t = 0; t < ntri; ++t) {
                uint32_t c[3] = {p.idx[3 * t], p.idx[3 * t + 1], p.idx[3 * t + 2]};
                if (mirror) std::swap(c[1], c[2]);
                float w[3][3];
                for (int k = 0; k < 3; ++k) xp(&p.pos[3 * c[k]], w[k]);
                const float e1[3] = {w[1][0] - w[0][0], w[1][1] - w[0][1], w[1][2] - w[0][2]};
                const float e2[3] = {w[2][0] - w[0][0], w[2][1] - w[0][1], w[2][2] - w[0][2]};
                const float gn[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                     e1[0] * e2[1] - e1[1] * e2[0]};
                const uint32_t base = (uint32_t)(m.positions.size() / 3);
                for (int k = 0; k < 3; ++k) {
                    m.positions.insert(m.positions.end(), w[k], w[k] + 3);
                    m.normals.insert(m.normals.end(), gn, gn + 3);
                    m.indices.push_back(base + (uint32_t)k);
                    put_uv_tan(c[k]);
                }
                m.tri_material.push_back(p.material);
            }
            m.stats.vertices += (uint32_t)(3 * ntri);
        }
        m.stats.triangles += (uint32_t)ntri;
    }

    std::vector<uint8_t> visited;  // nodes walked so far in the default scene

    vgi_status walk(vgi_gltf_model& m, uint32_t node, const Mat4& parent, int depth) {
        const JVal* nodes = doc.get("nodes");
        if (!nodes || nodes->kind != JVal::kArr || node >= nodes->arr.size())
            return set_error(VGI_E_INVALID, "vgi_gltf_import: node %u does not exist", node);
        if (depth > 1024) return set_error(VGI_E_INVALID, "vgi_gltf_import: node hierarchy deeper than 1024 (a cycle?)");
        // glTF node hierarchies are disjoint strict trees (each node has at
        // most one parent): a node reached twice would be baked twice, and a
        // chain of such nodes doubles the walk per level
        if (visited.size() != nodes->arr.size()) visited.assign(nodes->arr.size(), 0);
        if (visited[node]) return set_error(VGI_E_INVALID, "vgi_gltf_import: node %u is reached twice (nodes must form "
                                                           "strict trees)", node);
        visited[node] = 1;
        const JVal& n = nodes->arr[node];
        Mat4 local;
        vgi_status s = node_local(n, local);
        if (s != VGI_OK) return s;
        const Mat4 x = parent * local;
        uint32_t mesh = 0;
        if (as_u32(n.get("mesh"), mesh)) {
            if (mesh >= meshes.size()) return set_error(VGI_E_INVALID, "vgi_gltf_import: node %u: mesh %u", node, mesh);
            // cap what one scene may bake (the BVH's own limit is 2^31 / 48
            // triangles; vertex indices are u32)
            size_t add = 0, addv = 0;
            for (const Prim& p : meshes[mesh]) {
                add += p.idx.size() / 3;
                addv += p.has_normals ? p.pos.size() / 3 : p.idx.size();
            }
            if (m.stats.instances >= kMaxInstances || m.stats.triangles + add > kMaxTriangles ||
                (uint64_t)m.stats.vertices + addv > kMaxVertices)
                return set_error(VGI_E_UNSUPPORTED,
                                 "vgi_gltf_import: the scene exceeds %u instances, %u triangles or %llu vertices",
                                 kMaxInstances, kMaxTriangles, (unsigned long long)kMaxVertices);
            ++m.stats.instances;
            for (const Prim& p : meshes[mesh]) bake(m, p, x);
        }
        if (const JVal* ch = n.get("children")) {
            if (ch->kind != JVal::kArr) return set_error(VGI_E_INVALID, "vgi_gltf_import: node %u: children", node);
            for (const JVal& c : ch->arr) {
                uint32_t ci = 0;
                if (!as_u32(&c, ci)) return set_error(VGI_E_INVALID, "vgi_gltf_import: node %u: child index", node);
                if ((s = walk(m, ci, x, depth + 1)) != VGI_OK) return s;
            }
        }
        return VGI_OK;
    }

    vgi_status run(const char* path, float scale, vgi_gltf_model& m) {
        vgi_status s = load(path);
        if (s == VGI_OK) s = load_materials(m);
        if (s == VGI_OK) s = load_images(m);
        if (s == VGI_OK) s = load_meshes(m.stats);
        if (s != VGI_OK) return s;
        uint32_t scene = 0;
        if (!as_u32(doc.get("scene"), scene)) return set_error(VGI_E_INVALID, "vgi_gltf_import: no default scene");
        const JVal* scenes = doc.get("scenes");
        if (!scenes || scenes->kind != JVal::kArr || scene >= scenes->arr.size())
            return set_error(VGI_E_INVALID, "vgi_gltf_import: scene %u does not exist", scene);
        // glTF Y-up -> Z-up: rotation by the f32 angle pi/2 about X (columns
        // x, (0, cos, sin), (0, -sin, cos)), then the uniform scale
        const float ang = 1.57079637f, c = std::cos(ang), sn = std::sin(ang);
        Mat4 rot = Mat4::identity(), scl = Mat4::identity();
        rot.m[5] = c;
        rot.m[6] = sn;
        rot.m[9] = -sn;
        rot.m[10] = c;
        scl.m[0] = scl.m[5] = scl.m[10] = scale;
        const Mat4 root = rot * scl;
        if (const JVal* roots = scenes->arr[scene].get("nodes")) {
            if (roots->kind != JVal::kArr) return set_error(VGI_E_INVALID, "vgi_gltf_import: scene nodes");
            for (const JVal& r : roots->arr) {
                uint32_t ri = 0;
                if (!as_u32(&r, ri)) return set_error(VGI_E_INVALID, "vgi_gltf_import: scene node index");
                if ((s = walk(m, ri, root, 0)) != VGI_OK) return s;
            }
        }
        return VGI_OK;
    }
};

}  // namespace

extern "C" vgi_status vgi_gltf_default_config(vgi_gltf_config* cfg) {
    VGI_REQUIRE(cfg, "vgi_gltf_default_config: null");
    cfg->scale = 1.0f;
    return VGI_OK;
}

extern "C" vgi_status vgi_gltf_import(const char* path, const vgi_gltf_config* cfg, vgi_gltf_model** out) {
    VGI_REQUIRE(path && out, "vgi_gltf_import: null argument");
    *out = nullptr;
    const float scale = cfg ? cfg->scale : 1.0f;
    VGI_REQUIRE(std::isfinite(scale) && scale > 0.0f, "vgi_gltf_import: scale must be positive");
    std::unique_ptr<vgi_gltf_model> m(new (std::nothrow) vgi_gltf_model());
    if (!m) return set_error(VGI_E_NOMEM, "vgi_gltf_import: out of host memory");
    try {
        Importer im;
        vgi_status s = im.run(path, scale, *m);
        if (s != VGI_OK) return s;
    } catch (const std::bad_alloc&) {
        return set_error(VGI_E_NOMEM, "vgi_gltf_import: out of host memory");
    }
    VGI_REQUIRE(m->stats.triangles > 0, "vgi_gltf_import: the default scene has no triangles");
    *out = m.release();
    return VGI_OK;
}

extern "C" vgi_status vgi_gltf_model_desc(const vgi_gltf_model* m, vgi_scene_desc* desc) {
    VGI_REQUIRE(m && desc, "vgi_gltf_model_desc: null argument");
    desc->positions = m->positions.data();
    desc->normals = m->normals.data();
    desc->num_vertices = (uint32_t)(m->positions.size() / 3);
    desc->indices = m->indices.data();
    desc->num_triangles = (uint32_t)(m->indices.size() / 3);
    desc->triangle_material = m->tri_material.data();
    desc->materials = m->materials.data();
    desc->num_materials = (uint32_t)m->materials.size();
    desc->texcoords = m->has_texcoords ? m->texcoords.data() : nullptr;
    desc->tangents = m->has_tangents ? m->tangents.data() : nullptr;
    desc->textures = m->textures.data();
    desc->num_textures = (uint32_t)m->textures.size();
    return VGI_OK;
}

extern "C" vgi_status vgi_gltf_model_stats(const vgi_gltf_model* m, vgi_gltf_stats* out) {
    VGI_REQUIRE(m && out, "vgi_gltf_model_stats: null argument");
    *out = m->stats;
    return VGI_OK;
}

extern "C" vgi_status vgi_gltf_destroy(vgi_gltf_model* m) {
    delete m;
    return VGI_OK;
}
