/*
 * vgi.h — C ABI of the MI355X-native ReSTIR GI hot path ("vgi").
 *
 * Drop-in boundary for violet's render-graph passes.  violet records each
 * pass with a builder (crates/violet/src/render_graph.rs:603-763): a shader,
 * named resource bindings (.texture / .rw_texture / .buffer / .rw_buffer,
 * render_graph.rs:316-391), packed push constants (render_graph.rs:393-399,
 * 2559-2583) and a dispatch size.  Here every such pass is one entry point
 *
 *     vgi_status vgi_pass_<name>(const vgi_frame_params*, const <io>*,
 *                                const <push constants>*, void* stream);
 *
 * where <io> carries one device pointer per binding name of the HLSL pass
 * (same names as the shader declarations) and <push constants> is the HLSL
 * `PushConstants` struct field for field.  `stream` is a hipStream_t (NULL =
 * default stream).  Calls are asynchronous on that stream and never
 * synchronise the host.  Every image is a row-major width*height array;
 * G-buffer images are four such planes (Texture2DArray layers 0..3).
 * Per-pixel floating-point fields use the storage mode named in the IO
 * struct (VGI_STORAGE_F32: 3/2/1 x float; VGI_STORAGE_F16: 3/2/1 x binary16).
 *
 * Errors: every call returns VGI_OK (0) or a negative vgi_status and never
 * aborts; vgi_last_error() returns a thread-local message.  violet itself
 * warns and continues on a missing binding (render_graph.rs:1881-1886); here
 * a missing required pointer or a bad size is VGI_E_INVALID and nothing is
 * launched.  A NULL history input together with its has_* flag = 0 is the
 * equivalent of violet's dummy textures (restir_renderer.rs:332-350).
 */
#ifndef VGI_H
#define VGI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t vgi_status;
#define VGI_OK 0
#define VGI_E_INVALID (-1)
#define VGI_E_HIP (-2)
#define VGI_E_BUDGET (-3)
#define VGI_E_NOMEM (-4)
#define VGI_E_UNSUPPORTED (-5)
#define VGI_E_TIMEOUT (-6)

#define VGI_STORAGE_F32 0u
#define VGI_STORAGE_F16 1u

/* Thread-local description of the last error on this thread. */
const char* vgi_last_error(void);
/* Library version string. */
const char* vgi_version(void);
/* Device-to-device copy on `stream` (utility for hosts that need to read a
 * library-owned image, e.g. a denoiser output; no reference counterpart:
 * violet's render graph reads such images through descriptor bindings). */
vgi_status vgi_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);

/* Frame fence with a deadline: violet waits for the oldest frame's command
 * buffer fence for at most 5 s, warns above 33.3 ms and carries on after a
 * time-out (crates/violet-app/src/render_loop.rs:445-482).  A fence marks
 * everything issued on a stream so far; waiting never blocks past the
 * deadline.  VGI_FENCE_TIMEOUT_MS is violet's deadline. */
#define VGI_FENCE_TIMEOUT_MS 5000u
typedef struct vgi_fence vgi_fence;
vgi_status vgi_fence_create(vgi_fence** out);
vgi_status vgi_fence_destroy(vgi_fence* fence);
/* Marks the work issued on `stream` so far (replaces an earlier mark). */
vgi_status vgi_fence_record(vgi_fence* fence, void* stream);
/* VGI_OK once the marked work has finished (also for a fence never
 * recorded); VGI_E_TIMEOUT if it has not within timeout_ms (0 = poll once):
 * the work keeps running and the fence can be waited on again.  The time
 * spent waiting goes to *waited_ms when it is not NULL. */
vgi_status vgi_fence_wait(vgi_fence* fence, uint32_t timeout_ms, double* waited_ms);

/* Per-pass GPU timers: the counterpart of violet's per-pass timestamp
 * queries (crates/violet/src/gpu_profiling.rs:253-355, written around every
 * pass by render_graph.rs:2585-2591,2664-2669).  When enabled, every pass
 * entry point records a hipEvent pair around its kernel(s) on its stream.
 * vgi_profiler_read() waits for the recorded events, accumulates them per
 * pass name and returns up to `cap` entries. */
typedef struct vgi_profile_entry {
    char name[48];
    double total_ms;     /* summed kernel time */
    uint64_t count;      /* number of launches */
    double bytes;        /* summed algorithmic bytes of those launches */
    /* traversal counts of those launches (zero unless trace stats are on) */
    double rays_closest; /* closest-hit BVH traversals */
    double rays_any;     /* any-hit (shadow) BVH traversals */
    double nodes;        /* interior (4-wide) nodes visited (VGI_BVH_NODE_BYTES each) */
    double tris;         /* triangles tested (VGI_BVH_TRI_BYTES each) */
} vgi_profile_entry;
vgi_status vgi_profiler_enable(uint32_t enable);
/* Restrict the timers to the passes named `name` (NULL or "" = every pass):
 * a timed region can then carry the events of one kernel only. */
vgi_status vgi_profiler_only(const char* name);
vgi_status vgi_profiler_reset(void);
int32_t vgi_profiler_read(vgi_profile_entry* out, int32_t cap);

/* Launch shapes (no reference counterpart: violet's shaders fix them with
 * [numthreads], e.g. svgf_atrous_filter.hlsl's 8x8).  Results do not depend
 * on them; violet_amd/autotune.py times the candidates on the views to be
 * rendered and sets the fastest.  The setting is process-wide and applies to
 * launches issued after the call (set it before rendering, not while another
 * thread issues frames).
 *   atrous_block      a-trous lattice block edge staged in LDS per workgroup:
 *                     8 (64 threads), 16 (256, the default) or 32 (1024)
 *   trace_tile_width  pixel tile of one 64-lane wave in the four sample
 *                     passes: 8 (8x8 in 2x2 quads, the default), 16 (16x4)
 *                     or 32 (32x2) */
typedef struct vgi_tuning {
    uint32_t atrous_block;
    uint32_t trace_tile_width;
} vgi_tuning;
vgi_status vgi_tuning_set(const vgi_tuning* tuning); /* NULL = the defaults */
vgi_status vgi_tuning_get(vgi_tuning* tuning);

/* Traversal statistics (no reference counterpart; measurement only).  When
 * enabled, every ray-tracing pass launches an instrumented build of its
 * kernel that counts rays, interior nodes visited and triangles tested, and
 * waits for it, adding the counts to the pass's profiler entry (the
 * profiler must be enabled too).  Counts do not depend on scheduling.
 * Allocates a small counter buffer on the current device when turned on. */
#define VGI_BVH_NODE_BYTES 64
#define VGI_BVH_TRI_BYTES 48
vgi_status vgi_trace_stats_enable(uint32_t enable);

/* Dependent-fetch probe (no reference counterpart; measurement only): the
 * rate at which the GPU completes dependent 128-byte fetches when every
 * resident lane chases a random chain through a buffer of `bytes` (say, the
 * scene BVH's size), launched like the ray-tracing kernels (64-lane groups,
 * 8 waves per SIMD on every CU).  A step loads one 128-byte record (eight
 * 16-byte loads issued together) and follows the index in its first word, as
 * a traversal fetches a node and then one of its children.  Groups of
 * `lanes_per_chain` consecutive lanes (a power of two, 1..64) follow the
 * same chain: 64 is a wave of perfectly coherent rays, 1 a wave of rays
 * that share no node.  Writes the completed lane-steps per second.
 * Allocates and frees its own buffers and waits for the stream. */
vgi_status vgi_probe_dependent_fetch(size_t bytes, uint32_t steps, uint32_t lanes_per_chain, double* steps_per_s,
                                     void* stream);

/* FrameParams (shader/frame_bindings.hlsl:6-25; filled by
 * crates/violet-app/src/render_loop.rs:108-158).  384 bytes, matrices
 * column-major (glam / #pragma pack_matrix(column_major)). */
typedef struct vgi_frame_params {
    float view_proj[16];
    float inv_view_proj[16];
    float view_pos[4];
    float view_ray_top_left[4];
    float view_ray_right_shift[4];
    float view_ray_down_shift[4];
    float prev_view_proj[16];
    float prev_inv_view_proj[16];
    float prev_view_pos[4];
    float jitter[4]; /* xy: this frame, zw: previous frame (NDC units) */
    float sun_dir[4];
    float sun_inten[4];
} vgi_frame_params;

/* ------------------------------------------------------------------------
 * Denoiser passes (ReLAX/SVGF), shader/denoise/ (HLSL)
 * ---------------------------------------------------------------------- */

/* Bindings of denoise/svgf_temporal_filter.hlsl:16-39.
 * Replaces rg.new_compute("Temporal Filter") at
 * crates/violet-restir/src/denoising.rs:136-160. */
typedef struct vgi_denoise_temporal_io {
    uint32_t width, height, storage;
    const float* depth_texture;              /* f32 */
    const uint32_t* gbuffer_texture;         /* 4 layers of u32 */
    const void* diff_source_texture;         /* radiance */
    const void* spec_source_texture;         /* radiance */
    const void* spec_ray_len_texture;        /* scalar */
    const float* prev_depth_texture;         /* may be NULL if !has_history */
    const uint32_t* prev_gbuffer_texture;    /* may be NULL if !has_history */
    const void* diff_history_texture;        /* may be NULL if !has_history */
    const void* spec_history_texture;        /* may be NULL if !has_history */
    const void* moments_history_texture;     /* pair; may be NULL if !has_history */
    const uint8_t* history_len_texture;      /* R8 count; may be NULL if !has_history */
    const void* diff_fast_history_texture;   /* used only if fast_history */
    const void* spec_fast_history_texture;   /* used only if fast_history */
    void* rw_diff_texture;
    void* rw_spec_texture;
    void* rw_moments_texture;
    uint8_t* rw_history_len_texture;
    void* rw_diff_fast_history_texture;      /* written only if fast_history */
    void* rw_spec_fast_history_texture;      /* written only if fast_history */
    /* FP16 storage only, both NULL or both set: the low f16 halves of the
     * compensated moments (moment = hi + lo, hi in the moments texture,
     * lo = f16(m - hi) the rounding residual).  NULL = plain f16 moments.
     * No reference counterpart (the reference stores R16G16_SFLOAT). */
    const void* moments_comp_history_texture; /* pair; may be NULL if !has_history */
    void* rw_moments_comp_texture;            /* pair */
} vgi_denoise_temporal_io;

/* svgf_temporal_filter.hlsl:41-47 (host: denoising.rs:157-159). */
typedef struct vgi_pc_denoise_temporal {
    float buffer_size[2];
    uint32_t has_history;
    uint32_t fast_history;
} vgi_pc_denoise_temporal;

vgi_status vgi_pass_denoise_temporal(const vgi_frame_params* fp, const vgi_denoise_temporal_io* io,
                                     const vgi_pc_denoise_temporal* pc, void* stream);

/* Bindings of denoise/relax_disocclusion_fix.hlsl:16-25 and
 * denoise/svgf_spatial_fallback_filter.hlsl:14-23 (the fallback writes
 * rw_variance_texture in place of rw_moments_texture).
 * Replaces denoising.rs:172-183 and :228-239. */
typedef struct vgi_denoise_spatial_io {
    uint32_t width, height, storage;
    const float* depth_texture;
    const uint32_t* gbuffer_texture;
    const uint8_t* history_len_texture;
    const void* diff_texture;
    const void* spec_texture;
    const void* moments_texture;
    void* rw_diff_texture;
    void* rw_spec_texture;
    void* rw_moments_or_variance_texture;
    /* FP16 compensated moments (see vgi_denoise_temporal_io): read by both
     * passes when set; the disocclusion fix also writes the residual of its
     * output moments to rw_moments_comp_texture (NULL for the fallback). */
    const void* moments_comp_texture;
    void* rw_moments_comp_texture;
} vgi_denoise_spatial_io;

vgi_status vgi_pass_denoise_disocclusion_fix(const vgi_denoise_spatial_io* io, void* stream);
vgi_status vgi_pass_denoise_spatial_fallback(const vgi_denoise_spatial_io* io, void* stream);

/* Bindings of denoise/relax_fast_history_clamp.hlsl:5-11 (off by default,
 * denoising.rs:55; replaces denoising.rs:192-207).  rw_* are in/out. */
typedef struct vgi_denoise_fast_clamp_io {
    uint32_t width, height, storage;
    const void* diff_fast_history_texture;
    const void* spec_fast_history_texture;
    void* rw_diff_texture;
    void* rw_spec_texture;
} vgi_denoise_fast_clamp_io;

vgi_status vgi_pass_denoise_fast_history_clamp(const vgi_denoise_fast_clamp_io* io, void* stream);

/* Bindings of denoise/svgf_atrous_filter.hlsl:6-15; replaces one
 * "Spatial Filter: i" pass (denoising.rs:267-278). */
typedef struct vgi_denoise_atrous_io {
    uint32_t width, height, storage;
    const float* depth_texture;
    const uint32_t* gbuffer_texture;
    const void* diff_texture;
    const void* spec_texture;
    const void* variance_texture;
    void* rw_diff_texture;
    void* rw_spec_texture;
    void* rw_variance_texture;
} vgi_denoise_atrous_io;

/* svgf_atrous_filter.hlsl:17-21 */
typedef struct vgi_pc_denoise_atrous {
    int32_t step_size;
} vgi_pc_denoise_atrous;

vgi_status vgi_pass_denoise_atrous(const vgi_denoise_atrous_io* io, const vgi_pc_denoise_atrous* pc,
                                   void* stream);

/* ------------------------------------------------------------------------
 * Denoiser sequencer: crates/violet-restir/src/denoising.rs `Denoiser`.
 * Owns the temporal history (RGTemporal textures of denoising.rs:41-49)
 * and the transient ping-pong images, in HBM, per view.
 * ---------------------------------------------------------------------- */
typedef struct vgi_denoiser vgi_denoiser;

typedef struct vgi_denoiser_config {
    uint32_t disocclusion_fix;   /* default 1 (denoising.rs:54) */
    uint32_t fast_history_clamp; /* default 0 (denoising.rs:55) */
    uint32_t atrous_iterations;  /* default 3 (denoising.rs:56), clamped to 0..5 */
    uint32_t compensated_moments; /* FP16 storage: keep moments as hi + lo f16 pairs (no reference counterpart) */
} vgi_denoiser_config;

/* denoising.rs:10-18 `Input` */
typedef struct vgi_denoiser_input {
    const float* depth;
    const uint32_t* gbuffer;
    const float* prev_depth;
    const uint32_t* prev_gbuffer;
    const void* diffuse;
    const void* specular;
    const void* specular_ray_len;
} vgi_denoiser_input;

/* denoising.rs:27-31 `Output` (+ DebugViews :20-25).  Pointers reference
 * denoiser-owned device images valid until the next add_passes/destroy. */
typedef struct vgi_denoiser_output {
    const void* diffuse;
    const void* specular;
    const uint8_t* history_len;
    const void* variance;
} vgi_denoiser_output;

vgi_status vgi_denoiser_create(uint32_t width, uint32_t height, uint32_t storage,
                               const vgi_denoiser_config* cfg, vgi_denoiser** out);
vgi_status vgi_denoiser_destroy(vgi_denoiser* d);
/* denoising.rs:62-64 */
vgi_status vgi_denoiser_reset(vgi_denoiser* d);
vgi_status vgi_denoiser_set_config(vgi_denoiser* d, const vgi_denoiser_config* cfg);
/* denoising.rs:66-301 */
vgi_status vgi_denoiser_add_passes(vgi_denoiser* d, const vgi_frame_params* fp, const vgi_denoiser_input* in,
                                   vgi_denoiser_output* out, void* stream);

/* ------------------------------------------------------------------------
 * Allocation-free denoiser frame.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 *
 * vgi_denoise_frame_static runs the pass chain of vgi_denoiser_add_passes
 * (same kernels, same results) with the caller owning every buffer that
 * outlives the call: the inputs, the outputs and the history.  The history
 * is read from history_* (history_len NULL, or prev_depth/prev_gbuffer
 * NULL, = no history, like vgi_denoiser_reset) and written to rw_history_*,
 * which must not alias history_*: pass this frame's rw_history_* as the
 * next frame's history_*.  The frame's intermediate images live in one
 * fixed device array of VGI_DENOISE_STATIC_SCRATCH_BYTES per GPU, which the
 * library holds until the process ends.  They take 64 B/px with FP32
 * storage and 36 B/px with FP16 (32 without compensated moments), so at most
 * about 2.09 Mpx FP32 (1920x1080 fits) or 3.7 Mpx FP16 (2560x1440 fits);
 * larger images return VGI_E_INVALID.  cfg NULL = the defaults
 * {1, 0, 3, 1}; fast_history_clamp must be 0.  The call makes no device
 * allocation (with the profiler off, its default).  Calls on one GPU share
 * the array: calls that overlap on one GPU (two streams or two host
 * threads) race.
 * ---------------------------------------------------------------------- */
#define VGI_DENOISE_STATIC_SCRATCH_BYTES (128u << 20)

typedef struct vgi_denoise_frame_io {
    uint32_t width, height, storage;
    /* denoising.rs:10-18 Input */
    const float* depth;
    const uint32_t* gbuffer;
    const float* prev_depth;
    const uint32_t* prev_gbuffer;
    const void* diffuse;
    const void* specular;
    const void* specular_ray_len;
    /* history read (denoising.rs:41-49 History): radiance, radiance, pair,
     * FP16 moment residual (may be NULL), R8 count */
    const void* history_diffuse;
    const void* history_specular;
    const void* history_moments;
    const void* history_moments_comp;
    const uint8_t* history_len;
    /* history written; rw_history_moments_comp is required for FP16 storage
     * with compensated moments and unused otherwise */
    void* rw_history_diffuse;
    void* rw_history_specular;
    void* rw_history_moments;
    void* rw_history_moments_comp;
    uint8_t* rw_history_len;
    /* denoising.rs:27-31 Output; rw_variance (DebugViews.variance, the
     * spatial fallback's variance) may be NULL */
    void* rw_diffuse;
    void* rw_specular;
    void* rw_variance;
} vgi_denoise_frame_io;

vgi_status vgi_denoise_frame_static(const vgi_frame_params* fp, const vgi_denoise_frame_io* io,
                                    const vgi_denoiser_config* cfg, void* stream);

/* ------------------------------------------------------------------------
 * Scene and software ray tracing.
 * Replaces the BLAS/TLAS build of crates/violet/src/render_scene.rs:
 * 1170-1210 (BLAS per geometry group, UINT16 indices, f32x3 positions) and
 * :1557-1712 (TLAS over instances) with a host-built binary BVH in HBM, and
 * TraceRay (restir/raytrace.inc.hlsl:42-53,74-82) with HIP traversal.
 * Instances are baked to world space by the caller; indices are u32.
 * ---------------------------------------------------------------------- */
typedef struct vgi_scene vgi_scene;

/* Material textures (crates/violet/src/model.rs:800-945 with
 * --disable-tex-compress, render_scene.rs:344-346,768-833,920-935): RGBA8
 * texels with a full mip chain.  Level l is max(1, width >> l) x
 * max(1, height >> l) texels, rows packed, levels back to back (violet's
 * importer makes 1 + log2(next_pow2(max(width, height))) levels by halving).
 * srgb = 1 marks R8G8B8A8_SRGB (base colour: texels are decoded to linear
 * before filtering); 0 is UNORM.  A metallic-roughness texture carries
 * metallic in R and roughness in G (violet's importer swizzles glTF's B/G,
 * model.rs:811-814).  Sampling is violet's sampler_linear_wrap
 * (render_loop.rs:199-212): bilinear, linear between mips, REPEAT. */
typedef struct vgi_texture {
    uint32_t width, height;
    uint32_t levels; /* 1..16 */
    uint32_t srgb;
    const uint8_t* texels; /* host memory */
} vgi_texture;

/* Texture indices of a material, as violet's bindless table numbers them:
 * 0 = a white 1x1 texture, 1 = a flat normal map (128, 128, 255, 255), and
 * k >= 2 = vgi_scene_desc.textures[k - 2]. */
#define VGI_TEXTURE_WHITE 0u
#define VGI_TEXTURE_NORMAL_UP 1u
#define VGI_TEXTURE_FIRST 2u

/* MaterialParams (shader/scene_bindings.hlsl:24-51): the factors multiply
 * the sampled textures (geometry_rchit.inc.hlsl:79-86).  The texture words
 * pack two 16-bit indices as violet's do: base colour | metallic-roughness
 * << 16, and normal map in the low 16 bits (the high half, violet's
 * transmission map, is unused).  An all-zero word means white textures. */
typedef struct vgi_material {
    float base_color[4];
    float metallic;
    float roughness; /* perceptual roughness */
    uint32_t color_metalrough_texture;
    uint32_t normal_texture;
} vgi_material;

typedef struct vgi_scene_desc {
    const float* positions;          /* world space, 3 floats per vertex */
    const float* normals;            /* per-vertex normals (3 floats) or NULL = geometric normal */
    uint32_t num_vertices;
    const uint32_t* indices;         /* 3 per triangle */
    uint32_t num_triangles;
    const uint32_t* triangle_material; /* per triangle or NULL (= material 0) */
    const vgi_material* materials;
    uint32_t num_materials;
    /* Texturing (all optional).  texcoords: TEXCOORD_0, 2 floats per vertex
     * (NULL = (0, 0)).  tangents: 4 floats per vertex, xyz in world space
     * and w the bitangent sign (MikkTSpace convention); NULL disables the
     * tangent-space terms of normal mapping (violet's zero tangents).
     * textures: indexed from VGI_TEXTURE_FIRST by the materials. */
    const float* texcoords;
    const float* tangents;
    const vgi_texture* textures;
    uint32_t num_textures;
} vgi_scene_desc;

typedef struct vgi_scene_stats {
    uint32_t num_triangles, num_vertices, num_nodes, depth;
    uint64_t device_bytes;
} vgi_scene_stats;

/* Host arrays in, BVH built on the host, uploaded to the current device. */
vgi_status vgi_scene_create(const vgi_scene_desc* desc, vgi_scene** out);
vgi_status vgi_scene_destroy(vgi_scene* s);
vgi_status vgi_scene_info(const vgi_scene* s, vgi_scene_stats* out);

/* glTF 2.0 import (crates/violet/src/model.rs:389-720 import_gltf_uncached;
 * the scene's BLAS/TLAS of render_scene.rs:1170-1712 are replaced by one BVH
 * over the baked triangles).  Kept primitives: TRIANGLES mode with a
 * material, indices and POSITION; indices read as u16 (u32 indices are
 * truncated, as violet's importer truncates them).  The default scene
 * (`scene` property; none is an error) is flattened depth first with each
 * node's matrix or T*R*S after its parent's, under the root transform
 * rotation(+90 deg about X) * scale (glTF Y-up -> violet Z-up).  Instances
 * are baked to world space: positions through the affine transform, normals
 * through the inverse transpose without normalising, winding reversed under
 * a mirroring transform; a primitive without NORMAL gets per-corner vertices
 * with the geometric normal.  Materials keep the PBR factors (glTF defaults
 * 1) and their base-colour, metallic-roughness and normal textures: PNG
 * and baseline JPEG images are decoded to RGBA8 with violet's Lanczos3 mip
 * chain; other forms (progressive JPEG, ...) are counted in
 * stats.images_unsupported and leave violet's default textures in their
 * slots.  TEXCOORD_0 and TANGENT are kept
 * (tangents are generated from the texture coordinates when absent).
 * Sources: .gltf with relative-file or base64 data-URI buffers and images,
 * or .glb. */
typedef struct vgi_gltf_model vgi_gltf_model;
typedef struct vgi_gltf_config {
    float scale; /* LoadConfig::scale (model.rs:20-33), default 1 */
} vgi_gltf_config;
typedef struct vgi_gltf_stats {
    uint32_t meshes, primitives, primitives_skipped, instances, materials, textured_materials;
    uint32_t triangles, vertices, truncated_indices;
    uint32_t images;             /* images the materials use */
    uint32_t images_unsupported; /* of which not decodable here (their slots keep the defaults) */
    uint32_t textures;           /* textures made (decoded images) */
} vgi_gltf_stats;
vgi_status vgi_gltf_default_config(vgi_gltf_config* cfg);
vgi_status vgi_gltf_import(const char* path, const vgi_gltf_config* cfg, vgi_gltf_model** out);
/* Fills desc with the model's arrays (valid until vgi_gltf_destroy); pass it to vgi_scene_create. */
vgi_status vgi_gltf_model_desc(const vgi_gltf_model* m, vgi_scene_desc* desc);
vgi_status vgi_gltf_model_stats(const vgi_gltf_model* m, vgi_gltf_stats* out);
vgi_status vgi_gltf_destroy(vgi_gltf_model* m);

/* Batch ray query (TraceRay without shading).  rays: 8 floats per ray
 * {o.x,o.y,o.z,tmin, d.x,d.y,d.z,tmax} in device memory; hits: 4 words per
 * ray {t (f32), u (f32), v (f32), triangle id (u32, 0xFFFFFFFF = miss)}.
 * any_hit != 0 gives shadow-ray semantics (first accepted hit, t and
 * barycentrics unspecified, id = some hit triangle). */
vgi_status vgi_trace_rays(const vgi_scene* scene, const float* rays, uint32_t num_rays, uint32_t any_hit,
                          uint32_t* hits, void* stream);

/* ------------------------------------------------------------------------
 * ReSTIR GI passes (shader/restir/).
 * ---------------------------------------------------------------------- */

/* HashGridCell (restir/hash_grid.inc.hlsl:112-122): 16 bytes.  The
 * reference host allocates 20 B per cell (restir_renderer.rs:27-28); the
 * HLSL struct is 16 B and that packed layout is used here. */
typedef struct vgi_hash_grid_cell {
    uint32_t checksum; /* 0 = empty */
    float radiance[3];
} vgi_hash_grid_cell;

/* HashGridQuery (restir/hash_grid.inc.hlsl:186-207): 24 bytes. */
typedef struct vgi_hash_grid_query {
    float hit_position[3];
    uint32_t hit_normal_encoded; /* oct u32 of the geometric normal */
    uint32_t cell_hash;
    uint32_t cell_checksum;
} vgi_hash_grid_query;

#define VGI_HASH_GRID_NUM_CELLS (65536u * 16u)
#define VGI_HASH_GRID_MAX_QUERIES (128u * 1024u)

/* Resources of restir/raytrace.inc.hlsl:17-28 bound by every radiance-trace
 * pass (restir_renderer.rs:99-121 restir_raytrace_resrouces +
 * hg_cache_resources).  The sky cube is replaced by a constant sky
 * radiance.  prev_* may be NULL when the trace pass has has_prev_frame = 0. */
typedef struct vgi_trace_resources {
    const vgi_scene* scene_tlas;
    float sky_radiance[3];
    uint32_t storage;                           /* storage mode of prev_indirect_diffuse_texture */
    const void* prev_indirect_diffuse_texture;  /* radiance, previous frame's denoised diffuse */
    const float* prev_depth_texture;
    uint32_t prev_width, prev_height;           /* size of the two prev_* images */
    const vgi_hash_grid_cell* hash_grid_storage_buffer;
    vgi_hash_grid_query* rw_hash_grid_query_buffer;
    uint32_t* rw_hash_grid_query_counter_buffer;
    /* NULL: every query is appended, as in the reference.  Otherwise a
     * device word T written by vgi_pass_hash_grid_query_select_ordered: a
     * query is appended only if the hash of its 24 bytes is >= T, a
     * deterministic sample of about (2^32 - T) / 2^32 of the queries, so
     * the appended set stays below VGI_HASH_GRID_MAX_QUERIES and does not
     * depend on the order of the appends (no reference counterpart). */
    const uint32_t* hash_grid_query_threshold;
} vgi_trace_resources;

/* restir/ind_diff_sample_gen.hlsl:11-23; restir_renderer.rs:478-492 */
typedef struct vgi_ind_diff_sample_gen_io {
    uint32_t width, height, storage;
    const uint32_t* gbuffer_color;
    const float* gbuffer_depth;
    uint32_t* rw_hit_pos_normal_texture; /* uint2 per pixel */
    void* rw_hit_radiance_texture;       /* radiance */
} vgi_ind_diff_sample_gen_io;
typedef struct vgi_pc_sample_gen {
    uint32_t frame_rand;
    uint32_t has_prev_frame;
} vgi_pc_sample_gen;
vgi_status vgi_pass_ind_diff_sample_gen(const vgi_frame_params* fp, const vgi_trace_resources* rt,
                                        const vgi_ind_diff_sample_gen_io* io, const vgi_pc_sample_gen* pc,
                                        void* stream);

/* restir/ind_diff_sample_validate.hlsl:83-88; restir_renderer.rs:499-512
 * (depth_texture = previous frame's depth; rw_* = previous reservoir). */
typedef struct vgi_ind_diff_sample_validate_io {
    uint32_t width, height, storage;
    const float* depth_texture;
    uint32_t* rw_reservoir_texture;
    uint32_t* rw_hit_pos_normal_texture;
    void* rw_hit_radiance_texture;
} vgi_ind_diff_sample_validate_io;
vgi_status vgi_pass_ind_diff_sample_validate(const vgi_frame_params* fp, const vgi_trace_resources* rt,
                                             const vgi_ind_diff_sample_validate_io* io, void* stream);

/* restir/ind_spec_sample_gen.hlsl:12-26; restir_renderer.rs:685-701 */
typedef struct vgi_ind_spec_sample_gen_io {
    uint32_t width, height, storage;
    const uint32_t* gbuffer_color;
    const float* gbuffer_depth;
    float* rw_hit_pos_texture;     /* float4 per pixel */
    void* rw_hit_radiance_texture; /* radiance */
} vgi_ind_spec_sample_gen_io;
vgi_status vgi_pass_ind_spec_sample_gen(const vgi_frame_params* fp, const vgi_trace_resources* rt,
                                        const vgi_ind_spec_sample_gen_io* io, const vgi_pc_sample_gen* pc,
                                        void* stream);

/* restir/ind_spec_sample_validate.hlsl:140-142; restir_renderer.rs:722-737
 * (reads prev_depth_texture of the trace resources). */
typedef struct vgi_ind_spec_sample_validate_io {
    uint32_t width, height, storage;
    uint32_t* rw_prev_reservoir_texture; /* uint2 */
    float* rw_prev_hit_pos_texture;      /* float4 */
    void* rw_prev_hit_radiance_texture;  /* radiance */
} vgi_ind_spec_sample_validate_io;
vgi_status vgi_pass_ind_spec_sample_validate(const vgi_frame_params* fp, const vgi_trace_resources* rt,
                                             const vgi_ind_spec_sample_validate_io* io, void* stream);

/* restir/raytraced_ao_gen.hlsl:5-16; restir_renderer.rs:528-545 */
typedef struct vgi_ao_gen_io {
    uint32_t width, height, storage;
    const uint32_t* hit_pos_normal_texture; /* may be NULL if !has_new_sample */
    const float* depth_buffer;
    const void* prev_ao_texture;            /* pair (x, a); may be NULL if !has_prev_frame */
    void* rw_ao_texture;                    /* pair (x, a) */
} vgi_ao_gen_io;
typedef struct vgi_pc_ao_gen {
    uint32_t has_new_sample;
    uint32_t has_prev_frame;
    float radius_ws;
} vgi_pc_ao_gen;
vgi_status vgi_pass_raytraced_ao_gen(const vgi_frame_params* fp, const vgi_ao_gen_io* io, const vgi_pc_ao_gen* pc,
                                     void* stream);

/* restir/raytraced_ao_filter.hlsl:4-6; restir_renderer.rs:551-560 */
typedef struct vgi_ao_filter_io {
    uint32_t width, height, storage;
    const float* depth_buffer;
    const void* ao_texture;
    void* rw_filtered_ao_texture;
} vgi_ao_filter_io;
vgi_status vgi_pass_raytraced_ao_filter(const vgi_ao_filter_io* io, void* stream);

/* restir/ind_diff_temporal_resample.hlsl:15-36; restir_renderer.rs:589-607 */
typedef struct vgi_ind_diff_temporal_io {
    uint32_t width, height, storage;
    const float* prev_gbuffer_depth;
    const float* gbuffer_depth;
    const uint32_t* new_hit_pos_normal_texture;
    const void* new_hit_radiance_texture;
    const uint32_t* prev_reservoir_texture;
    const uint32_t* prev_hit_pos_normal_texture;
    const void* prev_hit_radiance_texture;
    uint32_t* rw_reservoir_texture;
    uint32_t* rw_hit_pos_normal_texture;
    void* rw_hit_radiance_texture;
} vgi_ind_diff_temporal_io;
/* PushConstants of ind_diff_temporal_resample.hlsl:31-36, field for field (12 B). */
typedef struct vgi_pc_ind_diff_temporal {
    uint32_t frame_rand;
    uint32_t has_prev_frame;
    uint32_t has_new_sample;
} vgi_pc_ind_diff_temporal;

/* Resampling constants the HLSL compiles in, exposed as one optional
 * argument of the four resampling passes (it is no part of any HLSL
 * PushConstants).  NULL, or a zero field, keeps the reference constant:
 *   temporal_m_cap     diffuse history M clamp M_MAX = 16
 *                      (ind_diff_temporal_resample.hlsl:136); the specular
 *                      clamp is uint(cap * roughness) (ind_spec_temporal_resample.hlsl:252)
 *   diff_spatial_taps  golden-spiral taps MAX_ITERATION = 9
 *                      (ind_diff_spatial_resample.hlsl:12), at most 16
 *   spatial_radius_px  initial search radius INIT_RADIUS_FACTOR * min(w, h)
 *                      = 0.05 * min(w, h) (ind_diff_spatial_resample.hlsl:13,
 *                      ind_spec_spatial_resample.hlsl:13), at most 16384 */
typedef struct vgi_resample_params {
    uint32_t temporal_m_cap;
    uint32_t diff_spatial_taps;
    float spatial_radius_px;
} vgi_resample_params;

vgi_status vgi_pass_ind_diff_temporal_resample(const vgi_frame_params* fp, const vgi_ind_diff_temporal_io* io,
                                               const vgi_pc_ind_diff_temporal* pc,
                                               const vgi_resample_params* params, void* stream);

/* restir/ind_diff_spatial_resample.hlsl:29-49; restir_renderer.rs:632-652 */
typedef struct vgi_ind_diff_spatial_io {
    uint32_t width, height, storage;
    const uint32_t* gbuffer_color;
    const float* gbuffer_depth;
    const void* ao_texture; /* the filtered AO pair; .x is used */
    const uint32_t* temp_reservoir_texture;
    const uint32_t* temp_hit_pos_normal_texture;
    const void* temp_hit_radiance_texture;
    uint32_t* rw_reservoir_texture;
    uint32_t* rw_hit_pos_normal_texture;
    void* rw_hit_radiance_texture;
    void* rw_lighting_texture;
} vgi_ind_diff_spatial_io;
/* PushConstants {uint frame_rand} of ind_diff_spatial_resample.hlsl:45-49 and
 * ind_spec_spatial_resample.hlsl:31-35 (4 B). */
typedef struct vgi_pc_frame_rand {
    uint32_t frame_rand;
} vgi_pc_frame_rand;
typedef vgi_pc_frame_rand vgi_pc_ind_diff_spatial;
vgi_status vgi_pass_ind_diff_spatial_resample(const vgi_frame_params* fp, const vgi_ind_diff_spatial_io* io,
                                              const vgi_pc_ind_diff_spatial* pc,
                                              const vgi_resample_params* params, void* stream);

/* restir/ind_spec_temporal_resample.hlsl:26-44; restir_renderer.rs:743-761
 * (rw_hit_pos/rw_hit_radiance hold the new sample on entry). */
typedef struct vgi_ind_spec_temporal_io {
    uint32_t width, height, storage;
    const uint32_t* prev_gbuffer_color;
    const float* prev_gbuffer_depth;
    const uint32_t* gbuffer_color;
    const float* gbuffer_depth;
    const uint32_t* prev_reservoir_texture; /* uint2 */
    const float* prev_hit_pos_texture;      /* float4 */
    const void* prev_hit_radiance_texture;
    uint32_t* rw_reservoir_texture;         /* uint2 */
    float* rw_hit_pos_texture;              /* float4 */
    void* rw_hit_radiance_texture;
} vgi_ind_spec_temporal_io;
/* PushConstants of ind_spec_temporal_resample.hlsl:39-44, field for field (8 B). */
typedef struct vgi_pc_ind_spec_temporal {
    uint32_t frame_rand;
    uint32_t has_prev_frame;
} vgi_pc_ind_spec_temporal;
vgi_status vgi_pass_ind_spec_temporal_resample(const vgi_frame_params* fp, const vgi_ind_spec_temporal_io* io,
                                               const vgi_pc_ind_spec_temporal* pc,
                                               const vgi_resample_params* params, void* stream);

/* restir/ind_spec_spatial_resample.hlsl:23-35; restir_renderer.rs:770-784 */
typedef struct vgi_ind_spec_spatial_io {
    uint32_t width, height, storage;
    const uint32_t* gbuffer_color;
    const float* gbuffer_depth;
    const uint32_t* reservoir_texture;
    const float* hit_pos_texture;
    const void* hit_radiance_texture;
    void* rw_lighting_texture;
    void* rw_ray_len_texture; /* scalar */
} vgi_ind_spec_spatial_io;
typedef vgi_pc_frame_rand vgi_pc_ind_spec_spatial;
vgi_status vgi_pass_ind_spec_spatial_resample(const vgi_frame_params* fp, const vgi_ind_spec_spatial_io* io,
                                              const vgi_pc_ind_spec_spatial* pc,
                                              const vgi_resample_params* params, void* stream);

/* restir/hash_grid_decay.hlsl:5-6; restir_renderer.rs:322-329 */
typedef struct vgi_hash_grid_decay_io {
    uint8_t* rw_decay_buffer;                 /* 1 byte per cell */
    vgi_hash_grid_cell* rw_storage_buffer;
} vgi_hash_grid_decay_io;
vgi_status vgi_pass_hash_grid_decay(const vgi_hash_grid_decay_io* io, void* stream);

/* restir/hash_grid_query_select.hlsl:9-20; restir_renderer.rs:366-373 */
typedef struct vgi_hash_grid_query_select_io {
    const vgi_hash_grid_query* hash_grid_query_buffer;
    const uint32_t* hash_grid_query_counter_buffer;
    vgi_hash_grid_cell* rw_hash_grid_storage_buffer;
    uint32_t* rw_hash_grid_query_selection_buffer; /* 1 word per cell */
} vgi_hash_grid_query_select_io;
typedef struct vgi_pc_frame_hash {
    uint32_t frame_hash;
} vgi_pc_frame_hash;
vgi_status vgi_pass_hash_grid_query_select(const vgi_hash_grid_query_select_io* io, const vgi_pc_frame_hash* pc,
                                           void* stream);

/* restir/hash_grid_query_compact.hlsl:63-66; restir_renderer.rs:384-391 */
typedef struct vgi_hash_grid_query_compact_io {
    const uint32_t* hash_grid_query_selection_buffer;
    uint32_t* rw_compact_query_index_buffer;
    uint32_t* rw_compact_query_cell_addr_buffer;
    uint32_t* rw_compact_query_counter_buffer;
} vgi_hash_grid_query_compact_io;
vgi_status vgi_pass_hash_grid_query_compact(const vgi_hash_grid_query_compact_io* io, void* stream);

/* restir/hash_grid_raygen.hlsl:125-139; restir_renderer.rs:405-426.  The
 * trace resources' query buffer/counter are the *new* query buffer; the
 * trace reads rw_hash_grid_storage_buffer (as in the reference). */
typedef struct vgi_hash_grid_raygen_io {
    vgi_hash_grid_cell* rw_hash_grid_storage_buffer;
    const vgi_hash_grid_query* hash_grid_query_buffer;
    const uint32_t* compact_query_counter_buffer;
    const uint32_t* compact_query_index_buffer;
    const uint32_t* compact_query_cell_addr_buffer;
    uint8_t* rw_hash_grid_decay_buffer;
} vgi_hash_grid_raygen_io;
vgi_status vgi_pass_hash_grid_raygen(const vgi_frame_params* fp, const vgi_trace_resources* rt,
                                     const vgi_hash_grid_raygen_io* io, const vgi_pc_frame_hash* pc, void* stream);

/* Order-independent cache update.  No counterpart in the reference: there
 * (and in the three passes above) the slot a new cell takes, the query a
 * cell selects (a random key drawn from the query's index in the append
 * order), the order of the compacted list and the cache values the cache
 * rays read all depend on how the GPU schedules the lanes, so two runs of
 * the same frames drift apart (about 1e-2 relative mean after a few frames).
 * These three variants give the same result for every schedule, so a view
 * renders the same alone and inside a batch, run after run:
 *
 *   select   the keys to insert into a bucket go in one at a time, in
 *            ascending checksum order, through the reference's find-or-
 *            insert (the reference's result for that schedule); a cell
 *            selects the query with the largest key drawn from the query's
 *            contents (uniform over the cell's queries, like the reference's
 *            key drawn from the query's index), and the query's index goes
 *            to the winner buffer (identical queries are interchangeable);
 *   compact  the selected cells in cell order, indices from the winner
 *            buffer, appended after the counter's value on entry;
 *   raygen   every cache ray reads the cache as it was before the pass:
 *            blends go to a staging buffer (4 floats per compacted query)
 *            and are written into the cells after all rays have traced.
 *
 * At 1920x1080 the frame's trace passes append about ten times as many
 * queries as the buffer holds; the reference keeps the first
 * VGI_HASH_GRID_MAX_QUERIES to arrive (an order set by the schedule).  With
 * the threshold word of vgi_trace_resources the appends are a deterministic
 * sample of about VGI_HASH_GRID_QUERY_TARGET queries instead.
 *
 * Left to scheduling: which keys get in when more than 16 keys of one
 * bucket are new in one frame, and which queries survive when the sampled
 * appends still exceed VGI_HASH_GRID_MAX_QUERIES. */
typedef struct vgi_hash_grid_ordered_select_io {
    const vgi_hash_grid_query* hash_grid_query_buffer;
    const uint32_t* hash_grid_query_counter_buffer;
    vgi_hash_grid_cell* rw_hash_grid_storage_buffer;
    uint32_t* rw_hash_grid_query_selection_buffer; /* 1 word per cell, zero on entry */
    uint32_t* rw_hash_grid_query_winner_buffer;    /* 1 word per cell, zero on entry, 16-byte aligned: on
                                                    * return the index of the query a selected cell chose
                                                    * (and zero elsewhere) */
    /* NULL, or the threshold word of vgi_trace_resources: replaced by the
     * threshold for the queries appended next, from the counter of the
     * queries selected here (those appended under the word's old value) */
    uint32_t* rw_hash_grid_query_threshold;
} vgi_hash_grid_ordered_select_io;
/* The query threshold word for a frame that expects to append about
 * `expected_queries` queries (before any thinning); 0 (keep every query)
 * when they fit.  Queries are sampled down to VGI_HASH_GRID_QUERY_TARGET. */
#define VGI_HASH_GRID_QUERY_TARGET (VGI_HASH_GRID_MAX_QUERIES / 2u)
uint32_t vgi_hash_grid_query_threshold_for(uint64_t expected_queries);
vgi_status vgi_pass_hash_grid_query_select_ordered(const vgi_hash_grid_ordered_select_io* io,
                                                   const vgi_pc_frame_hash* pc, void* stream);

#define VGI_HASH_GRID_COMPACT_WORKSPACE_WORDS 257u
typedef struct vgi_hash_grid_ordered_compact_io {
    /* the ordered select's outputs, 16-byte aligned; the compaction sets
     * every word it lists back to zero, so both buffers hold zeros again
     * (as the next ordered select needs) without a clear */
    uint32_t* rw_hash_grid_query_selection_buffer;
    uint32_t* rw_hash_grid_query_winner_buffer;
    uint32_t* rw_compact_query_index_buffer;
    uint32_t* rw_compact_query_cell_addr_buffer;
    uint32_t* rw_compact_query_counter_buffer;
    uint32_t* rw_workspace_buffer; /* VGI_HASH_GRID_COMPACT_WORKSPACE_WORDS words */
} vgi_hash_grid_ordered_compact_io;
vgi_status vgi_pass_hash_grid_query_compact_ordered(const vgi_hash_grid_ordered_compact_io* io, void* stream);

typedef struct vgi_hash_grid_ordered_raygen_io {
    vgi_hash_grid_raygen_io raygen;    /* as for vgi_pass_hash_grid_raygen */
    float* rw_radiance_staging_buffer; /* 4 floats x VGI_HASH_GRID_MAX_QUERIES, 16-byte aligned */
} vgi_hash_grid_ordered_raygen_io;
vgi_status vgi_pass_hash_grid_raygen_ordered(const vgi_frame_params* fp, const vgi_trace_resources* rt,
                                             const vgi_hash_grid_ordered_raygen_io* io, const vgi_pc_frame_hash* pc,
                                             void* stream);

/* util/clear_buffer.hlsl:3-17; util_passes.rs:9-41 */
vgi_status vgi_pass_clear_buffer(uint32_t* buffer, uint32_t length, uint32_t value, void* stream);

/* Primary-visibility G-buffer producer (input side of the hot path; stands
 * in for the raster pass gbuffer_pass.rs:45-137 / mesh_gbuffer.hlsl:62-126):
 * one ray per pixel centre, reverse-Z depth of the hit, G-buffer layers
 * encoded as gbuffer.hlsl:54-77 with fwidth_z / fwidth_n from differences
 * inside each aligned 2x2 quad (a partner without geometry counts as 0). */
typedef struct vgi_gbuffer_io {
    uint32_t width, height;
    float* rw_depth;       /* f32, 0 = no geometry */
    uint32_t* rw_gbuffer;  /* 4 layers */
} vgi_gbuffer_io;
vgi_status vgi_pass_gbuffer_primary(const vgi_frame_params* fp, const vgi_scene* scene, const vgi_gbuffer_io* io,
                                    void* stream);

/* ------------------------------------------------------------------------
 * Direct light and composition, the passes right after the GI path
 * (restir_renderer.rs:831-911).
 * ---------------------------------------------------------------------- */
/* raytraced_shadow.hlsl:9-53 (restir_renderer.rs:831-849): one sun shadow
 * ray per pixel from the depth-reconstructed position (+2^-24 depth bias),
 * t in [0.0005, 1000], any hit.  R8_UNORM semantics: 255 = lit, 0 =
 * shadowed.  Pixels without geometry (depth 0) get 0; the reference leaves
 * them unwritten. */
typedef struct vgi_raytraced_shadow_io {
    uint32_t width, height;
    const float* gbuffer_depth;
    uint8_t* rw_shadow;
} vgi_raytraced_shadow_io;
vgi_status vgi_pass_raytraced_shadow(const vgi_frame_params* fp, const vgi_scene* scene,
                                     const vgi_raytraced_shadow_io* io, void* stream);

/* restir/final_lighting.hlsl:16-62 (restir_renderer.rs:893-911): sun light
 * through the shadow mask plus the demodulated indirect lighting
 * re-modulated by the diffuse albedo and the GGX integral
 * (DEMODULATE_INDIRECT_SPECULAR_FOR_DENOISER, IND_SPEC_R_MIN = 0.002, floor
 * 1/65536; restir/config.inc.hlsl).  Pixels without geometry get
 * sky_radiance (the sky pass's output, constant here).  Colours in the
 * storage mode's radiance layout. */
typedef struct vgi_final_lighting_io {
    uint32_t width, height, storage;
    const uint32_t* gbuffer_color;          /* 4 layers */
    const uint8_t* shadow_mask_buffer;
    const void* indirect_diffuse_texture;   /* radiance */
    const void* indirect_specular_texture;  /* radiance */
    void* rw_color_buffer;                  /* radiance */
    float sky_radiance[3];
} vgi_final_lighting_io;
vgi_status vgi_pass_final_lighting(const vgi_frame_params* fp, const vgi_final_lighting_io* io, void* stream);

/* temporal_aa.hlsl:112-264 (restir_renderer.rs:914-948, "Temporal AA"):
 * Mitchell-filtered 3x3 source around the un-jittered pixel centre,
 * reprojection of the pixel's own depth with FrameParams' current and
 * previous jitter, bilinear clamp-to-edge history fetch, neighbourhood
 * min/max clamp, variance clip (gamma 1) and the anti-flicker weighted blend
 * (1/32).  Without history, or where the history position leaves the
 * previous view, the output is the filtered source.  Images in the storage
 * mode's radiance layout. */
typedef struct vgi_temporal_aa_io {
    uint32_t width, height, storage;
    const float* gbuffer_depth;
    const void* source_texture;   /* radiance */
    const void* history_texture;  /* radiance; may be NULL when has_history = 0 */
    void* rw_target;              /* radiance */
} vgi_temporal_aa_io;
typedef struct vgi_pc_temporal_aa {
    uint32_t has_history;
} vgi_pc_temporal_aa;
vgi_status vgi_pass_temporal_aa(const vgi_frame_params* fp, const vgi_temporal_aa_io* io,
                                const vgi_pc_temporal_aa* pc, void* stream);

/* ------------------------------------------------------------------------
 * Renderer sequencer: crates/violet-restir/src/restir_renderer.rs
 * `RestirRenderer::add_passes` (:258-1028).  Records the whole GI pass chain
 * of one frame in violet's order, with violet's FrameRand seed consumption
 * (restir_renderer.rs:130-143) and history flags, on per-view persistent
 * HBM state (the RGTemporal resources of restir_renderer.rs:146-184).
 * ---------------------------------------------------------------------- */
typedef struct vgi_renderer vgi_renderer;

typedef struct vgi_renderer_config {
    uint32_t storage;                /* VGI_STORAGE_F32 / VGI_STORAGE_F16 */
    uint32_t denoise;                /* default 1 (restir_renderer.rs:66) */
    uint32_t ind_diff_validate;      /* default 1 */
    uint32_t ind_spec_validate;      /* default 1 */
    uint32_t hash_grid_cache_decay;  /* default 1 */
    float ao_radius;                 /* default 2.0 */
    float sky_radiance[3];           /* constant sky (replaces the sky cube) */
    /* resampling parameters; 0 = the reference's constants */
    uint32_t temporal_m_cap;         /* diffuse M clamp; specular uint(cap * roughness); 0 = 16 */
    uint32_t diff_spatial_taps;      /* diffuse spatial golden-spiral taps; 0 = 9, at most 16 */
    float spatial_radius_px;         /* spatial search radius; 0 = 0.05 * min(w, h) */
    /* default 1: add_passes runs independent pass chains of a frame on a
     * second stream owned by the renderer (hash-grid cache update beside the
     * G-buffer, specular beside diffuse resampling), forked from and joined
     * back into the caller's stream; 0 = every pass on the caller's stream */
    uint32_t concurrent_lanes;
    /* default 0: 1 = after the denoiser also run the sun shadow mask and the
     * final lighting (restir_renderer.rs:831-911), the lit image in `color` */
    uint32_t compose;
    /* default 0: 1 = temporal AA of the lit image (needs compose = 1;
     * violet's RestirConfig::taa, restir_renderer.rs:51-74).  The caller
     * jitters FrameParams (jitter.xy this frame, .zw the previous one) by
     * Halton(2,3) of (frame_index mod jitter_cycle) + 1, render_loop.rs:89-106. */
    uint32_t taa;
    vgi_denoiser_config denoiser;    /* default {1, 0, 3} */
    /* default 1: the cache update runs the order-independent passes
     * (vgi_pass_hash_grid_*_ordered), so a view renders the same frames run
     * after run and alone or in a batch; 0 = the reference's passes, whose
     * cache contents depend on the GPU's scheduling */
    uint32_t ordered_cache_update;
} vgi_renderer_config;

/* Outputs of one frame: device images owned by the renderer, valid until
 * the next add_passes on this renderer. */
typedef struct vgi_renderer_output {
    const void* indirect_diffuse;   /* denoised (or raw when the denoiser did not run) */
    const void* indirect_specular;
    const float* depth;
    const uint32_t* gbuffer;
    uint32_t denoised;              /* 1 if the denoiser ran this frame */
    uint32_t validation_frame;      /* 1 if this was a diffuse validation frame */
    uint32_t hash_grid_updated;     /* 1 if the cache update passes ran */
    const void* color;              /* the lit image (radiance layout) when compose = 1 (after temporal
                                     * AA when taa = 1), else NULL */
} vgi_renderer_output;

vgi_status vgi_renderer_default_config(vgi_renderer_config* cfg);
vgi_status vgi_renderer_create(uint32_t width, uint32_t height, const vgi_renderer_config* cfg, vgi_renderer** out);
vgi_status vgi_renderer_destroy(vgi_renderer* r);
/* Drops all history (the state after a reset/eviction: every has_prev_* = 0,
 * hash grid cleared on the next frame). */
vgi_status vgi_renderer_reset(vgi_renderer* r);
/* restir_renderer.rs:258-1028 (G-buffer from primary rays on `scene`). */
vgi_status vgi_renderer_add_passes(vgi_renderer* r, const vgi_scene* scene, const vgi_frame_params* fp,
                                   uint32_t frame_index, vgi_renderer_output* out, void* stream);
/* Named per-view images (violet's debug views / temporal resources), for
 * inspection, tests and state export: "depth", "gbuffer", "prev_depth",
 * "diff_reservoir", "diff_hit_pos_normal", "diff_hit_radiance" (spatial
 * output), "diff_temporal_reservoir", "diff_temporal_hit_pos_normal",
 * "diff_temporal_hit_radiance", "diff_new_hit_pos_normal",
 * "diff_new_hit_radiance", "ao", "ao_filtered", "indirect_diffuse_raw",
 * "spec_reservoir", "spec_hit_pos", "spec_hit_radiance",
 * "indirect_specular_raw", "spec_ray_len", "hash_grid_storage",
 * "hash_grid_decay", "hash_grid_queries", "hash_grid_query_counter",
 * "hash_grid_query_threshold" (ordered_cache_update = 1 only). */
vgi_status vgi_renderer_get_image(const vgi_renderer* r, const char* name, void** ptr, size_t* bytes);
/* Per-view state snapshot (checkpoint/resume, serving eviction): copies all
 * temporal state to/from a host blob of vgi_renderer_state_size() bytes. */
size_t vgi_renderer_state_size(const vgi_renderer* r);
/* Device memory held by the view (its state plus per-frame scratch images). */
size_t vgi_renderer_device_bytes(const vgi_renderer* r);
vgi_status vgi_renderer_export_state(const vgi_renderer* r, void* host_blob, size_t bytes, void* stream);
vgi_status vgi_renderer_import_state(vgi_renderer* r, const void* host_blob, size_t bytes, void* stream);

/* ------------------------------------------------------------------------
 * Batched views.  No reference counterpart: violet renders one view per
 * frame (restir_render_loop.rs:262-268).  Each view is a vgi_renderer with
 * its own reservoirs, denoiser history and radiance cache; a batched frame
 * advances every given view by one frame exactly as vgi_renderer_add_passes
 * would, but runs each pass of all views as ONE persistent launch (a fixed
 * grid walking the tiles of every view), optionally replayed from a
 * captured HIP graph.  All views of a call share the scene and the storage
 * mode; a view appears at most once per call.  The views are externally
 * synchronized like single-view renderers.
 * ------------------------------------------------------------------------ */
typedef struct vgi_batch vgi_batch;
/* Group counts: N > 0 launches N persistent groups that stride over the
 * tiles of all views; 0 launches one group per tile (the grid covers the
 * largest tile count seen for the pass). */
typedef struct vgi_batch_config {
    uint32_t max_views;      /* views per call, 1..64 */
    uint32_t pixel_groups;   /* 64-lane groups of the pixel and cache passes */
    uint32_t trace_groups;   /* 64-lane groups of the ray-tracing passes */
    uint32_t atrous_groups;  /* 256-lane groups of the a-trous and spatial-fallback passes */
    uint32_t use_graph;      /* 1: capture the frame once as a HIP graph and replay it */
    uint32_t concurrent_lanes; /* 1: independent pass chains (hash-grid update beside the
                                * G-buffer, specular beside diffuse) on a second stream */
} vgi_batch_config;
/* Defaults: 64 views, one group per tile everywhere, no graph, two lanes. */
vgi_status vgi_batch_default_config(vgi_batch_config* cfg);
vgi_status vgi_batch_create(const vgi_batch_config* cfg, vgi_batch** out);
vgi_status vgi_batch_destroy(vgi_batch* b);
/* One frame of `num_views` views: views[i] renders fps[i] as frame
 * frame_indices[i] and reports into outs[i].  Ordered on `stream`. */
vgi_status vgi_batch_add_passes(vgi_batch* b, const vgi_scene* scene, uint32_t num_views, vgi_renderer* const* views,
                                const vgi_frame_params* fps, const uint32_t* frame_indices, vgi_renderer_output* outs,
                                void* stream);

#ifdef __cplusplus
}
#endif

#endif /* VGI_H */
