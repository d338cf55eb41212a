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
} vgi_profile_entry;
vgi_status vgi_profiler_enable(uint32_t enable);
vgi_status vgi_profiler_reset(void);
int32_t vgi_profiler_read(vgi_profile_entry* out, int32_t cap);

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
 * Scene and software ray tracing.
 * Replaces the BLAS/TLAS build of crates/violet/src/render_scene.rs:
 * 1170-1210 (BLAS per geometry group, UINT16 indices, f32x3 positions) and
 * :1557-1712 (TLAS over instances) with a host-built binary BVH in HBM, and
 * TraceRay (restir/raytrace.inc.hlsl:42-53,74-82) with HIP traversal.
 * Instances are baked to world space by the caller; indices are u32.
 * ---------------------------------------------------------------------- */
typedef struct vgi_scene vgi_scene;

/* MaterialParams factors (shader/scene_bindings.hlsl:24-51); textures are
 * not sampled (untextured scenes: every texture fetch returns 1). */
typedef struct vgi_material {
    float base_color[4];
    float metallic;
    float roughness; /* perceptual roughness */
    float pad[2];
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
} vgi_scene_desc;

typedef struct vgi_scene_stats {
    uint32_t num_triangles, num_vertices, num_nodes, depth;
    uint64_t device_bytes;
} vgi_scene_stats;

/* Host arrays in, BVH built on the host, uploaded to the current device. */
vgi_status vgi_scene_create(const vgi_scene_desc* desc, vgi_scene** out);
vgi_status vgi_scene_destroy(vgi_scene* s);
vgi_status vgi_scene_info(const vgi_scene* s, vgi_scene_stats* out);

/* Batch ray query (TraceRay without shading).  rays: 8 floats per ray
 * {o.x,o.y,o.z,tmin, d.x,d.y,d.z,tmax} in device memory; hits: 4 words per
 * ray {t (f32), u (f32), v (f32), triangle id (u32, 0xFFFFFFFF = miss)}.
 * any_hit != 0 gives shadow-ray semantics (first accepted hit, t and
 * barycentrics unspecified, id = some hit triangle). */
vgi_status vgi_trace_rays(const vgi_scene* scene, const float* rays, uint32_t num_rays, uint32_t any_hit,
                          uint32_t* hits, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* VGI_H */
