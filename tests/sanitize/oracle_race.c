/* ThreadSanitizer driver for the oracle's OpenMP twin (FP32 build): two
 * frames of the G-buffer, diffuse and specular sample generation (parallel
 * pixel loops whose cache lookups read shared state and whose query appends
 * go through one atomic counter) and the denoiser passes (temporal filter
 * without and with history, disocclusion fix, spatial fallback, two a-trous
 * iterations) on a scene and cameras read from the file named by argv[1]:
 *
 *   u32 w, h, num_vertices, num_triangles, num_materials
 *   f32 frame_params[2][96]
 *   f32 positions[3 * num_vertices]
 *   u32 indices[3 * num_triangles], triangle_material[num_triangles]
 *   f32 materials[8 * num_materials]
 *
 * The oracle's restir sources are compiled into this file (their trace-
 * resource struct is private to them); the denoiser sources link beside it.
 * Exit status 0; a race report from ThreadSanitizer fails the run. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_restir.c"

void oracle_denoise_temporal(const float* fpf, int w, int h, const real* depth, const uint32_t* gb,
                             const real* diff_src, const real* spec_src, const real* ray_len, const real* prev_depth,
                             const uint32_t* prev_gb, const real* diff_hist, const real* spec_hist,
                             const real* mom_hist, const uint8_t* len_hist, const real* diff_fast_hist,
                             const real* spec_fast_hist, int has_history, int fast_history, real* o_diff,
                             real* o_spec, real* o_mom, uint8_t* o_len, real* o_diff_fast, real* o_spec_fast);
void oracle_denoise_disocclusion(int w, int h, const real* depth, const uint32_t* gb, const uint8_t* len_tex,
                                 const real* diff, const real* spec, const real* mom, real* o_diff, real* o_spec,
                                 real* o_mom);
void oracle_denoise_fallback(int w, int h, const real* depth, const uint32_t* gb, const uint8_t* len_tex,
                             const real* diff, const real* spec, const real* mom, real* o_diff, real* o_spec,
                             real* o_var);
int oracle_num_threads(void);
void oracle_denoise_atrous(int w, int h, int step, const real* depth, const uint32_t* gb, const real* diff,
                           const real* spec, const real* var, real* o_diff, real* o_spec, real* o_var);

static void* need(size_t n) {
    void* p = calloc(n ? n : 1, 1);
    if (!p) {
        fprintf(stderr, "oracle_race: out of memory\n");
        exit(2);
    }
    return p;
}

static void take(FILE* f, void* dst, size_t n) {
    if (fread(dst, 1, n, f) != n) {
        fprintf(stderr, "oracle_race: short input file\n");
        exit(2);
    }
}

int main(int argc, char** argv) {
    if (argc != 2) return 2;
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    uint32_t hd[5];
    take(f, hd, sizeof(hd));
    const int w = (int)hd[0], h = (int)hd[1];
    const uint32_t nv = hd[2], nt = hd[3], nm = hd[4];
    float fp[2][96];
    take(f, fp, sizeof(fp));
    float* P = need(12 * (size_t)nv);
    uint32_t* I = need(12 * (size_t)nt);
    uint32_t* TM = need(4 * (size_t)nt);
    float* mats = need(32 * (size_t)nm);
    take(f, P, 12 * (size_t)nv);
    take(f, I, 12 * (size_t)nt);
    take(f, TM, 4 * (size_t)nt);
    take(f, mats, 32 * (size_t)nm);
    fclose(f);

    const size_t n = (size_t)w * h;
    oscene* s = oracle_scene_create(P, NULL, I, nt, TM, mats);
    real* depth[2];
    uint32_t* gb[2];
    for (int k = 0; k < 2; ++k) {
        depth[k] = need(sizeof(real) * n);
        gb[k] = need(16 * n);
        oracle_gbuffer(fp[k], s, w, h, depth[k], gb[k]);
    }
    uint32_t* cells = need(16u * 65536u * 16u);
    uint32_t* queries = need(24u * 131072u);
    uint32_t counter = 0;
    otrace_res tr;
    memset(&tr, 0, sizeof(tr));
    tr.scene = s;
    tr.sky[0] = tr.sky[1] = tr.sky[2] = 1.0f;
    tr.hg_cells = cells;
    tr.queries = queries;
    tr.query_counter = &counter;
    uint32_t* hit = need(8 * n);
    real* diff = need(3 * sizeof(real) * n);
    real* spec = need(3 * sizeof(real) * n);
    real* hitpos = need(4 * sizeof(real) * n);
    real* raylen = need(sizeof(real) * n);
    /* frame 0: no history; frame 1: the first frame's diffuse as the screen cache */
    oracle_diff_sample_gen(fp[0], &tr, w, h, gb[0], depth[0], 11u, 0, hit, diff);
    oracle_spec_sample_gen(fp[0], &tr, w, h, gb[0], depth[0], 12u, 0, hitpos, spec);
    real* prev_diff = need(3 * sizeof(real) * n);
    memcpy(prev_diff, diff, 3 * sizeof(real) * n);
    tr.prev_ind_diff = prev_diff;
    tr.prev_depth = depth[0];
    tr.pw = w;
    tr.ph = h;
    oracle_diff_sample_gen(fp[1], &tr, w, h, gb[1], depth[1], 21u, 1, hit, diff);
    oracle_spec_sample_gen(fp[1], &tr, w, h, gb[1], depth[1], 22u, 1, hitpos, spec);
    for (size_t i = 0; i < n; ++i) raylen[i] = hitpos[4 * i + 3];

    real *hd_[2], *hs_[2], *hm_[2];
    uint8_t* hl_[2];
    for (int k = 0; k < 2; ++k) {
        hd_[k] = need(3 * sizeof(real) * n);
        hs_[k] = need(3 * sizeof(real) * n);
        hm_[k] = need(2 * sizeof(real) * n);
        hl_[k] = need(n);
    }
    real* df = need(3 * sizeof(real) * n);
    real* sf = need(3 * sizeof(real) * n);
    oracle_denoise_temporal(fp[0], w, h, depth[0], gb[0], prev_diff, spec, raylen, NULL, NULL, NULL, NULL, NULL,
                            NULL, NULL, NULL, 0, 0, hd_[0], hs_[0], hm_[0], hl_[0], df, sf);
    oracle_denoise_temporal(fp[1], w, h, depth[1], gb[1], diff, spec, raylen, depth[0], gb[0], hd_[0], hs_[0],
                            hm_[0], hl_[0], NULL, NULL, 1, 0, hd_[1], hs_[1], hm_[1], hl_[1], df, sf);
    real* fd = need(3 * sizeof(real) * n);
    real* fs = need(3 * sizeof(real) * n);
    real* fm = need(2 * sizeof(real) * n);
    oracle_denoise_disocclusion(w, h, depth[1], gb[1], hl_[1], hd_[1], hs_[1], hm_[1], fd, fs, fm);
    real* var = need(2 * sizeof(real) * n);
    oracle_denoise_fallback(w, h, depth[1], gb[1], hl_[1], fd, fs, fm, hd_[0], hs_[0], var);
    real* var2 = need(2 * sizeof(real) * n);
    oracle_denoise_atrous(w, h, 1, depth[1], gb[1], hd_[0], hs_[0], var, fd, fs, var2);
    oracle_denoise_atrous(w, h, 2, depth[1], gb[1], fd, fs, var2, hd_[0], hs_[0], var);

    double sum = 0.0;
    for (size_t i = 0; i < 3 * n; ++i) sum += (double)hd_[0][i] + (double)hs_[0][i];
    printf("oracle_race: %dx%d, %u triangles, %u queries, %d threads, output sum %.6g\n", w, h, nt, counter,
           oracle_num_threads(), sum);
    oracle_scene_destroy(s);
    return 0;
}
