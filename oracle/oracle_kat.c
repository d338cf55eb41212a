/*
 * oracle_kat.c — ORACLE (test infrastructure only).  Exported wrappers of
 * the integer streams and encodings restated in oracle_common.h, so tests
 * can pin them against known-answer vectors and an independent numpy
 * restatement (rand.hlsl, enc.inc.hlsl, gbuffer.hlsl, restir/reservoir.hlsl).
 */
#include "oracle_common.h"

uint32_t oracle_jenkins(uint32_t x) { return o_jenkins(x); }
uint32_t oracle_lcg_init(uint32_t px, uint32_t py, uint32_t seed) { return o_lcg_init(px, py, seed); }
void oracle_lcg_stream(uint32_t state, int n, uint32_t* out_states, real* out_rand) {
    for (int i = 0; i < n; ++i) {
        out_rand[i] = o_rand(&state);
        out_states[i] = state;
    }
}
uint32_t oracle_xxh(const uint32_t* words, int n) {
    uint32_t h = o_xxh_init(words[0]);
    for (int i = 1; i < n; ++i) h = o_xxh_add(h, words[i]);
    return o_xxh_eval(h);
}
uint32_t oracle_f32_to_f16(float f) { return o_double_to_half((double)f); }
double oracle_f16_to_f64(uint32_t h) { return o_half_to_double(h); }
void oracle_oct_dec_u16(uint32_t e, real* n) { v3 v = o_oct_dec_u16(e); n[0] = v.x; n[1] = v.y; n[2] = v.z; }
uint32_t oracle_oct_enc_u16(const real* n) { return o_oct_enc_u16(V3(n[0], n[1], n[2])); }
void oracle_oct_dec_u32(uint32_t e, real* n) { v3 v = o_oct_dec_u32(e); n[0] = v.x; n[1] = v.y; n[2] = v.z; }
uint32_t oracle_oct_enc_u32(const real* n) { return o_oct_enc_u32(V3(n[0], n[1], n[2])); }
void oracle_gb_normal(uint32_t l1, real* n) { v3 v = o_gb_normal(l1); n[0] = v.x; n[1] = v.y; n[2] = v.z; }
uint32_t oracle_rsv_enc(uint32_t M, float W) { return o_rsv_enc(M, W); }
void oracle_rsv_dec(uint32_t u, uint32_t* M, float* W) { o_rsv_dec(u, M, W); }
void oracle_sincos(real x, real* s, real* c) { r_sincos(x, s, c); }
real oracle_log2(real x) { return R_LOG2(x); }
int oracle_real_bytes(void) { return (int)sizeof(real); }
#ifdef _OPENMP
#include <omp.h>
int oracle_num_threads(void) { return omp_get_max_threads(); }
#else
int oracle_num_threads(void) { return 1; }
#endif
