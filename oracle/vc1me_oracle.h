/*
 * vc1me_oracle.h — CPU restatement of libvc1's P-frame motion estimation.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker.  The product (libvc1_amd) never links, loads or falls back to it.
 *
 * Types are those of the product ABI (include/vc1me.h) so results compare
 * field by field.
 */
#ifndef VC1ME_ORACLE_H
#define VC1ME_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/vc1me.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Algorithmic operation counters (units of the VALU roofline):
 *   sad_px      one |a-b| accumulate per sampled pixel
 *   interp_taps one 4-tap filter application (separable 2D counts both passes)
 *   satd_ops    112 per 4x4 Hadamard tile
 *   sig_ops     one signature comparison per long-range index cell scanned
 *   frame_px    one |a-b| accumulate per sampled pixel of a frame SAD */
typedef struct oracle_ops {
    uint64_t sad_px;
    uint64_t interp_taps;
    uint64_t satd_ops;
    uint64_t sig_ops;
    uint64_t frame_px;
} oracle_ops;

/* Whole-frame P analysis: Vc1Encoder::analyze_p_picture_core(cur, ref, mode,
 * allow_intra, nullptr, nullptr), src/encoder_motion.cpp:47-442.
 * Returns 0, or -1 for an unsupported configuration (me_quality SAD/RD). */
int oracle_analyze_p_frame(const vc1me_params* p, const vc1me_planes* cur,
                           const vc1me_planes* ref, vc1me_p_result* out, oracle_ops* ops);

/* The seeded core: analyze_p_picture_core(cur, ref, mode, allow_intra,
 * &seed, nullptr) (src/encoder_motion.cpp:97-101, :120-130, :432-434): the
 * seed's 1-MV vectors start each macroblock's staged refinement and its
 * distant-match fields are copied.  global_* of the result are zero. */
int oracle_analyze_p_frame_seeded(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                                  const vc1me_p_result* seed, vc1me_p_result* out, oracle_ops* ops);

/* Batch scorer semantics of VulkanComputeContext::score_motion_batch as the
 * GLSL kernel computes it (src/shaders/vulkan_motion_cost.comp:109-183):
 * clamped fetch, sampled SAD or SATD4x4, no hidden-padding veto. */
void oracle_score_batch(const uint8_t* cur, const uint8_t* ref, int width, int height,
                        int x0, int y0, int bw, int bh, int sample_stride, int bilinear,
                        int rnd, int metric, const vc1me_candidate* c, size_t n,
                        uint64_t* costs);

/* Primitives exposed for unit tests (same laws as the reference). */
int oracle_luma_mc_sample(const uint8_t* plane, int w, int h, int x, int y, int xq, int yq,
                          int mode, int rnd);
uint64_t oracle_mb_sad(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                       int mx, int my, int dx, int dy, int stride);
uint64_t oracle_block_cost(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                           int x0, int y0, int bw, int bh, int xq, int yq, int mode, int metric);
uint64_t oracle_frame_sad(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                          int dx, int dy, int stride);
void oracle_global_motion(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                          int range, int32_t* dx, int32_t* dy, uint64_t* sad);
/* MV rate in bits: motion_vector_rate_bits (src/encoder_motion.cpp:3214-3225). */
/* 20 * (SAD + lambda * R) per macroblock of a finished analysis at the given
 * 1-MV vectors (mbs x 2), with the result's block vectors (mbs x 4 x 2) and
 * intra flags as the predictor field. */
void oracle_mb_costs20(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                       const int32_t* one_mvs, const int32_t* block_mvs, const uint8_t* use_intra, uint64_t* out);
int oracle_mv_rate_bits(const vc1me_params* p, int dxq, int dyq, int pxq, int pyq, int mode, int more);
/* x87 long-double score compare used by the staged sub-pel finalists:
 * returns -1/0/1 for (sad_a + lambda*rate_a) <=> (sad_b + lambda*rate_b)
 * evaluated exactly as the reference does (src/encoder_motion.cpp:3465-3471). */
int oracle_score_cmp(int pqindex, uint64_t sad_a, uint64_t rate_a, uint64_t sad_b, uint64_t rate_b);
/* motion_uses_hidden_padding (src/encoder_motion.cpp:2191-2195) */
int oracle_hidden_padding(const vc1me_params* p, int x0, int y0, int bw, int bh, int xq, int yq, int mode, int chroma);
/* distant_match_decision (src/encoder_motion.cpp:1642-1656) */
int oracle_distant_decision(const vc1me_params* p, uint64_t cand, uint64_t local, int mx, int my);
/* integer_motion_local with a bare predictor (pre only, A/C invalid) and a
 * given global vector (src/encoder_motion_diamond.cpp:5-11). */
void oracle_integer_search(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref, int mx, int my,
                           int pre_xq, int pre_yq, int gdx, int gdy, int range, int32_t* dx, int32_t* dy,
                           uint64_t* sad);

/* intensity_map_y / intensity_map_uv (src/encoder_motion_compensation.cpp:175-193) */
uint8_t oracle_intensity_map_y(uint8_t v, const vc1me_intensity* ic);
uint8_t oracle_intensity_map_uv(uint8_t v, const vc1me_intensity* ic);
/* intensity_compensated_reference (src/encoder_motion_compensation.cpp:195-294) */
void oracle_intensity_compensate(const vc1me_planes* ref, const vc1me_intensity* ic, uint8_t* y, uint8_t* u,
                                 uint8_t* v);
/* estimate_intensity_comp (src/encoder_motion_compensation.cpp:296-348); mvs:
 * mbw * mbh quarter-pel (x, y) pairs */
int oracle_estimate_intensity_comp(const vc1me_planes* cur, const vc1me_planes* ref, const int32_t* mvs,
                                   vc1me_intensity* out);

/* analyze_b_picture_mode (src/encoder_motion.cpp:753-1004) for the 1-MV mode
 * p->mode; -1 for an unsupported fraction or mode */
int oracle_analyze_b_frame(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* past,
                           const vc1me_planes* future, const vc1me_b_params* b, vc1me_b_result* out);

/* scene_change_score (src/encoder_motion.cpp:599-676) under the ME settings
 * p, fade_compensation and scene_threshold of the encoder */
int oracle_scene_change_score(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                              int fade_compensation, double scene_threshold, double* out);

/* has_good_motion_match (src/encoder_motion.cpp:678-751): *out = 1 when some
 * macroblock has a usable motion match; field_out (or NULL): every
 * macroblock's refined vector, the scan continuing past the first match */
int oracle_has_good_motion_match(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                                 int fade_compensation, int* out, int32_t* field_out);

/* refine_qpel_block (src/encoder_motion.cpp:3596-3685) on one block */
int oracle_refine_qpel_block(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref, int x0, int y0,
                             int bw, int bh, int seed_xq, int seed_yq, int range, uint64_t known_seed_sad,
                             int32_t* xq, int32_t* yq, uint64_t* sad);

/* the P prediction picture of an analysis (motion_compensate_mb /
 * motion_compensate_4mv, src/encoder_motion_compensation.cpp:474-571) */
void oracle_predict_p_frame(const vc1me_planes* ref, int mode, const int32_t* mvs, const int32_t* block_mvs,
                            const uint8_t* use_4mv, uint8_t* y, uint8_t* u, uint8_t* v);

/* the bi-directional prediction picture (motion_compensate_bi_mb,
 * src/encoder_motion_compensation.cpp:659-711, every macroblock) */
void oracle_predict_bi_frame(const vc1me_planes* past, const vc1me_planes* future, int mode, const int32_t* fmvs,
                             const int32_t* bmvs, uint8_t* y, uint8_t* u, uint8_t* v);

/* the padded P prediction picture (motion_compensate_mb_padded /
 * motion_compensate_4mv_padded, src/encoder_motion_compensation.cpp:573-657):
 * an mbw*16 x mbh*16 picture, samples clamped to ref's own raster */
void oracle_predict_p_frame_padded(const vc1me_planes* ref, int mbw, int mbh, int mode, const int32_t* mvs,
                                   const int32_t* block_mvs, const uint8_t* use_4mv, uint8_t* y, uint8_t* u,
                                   uint8_t* v);

/* the padded bi-directional picture (motion_compensate_bi_mb_padded,
 * src/encoder_motion_compensation.cpp:712-763) */
void oracle_predict_bi_frame_padded(const vc1me_planes* past, const vc1me_planes* future, int mbw, int mbh, int mode,
                                    const int32_t* fmvs, const int32_t* bmvs, uint8_t* y, uint8_t* u, uint8_t* v);

#ifdef __cplusplus
}
#endif

#endif
