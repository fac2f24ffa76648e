/*
 * vc1me.h — C ABI of the MI355X (gfx950) motion-estimation backend for libvc1.
 *
 * The backend occupies the encoder's motion-cost compute slot
 * (libvc1 src/encoder_internal.h:286-289, EncoderConfig::vulkan_compute and
 * vulkan_min_batch*) that the experimental Vulkan backend fills today, and adds
 * a whole-frame entry point in front of Vc1Encoder::analyze_p_picture_core
 * (src/encoder_motion.cpp:47-442).  Only plain C types cross this boundary:
 * raw pointers, sizes and POD structs.  No C++ or torch types.
 *
 * Status convention: every int-returning entry point returns VC1ME_OK (0) on
 * success and a negative VC1ME_E* code on failure.  A failure is recoverable:
 * the caller re-scores on the CPU exactly as the reference does when
 * VulkanComputeContext::score_motion_batch returns false
 * (src/vulkan_compute.cpp:486-489, callers e.g. src/encoder_motion.cpp:1938-1942).
 * The message of the last failure on the calling thread is returned by
 * vc1me_last_error() (thread-local, like vc1_encoder_last_error,
 * include/libvc1.h:653).
 *
 * Thread safety: one context may be shared by several host threads; every
 * call that touches the device is serialised by a per-context mutex
 * (the reference serialises on one mutex too, src/vulkan_compute.cpp:385).
 */
#ifndef VC1ME_H
#define VC1ME_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VC1ME_ABI_VERSION 1

enum {
    VC1ME_OK = 0,
    VC1ME_EINVAL = -1,     /* bad argument (null pointer, size out of range) */
    VC1ME_EDEVICE = -2,    /* HIP runtime / device failure */
    VC1ME_ENOMEM = -3,     /* device allocation failed */
    VC1ME_ELIMIT = -4,     /* request exceeds a documented limit */
    VC1ME_EINJECTED = -5,  /* failure forced by VC1ME_FAIL_AT (fault injection) */
    VC1ME_EKERNEL = -6     /* a kernel reported an internal error flag */
};

/* Candidate motion vector in quarter-pel units; mirrors VulkanMotionCandidate
 * (src/vulkan_compute.h:11-14). */
typedef struct vc1me_candidate {
    int32_t xq;
    int32_t yq;
} vc1me_candidate;

/* Cost metric; mirrors VulkanMotionMetric {Sad=0,Satd=1} (src/vulkan_compute.h:16). */
enum { VC1ME_METRIC_SAD = 0, VC1ME_METRIC_SATD = 1 };

/* Progressive MV modes; mirrors Vc1Encoder::ProgressiveMvMode
 * (src/encoder_internal.h:338). */
enum {
    VC1ME_MODE_ONE_MV_QPEL = 0,
    VC1ME_MODE_ONE_MV_HPEL = 1,
    VC1ME_MODE_ONE_MV_HPEL_BILINEAR = 2,
    VC1ME_MODE_MIXED_MV = 3
};

/* Motion-estimation parameters: the EncoderConfig fields the P analysis reads
 * (src/encoder_internal.h:223-315; ME fields :246-253, :289-302) plus the two
 * call arguments of analyze_p_picture_core (mode, allow_intra). */
typedef struct vc1me_params {
    int32_t width;                      /* coded even raster width  (EncoderConfig::width) */
    int32_t height;                     /* coded even raster height (EncoderConfig::height) */
    int32_t display_width;              /* 0 = width  (EncoderConfig::display_width) */
    int32_t display_height;             /* 0 = height (EncoderConfig::display_height) */
    int32_t pqindex;                    /* 1..31 */
    int32_t motion_search_range;        /* 0..1024 px */
    int32_t motion_local_search_range;  /* px, default 32 */
    int32_t me_algorithm;               /* 0 = UMH, 1 = diamond (vc1_me_algorithm_e) */
    int32_t me_quality;                 /* 0 = SAD, 1 = RATE, 2 = SATD (RD is not offered) */
    int32_t subpixel_refinement;        /* bool */
    int32_t fast_subpixel_refinement;   /* bool */
    int32_t subpixel_refinement_range;  /* px, default 1024 */
    int32_t long_range_search_mode;     /* 0 compare, 1 local-good skip, 2 legacy distant-first */
    int32_t mode;                       /* VC1ME_MODE_* */
    int32_t allow_intra;                /* bool */
    int32_t debug_disable_p_intra;      /* bool */
    double distant_match_max_mae;       /* default 255.0 */
    double inter_intra_threshold;       /* default 0.80 */
} vc1me_params;

/* Fill p with the reference's defaults for a width x height raster
 * (src/encoder_internal.h:243-253, :295, :289-302). */
void vc1me_params_default(vc1me_params* p, int32_t width, int32_t height);

/* Per-frame P analysis output; the fields of Vc1Encoder::PAnalysis that ME
 * produces (src/encoder_internal.h:346-370).  Arrays are caller-owned and
 * hold one entry per macroblock (mbw*mbh, raster order) unless noted. */
typedef struct vc1me_p_result {
    int32_t* mvs;                    /* [mbs][2]  decoder block-0 / 1-MV vector, qpel */
    int32_t* block_mvs;              /* [mbs][4][2] four luma block vectors, qpel */
    uint8_t* use_4mv;                /* [mbs] */
    uint8_t* use_intra;              /* [mbs] */
    uint64_t* screen_sads;           /* [mbs] */
    double* inter_intra_cost_ratio;  /* [mbs]  -1 when unavailable */
    double* distant_local_mae;       /* [mbs]  -1 when unavailable */
    double* distant_candidate_mae;   /* [mbs]  -1 when unavailable */
    uint8_t* distant_match_decision; /* [mbs]  vc1_distant_match_decision_e */
    /* scalar outputs */
    double mean_abs_residual;
    double rate_complexity;
    uint64_t motion_sad;
    uint64_t moved_macroblocks;
    uint64_t fractional_chroma_macroblocks;
    uint64_t four_mv_macroblocks;
    uint64_t intra_macroblocks;
    /* diagnostics (not in PAnalysis): the frame translation found by the
     * global search (src/encoder_motion.cpp:93-101) */
    int32_t global_dx;
    int32_t global_dy;
    uint64_t global_sad;
} vc1me_p_result;

/* One 8-bit 4:2:0 picture on the host: tight planes, stride = width
 * (libvc1 Frame, src/encoder_internal.h:67-70). */
typedef struct vc1me_planes {
    const uint8_t* y;
    const uint8_t* u;
    const uint8_t* v;
    int32_t width;
    int32_t height;
} vc1me_planes;

/* Per-context counters; mirrors VulkanComputeCounters (src/vulkan_compute.h:18-24)
 * and adds whole-frame work. */
typedef struct vc1me_counters {
    uint64_t dispatches;
    uint64_t candidates;
    uint64_t frame_uploads;
    uint64_t frame_upload_bytes;
    uint64_t submit_wait_nanoseconds;
    uint64_t frames_analyzed;
    uint64_t macroblocks_analyzed;
} vc1me_counters;

typedef struct vc1me_ctx vc1me_ctx;
typedef struct vc1me_frame vc1me_frame;

/* ---- lifecycle (replaces VulkanComputeContext::create, src/vulkan_compute.cpp:493-505) */
int vc1me_create(int device, vc1me_ctx** out);
void vc1me_destroy(vc1me_ctx* ctx);
const char* vc1me_last_error(void);
int vc1me_abi_version(void);
/* Device name of the context (VulkanComputeContext::device_name, src/vulkan_compute.h:33). */
const char* vc1me_device_name(const vc1me_ctx* ctx);

/* ---- (i) batch scorer: exactly VulkanComputeContext::score_motion_batch
 * (src/vulkan_compute.h:43-51, src/vulkan_compute.cpp:370-490) with a raw
 * costs array of n entries.  Planes are luma only, tight (stride = width).
 * A null plane, non-positive size or n == 0 succeeds with nothing written
 * (src/vulkan_compute.cpp:376-379).  Limits: n <= 4,194,240 and
 * width*height <= INT32_MAX (src/vulkan_compute.cpp:380-383).
 * Host planes are uploaded once per (pointer, size, content checksum). */
int vc1me_score_batch(vc1me_ctx* ctx, const uint8_t* cur, const uint8_t* ref,
                      int width, int height, int x0, int y0, int bw, int bh,
                      int sample_stride, int bilinear, int rnd, int metric,
                      const vc1me_candidate* candidates, size_t n,
                      uint64_t* costs_out);

/* vc1me_score_batch with the device copies of the planes keyed by a
 * generation number the caller keeps per plane instead of a checksum: no pass
 * over the planes per call (the checksum of a 1080p plane costs about as
 * much as scoring a batch).  The caller must change a plane's generation
 * whenever it writes the plane (the reference's pointer-and-fingerprint
 * cache, src/vulkan_compute.cpp:194-231, misses such writes when the sampled
 * bytes stay the same; a generation does not).  An encoder can pass the
 * picture's coding index.  No other difference. */
int vc1me_score_batch_gen(vc1me_ctx* ctx, const uint8_t* cur, uint64_t cur_generation,
                          const uint8_t* ref, uint64_t ref_generation,
                          int width, int height, int x0, int y0, int bw, int bh,
                          int sample_stride, int bilinear, int rnd, int metric,
                          const vc1me_candidate* candidates, size_t n,
                          uint64_t* costs_out);

/* ---- frames resident in HBM (explicit handles instead of the reference's
 * pointer+fingerprint cache, src/vulkan_compute.cpp:194-231).  A handle may
 * be used in the jobs of every context of the same device (so that several
 * contexts keep steps in flight over one set of frames); the context that
 * uploaded it releases it, once no step that reads it is running. */
int vc1me_frame_upload(vc1me_ctx* ctx, const vc1me_planes* planes, vc1me_frame** out);
void vc1me_frame_release(vc1me_ctx* ctx, vc1me_frame* frame);

/* Same contract as vc1me_score_batch on resident frames. */
int vc1me_score_batch_frames(vc1me_ctx* ctx, const vc1me_frame* cur, const vc1me_frame* ref,
                             int x0, int y0, int bw, int bh, int sample_stride,
                             int bilinear, int rnd, int metric,
                             const vc1me_candidate* candidates, size_t n,
                             uint64_t* costs_out);

/* Same contract as vc1me_score_batch_frames, but the call allocates nothing:
 * no hipMalloc, no hipFree, no allocating library routine.  No counterpart
 * in the reference; it is for callers that cannot allocate device memory
 * (code around HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The candidate list and its 32-bit costs are staged in a
 * fixed scratch array of the library (768 KiB per GPU, one copy per GPU,
 * held until the process ends), so n is limited to
 * VC1ME_NOALLOC_MAX_CANDIDATES (VC1ME_ELIMIT above it, checked before any
 * device work).  Every call on one GPU shares that GPU's array, which keeps
 * the last call's contents (each call overwrites all it reads); the library
 * serialises such calls with a per-GPU lock, so calls that would overlap on
 * one GPU wait for each other instead of racing. */
#define VC1ME_NOALLOC_MAX_CANDIDATES 65536
int vc1me_score_batch_frames_noalloc(vc1me_ctx* ctx, const vc1me_frame* cur, const vc1me_frame* ref,
                                     int x0, int y0, int bw, int bh, int sample_stride,
                                     int bilinear, int rnd, int metric,
                                     const vc1me_candidate* candidates, size_t n,
                                     uint64_t* costs_out);

/* ---- (ii) whole-frame plain P motion estimation: the behaviour of
 * Vc1Encoder::analyze_p_picture_core(cur, ref, mode, allow_intra,
 * shared_seed=nullptr, b_reuse=nullptr) (src/encoder_motion.cpp:47-442).
 * Results are bit-exact with the reference CPU path. */
int vc1me_analyze_p_frame(vc1me_ctx* ctx, const vc1me_frame* cur, const vc1me_frame* ref,
                          const vc1me_params* params, vc1me_p_result* out);

/* Costs of n candidates for one 16x16 or 8x8 block of resident frames under
 * the search's own laws: metric 0 = block_sad_motion (hidden-padding veto
 * included, src/encoder_motion.cpp:2670-2933), metric 1 = block_satd_motion
 * (src/encoder_motion.cpp:2962-3042), for progressive mode `mode`.  Runs the
 * same device code as the sub-pel stages of vc1me_analyze_p_frame.  No
 * reference counterpart (a finer-grained form of vc1me_score_batch). */
int vc1me_block_costs(vc1me_ctx* ctx, const vc1me_frame* cur, const vc1me_frame* ref, const vc1me_params* params,
                      int x0, int y0, int bw, int bh, int mode, int metric, const vc1me_candidate* candidates,
                      size_t n, uint64_t* costs_out);

/* ---- (iii) batched multi-stream step: n independent (cur, ref) pairs of
 * possibly different rasters analysed in one launch sequence on the
 * context's stream.  Equivalent to n calls of vc1me_analyze_p_frame. */
typedef struct vc1me_job {
    const vc1me_frame* cur;
    const vc1me_frame* ref;
    const vc1me_params* params;
    vc1me_p_result* out;      /* may be NULL when results are not wanted on the host */
} vc1me_job;
int vc1me_analyze_step(vc1me_ctx* ctx, const vc1me_job* jobs, int n);

/* ---- seeded core: Vc1Encoder::analyze_p_picture_core with a shared seed
 * (src/encoder_motion.cpp:47-442; the seeded branch at :97-101 and :120-130,
 * the global MAD at :432-434).  Each macroblock's staged quarter-pel
 * refinement starts from the seed's 1-MV vector (seed->mvs, clamped to the
 * search bounds) instead of the global and integer search; the seed's
 * distant_local_mae, distant_candidate_mae and distant_match_decision are
 * copied into the result; global_dx/dy/sad are 0 and the global MAD is the
 * local one.  4-MV, intra and the summary run as in vc1me_analyze_p_frame.
 * The seed is normally the 1-MV quarter-pel result of the same pair (the
 * reference's shared ME pass of an I/P/B group); its arrays must all be set
 * (VC1ME_EINVAL otherwise).  A motion_search_range of 0 ignores the seed. */
int vc1me_analyze_p_frame_seeded(vc1me_ctx* ctx, const vc1me_frame* cur, const vc1me_frame* ref,
                                 const vc1me_params* params, const vc1me_p_result* seed, vc1me_p_result* out);
/* n seeded pairs in one step: seeds[i] seeds jobs[i]. */
int vc1me_analyze_step_seeded(vc1me_ctx* ctx, const vc1me_job* jobs, const vc1me_p_result* const* seeds, int n);

/* ---- intensity (fade) compensation of the P analysis
 * (Vc1Encoder::IntensityComp, src/encoder_internal.h:341-345): LUMSCALE and
 * LUMSHIFT as coded (6 bits each); enabled = 0 means "no compensation". */
typedef struct vc1me_intensity {
    int32_t enabled;
    int32_t lumscale;
    int32_t lumshift;
} vc1me_intensity;

/* Vc1Encoder::estimate_intensity_comp(cur, ref, mvs)
 * (src/encoder_motion_compensation.cpp:296-348): least-squares fit of the
 * current luma against the reference luma displaced by each macroblock's
 * 1-MV vector rounded to whole pixels, sampled every 4th pixel in x and y,
 * then the (LUMSCALE, LUMSHIFT) of smallest sampled SAD within +-3 of the
 * fit.  mvs: mbw * mbh quarter-pel (x, y) pairs of cur's raster (the
 * PAnalysis::mvs of the chosen pass).  cur and ref must share a raster. */
int vc1me_estimate_intensity_comp(vc1me_ctx* ctx, const vc1me_frame* cur, const vc1me_frame* ref,
                                  const int32_t* mvs, vc1me_intensity* out);

/* Vc1Encoder::intensity_compensated_reference(ref, ic)
 * (src/encoder_motion_compensation.cpp:195-294, the maps intensity_map_y /
 * intensity_map_uv at :175-193): a new resident frame whose luma and chroma
 * are ref's under the compensation (a plain copy when ic->enabled is 0).
 * Release it with vc1me_frame_release. */
int vc1me_frame_intensity_compensated(vc1me_ctx* ctx, const vc1me_frame* ref, const vc1me_intensity* ic,
                                      vc1me_frame** out);

/* ---- B-picture analysis: Vc1Encoder::analyze_b_picture_mode
 * (src/encoder_motion.cpp:753-1004) for one progressive 1-MV mode.  Two
 * core passes without intra (forward: cur against past, intensity
 * compensated when forward_intensity is enabled; backward: cur against
 * future), then per macroblock in raster order the choice among direct
 * (the future anchor's collocated vector scaled by the temporal fraction),
 * forward, backward and interpolated prediction by SAD + lambda * MV bits
 * (lambda = max(1, 0.45 * PQINDEX)), and the B-intra decision.  Modes as
 * BAnalysis::modes (src/encoder_internal.h:407). */
enum { VC1ME_B_FORWARD = 0, VC1ME_B_BACKWARD = 1, VC1ME_B_INTERPOLATED = 2, VC1ME_B_DIRECT = 3, VC1ME_B_INTRA = 4 };

typedef struct vc1me_b_params {
    const int32_t* future_anchor_mvs;   /* mbw * mbh quarter-pel (x, y): the future P anchor's mvs */
    const uint8_t* future_anchor_4mv;   /* mbw * mbh: the anchor's use_4mv */
    vc1me_intensity forward_intensity;  /* the anchor's LUMSCALE / LUMSHIFT, applied to past */
    int32_t fraction_num;               /* temporal position: 1/2, 1/3 or 2/3 */
    int32_t fraction_den;
    int32_t allow_intra;                /* !EncoderConfig::debug_disable_b_intra */
    /* the encoder's minimal zero-motion configuration (search range 0 and no
     * fade compensation, loop filter, variable transforms, trellis or
     * adaptive quality, src/encoder_motion.cpp:775-787): every macroblock
     * forward with zero vectors, no search */
    int32_t zero_motion_shortcut;
} vc1me_b_params;

/* BAnalysis (src/encoder_internal.h:408-434), the motion fields; array
 * pointers may be NULL when not wanted.  Under the zero-motion shortcut the
 * distant arrays hold -1, -1, 0. */
typedef struct vc1me_b_result {
    uint8_t* modes;
    int32_t* forward_mvs;               /* mbw * mbh quarter-pel (x, y) */
    int32_t* backward_mvs;
    double* inter_intra_cost_ratio;
    double* forward_distant_local_mae;
    double* forward_distant_candidate_mae;
    uint8_t* forward_distant_match_decision;
    double* backward_distant_local_mae;
    double* backward_distant_candidate_mae;
    uint8_t* backward_distant_match_decision;
    uint64_t forward_macroblocks;
    uint64_t backward_macroblocks;
    uint64_t interpolated_macroblocks;
    uint64_t direct_macroblocks;
    uint64_t intra_macroblocks;
    uint64_t moved_macroblocks;
    uint64_t fractional_chroma_macroblocks;
    double mean_abs_residual;
    double rate_complexity;
    uint64_t motion_sad;
    uint64_t motion_bits;
} vc1me_b_result;

/* params: the encoder's ME settings; params->mode is the B picture's mode
 * (VC1ME_MODE_ONE_MV_QPEL, _HPEL or _HPEL_BILINEAR; Mixed-MV is
 * VC1ME_EINVAL), params->allow_intra is not used (the core passes run
 * without intra).  cur, past and future share one raster. */
int vc1me_analyze_b_frame(vc1me_ctx* ctx, const vc1me_frame* cur, const vc1me_frame* past, const vc1me_frame* future,
                          const vc1me_params* params, const vc1me_b_params* b, vc1me_b_result* out);

/* n B pictures in one step (their 2n core passes run as one batched step,
 * then every picture's decision): equivalent to n calls of
 * vc1me_analyze_b_frame; at most 127 pictures. */
typedef struct vc1me_b_job {
    const vc1me_frame* cur;
    const vc1me_frame* past;
    const vc1me_frame* future;
    const vc1me_params* params;
    const vc1me_b_params* b;
    vc1me_b_result* out;
} vc1me_b_job;
int vc1me_analyze_b_step(vc1me_ctx* ctx, const vc1me_b_job* jobs, int n);

/* ---- scene cuts: Vc1Encoder::scene_change_score (src/encoder_motion.cpp:599-676).
 * The global-translation MAD of cur against ref (search range
 * params->motion_search_range, clamped to 32 by the search), the
 * intensity-compensated one when fade_compensation is set, and for a score
 * at or above scene_threshold the 1-MV quarter-pel core's MAD if the
 * interior vector field is a coherent zoom (translation + isotropic scale
 * within 2.5 px for 60% of vectors, 0.005 <= |scale| <= 0.08) and that MAD
 * is below 0.72 x the score and 0.86 x the threshold.  The encoder inserts
 * an I picture when the score reaches its scene threshold (default 28). */
int vc1me_scene_change_score(vc1me_ctx* ctx, const vc1me_frame* cur, const vc1me_frame* ref,
                             const vc1me_params* params, int fade_compensation, double scene_threshold,
                             double* score);

/* ---- motion failure: Vc1Encoder::has_good_motion_match
 * (src/encoder_motion.cpp:678-751).  *good = 1 when some macroblock of cur
 * has a usable match in ref (against the fade-compensated reference when
 * fade_compensation finds a code on the global field): the P search's
 * integer and long-range stages refined by the SAD-only refine_qpel_block
 * (:3596-3685), usable at <= 12 SAD per pixel or <= 90% of the zero-vector
 * SAD.  The encoder inserts an emergency I picture when it is 0.
 * field (NULL when not wanted; no reference counterpart): every
 * macroblock's refined quarter-pel vector, mbw * mbh (x, y) pairs, as the
 * reference's scan would compute them if it did not stop at the first
 * usable match. */
int vc1me_has_good_motion_match(vc1me_ctx* ctx, const vc1me_frame* cur, const vc1me_frame* ref,
                                const vc1me_params* params, int fade_compensation, int* good, int32_t* field);

/* ---- motion compensation for coding: the P prediction picture of an
 * analysis, motion_compensate_4mv (src/encoder_motion_compensation.cpp:536-571)
 * for macroblocks with use_4mv (four bicubic luma blocks at block_mvs,
 * chroma at their median-4 vector) and motion_compensate_mb (:474-521) with
 * params->mode at mvs for the others (intra macroblocks carry zero vectors,
 * so they get the co-located reference).  A new resident frame; release it
 * with vc1me_frame_release. */
int vc1me_predict_p_frame(vc1me_ctx* ctx, const vc1me_frame* ref, const vc1me_params* params,
                          const vc1me_p_result* analysis, vc1me_frame** out);

/* The bi-directional prediction picture: motion_compensate_bi_mb
 * (src/encoder_motion_compensation.cpp:659-711) for every macroblock, with
 * fmvs against past and bmvs against future (mbw * mbh quarter-pel (x, y)
 * each; params->mode a 1-MV mode).  The interpolated and direct B
 * macroblocks' prediction; a new resident frame. */
int vc1me_predict_bi_frame(vc1me_ctx* ctx, const vc1me_frame* past, const vc1me_frame* future,
                           const vc1me_params* params, const int32_t* fmvs, const int32_t* bmvs, vc1me_frame** out);

/* The padded prediction pictures of the loop reconstruction
 * (src/encoder_encode.cpp:1183, :1195, :1681-1687): the same pictures as
 * vc1me_predict_p_frame / vc1me_predict_bi_frame built by
 * motion_compensate_4mv_padded, motion_compensate_mb_padded
 * (src/encoder_motion_compensation.cpp:573-657) and
 * motion_compensate_bi_mb_padded (:712-763).  The new resident frame covers
 * the whole macroblock raster of params (mbw*16 x mbh*16, the encoder's
 * empty_padded_frame, src/encoder_encode.cpp:1949-1956), every sample of
 * every macroblock written; each reference is sampled with clamping to its
 * own raster, which may be the coded raster or the padded one (the
 * encoder's padded loop reference).  A B picture's padded reconstruction
 * takes forward and backward macroblocks from vc1me_predict_p_frame_padded
 * of each reference and interpolated or direct ones from
 * vc1me_predict_bi_frame_padded. */
int vc1me_predict_p_frame_padded(vc1me_ctx* ctx, const vc1me_frame* ref, const vc1me_params* params,
                                 const vc1me_p_result* analysis, vc1me_frame** out);
int vc1me_predict_bi_frame_padded(vc1me_ctx* ctx, const vc1me_frame* past, const vc1me_frame* future,
                                  const vc1me_params* params, const int32_t* fmvs, const int32_t* bmvs,
                                  vc1me_frame** out);

/* Copies a resident frame's planes back to the host (width x height luma,
 * two width/2 x height/2 chroma planes).  No reference counterpart. */
int vc1me_frame_download(vc1me_ctx* ctx, const vc1me_frame* frame, uint8_t* y, uint8_t* u, uint8_t* v);

/* Timing of the last vc1me_analyze_step / vc1me_analyze_p_frame: device
 * milliseconds between HIP events recorded on the context stream around the
 * whole launch sequence, and around every launch of the macroblock search
 * kernel (sum and count).  graph_replay: 0 = launched one by one, 1 = a
 * cached HIP graph of the step was replayed, 2 = the graph was captured by
 * this step (vc1me_set_graphs).  schedule: the macroblock schedule the step
 * ran (VC1ME_SCHEDULE_*); relax_passes and relax_analyses: passes of the
 * relaxation schedule and macroblock analyses over all of them (equal to the
 * macroblock count when one pass settles), 0 under the diagonal sweep. */
typedef struct vc1me_timing {
    double step_ms;
    double mb_kernel_ms_total;
    uint64_t mb_kernel_launches;
    double global_kernel_ms_total;
    uint64_t global_kernel_launches;
    uint64_t graph_replay;
    uint64_t schedule;
    uint64_t relax_passes;
    uint64_t relax_analyses;
    /* with profiling on: device ms of the macroblock pre-pass (the
     * relaxation's guess pass, or the extended long-range pass ahead of the
     * sweep), which is included in mb_kernel_ms_total / mb_kernel_launches */
    double prepass_ms;
} vc1me_timing;
int vc1me_last_timing(vc1me_ctx* ctx, vc1me_timing* out);
/* Record HIP events around every macroblock-kernel launch so that
 * vc1me_timing.mb_kernel_ms_total is filled (off by default: the events add
 * a little launch overhead).  No reference counterpart (profiling aid,
 * like the reference's SpeedProfiler, src/encoder_internal.h:182-221). */
int vc1me_set_profiling(vc1me_ctx* ctx, int on);
/* Run each plain analysis step as one HIP graph: the step's whole launch
 * sequence (counter clearing, phase planes, global search, signature index,
 * then either the relaxation's guess pass and its device-driven passes or
 * the long-range pre-pass and one launch per anti-diagonal, then the step's
 * summary and packed results) is captured the first time a launch shape
 * (job count, rasters, ranges, modes, results wanted, relaxation passes) is
 * seen and replayed afterwards; the job table is uploaded before each replay
 * and the results come back in one copy after it.  Results are identical to
 * launching one by one.  Not used while profiling is on or for the learned
 * forms.  No reference counterpart (replaces the per-batch dispatches of
 * src/vulkan_compute.cpp:370-490). */
int vc1me_set_graphs(vc1me_ctx* ctx, int on);

/* Macroblock schedule of the plain and learned steps.  Both give the
 * raster-order result of analyze_p_picture_core's macroblock loop
 * (src/encoder_motion.cpp:113-428) bit for bit.
 *   VC1ME_SCHEDULE_DIAGONAL: one launch per anti-diagonal mx + 2*my = t, so
 *     every macroblock runs once, after its left, top-left, top and
 *     top-right neighbours.
 *   VC1ME_SCHEDULE_RELAX (default): every macroblock at once from a guessed
 *     neighbour state (the global vector), then passes that re-analyse only
 *     the macroblocks whose neighbours changed, until a pass changes
 *     nothing; the dependency graph is acyclic, so that fixed point is the
 *     raster-order result.  The passes run from device-side worklists
 *     without host reads: a step launches as many passes as its launch
 *     shape has needed before (plus two; 24 for a new shape) and launches
 *     more only if the last one still changed something.  After 96 passes
 *     without settling the step finishes with the diagonal sweep.
 * Both schedules are captured into HIP graphs with vc1me_set_graphs.  No
 * reference counterpart (the reference runs its macroblock loop on one
 * thread). */
enum { VC1ME_SCHEDULE_DIAGONAL = 0, VC1ME_SCHEDULE_RELAX = 1 };
int vc1me_set_schedule(vc1me_ctx* ctx, int schedule);
/* Per-pass record of the last step's relaxation: macroblock analyses of each
 * pass summed over the jobs and, when profiling is on, the pass's device
 * milliseconds (-1 otherwise), for the first `cap` passes.  Returns the pass
 * count (0 under the diagonal sweep) or a negative status.  No reference
 * counterpart (profiling aid). */
int vc1me_last_relax_passes(vc1me_ctx* ctx, int64_t* analyses, double* ms, int cap);

/* Algorithmic work of the last step, summed over its jobs, in the units of
 * the roofline: one |a-b| accumulate per sampled pixel of a block SAD
 * (sad_px), one 4-tap filter application per interpolated sample (1D = 1,
 * separable 2D = 5; interp_taps), 112 per 4x4 Hadamard tile (satd_ops), one
 * long-range signature comparison per index cell (sig_ops), one |a-b| per
 * sampled pixel of a frame SAD (frame_px).  These are the operations the
 * reference algorithm performs on the same inputs.  No reference
 * counterpart (profiling aid). */
typedef struct vc1me_work {
    uint64_t sad_px;
    uint64_t interp_taps;
    uint64_t satd_ops;
    uint64_t sig_ops;
    uint64_t frame_px;
    uint64_t seeded_macroblocks;   /* learned forms: macroblocks searched from proposals */
} vc1me_work;
int vc1me_last_work(vc1me_ctx* ctx, vc1me_work* out);

/* ---- counters (VulkanComputeContext::counters / reset_counters,
 * src/vulkan_compute.h:53-54) */
int vc1me_counters_get(vc1me_ctx* ctx, vc1me_counters* out);
void vc1me_counters_reset(vc1me_ctx* ctx);

/* ---- learned proposals.  No counterpart in the reference: a small network
 * proposes, per macroblock, two integer-search seeds and a confidence; the
 * search still judges every vector with exact costs.
 *
 * Network: 1/4-resolution luma of cur and ref (2 channels), four 3x3
 * convolutions to 32 channels (dilations 1, 2, 4, 8) with bias and ReLU, a
 * mean pool per macroblock and a 1x1 head to 5 outputs per macroblock:
 * (dx1, dy1, dx2, dy2) in units of 64 pixels and a confidence logit.
 * Parameters are fp32 masters (layout: per convolution 32 x 9 x cin weights
 * [out][tap][in] then 32 biases; then 5 x 32 head weights and 5 biases);
 * the kernels compute in bf16 with fp32 accumulation. */
typedef struct vc1me_nn vc1me_nn;
#define VC1ME_NN_OUTPUTS 5
int vc1me_nn_param_count(void);
int vc1me_nn_create(vc1me_ctx* ctx, vc1me_nn** out);
void vc1me_nn_destroy(vc1me_nn* nn);
/* Copy parameters in (resets the optimizer state) or out, kParams floats. */
int vc1me_nn_set_params(vc1me_nn* nn, const float* params);
int vc1me_nn_get_params(vc1me_nn* nn, float* params);
/* Forward pass on n frame pairs of one raster; out receives
 * n x mbh x mbw x 5 floats.  globals (n x 2 integer pixels, or null for
 * zero) is each pair's global vector: the network sees ref displaced by it
 * and its two candidates are residuals from it. */
int vc1me_nn_forward_frames(vc1me_nn* nn, const vc1me_frame* const* curs, const vc1me_frame* const* refs,
                            const int32_t* globals, int n, float* out);
/* Forward and backward on a given network input (n x h4 x w4 x 2 floats, as
 * the downsampling would produce) with a given output gradient
 * (n x mbh x mbw x 5, mbh = ceil(h4/4), mbw = ceil(w4/4)); returns the outputs
 * and the parameter gradients.  For gradient checks. */
int vc1me_nn_forward_backward(vc1me_nn* nn, const float* input, int n, int h4, int w4, const float* dout, float* out,
                              float* grads);
/* One optimizer step (Adam) on n frame pairs of one raster: targets are the
 * plain search's 1-MV vectors (n x mbs x 2, quarter-pel), valid marks the
 * macroblocks that count (n x mbs), globals as for vc1me_nn_forward_frames
 * (the plain search's global vectors).  loss receives the batch loss. */
int vc1me_nn_train_step(vc1me_nn* nn, const vc1me_frame* const* curs, const vc1me_frame* const* refs,
                        const int32_t* globals, const int32_t* targets, const uint8_t* valid, int n, float lr,
                        float* loss);
/* vc1me_analyze_step with learned seeding: the network runs on every pair
 * right after the global search (so it sees ref aligned by the global vector);
 * macroblocks whose confidence reaches `threshold` replace the integer search
 * (UMH stages and the long-range stages) by a search seeded with the two
 * proposals plus the predictor, zero and the global vector, ending in a unit
 * diamond; sub-pel refinement, 4-MV and intra run as in the plain search.
 * The others run the plain search unchanged. */
int vc1me_analyze_step_learned(vc1me_ctx* ctx, vc1me_nn* nn, const vc1me_job* jobs, int n, float threshold);
/* The same seeding with proposals from the caller: proposals[i] holds
 * mbw x mbh x 5 floats for job i in the network's output layout (two vectors
 * in units of 64 pixels relative to the job's global vector, then a
 * confidence logit). */
int vc1me_analyze_step_proposals(vc1me_ctx* ctx, const vc1me_job* jobs, int n, const float* const* proposals,
                                 float threshold);
/* Learned seeding without the proposals' vectors (on = 0): confident
 * macroblocks of vc1me_analyze_step_learned run the seeded search from the
 * predictor, zero and the global vector only, the network acting as a
 * confidence gate; on = 1 (the default) uses its two candidates too.  Only
 * vc1me_analyze_step_learned reads this setting: vc1me_analyze_step_proposals
 * always uses the caller's vectors.  No reference counterpart. */
int vc1me_set_learned_candidates(vc1me_ctx* ctx, int on);

/* ---- cost report (no reference counterpart; judges the learned form).
 * While on, every analysis records per macroblock C20 = 20 * (SAD + lambda *
 * R) of its final 1-MV vector: SAD by block_sad_motion at that vector, R the
 * MV rate against predictor_info_mixed_1mv of the final field including the
 * HYBRIDPRED bit, lambda = max(1, 0.35 * PQINDEX) (the rate-winner score of
 * src/encoder_motion.cpp:3458-3471; 20 * lambda = max(20, 7 * PQINDEX) is an
 * integer).  vc1me_last_mb_costs copies job `job` of the last step: mbs
 * values, and optionally the 1-MV vectors (mbs x 2 quarter-pel). */
int vc1me_set_cost_report(vc1me_ctx* ctx, int on);
int vc1me_last_mb_costs(vc1me_ctx* ctx, int job, uint64_t* cost20, int32_t* one_mvs);

#ifdef __cplusplus
}
#endif

#endif /* VC1ME_H */
