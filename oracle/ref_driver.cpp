// ref_driver.cpp — test infrastructure only.
//
// Binds the real libvc1 reference (compiled from the read-only reference
// sources by oracle/Makefile into oracle/_ref/) to a small C ABI so that the
// tests can check the C restatement in oracle/vc1me_oracle.c against it, and
// so that bench.py can time the reference's own multithreaded SIMD path as
// the CPU baseline.  Nothing in the product links this file.
//
// Every entry point calls a public member of libvc1::Vc1Encoder:
//   analyze_p_picture_core   src/encoder_motion.cpp:47-442
//   mb_sad                   src/encoder_motion.cpp:2197-2284
//   block_sad_motion         src/encoder_motion.cpp:2670-2933
//   block_satd_motion        src/encoder_motion.cpp:2962-3042
//   global_motion_search     src/encoder_motion.cpp:2076-2141
//   frame_sad                src/encoder_motion.cpp:1499-1576
//   luma_mc_sample_mode      src/encoder_motion_compensation.cpp:398-415
//   estimate_intensity_comp  src/encoder_motion_compensation.cpp:296-348
//   intensity_compensated_reference  src/encoder_motion_compensation.cpp:195-294
//   analyze_b_picture_mode   src/encoder_motion.cpp:753-1004
//   scene_change_score       src/encoder_motion.cpp:599-676
//   has_good_motion_match    src/encoder_motion.cpp:678-751
//   motion_compensate_mb / _4mv  src/encoder_motion_compensation.cpp:474-571
//   motion_compensate_bi_mb  src/encoder_motion_compensation.cpp:659-711
//   motion_compensate_mb_padded / _4mv_padded  src/encoder_motion_compensation.cpp:573-657
//   motion_compensate_bi_mb_padded  src/encoder_motion_compensation.cpp:712-763
//   empty_padded_frame       src/encoder_encode.cpp:1949-1956
#include "encoder_internal.h"

#include "../include/vc1me.h"

#include <array>
#include <atomic>
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

using namespace libvc1;

namespace {

EncoderConfig make_config(const vc1me_params* p, bool simd_auto) {
    EncoderConfig c;
    c.width = p->width;
    c.height = p->height;
    c.display_width = p->display_width;
    c.display_height = p->display_height;
    c.pqindex = p->pqindex;
    c.motion_search_range = p->motion_search_range;
    c.motion_local_search_range = p->motion_local_search_range;
    c.me_algorithm = static_cast<vc1_me_algorithm_e>(p->me_algorithm);
    c.me_quality = static_cast<vc1_me_quality_e>(p->me_quality);
    c.subpixel_refinement = p->subpixel_refinement != 0;
    c.fast_subpixel_refinement = p->fast_subpixel_refinement != 0;
    c.subpixel_refinement_range = p->subpixel_refinement_range;
    c.long_range_search_mode = static_cast<vc1_long_range_search_mode_e>(p->long_range_search_mode);
    c.distant_match_max_mae = p->distant_match_max_mae;
    c.inter_intra_threshold = p->inter_intra_threshold;
    c.debug_disable_p_intra = p->debug_disable_p_intra != 0;
    c.syntax = StreamSyntax::Advanced;
    c.rndctrl = false;
    if (simd_auto) {
        static const SimdDispatch dispatch = auto_select_simd(false, false).dispatch;
        c.simd_dispatch = dispatch;
    }
    return c;
}

// an encoder configuration that only fixes the raster (and SIMD tier)
EncoderConfig raster_config(int width, int height, bool simd_auto) {
    EncoderConfig c;
    c.width = width;
    c.height = height;
    c.syntax = StreamSyntax::Advanced;
    if (simd_auto) {
        static const SimdDispatch dispatch = auto_select_simd(false, false).dispatch;
        c.simd_dispatch = dispatch;
    }
    return c;
}

Frame make_frame(const vc1me_planes* pl) {
    Frame f;
    f.width = pl->width;
    f.height = pl->height;
    const size_t ys = static_cast<size_t>(pl->width) * pl->height;
    const size_t cs = static_cast<size_t>(pl->width / 2) * (pl->height / 2);
    f.y.assign(pl->y, pl->y + ys);
    if (pl->u) f.u.assign(pl->u, pl->u + cs); else f.u.assign(cs, 128);
    if (pl->v) f.v.assign(pl->v, pl->v + cs); else f.v.assign(cs, 128);
    return f;
}

void copy_result(const Vc1Encoder::PAnalysis& a, vc1me_p_result* out) {
    const size_t n = a.mvs.size();
    for (size_t i = 0; i < n; ++i) {
        if (out->mvs) { out->mvs[2 * i] = a.mvs[i].xq; out->mvs[2 * i + 1] = a.mvs[i].yq; }
        if (out->block_mvs)
            for (int k = 0; k < 4; ++k) {
                out->block_mvs[8 * i + 2 * k] = a.block_mvs[i][k].xq;
                out->block_mvs[8 * i + 2 * k + 1] = a.block_mvs[i][k].yq;
            }
        if (out->use_4mv) out->use_4mv[i] = a.use_4mv[i];
        if (out->use_intra) out->use_intra[i] = a.use_intra[i];
        if (out->screen_sads) out->screen_sads[i] = a.screen_sads[i];
        if (out->inter_intra_cost_ratio) out->inter_intra_cost_ratio[i] = a.inter_intra_cost_ratio[i];
        if (out->distant_local_mae) out->distant_local_mae[i] = a.distant_local_mae[i];
        if (out->distant_candidate_mae) out->distant_candidate_mae[i] = a.distant_candidate_mae[i];
        if (out->distant_match_decision) out->distant_match_decision[i] = a.distant_match_decision[i];
    }
    out->mean_abs_residual = a.mean_abs_residual;
    out->rate_complexity = a.rate_complexity;
    out->motion_sad = a.motion_sad;
    out->moved_macroblocks = a.moved_macroblocks;
    out->fractional_chroma_macroblocks = a.fractional_chroma_macroblocks;
    out->four_mv_macroblocks = a.four_mv_macroblocks;
    out->intra_macroblocks = a.intra_macroblocks;
}

thread_local std::string g_error;

}  // namespace

extern "C" {

const char* ref_last_error(void) { return g_error.c_str(); }

int ref_analyze_p_core(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                       vc1me_p_result* out, int simd_auto) {
    try {
        Vc1Encoder enc(make_config(p, simd_auto != 0));
        const Frame f = make_frame(cur), r = make_frame(ref);
        const auto a = enc.analyze_p_picture_core(
            f, r, static_cast<Vc1Encoder::ProgressiveMvMode>(p->mode), p->allow_intra != 0, nullptr, nullptr);
        copy_result(a, out);
        const int local = std::min(p->motion_search_range, std::max(0, p->motion_local_search_range));
        if (p->motion_search_range > 0) {
            const auto g = enc.global_motion_search(f, r, local);
            out->global_dx = g.dx; out->global_dy = g.dy; out->global_sad = g.full_sad;
        } else {
            out->global_dx = 0; out->global_dy = 0; out->global_sad = 0;
        }
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

// analyze_p_picture_core with a shared seed built from `seed` (its 1-MV
// vectors and distant-match fields, src/encoder_motion.cpp:97-101, :120-130)
int ref_analyze_p_core_seeded(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                              const vc1me_p_result* seed, vc1me_p_result* out, int simd_auto) {
    try {
        Vc1Encoder enc(make_config(p, simd_auto != 0));
        const Frame f = make_frame(cur), r = make_frame(ref);
        const size_t mbs = static_cast<size_t>((p->width + 15) / 16) * ((p->height + 15) / 16);
        Vc1Encoder::PAnalysis s;
        s.mv_mode = Vc1Encoder::ProgressiveMvMode::OneMvQpel;
        s.mvs.resize(mbs);
        s.distant_local_mae.resize(mbs);
        s.distant_candidate_mae.resize(mbs);
        s.distant_match_decision.resize(mbs);
        for (size_t i = 0; i < mbs; ++i) {
            s.mvs[i].xq = seed->mvs[2 * i];
            s.mvs[i].yq = seed->mvs[2 * i + 1];
            s.distant_local_mae[i] = seed->distant_local_mae[i];
            s.distant_candidate_mae[i] = seed->distant_candidate_mae[i];
            s.distant_match_decision[i] = seed->distant_match_decision[i];
        }
        const auto a = enc.analyze_p_picture_core(
            f, r, static_cast<Vc1Encoder::ProgressiveMvMode>(p->mode), p->allow_intra != 0, &s, nullptr);
        copy_result(a, out);
        out->global_dx = 0; out->global_dy = 0; out->global_sad = 0;
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

uint64_t ref_mb_sad(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                    int mx, int my, int dx, int dy, int stride) {
    Vc1Encoder enc(make_config(p, false));
    return enc.mb_sad(make_frame(cur), make_frame(ref), mx, my, dx, dy, stride);
}

/* Costs of n candidates for one block: metric 0 = block_sad_motion, 1 = block_satd_motion. */
int ref_block_costs(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                    int x0, int y0, int bw, int bh, int mode, int metric,
                    const vc1me_candidate* c, size_t n, uint64_t* out) {
    try {
        Vc1Encoder enc(make_config(p, false));
        const Frame f = make_frame(cur), r = make_frame(ref);
        const auto m = static_cast<Vc1Encoder::ProgressiveMvMode>(mode);
        for (size_t i = 0; i < n; ++i) {
            const Vc1Encoder::MotionVector mv{c[i].xq, c[i].yq};
            out[i] = metric ? enc.block_satd_motion(f, r, x0, y0, bw, bh, mv, m)
                            : enc.block_sad_motion(f, r, x0, y0, bw, bh, mv, m);
        }
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

uint64_t ref_frame_sad(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                       int dx, int dy, int stride) {
    Vc1Encoder enc(make_config(p, false));
    return enc.frame_sad(make_frame(cur), make_frame(ref), dx, dy, stride);
}

int ref_global_motion(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                      int range, int32_t* dx, int32_t* dy, uint64_t* sad) {
    Vc1Encoder enc(make_config(p, false));
    const auto g = enc.global_motion_search(make_frame(cur), make_frame(ref), range);
    *dx = g.dx; *dy = g.dy; *sad = g.full_sad;
    return 0;
}

int ref_estimate_intensity_comp(const vc1me_planes* cur, const vc1me_planes* ref, const int32_t* mvs,
                                vc1me_intensity* out) {
    try {
        Vc1Encoder enc(raster_config(cur->width, cur->height, false));
        const size_t mbs = static_cast<size_t>((cur->width + 15) / 16) * ((cur->height + 15) / 16);
        std::vector<Vc1Encoder::MotionVector> mv(mbs);
        for (size_t i = 0; i < mbs; ++i) { mv[i].xq = mvs[2 * i]; mv[i].yq = mvs[2 * i + 1]; }
        const auto ic = enc.estimate_intensity_comp(make_frame(cur), make_frame(ref), mv);
        out->enabled = ic.enabled ? 1 : 0;
        out->lumscale = ic.lumscale;
        out->lumshift = ic.lumshift;
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int ref_intensity_compensate(const vc1me_planes* ref, const vc1me_intensity* ic, uint8_t* y, uint8_t* u,
                             uint8_t* v, int simd_auto) {
    try {
        Vc1Encoder enc(raster_config(ref->width, ref->height, simd_auto != 0));
        Vc1Encoder::IntensityComp c;
        c.enabled = ic->enabled != 0;
        c.lumscale = static_cast<uint8_t>(ic->lumscale);
        c.lumshift = static_cast<uint8_t>(ic->lumshift);
        const Frame out = enc.intensity_compensated_reference(make_frame(ref), c);
        std::memcpy(y, out.y.data(), out.y.size());
        std::memcpy(u, out.u.data(), out.u.size());
        std::memcpy(v, out.v.data(), out.v.size());
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int ref_analyze_b(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* past,
                  const vc1me_planes* future, const vc1me_b_params* b, vc1me_b_result* out) {
    try {
        EncoderConfig cfg = make_config(p, false);
        cfg.debug_disable_b_intra = b->allow_intra == 0;
        if (b->zero_motion_shortcut) {
            cfg.fade_compensation = false;
            cfg.loop_filter = false;
            cfg.variable_transforms = false;
            cfg.trellis = 0;
            cfg.adaptive_quality = false;
        }
        Vc1Encoder enc(cfg);
        const size_t mbs = static_cast<size_t>((p->width + 15) / 16) * ((p->height + 15) / 16);
        std::vector<Vc1Encoder::MotionVector> amv(mbs);
        std::vector<uint8_t> a4(mbs);
        for (size_t i = 0; i < mbs; ++i) {
            amv[i].xq = b->future_anchor_mvs[2 * i];
            amv[i].yq = b->future_anchor_mvs[2 * i + 1];
            a4[i] = b->future_anchor_4mv[i];
        }
        Vc1Encoder::IntensityComp ic;
        ic.enabled = b->forward_intensity.enabled != 0;
        ic.lumscale = static_cast<uint8_t>(b->forward_intensity.lumscale);
        ic.lumshift = static_cast<uint8_t>(b->forward_intensity.lumshift);
        const auto a = enc.analyze_b_picture_mode(make_frame(cur), make_frame(past), make_frame(future), amv, a4, ic,
                                                  b->fraction_num, b->fraction_den,
                                                  static_cast<Vc1Encoder::ProgressiveMvMode>(p->mode), nullptr,
                                                  nullptr, nullptr);
        auto dbl = [mbs](const std::vector<double>& v, double* o) {
            if (o) for (size_t i = 0; i < mbs; ++i) o[i] = v.size() == mbs ? v[i] : -1.0;
        };
        auto u8 = [mbs](const std::vector<uint8_t>& v, uint8_t* o) {
            if (o) for (size_t i = 0; i < mbs; ++i) o[i] = v.size() == mbs ? v[i] : 0;
        };
        for (size_t i = 0; i < mbs; ++i) {
            if (out->modes) out->modes[i] = static_cast<uint8_t>(a.modes[i]);
            if (out->forward_mvs) { out->forward_mvs[2 * i] = a.forward_mvs[i].xq; out->forward_mvs[2 * i + 1] = a.forward_mvs[i].yq; }
            if (out->backward_mvs) { out->backward_mvs[2 * i] = a.backward_mvs[i].xq; out->backward_mvs[2 * i + 1] = a.backward_mvs[i].yq; }
        }
        dbl(a.inter_intra_cost_ratio, out->inter_intra_cost_ratio);
        dbl(a.forward_distant_local_mae, out->forward_distant_local_mae);
        dbl(a.forward_distant_candidate_mae, out->forward_distant_candidate_mae);
        u8(a.forward_distant_match_decision, out->forward_distant_match_decision);
        dbl(a.backward_distant_local_mae, out->backward_distant_local_mae);
        dbl(a.backward_distant_candidate_mae, out->backward_distant_candidate_mae);
        u8(a.backward_distant_match_decision, out->backward_distant_match_decision);
        out->forward_macroblocks = a.forward_macroblocks;
        out->backward_macroblocks = a.backward_macroblocks;
        out->interpolated_macroblocks = a.interpolated_macroblocks;
        out->direct_macroblocks = a.direct_macroblocks;
        out->intra_macroblocks = a.intra_macroblocks;
        out->moved_macroblocks = a.moved_macroblocks;
        out->fractional_chroma_macroblocks = a.fractional_chroma_macroblocks;
        out->mean_abs_residual = a.mean_abs_residual;
        out->rate_complexity = a.rate_complexity;
        out->motion_sad = a.motion_sad;
        out->motion_bits = a.motion_bits;
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

// seconds of one analyze_b_picture_mode call on this thread, SIMD tier
// auto-selected (the CPU baseline of the B-picture path)
double ref_time_b_frame(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* past,
                        const vc1me_planes* future, const vc1me_b_params* b) {
    try {
        Vc1Encoder enc(make_config(p, true));
        const size_t mbs = static_cast<size_t>((p->width + 15) / 16) * ((p->height + 15) / 16);
        std::vector<Vc1Encoder::MotionVector> amv(mbs);
        std::vector<uint8_t> a4(mbs);
        for (size_t i = 0; i < mbs; ++i) {
            amv[i].xq = b->future_anchor_mvs[2 * i];
            amv[i].yq = b->future_anchor_mvs[2 * i + 1];
            a4[i] = b->future_anchor_4mv[i];
        }
        Vc1Encoder::IntensityComp ic;
        const Frame f = make_frame(cur), pa = make_frame(past), fu = make_frame(future);
        const auto t0 = std::chrono::steady_clock::now();
        const auto a = enc.analyze_b_picture_mode(f, pa, fu, amv, a4, ic, b->fraction_num, b->fraction_den,
                                                  static_cast<Vc1Encoder::ProgressiveMvMode>(p->mode), nullptr,
                                                  nullptr, nullptr);
        const auto t1 = std::chrono::steady_clock::now();
        (void)a;
        return std::chrono::duration<double>(t1 - t0).count();
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1.0;
    }
}

int ref_scene_change_score(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                           int fade_compensation, double scene_threshold, double* out) {
    try {
        EncoderConfig cfg = make_config(p, false);
        cfg.fade_compensation = fade_compensation != 0;
        cfg.scene_threshold = scene_threshold;
        Vc1Encoder enc(cfg);
        *out = enc.scene_change_score(make_frame(cur), make_frame(ref));
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int ref_has_good_motion_match(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                              int fade_compensation, int* out) {
    try {
        EncoderConfig cfg = make_config(p, false);
        cfg.fade_compensation = fade_compensation != 0;
        Vc1Encoder enc(cfg);
        *out = enc.has_good_motion_match(make_frame(cur), make_frame(ref)) ? 1 : 0;
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int ref_refine_qpel_block(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref, int x0, int y0,
                          int bw, int bh, int seed_xq, int seed_yq, int range, uint64_t known_seed_sad,
                          int32_t* xq, int32_t* yq, uint64_t* sad) {
    try {
        Vc1Encoder enc(make_config(p, false));
        const Vc1Encoder::MotionVector m = enc.refine_qpel_block(make_frame(cur), make_frame(ref), x0, y0, bw, bh,
                                                                  Vc1Encoder::MotionVector{seed_xq, seed_yq}, range,
                                                                  sad, known_seed_sad);
        *xq = m.xq;
        *yq = m.yq;
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int ref_predict_p_frame(const vc1me_params* p, const vc1me_planes* ref, const int32_t* mvs, const int32_t* block_mvs,
                        const uint8_t* use_4mv, uint8_t* y, uint8_t* u, uint8_t* v, int simd_auto) {
    try {
        Vc1Encoder enc(make_config(p, simd_auto != 0));
        const Frame r = make_frame(ref);
        Frame dst = r;
        const int mbw = (p->width + 15) / 16, mbh = (p->height + 15) / 16;
        for (int my = 0; my < mbh; ++my)
            for (int mx = 0; mx < mbw; ++mx) {
                const size_t pos = static_cast<size_t>(my) * mbw + mx;
                if (use_4mv[pos]) {
                    std::array<Vc1Encoder::MotionVector, 4> b;
                    for (int k = 0; k < 4; ++k) b[k] = {block_mvs[8 * pos + 2 * k], block_mvs[8 * pos + 2 * k + 1]};
                    enc.motion_compensate_4mv(dst, r, mx, my, b);
                } else {
                    enc.motion_compensate_mb(dst, r, mx, my, {mvs[2 * pos], mvs[2 * pos + 1]},
                                             static_cast<Vc1Encoder::ProgressiveMvMode>(p->mode));
                }
            }
        std::memcpy(y, dst.y.data(), dst.y.size());
        std::memcpy(u, dst.u.data(), dst.u.size());
        std::memcpy(v, dst.v.data(), dst.v.size());
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int ref_predict_bi_frame(const vc1me_params* p, const vc1me_planes* past, const vc1me_planes* future,
                         const int32_t* fmvs, const int32_t* bmvs, uint8_t* y, uint8_t* u, uint8_t* v, int simd_auto) {
    try {
        Vc1Encoder enc(make_config(p, simd_auto != 0));
        const Frame pa = make_frame(past), fu = make_frame(future);
        Frame dst = pa;
        const int mbw = (p->width + 15) / 16, mbh = (p->height + 15) / 16;
        for (int my = 0; my < mbh; ++my)
            for (int mx = 0; mx < mbw; ++mx) {
                const size_t pos = static_cast<size_t>(my) * mbw + mx;
                enc.motion_compensate_bi_mb(dst, pa, fu, mx, my, {fmvs[2 * pos], fmvs[2 * pos + 1]},
                                            {bmvs[2 * pos], bmvs[2 * pos + 1]},
                                            static_cast<Vc1Encoder::ProgressiveMvMode>(p->mode));
            }
        std::memcpy(y, dst.y.data(), dst.y.size());
        std::memcpy(u, dst.u.data(), dst.u.size());
        std::memcpy(v, dst.v.data(), dst.v.size());
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

/* The loop-reconstruction prediction of a P picture: every macroblock of
 * p's raster through motion_compensate_4mv_padded or
 * motion_compensate_mb_padded into the encoder's own empty padded frame
 * (mbw*16 x mbh*16), from `ref` at whatever raster it has. */
int ref_predict_p_frame_padded(const vc1me_params* p, const vc1me_planes* ref, const int32_t* mvs,
                               const int32_t* block_mvs, const uint8_t* use_4mv, uint8_t* y, uint8_t* u, uint8_t* v,
                               int simd_auto) {
    try {
        Vc1Encoder enc(make_config(p, simd_auto != 0));
        const Frame r = make_frame(ref);
        const int mbw = (p->width + 15) / 16, mbh = (p->height + 15) / 16;
        Frame dst = enc.empty_padded_frame(mbw, mbh);
        for (int my = 0; my < mbh; ++my)
            for (int mx = 0; mx < mbw; ++mx) {
                const size_t pos = static_cast<size_t>(my) * mbw + mx;
                if (use_4mv[pos]) {
                    std::array<Vc1Encoder::MotionVector, 4> b;
                    for (int k = 0; k < 4; ++k) b[k] = {block_mvs[8 * pos + 2 * k], block_mvs[8 * pos + 2 * k + 1]};
                    enc.motion_compensate_4mv_padded(dst, r, mx, my, b);
                } else {
                    enc.motion_compensate_mb_padded(dst, r, mx, my, {mvs[2 * pos], mvs[2 * pos + 1]},
                                                    static_cast<Vc1Encoder::ProgressiveMvMode>(p->mode));
                }
            }
        std::memcpy(y, dst.y.data(), dst.y.size());
        std::memcpy(u, dst.u.data(), dst.u.size());
        std::memcpy(v, dst.v.data(), dst.v.size());
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

/* The loop-reconstruction bi-directional prediction: every macroblock
 * through motion_compensate_bi_mb_padded into an empty padded frame. */
int ref_predict_bi_frame_padded(const vc1me_params* p, const vc1me_planes* past, const vc1me_planes* future,
                                const int32_t* fmvs, const int32_t* bmvs, uint8_t* y, uint8_t* u, uint8_t* v,
                                int simd_auto) {
    try {
        Vc1Encoder enc(make_config(p, simd_auto != 0));
        const Frame pa = make_frame(past), fu = make_frame(future);
        const int mbw = (p->width + 15) / 16, mbh = (p->height + 15) / 16;
        Frame dst = enc.empty_padded_frame(mbw, mbh);
        for (int my = 0; my < mbh; ++my)
            for (int mx = 0; mx < mbw; ++mx) {
                const size_t pos = static_cast<size_t>(my) * mbw + mx;
                enc.motion_compensate_bi_mb_padded(dst, pa, fu, mx, my, {fmvs[2 * pos], fmvs[2 * pos + 1]},
                                                   {bmvs[2 * pos], bmvs[2 * pos + 1]},
                                                   static_cast<Vc1Encoder::ProgressiveMvMode>(p->mode));
            }
        std::memcpy(y, dst.y.data(), dst.y.size());
        std::memcpy(u, dst.u.data(), dst.u.size());
        std::memcpy(v, dst.v.data(), dst.v.size());
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

int ref_luma_mc_sample(const uint8_t* plane, int w, int h, int x, int y, int xq, int yq, int mode, int rnd) {
    std::vector<uint8_t> p(plane, plane + static_cast<size_t>(w) * h);
    return Vc1Encoder::luma_mc_sample_mode(p, w, h, x, y, xq, yq,
                                           static_cast<Vc1Encoder::ProgressiveMvMode>(mode), rnd != 0);
}

/* CPU baseline: analyse n independent (cur, ref) pairs with `threads` workers,
 * one whole frame pair per task (the reference parallelises only across
 * GOPs/frames, src/encoder_core.cpp:2180-2225).  Returns wall seconds of the
 * analysis calls (frames already resident in host memory), or -1 on error. */
double ref_time_p_frames(const vc1me_params* p, int n, const vc1me_planes* curs, const vc1me_planes* refs,
                         int threads, int simd_auto) {
    try {
        const EncoderConfig cfg = make_config(p, simd_auto != 0);
        std::vector<Frame> fc, fr;
        fc.reserve(n); fr.reserve(n);
        for (int i = 0; i < n; ++i) { fc.push_back(make_frame(&curs[i])); fr.push_back(make_frame(&refs[i])); }
        std::atomic<int> next{0};
        std::atomic<bool> failed{false};
        const auto mode = static_cast<Vc1Encoder::ProgressiveMvMode>(p->mode);
        const bool allow_intra = p->allow_intra != 0;
        auto worker = [&]() {
            try {
                Vc1Encoder enc(cfg);
                for (int i = next.fetch_add(1); i < n; i = next.fetch_add(1)) {
                    const auto a = enc.analyze_p_picture_core(fc[i], fr[i], mode, allow_intra, nullptr, nullptr);
                    if (a.mvs.empty()) failed = true;
                }
            } catch (...) {
                failed = true;
            }
        };
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<std::thread> pool;
        for (int t = 0; t < std::max(1, threads); ++t) pool.emplace_back(worker);
        for (auto& t : pool) t.join();
        const auto t1 = std::chrono::steady_clock::now();
        if (failed) { g_error = "reference analysis failed"; return -1.0; }
        return std::chrono::duration<double>(t1 - t0).count();
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1.0;
    }
}

/* SIMD tier the reference's auto-selector picks for the ME primitives
 * (src/encoder_simd.cpp:1374-1419); -1 if unknown. */
int ref_simd_tier(int primitive) {
    static const SimdDispatch dispatch = auto_select_simd(false, false).dispatch;
    if (primitive < 0 || primitive >= static_cast<int>(kSimdPrimitiveCount)) return -1;
    return static_cast<int>(dispatch.tier[static_cast<size_t>(primitive)]);
}

}  // extern "C"
