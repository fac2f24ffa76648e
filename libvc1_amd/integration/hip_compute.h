// hip_compute.h — the adapter a libvc1 maintainer adds to the encoder to use
// this backend: a HipComputeContext with the interface of libvc1's
// VulkanComputeContext (src/vulkan_compute.h:26-60), so the encoder's
// compute slot (EncoderConfig::vulkan_compute and the vulkan_min_batch*
// thresholds, src/encoder_internal.h:286-289) and its CPU fallback stay as
// they are, plus the whole-frame analysis in front of
// Vc1Encoder::analyze_p_picture_core (src/encoder_motion.cpp:47-442).
//
// It is written as it would sit in libvc1's src/ (namespace libvc1) and is
// compiled by tests/adapter against this repository's include/vc1me.h.  It
// needs in scope: Frame, EncoderConfig, Vc1Encoder (MotionVector,
// ProgressiveMvMode, PAnalysis), VulkanMotionCandidate, VulkanMotionMetric
// and VulkanComputeCounters.
#pragma once

#include <array>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "vc1me.h"

namespace libvc1 {

class HipComputeContext {
public:
    // VulkanComputeContext::create (src/vulkan_compute.cpp:493-505): null and
    // a message when no device can be used; the encoder then stays on the CPU
    static std::shared_ptr<HipComputeContext> create(int device, std::string& error) {
        vc1me_ctx* c = nullptr;
        if (vc1me_create(device, &c) != VC1ME_OK) {
            error = vc1me_last_error();
            return {};
        }
        return std::shared_ptr<HipComputeContext>(new HipComputeContext(c));
    }
    ~HipComputeContext() { vc1me_destroy(ctx_); }
    HipComputeContext(const HipComputeContext&) = delete;
    HipComputeContext& operator=(const HipComputeContext&) = delete;

    std::string device_name() const { return vc1me_device_name(ctx_); }

    // Same contract as VulkanComputeContext::score_motion_batch
    // (src/vulkan_compute.h:43-51): false (with the reason) means "re-score
    // on the CPU", and the bitstream does not change.
    bool score_motion_batch(const uint8_t* cur, const uint8_t* ref, int width, int height, int x0, int y0, int bw,
                            int bh, int sample_stride, bool bilinear, bool rnd, VulkanMotionMetric metric,
                            const VulkanMotionCandidate* cands, size_t n, std::vector<uint64_t>& costs,
                            std::string* error = nullptr) {
        static_assert(sizeof(VulkanMotionCandidate) == sizeof(vc1me_candidate), "candidate layout");
        costs.resize(n);
        const int rc = vc1me_score_batch(ctx_, cur, ref, width, height, x0, y0, bw, bh, sample_stride,
                                         bilinear ? 1 : 0, rnd ? 1 : 0, static_cast<int>(metric),
                                         reinterpret_cast<const vc1me_candidate*>(cands), n, costs.data());
        if (rc != VC1ME_OK) {
            if (error) *error = vc1me_last_error();
            return false;
        }
        return true;
    }

    // Whole-frame P analysis: analyze_p_picture_core(f, ref, mode,
    // allow_intra, shared_seed, nullptr).  shared_seed null: the plain core
    // (vc1me_analyze_p_frame); otherwise the seeded core with the seed's
    // vectors and distant fields (vc1me_analyze_p_frame_seeded).  False (with
    // the reason) means "run the CPU loop".
    bool analyze_p(const Frame& f, const Frame& ref, const EncoderConfig& c, Vc1Encoder::ProgressiveMvMode mode,
                   bool allow_intra, const Vc1Encoder::PAnalysis* shared_seed, Vc1Encoder::PAnalysis& out,
                   std::string* error = nullptr) {
        static_assert(sizeof(Vc1Encoder::MotionVector) == 2 * sizeof(int32_t), "MotionVector is two int32");
        static_assert(sizeof(std::array<Vc1Encoder::MotionVector, 4>) == 8 * sizeof(int32_t), "block vectors");
        vc1me_params p;
        vc1me_params_default(&p, c.width, c.height);
        p.display_width = c.display_width;
        p.display_height = c.display_height;
        p.pqindex = c.pqindex;
        p.motion_search_range = c.motion_search_range;
        p.motion_local_search_range = c.motion_local_search_range;
        p.me_algorithm = static_cast<int32_t>(c.me_algorithm);
        p.me_quality = static_cast<int32_t>(c.me_quality);
        p.subpixel_refinement = c.subpixel_refinement ? 1 : 0;
        p.fast_subpixel_refinement = c.fast_subpixel_refinement ? 1 : 0;
        p.subpixel_refinement_range = c.subpixel_refinement_range;
        p.long_range_search_mode = static_cast<int32_t>(c.long_range_search_mode);
        p.distant_match_max_mae = c.distant_match_max_mae;
        p.inter_intra_threshold = c.inter_intra_threshold;
        p.debug_disable_p_intra = c.debug_disable_p_intra ? 1 : 0;
        p.mode = static_cast<int32_t>(mode);
        p.allow_intra = allow_intra ? 1 : 0;
        const size_t mbs = static_cast<size_t>((c.width + 15) / 16) * ((c.height + 15) / 16);
        out.mvs.assign(mbs, {});
        out.block_mvs.assign(mbs, {});
        out.use_4mv.assign(mbs, 0);
        out.use_intra.assign(mbs, 0);
        out.screen_sads.assign(mbs, 0);
        out.inter_intra_cost_ratio.assign(mbs, -1.0);
        out.distant_local_mae.assign(mbs, -1.0);
        out.distant_candidate_mae.assign(mbs, -1.0);
        out.distant_match_decision.assign(mbs, 0);
        vc1me_p_result r{};
        point(r, out);
        vc1me_frame* fc = nullptr;
        vc1me_frame* fr = nullptr;
        int rc = upload(f, &fc);
        if (rc == VC1ME_OK) rc = upload(ref, &fr);
        if (rc == VC1ME_OK) {
            if (shared_seed) {
                vc1me_p_result s{};
                point(s, *const_cast<Vc1Encoder::PAnalysis*>(shared_seed));
                rc = vc1me_analyze_p_frame_seeded(ctx_, fc, fr, &p, &s, &r);
            } else {
                rc = vc1me_analyze_p_frame(ctx_, fc, fr, &p, &r);
            }
        }
        if (rc != VC1ME_OK && error) *error = vc1me_last_error();
        vc1me_frame_release(ctx_, fc);
        vc1me_frame_release(ctx_, fr);
        if (rc != VC1ME_OK) return false;
        out.mean_abs_residual = r.mean_abs_residual;
        out.rate_complexity = r.rate_complexity;
        out.motion_sad = r.motion_sad;
        out.moved_macroblocks = static_cast<size_t>(r.moved_macroblocks);
        out.fractional_chroma_macroblocks = static_cast<size_t>(r.fractional_chroma_macroblocks);
        out.four_mv_macroblocks = static_cast<size_t>(r.four_mv_macroblocks);
        out.intra_macroblocks = static_cast<size_t>(r.intra_macroblocks);
        out.mv_mode = mode;
        return true;
    }

    // VulkanComputeContext::counters / reset_counters (src/vulkan_compute.h:53-54)
    VulkanComputeCounters counters() const {
        vc1me_counters k{};
        vc1me_counters_get(ctx_, &k);
        VulkanComputeCounters o;
        o.dispatches = k.dispatches;
        o.candidates = k.candidates;
        o.frame_uploads = k.frame_uploads;
        o.frame_upload_bytes = k.frame_upload_bytes;
        o.submit_wait_nanoseconds = k.submit_wait_nanoseconds;
        return o;
    }
    void reset_counters() { vc1me_counters_reset(ctx_); }

    vc1me_ctx* raw() const { return ctx_; }

private:
    explicit HipComputeContext(vc1me_ctx* c) : ctx_(c) {}

    static void point(vc1me_p_result& r, Vc1Encoder::PAnalysis& a) {
        r.mvs = reinterpret_cast<int32_t*>(a.mvs.data());
        r.block_mvs = reinterpret_cast<int32_t*>(a.block_mvs.data());
        r.use_4mv = a.use_4mv.data();
        r.use_intra = a.use_intra.data();
        r.screen_sads = a.screen_sads.data();
        r.inter_intra_cost_ratio = a.inter_intra_cost_ratio.data();
        r.distant_local_mae = a.distant_local_mae.data();
        r.distant_candidate_mae = a.distant_candidate_mae.data();
        r.distant_match_decision = a.distant_match_decision.data();
    }

    int upload(const Frame& f, vc1me_frame** out) {
        vc1me_planes pl{f.y.data(), f.u.empty() ? nullptr : f.u.data(), f.v.empty() ? nullptr : f.v.data(), f.width,
                        f.height};
        return vc1me_frame_upload(ctx_, &pl, out);
    }

    vc1me_ctx* ctx_;
};

}  // namespace libvc1
