// libvc1_shapes.h — the libvc1 types the adapter (libvc1_amd/integration/
// hip_compute.h) touches, with the field names and types libvc1 gives them,
// so that the adapter compiles and runs outside libvc1 in this test:
//   Frame                      src/encoder_internal.h:67-70
//   EncoderConfig (ME fields)  src/encoder_internal.h:223-315
//   Vc1Encoder::MotionVector, ProgressiveMvMode, PAnalysis
//                              src/encoder_internal.h:336-370
//   VulkanMotionCandidate, VulkanMotionMetric, VulkanComputeCounters
//                              src/vulkan_compute.h:11-24
// Only what the adapter reads or writes is here; the enums of libvc1's
// public API (vc1_me_algorithm_e and the like) are plain ints with the same
// values.
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <vector>

namespace libvc1 {

struct Frame {
    int width = 0, height = 0;
    std::vector<uint8_t> y, u, v;
};

struct EncoderConfig {
    int width = 0, height = 0;
    int display_width = 0, display_height = 0;
    int pqindex = 9;
    int motion_search_range = 1024;
    int motion_local_search_range = 32;
    int me_algorithm = 0;            // vc1_me_algorithm_e: 0 UMH, 1 diamond
    int me_quality = 2;              // vc1_me_quality_e: 1 RATE, 2 SATD
    bool subpixel_refinement = true;
    bool fast_subpixel_refinement = false;
    int subpixel_refinement_range = 1024;
    int long_range_search_mode = 0;  // vc1_long_range_search_mode_e
    double distant_match_max_mae = 255.0;
    double inter_intra_threshold = 0.80;
    bool debug_disable_p_intra = false;
};

struct Vc1Encoder {
    struct MotionVector { int xq = 0, yq = 0; };
    enum class ProgressiveMvMode : uint8_t { OneMvQpel = 0, OneMvHpel = 1, OneMvHpelBilinear = 2, MixedMv = 3 };
    struct PAnalysis {
        std::vector<MotionVector> mvs;
        std::vector<std::array<MotionVector, 4>> block_mvs;
        std::vector<uint8_t> use_4mv;
        std::vector<uint8_t> use_intra;
        std::vector<uint64_t> screen_sads;
        std::vector<double> inter_intra_cost_ratio;
        std::vector<double> distant_local_mae;
        std::vector<double> distant_candidate_mae;
        std::vector<uint8_t> distant_match_decision;
        double mean_abs_residual = 0.0;
        double rate_complexity = 0.0;
        size_t moved_macroblocks = 0;
        size_t fractional_chroma_macroblocks = 0;
        size_t four_mv_macroblocks = 0;
        size_t intra_macroblocks = 0;
        ProgressiveMvMode mv_mode = ProgressiveMvMode::OneMvQpel;
        uint64_t motion_sad = 0;
    };
};

struct VulkanMotionCandidate {
    int32_t xq = 0;
    int32_t yq = 0;
};

enum class VulkanMotionMetric : uint8_t { Sad = 0, Satd = 1 };

struct VulkanComputeCounters {
    uint64_t dispatches = 0;
    uint64_t candidates = 0;
    uint64_t frame_uploads = 0;
    uint64_t frame_upload_bytes = 0;
    uint64_t submit_wait_nanoseconds = 0;
};

}  // namespace libvc1
