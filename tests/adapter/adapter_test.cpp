// adapter_test.cpp — drives the libvc1 adapter (libvc1_amd/integration/
// hip_compute.h) the way the encoder's compute slot does, through
// include/vc1me.h only.  Run by tests/test_gpu_adapter.py.
//
//   adapter_test full     scorer against the CPU law, whole-frame analysis
//                         and its seeded form against the C ABI called
//                         directly, one context shared by two host threads
//   adapter_test inject   with VC1ME_FAIL_AT set by the caller: the scorer
//                         returns false with the injected failure's message,
//                         the caller re-scores on the CPU and nothing changes
//
// Prints one line per check and "ADAPTER OK" at the end; exits non-zero on
// the first mismatch.
#include "libvc1_shapes.h"
#include "hip_compute.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

using namespace libvc1;

namespace {

int g_failures = 0;

void check(bool ok, const std::string& what) {
    std::printf("%s %s\n", ok ? "ok  " : "FAIL", what.c_str());
    std::fflush(stdout);
    if (!ok) ++g_failures;
}

// a smooth textured picture shifted by (sx, sy) pixels (deterministic)
Frame picture(int w, int h, int sx, int sy, uint32_t seed) {
    const int pw = w + 64, ph = h + 64;
    std::vector<int> noise(static_cast<size_t>(pw) * ph);
    uint32_t s = seed;
    for (auto& v : noise) { s = s * 1664525u + 1013904223u; v = static_cast<int>(s >> 24); }
    Frame f;
    f.width = w;
    f.height = h;
    f.y.resize(static_cast<size_t>(w) * h);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int acc = 0;
            for (int dy = 0; dy < 3; ++dy)
                for (int dx = 0; dx < 3; ++dx) acc += noise[static_cast<size_t>(y + 32 + sy + dy) * pw + x + 32 + sx + dx];
            f.y[static_cast<size_t>(y) * w + x] = static_cast<uint8_t>(acc / 9);
        }
    f.u.assign(static_cast<size_t>(w / 2) * (h / 2), 128);
    f.v.assign(f.u.size(), 128);
    return f;
}

// The CPU re-score of the reference's callers for integer vectors: SAD over
// the active block at sample stride s with clamped fetch, times s^2 (the law
// of the reference's sampled-SAD test, tests/vulkan_compute_runtime.cpp:17-29)
uint64_t cpu_sad(const Frame& cur, const Frame& ref, int x0, int y0, int bw, int bh, int dx, int dy, int s) {
    const int aw = std::max(0, std::min(bw, cur.width - x0)), ah = std::max(0, std::min(bh, cur.height - y0));
    uint64_t sad = 0;
    for (int y = 0; y < ah; y += s)
        for (int x = 0; x < aw; x += s) {
            const int rx = std::min(std::max(x0 + x + dx, 0), ref.width - 1);
            const int ry = std::min(std::max(y0 + y + dy, 0), ref.height - 1);
            sad += static_cast<uint64_t>(std::abs(static_cast<int>(cur.y[static_cast<size_t>(y0 + y) * cur.width + x0 + x]) -
                                                  static_cast<int>(ref.y[static_cast<size_t>(ry) * ref.width + rx])));
        }
    return sad * static_cast<uint64_t>(s * s);
}

std::vector<VulkanMotionCandidate> integer_candidates(int n, uint32_t seed) {
    std::vector<VulkanMotionCandidate> c(n);
    uint32_t s = seed;
    for (auto& v : c) {
        s = s * 1103515245u + 12345u;
        v.xq = 4 * (static_cast<int>((s >> 8) % 41) - 20);
        s = s * 1103515245u + 12345u;
        v.yq = 4 * (static_cast<int>((s >> 8) % 41) - 20);
    }
    return c;
}

// the encoder's call: the GPU scorer, or on failure the CPU re-score
std::vector<uint64_t> score_or_rescore(HipComputeContext& hc, const Frame& cur, const Frame& ref, int x0, int y0,
                                       int bw, int bh, int s, const std::vector<VulkanMotionCandidate>& c,
                                       bool* used_gpu, std::string* why) {
    std::vector<uint64_t> costs;
    if (hc.score_motion_batch(cur.y.data(), ref.y.data(), cur.width, cur.height, x0, y0, bw, bh, s, false, false,
                              VulkanMotionMetric::Sad, c.data(), c.size(), costs, why)) {
        *used_gpu = true;
        return costs;
    }
    *used_gpu = false;
    costs.resize(c.size());
    for (size_t i = 0; i < c.size(); ++i) costs[i] = cpu_sad(cur, ref, x0, y0, bw, bh, c[i].xq / 4, c[i].yq / 4, s);
    return costs;
}

bool same(const Vc1Encoder::PAnalysis& a, const Vc1Encoder::PAnalysis& b) {
    if (a.mvs.size() != b.mvs.size()) return false;
    for (size_t i = 0; i < a.mvs.size(); ++i) {
        if (a.mvs[i].xq != b.mvs[i].xq || a.mvs[i].yq != b.mvs[i].yq) return false;
        for (int k = 0; k < 4; ++k)
            if (a.block_mvs[i][k].xq != b.block_mvs[i][k].xq || a.block_mvs[i][k].yq != b.block_mvs[i][k].yq) return false;
    }
    return a.use_4mv == b.use_4mv && a.use_intra == b.use_intra && a.screen_sads == b.screen_sads &&
           std::memcmp(a.inter_intra_cost_ratio.data(), b.inter_intra_cost_ratio.data(), 8 * a.mvs.size()) == 0 &&
           std::memcmp(a.distant_local_mae.data(), b.distant_local_mae.data(), 8 * a.mvs.size()) == 0 &&
           std::memcmp(a.distant_candidate_mae.data(), b.distant_candidate_mae.data(), 8 * a.mvs.size()) == 0 &&
           a.distant_match_decision == b.distant_match_decision && a.motion_sad == b.motion_sad &&
           a.mean_abs_residual == b.mean_abs_residual && a.rate_complexity == b.rate_complexity &&
           a.moved_macroblocks == b.moved_macroblocks && a.four_mv_macroblocks == b.four_mv_macroblocks &&
           a.intra_macroblocks == b.intra_macroblocks &&
           a.fractional_chroma_macroblocks == b.fractional_chroma_macroblocks;
}

// the same analysis through the C ABI directly
Vc1Encoder::PAnalysis direct(vc1me_ctx* ctx, const Frame& f, const Frame& r, const EncoderConfig& c) {
    vc1me_params p;
    vc1me_params_default(&p, c.width, c.height);
    p.pqindex = c.pqindex;
    p.motion_search_range = c.motion_search_range;
    p.motion_local_search_range = c.motion_local_search_range;
    p.me_quality = c.me_quality;
    p.mode = VC1ME_MODE_MIXED_MV;
    p.allow_intra = 1;
    const size_t mbs = static_cast<size_t>((c.width + 15) / 16) * ((c.height + 15) / 16);
    Vc1Encoder::PAnalysis a;
    a.mvs.resize(mbs); a.block_mvs.resize(mbs); a.use_4mv.resize(mbs); a.use_intra.resize(mbs);
    a.screen_sads.resize(mbs); a.inter_intra_cost_ratio.resize(mbs); a.distant_local_mae.resize(mbs);
    a.distant_candidate_mae.resize(mbs); a.distant_match_decision.resize(mbs);
    vc1me_p_result res{};
    res.mvs = reinterpret_cast<int32_t*>(a.mvs.data());
    res.block_mvs = reinterpret_cast<int32_t*>(a.block_mvs.data());
    res.use_4mv = a.use_4mv.data(); res.use_intra = a.use_intra.data(); res.screen_sads = a.screen_sads.data();
    res.inter_intra_cost_ratio = a.inter_intra_cost_ratio.data(); res.distant_local_mae = a.distant_local_mae.data();
    res.distant_candidate_mae = a.distant_candidate_mae.data();
    res.distant_match_decision = a.distant_match_decision.data();
    vc1me_planes pc{f.y.data(), f.u.data(), f.v.data(), f.width, f.height};
    vc1me_planes pr{r.y.data(), r.u.data(), r.v.data(), r.width, r.height};
    vc1me_frame *fc = nullptr, *fr = nullptr;
    if (vc1me_frame_upload(ctx, &pc, &fc) || vc1me_frame_upload(ctx, &pr, &fr) ||
        vc1me_analyze_p_frame(ctx, fc, fr, &p, &res))
        std::printf("direct analysis failed: %s\n", vc1me_last_error());
    vc1me_frame_release(ctx, fc);
    vc1me_frame_release(ctx, fr);
    a.mean_abs_residual = res.mean_abs_residual; a.rate_complexity = res.rate_complexity;
    a.motion_sad = res.motion_sad; a.moved_macroblocks = res.moved_macroblocks;
    a.fractional_chroma_macroblocks = res.fractional_chroma_macroblocks;
    a.four_mv_macroblocks = res.four_mv_macroblocks; a.intra_macroblocks = res.intra_macroblocks;
    return a;
}

}  // namespace

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "full";
    std::string error;
    auto hc = HipComputeContext::create(0, error);
    if (!hc) {
        std::printf("FAIL create: %s\n", error.c_str());
        return 2;
    }
    const Frame ref = picture(160, 96, 0, 0, 7);
    const Frame cur = picture(160, 96, 3, -2, 7);
    const auto cands = integer_candidates(36, 11);
    if (mode == "inject") {
        // VC1ME_FAIL_AT=1: the first device call of the process fails
        bool gpu = true;
        std::string why;
        const auto costs = score_or_rescore(*hc, cur, ref, 32, 16, 16, 16, 1, cands, &gpu, &why);
        check(!gpu && why.find("VC1ME_FAIL_AT") != std::string::npos, "injected failure reported: " + why);
        bool again = false;
        const auto later = score_or_rescore(*hc, cur, ref, 32, 16, 16, 16, 1, cands, &again, &why);
        check(again && later == costs, "next call on the GPU equals the CPU re-score");
        std::printf(g_failures ? "ADAPTER FAILED\n" : "ADAPTER OK\n");
        return g_failures ? 1 : 0;
    }
    // scorer against the CPU law, stride 1 and 4, a block at the edge
    for (const int s : {1, 4})
        for (const auto& b : {std::array<int, 4>{32, 16, 16, 16}, std::array<int, 4>{152, 88, 16, 16},
                              std::array<int, 4>{8, 40, 8, 8}}) {
            bool gpu = false;
            std::string why;
            const auto got = score_or_rescore(*hc, cur, ref, b[0], b[1], b[2], b[3], s, cands, &gpu, &why);
            bool eq = gpu;
            for (size_t i = 0; eq && i < cands.size(); ++i)
                eq = got[i] == cpu_sad(cur, ref, b[0], b[1], b[2], b[3], cands[i].xq / 4, cands[i].yq / 4, s);
            check(eq, "score_motion_batch equals the CPU law, stride " + std::to_string(s) + " block at (" +
                          std::to_string(b[0]) + "," + std::to_string(b[1]) + ")");
        }
    std::vector<uint64_t> none;
    check(hc->score_motion_batch(cur.y.data(), ref.y.data(), 0, 0, 0, 0, 16, 16, 1, false, false,
                                 VulkanMotionMetric::Sad, cands.data(), cands.size(), none, &error),
          "degenerate request succeeds");
    const VulkanComputeCounters k = hc->counters();
    check(k.dispatches == 6 && k.candidates == 6 * cands.size(), "counters count the dispatches");
    // whole-frame analysis against the C ABI called directly
    EncoderConfig cfg;
    cfg.width = 160; cfg.height = 96; cfg.pqindex = 8; cfg.motion_search_range = 128;
    Vc1Encoder::PAnalysis a;
    check(hc->analyze_p(cur, ref, cfg, Vc1Encoder::ProgressiveMvMode::MixedMv, true, nullptr, a, &error),
          "analyze_p runs");
    check(same(a, direct(hc->raw(), cur, ref, cfg)), "analyze_p equals vc1me_analyze_p_frame");
    check(a.motion_sad > 0 && a.moved_macroblocks > 0, "the pan is found");
    Vc1Encoder::PAnalysis q, seeded, seeded2;
    check(hc->analyze_p(cur, ref, cfg, Vc1Encoder::ProgressiveMvMode::OneMvQpel, false, nullptr, q, &error) &&
              hc->analyze_p(cur, ref, cfg, Vc1Encoder::ProgressiveMvMode::MixedMv, true, &q, seeded, &error) &&
              hc->analyze_p(cur, ref, cfg, Vc1Encoder::ProgressiveMvMode::MixedMv, true, &q, seeded2, &error) &&
              same(seeded, seeded2),
          "seeded analysis runs and repeats");
    // one context shared by two host threads: every result equals the
    // sequential one
    const Frame cur2 = picture(160, 96, -5, 4, 9), ref2 = picture(160, 96, 0, 0, 9);
    Vc1Encoder::PAnalysis want2;
    hc->analyze_p(cur2, ref2, cfg, Vc1Encoder::ProgressiveMvMode::MixedMv, true, nullptr, want2, &error);
    bool gpu = false;
    std::string why;
    const auto want_costs = score_or_rescore(*hc, cur2, ref2, 48, 32, 16, 16, 1, cands, &gpu, &why);
    bool ok1 = true, ok2 = true;
    std::thread t1([&] {
        for (int it = 0; it < 6; ++it) {
            Vc1Encoder::PAnalysis r;
            std::string e;
            ok1 = ok1 && hc->analyze_p(cur, ref, cfg, Vc1Encoder::ProgressiveMvMode::MixedMv, true, nullptr, r, &e) &&
                  same(r, a);
        }
    });
    std::thread t2([&] {
        for (int it = 0; it < 6; ++it) {
            Vc1Encoder::PAnalysis r;
            std::string e;
            ok2 = ok2 && hc->analyze_p(cur2, ref2, cfg, Vc1Encoder::ProgressiveMvMode::MixedMv, true, nullptr, r, &e) &&
                  same(r, want2);
            bool g = false;
            ok2 = ok2 && score_or_rescore(*hc, cur2, ref2, 48, 32, 16, 16, 1, cands, &g, &e) == want_costs && g;
        }
    });
    t1.join();
    t2.join();
    check(ok1 && ok2, "two threads sharing one context get the sequential results");
    std::printf(g_failures ? "ADAPTER FAILED\n" : "ADAPTER OK\n");
    return g_failures ? 1 : 0;
}
