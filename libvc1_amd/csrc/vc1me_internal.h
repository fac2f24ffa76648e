// vc1me_internal.h — types shared by the HIP kernels and the host layer of
// the MI355X motion-estimation backend.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vc1me {

constexpr uint64_t kBadSad = UINT64_MAX / 4;   // hidden-padding veto cost (src/encoder_motion.cpp:2201-2203)
constexpr int kWave = 64;
constexpr int kMaxLocalRange = 64;             // LDS window bound of the integer search
constexpr int kGlobalMaxPts = 512;

// Frame-level global translation search state (src/encoder_motion.cpp:2025-2141).
struct GlobalState {
    int32_t n;                  // points of the current stage
    int32_t stride;             // frame-SAD sample stride of the stage
    int32_t pts[2 * kGlobalMaxPts];
    uint64_t cost[kGlobalMaxPts];
    // sorted candidate list of the extended search ("top", <= 8 entries)
    int32_t tn;
    uint64_t tcost[9];
    int32_t tdx[9], tdy[9];
    // result
    int32_t dx, dy;
    uint64_t full_sad;
    int32_t pad;
};

// One (cur, ref) pair of the batch with its parameters and device state.
struct JobDesc {
    const uint8_t* cy; const uint8_t* cu; const uint8_t* cv;
    const uint8_t* ry; const uint8_t* ru; const uint8_t* rv;
    int32_t W, H, vw, vh, mbw, mbh;
    int32_t pq, range, local_range, algo, quality, subpel, fast, subpel_range, lr_mode, mode, allow_intra;
    int32_t mvrange;
    double lambda;              // max(1, 0.35*pq) as the reference computes it
    uint64_t lam_m;             // lambda == lam_m * 2^-lam_k exactly
    int32_t lam_k;
    int32_t ncx, ncy;           // long-range index lattice
    double intra_thr;
    // distant_match_max_mae threshold per active pixel count: reject iff
    // sad > rej_thr[px] (rej_always[px] => always reject)
    uint64_t rej_thr[257];
    uint8_t rej_always[257];
    // state / outputs (device), one entry per macroblock
    int2* mvs; int2* bmv; uint8_t* flags; uint8_t* prop;
    uint64_t* screen; double* ratio; double* dl_mae; double* dc_mae; uint8_t* decision;
    uint64_t* chosen; uint64_t* temporal;
    uint64_t* sig;              // long-range signatures (ncx*ncy)
    GlobalState* g;
    int32_t* err;
    // algorithmic operation counters of this job (sad_px, interp_taps,
    // satd_ops, sig_ops, frame_px), in the units of the CPU oracle
    unsigned long long* ops;
    int32_t max_diag;           // last anti-diagonal index mbw + 2*mbh - 3
    // extended long-range search result per macroblock, kept for the step
    // once computed (ext_done): best SAD, its vector, and the op counts
    // (signature cells, SAD pixels)
    uint64_t* ext_sad; int2* ext_mv; uint32_t* ext_ops; uint8_t* ext_done;
    // interpolated reference: 16 bicubic (qp) or 4 bilinear (bp) phase planes
    // of (W+3) x (H+3) samples at pitch pp, row count ph = H + 3
    const uint8_t* qp;
    const uint8_t* bp;
    int32_t pp, ph;
    // learned seeding: proposals per macroblock (5 floats: two candidates in
    // 64-px units, confidence logit) or null, and the logit threshold
    const float* nn_out;
    float nn_logit;
    int32_t nn_candidates;      // 0: the proposals' vectors are not used (confidence gate only)
    // cost report (null = off): per macroblock 20 * (SAD + lambda * R) of the
    // final 1-MV vector and that vector; lam20 = 20 * lambda = max(20, 7 * pq)
    uint64_t* cost20;
    int2* one_mv;
    int32_t lam20;
    // this job computes its reference's phase planes in the step (the first
    // job of the step with that reference and filter)
    int32_t phase_owner;
    // algorithmic operation counts of each macroblock's last analysis:
    // sad_px, interp_taps, satd_ops, sig_ops, seeded (5 per macroblock),
    // summed into ops[] once the macroblocks are final
    uint64_t* mbops;
    // relaxation schedule: the queue stamp of each macroblock (p = already
    // queued for pass p); the lists themselves are the step's RelaxQueue
    uint32_t* mark;
    // packed copy of the per-macroblock outputs written at the end of the
    // step (PackedOut layout), or null when the results stay on the device
    uint8_t* pack;
    // seeded core (null: the full search): each macroblock's 1-MV start
    // vector, quarter-pel (analyze_p_picture_core's shared seed)
    const int2* seed_mvs;
    // motion-failure check (has_good_motion_match): the integer and long-range
    // stages, then refine_qpel_block; flags carry kFlagGoodMatch
    int32_t motion_check;
};

// Longest relaxation run before the step falls back to the anti-diagonal
// sweep (which is exact from any starting state).
constexpr int kMaxRelaxPasses = 96;

// A worklist entry: job (0..254) in the top 8 bits, macroblock index (below
// 2^24) in the low 24, packed and unpacked as unsigned (a job of 128 or more
// sets the top bit).
constexpr int kMaxRelaxMacroblocks = 1 << 24;
__host__ __device__ inline uint32_t relax_entry(int job, int mb) {
    return (static_cast<uint32_t>(job) << 24) | static_cast<uint32_t>(mb);
}
__host__ __device__ inline int relax_entry_job(uint32_t e) { return static_cast<int>(e >> 24); }
__host__ __device__ inline int relax_entry_pos(uint32_t e) { return static_cast<int>(e & 0xffffffu); }

// XCD bands of a step's (macroblocks x jobs) grid of `total` items in
// job-major order (item = job * per + macroblock): the dispatcher deals a
// grid's workgroups to the eight XCDs in turn, and the whole-frame kernels
// give the workgroups on XCD x the contiguous items of band x.  Bands
// 0..r-1 hold q + 1 items, the others q (total = 8q + r); xcd_band_of is the
// inverse map.
constexpr int kXcds = 8;
__host__ __device__ inline int xcd_band_start(int total, int x) {
    const int q = total / kXcds, r = total % kXcds;
    return x * q + (x < r ? x : r);
}
__host__ __device__ inline int xcd_band_size(int total, int x) { return total / kXcds + (x < total % kXcds ? 1 : 0); }
__host__ __device__ inline int xcd_band_of(int total, int item) {
    const int q = total / kXcds, r = total % kXcds, big = r * (q + 1);
    return item < big ? item / (q + 1) : r + (item - big) / q;   // q > 0 whenever item >= big
}

// The relaxation's worklists over all jobs of a step: pass p reads cnt[p]
// entries of wl[p & 1] (relax_entry(job, macroblock)) by ticket (tk[p]) and
// appends to wl[(p + 1) & 1]; cnt and tk hold kMaxRelaxPasses + 1 entries
// each, cap is the step's macroblock count.
struct RelaxQueue {
    int32_t* cnt;
    int32_t* tk;
    int32_t* wl[2];
    int32_t cap;
};

// Per-job record of a finished step (k_step_end): error flag, global vector,
// op totals (sad_px, interp_taps, satd_ops, sig_ops, frame_px, seeded) and
// the frame aggregates as exact integers.
struct StepSummary {
    int32_t err;
    int32_t gdx, gdy;
    int32_t pad;
    uint64_t global_sad;
    unsigned long long ops[8];
    uint64_t total_sad, rate_total, total_px, moved, frac, four, intra;
    uint64_t pad2;
};

// Packed per-macroblock outputs of one job: each field contiguous, 256-byte
// aligned, in the order below (layout shared by the kernel and the host).
struct PackedOut {
    int2* mvs;
    int2* bmv;
    uint64_t* screen;
    double* ratio;
    double* dl;
    double* dc;
    uint8_t* four;
    uint8_t* intra;
    uint8_t* decision;
    size_t bytes;
};
__host__ __device__ inline size_t pack_align(size_t v) { return (v + 255) & ~static_cast<size_t>(255); }
__host__ __device__ inline PackedOut pack_layout(uint8_t* base, int mbs) {
    PackedOut p;
    const size_t n = static_cast<size_t>(mbs);
    size_t o = 0;
    p.mvs = reinterpret_cast<int2*>(base + o); o += pack_align(n * 8);
    p.bmv = reinterpret_cast<int2*>(base + o); o += pack_align(n * 32);
    p.screen = reinterpret_cast<uint64_t*>(base + o); o += pack_align(n * 8);
    p.ratio = reinterpret_cast<double*>(base + o); o += pack_align(n * 8);
    p.dl = reinterpret_cast<double*>(base + o); o += pack_align(n * 8);
    p.dc = reinterpret_cast<double*>(base + o); o += pack_align(n * 8);
    p.four = base + o; o += pack_align(n);
    p.intra = base + o; o += pack_align(n);
    p.decision = base + o; o += pack_align(n);
    p.bytes = o;
    return p;
}

// Phase-plane geometry of a W x H reference: sample (px, py), px in [-2, W],
// py in [-2, H], lives at row py + 2, column px + 2.  The pitch leaves room
// for the 8-byte reads of the block scorer.
inline int32_t phase_pitch(int w) { return (w + 3 + 8 + 15) & ~15; }
inline int32_t phase_rows(int h) { return h + 3; }

// Flag bits of JobDesc::flags
constexpr uint8_t kFlag4mvChosen = 1;
constexpr uint8_t kFlagIntra = 2;
constexpr uint8_t kFlagGoodMatch = 4;   // motion-failure check: a usable match

// Kernel launchers (vc1me_kernels.hip).
hipError_t launch_global_search(const JobDesc* d_jobs, int njobs, const JobDesc* h_jobs, hipStream_t s,
                                hipEvent_t ev0, hipEvent_t ev1);
hipError_t launch_signatures(const JobDesc* d_jobs, int njobs, int max_ncx, int max_ncy, hipStream_t s);
hipError_t launch_mb_diagonal(const JobDesc* d_jobs, int njobs, int diag, int max_per_diag, int lds_bytes,
                              hipStream_t s);
// extended long-range search of every macroblock of the plain jobs, ahead of
// the anti-diagonal sweep (the relaxation schedule computes it on demand)
hipError_t launch_extended(const JobDesc* d_jobs, int njobs, int max_mbs, int lds_bytes, hipStream_t s);
// relaxation schedule: starting state, guess pass (by rows, as pass 0), pass 0 over every
// macroblock (launch `phase` of `phases` by rows; max_items = max over jobs
// of mbw * ceil(mbh / phases)), and a work-queue pass over pass p's list
// (grid: resident workgroups, relax_queue_grid)
hipError_t launch_relax_init(const JobDesc* d_jobs, int njobs, int max_mbs, hipStream_t s);
hipError_t launch_mb_guess(const JobDesc* d_jobs, int njobs, int max_items, int phase, int phases, int lds_bytes,
                           hipStream_t s);
hipError_t launch_mb_relax0(const JobDesc* d_jobs, int njobs, int max_items, int phase, int phases, int lds_bytes,
                            const RelaxQueue& q, hipStream_t s);
hipError_t launch_mb_relaxq(const JobDesc* d_jobs, int pass, int grid, int lds_bytes, const RelaxQueue& q,
                            hipStream_t s);
int relax_queue_grid(int lds_bytes, int cu_count);
// start of a step (clears the per-job flags, counters and marks and the
// first `ncounters` ints at `counters`) and end of a step (per-job summary
// and, where J->pack is set, the packed outputs)
hipError_t launch_step_begin(const JobDesc* d_jobs, int njobs, int max_mbs, int32_t* counters, int ncounters,
                             hipStream_t s);
hipError_t launch_step_end(const JobDesc* d_jobs, int njobs, StepSummary* out, hipStream_t s);
// the cost report of the final field (jobs with cost20 set)
hipError_t launch_cost_final(const JobDesc* d_jobs, int njobs, int max_mbs, int lds_bytes, hipStream_t s);
int mb_kernel_lds_bytes(int local_range);
// phase planes of the step's owning jobs (k_phase_planes_jobs)
hipError_t launch_phase_planes_jobs(const JobDesc* d_jobs, int njobs, int max_w, int max_h, hipStream_t s);
hipError_t launch_phase_planes(const uint8_t* ref, int width, int height, uint8_t* out, int pitch, int rows,
                               int bilinear, hipStream_t s);
hipError_t launch_block_costs(const JobDesc* d_job, int x0, int y0, int bw, int bh, int mode, int metric,
                              const int2* cands, int n, uint64_t* out, int lds_bytes, hipStream_t s);
hipError_t launch_score_batch(const uint8_t* cur, const uint8_t* ref, int width, int height, int x0, int y0,
                              int bw, int bh, int sample_stride, int bilinear, int rnd, int metric,
                              const int2* cands, int n, uint32_t* costs, hipStream_t s);

// B-picture decision (analyze_b_picture_mode after its two core passes):
// for picture i, job 2i of the step is the forward pass (cur vs past), job
// 2i + 1 the backward pass (cur vs future); their vectors, screen SADs and
// phase planes are read in place.
struct BDesc {
    const int2* anchor_mvs;      // future anchor's 1-MV field
    const uint8_t* anchor_4mv;
    int32_t mode, bfrac, below_half, allow_intra;
    double lambda;               // max(1, 0.45 * PQINDEX)
    uint64_t intra_guard;        // llround(lambda * 72)
    double intra_thr;
    // per macroblock: costs of direct, forward, backward, interpolated
    // (invalid markers and the 4-MV anchor penalty applied), the intra
    // proxy SAD, the direct vectors (forward, backward)
    uint64_t* cost;
    uint64_t* intra_sad;
    int2* direct;
    // decisions: mode, the final forward and backward fields, the intra
    // ratio, and the per-macroblock terms of the picture totals
    uint8_t* modes;
    int2* fmv;
    int2* bmv;
    double* ratio;
    uint64_t* sel_sad;
    uint64_t* rate_sad;
    uint32_t* sel_rate;
};
// nb B pictures: picture i uses jobs 2i (forward) and 2i + 1 (backward) and
// descriptor i; the decision runs one workgroup per picture with one thread
// per macroblock row
constexpr int kBMaxRows = 1024;
hipError_t launch_b_costs(const JobDesc* d_jobs, const BDesc* d_b, int nb, int max_mbs, hipStream_t s);
hipError_t launch_b_decide(const JobDesc* d_jobs, const BDesc* d_b, int nb, int max_mbh, hipStream_t s);

// The P prediction picture of an analysis (k_predict_luma, k_predict_chroma):
// a DW x DH picture (the coded raster, or the macroblock raster for the
// padded picture of the loop reconstruction) sampled from a reference of
// raster RW x RH.
hipError_t launch_predict_p(const uint8_t* ry, const uint8_t* ru, const uint8_t* rv, int RW, int RH, int DW, int DH,
                            int mode, const int2* mvs, const int2* bmv, const uint8_t* four, uint8_t* dy, uint8_t* du,
                            uint8_t* dv, hipStream_t s);

// The bi-directional prediction picture (k_predict_bi); past and fut are
// frame buffers with the two chroma planes after luma, of rasters PW x PH
// and FW x FH; the picture is DW x DH.
hipError_t launch_predict_bi(const uint8_t* past, int PW, int PH, const uint8_t* fut, int FW, int FH, int DW, int DH,
                             int mode, const int2* fmv, const int2* bmv, uint8_t* dy, uint8_t* du, uint8_t* dv,
                             hipStream_t s);

// Intensity compensation (vc1me_intensity.hip): the luma table at bytes
// 0..255 and the chroma table at 256..511.
struct IntensityLuts {
    uint32_t w[128];
};
// The sampler and the SAD grid write one row of partial sums per block (4
// and 49 values); ic_*_blocks give the block counts for n samples.
int ic_sample_blocks(int n);
int ic_sad_blocks(int n);
hipError_t launch_ic_samples(const uint8_t* cy, const uint8_t* ry, int W, int H, const int2* mvs, int mbw,
                             uint16_t* samples, unsigned long long* sums, hipStream_t s);
hipError_t launch_ic_sads(const uint16_t* samples, int n, int l0, int z0, unsigned long long* sads, hipStream_t s);
hipError_t launch_ic_remap(const uint8_t* src, uint8_t* dst, size_t bytes, size_t luma_bytes,
                           const IntensityLuts& luts, hipStream_t s);

}  // namespace vc1me
