// vc1me_api.cpp — host side of the C ABI declared in include/vc1me.h.
//
// Owns the HIP device, stream, resident frames and per-job device state, and
// drives the kernels of vc1me_kernels.hip.  Everything here is plain HIP
// runtime + C++17; no torch types.
#include "../../include/vc1me.h"
#include "vc1me_internal.h"
#include "vc1me_nn_internal.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

using vc1me::GlobalState;
using vc1me::JobDesc;

struct vc1me_frame {
    uint8_t* y = nullptr;
    uint8_t* u = nullptr;
    uint8_t* v = nullptr;
    int32_t w = 0, h = 0;
    // interpolated phase planes of this frame as a reference (bicubic 16,
    // bilinear 4) for vc1me_block_costs only, allocated on its first use;
    // analysis steps use the context's step pool instead
    mutable uint8_t* qp = nullptr;
    mutable uint8_t* bp = nullptr;
};

namespace {

thread_local std::string g_last_error;

// (cur, ref) pairs of one step
constexpr int kMaxJobs = 255;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(VC1ME_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Fault injection (VC1ME_FAIL_AT=N makes the N-th device call fail).
std::atomic<long long> g_calls{0};
bool injected_failure() {
    static const long long at = [] {
        const char* e = std::getenv("VC1ME_FAIL_AT");
        return e && *e ? std::atoll(e) : 0LL;
    }();
    const long long n = ++g_calls;
    return at > 0 && n == at;
}

// Pass limit of the relaxation schedule; VC1ME_RELAX_MAX_PASSES lowers it
// (read at every step) so that the diagonal fallback can be exercised.
int relax_pass_limit() {
    const char* e = std::getenv("VC1ME_RELAX_MAX_PASSES");
    const int v = e && *e ? std::atoi(e) : vc1me::kMaxRelaxPasses;
    return std::max(1, std::min(v, vc1me::kMaxRelaxPasses));
}

// Relaxation passes a launch shape starts with before its need is known
// (VC1ME_RELAX_FIRST_PASSES lowers or raises it; read at every step), so
// that the continuation of an unsettled step can be exercised.
int relax_first_passes() {
    const char* e = std::getenv("VC1ME_RELAX_FIRST_PASSES");
    const int v = e && *e ? std::atoi(e) : 24;
    return std::max(1, std::min(v, vc1me::kMaxRelaxPasses));
}

// Launches of the relaxation's pass 0, by rows (VC1ME_RELAX_PASS0_PHASES;
// read at every step): launch r analyses the rows y with y % phases == r, so
// that the macroblocks of later launches read their top neighbours' pass-0
// state and are not queued again for it.
int relax_pass0_phases() {
    const char* e = std::getenv("VC1ME_RELAX_PASS0_PHASES");
    const int v = e && *e ? std::atoi(e) : 4;
    return std::max(1, std::min(v, 16));
}

// Launches of the guess pass, by rows as pass 0 (VC1ME_RELAX_GUESS_PHASES;
// read at every step).
int relax_guess_phases() {
    const char* e = std::getenv("VC1ME_RELAX_GUESS_PHASES");
    const int v = e && *e ? std::atoi(e) : 4;
    return std::max(1, std::min(v, 16));
}

// Guess pass before the relaxation's pass 0 (VC1ME_RELAX_GUESS=0 turns it off).
bool relax_guess() {
    const char* e = std::getenv("VC1ME_RELAX_GUESS");
    return !(e && *e == '0');
}

// ---- scratch of the allocation-free batch scorer
// (vc1me_score_batch_frames_noalloc): the candidate list staged for the
// kernel, then its 32-bit costs.  One copy per GPU, held until the process
// ends; host code never uses the array's name as an address, only the
// pointer hipGetSymbolAddress returned for the current device.
constexpr size_t kNoAllocCandBytes = static_cast<size_t>(VC1ME_NOALLOC_MAX_CANDIDATES) * sizeof(int2);
constexpr size_t kNoAllocScratchBytes =
    kNoAllocCandBytes + static_cast<size_t>(VC1ME_NOALLOC_MAX_CANDIDATES) * sizeof(uint32_t);
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kMaxScratchDevices = 64;
struct DeviceScratch {
    std::mutex mu;          // serialises the allocation-free calls on this GPU
    void* ptr = nullptr;    // device address of this GPU's copy of `scratch`
    bool ready = false;
};
DeviceScratch g_scratch[kMaxScratchDevices];

// Looks up this device's copy of the scratch array once (the device must be
// current).  Returns its address, or null with the error in `err`.
void* scratch_for_device(int device, hipError_t* err) {
    *err = hipSuccess;
    if (device < 0 || device >= kMaxScratchDevices) { *err = hipErrorInvalidDevice; return nullptr; }
    DeviceScratch& d = g_scratch[device];
    std::lock_guard<std::mutex> lock(d.mu);
    if (!d.ready) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) { *err = e != hipSuccess ? e : hipErrorInvalidSymbol; return nullptr; }
        d.ptr = ptr;
        d.ready = true;
    }
    return d.ptr;
}

uint64_t fnv64(const uint8_t* p, size_t n) {
    uint64_t h = 1469598103934665603ull;
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        std::memcpy(&w, p + i, 8);
        h = (h ^ w) * 1099511628211ull;
    }
    for (; i < n; ++i) h = (h ^ p[i]) * 1099511628211ull;
    return h;
}

struct JobSlot {
    size_t mbs_cap = 0, cells_cap = 0;
    void* mem = nullptr;  // one device allocation carved into the arrays below
    size_t bytes = 0;
    int2* mvs = nullptr;
    int2* bmv = nullptr;
    uint8_t* flags = nullptr;
    uint8_t* prop = nullptr;
    uint64_t* screen = nullptr;
    double* ratio = nullptr;
    double* dl = nullptr;
    double* dc = nullptr;
    uint8_t* decision = nullptr;
    uint64_t* chosen = nullptr;
    uint64_t* temporal = nullptr;
    uint64_t* sig = nullptr;
    GlobalState* g = nullptr;
    int32_t* err = nullptr;
    unsigned long long* ops = nullptr;
    uint64_t* ext_sad = nullptr;
    int2* ext_mv = nullptr;
    uint32_t* ext_ops = nullptr;
    uint8_t* ext_done = nullptr;
    uint64_t* cost20 = nullptr;     // cost report
    int2* one_mv = nullptr;
    uint64_t* mbops = nullptr;      // per-macroblock op counts of the last analysis
    uint32_t* mark = nullptr;       // relaxation queue stamps
    int2* seed_mvs = nullptr;       // seeded core: the seed's 1-MV vectors
    int mbs_last = 0;               // macroblocks of the last step's job
};

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

hipError_t ensure_slot(JobSlot& s, size_t mbs, size_t cells) {
    if (s.mem && s.mbs_cap >= mbs && s.cells_cap >= cells) return hipSuccess;
    if (s.mem) { (void)hipFree(s.mem); s.mem = nullptr; }
    mbs = std::max(mbs, s.mbs_cap);
    cells = std::max(cells, s.cells_cap);
    size_t off = 0;
    auto carve = [&](size_t n) { const size_t o = off; off = align_up(off + n, 256); return o; };
    const size_t o_mvs = carve(mbs * sizeof(int2));
    const size_t o_bmv = carve(mbs * 4 * sizeof(int2));
    const size_t o_flags = carve(mbs);
    const size_t o_prop = carve(mbs);
    const size_t o_screen = carve(mbs * 8);
    const size_t o_ratio = carve(mbs * 8);
    const size_t o_dl = carve(mbs * 8);
    const size_t o_dc = carve(mbs * 8);
    const size_t o_dec = carve(mbs);
    const size_t o_chosen = carve(mbs * 8);
    const size_t o_temp = carve(mbs * 8);
    const size_t o_sig = carve(std::max<size_t>(cells, 1) * 8);
    const size_t o_g = carve(sizeof(GlobalState));
    const size_t o_err = carve(sizeof(int32_t));
    const size_t o_ops = carve(8 * sizeof(unsigned long long));
    const size_t o_exs = carve(mbs * 8);
    const size_t o_exm = carve(mbs * sizeof(int2));
    const size_t o_exo = carve(mbs * 2 * sizeof(uint32_t));
    const size_t o_exd = carve(mbs);
    const size_t o_c20 = carve(mbs * 8);
    const size_t o_one = carve(mbs * sizeof(int2));
    const size_t o_mbops = carve(mbs * 5 * sizeof(uint64_t));
    const size_t o_mark = carve(mbs * sizeof(uint32_t));
    const size_t o_seed = carve(mbs * sizeof(int2));
    hipError_t e = hipMalloc(&s.mem, off);
    if (e != hipSuccess) { s.mem = nullptr; return e; }
    uint8_t* b = static_cast<uint8_t*>(s.mem);
    s.bytes = off;
    s.mvs = reinterpret_cast<int2*>(b + o_mvs);
    s.bmv = reinterpret_cast<int2*>(b + o_bmv);
    s.flags = b + o_flags;
    s.prop = b + o_prop;
    s.screen = reinterpret_cast<uint64_t*>(b + o_screen);
    s.ratio = reinterpret_cast<double*>(b + o_ratio);
    s.dl = reinterpret_cast<double*>(b + o_dl);
    s.dc = reinterpret_cast<double*>(b + o_dc);
    s.decision = b + o_dec;
    s.chosen = reinterpret_cast<uint64_t*>(b + o_chosen);
    s.temporal = reinterpret_cast<uint64_t*>(b + o_temp);
    s.sig = reinterpret_cast<uint64_t*>(b + o_sig);
    s.g = reinterpret_cast<GlobalState*>(b + o_g);
    s.err = reinterpret_cast<int32_t*>(b + o_err);
    s.ops = reinterpret_cast<unsigned long long*>(b + o_ops);
    s.ext_sad = reinterpret_cast<uint64_t*>(b + o_exs);
    s.ext_mv = reinterpret_cast<int2*>(b + o_exm);
    s.ext_ops = reinterpret_cast<uint32_t*>(b + o_exo);
    s.ext_done = b + o_exd;
    s.cost20 = reinterpret_cast<uint64_t*>(b + o_c20);
    s.one_mv = reinterpret_cast<int2*>(b + o_one);
    s.mbops = reinterpret_cast<uint64_t*>(b + o_mbops);
    s.mark = reinterpret_cast<uint32_t*>(b + o_mark);
    s.seed_mvs = reinterpret_cast<int2*>(b + o_seed);
    s.mbs_cap = mbs;
    s.cells_cap = cells;
    return hipSuccess;
}

struct CachedPlane {
    const uint8_t* host = nullptr;
    int w = 0, h = 0;
    bool by_generation = false;     // sum is the caller's generation, not a checksum
    uint64_t sum = 0;
    uint8_t* dev = nullptr;
    uint64_t last_use = 0;
};

}  // namespace

struct vc1me_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream of a step: work that depends on the frames alone (phase
    // planes, signature index) runs there beside the global search, forked
    // from and joined back to `stream` by the two events
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    std::mutex mu;
    std::string name;
    vc1me_counters counters{};
    vc1me_timing timing{};
    vc1me_work work{};
    bool profiling = false;
    bool cost_report = false;
    bool use_graphs = false;
    std::vector<std::pair<uint8_t*, size_t>> phase_pool;   // phase planes of the step's references
    std::vector<std::pair<std::string, hipGraphExec_t>> graphs;   // captured steps by launch shape
    int last_jobs = 0;
    std::vector<JobSlot> slots;
    JobDesc* d_jobs = nullptr;
    size_t d_jobs_cap = 0;
    std::vector<JobDesc> h_jobs;
    // events: 0/1 whole step, 2/3 global search, then pairs per diagonal
    std::vector<hipEvent_t> events;
    // score-batch resources
    CachedPlane planes[8];
    uint64_t use_clock = 0;
    int2* d_cands = nullptr;
    uint32_t* d_costs = nullptr;
    size_t cand_cap = 0;
    // pinned staging of the scorer's candidates and costs (the allocating
    // forms grow it; the allocation-free form uses a fixed one made at
    // creation, as large as its scratch array)
    uint8_t* h_stage = nullptr;
    size_t stage_cap = 0;
    uint8_t* h_noalloc = nullptr;
    int cu_count = 256;
    // proposals handed in by the host (vc1me_analyze_step_proposals)
    float* d_props = nullptr;
    size_t props_cap = 0;
    // macroblock schedule (VC1ME_SCHEDULE_*); the relaxation's two
    // worklists over all jobs of a step (2 x relax_wl_cap entries), its
    // work-queue grid, and the pass count each launch shape last needed
    int schedule = VC1ME_SCHEDULE_RELAX;
    int32_t* d_relax_wl = nullptr;
    size_t relax_wl_cap = 0;
    int relax_grid = 0, relax_grid_lds = -1;
    std::vector<std::pair<std::string, int>> pass_hints;
    // a step's results: the relaxation counters, one summary per job and the
    // packed per-macroblock outputs, on the device and in pinned host memory
    // (one copy per step)
    uint8_t* d_out = nullptr;
    uint8_t* h_out = nullptr;
    size_t out_cap = 0;
    // intensity estimation: vectors, samples, the four sums and 49 SADs
    uint8_t* d_ic = nullptr;
    size_t ic_cap = 0;
    // learned seeding: use the proposals' vectors (1) or only their
    // confidence gate (0)
    int learned_candidates = 1;
    // B-picture decision buffers
    uint8_t* d_b = nullptr;
    size_t b_cap = 0;
    // last step's relaxation passes: analyses per pass (summed over jobs) and,
    // with profiling on, device milliseconds per pass
    std::vector<int64_t> pass_analyses;
    std::vector<double> pass_ms;
    // this device's copy of the allocation-free scorer's scratch array (null
    // if its lookup failed at creation; scratch_error says why)
    void* scratch = nullptr;
    hipError_t scratch_error = hipSuccess;
};

namespace {

int set_device(vc1me_ctx* ctx) {
    HIP_TRY(hipSetDevice(ctx->device));
    return VC1ME_OK;
}

int ensure_events(vc1me_ctx* ctx, size_t n) {
    while (ctx->events.size() < n) {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        ctx->events.push_back(e);
    }
    return VC1ME_OK;
}

// Device copy of a host luma plane for the batch scorer; keyed by pointer,
// size and either a checksum of the whole plane (so in-place edits are never
// missed) or, with by_generation, the caller's generation number of it (no
// pass over the plane: the caller changes the generation when it writes the
// plane).
int plane_for(vc1me_ctx* ctx, const uint8_t* host, int w, int h, uint8_t** out, bool by_generation = false,
              uint64_t generation = 0) {
    const size_t n = static_cast<size_t>(w) * h;
    const uint64_t sum = by_generation ? generation : fnv64(host, n);
    ++ctx->use_clock;
    for (auto& p : ctx->planes)
        if (p.dev && p.host == host && p.w == w && p.h == h && p.by_generation == by_generation && p.sum == sum) {
            p.last_use = ctx->use_clock;
            *out = p.dev;
            return VC1ME_OK;
        }
    CachedPlane* victim = &ctx->planes[0];
    for (auto& p : ctx->planes) {
        if (!p.dev) { victim = &p; break; }
        if (p.last_use < victim->last_use) victim = &p;
    }
    if (victim->dev && (static_cast<size_t>(victim->w) * victim->h < n)) {
        (void)hipFree(victim->dev);
        victim->dev = nullptr;
    }
    if (!victim->dev) HIP_TRY(hipMalloc(&victim->dev, n));
    HIP_TRY(hipMemcpyAsync(victim->dev, host, n, hipMemcpyHostToDevice, ctx->stream));
    victim->host = host; victim->w = w; victim->h = h; victim->sum = sum; victim->last_use = ctx->use_clock;
    victim->by_generation = by_generation;
    ctx->counters.frame_uploads++;
    ctx->counters.frame_upload_bytes += n;
    *out = victim->dev;
    return VC1ME_OK;
}

int score_on_device(vc1me_ctx* ctx, const uint8_t* dcur, const uint8_t* dref, int width, int height, int x0, int y0,
                    int bw, int bh, int sample_stride, int bilinear, int rnd, int metric,
                    const vc1me_candidate* candidates, size_t n, uint64_t* costs_out) {
    if (n > ctx->cand_cap) {
        if (ctx->d_cands) (void)hipFree(ctx->d_cands);
        if (ctx->d_costs) (void)hipFree(ctx->d_costs);
        ctx->d_cands = nullptr; ctx->d_costs = nullptr; ctx->cand_cap = 0;
        const size_t cap = std::max<size_t>(n, 4096);
        HIP_TRY(hipMalloc(&ctx->d_cands, cap * sizeof(int2)));
        HIP_TRY(hipMalloc(&ctx->d_costs, cap * sizeof(uint32_t)));
        ctx->cand_cap = cap;
    }
    const size_t stage = n * (sizeof(int2) + sizeof(uint32_t));
    if (ctx->stage_cap < stage) {
        if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
        ctx->h_stage = nullptr;
        ctx->stage_cap = 0;
        const size_t cap = std::max(stage, size_t(4096) * (sizeof(int2) + sizeof(uint32_t)));
        HIP_TRY(hipHostMalloc(&ctx->h_stage, cap, hipHostMallocDefault));
        ctx->stage_cap = cap;
    }
    int2* h_c = reinterpret_cast<int2*>(ctx->h_stage);
    uint32_t* h_k = reinterpret_cast<uint32_t*>(ctx->h_stage + n * sizeof(int2));
    const auto t0 = std::chrono::steady_clock::now();
    std::memcpy(h_c, candidates, n * sizeof(int2));
    HIP_TRY(hipMemcpyAsync(ctx->d_cands, h_c, n * sizeof(int2), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(vc1me::launch_score_batch(dcur, dref, width, height, x0, y0, bw, bh, sample_stride, bilinear, rnd, metric,
                                      ctx->d_cands, static_cast<int>(n), ctx->d_costs, ctx->stream));
    HIP_TRY(hipMemcpyAsync(h_k, ctx->d_costs, n * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const auto t1 = std::chrono::steady_clock::now();
    for (size_t i = 0; i < n; ++i) costs_out[i] = h_k[i];
    ctx->counters.dispatches++;
    ctx->counters.candidates += n;
    ctx->counters.submit_wait_nanoseconds +=
        static_cast<uint64_t>(std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count());
    return VC1ME_OK;
}

int validate_score_args(int width, int height, size_t n) {
    if (n > 64u * 65535u) return fail(VC1ME_ELIMIT, "motion batches above 4,194,240 candidates are not supported");
    if (static_cast<uint64_t>(width) * static_cast<uint64_t>(height) > static_cast<uint64_t>(INT32_MAX))
        return fail(VC1ME_ELIMIT, "luma plane exceeds 32-bit addressing");
    return VC1ME_OK;
}

struct Bounds4 { int xmin, xmax, ymin, ymax; };
// the widest legal quarter-pel box (motion_search_bounds_qpel at range 1024)
Bounds4 qpel_limits() { return {-4096, 4095, -1024, 1023}; }

int mvrange_of(int r) { return r <= 31 ? 0 : r <= 63 ? 1 : r <= 127 ? 2 : 3; }

int validate_params(const vc1me_params* p, const vc1me_frame* cur, const vc1me_frame* ref) {
    if (!p || !cur || !ref) return fail(VC1ME_EINVAL, "null argument");
    if (p->width <= 0 || p->height <= 0) return fail(VC1ME_EINVAL, "raster must be positive");
    if ((p->width & 1) || (p->height & 1)) return fail(VC1ME_EINVAL, "coded raster must be even (4:2:0)");
    if (static_cast<int64_t>((p->width + 15) / 16) * ((p->height + 15) / 16) >= vc1me::kMaxRelaxMacroblocks)
        return fail(VC1ME_ELIMIT, "at most 2^24 - 1 macroblocks per picture");
    if (cur->w != p->width || cur->h != p->height || ref->w != p->width || ref->h != p->height)
        return fail(VC1ME_EINVAL, "P-picture reference dimensions do not match encoder configuration");
    if (p->pqindex < 1 || p->pqindex > 31) return fail(VC1ME_EINVAL, "PQINDEX must be 1..31");
    if (p->motion_search_range < 0 || p->motion_search_range > 1024)
        return fail(VC1ME_EINVAL, "motion search range must be 0..1024 pixels");
    const int local = std::min(p->motion_search_range, std::max(0, p->motion_local_search_range));
    if (local > vc1me::kMaxLocalRange)
        return fail(VC1ME_ELIMIT, "local search range above 64 pixels is not supported by the GPU window");
    if (p->me_quality != 1 && p->me_quality != 2)
        return fail(VC1ME_EINVAL, "me_quality must be RATE (1) or SATD (2) on the GPU backend");
    if (p->me_algorithm != 0 && p->me_algorithm != 1) return fail(VC1ME_EINVAL, "me_algorithm must be UMH or diamond");
    if (p->mode < 0 || p->mode > 3) return fail(VC1ME_EINVAL, "invalid progressive MV mode");
    if (p->long_range_search_mode < 0 || p->long_range_search_mode > 2)
        return fail(VC1ME_EINVAL, "invalid long-range search mode");
    if (p->subpixel_refinement_range < 0) return fail(VC1ME_EINVAL, "negative sub-pixel refinement range");
    if (!(p->inter_intra_threshold == p->inter_intra_threshold)) return fail(VC1ME_EINVAL, "NaN intra threshold");
    return VC1ME_OK;
}

void fill_job(JobDesc& d, const vc1me_params* p, const vc1me_frame* cur, const vc1me_frame* ref, const JobSlot& s) {
    std::memset(&d, 0, sizeof d);
    d.cy = cur->y; d.cu = cur->u; d.cv = cur->v;
    d.ry = ref->y; d.ru = ref->u; d.rv = ref->v;
    d.W = p->width; d.H = p->height;
    d.vw = p->display_width > 0 ? p->display_width : p->width;
    d.vh = p->display_height > 0 ? p->display_height : p->height;
    d.mbw = (p->width + 15) / 16;
    d.mbh = (p->height + 15) / 16;
    d.pq = p->pqindex;
    d.range = p->motion_search_range;
    d.local_range = std::min(p->motion_search_range, std::max(0, p->motion_local_search_range));
    d.algo = p->me_algorithm;
    d.quality = p->me_quality;
    d.subpel = p->subpixel_refinement ? 1 : 0;
    d.fast = p->fast_subpixel_refinement ? 1 : 0;
    d.subpel_range = p->subpixel_refinement_range;
    d.lr_mode = p->long_range_search_mode;
    d.mode = p->mode;
    d.allow_intra = (p->allow_intra && !p->debug_disable_p_intra) ? 1 : 0;
    d.mvrange = mvrange_of(p->motion_search_range);
    // lambda_me = max(1, 0.35*pq) (src/encoder_motion.cpp:3458 and :246)
    const double l = 0.35 * static_cast<double>(p->pqindex);
    d.lambda = std::max(1.0, l);
    int e = 0;
    const double f = std::frexp(d.lambda, &e);            // lambda = f * 2^e, f in [0.5,1)
    uint64_t m = static_cast<uint64_t>(std::ldexp(f, 53));  // exact: f has 53 significant bits
    int k = 53 - e;
    while (k > 0 && !(m & 1)) { m >>= 1; --k; }
    d.lam_m = m;
    d.lam_k = k;
    d.ncx = (p->width + 7) / 8;
    d.ncy = (p->height + 7) / 8;
    d.intra_thr = p->inter_intra_threshold;
    // distant_match_decision's long-double rejection threshold, per active
    // pixel count: reject iff (long double)sad > px * max_mae (exact product).
    for (int px = 0; px <= 256; ++px) {
        const long double lim = static_cast<long double>(px) * static_cast<long double>(p->distant_match_max_mae);
        if (!(lim >= 0.0L)) { d.rej_always[px] = 1; d.rej_thr[px] = 0; }
        else if (lim >= 18446744073709551615.0L) { d.rej_always[px] = 0; d.rej_thr[px] = UINT64_MAX; }
        else { d.rej_always[px] = 0; d.rej_thr[px] = static_cast<uint64_t>(floorl(lim)); }
    }
    d.mvs = s.mvs; d.bmv = s.bmv; d.flags = s.flags; d.prop = s.prop;
    d.screen = s.screen; d.ratio = s.ratio; d.dl_mae = s.dl; d.dc_mae = s.dc; d.decision = s.decision;
    d.chosen = s.chosen; d.temporal = s.temporal;
    d.sig = s.sig;
    d.g = s.g;
    d.err = s.err;
    d.ops = s.ops;
    d.ext_sad = s.ext_sad; d.ext_mv = s.ext_mv; d.ext_ops = s.ext_ops; d.ext_done = s.ext_done;
    d.mbops = s.mbops;
    d.mark = s.mark;
    d.pack = nullptr;
    d.nn_out = nullptr; d.nn_logit = 0.f; d.nn_candidates = 1;
    d.seed_mvs = nullptr;
    d.motion_check = 0;
    d.cost20 = nullptr; d.one_mv = nullptr;
    d.lam20 = std::max(20, 7 * p->pqindex);
    d.max_diag = d.mbw + 2 * d.mbh - 3;
    d.qp = ref->qp;
    d.bp = ref->bp;
    d.pp = vc1me::phase_pitch(ref->w);
    d.ph = vc1me::phase_rows(ref->h);
}

// Phase planes of `ref` for vc1me_block_costs: allocated here on first use,
// filled by launch_phases in the call itself.
hipError_t ensure_phase_alloc(const vc1me_frame* ref, int mode) {
    const bool bil = mode == 2;
    uint8_t*& buf = bil ? ref->bp : ref->qp;
    if (buf) return hipSuccess;
    const size_t psz = static_cast<size_t>(vc1me::phase_pitch(ref->w)) * vc1me::phase_rows(ref->h);
    return hipMalloc(&buf, psz * (bil ? 4 : 16));
}

hipError_t launch_phases(const vc1me_frame* ref, int mode, hipStream_t st) {
    const bool bil = mode == 2;
    return vc1me::launch_phase_planes(ref->y, ref->w, ref->h, bil ? ref->bp : ref->qp, vc1me::phase_pitch(ref->w),
                                      vc1me::phase_rows(ref->h), bil ? 1 : 0, st);
}

int mbs_on_diag(int mbw, int mbh, int t) {
    const int lo = std::max(0, (t - (mbw - 1) + 1) / 2);
    const int hi = std::min(mbh - 1, t / 2);
    return std::max(0, hi - lo + 1);
}

// Fill one job's caller arrays from the step's packed copy and form the
// frame scalars from its exact integer aggregates
// (src/encoder_motion.cpp:416-441).  `seed`: the seeded core's seed (its
// distant-match fields are the result's, src/encoder_motion.cpp:97-101), or
// null.
void scatter_result(const JobDesc& d, const vc1me::StepSummary& sum, const uint8_t* packed, vc1me_p_result* out,
                    const vc1me_p_result* seed) {
    const int mbs = d.mbw * d.mbh;
    const size_t n = static_cast<size_t>(mbs);
    const vc1me::PackedOut P = vc1me::pack_layout(const_cast<uint8_t*>(packed), mbs);
    if (out->mvs) std::memcpy(out->mvs, P.mvs, n * 8);
    if (out->block_mvs) std::memcpy(out->block_mvs, P.bmv, n * 32);
    if (out->use_4mv) std::memcpy(out->use_4mv, P.four, n);
    if (out->use_intra) std::memcpy(out->use_intra, P.intra, n);
    if (out->screen_sads) std::memcpy(out->screen_sads, P.screen, n * 8);
    if (out->inter_intra_cost_ratio) std::memcpy(out->inter_intra_cost_ratio, P.ratio, n * 8);
    const bool from_seed = seed && d.range != 0;
    if (out->distant_local_mae) std::memcpy(out->distant_local_mae, from_seed ? seed->distant_local_mae : P.dl, n * 8);
    if (out->distant_candidate_mae)
        std::memcpy(out->distant_candidate_mae, from_seed ? seed->distant_candidate_mae : P.dc, n * 8);
    if (out->distant_match_decision)
        std::memcpy(out->distant_match_decision, from_seed ? seed->distant_match_decision : P.decision, n);
    const uint64_t total_px = sum.total_px;
    if (d.range == 0) {
        out->mean_abs_residual = total_px ? static_cast<double>(sum.total_sad) / static_cast<double>(total_px) : 0.0;
        out->rate_complexity = out->mean_abs_residual;
        out->motion_sad = 0;
        out->moved_macroblocks = out->fractional_chroma_macroblocks = out->four_mv_macroblocks = out->intra_macroblocks = 0;
        out->global_dx = 0; out->global_dy = 0; out->global_sad = 0;
        return;
    }
    const double local_mad = total_px ? static_cast<double>(sum.total_sad) / static_cast<double>(total_px) : 0.0;
    // the seeded core has no global search: its global MAD is the local one
    // (src/encoder_motion.cpp:432-434)
    const double global_mad = seed ? local_mad
                                   : static_cast<double>(sum.global_sad) /
                                         static_cast<double>(static_cast<uint64_t>(d.W) * d.H);
    out->mean_abs_residual = std::min(local_mad, global_mad);
    const double rate_local = total_px ? static_cast<double>(sum.rate_total) / static_cast<double>(total_px) : 0.0;
    out->rate_complexity = std::min(rate_local, global_mad);
    out->motion_sad = sum.total_sad;
    out->moved_macroblocks = sum.moved;
    out->fractional_chroma_macroblocks = sum.frac;
    out->four_mv_macroblocks = sum.four;
    out->intra_macroblocks = sum.intra;
    out->global_dx = seed ? 0 : sum.gdx;
    out->global_dy = seed ? 0 : sum.gdy;
    out->global_sad = seed ? 0 : sum.global_sad;
}

constexpr int kRelaxCounters = 2 * (vc1me::kMaxRelaxPasses + 1);   // list lengths, then tickets
constexpr size_t kOutSummaryAt = 1024;                              // byte offset of the summaries
static_assert(kRelaxCounters * sizeof(int32_t) <= kOutSummaryAt, "relaxation counters overflow their region");

// after_global, if set, is called once the global search is queued (the
// learned form runs its network there, on the global vectors in the slots).
int run_jobs(vc1me_ctx* ctx, const vc1me_job* jobs, int n, const float* const* nn_out = nullptr, float nn_logit = 0.f,
             const std::function<int(hipStream_t)>& after_global = {},
             const vc1me_p_result* const* seeds = nullptr, bool motion_check = false, bool nn_candidates = true,
             const std::string& graph_tag = std::string()) {
    if (n <= 0) return VC1ME_OK;
    if (!jobs) return fail(VC1ME_EINVAL, "null jobs");
    if (n > kMaxJobs) return fail(VC1ME_ELIMIT, "at most " + std::to_string(kMaxJobs) + " jobs per step");
    for (int i = 0; i < n; ++i) {
        const int rc = validate_params(jobs[i].params, jobs[i].cur, jobs[i].ref);
        if (rc) return rc;
    }
    if (injected_failure()) return fail(VC1ME_EINJECTED, "failure injected by VC1ME_FAIL_AT");
    if (ctx->slots.size() < static_cast<size_t>(n)) ctx->slots.resize(n);
    ctx->h_jobs.resize(n);
    hipStream_t st = ctx->stream;
    int max_local = 0, max_diag = 0;
    size_t max_cells = 0, total_mbs = 0;
    int max_ncx = 0, max_ncy = 0;        // signature lattice of the jobs that search long range
    for (int i = 0; i < n; ++i) {
        const vc1me_params* p = jobs[i].params;
        const size_t mbs = static_cast<size_t>((p->width + 15) / 16) * ((p->height + 15) / 16);
        const size_t cells = static_cast<size_t>((p->width + 7) / 8) * ((p->height + 7) / 8);
        hipError_t e = ensure_slot(ctx->slots[i], mbs, cells);
        if (e != hipSuccess) return fail(VC1ME_ENOMEM, std::string("device state allocation: ") + hipGetErrorString(e));
        fill_job(ctx->h_jobs[i], p, jobs[i].cur, jobs[i].ref, ctx->slots[i]);
        ctx->h_jobs[i].motion_check = motion_check ? 1 : 0;
        if (nn_out) {
            ctx->h_jobs[i].nn_out = nn_out[i];
            ctx->h_jobs[i].nn_logit = nn_logit;
            ctx->h_jobs[i].nn_candidates = nn_candidates ? 1 : 0;
        }
        if (ctx->cost_report) {
            ctx->h_jobs[i].cost20 = ctx->slots[i].cost20;
            ctx->h_jobs[i].one_mv = ctx->slots[i].one_mv;
        }
        ctx->slots[i].mbs_last = static_cast<int>(mbs);
        total_mbs += mbs;
        max_local = std::max(max_local, ctx->h_jobs[i].local_range);
        max_diag = std::max(max_diag, ctx->h_jobs[i].max_diag);
        if (seeds) {
            const vc1me_p_result* sd = seeds[i];
            if (!sd || !sd->mvs || !sd->distant_local_mae || !sd->distant_candidate_mae || !sd->distant_match_decision)
                return fail(VC1ME_EINVAL, "seed " + std::to_string(i) + " lacks its vectors or distant-match fields");
            HIP_TRY(hipMemcpyAsync(ctx->slots[i].seed_mvs, sd->mvs, mbs * sizeof(int2), hipMemcpyHostToDevice, st));
            ctx->h_jobs[i].seed_mvs = ctx->slots[i].seed_mvs;
        }
        if (ctx->h_jobs[i].range > ctx->h_jobs[i].local_range && !seeds) {
            max_cells = std::max(max_cells, cells);
            max_ncx = std::max(max_ncx, ctx->h_jobs[i].ncx);
            max_ncy = std::max(max_ncy, ctx->h_jobs[i].ncy);
        }
    }
    if (ctx->d_jobs_cap < static_cast<size_t>(n)) {
        if (ctx->d_jobs) (void)hipFree(ctx->d_jobs);
        ctx->d_jobs = nullptr;
        HIP_TRY(hipMalloc(&ctx->d_jobs, sizeof(JobDesc) * n));
        ctx->d_jobs_cap = n;
    }
    // pass 0 by rows: the grid of one launch covers the largest job's rows of one phase
    const int phases0 = relax_pass0_phases(), phasesg = relax_guess_phases();
    int max_items0 = 0, max_itemsg = 0;
    for (int i = 0; i < n; ++i) {
        max_items0 = std::max(max_items0, ctx->h_jobs[i].mbw * ((ctx->h_jobs[i].mbh + phases0 - 1) / phases0));
        max_itemsg = std::max(max_itemsg, ctx->h_jobs[i].mbw * ((ctx->h_jobs[i].mbh + phasesg - 1) / phasesg));
    }
    if (int rc = ensure_events(ctx, 4 + 2 * static_cast<size_t>(max_diag + 1 + vc1me::kMaxRelaxPasses + phases0 +
                                                                phasesg)))
        return rc;
    // interpolated references: one owning job per distinct (frame, filter),
    // each with a buffer of the context's step pool (planes live for the step)
    int max_pw = 0, max_ph = 0;
    {
        std::vector<std::pair<const vc1me_frame*, bool>> done;
        for (int i = 0; i < n; ++i) {
            JobDesc& d = ctx->h_jobs[i];
            const bool bil = jobs[i].params->mode == 2;
            const std::pair<const vc1me_frame*, bool> key(jobs[i].ref, bil);
            const auto it = std::find(done.begin(), done.end(), key);
            size_t k = static_cast<size_t>(it - done.begin());
            d.phase_owner = it == done.end() ? 1 : 0;
            if (d.phase_owner) {
                done.push_back(key);
                const size_t need = static_cast<size_t>(d.pp) * d.ph * (bil ? 4 : 16);
                if (ctx->phase_pool.size() <= k) ctx->phase_pool.resize(k + 1, {nullptr, 0});
                auto& slot = ctx->phase_pool[k];
                if (slot.second < need) {
                    if (slot.first) (void)hipFree(slot.first);
                    slot = {nullptr, 0};
                    const hipError_t e = hipMalloc(&slot.first, need);
                    if (e != hipSuccess) return fail(VC1ME_ENOMEM, std::string("phase plane allocation: ") + hipGetErrorString(e));
                    slot.second = need;
                }
                max_pw = std::max(max_pw, d.W);
                max_ph = std::max(max_ph, d.H);
            }
            uint8_t* buf = ctx->phase_pool[k].first;
            if (bil) d.bp = buf; else d.qp = buf;
        }
    }
    // the step's result buffer: relaxation counters, summaries, then the
    // packed outputs of every job whose results the caller wants
    const size_t sum_at = kOutSummaryAt;
    const size_t pack_at = vc1me::pack_align(sum_at + sizeof(vc1me::StepSummary) * n);
    std::vector<size_t> pack_off(n, 0);
    size_t out_bytes = pack_at;
    bool want_any = false;
    for (int i = 0; i < n; ++i) {
        if (!jobs[i].out) continue;
        want_any = true;
        pack_off[i] = out_bytes;
        out_bytes += vc1me::pack_layout(nullptr, ctx->h_jobs[i].mbw * ctx->h_jobs[i].mbh).bytes;
    }
    if (ctx->out_cap < out_bytes) {
        const size_t cap = std::max(out_bytes, ctx->out_cap * 2);
        if (ctx->d_out) (void)hipFree(ctx->d_out);
        if (ctx->h_out) (void)hipHostFree(ctx->h_out);
        ctx->d_out = nullptr;
        ctx->h_out = nullptr;
        ctx->out_cap = 0;
        const hipError_t e = hipMalloc(&ctx->d_out, cap);
        if (e != hipSuccess) return fail(VC1ME_ENOMEM, std::string("result buffer: ") + hipGetErrorString(e));
        HIP_TRY(hipHostMalloc(&ctx->h_out, cap, hipHostMallocDefault));
        ctx->out_cap = cap;
    }
    for (int i = 0; i < n; ++i) ctx->h_jobs[i].pack = jobs[i].out ? ctx->d_out + pack_off[i] : nullptr;
    int32_t* d_cnt = reinterpret_cast<int32_t*>(ctx->d_out);
    auto* d_sum = reinterpret_cast<vc1me::StepSummary*>(ctx->d_out + sum_at);
    const bool relax = ctx->schedule == VC1ME_SCHEDULE_RELAX;
    int max_mbs = 0, max_mbs_all = 0;
    for (int i = 0; i < n; ++i) {
        const int m = ctx->h_jobs[i].mbw * ctx->h_jobs[i].mbh;
        max_mbs_all = std::max(max_mbs_all, m);
        if (ctx->h_jobs[i].range > ctx->h_jobs[i].local_range) max_mbs = std::max(max_mbs, m);
    }
    // the whole-frame kernels' (macroblocks x jobs) grid is indexed in int
    if (static_cast<int64_t>(max_mbs_all) * n >= (int64_t(1) << 30))
        return fail(VC1ME_ELIMIT, "a step holds at most 2^30 macroblocks (largest picture's count x jobs)");
    vc1me::RelaxQueue Q{};
    const int lds = vc1me::mb_kernel_lds_bytes(max_local);
    if (relax) {
        if (ctx->relax_wl_cap < total_mbs) {
            if (ctx->d_relax_wl) (void)hipFree(ctx->d_relax_wl);
            ctx->d_relax_wl = nullptr;
            ctx->relax_wl_cap = 0;
            const hipError_t e = hipMalloc(&ctx->d_relax_wl, 2 * total_mbs * sizeof(int32_t));
            if (e != hipSuccess) return fail(VC1ME_ENOMEM, std::string("relaxation lists: ") + hipGetErrorString(e));
            ctx->relax_wl_cap = total_mbs;
        }
        if (ctx->relax_grid_lds != lds) {
            ctx->relax_grid = vc1me::relax_queue_grid(lds, ctx->cu_count);
            ctx->relax_grid_lds = lds;
        }
        Q.cnt = d_cnt;
        Q.tk = d_cnt + (vc1me::kMaxRelaxPasses + 1);
        Q.wl[0] = ctx->d_relax_wl;
        Q.wl[1] = ctx->d_relax_wl + total_mbs;
        Q.cap = static_cast<int32_t>(total_mbs);
    }
    HIP_TRY(hipMemcpyAsync(ctx->d_jobs, ctx->h_jobs.data(), sizeof(JobDesc) * n, hipMemcpyHostToDevice, st));
    ctx->last_jobs = n;
    auto per_diag = [&](int t) {
        int per = 0;
        for (int i = 0; i < n; ++i)
            if (t <= ctx->h_jobs[i].max_diag) per = std::max(per, mbs_on_diag(ctx->h_jobs[i].mbw, ctx->h_jobs[i].mbh, t));
        return per;
    };
    // the launch shape: rasters, ranges and modes of the jobs (graphs and
    // the relaxation's pass count are kept per shape)
    std::string shape;
    {
        auto put = [&shape](long long v) { shape += std::to_string(v); shape += ','; };
        put(n); put(lds); put(static_cast<long long>(max_cells)); put(max_mbs); put(max_diag); put(max_pw); put(max_ph);
        put(relax ? 1 : 0);
        for (int i = 0; i < n; ++i) {
            const JobDesc& d = ctx->h_jobs[i];
            put(d.W); put(d.H); put(d.range); put(d.local_range); put(d.mbw); put(d.mbh); put(d.max_diag);
            put(d.mode); put(d.phase_owner); put(d.seed_mvs ? 1 : 0); put(d.motion_check); put(d.pack ? 1 : 0);
            put(d.nn_out ? 1 : 0); put(d.cost20 ? 1 : 0);
        }
    }
    // passes launched without a host read: the most this shape has needed
    // so far plus two (each pass after the state settles costs a launch of
    // empty workgroups), or a fixed start for a new shape
    const int limit = relax_pass_limit();
    int passes = std::min(limit, relax_first_passes());
    int seen_need = 0;
    for (const auto& h : ctx->pass_hints)
        if (h.first == shape) seen_need = h.second;
    if (seen_need) passes = std::min(limit, seen_need + 2);
    // launches with events around them when profiling
    int mb_launches = 0;
    std::vector<std::pair<int, int>> mb_events;
    int prepass_event = -1;   // mb_events entry of the guess pass (or the sweep's extended pre-pass)
    // one event pair around `launch`, which makes `count` kernel launches
    auto timed = [&](bool events, const std::function<hipError_t()>& launch, int count = 1) -> int {
        const int ev = 4 + 2 * mb_launches;
        const bool rec = events && ctx->profiling && static_cast<size_t>(ev + 1) < ctx->events.size();
        if (rec) HIP_TRY(hipEventRecord(ctx->events[ev], st));
        HIP_TRY(launch());
        if (rec) {
            HIP_TRY(hipEventRecord(ctx->events[ev + 1], st));
            mb_events.emplace_back(ev, ev + 1);
        }
        mb_launches += count;
        return VC1ME_OK;
    };
    // one launch per anti-diagonal: every macroblock after its neighbours
    auto diagonal_all = [&](bool events) -> int {
        for (int t = 0; t <= max_diag; ++t) {
            const int per = per_diag(t);
            if (!per) continue;
            if (int rc = timed(events, [&] { return vc1me::launch_mb_diagonal(ctx->d_jobs, n, t, per, lds, st); }))
                return rc;
        }
        return VC1ME_OK;
    };
    // relaxation passes [p0, p1): work-queue passes over the device lists,
    // no host read in between
    auto relax_passes = [&](bool events, int p0, int p1) -> int {
        for (int p = p0; p < p1; ++p) {
            if (p == 0) {
                for (int r = 0; r < phases0; ++r)
                    if (int rc = timed(events, [&] {
                            return vc1me::launch_mb_relax0(ctx->d_jobs, n, max_items0, r, phases0, lds, Q, st);
                        }))
                        return rc;
                continue;
            }
            if (int rc = timed(events, [&] { return vc1me::launch_mb_relaxq(ctx->d_jobs, p, ctx->relax_grid, lds, Q, st); }))
                return rc;
        }
        return VC1ME_OK;
    };
    // end of the step: the cost report of the final field, then the summaries
    auto finish = [&]() -> int {
        if (ctx->cost_report) HIP_TRY(vc1me::launch_cost_final(ctx->d_jobs, n, max_mbs_all, lds, st));
        HIP_TRY(vc1me::launch_step_end(ctx->d_jobs, n, d_sum, st));
        return VC1ME_OK;
    };
    // the step's launch sequence; every launch takes the job table and scalars only
    auto launch_all = [&](bool events) -> int {
        HIP_TRY(vc1me::launch_step_begin(ctx->d_jobs, n, max_mbs_all, d_cnt, kRelaxCounters, st));
        // the phase planes and the signature index depend on the frames
        // alone: they run on the side stream beside the global search (and
        // the network), and the macroblock passes wait for both (in a
        // captured step, two branches of the graph)
        HIP_TRY(hipEventRecord(ctx->fork, st));
        HIP_TRY(hipStreamWaitEvent(ctx->side, ctx->fork, 0));
        HIP_TRY(vc1me::launch_phase_planes_jobs(ctx->d_jobs, n, max_pw, max_ph, ctx->side));
        HIP_TRY(vc1me::launch_signatures(ctx->d_jobs, n, max_ncx, max_ncy, ctx->side));
        HIP_TRY(hipEventRecord(ctx->join, ctx->side));
        HIP_TRY(vc1me::launch_global_search(ctx->d_jobs, n, ctx->h_jobs.data(), st,
                                            events ? ctx->events[2] : nullptr, events ? ctx->events[3] : nullptr));
        if (after_global)
            if (int rc = after_global(st)) return rc;
        HIP_TRY(hipStreamWaitEvent(st, ctx->join, 0));
        if (relax) {
            HIP_TRY(vc1me::launch_relax_init(ctx->d_jobs, n, max_mbs_all, st));
            if (relax_guess()) {
                const size_t at = mb_events.size();
                // the guess's launches by rows are timed as one interval
                if (int rc = timed(
                        events,
                        [&] {
                            for (int r = 0; r < phasesg; ++r)
                                if (hipError_t e = vc1me::launch_mb_guess(ctx->d_jobs, n, max_itemsg, r, phasesg, lds, st))
                                    return e;
                            return hipSuccess;
                        },
                        phasesg))
                    return rc;
                if (mb_events.size() > at) prepass_event = static_cast<int>(at);
            }
            if (int rc = relax_passes(events, 0, passes)) return rc;
        } else {
            const size_t at = mb_events.size();
            if (int rc = timed(events, [&] { return vc1me::launch_extended(ctx->d_jobs, n, max_mbs, lds, st); }))
                return rc;
            if (mb_events.size() > at) prepass_event = static_cast<int>(at);
            if (int rc = diagonal_all(events)) return rc;
        }
        return finish();
    };
    ctx->timing.graph_replay = 0;
    // a step with work after the global search is capturable when its caller
    // tags the buffers that work uses (the learned step's network)
    const bool graph = ctx->use_graphs && !ctx->profiling && (!after_global || !graph_tag.empty());
    if (graph) {
        // replay the step as one HIP graph, captured once per launch shape
        const std::string key = shape + "p" + std::to_string(relax ? passes : 0) + "r" + std::to_string(phases0) + "g" + std::to_string(phasesg) + "@" +
                                std::to_string(reinterpret_cast<long long>(ctx->d_jobs)) + "," +
                                std::to_string(reinterpret_cast<long long>(ctx->d_out)) + "," +
                                std::to_string(reinterpret_cast<long long>(ctx->d_relax_wl)) + graph_tag;
        hipGraphExec_t exec = nullptr;
        for (auto& g : ctx->graphs)
            if (g.first == key) exec = g.second;
        if (!exec) {
            HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const int rc = launch_all(false);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(st, &g);
            if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
            HIP_TRY(ec);
            const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIP_TRY(ei);
            if (ctx->graphs.size() >= 16) {
                (void)hipGraphExecDestroy(ctx->graphs.front().second);
                ctx->graphs.erase(ctx->graphs.begin());
            }
            ctx->graphs.emplace_back(key, exec);
            ctx->timing.graph_replay = 2;
        } else {
            ctx->timing.graph_replay = 1;
        }
        HIP_TRY(hipEventRecord(ctx->events[0], st));
        HIP_TRY(hipGraphLaunch(exec, st));
    } else {
        HIP_TRY(hipEventRecord(ctx->events[0], st));
        if (int rc = launch_all(true)) return rc;
    }
    // one copy of the counters and summaries (and the packed outputs)
    const size_t head = pack_at;
    HIP_TRY(hipMemcpyAsync(ctx->h_out, ctx->d_out, want_any ? out_bytes : head, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipEventRecord(ctx->events[1], st));
    HIP_TRY(hipEventSynchronize(ctx->events[1]));
    const int32_t* cnt = reinterpret_cast<const int32_t*>(ctx->h_out);
    bool relax_fell_back = false;
    int used = 0;   // relaxation passes until the state settled
    if (relax) {
        // pass p's list is cnt[p]; the state settled once a list is empty
        auto settled_at = [&](int upto) {
            for (int p = 1; p <= upto; ++p)
                if (cnt[p] == 0) return p;
            return 0;
        };
        int launched = passes;
        used = settled_at(launched);
        while (!used && launched < limit) {
            // not settled: more passes, eight at a time, then look again
            const int more = std::min(limit, launched + 8);
            if (int rc = relax_passes(true, launched, more)) return rc;
            launched = more;
            HIP_TRY(hipMemcpyAsync(ctx->h_out, ctx->d_out, head, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            used = settled_at(launched);
        }
        if (!used) {
            // not settled within the limit: the diagonal sweep is exact from any state
            relax_fell_back = true;
            if (int rc = diagonal_all(true)) return rc;
        }
        if (!used || launched > passes) {
            if (int rc = finish()) return rc;
            HIP_TRY(hipMemcpyAsync(ctx->h_out, ctx->d_out, want_any ? out_bytes : head, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipEventRecord(ctx->events[1], st));
            HIP_TRY(hipEventSynchronize(ctx->events[1]));
        }
        // the most passes this shape has needed (a rise is a new captured
        // graph at the next step; a fall changes nothing)
        const int need = used ? used : limit;
        bool found = false;
        for (auto& h : ctx->pass_hints)
            if (h.first == shape) { found = true; h.second = std::max(h.second, need); }
        if (!found) {
            if (ctx->pass_hints.size() >= 64) ctx->pass_hints.erase(ctx->pass_hints.begin());
            ctx->pass_hints.emplace_back(shape, need);
        }
    }
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->events[0], ctx->events[1]));
    ctx->timing.step_ms = ms;
    float gms = 0.f;
    if (!ctx->timing.graph_replay) HIP_TRY(hipEventElapsedTime(&gms, ctx->events[2], ctx->events[3]));
    ctx->timing.global_kernel_ms_total = gms;
    ctx->timing.global_kernel_launches = 9;
    ctx->pass_analyses.clear();
    ctx->pass_ms.clear();
    uint64_t analyses = 0;
    if (relax) {
        const int np = used ? used : limit;
        for (int p = 0; p < np; ++p) {
            const int64_t a = p == 0 ? static_cast<int64_t>(total_mbs) : cnt[p];
            ctx->pass_analyses.push_back(a);
            analyses += static_cast<uint64_t>(a);
        }
    }
    double mbms = 0.0;
    double prepass_ms = 0.0;
    for (size_t k = 0; k < mb_events.size(); ++k) {
        float x = 0.f;
        HIP_TRY(hipEventElapsedTime(&x, ctx->events[mb_events[k].first], ctx->events[mb_events[k].second]));
        mbms += x;
        if (static_cast<int>(k) == prepass_event) { prepass_ms = x; continue; }
        // the relaxation's pass 0 is its first phases0 launches
        const size_t rel = prepass_event >= 0 && static_cast<int>(k) > prepass_event ? k - 1 : k;
        const size_t pass = !relax ? rel : rel < static_cast<size_t>(phases0) ? 0 : rel - phases0 + 1;
        if (relax && pass < ctx->pass_analyses.size()) {
            if (rel > 0 && rel < static_cast<size_t>(phases0) && !ctx->pass_ms.empty()) ctx->pass_ms.back() += x;
            else ctx->pass_ms.push_back(x);
        }
    }
    if (ctx->timing.graph_replay) {
        mb_launches = 0;
        if (relax) mb_launches = passes + phases0 - 1 + (relax_guess() ? phasesg : 0);
        else
            for (int t = 0; t <= max_diag; ++t) mb_launches += per_diag(t) ? 1 : 0;
        if (!relax) ++mb_launches;   // the extended pre-pass
    }
    ctx->timing.mb_kernel_ms_total = mbms;
    ctx->timing.prepass_ms = prepass_ms;
    ctx->timing.mb_kernel_launches = static_cast<uint64_t>(mb_launches);
    ctx->timing.schedule = relax && !relax_fell_back ? VC1ME_SCHEDULE_RELAX : VC1ME_SCHEDULE_DIAGONAL;
    ctx->timing.relax_passes = relax ? static_cast<uint64_t>(used ? used : limit) : 0;
    ctx->timing.relax_analyses = relax ? analyses : 0;
    // kernel-side error flags and work counters, from the one copy
    const auto* sums = reinterpret_cast<const vc1me::StepSummary*>(ctx->h_out + sum_at);
    ctx->work = vc1me_work{};
    for (int i = 0; i < n; ++i) {
        if (sums[i].err) return fail(VC1ME_EKERNEL, "kernel reported internal error code " + std::to_string(sums[i].err));
        ctx->work.sad_px += sums[i].ops[0];
        ctx->work.interp_taps += sums[i].ops[1];
        ctx->work.satd_ops += sums[i].ops[2];
        ctx->work.sig_ops += sums[i].ops[3];
        ctx->work.frame_px += sums[i].ops[4];
        ctx->work.seeded_macroblocks += sums[i].ops[5];
    }
    for (int i = 0; i < n; ++i) {
        ctx->counters.frames_analyzed++;
        ctx->counters.macroblocks_analyzed += static_cast<uint64_t>(ctx->h_jobs[i].mbw) * ctx->h_jobs[i].mbh;
    }
    if (want_any) {
        // the caller's arrays from the pinned copy; large steps spread the
        // copying over a few host threads
        auto one = [&](int i) {
            if (jobs[i].out)
                scatter_result(ctx->h_jobs[i], sums[i], ctx->h_out + pack_off[i], jobs[i].out, seeds ? seeds[i] : nullptr);
        };
        const int workers = out_bytes > (size_t(8) << 20) ? std::min(8, n) : 1;
        if (workers <= 1) {
            for (int i = 0; i < n; ++i) one(i);
        } else {
            std::vector<std::thread> pool;
            for (int w = 0; w < workers; ++w)
                pool.emplace_back([&, w] { for (int i = w; i < n; i += workers) one(i); });
            for (auto& t : pool) t.join();
        }
    }
    return VC1ME_OK;
}


// ------------------------------------------------------------------ proposal network

struct NnBuffers {
    // capacities: images (gxy), samples over all images (x0, act, dz, in32)
    // and macroblock cells over all images (pooled, out, dout, tgt, valid);
    // kept apart because a raster with fewer samples can have more cells
    size_t nb = 0, px = 0, mbs = 0;
    uint16_t* x0 = nullptr;            // [nb][h4][w4][2]
    uint16_t* act[pnet::kLayers] = {};
    uint16_t* dz[2] = {};
    float* pooled = nullptr;
    float* out = nullptr;
    float* dout = nullptr;
    float* tgt = nullptr;
    uint8_t* valid = nullptr;
    float* in32 = nullptr;
    float* loss = nullptr;
    int32_t* gxy = nullptr;            // [nb][2] global vectors handed in by the host
};

}  // namespace

struct vc1me_nn {
    vc1me_ctx* ctx = nullptr;
    float* p = nullptr;      // fp32 master parameters
    float* g = nullptr;      // gradients
    float* m = nullptr;      // Adam moments
    float* v = nullptr;
    int step = 0;
    uint16_t* wt[pnet::kLayers] = {};    // bf16 [co][9][ci]
    uint16_t* wtt[pnet::kLayers] = {};   // bf16 [9][ci][co] (layers with cin 32)
    bool packed = false;
    NnBuffers b;
    // proposals of a learned step: 5 floats per macroblock, each raster
    // group's jobs contiguous in one buffer (props[i] points at job i's)
    float* props_all = nullptr;
    size_t props_cap = 0;
    std::vector<float*> props;
};

namespace {

int nn_dims(int W, int H, int& h4, int& w4, int& mbh, int& mbw) {
    h4 = (H + 3) / 4; w4 = (W + 3) / 4;
    mbh = (h4 + 3) / 4; mbw = (w4 + 3) / 4;
    return h4 * w4;
}

hipError_t nn_reserve(vc1me_nn* nn, int nb_, int h4, int w4) {
    NnBuffers& b = nn->b;
    const size_t nb = static_cast<size_t>(nb_);
    size_t px = nb * h4 * w4;
    size_t mbs = nb * ((h4 + 3) / 4) * ((w4 + 3) / 4);
    if (b.x0 && nb <= b.nb && px <= b.px && mbs <= b.mbs) return hipSuccess;
    auto rel = [](auto*& q) { if (q) (void)hipFree(q); q = nullptr; };
    rel(b.x0); for (auto& a : b.act) rel(a); rel(b.dz[0]); rel(b.dz[1]); rel(b.pooled); rel(b.out); rel(b.dout);
    rel(b.tgt); rel(b.valid); rel(b.in32); rel(b.loss); rel(b.gxy);
    // never shrink: the next request may need the old capacity again
    const size_t nbc = std::max(nb, b.nb);
    px = std::max(px, b.px);
    mbs = std::max(mbs, b.mbs);
    b.nb = b.px = b.mbs = 0;
    hipError_t e = hipMalloc(&b.x0, px * 2 * 2);
    for (auto& a : b.act) if (e == hipSuccess) e = hipMalloc(&a, px * 32 * 2);
    if (e == hipSuccess) e = hipMalloc(&b.dz[0], px * 32 * 2);
    if (e == hipSuccess) e = hipMalloc(&b.dz[1], px * 32 * 2);
    if (e == hipSuccess) e = hipMalloc(&b.pooled, mbs * 32 * 4);
    if (e == hipSuccess) e = hipMalloc(&b.out, mbs * 5 * 4);
    if (e == hipSuccess) e = hipMalloc(&b.dout, mbs * 5 * 4);
    if (e == hipSuccess) e = hipMalloc(&b.tgt, mbs * 2 * 4);
    if (e == hipSuccess) e = hipMalloc(&b.valid, mbs);
    if (e == hipSuccess) e = hipMalloc(&b.in32, px * 2 * 4);
    if (e == hipSuccess) e = hipMalloc(&b.loss, 4);
    if (e == hipSuccess) e = hipMalloc(&b.gxy, sizeof(int32_t) * 2 * nbc);
    if (e != hipSuccess) return e;
    b.nb = nbc; b.px = px; b.mbs = mbs;
    return hipSuccess;
}

hipError_t nn_pack(vc1me_nn* nn, hipStream_t st) {
    if (nn->packed) return hipSuccess;
    for (int l = 0; l < pnet::kLayers; ++l) {
        const hipError_t e = pnet::launch_pack_weights(nn->p + pnet::layer_offset(l), nn->wt[l],
                                                            l ? nn->wtt[l] : nullptr, pnet::layer_cin(l), st);
        if (e != hipSuccess) return e;
    }
    nn->packed = true;
    return hipSuccess;
}

// forward of nb images already in b.x0; outputs in out (null: b.out).  The
// weights must be packed already when the launches are being captured.
hipError_t nn_forward(vc1me_nn* nn, int nb, int h4, int w4, hipStream_t st, float* out = nullptr) {
    hipError_t e = nn_pack(nn, st);
    const uint16_t* in = nn->b.x0;
    for (int l = 0; l < pnet::kLayers && e == hipSuccess; ++l) {
        const int cin = pnet::layer_cin(l);
        const float* bias = nn->p + pnet::layer_offset(l) + 32 * 9 * cin;
        e = pnet::launch_conv_fwd(in, nn->wt[l], bias, nn->b.act[l], nb, h4, w4, cin, pnet::layer_dil(l), st);
        in = nn->b.act[l];
    }
    if (e == hipSuccess)
        e = pnet::launch_pool_head(nn->b.act[pnet::kLayers - 1], nn->p + pnet::head_offset(), nn->b.pooled,
                                       out ? out : nn->b.out, nb, h4, w4, (h4 + 3) / 4, (w4 + 3) / 4, st);
    return e;
}

// backward from b.dout into nn->g (zeroed here)
hipError_t nn_backward(vc1me_nn* nn, int nb, int h4, int w4, hipStream_t st) {
    using namespace pnet;
    hipError_t e = hipMemsetAsync(nn->g, 0, sizeof(float) * kParams, st);
    if (e == hipSuccess)
        e = launch_pool_head_bwd(nn->b.dout, nn->b.pooled, nn->p + head_offset(), nn->b.act[kLayers - 1],
                                 nn->g + head_offset(), nn->b.dz[0], nb, h4, w4, (h4 + 3) / 4, (w4 + 3) / 4, st);
    int cur = 0;
    for (int l = kLayers - 1; l >= 0 && e == hipSuccess; --l) {
        const int cin = layer_cin(l);
        const uint16_t* xin = l ? nn->b.act[l - 1] : nn->b.x0;
        float* dw = nn->g + layer_offset(l);
        e = launch_conv_bwd_weight(nn->b.dz[cur], xin, dw, dw + 32 * 9 * cin, nb, h4, w4, cin, layer_dil(l), st);
        if (e == hipSuccess && l > 0) {
            e = launch_conv_bwd_data(nn->b.dz[cur], nn->wtt[l], nn->b.act[l - 1], nn->b.dz[1 - cur], nb, h4, w4,
                                     layer_dil(l), st);
            cur = 1 - cur;
        }
    }
    return e;
}

// downsampled inputs of n frame pairs of one raster into b.x0; gdev[i] is
// pair i's global vector in device memory (null: none), or host_g (n x 2,
// null: none) is uploaded to b.gxy first
int nn_inputs(vc1me_nn* nn, const vc1me_frame* const* curs, const vc1me_frame* const* refs, int n, int& h4, int& w4,
              int& mbh, int& mbw, const int32_t* host_g, const int32_t* const* gdev = nullptr) {
    if (!curs || !refs || n <= 0) return fail(VC1ME_EINVAL, "no frames");
    const int W = curs[0]->w, H = curs[0]->h;
    for (int i = 0; i < n; ++i)
        if (!curs[i] || !refs[i] || curs[i]->w != W || refs[i]->w != W || curs[i]->h != H || refs[i]->h != H)
            return fail(VC1ME_EINVAL, "frames of one batch must share one raster");
    nn_dims(W, H, h4, w4, mbh, mbw);
    hipStream_t st = nn->ctx->stream;
    HIP_TRY(nn_reserve(nn, n, h4, w4));
    if (host_g) HIP_TRY(hipMemcpyAsync(nn->b.gxy, host_g, sizeof(int32_t) * 2 * n, hipMemcpyHostToDevice, st));
    for (int i = 0; i < n; ++i) {
        const int32_t* g = gdev ? gdev[i] : host_g ? nn->b.gxy + 2 * i : nullptr;
        HIP_TRY(pnet::launch_downsample(curs[i]->y, refs[i]->y, W, H, h4, w4, g,
                                            nn->b.x0 + static_cast<size_t>(i) * h4 * w4 * 2, st));
    }
    return VC1ME_OK;
}

}  // namespace

extern "C" {

void vc1me_params_default(vc1me_params* p, int32_t width, int32_t height) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->width = width;
    p->height = height;
    p->pqindex = 9;
    p->motion_search_range = 1024;
    p->motion_local_search_range = 32;
    p->me_algorithm = 0;
    p->me_quality = 2;
    p->subpixel_refinement = 1;
    p->fast_subpixel_refinement = 0;
    p->subpixel_refinement_range = 1024;
    p->long_range_search_mode = 0;
    p->mode = VC1ME_MODE_ONE_MV_QPEL;
    p->allow_intra = 1;
    p->debug_disable_p_intra = 0;
    p->distant_match_max_mae = 255.0;
    p->inter_intra_threshold = 0.80;
}

int vc1me_abi_version(void) { return VC1ME_ABI_VERSION; }

const char* vc1me_last_error(void) { return g_last_error.c_str(); }

int vc1me_create(int device, vc1me_ctx** out) {
    if (!out) return fail(VC1ME_EINVAL, "null output pointer");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) return fail(VC1ME_EDEVICE, "no HIP device available");
    if (device < -1 || device >= count) return fail(VC1ME_EINVAL, "device ordinal out of range");
    auto* ctx = new vc1me_ctx();
    ctx->device = device < 0 ? 0 : device;
    if (int rc = set_device(ctx)) { delete ctx; return rc; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) {
        ctx->name = std::string(prop.name) + " (" + prop.gcnArchName + ")";
        if (prop.multiProcessorCount > 0) ctx->cu_count = prop.multiProcessorCount;
    }
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming);
    if (e != hipSuccess) {
        if (ctx->join) (void)hipEventDestroy(ctx->join);
        if (ctx->fork) (void)hipEventDestroy(ctx->fork);
        if (ctx->side) (void)hipStreamDestroy(ctx->side);
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return fail(VC1ME_EDEVICE, std::string("stream set-up: ") + hipGetErrorString(e));
    }
    // the allocation-free scorer's scratch array on this device (looked up
    // once per device, with the device current)
    ctx->scratch = scratch_for_device(ctx->device, &ctx->scratch_error);
    if (hipHostMalloc(&ctx->h_noalloc, kNoAllocScratchBytes, hipHostMallocDefault) != hipSuccess)
        ctx->h_noalloc = nullptr;
    *out = ctx;
    return VC1ME_OK;
}

void vc1me_destroy(vc1me_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& s : ctx->slots)
        if (s.mem) (void)hipFree(s.mem);
    for (auto& p : ctx->planes)
        if (p.dev) (void)hipFree(p.dev);
    if (ctx->d_jobs) (void)hipFree(ctx->d_jobs);
    if (ctx->d_cands) (void)hipFree(ctx->d_cands);
    if (ctx->d_costs) (void)hipFree(ctx->d_costs);
    if (ctx->d_props) (void)hipFree(ctx->d_props);
    if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
    if (ctx->h_noalloc) (void)hipHostFree(ctx->h_noalloc);
    if (ctx->d_relax_wl) (void)hipFree(ctx->d_relax_wl);
    if (ctx->d_out) (void)hipFree(ctx->d_out);
    if (ctx->h_out) (void)hipHostFree(ctx->h_out);
    if (ctx->d_ic) (void)hipFree(ctx->d_ic);
    if (ctx->d_b) (void)hipFree(ctx->d_b);
    for (auto& g : ctx->graphs) (void)hipGraphExecDestroy(g.second);
    for (auto& b : ctx->phase_pool) if (b.first) (void)hipFree(b.first);
    for (auto e : ctx->events) (void)hipEventDestroy(e);
    (void)hipEventDestroy(ctx->join);
    (void)hipEventDestroy(ctx->fork);
    (void)hipStreamSynchronize(ctx->side);
    (void)hipStreamDestroy(ctx->side);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* vc1me_device_name(const vc1me_ctx* ctx) { return ctx ? ctx->name.c_str() : ""; }

int vc1me_score_batch(vc1me_ctx* ctx, const uint8_t* cur, const uint8_t* ref, int width, int height, int x0, int y0,
                      int bw, int bh, int sample_stride, int bilinear, int rnd, int metric,
                      const vc1me_candidate* candidates, size_t n, uint64_t* costs_out) {
    if (!ctx) return fail(VC1ME_EINVAL, "null context");
    if (!cur || !ref || width <= 0 || height <= 0 || bw <= 0 || bh <= 0 || !candidates || !n) return VC1ME_OK;
    if (!costs_out) return fail(VC1ME_EINVAL, "null costs");
    if (int rc = validate_score_args(width, height, n)) return rc;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    if (injected_failure()) return fail(VC1ME_EINJECTED, "failure injected by VC1ME_FAIL_AT");
    uint8_t *dc = nullptr, *dr = nullptr;
    if (int rc = plane_for(ctx, cur, width, height, &dc)) return rc;
    if (int rc = plane_for(ctx, ref, width, height, &dr)) return rc;
    return score_on_device(ctx, dc, dr, width, height, x0, y0, bw, bh, sample_stride, bilinear, rnd, metric,
                           candidates, n, costs_out);
}

int vc1me_score_batch_gen(vc1me_ctx* ctx, const uint8_t* cur, uint64_t cur_generation, const uint8_t* ref,
                          uint64_t ref_generation, int width, int height, int x0, int y0, int bw, int bh,
                          int sample_stride, int bilinear, int rnd, int metric, const vc1me_candidate* candidates,
                          size_t n, uint64_t* costs_out) {
    if (!ctx) return fail(VC1ME_EINVAL, "null context");
    if (!cur || !ref || width <= 0 || height <= 0 || bw <= 0 || bh <= 0 || !candidates || !n) return VC1ME_OK;
    if (!costs_out) return fail(VC1ME_EINVAL, "null costs");
    if (int rc = validate_score_args(width, height, n)) return rc;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    if (injected_failure()) return fail(VC1ME_EINJECTED, "failure injected by VC1ME_FAIL_AT");
    uint8_t *dc = nullptr, *dr = nullptr;
    if (int rc = plane_for(ctx, cur, width, height, &dc, true, cur_generation)) return rc;
    if (int rc = plane_for(ctx, ref, width, height, &dr, true, ref_generation)) return rc;
    return score_on_device(ctx, dc, dr, width, height, x0, y0, bw, bh, sample_stride, bilinear, rnd, metric,
                           candidates, n, costs_out);
}

int vc1me_frame_upload(vc1me_ctx* ctx, const vc1me_planes* planes, vc1me_frame** out) {
    if (!ctx || !planes || !out || !planes->y) return fail(VC1ME_EINVAL, "null argument");
    if (planes->width <= 0 || planes->height <= 0 || (planes->width & 1) || (planes->height & 1))
        return fail(VC1ME_EINVAL, "frame raster must be positive and even");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    const size_t ys = static_cast<size_t>(planes->width) * planes->height;
    const size_t cs = static_cast<size_t>(planes->width / 2) * (planes->height / 2);
    auto* f = new vc1me_frame();
    f->w = planes->width;
    f->h = planes->height;
    uint8_t* mem = nullptr;
    hipError_t e = hipMalloc(&mem, ys + 2 * cs + 64);
    if (e != hipSuccess) { delete f; return fail(VC1ME_ENOMEM, std::string("frame allocation: ") + hipGetErrorString(e)); }
    f->y = mem;
    f->u = mem + ys;
    f->v = mem + ys + cs;
    std::vector<uint8_t> grey;
    const uint8_t* u = planes->u;
    const uint8_t* v = planes->v;
    if (!u || !v) { grey.assign(cs, 128); if (!u) u = grey.data(); if (!v) v = grey.data(); }
    e = hipMemcpy(f->y, planes->y, ys, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(f->u, u, cs, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(f->v, v, cs, hipMemcpyHostToDevice);
    if (e != hipSuccess) { (void)hipFree(mem); delete f; return fail(VC1ME_EDEVICE, std::string("frame upload: ") + hipGetErrorString(e)); }
    ctx->counters.frame_uploads++;
    ctx->counters.frame_upload_bytes += ys + 2 * cs;
    *out = f;
    return VC1ME_OK;
}

void vc1me_frame_release(vc1me_ctx* ctx, vc1me_frame* frame) {
    if (!frame) return;
    if (ctx) {
        std::lock_guard<std::mutex> lock(ctx->mu);
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
    }
    (void)hipFree(frame->y);
    if (frame->qp) (void)hipFree(frame->qp);
    if (frame->bp) (void)hipFree(frame->bp);
    delete frame;
}

int vc1me_score_batch_frames(vc1me_ctx* ctx, const vc1me_frame* cur, const vc1me_frame* ref, int x0, int y0, int bw,
                             int bh, int sample_stride, int bilinear, int rnd, int metric,
                             const vc1me_candidate* candidates, size_t n, uint64_t* costs_out) {
    if (!ctx) return fail(VC1ME_EINVAL, "null context");
    if (!cur || !ref || bw <= 0 || bh <= 0 || !candidates || !n) return VC1ME_OK;
    if (cur->w != ref->w || cur->h != ref->h) return fail(VC1ME_EINVAL, "frame size mismatch");
    if (!costs_out) return fail(VC1ME_EINVAL, "null costs");
    if (int rc = validate_score_args(cur->w, cur->h, n)) return rc;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    if (injected_failure()) return fail(VC1ME_EINJECTED, "failure injected by VC1ME_FAIL_AT");
    return score_on_device(ctx, cur->y, ref->y, cur->w, cur->h, x0, y0, bw, bh, sample_stride, bilinear, rnd, metric,
                           candidates, n, costs_out);
}

int vc1me_score_batch_frames_noalloc(vc1me_ctx* ctx, const vc1me_frame* cur, const vc1me_frame* ref, int x0, int y0,
                                     int bw, int bh, int sample_stride, int bilinear, int rnd, int metric,
                                     const vc1me_candidate* candidates, size_t n, uint64_t* costs_out) {
    if (!ctx) return fail(VC1ME_EINVAL, "null context");
    if (!cur || !ref || bw <= 0 || bh <= 0 || !candidates || !n) return VC1ME_OK;
    if (cur->w != ref->w || cur->h != ref->h) return fail(VC1ME_EINVAL, "frame size mismatch");
    if (!costs_out) return fail(VC1ME_EINVAL, "null costs");
    if (n > static_cast<size_t>(VC1ME_NOALLOC_MAX_CANDIDATES))
        return fail(VC1ME_ELIMIT, "the allocation-free scorer takes at most " +
                                      std::to_string(VC1ME_NOALLOC_MAX_CANDIDATES) + " candidates per call");
    if (int rc = validate_score_args(cur->w, cur->h, n)) return rc;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    if (injected_failure()) return fail(VC1ME_EINJECTED, "failure injected by VC1ME_FAIL_AT");
    if (!ctx->scratch)
        return fail(VC1ME_EDEVICE, std::string("scratch array unavailable: ") + hipGetErrorString(ctx->scratch_error));
    // every call on this GPU shares its scratch array: one at a time
    std::lock_guard<std::mutex> dev_lock(g_scratch[ctx->device].mu);
    int2* d_cands = static_cast<int2*>(ctx->scratch);
    uint32_t* d_costs = reinterpret_cast<uint32_t*>(static_cast<uint8_t*>(ctx->scratch) + kNoAllocCandBytes);
    // candidates and costs pass through the context's fixed pinned staging
    // (made at creation, the size of the scratch array)
    if (!ctx->h_noalloc) return fail(VC1ME_EDEVICE, "pinned staging unavailable");
    int2* h_c = reinterpret_cast<int2*>(ctx->h_noalloc);
    uint32_t* staged = reinterpret_cast<uint32_t*>(ctx->h_noalloc + kNoAllocCandBytes);
    hipStream_t st = ctx->stream;
    const auto t0 = std::chrono::steady_clock::now();
    std::memcpy(h_c, candidates, n * sizeof(int2));
    HIP_TRY(hipMemcpyAsync(d_cands, h_c, n * sizeof(int2), hipMemcpyHostToDevice, st));
    HIP_TRY(vc1me::launch_score_batch(cur->y, ref->y, cur->w, cur->h, x0, y0, bw, bh, sample_stride, bilinear, rnd,
                                      metric, d_cands, static_cast<int>(n), d_costs, st));
    HIP_TRY(hipMemcpyAsync(staged, d_costs, n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    const auto t1 = std::chrono::steady_clock::now();
    for (size_t i = 0; i < n; ++i) costs_out[i] = staged[i];
    ctx->counters.dispatches++;
    ctx->counters.candidates += n;
    ctx->counters.submit_wait_nanoseconds +=
        static_cast<uint64_t>(std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count());
    return VC1ME_OK;
}

int vc1me_block_costs(vc1me_ctx* ctx, const vc1me_frame* cur, const vc1me_frame* ref, const vc1me_params* params,
                      int x0, int y0, int bw, int bh, int mode, int metric, const vc1me_candidate* candidates,
                      size_t n, uint64_t* costs_out) {
    if (!ctx) return fail(VC1ME_EINVAL, "null context");
    if (int rc = validate_params(params, cur, ref)) return rc;
    if (!((bw == 16 && bh == 16) || (bw == 8 && bh == 8))) return fail(VC1ME_EINVAL, "block must be 16x16 or 8x8");
    if (x0 < 0 || y0 < 0 || x0 >= params->width || y0 >= params->height) return fail(VC1ME_EINVAL, "block outside the raster");
    if (mode < 0 || mode > 3 || (metric != 0 && metric != 1)) return fail(VC1ME_EINVAL, "bad mode or metric");
    if (!n) return VC1ME_OK;
    if (!candidates || !costs_out) return fail(VC1ME_EINVAL, "null candidates or costs");
    const Bounds4 qb = qpel_limits();
    for (size_t i = 0; i < n; ++i)
        if (candidates[i].xq < qb.xmin || candidates[i].xq > qb.xmax || candidates[i].yq < qb.ymin || candidates[i].yq > qb.ymax)
            return fail(VC1ME_ELIMIT, "candidate outside the legal quarter-pel range");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    if (ctx->slots.empty()) ctx->slots.resize(1);
    const vc1me_params* p = params;
    const size_t mbs = static_cast<size_t>((p->width + 15) / 16) * ((p->height + 15) / 16);
    const size_t cells = static_cast<size_t>((p->width + 7) / 8) * ((p->height + 7) / 8);
    hipError_t e = ensure_slot(ctx->slots[0], mbs, cells);
    if (e != hipSuccess) return fail(VC1ME_ENOMEM, "device state allocation");
    e = ensure_phase_alloc(ref, mode);
    if (e != hipSuccess) return fail(VC1ME_ENOMEM, "phase plane allocation");
    HIP_TRY(launch_phases(ref, mode, ctx->stream));
    JobDesc d;
    fill_job(d, p, cur, ref, ctx->slots[0]);
    if (!ctx->d_jobs) {
        HIP_TRY(hipMalloc(&ctx->d_jobs, sizeof(JobDesc)));
        ctx->d_jobs_cap = 1;
    }
    int2* dc = nullptr;
    uint64_t* dout = nullptr;
    HIP_TRY(hipMalloc(&dc, n * sizeof(int2)));
    e = hipMalloc(&dout, n * sizeof(uint64_t));
    if (e != hipSuccess) { (void)hipFree(dc); return fail(VC1ME_ENOMEM, "cost buffer"); }
    hipStream_t st = ctx->stream;
    e = hipMemcpyAsync(ctx->d_jobs, &d, sizeof d, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dc, candidates, n * sizeof(int2), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(ctx->slots[0].err, 0, sizeof(int32_t), st);
    if (e == hipSuccess)
        e = vc1me::launch_block_costs(ctx->d_jobs, x0, y0, bw, bh, mode, metric, dc, static_cast<int>(n), dout,
                                      vc1me::mb_kernel_lds_bytes(d.local_range), st);
    if (e == hipSuccess) e = hipMemcpyAsync(costs_out, dout, n * sizeof(uint64_t), hipMemcpyDeviceToHost, st);
    int32_t err = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&err, ctx->slots[0].err, sizeof err, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(dc);
    (void)hipFree(dout);
    if (e != hipSuccess) return fail(VC1ME_EDEVICE, std::string("block costs: ") + hipGetErrorString(e));
    if (err) return fail(VC1ME_EKERNEL, "kernel reported internal error code " + std::to_string(err));
    return VC1ME_OK;
}

int vc1me_analyze_p_frame(vc1me_ctx* ctx, const vc1me_frame* cur, const vc1me_frame* ref, const vc1me_params* params,
                          vc1me_p_result* out) {
    if (!ctx) return fail(VC1ME_EINVAL, "null context");
    vc1me_job j{cur, ref, params, out};
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    return run_jobs(ctx, &j, 1);
}

int vc1me_analyze_step_seeded(vc1me_ctx* ctx, const vc1me_job* jobs, const vc1me_p_result* const* seeds, int n) {
    if (!ctx) return fail(VC1ME_EINVAL, "null context");
    if (n > 0 && !seeds) return fail(VC1ME_EINVAL, "null seeds");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    return run_jobs(ctx, jobs, n, nullptr, 0.f, {}, seeds);
}

int vc1me_analyze_p_frame_seeded(vc1me_ctx* ctx, const vc1me_frame* cur, const vc1me_frame* ref,
                                 const vc1me_params* params, const vc1me_p_result* seed, vc1me_p_result* out) {
    if (!ctx) return fail(VC1ME_EINVAL, "null context");
    vc1me_job j{cur, ref, params, out};
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    return run_jobs(ctx, &j, 1, nullptr, 0.f, {}, &seed);
}

// ---- intensity compensation

namespace {

// (scale, shift) of intensity_map_y (src/encoder_motion_compensation.cpp:175-187)
void intensity_scale_shift(const vc1me_intensity& ic, int* scale, int* shift) {
    if (ic.lumscale == 0) {
        *scale = -64;
        *shift = (255 - ic.lumshift * 2) * 64;
        if (ic.lumshift > 31) *shift += 128 * 64;
    } else {
        *scale = ic.lumscale + 32;
        *shift = (ic.lumshift > 31 ? ic.lumshift - 64 : ic.lumshift) * 64;
    }
}

// the luma and chroma maps of intensity_map_y / intensity_map_uv
// (src/encoder_motion_compensation.cpp:175-193) as byte tables
vc1me::IntensityLuts intensity_luts(const vc1me_intensity& ic) {
    vc1me::IntensityLuts t{};
    uint8_t* b = reinterpret_cast<uint8_t*>(t.w);
    int scale = 64, shift = 0;
    if (ic.enabled) intensity_scale_shift(ic, &scale, &shift);
    const int uv_scale = ic.lumscale == 0 ? -64 : ic.lumscale + 32;
    for (int v = 0; v < 256; ++v) {
        b[v] = ic.enabled ? static_cast<uint8_t>(std::clamp((scale * v + shift + 32) >> 6, 0, 255))
                          : static_cast<uint8_t>(v);
        b[256 + v] = ic.enabled ? static_cast<uint8_t>(std::clamp((uv_scale * (v - 128) + 128 * 64 + 32) >> 6, 0, 255))
                                : static_cast<uint8_t>(v);
    }
    return t;
}

}  // namespace

namespace {

// The fade line cur ~ gain * ref + offset fitted by least squares to the
// sample moments m = {sum ref, sum cur, sum ref^2, sum ref*cur} of `count`
// samples; a singular system keeps unit gain and takes the mean difference
// as the offset.  The long-double expressions are evaluated in the order
// that fixes the rounding of the coded parameters.
struct FadeLine {
    long double gain, offset;
};

FadeLine fit_fade_line(const unsigned long long m[4], int count) {
    const long double k = static_cast<long double>(count);
    const long double s_ref = static_cast<long double>(m[0]), s_cur = static_cast<long double>(m[1]);
    const long double det = k * static_cast<long double>(m[2]) - s_ref * s_ref;
    if (std::fabs(det) > 1e-9L) {
        const long double gain = (k * static_cast<long double>(m[3]) - s_ref * s_cur) / det;
        return {gain, (s_cur - gain * s_ref) / k};
    }
    return {1.0L, (s_cur - s_ref) / k};
}

// estimate_intensity_comp on resident frames (context lock held)
int estimate_ic(vc1me_ctx* ctx, const vc1me_frame* cur, const vc1me_frame* ref, const int32_t* mvs,
                vc1me_intensity* out) {
    out->enabled = 0; out->lumscale = 32; out->lumshift = 0;
    const int W = cur->w, H = cur->h, mbw = (W + 15) / 16, mbh = (H + 15) / 16;
    const size_t mbs = static_cast<size_t>(mbw) * mbh;
    const int n = ((W + 3) / 4) * ((H + 3) / 4);
    if (n < 16) return VC1ME_OK;      // too few samples: no compensation (:318)
    const int sb = vc1me::ic_sample_blocks(n), gb = vc1me::ic_sad_blocks(n);
    const size_t o_mv = 0, o_smp = (mbs * sizeof(int2) + 255) & ~size_t(255);
    const size_t o_sum = (o_smp + n * sizeof(uint16_t) + 255) & ~size_t(255);
    const size_t o_sad = o_sum + 4 * sizeof(unsigned long long) * sb;
    const size_t need = o_sad + 49 * sizeof(unsigned long long) * gb;
    if (ctx->ic_cap < need) {
        if (ctx->d_ic) (void)hipFree(ctx->d_ic);
        ctx->d_ic = nullptr;
        ctx->ic_cap = 0;
        const hipError_t e = hipMalloc(&ctx->d_ic, need);
        if (e != hipSuccess) return fail(VC1ME_ENOMEM, std::string("intensity buffers: ") + hipGetErrorString(e));
        ctx->ic_cap = need;
    }
    int2* d_mv = reinterpret_cast<int2*>(ctx->d_ic + o_mv);
    uint16_t* d_smp = reinterpret_cast<uint16_t*>(ctx->d_ic + o_smp);
    auto* d_sum = reinterpret_cast<unsigned long long*>(ctx->d_ic + o_sum);
    auto* d_sad = reinterpret_cast<unsigned long long*>(ctx->d_ic + o_sad);
    hipStream_t st = ctx->stream;
    HIP_TRY(hipMemcpyAsync(d_mv, mvs, mbs * sizeof(int2), hipMemcpyHostToDevice, st));
    HIP_TRY(vc1me::launch_ic_samples(cur->y, ref->y, W, H, d_mv, mbw, d_smp, d_sum, st));
    std::vector<unsigned long long> part(std::max<size_t>(4 * sb, 49 * gb));
    HIP_TRY(hipMemcpyAsync(part.data(), d_sum, 4 * sizeof(unsigned long long) * sb, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    unsigned long long sums[4] = {0, 0, 0, 0};
    for (int blk = 0; blk < sb; ++blk)
        for (int k = 0; k < 4; ++k) sums[k] += part[4 * blk + k];
    // cur ~ gain * ref + offset over the samples, in long double on the exact
    // integer moments (src/encoder_motion_compensation.cpp:319-328), coded as
    // LUMSCALE/LUMSHIFT and refined over the 7 x 7 codes around them
    const FadeLine line = fit_fade_line(sums, n);
    const int scale0 = std::clamp(static_cast<int>(std::llround(line.gain * 64.0L - 32.0L)), 1, 63);
    const int shift0 = std::clamp(static_cast<int>(std::llround(line.offset)), -32, 31);
    const int l0 = std::max(1, scale0 - 3), l1 = std::min(63, scale0 + 3);
    const int z0 = std::max(-32, shift0 - 3), z1 = std::min(31, shift0 + 3);
    HIP_TRY(vc1me::launch_ic_sads(d_smp, n, l0, z0, d_sad, st));
    HIP_TRY(hipMemcpyAsync(part.data(), d_sad, 49 * sizeof(unsigned long long) * gb, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    unsigned long long sads[49] = {};
    for (int blk = 0; blk < gb; ++blk)
        for (int k = 0; k < 49; ++k) sads[k] += part[49 * blk + k];
    // the code pair of least sampled SAD; equal SADs go to the pair nearer
    // the identity (scale code 32, shift 0) in L1 distance, then to the first
    // in (scale, shift) scan order (:337-345)
    std::tuple<uint64_t, int, int> least{UINT64_MAX, 0, 0};
    int scale = 32, shift = 0;
    for (int l = l0; l <= l1; ++l)
        for (int z = z0; z <= z1; ++z) {
            const int scan = (l - l0) * 7 + (z - z0);
            const std::tuple<uint64_t, int, int> key{sads[scan], std::abs(l - 32) + std::abs(z), scan};
            if (key < least) { least = key; scale = l; shift = z; }
        }
    if (scale == 32 && shift == 0) return VC1ME_OK;
    out->enabled = 1;
    out->lumscale = scale;
    out->lumshift = shift < 0 ? shift + 64 : shift;
    return VC1ME_OK;
}

}  // namespace

int vc1me_estimate_intensity_comp(vc1me_ctx* ctx, const vc1me_frame* cur, const vc1me_frame* ref,
                                  const int32_t* mvs, vc1me_intensity* out) {
    if (!ctx || !cur || !ref || !mvs || !out) return fail(VC1ME_EINVAL, "null argument");
    if (cur->w != ref->w || cur->h != ref->h) return fail(VC1ME_EINVAL, "cur and ref rasters differ");
    out->enabled = 0; out->lumscale = 32; out->lumshift = 0;
    if (injected_failure()) return fail(VC1ME_EINJECTED, "failure injected by VC1ME_FAIL_AT");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    return estimate_ic(ctx, cur, ref, mvs, out);
}

namespace {

// a new resident frame holding ref under the compensation (context lock held)
int compensated_frame(vc1me_ctx* ctx, const vc1me_frame* ref, const vc1me_intensity* ic, vc1me_frame** out) {
    const size_t ys = static_cast<size_t>(ref->w) * ref->h;
    const size_t cs = static_cast<size_t>(ref->w / 2) * (ref->h / 2);
    auto* f = new vc1me_frame();
    f->w = ref->w;
    f->h = ref->h;
    uint8_t* mem = nullptr;
    // the same 64-byte tail as an uploaded frame: the remap's 16-byte words
    // may run past the last plane into it
    hipError_t e = hipMalloc(&mem, ys + 2 * cs + 64);
    if (e != hipSuccess) { delete f; return fail(VC1ME_ENOMEM, std::string("frame allocation: ") + hipGetErrorString(e)); }
    f->y = mem;
    f->u = mem + ys;
    f->v = mem + ys + cs;
    e = vc1me::launch_ic_remap(ref->y, mem, ys + 2 * cs, ys, intensity_luts(*ic), ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) { (void)hipFree(mem); delete f; return fail(VC1ME_EDEVICE, std::string("intensity remap: ") + hipGetErrorString(e)); }
    *out = f;
    return VC1ME_OK;
}

bool intensity_codes_ok(const vc1me_intensity& ic) {
    return !ic.enabled || (ic.lumscale >= 0 && ic.lumscale <= 63 && ic.lumshift >= 0 && ic.lumshift <= 63);
}

}  // namespace

int vc1me_frame_intensity_compensated(vc1me_ctx* ctx, const vc1me_frame* ref, const vc1me_intensity* ic,
                                      vc1me_frame** out) {
    if (!ctx || !ref || !ic || !out) return fail(VC1ME_EINVAL, "null argument");
    if (!intensity_codes_ok(*ic)) return fail(VC1ME_EINVAL, "LUMSCALE and LUMSHIFT are 6-bit codes");
    if (injected_failure()) return fail(VC1ME_EINJECTED, "failure injected by VC1ME_FAIL_AT");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    return compensated_frame(ctx, ref, ic, out);
}

// ---- motion compensation

namespace {

// a new resident frame of w x h (chroma planes after luma, the 64-byte tail
// of an uploaded frame)
int new_frame(int w, int h, vc1me_frame** out) {
    const size_t ys = static_cast<size_t>(w) * h, cs = static_cast<size_t>(w / 2) * (h / 2);
    auto* f = new vc1me_frame();
    f->w = w;
    f->h = h;
    uint8_t* mem = nullptr;
    const hipError_t e = hipMalloc(&mem, ys + 2 * cs + 64);
    if (e != hipSuccess) { delete f; return fail(VC1ME_ENOMEM, std::string("frame allocation: ") + hipGetErrorString(e)); }
    f->y = mem;
    f->u = mem + ys;
    f->v = mem + ys + cs;
    *out = f;
    return VC1ME_OK;
}

void drop_frame(vc1me_frame* f) {
    (void)hipFree(f->y);
    delete f;
}

// the context's prediction scratch of at least `need` bytes
int ensure_b_scratch(vc1me_ctx* ctx, size_t need) {
    if (ctx->b_cap >= need) return VC1ME_OK;
    if (ctx->d_b) (void)hipFree(ctx->d_b);
    ctx->d_b = nullptr;
    ctx->b_cap = 0;
    const hipError_t e = hipMalloc(&ctx->d_b, need);
    if (e != hipSuccess) return fail(VC1ME_ENOMEM, std::string("prediction buffers: ") + hipGetErrorString(e));
    ctx->b_cap = need;
    return VC1ME_OK;
}

// The P prediction picture (padded = false: the coded raster of params;
// padded = true: its macroblock raster), context lock held.
int predict_p(vc1me_ctx* ctx, const vc1me_frame* ref, const vc1me_params* params, const vc1me_p_result* analysis,
              bool padded, vc1me_frame** out) {
    const int W = params->width, H = params->height;
    const size_t mbs = static_cast<size_t>((W + 15) / 16) * ((H + 15) / 16);
    const int DW = padded ? (W + 15) / 16 * 16 : W, DH = padded ? (H + 15) / 16 * 16 : H;
    const size_t o_mv = 0, o_bmv = (mbs * 8 + 255) & ~size_t(255), o_four = o_bmv + ((mbs * 32 + 255) & ~size_t(255));
    if (int rc = ensure_b_scratch(ctx, o_four + mbs)) return rc;
    vc1me_frame* f = nullptr;
    if (int rc = new_frame(DW, DH, &f)) return rc;
    hipStream_t st = ctx->stream;
    uint8_t* base = ctx->d_b;
    hipError_t e = hipMemcpyAsync(base + o_mv, analysis->mvs, mbs * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(base + o_bmv, analysis->block_mvs, mbs * 32, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(base + o_four, analysis->use_4mv, mbs, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = vc1me::launch_predict_p(ref->y, ref->u, ref->v, ref->w, ref->h, DW, DH, params->mode,
                                    reinterpret_cast<const int2*>(base + o_mv),
                                    reinterpret_cast<const int2*>(base + o_bmv), base + o_four, f->y, f->u, f->v, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) { drop_frame(f); return fail(VC1ME_EDEVICE, std::string("prediction: ") + hipGetErrorString(e)); }
    *out = f;
    return VC1ME_OK;
}

int check_predict_p(vc1me_ctx* ctx, const vc1me_frame* ref, const vc1me_params* params, const vc1me_p_result* analysis,
                    vc1me_frame** out) {
    if (!ctx || !ref || !params || !analysis || !out) return fail(VC1ME_EINVAL, "null argument");
    if (!analysis->mvs || !analysis->block_mvs || !analysis->use_4mv)
        return fail(VC1ME_EINVAL, "the analysis lacks mvs, block_mvs or use_4mv");
    if (params->width <= 0 || params->height <= 0 || (params->width & 1) || (params->height & 1))
        return fail(VC1ME_EINVAL, "raster must be positive and even");
    if (params->mode < VC1ME_MODE_ONE_MV_QPEL || params->mode > VC1ME_MODE_MIXED_MV) return fail(VC1ME_EINVAL, "bad mode");
    return VC1ME_OK;
}

// The bi-directional prediction picture, context lock held.
int predict_bi(vc1me_ctx* ctx, const vc1me_frame* past, const vc1me_frame* future, const vc1me_params* params,
               const int32_t* fmvs, const int32_t* bmvs, bool padded, vc1me_frame** out) {
    const int W = params->width, H = params->height;
    const size_t mbs = static_cast<size_t>((W + 15) / 16) * ((H + 15) / 16);
    const int DW = padded ? (W + 15) / 16 * 16 : W, DH = padded ? (H + 15) / 16 * 16 : H;
    const size_t o_f = 0, o_b = (mbs * 8 + 255) & ~size_t(255);
    if (int rc = ensure_b_scratch(ctx, o_b + mbs * 8)) return rc;
    vc1me_frame* f = nullptr;
    if (int rc = new_frame(DW, DH, &f)) return rc;
    hipStream_t st = ctx->stream;
    uint8_t* base = ctx->d_b;
    hipError_t e = hipMemcpyAsync(base + o_f, fmvs, mbs * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(base + o_b, bmvs, mbs * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = vc1me::launch_predict_bi(past->y, past->w, past->h, future->y, future->w, future->h, DW, DH, params->mode,
                                     reinterpret_cast<const int2*>(base + o_f),
                                     reinterpret_cast<const int2*>(base + o_b), f->y, f->u, f->v, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) { drop_frame(f); return fail(VC1ME_EDEVICE, std::string("prediction: ") + hipGetErrorString(e)); }
    *out = f;
    return VC1ME_OK;
}

int check_predict_bi(vc1me_ctx* ctx, const vc1me_frame* past, const vc1me_frame* future, const vc1me_params* params,
                     const int32_t* fmvs, const int32_t* bmvs, vc1me_frame** out) {
    if (!ctx || !past || !future || !params || !fmvs || !bmvs || !out) return fail(VC1ME_EINVAL, "null argument");
    if (params->width <= 0 || params->height <= 0 || (params->width & 1) || (params->height & 1))
        return fail(VC1ME_EINVAL, "raster must be positive and even");
    if (params->mode < VC1ME_MODE_ONE_MV_QPEL || params->mode > VC1ME_MODE_ONE_MV_HPEL_BILINEAR)
        return fail(VC1ME_EINVAL, "bi-directional prediction takes a 1-MV mode");
    return VC1ME_OK;
}

}  // namespace

int vc1me_predict_p_frame(vc1me_ctx* ctx, const vc1me_frame* ref, const vc1me_params* params,
                          const vc1me_p_result* analysis, vc1me_frame** out) {
    if (int rc = check_predict_p(ctx, ref, params, analysis, out)) return rc;
    if (params->width != ref->w || params->height != ref->h) return fail(VC1ME_EINVAL, "raster differs from the parameters");
    if (injected_failure()) return fail(VC1ME_EINJECTED, "failure injected by VC1ME_FAIL_AT");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    return predict_p(ctx, ref, params, analysis, false, out);
}

int vc1me_predict_p_frame_padded(vc1me_ctx* ctx, const vc1me_frame* ref, const vc1me_params* params,
                                 const vc1me_p_result* analysis, vc1me_frame** out) {
    if (int rc = check_predict_p(ctx, ref, params, analysis, out)) return rc;
    if (injected_failure()) return fail(VC1ME_EINJECTED, "failure injected by VC1ME_FAIL_AT");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    return predict_p(ctx, ref, params, analysis, true, out);
}

int vc1me_predict_bi_frame(vc1me_ctx* ctx, const vc1me_frame* past, const vc1me_frame* future,
                           const vc1me_params* params, const int32_t* fmvs, const int32_t* bmvs, vc1me_frame** out) {
    if (int rc = check_predict_bi(ctx, past, future, params, fmvs, bmvs, out)) return rc;
    if (params->width != past->w || params->height != past->h || future->w != past->w || future->h != past->h)
        return fail(VC1ME_EINVAL, "raster differs from the parameters");
    if (injected_failure()) return fail(VC1ME_EINJECTED, "failure injected by VC1ME_FAIL_AT");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    return predict_bi(ctx, past, future, params, fmvs, bmvs, false, out);
}

int vc1me_predict_bi_frame_padded(vc1me_ctx* ctx, const vc1me_frame* past, const vc1me_frame* future,
                                  const vc1me_params* params, const int32_t* fmvs, const int32_t* bmvs,
                                  vc1me_frame** out) {
    if (int rc = check_predict_bi(ctx, past, future, params, fmvs, bmvs, out)) return rc;
    if (injected_failure()) return fail(VC1ME_EINJECTED, "failure injected by VC1ME_FAIL_AT");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    return predict_bi(ctx, past, future, params, fmvs, bmvs, true, out);
}

// ---- B pictures

namespace {

int check_b_job(const vc1me_b_job& j) {
    if (!j.cur || !j.past || !j.future || !j.params || !j.b || !j.out) return fail(VC1ME_EINVAL, "null argument");
    const vc1me_b_params* b = j.b;
    if (!b->future_anchor_mvs || !b->future_anchor_4mv) return fail(VC1ME_EINVAL, "null future anchor fields");
    const vc1me_frame* cur = j.cur;
    if (j.past->w != cur->w || j.past->h != cur->h || j.future->w != cur->w || j.future->h != cur->h ||
        j.params->width != cur->w || j.params->height != cur->h)
        return fail(VC1ME_EINVAL, "cur, past, future and params must share one raster");
    const int fn = b->fraction_num, fd = b->fraction_den;
    if (!((fd == 2 && fn == 1) || (fd == 3 && (fn == 1 || fn == 2))))
        return fail(VC1ME_EINVAL, "unsupported B-picture temporal fraction");
    if (j.params->mode < VC1ME_MODE_ONE_MV_QPEL || j.params->mode > VC1ME_MODE_ONE_MV_HPEL_BILINEAR)
        return fail(VC1ME_EINVAL, "B pictures take a 1-MV mode");
    if (!intensity_codes_ok(b->forward_intensity)) return fail(VC1ME_EINVAL, "LUMSCALE and LUMSHIFT are 6-bit codes");
    if ((cur->h + 15) / 16 > vc1me::kBMaxRows)
        return fail(VC1ME_ELIMIT, "B pictures are limited to " + std::to_string(vc1me::kBMaxRows * 16) + " rows");
    return VC1ME_OK;
}

void b_reset_counts(vc1me_b_result* out) {
    out->forward_macroblocks = out->backward_macroblocks = out->interpolated_macroblocks = 0;
    out->direct_macroblocks = out->intra_macroblocks = out->moved_macroblocks = 0;
    out->fractional_chroma_macroblocks = 0;
    out->mean_abs_residual = out->rate_complexity = 0.0;
    out->motion_sad = out->motion_bits = 0;
}

// the encoder's minimal zero-motion configuration: every macroblock forward
// with zero vectors (src/encoder_motion.cpp:775-787)
void b_zero_motion(vc1me_b_result* out, size_t mbs) {
    for (size_t i = 0; i < mbs; ++i) {
        if (out->modes) out->modes[i] = VC1ME_B_FORWARD;
        if (out->forward_mvs) out->forward_mvs[2 * i] = out->forward_mvs[2 * i + 1] = 0;
        if (out->backward_mvs) out->backward_mvs[2 * i] = out->backward_mvs[2 * i + 1] = 0;
        if (out->inter_intra_cost_ratio) out->inter_intra_cost_ratio[i] = -1.0;
        if (out->forward_distant_local_mae) out->forward_distant_local_mae[i] = -1.0;
        if (out->forward_distant_candidate_mae) out->forward_distant_candidate_mae[i] = -1.0;
        if (out->forward_distant_match_decision) out->forward_distant_match_decision[i] = 0;
        if (out->backward_distant_local_mae) out->backward_distant_local_mae[i] = -1.0;
        if (out->backward_distant_candidate_mae) out->backward_distant_candidate_mae[i] = -1.0;
        if (out->backward_distant_match_decision) out->backward_distant_match_decision[i] = 0;
    }
    out->forward_macroblocks = mbs;
}

// Device layout of one B picture's decision state in the context's B buffer.
struct BLayout {
    size_t amv, a4, cost, isad, dir, fmv, bmv, ratio, ssad, rsad, srate, modes, end;
    explicit BLayout(size_t base, size_t mbs) {
        auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
        amv = base;
        a4 = up(amv + mbs * 8);
        cost = up(a4 + mbs);
        isad = cost + mbs * 32;
        dir = isad + mbs * 8;
        fmv = dir + mbs * 16;
        bmv = fmv + mbs * 8;
        ratio = bmv + mbs * 8;
        ssad = ratio + mbs * 8;
        rsad = ssad + mbs * 8;
        srate = rsad + mbs * 8;
        modes = srate + mbs * 4;
        end = up(modes + mbs);
    }
};

}  // namespace

int vc1me_analyze_b_step(vc1me_ctx* ctx, const vc1me_b_job* bjobs, int n) {
    if (!ctx) return fail(VC1ME_EINVAL, "null context");
    if (n < 0 || (n > 0 && !bjobs)) return fail(VC1ME_EINVAL, "bad B job list");
    if (2 * n > kMaxJobs) return fail(VC1ME_ELIMIT, "at most " + std::to_string(kMaxJobs / 2) + " B pictures per step");
    for (int i = 0; i < n; ++i)
        if (int rc = check_b_job(bjobs[i])) return rc;
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    // pictures that run: not under the zero-motion shortcut
    std::vector<int> run;
    for (int i = 0; i < n; ++i) {
        const vc1me_b_job& j = bjobs[i];
        b_reset_counts(j.out);
        const size_t mbs = static_cast<size_t>((j.cur->w + 15) / 16) * ((j.cur->h + 15) / 16);
        if (j.b->zero_motion_shortcut && j.params->motion_search_range == 0) b_zero_motion(j.out, mbs);
        else run.push_back(i);
    }
    if (run.empty()) return VC1ME_OK;
    const int nb = static_cast<int>(run.size());
    // forward references carry the anchor's intensity compensation; one
    // compensated frame per distinct (past, code)
    struct Comp { const vc1me_frame* past; vc1me_intensity ic; vc1me_frame* frame; };
    struct Comps {
        std::vector<Comp> v;
        ~Comps() { for (auto& c : v) if (c.frame) { (void)hipFree(c.frame->y); delete c.frame; } }
    } comps;
    std::vector<const vc1me_frame*> fwd_ref(nb);
    for (int k = 0; k < nb; ++k) {
        const vc1me_b_job& j = bjobs[run[k]];
        fwd_ref[k] = j.past;
        const vc1me_intensity& ic = j.b->forward_intensity;
        if (!ic.enabled) continue;
        const vc1me_frame* found = nullptr;
        for (const auto& c : comps.v)
            if (c.past == j.past && c.ic.lumscale == ic.lumscale && c.ic.lumshift == ic.lumshift) found = c.frame;
        if (!found) {
            vc1me_frame* f = nullptr;
            if (int rc = compensated_frame(ctx, j.past, &ic, &f)) return rc;
            comps.v.push_back({j.past, ic, f});
            found = f;
        }
        fwd_ref[k] = found;
    }
    // the core passes of every picture in one step: job 2k forward, 2k + 1 backward
    std::vector<vc1me_params> qs(nb);
    std::vector<vc1me_p_result> pres(2 * static_cast<size_t>(nb));
    std::vector<vc1me_job> jobs(2 * static_cast<size_t>(nb));
    for (int k = 0; k < nb; ++k) {
        const vc1me_b_job& j = bjobs[run[k]];
        qs[k] = *j.params;
        qs[k].allow_intra = 0;
        vc1me_p_result& fo = pres[2 * k];
        vc1me_p_result& bo = pres[2 * k + 1];
        fo = vc1me_p_result{};
        bo = vc1me_p_result{};
        fo.distant_local_mae = j.out->forward_distant_local_mae;
        fo.distant_candidate_mae = j.out->forward_distant_candidate_mae;
        fo.distant_match_decision = j.out->forward_distant_match_decision;
        bo.distant_local_mae = j.out->backward_distant_local_mae;
        bo.distant_candidate_mae = j.out->backward_distant_candidate_mae;
        bo.distant_match_decision = j.out->backward_distant_match_decision;
        jobs[2 * k] = {j.cur, fwd_ref[k], &qs[k], &fo};
        jobs[2 * k + 1] = {j.cur, j.future, &qs[k], &bo};
    }
    if (int rc = run_jobs(ctx, jobs.data(), 2 * nb)) return rc;
    // the decision reads the step's jobs in place (vectors, screen SADs, phase
    // planes), so it runs before anything else on the context
    const size_t o_desc = 0;
    size_t off = ((sizeof(vc1me::BDesc) * nb) + 255) & ~size_t(255);
    std::vector<BLayout> lay;
    int max_mbs = 0, max_mbh = 0;
    for (int k = 0; k < nb; ++k) {
        const vc1me_frame* cur = bjobs[run[k]].cur;
        const int mbw = (cur->w + 15) / 16, mbh = (cur->h + 15) / 16;
        lay.emplace_back(off, static_cast<size_t>(mbw) * mbh);
        off = lay.back().end;
        max_mbs = std::max(max_mbs, mbw * mbh);
        max_mbh = std::max(max_mbh, mbh);
    }
    if (ctx->b_cap < off) {
        if (ctx->d_b) (void)hipFree(ctx->d_b);
        ctx->d_b = nullptr;
        ctx->b_cap = 0;
        const hipError_t e = hipMalloc(&ctx->d_b, off);
        if (e != hipSuccess) return fail(VC1ME_ENOMEM, std::string("B-picture buffers: ") + hipGetErrorString(e));
        ctx->b_cap = off;
    }
    uint8_t* base = ctx->d_b;
    hipStream_t st = ctx->stream;
    std::vector<vc1me::BDesc> descs(nb);
    for (int k = 0; k < nb; ++k) {
        const vc1me_b_job& j = bjobs[run[k]];
        const BLayout& L = lay[k];
        const size_t mbs = static_cast<size_t>((j.cur->w + 15) / 16) * ((j.cur->h + 15) / 16);
        const int fn = j.b->fraction_num, fd = j.b->fraction_den;
        vc1me::BDesc& d = descs[k];
        d = vc1me::BDesc{};
        d.anchor_mvs = reinterpret_cast<const int2*>(base + L.amv);
        d.anchor_4mv = base + L.a4;
        d.mode = j.params->mode;
        d.bfrac = fd == 2 ? 128 : (fn == 1 ? 85 : 170);
        d.below_half = 2 * fn < fd ? 1 : 0;
        d.allow_intra = j.b->allow_intra ? 1 : 0;
        d.lambda = std::max(1.0, 0.45 * static_cast<double>(j.params->pqindex));
        d.intra_guard = static_cast<uint64_t>(std::llround(d.lambda * 72.0));
        d.intra_thr = j.params->inter_intra_threshold;
        d.cost = reinterpret_cast<uint64_t*>(base + L.cost);
        d.intra_sad = reinterpret_cast<uint64_t*>(base + L.isad);
        d.direct = reinterpret_cast<int2*>(base + L.dir);
        d.fmv = reinterpret_cast<int2*>(base + L.fmv);
        d.bmv = reinterpret_cast<int2*>(base + L.bmv);
        d.ratio = reinterpret_cast<double*>(base + L.ratio);
        d.sel_sad = reinterpret_cast<uint64_t*>(base + L.ssad);
        d.rate_sad = reinterpret_cast<uint64_t*>(base + L.rsad);
        d.sel_rate = reinterpret_cast<uint32_t*>(base + L.srate);
        d.modes = base + L.modes;
        HIP_TRY(hipMemcpyAsync(base + L.amv, j.b->future_anchor_mvs, mbs * sizeof(int2), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(base + L.a4, j.b->future_anchor_4mv, mbs, hipMemcpyHostToDevice, st));
    }
    HIP_TRY(hipMemcpyAsync(base + o_desc, descs.data(), sizeof(vc1me::BDesc) * nb, hipMemcpyHostToDevice, st));
    const auto* d_desc = reinterpret_cast<const vc1me::BDesc*>(base + o_desc);
    HIP_TRY(vc1me::launch_b_costs(ctx->d_jobs, d_desc, nb, max_mbs, st));
    HIP_TRY(vc1me::launch_b_decide(ctx->d_jobs, d_desc, nb, max_mbh, st));
    // results: one copy of each picture's decision block
    for (int k = 0; k < nb; ++k) {
        const vc1me_b_job& j = bjobs[run[k]];
        const BLayout& L = lay[k];
        const int mbw = (j.cur->w + 15) / 16;
        const size_t mbs = static_cast<size_t>(mbw) * ((j.cur->h + 15) / 16);
        std::vector<uint8_t> host(L.end - L.fmv);
        HIP_TRY(hipMemcpyAsync(host.data(), base + L.fmv, host.size(), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        auto at = [&](size_t o) { return host.data() + (o - L.fmv); };
        const int2* fmv = reinterpret_cast<const int2*>(at(L.fmv));
        const int2* bmv = reinterpret_cast<const int2*>(at(L.bmv));
        const double* ratio = reinterpret_cast<const double*>(at(L.ratio));
        const uint64_t* ssad = reinterpret_cast<const uint64_t*>(at(L.ssad));
        const uint64_t* rsad = reinterpret_cast<const uint64_t*>(at(L.rsad));
        const uint32_t* srate = reinterpret_cast<const uint32_t*>(at(L.srate));
        const uint8_t* modes = at(L.modes);
        vc1me_b_result* out = j.out;
        // picture totals (src/encoder_motion.cpp:933-1003)
        uint64_t sel_sad = 0, rate_sad = 0, sel_rate = 1, sel_px = 0;
        for (size_t i = 0; i < mbs; ++i) {
            const int mx = static_cast<int>(i % mbw), my = static_cast<int>(i / mbw);
            sel_sad += ssad[i];
            rate_sad += rsad[i];
            sel_rate += srate[i];
            sel_px += static_cast<uint64_t>(std::max(0, std::min(16, j.cur->w - mx * 16))) *
                      static_cast<uint64_t>(std::max(0, std::min(16, j.cur->h - my * 16)));
            switch (modes[i]) {
                case VC1ME_B_FORWARD: ++out->forward_macroblocks; break;
                case VC1ME_B_BACKWARD: ++out->backward_macroblocks; break;
                case VC1ME_B_INTERPOLATED: ++out->interpolated_macroblocks; break;
                case VC1ME_B_DIRECT: ++out->direct_macroblocks; break;
                default: ++out->intra_macroblocks; break;
            }
            const int2 f = fmv[i], g = bmv[i];
            if (f.x || f.y || g.x || g.y) ++out->moved_macroblocks;
            auto frac_chroma = [](int2 v) { return (v.x & 3) || (v.y & 3) || ((v.x / 4) & 1) || ((v.y / 4) & 1); };
            if (frac_chroma(f) || frac_chroma(g)) ++out->fractional_chroma_macroblocks;
            if (out->modes) out->modes[i] = modes[i];
            if (out->forward_mvs) { out->forward_mvs[2 * i] = f.x; out->forward_mvs[2 * i + 1] = f.y; }
            if (out->backward_mvs) { out->backward_mvs[2 * i] = g.x; out->backward_mvs[2 * i + 1] = g.y; }
            if (out->inter_intra_cost_ratio) out->inter_intra_cost_ratio[i] = ratio[i];
        }
        out->mean_abs_residual = sel_px ? static_cast<double>(sel_sad) / static_cast<double>(sel_px) : 0.0;
        out->rate_complexity = sel_px ? static_cast<double>(rate_sad) / static_cast<double>(sel_px) : 0.0;
        out->motion_sad = sel_sad;
        out->motion_bits = sel_rate;
    }
    return VC1ME_OK;
}

int vc1me_analyze_b_frame(vc1me_ctx* ctx, const vc1me_frame* cur, const vc1me_frame* past, const vc1me_frame* future,
                          const vc1me_params* params, const vc1me_b_params* b, vc1me_b_result* out) {
    const vc1me_b_job j{cur, past, future, params, b, out};
    return vc1me_analyze_b_step(ctx, &j, 1);
}

// ---- scene cuts

namespace {

// global_motion_search(cur, ref, range) on its own (range clamped to 32 by the
// search itself, src/encoder_motion.cpp:2076-2141); context lock held
int global_only(vc1me_ctx* ctx, const vc1me_frame* cur, const vc1me_frame* ref, const vc1me_params* p, int range,
                vc1me::GlobalState* out) {
    vc1me_params q = *p;
    // a job whose global search runs at min(range, 32); range 0 is the plain
    // frame SAD at (0, 0)
    q.motion_search_range = std::max(range, 1);
    q.motion_local_search_range = std::max(range, 0);
    const size_t mbs = static_cast<size_t>((cur->w + 15) / 16) * ((cur->h + 15) / 16);
    const size_t cells = static_cast<size_t>((cur->w + 7) / 8) * ((cur->h + 7) / 8);
    if (ctx->slots.empty()) ctx->slots.resize(1);
    if (ctx->h_jobs.empty()) ctx->h_jobs.resize(1);
    hipError_t e = ensure_slot(ctx->slots[0], mbs, cells);
    if (e != hipSuccess) return fail(VC1ME_ENOMEM, std::string("device state allocation: ") + hipGetErrorString(e));
    if (ctx->d_jobs_cap < 1) {
        HIP_TRY(hipMalloc(&ctx->d_jobs, sizeof(JobDesc)));
        ctx->d_jobs_cap = 1;
    }
    JobDesc& d = ctx->h_jobs[0];
    fill_job(d, &q, cur, ref, ctx->slots[0]);
    hipStream_t st = ctx->stream;
    HIP_TRY(hipMemcpyAsync(ctx->d_jobs, &d, sizeof(JobDesc), hipMemcpyHostToDevice, st));
    HIP_TRY(vc1me::launch_global_search(ctx->d_jobs, 1, &d, st, nullptr, nullptr));
    HIP_TRY(hipMemcpyAsync(out, ctx->slots[0].g, sizeof(vc1me::GlobalState), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    ctx->last_jobs = 0;
    return VC1ME_OK;
}

// Whether the interior of a 1-MV quarter-pel field (macroblocks off the
// picture border) moves as one translation plus an isotropic zoom about the
// centroid of their centres (src/encoder_motion.cpp:637-673): the zoom is
// the least-squares slope of the centred displacements (pixels) against the
// centred positions; the field is coherent when at least 60% of at least 12
// interior macroblocks lie within 2.5 px of the fitted motion and the zoom's
// magnitude is in [0.005, 0.08].  Sums run in raster order in long double,
// the order that fixes their rounding.
bool zoom_coherent(const int32_t* mv, int mbw, int mbh) {
    struct Sample {
        long double x, y, u, v;   // centre (pixels) and displacement (pixels)
    };
    std::vector<Sample> in;
    for (int by = 1; by < mbh - 1; ++by)
        for (int bx = 1; bx < mbw - 1; ++bx) {
            const size_t at = 2 * (static_cast<size_t>(by) * mbw + bx);
            in.push_back({bx * 16.0L + 8.0L, by * 16.0L + 8.0L, mv[at] / 4.0L, mv[at + 1] / 4.0L});
        }
    if (in.size() < 12) return false;
    Sample mean{0.0L, 0.0L, 0.0L, 0.0L};
    for (const Sample& e : in) { mean.x += e.x; mean.y += e.y; mean.u += e.u; mean.v += e.v; }
    const size_t count = in.size();
    mean.x /= count; mean.y /= count; mean.u /= count; mean.v /= count;
    long double cross = 0.0L, spread = 0.0L;
    for (const Sample& e : in) {
        const long double x = e.x - mean.x, y = e.y - mean.y;
        cross += x * (e.u - mean.u) + y * (e.v - mean.v);
        spread += x * x + y * y;
    }
    const long double zoom = spread > 0.0L ? cross / spread : 0.0L;
    size_t agree = 0;
    for (const Sample& e : in) {
        const long double x = e.x - mean.x, y = e.y - mean.y;
        const long double ru = e.u - (mean.u + zoom * x), rv = e.v - (mean.v + zoom * y);
        agree += ru * ru + rv * rv <= 6.25L ? 1 : 0;
    }
    const long double mag = std::fabs(zoom);
    return agree * 5 >= count * 3 && mag >= 0.005L && mag <= 0.08L;
}

}  // namespace

int vc1me_scene_change_score(vc1me_ctx* ctx, const vc1me_frame* cur, const vc1me_frame* ref,
                             const vc1me_params* params, int fade_compensation, double scene_threshold,
                             double* score) {
    if (!ctx || !cur || !ref || !params || !score) return fail(VC1ME_EINVAL, "null argument");
    if (cur->w != ref->w || cur->h != ref->h || params->width != cur->w || params->height != cur->h)
        return fail(VC1ME_EINVAL, "scene-analysis dimensions do not match the parameters");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    const int W = cur->w, H = cur->h, mbw = (W + 15) / 16, mbh = (H + 15) / 16;
    const size_t mbs = static_cast<size_t>(mbw) * mbh;
    const uint64_t pixels = static_cast<uint64_t>(W) * H;
    vc1me::GlobalState g;
    if (int rc = global_only(ctx, cur, ref, params, params->motion_search_range, &g)) return rc;
    double best = pixels ? static_cast<double>(g.full_sad) / static_cast<double>(pixels) : 0.0;
    if (fade_compensation) {
        // the fade fit on the uniform global field, then the global search on
        // the compensated reference (src/encoder_motion.cpp:606-615)
        std::vector<int32_t> uni(2 * mbs);
        for (size_t i = 0; i < mbs; ++i) { uni[2 * i] = g.dx * 4; uni[2 * i + 1] = g.dy * 4; }
        vc1me_intensity ic;
        if (int rc = estimate_ic(ctx, cur, ref, uni.data(), &ic)) return rc;
        if (ic.enabled) {
            vc1me_frame* comp = nullptr;
            if (int rc = compensated_frame(ctx, ref, &ic, &comp)) return rc;
            vc1me::GlobalState wg;
            const int rc = global_only(ctx, cur, comp, params, params->motion_search_range, &wg);
            (void)hipFree(comp->y);
            delete comp;
            if (rc) return rc;
            const double wm = pixels ? static_cast<double>(wg.full_sad) / static_cast<double>(pixels) : 0.0;
            best = std::min(best, wm);
        }
    }
    if (best >= scene_threshold) {
        // a zoom rescue: the 1-MV quarter-pel core, its interior field fitted
        // to translation + isotropic zoom (src/encoder_motion.cpp:627-673)
        vc1me_params q = *params;
        q.mode = VC1ME_MODE_ONE_MV_QPEL;
        q.allow_intra = 0;
        std::vector<int32_t> mv(2 * mbs);
        vc1me_p_result local{};
        local.mvs = mv.data();
        const vc1me_job job{cur, ref, &q, &local};
        if (int rc = run_jobs(ctx, &job, 1)) return rc;
        const double lm = local.mean_abs_residual;
        const bool coherent = zoom_coherent(mv.data(), mbw, mbh);
        if (coherent && lm < best * 0.72 && lm < scene_threshold * 0.86) best = lm;
    }
    *score = best;
    return VC1ME_OK;
}

int vc1me_has_good_motion_match(vc1me_ctx* ctx, const vc1me_frame* cur, const vc1me_frame* ref,
                                const vc1me_params* params, int fade_compensation, int* good, int32_t* field) {
    if (!ctx || !cur || !ref || !params || !good) return fail(VC1ME_EINVAL, "null argument");
    if (cur->w != ref->w || cur->h != ref->h || params->width != cur->w || params->height != cur->h)
        return fail(VC1ME_EINVAL, "motion-failure analysis dimensions do not match the parameters");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    *good = 0;
    const int W = cur->w, H = cur->h, mbw = (W + 15) / 16, mbh = (H + 15) / 16;
    const size_t mbs = static_cast<size_t>(mbw) * mbh;
    const int range = std::max(0, params->motion_search_range);
    const int local_range = std::min(range, std::max(0, params->motion_local_search_range));
    // the fade fit on the uniform global field picks the prediction reference
    // (src/encoder_motion.cpp:684-697); the step's own global search then runs
    // on that reference, as the reference's second global search does
    vc1me_frame* comp = nullptr;
    if (fade_compensation) {
        vc1me::GlobalState g;
        if (int rc = global_only(ctx, cur, ref, params, local_range, &g)) return rc;
        std::vector<int32_t> uni(2 * mbs);
        for (size_t i = 0; i < mbs; ++i) { uni[2 * i] = g.dx * 4; uni[2 * i + 1] = g.dy * 4; }
        vc1me_intensity ic;
        if (int rc = estimate_ic(ctx, cur, ref, uni.data(), &ic)) return rc;
        if (ic.enabled)
            if (int rc = compensated_frame(ctx, ref, &ic, &comp)) return rc;
    }
    struct Release {
        vc1me_frame* f;
        ~Release() { if (f) { (void)hipFree(f->y); delete f; } }
    } release{comp};
    vc1me_params q = *params;
    q.mode = VC1ME_MODE_ONE_MV_QPEL;
    q.allow_intra = 0;
    q.motion_search_range = range;
    const vc1me_job job{cur, comp ? comp : ref, &q, nullptr};
    if (int rc = run_jobs(ctx, &job, 1, nullptr, 0.f, {}, nullptr, true)) return rc;
    // a usable match anywhere: the raster scan's early return finds the first
    // one, and every macroblock's analysis equals the scan's up to there
    std::vector<uint8_t> flags(mbs);
    HIP_TRY(hipMemcpy(flags.data(), ctx->slots[0].flags, mbs, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < mbs && !*good; ++i) *good = (flags[i] & vc1me::kFlagGoodMatch) ? 1 : 0;
    if (field) HIP_TRY(hipMemcpy(field, ctx->slots[0].mvs, mbs * sizeof(int2), hipMemcpyDeviceToHost));
    return VC1ME_OK;
}

int vc1me_frame_download(vc1me_ctx* ctx, const vc1me_frame* frame, uint8_t* y, uint8_t* u, uint8_t* v) {
    if (!ctx || !frame || !y || !u || !v) return fail(VC1ME_EINVAL, "null argument");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    const size_t ys = static_cast<size_t>(frame->w) * frame->h;
    const size_t cs = static_cast<size_t>(frame->w / 2) * (frame->h / 2);
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipMemcpy(y, frame->y, ys, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(u, frame->u, cs, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(v, frame->v, cs, hipMemcpyDeviceToHost));
    return VC1ME_OK;
}

int vc1me_analyze_step(vc1me_ctx* ctx, const vc1me_job* jobs, int n) {
    if (!ctx) return fail(VC1ME_EINVAL, "null context");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    return run_jobs(ctx, jobs, n);
}

int vc1me_last_timing(vc1me_ctx* ctx, vc1me_timing* out) {
    if (!ctx || !out) return fail(VC1ME_EINVAL, "null argument");
    std::lock_guard<std::mutex> lock(ctx->mu);
    *out = ctx->timing;
    return VC1ME_OK;
}

int vc1me_last_work(vc1me_ctx* ctx, vc1me_work* out) {
    if (!ctx || !out) return fail(VC1ME_EINVAL, "null argument");
    std::lock_guard<std::mutex> lock(ctx->mu);
    *out = ctx->work;
    return VC1ME_OK;
}

int vc1me_set_profiling(vc1me_ctx* ctx, int on) {
    if (!ctx) return fail(VC1ME_EINVAL, "null context");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->profiling = on != 0;
    return VC1ME_OK;
}

int vc1me_set_learned_candidates(vc1me_ctx* ctx, int on) {
    if (!ctx) return fail(VC1ME_EINVAL, "null context");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->learned_candidates = on ? 1 : 0;
    return VC1ME_OK;
}

int vc1me_set_graphs(vc1me_ctx* ctx, int on) {
    if (!ctx) return fail(VC1ME_EINVAL, "null context");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->use_graphs = on != 0;
    return VC1ME_OK;
}

int vc1me_last_relax_passes(vc1me_ctx* ctx, int64_t* analyses, double* ms, int cap) {
    if (!ctx || cap < 0) return fail(VC1ME_EINVAL, "bad argument");
    std::lock_guard<std::mutex> lock(ctx->mu);
    const int np = static_cast<int>(ctx->pass_analyses.size());
    for (int k = 0; k < np && k < cap; ++k) {
        if (analyses) analyses[k] = ctx->pass_analyses[k];
        if (ms) ms[k] = k < static_cast<int>(ctx->pass_ms.size()) ? ctx->pass_ms[k] : -1.0;
    }
    return np;
}

int vc1me_set_schedule(vc1me_ctx* ctx, int schedule) {
    if (!ctx) return fail(VC1ME_EINVAL, "null context");
    if (schedule != VC1ME_SCHEDULE_DIAGONAL && schedule != VC1ME_SCHEDULE_RELAX)
        return fail(VC1ME_EINVAL, "unknown schedule " + std::to_string(schedule));
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->schedule = schedule;
    return VC1ME_OK;
}

int vc1me_set_cost_report(vc1me_ctx* ctx, int on) {
    if (!ctx) return fail(VC1ME_EINVAL, "null context");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->cost_report = on != 0;
    return VC1ME_OK;
}

int vc1me_last_mb_costs(vc1me_ctx* ctx, int job, uint64_t* cost20, int32_t* one_mvs) {
    if (!ctx || !cost20) return fail(VC1ME_EINVAL, "null argument");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!ctx->cost_report) return fail(VC1ME_EINVAL, "cost report is off");
    if (job < 0 || job >= ctx->last_jobs) return fail(VC1ME_EINVAL, "no such job in the last step");
    if (int rc = set_device(ctx)) return rc;
    const JobSlot& s = ctx->slots[job];
    HIP_TRY(hipMemcpy(cost20, s.cost20, sizeof(uint64_t) * s.mbs_last, hipMemcpyDeviceToHost));
    if (one_mvs) HIP_TRY(hipMemcpy(one_mvs, s.one_mv, sizeof(int2) * s.mbs_last, hipMemcpyDeviceToHost));
    return VC1ME_OK;
}

int vc1me_counters_get(vc1me_ctx* ctx, vc1me_counters* out) {
    if (!ctx || !out) return fail(VC1ME_EINVAL, "null argument");
    std::lock_guard<std::mutex> lock(ctx->mu);
    *out = ctx->counters;
    return VC1ME_OK;
}

void vc1me_counters_reset(vc1me_ctx* ctx) {
    if (!ctx) return;
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->counters = vc1me_counters{};
}


int vc1me_nn_param_count(void) { return pnet::kParams; }

int vc1me_nn_create(vc1me_ctx* ctx, vc1me_nn** out) {
    if (!ctx || !out) return fail(VC1ME_EINVAL, "null argument");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    vc1me_nn* nn = new vc1me_nn();
    nn->ctx = ctx;
    const size_t pb = sizeof(float) * pnet::kParams;
    hipError_t e = hipMalloc(&nn->p, pb);
    if (e == hipSuccess) e = hipMalloc(&nn->g, pb);
    if (e == hipSuccess) e = hipMalloc(&nn->m, pb);
    if (e == hipSuccess) e = hipMalloc(&nn->v, pb);
    for (int l = 0; l < pnet::kLayers && e == hipSuccess; ++l) {
        e = hipMalloc(&nn->wt[l], 2 * 32 * 9 * pnet::layer_cin(l));
        if (e == hipSuccess && l) e = hipMalloc(&nn->wtt[l], 2 * 32 * 9 * 32);
    }
    if (e == hipSuccess) e = hipMemset(nn->p, 0, pb);
    if (e == hipSuccess) e = hipMemset(nn->m, 0, pb);
    if (e == hipSuccess) e = hipMemset(nn->v, 0, pb);
    if (e != hipSuccess) {
        vc1me_nn_destroy(nn);
        return fail(VC1ME_ENOMEM, std::string("network allocation: ") + hipGetErrorString(e));
    }
    *out = nn;
    return VC1ME_OK;
}

void vc1me_nn_destroy(vc1me_nn* nn) {
    if (!nn) return;
    if (nn->ctx) {
        std::lock_guard<std::mutex> lock(nn->ctx->mu);
        (void)hipStreamSynchronize(nn->ctx->stream);
    }
    auto rel = [](auto*& q) { if (q) (void)hipFree(q); q = nullptr; };
    rel(nn->p); rel(nn->g); rel(nn->m); rel(nn->v);
    for (int l = 0; l < pnet::kLayers; ++l) { rel(nn->wt[l]); rel(nn->wtt[l]); }
    NnBuffers& b = nn->b;
    rel(b.x0); for (auto& a : b.act) rel(a); rel(b.dz[0]); rel(b.dz[1]); rel(b.pooled); rel(b.out); rel(b.dout);
    rel(b.tgt); rel(b.valid); rel(b.in32); rel(b.loss); rel(b.gxy);
    rel(nn->props_all);
    delete nn;
}

int vc1me_nn_set_params(vc1me_nn* nn, const float* params) {
    if (!nn || !params) return fail(VC1ME_EINVAL, "null argument");
    std::lock_guard<std::mutex> lock(nn->ctx->mu);
    if (int rc = set_device(nn->ctx)) return rc;
    const size_t pb = sizeof(float) * pnet::kParams;
    HIP_TRY(hipMemcpy(nn->p, params, pb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(nn->m, 0, pb));
    HIP_TRY(hipMemset(nn->v, 0, pb));
    nn->step = 0;
    nn->packed = false;
    return VC1ME_OK;
}

int vc1me_nn_get_params(vc1me_nn* nn, float* params) {
    if (!nn || !params) return fail(VC1ME_EINVAL, "null argument");
    std::lock_guard<std::mutex> lock(nn->ctx->mu);
    if (int rc = set_device(nn->ctx)) return rc;
    HIP_TRY(hipStreamSynchronize(nn->ctx->stream));
    HIP_TRY(hipMemcpy(params, nn->p, sizeof(float) * pnet::kParams, hipMemcpyDeviceToHost));
    return VC1ME_OK;
}

int vc1me_nn_forward_frames(vc1me_nn* nn, const vc1me_frame* const* curs, const vc1me_frame* const* refs,
                            const int32_t* globals, int n, float* out) {
    if (!nn || !out) return fail(VC1ME_EINVAL, "null argument");
    std::lock_guard<std::mutex> lock(nn->ctx->mu);
    if (int rc = set_device(nn->ctx)) return rc;
    int h4, w4, mbh, mbw;
    if (int rc = nn_inputs(nn, curs, refs, n, h4, w4, mbh, mbw, globals)) return rc;
    hipStream_t st = nn->ctx->stream;
    HIP_TRY(nn_forward(nn, n, h4, w4, st));
    HIP_TRY(hipMemcpyAsync(out, nn->b.out, sizeof(float) * 5 * n * mbh * mbw, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return VC1ME_OK;
}

int vc1me_nn_forward_backward(vc1me_nn* nn, const float* input, int n, int h4, int w4, const float* dout, float* out,
                              float* grads) {
    if (!nn || !input || !dout || !out || !grads || n <= 0 || h4 <= 0 || w4 <= 0)
        return fail(VC1ME_EINVAL, "bad argument");
    std::lock_guard<std::mutex> lock(nn->ctx->mu);
    if (int rc = set_device(nn->ctx)) return rc;
    hipStream_t st = nn->ctx->stream;
    const int mbh = (h4 + 3) / 4, mbw = (w4 + 3) / 4;
    const size_t ni = static_cast<size_t>(n) * h4 * w4 * 2, no = static_cast<size_t>(n) * mbh * mbw * 5;
    HIP_TRY(nn_reserve(nn, n, h4, w4));
    HIP_TRY(hipMemcpyAsync(nn->b.in32, input, ni * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(pnet::launch_to_bf16(nn->b.in32, nn->b.x0, static_cast<int>(ni), st));
    HIP_TRY(nn_forward(nn, n, h4, w4, st));
    HIP_TRY(hipMemcpyAsync(nn->b.dout, dout, no * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(nn_backward(nn, n, h4, w4, st));
    HIP_TRY(hipMemcpyAsync(out, nn->b.out, no * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(grads, nn->g, sizeof(float) * pnet::kParams, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return VC1ME_OK;
}

int vc1me_nn_train_step(vc1me_nn* nn, const vc1me_frame* const* curs, const vc1me_frame* const* refs,
                        const int32_t* globals, const int32_t* targets, const uint8_t* valid, int n, float lr,
                        float* loss) {
    if (!nn || !targets || !valid || !loss) return fail(VC1ME_EINVAL, "null argument");
    std::lock_guard<std::mutex> lock(nn->ctx->mu);
    if (int rc = set_device(nn->ctx)) return rc;
    int h4, w4, mbh, mbw;
    if (int rc = nn_inputs(nn, curs, refs, n, h4, w4, mbh, mbw, globals)) return rc;
    hipStream_t st = nn->ctx->stream;
    const int per = mbh * mbw, mbs = n * per;
    std::vector<float> t(static_cast<size_t>(mbs) * 2);
    int count = 0;
    for (int i = 0; i < mbs; ++i) {
        // quarter-pel -> residual from the pair's global vector in 64-px units
        const int gx = globals ? globals[2 * (i / per)] : 0, gy = globals ? globals[2 * (i / per) + 1] : 0;
        t[2 * i] = static_cast<float>(targets[2 * i] - 4 * gx) / 256.0f;
        t[2 * i + 1] = static_cast<float>(targets[2 * i + 1] - 4 * gy) / 256.0f;
        count += valid[i] ? 1 : 0;
    }
    HIP_TRY(hipMemcpyAsync(nn->b.tgt, t.data(), t.size() * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(nn->b.valid, valid, mbs, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(nn->b.loss, 0, 4, st));
    HIP_TRY(nn_forward(nn, n, h4, w4, st));
    HIP_TRY(pnet::launch_loss(nn->b.out, nn->b.tgt, nn->b.valid, nn->b.dout, nn->b.loss, mbs,
                                  count ? 1.0f / static_cast<float>(count) : 0.f, st));
    HIP_TRY(nn_backward(nn, n, h4, w4, st));
    ++nn->step;
    const float b1 = 0.9f, b2 = 0.999f;
    HIP_TRY(pnet::launch_adam(nn->p, nn->g, nn->m, nn->v, pnet::kParams, lr, b1, b2,
                                  1.0f - std::pow(b1, static_cast<float>(nn->step)),
                                  1.0f - std::pow(b2, static_cast<float>(nn->step)), 1e-8f, st));
    nn->packed = false;
    HIP_TRY(hipMemcpyAsync(loss, nn->b.loss, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return VC1ME_OK;
}

int vc1me_analyze_step_proposals(vc1me_ctx* ctx, const vc1me_job* jobs, int n, const float* const* proposals,
                                 float threshold) {
    if (!ctx || !proposals) return fail(VC1ME_EINVAL, "null argument");
    if (!(threshold > 0.f && threshold < 1.f)) return fail(VC1ME_EINVAL, "threshold must lie in (0, 1)");
    if (n <= 0) return VC1ME_OK;
    if (!jobs || n > 255) return fail(VC1ME_EINVAL, "bad job list");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    size_t total = 0;
    std::vector<size_t> off(n);
    for (int i = 0; i < n; ++i) {
        if (!jobs[i].params || !proposals[i]) return fail(VC1ME_EINVAL, "null job field");
        off[i] = total;
        total += static_cast<size_t>((jobs[i].params->width + 15) / 16) * ((jobs[i].params->height + 15) / 16) * 5;
    }
    if (ctx->props_cap < total) {
        if (ctx->d_props) (void)hipFree(ctx->d_props);
        ctx->d_props = nullptr;
        ctx->props_cap = 0;
        HIP_TRY(hipMalloc(&ctx->d_props, total * sizeof(float)));
        ctx->props_cap = total;
    }
    std::vector<const float*> dp(n);
    for (int i = 0; i < n; ++i) {
        const size_t cnt = (i + 1 < n ? off[i + 1] : total) - off[i];
        HIP_TRY(hipMemcpyAsync(ctx->d_props + off[i], proposals[i], cnt * sizeof(float), hipMemcpyHostToDevice,
                               ctx->stream));
        dp[i] = ctx->d_props + off[i];
    }
    const float logit = std::log(threshold / (1.0f - threshold));
    return run_jobs(ctx, jobs, n, dp.data(), logit);
}

int vc1me_analyze_step_learned(vc1me_ctx* ctx, vc1me_nn* nn, const vc1me_job* jobs, int n, float threshold) {
    if (!ctx || !nn || nn->ctx != ctx) return fail(VC1ME_EINVAL, "network and context do not match");
    if (!(threshold > 0.f && threshold < 1.f)) return fail(VC1ME_EINVAL, "threshold must lie in (0, 1)");
    if (n <= 0) return VC1ME_OK;
    if (!jobs || n > 255) return fail(VC1ME_EINVAL, "bad job list");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (int rc = set_device(ctx)) return rc;
    for (int i = 0; i < n; ++i)
        if (!jobs[i].cur || !jobs[i].ref || !jobs[i].params) return fail(VC1ME_EINVAL, "null job field");
    // the jobs by raster: one batched forward pass per group, whose outputs
    // land contiguously in the step's proposal buffer
    struct Group { std::vector<int> jobs; int h4, w4, mbh, mbw; size_t at; };
    std::vector<Group> groups;
    {
        std::vector<bool> done(n, false);
        for (int i = 0; i < n; ++i) {
            if (done[i]) continue;
            Group g{};
            for (int k = i; k < n; ++k)
                if (!done[k] && jobs[k].cur->w == jobs[i].cur->w && jobs[k].cur->h == jobs[i].cur->h) {
                    g.jobs.push_back(k);
                    done[k] = true;
                }
            nn_dims(jobs[i].cur->w, jobs[i].cur->h, g.h4, g.w4, g.mbh, g.mbw);
            groups.push_back(std::move(g));
        }
    }
    size_t total = 0;
    for (auto& g : groups) {
        g.at = total;
        total += g.jobs.size() * static_cast<size_t>(g.mbh) * g.mbw * 5;
    }
    if (nn->props_cap < total) {
        if (nn->props_all) (void)hipFree(nn->props_all);
        nn->props_all = nullptr;
        nn->props_cap = 0;
        HIP_TRY(hipMalloc(&nn->props_all, total * sizeof(float)));
        nn->props_cap = total;
    }
    nn->props.assign(n, nullptr);
    for (const auto& g : groups)
        for (size_t j = 0; j < g.jobs.size(); ++j)
            nn->props[g.jobs[j]] = nn->props_all + g.at + j * static_cast<size_t>(g.mbh) * g.mbw * 5;
    // everything the forward passes use is allocated and packed before the
    // step, so that the step (forward included) can be captured as a graph
    for (const auto& g : groups) HIP_TRY(nn_reserve(nn, static_cast<int>(g.jobs.size()), g.h4, g.w4));
    HIP_TRY(nn_pack(nn, ctx->stream));
    // (the inputs are read through the step's job table: frames and global
    // vectors, as every other kernel of the step)
    auto forward = [&](hipStream_t st) -> int {
        for (const auto& g : groups) {
            pnet::JobIndex idx{};
            for (size_t j = 0; j < g.jobs.size(); ++j) idx.k[j] = static_cast<uint8_t>(g.jobs[j]);
            const int gn = static_cast<int>(g.jobs.size());
            HIP_TRY(pnet::launch_downsample_jobs(ctx->d_jobs, idx, gn, g.h4, g.w4, nn->b.x0, st));
            HIP_TRY(nn_forward(nn, gn, g.h4, g.w4, st, nn->props_all + g.at));
        }
        return VC1ME_OK;
    };
    // the graph of such a step also depends on the network's buffers
    std::string tag = "nn";
    for (const void* q : {static_cast<const void*>(nn->b.x0), static_cast<const void*>(nn->b.act[0]),
                          static_cast<const void*>(nn->b.act[1]), static_cast<const void*>(nn->b.act[2]),
                          static_cast<const void*>(nn->b.act[3]), static_cast<const void*>(nn->b.pooled),
                          static_cast<const void*>(nn->props_all), static_cast<const void*>(nn->p),
                          static_cast<const void*>(nn->wt[0]), static_cast<const void*>(nn->wt[1]),
                          static_cast<const void*>(nn->wt[2]), static_cast<const void*>(nn->wt[3])})
        tag += "," + std::to_string(reinterpret_cast<uintptr_t>(q));
    const float logit = std::log(threshold / (1.0f - threshold));
    return run_jobs(ctx, jobs, n, nn->props.data(), logit, forward, nullptr, false, ctx->learned_candidates != 0, tag);
}

}  // extern "C"
