// vc1me_kernels.hip — gfx950 kernels of the libvc1 motion-estimation backend.
//
// Frame level:
//   k_frame_sads / k_global_*   global translation search
//                               (Vc1Encoder::global_motion_search, src/encoder_motion.cpp:2025-2141)
//   k_signatures                long-range content index
//                               (build_long_range_index, src/encoder_motion.cpp:1615-1625)
// Macroblock level:
//   k_mb_diagonal               one 64-lane wave per macroblock; macroblocks of one
//                               anti-diagonal t = mx + 2*my run concurrently, so the
//                               left / top / top-left / top-right dependencies of
//                               analyze_p_picture_core (src/encoder_motion.cpp:113-428)
//                               are always final.  Candidates of every search stage are
//                               scored lane-parallel from a reference window staged in
//                               LDS; stage results merge through wave-wide minimum
//                               reductions of packed (cost, |mv|, y, x) keys, which is the
//                               reference's strict total order.
// Batch scorer:
//   k_score_batch               VulkanComputeContext::score_motion_batch semantics
//                               (src/shaders/vulkan_motion_cost.comp).
//
// Results are bit-exact with the reference CPU path: every cost is an exact
// integer, the x87 long-double finalist scores are reproduced by exact 128-bit
// rounding to a 64-bit significand, and the double decisions are evaluated in
// IEEE double without contraction.
#include "vc1me_internal.h"

// lambdas of device code are inlined like the functions around them (a lambda
// left out of line passes its captures through scratch memory)
#define VC1ME_INLINE __attribute__((always_inline))

namespace vc1me {

namespace {

constexpr uint64_t kNone = ~0ull;

// ------------------------------------------------------------------ small helpers

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ int iabs(int v) { return v < 0 ? -v : v; }
__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int floor_div4(int v) { return v >> 2; }  // arithmetic shift == floor division by 4
__device__ __forceinline__ int arshift(int v, int n) { return v >> n; }  // floor shift (src/encoder_entropy.cpp:377-381)

// DPP cross-lane steps (gfx9 DPP controls): quad_perm, row mirrors and the
// row broadcasts.  update_dpp(old, src, ...) leaves lanes outside row_mask at
// `old`.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ uint32_t dpp_mov(uint32_t old, uint32_t src) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(old), static_cast<int>(src), CTRL,
                                                             ROWMASK, 0xf, false));
}
constexpr int kDppXor1 = 0xB1;        // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141; // row_half_mirror
constexpr int kDppMirror = 0x140;     // row_mirror
constexpr int kDppBcast15 = 0x142;    // row_bcast:15
constexpr int kDppBcast31 = 0x143;    // row_bcast:31
constexpr int kDppShr1 = 0x111;       // row_shr:1 (lane i takes lane i - 1 of its row)
constexpr int kDppShr2 = 0x112;
constexpr int kDppShr4 = 0x114;
constexpr int kDppShr8 = 0x118;

template <int CTRL, int ROWMASK>
__device__ __forceinline__ uint64_t min64_step(uint64_t v) {
    const uint32_t lo = static_cast<uint32_t>(v), hi = static_cast<uint32_t>(v >> 32);
    const uint32_t plo = dpp_mov<CTRL, ROWMASK>(lo, lo), phi = dpp_mov<CTRL, ROWMASK>(hi, hi);
    const uint64_t p = (static_cast<uint64_t>(phi) << 32) | plo;
    return p < v ? p : v;
}
// min(v, the value of lane i - k of the row), where that lane exists
template <int CTRL>
__device__ __forceinline__ uint64_t min64_shr(uint64_t v) {
    const uint32_t lo = static_cast<uint32_t>(v), hi = static_cast<uint32_t>(v >> 32);
    const uint64_t p = (static_cast<uint64_t>(dpp_mov<CTRL, 0xf>(~0u, hi)) << 32) | dpp_mov<CTRL, 0xf>(~0u, lo);
    return p < v ? p : v;
}
// Inclusive prefix minimum within each 16-lane row (every lane active).
__device__ __forceinline__ uint64_t row_prefix_min64(uint64_t v) {
    v = min64_shr<kDppShr1>(v);
    v = min64_shr<kDppShr2>(v);
    v = min64_shr<kDppShr4>(v);
    return min64_shr<kDppShr8>(v);
}
// The value of lane i - 1 of the row (~0 in the row's first lane).
__device__ __forceinline__ uint64_t row_prev64(uint64_t v) {
    const uint32_t lo = static_cast<uint32_t>(v), hi = static_cast<uint32_t>(v >> 32);
    return (static_cast<uint64_t>(dpp_mov<kDppShr1, 0xf>(~0u, hi)) << 32) | dpp_mov<kDppShr1, 0xf>(~0u, lo);
}
// Minimum over the wave, returned uniformly (read from lane 63).
__device__ __forceinline__ uint64_t wave_min64(uint64_t v) {
    v = min64_step<kDppXor1, 0xf>(v);
    v = min64_step<kDppXor2, 0xf>(v);
    v = min64_step<kDppHalfMirror, 0xf>(v);
    v = min64_step<kDppMirror, 0xf>(v);
    v = min64_step<kDppBcast15, 0xa>(v);
    v = min64_step<kDppBcast31, 0xc>(v);
    const uint32_t lo = __builtin_amdgcn_readlane(static_cast<int>(v), 63);
    const uint32_t hi = __builtin_amdgcn_readlane(static_cast<int>(v >> 32), 63);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}
// Sum over the wave, returned uniformly.
__device__ __forceinline__ uint32_t wave_sum32(uint32_t v) {
    v += dpp_mov<kDppXor1, 0xf>(0, v);
    v += dpp_mov<kDppXor2, 0xf>(0, v);
    v += dpp_mov<kDppHalfMirror, 0xf>(0, v);
    v += dpp_mov<kDppMirror, 0xf>(0, v);
    v += dpp_mov<kDppBcast15, 0xa>(0, v);
    v += dpp_mov<kDppBcast31, 0xc>(0, v);
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}
// Sum over aligned groups of G lanes (G = 1, 2, 4, 8, 16); every lane of a group gets it.
__device__ __forceinline__ uint32_t group_sum32(uint32_t v, int G) {
    if (G >= 2) v += dpp_mov<kDppXor1, 0xf>(0, v);
    if (G >= 4) v += dpp_mov<kDppXor2, 0xf>(0, v);
    if (G >= 8) v += dpp_mov<kDppHalfMirror, 0xf>(0, v);
    if (G >= 16) v += dpp_mov<kDppMirror, 0xf>(0, v);
    return v;
}
// Value of lane q of each quad, in every lane of the quad.
template <int Q>
__device__ __forceinline__ int quad_bcast(int v) {
    return static_cast<int>(dpp_mov<Q | (Q << 2) | (Q << 4) | (Q << 6), 0xf>(0, static_cast<uint32_t>(v)));
}
__device__ __forceinline__ int lane_prefix(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0));
}

// Values of another lane: readlane (uniform lane index) and shuffles.
__device__ __forceinline__ uint64_t lane_read64(uint64_t v, int l) {
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), l));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v >> 32), l));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ __forceinline__ unsigned __int128 lane_read128(unsigned __int128 v, int l) {
    const uint64_t lo = lane_read64(static_cast<uint64_t>(v), l);
    const uint64_t hi = lane_read64(static_cast<uint64_t>(v >> 64), l);
    return (static_cast<unsigned __int128>(hi) << 64) | lo;
}
__device__ __forceinline__ uint64_t lane_shfl64(uint64_t v, int src) {
    const uint32_t lo = static_cast<uint32_t>(__shfl(static_cast<int>(v), src));
    const uint32_t hi = static_cast<uint32_t>(__shfl(static_cast<int>(v >> 32), src));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// An entry of a winner search, ordered lexicographically by (cls, s, a, b).
struct WinKey {
    uint32_t cls;
    unsigned __int128 s;
    uint64_t a, b;
};
template <int CTRL>
__device__ __forceinline__ uint64_t dpp64(uint64_t v) {
    const uint32_t lo = static_cast<uint32_t>(v), hi = static_cast<uint32_t>(v >> 32);
    return (static_cast<uint64_t>(dpp_mov<CTRL, 0xf>(hi, hi)) << 32) | dpp_mov<CTRL, 0xf>(lo, lo);
}
template <int CTRL>
__device__ __forceinline__ void win_step(WinKey& k) {
    WinKey p;
    p.cls = dpp_mov<CTRL, 0xf>(k.cls, k.cls);
    p.s = (static_cast<unsigned __int128>(dpp64<CTRL>(static_cast<uint64_t>(k.s >> 64))) << 64) |
          dpp64<CTRL>(static_cast<uint64_t>(k.s));
    p.a = dpp64<CTRL>(k.a);
    p.b = dpp64<CTRL>(k.b);
    bool less;
    if (p.cls != k.cls) less = p.cls < k.cls;
    else if (p.s != k.s) less = p.s < k.s;
    else if (p.a != k.a) less = p.a < k.a;
    else less = p.b < k.b;
    if (less) k = p;
}
// One tournament step on the leading fields (cls, s) only.
template <int CTRL>
__device__ __forceinline__ void win_step_cs(uint32_t& cls, unsigned __int128& s) {
    const uint32_t pc = dpp_mov<CTRL, 0xf>(cls, cls);
    const unsigned __int128 ps = (static_cast<unsigned __int128>(dpp64<CTRL>(static_cast<uint64_t>(s >> 64))) << 64) |
                                 dpp64<CTRL>(static_cast<uint64_t>(s));
    const bool less = pc != cls ? pc < cls : ps < s;
    if (less) { cls = pc; s = ps; }
}
// The least entry of lanes 0..7 (every lane active), in lane 0: a tournament
// over lane pairs, quads and the two halves of the eight on (cls, s); the
// entry's other fields come from the one lane holding that (cls, s), and
// only a tie on it runs the full tournament among the tying lanes.
__device__ __forceinline__ WinKey lanes8_min(WinKey k) {
    uint32_t cls = k.cls;
    unsigned __int128 s = k.s;
    win_step_cs<kDppXor1>(cls, s);
    win_step_cs<kDppXor2>(cls, s);
    win_step_cs<kDppHalfMirror>(cls, s);
    const int lane = static_cast<int>(threadIdx.x & (kWave - 1));
    const uint64_t tied = __ballot(lane < 8 && k.cls == cls && k.s == s) & 0xffull;
    if (__popcll(tied) == 1) {
        const int l = __ffsll(static_cast<long long>(tied)) - 1;
        return WinKey{cls, s, lane_read64(k.a, l), lane_read64(k.b, l)};
    }
    WinKey t = ((tied >> lane) & 1ull) ? k : WinKey{~0u, ~static_cast<unsigned __int128>(0), ~0ull, ~0ull};
    win_step<kDppXor1>(t);
    win_step<kDppXor2>(t);
    win_step<kDppHalfMirror>(t);
    return t;
}

// Global-memory view of a plane pointer (lets loads be scheduled past LDS stores).
typedef const __attribute__((address_space(1))) uint8_t* gptr_u8;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ gptr_u8 as_global(const uint8_t* p) { return (gptr_u8)(p); }

// llround without library dependence: round half away from zero; x - trunc(x) is exact.
__device__ __forceinline__ long long llround_exact(double x) {
    const double t = __builtin_trunc(x);
    const double f = x - t;
    long long r = static_cast<long long>(t);
    if (f >= 0.5) r += 1;
    else if (f <= -0.5) r -= 1;
    return r;
}

// Packed keys of the strict total order (cost, |dx|+|dy|, dy, dx) used by every
// integer search (src/encoder_motion.cpp:1788-1794).  Costs are < 2^20; the
// hidden-padding veto (2^62-1) maps to the largest 21-bit value.
__device__ __forceinline__ uint64_t cost21(uint64_t c) { return c >= (1u << 21) - 1 ? (1u << 21) - 1 : c; }
__device__ __forceinline__ uint64_t key_int(uint64_t cost, int dx, int dy) {
    return (cost21(cost) << 34) | (static_cast<uint64_t>(iabs(dx) + iabs(dy)) << 22) |
           (static_cast<uint64_t>(dy + 1024) << 11) | static_cast<uint64_t>(dx + 1024);
}
__device__ __forceinline__ int key_dx(uint64_t k) { return static_cast<int>(k & 2047) - 1024; }
__device__ __forceinline__ int key_dy(uint64_t k) { return static_cast<int>((k >> 11) & 2047) - 1024; }
__device__ __forceinline__ uint64_t key_cost(uint64_t k) {
    const uint64_t c = k >> 34;
    return c == (1u << 21) - 1 ? kBadSad : c;
}
// (cost, |dx|+|dy|, sequence) for the first-found-wins loops of the long-range searches
__device__ __forceinline__ uint64_t key_seq(uint64_t cost, int man, int seq) {
    return (cost21(cost) << 34) | (static_cast<uint64_t>(man) << 16) | static_cast<uint64_t>(seq);
}

// quarter-pel tie order (|x|+|y|, y, x) of the sub-pel stages (src/encoder_motion.cpp:3320-3325)
__device__ __forceinline__ bool tie_less(int ax, int ay, int bx, int by) {
    const int am = iabs(ax) + iabs(ay), bm = iabs(bx) + iabs(by);
    if (am != bm) return am < bm;
    if (ay != by) return ay < by;
    return ax < bx;
}

__device__ __forceinline__ bool mode_halfpel(int m) { return m == 1 || m == 2; }
__device__ __forceinline__ bool mode_bilinear(int m) { return m == 2; }

// MVDATA code lengths of MV table 0 (ST 421; src/vc1_groupa_tables.h:26-27 row 0),
// held as sixteen 4-bit entries per 64-bit word so that a lookup is register
// arithmetic.
constexpr uint8_t kMvBits0[73] = {
    6, 7, 7, 8, 14, 6, 5, 6, 7, 14, 14, 6, 6, 6, 8, 9, 10, 9, 7, 7, 7, 14, 14, 10, 9,
    10, 8, 10, 10, 14, 14, 14, 13, 13, 13, 6, 3, 5, 6, 8, 9, 13, 5, 4, 4, 5, 7, 9, 6, 5,
    5, 5, 6, 9, 8, 5, 5, 5, 7, 10, 7, 7, 7, 7, 7, 8, 10, 9, 8, 9, 9, 9, 3};
constexpr uint64_t mv_bits_word(int w) {
    uint64_t r = 0;
    for (int i = 0; i < 16; ++i)
        if (w * 16 + i < 73) r |= static_cast<uint64_t>(kMvBits0[w * 16 + i]) << (4 * i);
    return r;
}
__device__ __forceinline__ uint32_t mv_bits0(int sym) {
    constexpr uint64_t w0 = mv_bits_word(0), w1 = mv_bits_word(1), w2 = mv_bits_word(2), w3 = mv_bits_word(3),
                       w4 = mv_bits_word(4);
    const uint64_t w = sym < 32 ? (sym < 16 ? w0 : w1) : (sym < 48 ? w2 : (sym < 64 ? w3 : w4));
    return static_cast<uint32_t>((w >> (4 * (sym & 15))) & 15u);
}

// ------------------------------------------------------------------ hidden padding
// motion_uses_hidden_luma / _chroma (src/encoder_motion.cpp:2143-2189)

struct Geo {
    int W, H, vw, vh, mbw, mbh;
};

__device__ __forceinline__ bool hidden_luma(const Geo& g, int x0, int y0, int bw, int bh, int xq, int yq, int mode) {
    const bool rp = g.vw < g.mbw * 16, bp = g.vh < g.mbh * 16;
    if (!rp && !bp) return false;
    const int aw = imax(0, imin(bw, g.vw - x0)), ah = imax(0, imin(bh, g.vh - y0));
    if (!aw || !ah) return false;
    const int ix = floor_div4(xq), iy = floor_div4(yq);
    const int fx = xq - ix * 4, fy = yq - iy * 4;
    const int right = fx ? (mode_bilinear(mode) ? 1 : 2) : 0;
    const int bottom = fy ? (mode_bilinear(mode) ? 1 : 2) : 0;
    return (rp && x0 + aw - 1 + ix + right >= g.vw) || (bp && y0 + ah - 1 + iy + bottom >= g.vh);
}
__device__ __forceinline__ bool hidden_chroma(const Geo& g, int x0, int y0, int bw, int bh, int xq, int yq) {
    const int cw = (g.vw + 1) / 2, ch = (g.vh + 1) / 2;
    const bool rp = cw < g.mbw * 8, bp = ch < g.mbh * 8;
    if (!rp && !bp) return false;
    const int cx = x0 / 2, cy = y0 / 2;
    const int aw = imax(0, imin(bw / 2, cw - cx)), ah = imax(0, imin(bh / 2, ch - cy));
    if (!aw || !ah) return false;
    int uvx = arshift(xq + ((xq & 3) == 3), 1), uvy = arshift(yq + ((yq & 3) == 3), 1);
    uvx = uvx < 0 ? uvx + (uvx & 1) : uvx - (uvx & 1);
    uvy = uvy < 0 ? uvy + (uvy & 1) : uvy - (uvy & 1);
    const int ix = floor_div4(uvx), iy = floor_div4(uvy);
    const bool fx = (uvx - ix * 4) != 0, fy = (uvy - iy * 4) != 0;
    return (rp && cx + aw - 1 + ix + (fx ? 1 : 0) >= cw) || (bp && cy + ah - 1 + iy + (fy ? 1 : 0) >= ch);
}
__device__ __forceinline__ bool hidden_padding(const Geo& g, int x0, int y0, int bw, int bh, int xq, int yq,
                                               int mode, bool chroma) {
    // Early out: neither test can hold when the block, moved by the integer
    // part of the vector plus the filter's right and bottom reach (at most 2),
    // ends before the visible width and height; the chroma block's reach in
    // chroma samples, cx + 8 + floor(uvx / 4), stays below ceil(vw / 2) under
    // the same bound (floor(uvx / 4) <= (floor(xq / 4) + 1) / 2).
    if (x0 + bw + floor_div4(xq) + 1 < g.vw && y0 + bh + floor_div4(yq) + 1 < g.vh) return false;
    return hidden_luma(g, x0, y0, bw, bh, xq, yq, mode) || (chroma && hidden_chroma(g, x0, y0, bw, bh, xq, yq));
}
__device__ __forceinline__ bool chroma_block(int x0, int y0, int bw, int bh) {
    return bw == 16 && bh == 16 && !(x0 & 15) && !(y0 & 15);
}

// ------------------------------------------------------------------ x87 score emulation
// score = (long double)cost + (long double)lambda * rate, rounded once to a
// 64-bit significand (round to nearest even).  lambda = m * 2^-k exactly, so
// the exact sum is N * 2^-k with N = cost*2^k + m*rate; all scores of one MB
// share k and compare as rounded N.
__device__ __forceinline__ unsigned __int128 x87_round(unsigned __int128 n) {
    const uint64_t hi = static_cast<uint64_t>(n >> 64);
    if (!hi) return n;
    const int bits = 128 - __builtin_clzll(hi);  // bit length of n
    const int sh = bits - 64;
    unsigned __int128 q = n >> sh;
    const unsigned __int128 rem = n - (q << sh);
    const unsigned __int128 half = static_cast<unsigned __int128>(1) << (sh - 1);
    if (rem > half || (rem == half && (q & 1))) q += 1;
    return q << sh;  // q may have become 2^64: still the correctly rounded value
}
__device__ __forceinline__ unsigned __int128 x87_score(uint64_t cost, uint64_t rate, uint64_t m, int k) {
    return x87_round((static_cast<unsigned __int128>(cost) << k) + static_cast<unsigned __int128>(m) * rate);
}

// ------------------------------------------------------------------ LDS layout

// dedupe bitset of the integer searches: one bit per point of [-R, R]^2
__host__ __device__ inline int seen_words(int local_range) {
    return ((2 * local_range + 1) * (2 * local_range + 1) + 31) / 32;
}
constexpr int kSmallWin = 80 * 96;   // windows of the far (long-range) candidates

// Phase window: the interpolated reference (every phase) around a macroblock's
// integer vector, staged in LDS for the sub-pel stages.  kPwSide samples
// square per phase; rows hold kPwPitch bytes so that 4-byte aligned copies
// of the phase planes (plus the 8-byte read of the block scorer) fit whatever
// the alignment of the window origin.
// The window starts kPwMargin samples before the integer vector and covers
// kPwSide = 21 samples: the 16x16 sub-pel candidates (within 3 quarter-pels of
// the vector) and every 4-MV seed (within 2 pels of the 1-MV vector) stay
// inside it; a candidate outside reads the phase planes.
constexpr int kPwMargin = 3;
constexpr int kPwSide = 21;
constexpr int kPwPitch = 28;
static_assert(kPwSide + 3 <= kPwPitch && ((kPwSide - 4 + 3) & ~3) + 8 <= kPwPitch, "phase-window rows too short");
constexpr int kPwBytes = (16 * kPwSide * kPwPitch + 255) & ~255;   // whole 256-byte copy blocks

__host__ __device__ inline int win_capacity(int local_range) {
    const int side = 16 + 2 * local_range;
    const int pitch = (side + 3 + 8 + 3) & ~3;     // 4-byte alignment of the origin plus read slack
    const int need = (side + 1) * pitch;
    return need > kSmallWin ? need : kSmallWin;
}

struct Lds {
    uint8_t* cur;        // 16x16 current block, zero outside the active area
    uint8_t* win;        // reference window (4-byte aligned rows)
    uint8_t* pw;         // phase window
    uint32_t* seen_s;    // sampled-SAD dedupe bitset of the UMH search
    uint32_t* seen_f;    // full-SAD dedupe bitset
    int2* own;           // current MB's four block vectors
    uint32_t* nb;        // neighbour snapshot: kNbWords dwords each of left, top, top-left, top-right
    uint32_t* cs;        // sampled current rows: stride 2 (8 rows x 2 dwords), then stride 4 (4 rows x 1 dword)
    int* sd;             // a short seed list read by lane index (8 ints)
};

// Neighbour snapshot per neighbour: 1-MV vector (2 dwords), the four block
// vectors (8), flags (1), propagated flag (1).
constexpr int kNbWords = 12;
enum { NB_LEFT = 0, NB_TOP = 1, NB_TOPLEFT = 2, NB_TOPRIGHT = 3 };

// The integer searches' window and dedupe bitsets are dead once the sub-pel
// stages start, so the phase window shares their bytes; so does the extended
// search's staging (kXwRoundBytes, defined with it below).
__host__ __device__ inline int int_scratch_bytes(int local_range) {
    const int off = (win_capacity(local_range) + 15) & ~15;  // win
    return off + seen_words(local_range) * 4 * 2;             // seen_s, seen_f
}

__host__ __device__ inline int shared_bytes(int local_range);

__host__ __device__ inline int lds_bytes_for(int local_range) {
    // cur, shared region, own, neighbour snapshot, sampled rows, seed list, slack
    return 256 + shared_bytes(local_range) + 4 * 8 + 4 * 12 * 4 + 20 * 4 + 8 * 4 + 64;
}

__device__ __forceinline__ Lds carve(uint8_t* smem, int local_range) {
    Lds L;
    L.cur = smem;
    uint8_t* const region = smem + 256;
    int off = 0;
    L.win = region + off; off += (win_capacity(local_range) + 15) & ~15;
    L.seen_s = reinterpret_cast<uint32_t*>(region + off); off += seen_words(local_range) * 4;
    L.seen_f = reinterpret_cast<uint32_t*>(region + off);
    L.pw = region;
    const int shared = shared_bytes(local_range);
    L.own = reinterpret_cast<int2*>(region + shared);
    L.nb = reinterpret_cast<uint32_t*>(region + shared + 4 * 8);
    L.cs = reinterpret_cast<uint32_t*>(region + shared + 4 * 8 + 4 * 12 * 4);
    L.sd = reinterpret_cast<int*>(region + shared + 4 * 8 + 4 * 12 * 4 + 20 * 4);
    return L;
}

// ------------------------------------------------------------------ per-MB context

struct Ctx {
    const JobDesc* J;
    Geo g;
    const uint8_t* ry;
    int mx, my, x0, y0, aw, ah, pos, lane;
    Lds L;
    int winCap;
    // current window
    int wx0, wy0, ww, wh, wp;
    // phase window: plane coordinates of its sample (0,0), the byte offset of
    // that sample in each LDS row, and which filter's phases it holds
    int pwx, pwy, pwsh;
    bool pw_ok, pw_bil;
    // algorithmic operation counts of this macroblock: uniform totals, plus
    // per-lane partial counts of the lane-parallel block scoring (summed
    // over the wave when the macroblock is done)
    uint64_t n_sad, n_taps, n_satd, n_sig;
    uint32_t ln_sad, ln_taps, ln_satd;
};

__device__ __forceinline__ uint64_t taps_per_sample(int hm, int vm, bool bil) {
    if (!hm && !vm) return 0;
    if (bil) return 1;
    return (hm && vm) ? 5 : 1;
}

__device__ __forceinline__ void wsync() { __syncthreads(); }

__device__ __forceinline__ void set_err(Ctx& C, int code) {
    if (C.lane == 0) atomicMax(C.J->err, code);
}

// Stage a clamped reference window in LDS (sample_clamped semantics,
// src/encoder_motion_compensation.cpp:350-353) covering frame columns
// [wx0, wx0+ww) and rows [wy0, wy0+wh).  The stored origin is rounded down to
// 16 bytes so whole rows move as 16-byte loads.  Returns false if the window
// does not fit the buffer.
//
// win_issue only starts the direct copy (a window inside the picture's
// columns) and returns whether it did; the next wsync waits for it.  An
// analysis issues its integer-search window that way before it loads the
// neighbours' state, so that the two loads' latencies overlap.
__device__ __forceinline__ bool win_issue(Ctx& C, int wx0, int wy0, int ww, int wh) {
    const int ax0 = wx0 & ~3;
    const int aw = wx0 + ww - ax0;
    const int wp = (aw + 8 + 3) & ~3;
    const int W = C.g.W, H = C.g.H;
    if (ww <= 0 || wh <= 0 || wp * wh > C.winCap || (W & 3) != 0 || ax0 < 0 || ax0 + wp > W) return false;
    wsync();
    const int cpr = wp >> 2;
    const int total = cpr * wh;
    gptr_u8 ry = as_global(C.ry);
    uint32_t* w4 = reinterpret_cast<uint32_t*>(C.L.win);
    typedef const __attribute__((address_space(1))) void* gvoid;
    typedef __attribute__((address_space(3))) void* lvoid;
    const int dr = kWave / cpr, dc = kWave - dr * cpr;
    int r = C.lane / cpr, c = C.lane - (C.lane / cpr) * cpr;
    for (int base = 0; base < total; base += kWave) {
        const int sy = clampi(wy0 + r, 0, H - 1);
        // (rows and the picture width are below 2^24: a 24-bit product)
        const gptr_u8 src = ry + (__umul24(static_cast<uint32_t>(sy), static_cast<uint32_t>(W)) + static_cast<uint32_t>(ax0 + c * 4));
        if (C.lane < total - base) __builtin_amdgcn_global_load_lds((gvoid)src, (lvoid)(w4 + base), 4, 0, 0);
        c += dc;
        r += dr;
        if (c >= cpr) { c -= cpr; ++r; }
    }
    C.wx0 = ax0; C.wy0 = wy0; C.ww = aw; C.wh = wh; C.wp = wp;
    return true;
}
// whether the window staged last (issued or loaded) is exactly this one
__device__ __forceinline__ bool win_is(const Ctx& C, int wx0, int wy0, int ww, int wh) {
    const int ax0 = wx0 & ~3;
    return C.wx0 == ax0 && C.wy0 == wy0 && C.ww == wx0 + ww - ax0 && C.wh == wh;
}

__device__ __forceinline__ bool win_load(Ctx& C, int wx0, int wy0, int ww, int wh) {
    // rows start at the dword holding column wx0 and carry 8 bytes of read
    // slack for the row SAD; copied one dword per lane
    const int ax0 = wx0 & ~3;
    const int aw = wx0 + ww - ax0;
    const int wp = (aw + 8 + 3) & ~3;
    if (ww <= 0 || wh <= 0 || wp * wh > C.winCap) return false;
    wsync();
    const int W = C.g.W, H = C.g.H;
    const int cpr = wp >> 2;
    const int total = cpr * wh;
    const bool aligned_rows = (W & 3) == 0;
    gptr_u8 ry = as_global(C.ry);
    uint32_t* w4 = reinterpret_cast<uint32_t*>(C.L.win);
    if (aligned_rows && ax0 >= 0 && ax0 + wp <= W) {
        // every dword of every row (slack included) lies inside the plane's
        // columns: copy global -> LDS directly, 64 consecutive dwords of the
        // window per instruction (dword i lands at byte 4 i), all in flight
        // at once; rows clamp to the plane.  Lane l copies dword base + l,
        // whose row and column advance by (dr, dc) per instruction.
        typedef const __attribute__((address_space(1))) void* gvoid;
        typedef __attribute__((address_space(3))) void* lvoid;
        const int dr = kWave / cpr, dc = kWave - dr * cpr;
        int r = C.lane / cpr, c = C.lane - (C.lane / cpr) * cpr;
        for (int base = 0; base < total; base += kWave) {
            const int sy = clampi(wy0 + r, 0, H - 1);
            const gptr_u8 src = ry + (__umul24(static_cast<uint32_t>(sy), static_cast<uint32_t>(W)) + static_cast<uint32_t>(ax0 + c * 4));
            if (C.lane < total - base) __builtin_amdgcn_global_load_lds((gvoid)src, (lvoid)(w4 + base), 4, 0, 0);
            c += dc;
            r += dr;
            if (c >= cpr) { c -= cpr; ++r; }
        }
        C.wx0 = ax0; C.wy0 = wy0; C.ww = aw; C.wh = wh; C.wp = wp;
        wsync();
        return true;
    }
#pragma unroll 4
    for (int i = C.lane; i < total; i += kWave) {
        const int r = i / cpr, c = i - r * cpr;
        const int sy = clampi(wy0 + r, 0, H - 1);
        const int x = ax0 + c * 4;
        gptr_u8 row = ry + static_cast<size_t>(sy) * W;
        uint32_t v;
        if (aligned_rows && x >= 0 && x + 4 <= W) {
            v = *reinterpret_cast<const __attribute__((address_space(1))) uint32_t*>(row + x);
        } else {
            v = 0;
#pragma unroll
            for (int b = 0; b < 4; ++b) v |= static_cast<uint32_t>(row[clampi(x + b, 0, W - 1)]) << (8 * b);
        }
        w4[i] = v;
    }
    C.wx0 = ax0; C.wy0 = wy0; C.ww = aw; C.wh = wh; C.wp = wp;
    wsync();
    return true;
}

// ------------------------------------------------------------------ integer SAD from the window

// Partial SAD of sampled rows part, part+G, ... of the 16x16 block at integer
// offset (dx,dy).  Rows at or beyond the active height and columns at or
// beyond the active width contribute nothing (mb_sad, src/encoder_motion.cpp:2274-2282).
// The sampled current rows of the stride-2 and stride-4 SADs, packed once per
// analysis from the current block (zero outside the active area): dword
// 2 r + h of stride 2 holds row 2 r, columns 8 h + 0, 2, 4, 6; dword 16 + r of
// stride 4 holds row 4 r, columns 0, 4, 8, 12.
__device__ __forceinline__ void pack_sampled_rows(Ctx& C) {
    const int l = C.lane;
    if (l < 20) {
        const uint8_t* row = C.L.cur + (l < 16 ? (l >> 1) * 2 * 16 + (l & 1) * 8 : (l - 16) * 4 * 16);
        const int step = l < 16 ? 2 : 4;
        C.L.cs[l] = static_cast<uint32_t>(row[0]) | (static_cast<uint32_t>(row[step]) << 8) |
                    (static_cast<uint32_t>(row[2 * step]) << 16) | (static_cast<uint32_t>(row[3 * step]) << 24);
    }
}

// Byte mask of the packed samples at columns c0, c0 + step, c0 + 2 step,
// c0 + 3 step that lie inside the active width.
__device__ __forceinline__ uint32_t sample_mask(int aw, int c0, int step) {
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (c0 + k * step < aw) m |= 0xffu << (8 * k);
    return m;
}

// SAD of the block at window offset (dx, dy) over rows part, part + G, ...
// of the sampled rows (stride 1, 2 or 4 in both axes).  Stride 1 reads four
// window dwords per row against the current row; strides 2 and 4 pack the
// sampled window bytes of a row (alignbyte then v_perm, or v_perm straight
// from the window dwords) and compare them with the packed current rows.
// The current block is zero outside the active area, so only the window
// side is masked (and only when the block is not whole).
__device__ __forceinline__ uint32_t sad_int_part(const Ctx& C, int dx, int dy, int stride, int part, int G) {
    const int ox = C.x0 + dx - C.wx0, oy = C.y0 + dy - C.wy0;
    const int nrows = (C.ah + stride - 1) / stride;
    const int base = ox & ~3, sh = ox & 3;
    // (the point lies inside the window: its row and column offsets are
    // non-negative, and with the pitch below 2^24 a row's offset is a 24-bit
    // product)
    const uint32_t uwp = static_cast<uint32_t>(C.wp), ubase = static_cast<uint32_t>(base);
    const bool whole = C.aw == 16;
    uint32_t acc = 0;
    if (stride == 1) {
        uint32_t m[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) m[j] = sample_mask(C.aw, 4 * j, 1);
        for (int r = part; r < nrows; r += G) {
            const uint32_t* wr = reinterpret_cast<const uint32_t*>(C.L.win + (__umul24(static_cast<uint32_t>(oy + r), uwp) + ubase));
            const uint32_t* cr = reinterpret_cast<const uint32_t*>(C.L.cur + r * 16);
            const uint32_t d0 = wr[0], d1 = wr[1], d2 = wr[2], d3 = wr[3], d4 = wr[4];
            uint32_t r0 = __builtin_amdgcn_alignbyte(d1, d0, sh), r1 = __builtin_amdgcn_alignbyte(d2, d1, sh);
            uint32_t r2 = __builtin_amdgcn_alignbyte(d3, d2, sh), r3 = __builtin_amdgcn_alignbyte(d4, d3, sh);
            if (!whole) { r0 &= m[0]; r1 &= m[1]; r2 &= m[2]; r3 &= m[3]; }
            acc = __builtin_amdgcn_sad_u8(r0, cr[0], acc);
            acc = __builtin_amdgcn_sad_u8(r1, cr[1], acc);
            acc = __builtin_amdgcn_sad_u8(r2, cr[2], acc);
            acc = __builtin_amdgcn_sad_u8(r3, cr[3], acc);
        }
    } else if (stride == 2) {
        const uint32_t pm0 = sample_mask(C.aw, 0, 2), pm1 = sample_mask(C.aw, 8, 2);
        for (int r = part; r < nrows; r += G) {
            const uint32_t* wr = reinterpret_cast<const uint32_t*>(C.L.win + (__umul24(static_cast<uint32_t>(oy + 2 * r), uwp) + ubase));
            const uint32_t d0 = wr[0], d1 = wr[1], d2 = wr[2], d3 = wr[3], d4 = wr[4];
            const uint32_t r0 = __builtin_amdgcn_alignbyte(d1, d0, sh), r1 = __builtin_amdgcn_alignbyte(d2, d1, sh);
            const uint32_t r2 = __builtin_amdgcn_alignbyte(d3, d2, sh), r3 = __builtin_amdgcn_alignbyte(d4, d3, sh);
            // bytes 0 and 2 of each: columns 0, 2, 4, 6 and 8, 10, 12, 14
            const uint32_t e0 = __builtin_amdgcn_perm(r1, r0, 0x06040200u) & pm0;
            const uint32_t e1 = __builtin_amdgcn_perm(r3, r2, 0x06040200u) & pm1;
            acc = __builtin_amdgcn_sad_u8(e0, C.L.cs[2 * r], acc);
            acc = __builtin_amdgcn_sad_u8(e1, C.L.cs[2 * r + 1], acc);
        }
    } else {
        // byte sh of each of the four window dwords: columns 0, 4, 8, 12
        const uint32_t lo_sel = static_cast<uint32_t>(sh) | (static_cast<uint32_t>(sh + 4) << 8) | 0x0c0c0000u;
        const uint32_t hi_sel = 0x00000c0cu | (static_cast<uint32_t>(sh) << 16) | (static_cast<uint32_t>(sh + 4) << 24);
        const uint32_t pm = sample_mask(C.aw, 0, 4);
        for (int r = part; r < nrows; r += G) {
            const uint32_t* wr = reinterpret_cast<const uint32_t*>(C.L.win + (__umul24(static_cast<uint32_t>(oy + 4 * r), uwp) + ubase));
            const uint32_t d0 = wr[0], d1 = wr[1], d2 = wr[2], d3 = wr[3];
            const uint32_t e = (__builtin_amdgcn_perm(d1, d0, lo_sel) | __builtin_amdgcn_perm(d3, d2, hi_sel)) & pm;
            acc = __builtin_amdgcn_sad_u8(e, C.L.cs[16 + r], acc);
        }
    }
    return acc;
}

__device__ __forceinline__ bool in_window(const Ctx& C, int dx, int dy) {
    const int ox = C.x0 + dx - C.wx0, oy = C.y0 + dy - C.wy0;
    return ox >= 0 && oy >= 0 && ox + 16 <= C.ww && oy + 16 <= C.wh;
}

enum { KEY_TOTAL = 0, KEY_SEQ_MAN = 1, KEY_SEQ = 2 };

// The four least keys considered so far, ascending (kNone where fewer):
// wave-uniform values.
struct Top4 {
    uint64_t k0, k1, k2, k3;
};
__device__ __forceinline__ void top4_insert(Top4& t, uint64_t m) {
    if (m < t.k0) { t.k3 = t.k2; t.k2 = t.k1; t.k1 = t.k0; t.k0 = m; }
    else if (m < t.k1) { t.k3 = t.k2; t.k2 = t.k1; t.k1 = m; }
    else if (m < t.k2) { t.k3 = t.k2; t.k2 = m; }
    else if (m < t.k3) t.k3 = m;
}
// The wave's four least keys from every lane's own four least (keys are
// distinct across lanes: each scored point is one lane's), ascending: four
// rounds of a wave minimum, the lane holding it dropping its head.
__device__ __forceinline__ Top4 top4_wave(Top4 t) {
    Top4 w{kNone, kNone, kNone, kNone};
    for (int r = 0; r < 4; ++r) {
        const uint64_t m = wave_min64(t.k0);
        if (r == 0) w.k0 = m; else if (r == 1) w.k1 = m; else if (r == 2) w.k2 = m; else w.k3 = m;
        if (m == kNone) break;
        if (t.k0 == m) { t.k0 = t.k1; t.k1 = t.k2; t.k2 = t.k3; t.k3 = kNone; }
    }
    return w;
}

// Score n integer candidates (pt(i) -> (dx,dy), already clamped) at the given
// sample stride from the current window, G lanes per candidate.  seen: dedupe
// bitset or null.  top: optional per-lane four least keys of the points this
// lane scored, to which its keys of the batch are added (with the bitset,
// every key is new; top4_wave merges the lanes' lists).  Returns the minimum
// key of the batch (kNone if no candidate was scored).
template <class PtFn>
__device__ __forceinline__ uint64_t int_batch(Ctx& C, int n, int stride, uint32_t* seen, int keymode, int seq0, PtFn pt,
                              Top4* top = nullptr) {
    const int rows = (C.ah + stride - 1) / stride;
    int G = 1;
    while (G * 2 * n <= kWave && G * 2 <= rows && G < 16) G *= 2;
    const int per = kWave / G;
    const int part = C.lane & (G - 1);
    // A scored point's block lies inside the staged window; when the window
    // ends before the visible width and height, no point can read hidden
    // padding (the luma block's last column is at most the window's last,
    // E < vw, and the chroma block's at most ceil((E - 1) / 2) < ceil(vw / 2);
    // rows likewise), and the per-point test is skipped.
    const bool may_hide = C.ww <= 0 || C.wx0 + C.ww > C.g.vw || C.wy0 + C.wh > C.g.vh;
    uint64_t best = kNone;        // with `top`: the least key, merged per batch
    uint64_t lane_best = kNone;   // without: this lane's least key
    for (int base = 0; base < n; base += per) {
        const int i = base + C.lane / G;
        const bool act = i < n;
        int dx = 0, dy = 0;
        if (act) pt(i, dx, dy);
        int valid = act ? 1 : 0;
        if (seen) {
            if (act && part == 0) {
                const int R = C.J->local_range;   // every deduped point lies in [-R, R]^2
                const int bit = (dy + R) * (2 * R + 1) + (dx + R);
                const uint32_t old = atomicOr(&seen[bit >> 5], 1u << (bit & 31));
                valid = (old >> (bit & 31)) & 1 ? 0 : 1;
            }
            valid = __shfl(valid, C.lane - part);
        }
        uint32_t sum = 0;
        bool hid = false;
        if (valid) {
            hid = may_hide && hidden_padding(C.g, C.x0, C.y0, 16, 16, dx * 4, dy * 4, 0, true);
            if (!hid) {
                if (!in_window(C, dx, dy)) {
                    set_err(C, 2);
                    hid = true;
                } else {
                    sum = sad_int_part(C, dx, dy, stride, part, G);
                }
            }
        }
        sum = group_sum32(sum, G);
        {
            const uint64_t counted = __popcll(__ballot(valid && !hid && part == 0));
            const uint64_t npx = static_cast<uint64_t>((C.aw + stride - 1) / stride) * ((C.ah + stride - 1) / stride);
            C.n_sad += counted * npx;
        }
        uint64_t key = kNone;
        const bool leader = valid && part == 0;
        if (leader) {
            const uint64_t cost = hid ? kBadSad : static_cast<uint64_t>(sum) * static_cast<uint64_t>(stride * stride);
            if (keymode == KEY_TOTAL) key = key_int(cost, dx, dy);
            else if (keymode == KEY_SEQ_MAN) key = key_seq(cost, iabs(dx) + iabs(dy), seq0 + i);
            else key = key_seq(cost, 0, seq0 + i);
        }
        if (top) {
            const uint64_t mk = wave_min64(key);
            best = mk < best ? mk : best;
            top4_insert(*top, key);
        } else {
            lane_best = key < lane_best ? key : lane_best;
        }
    }
    return top ? best : wave_min64(lane_best);
}

// ------------------------------------------------------------------ geometry helpers

__device__ __forceinline__ uint64_t active_pixels(const Ctx& C) {
    return static_cast<uint64_t>(C.aw) * static_cast<uint64_t>(C.ah);
}
// motion_match_good, src/encoder_motion.cpp:1633-1640
__device__ __forceinline__ bool match_good(const Ctx& C, uint64_t sad) {
    const uint64_t px = active_pixels(C);
    if (!px || sad >= UINT64_MAX / 8) return false;
    return sad <= px * 6u;
}
// distant_match_decision, src/encoder_motion.cpp:1642-1656
__device__ __forceinline__ int distant_decision(const Ctx& C, uint64_t cand, uint64_t local) {
    const uint64_t px = active_pixels(C);
    if (!px || cand >= UINT64_MAX / 8) return 0;
    if (cand <= px * 6u) return 2;
    if (C.J->rej_always[px] || cand > C.J->rej_thr[px]) return 1;
    if (local >= UINT64_MAX / 8) return 3;
    return cand * 4u <= local * 3u ? 3 : 1;
}

// round_qpel_to_pixel, src/encoder_motion.cpp:1295-1297
__device__ __forceinline__ int round_qpel_to_pixel(int q) { return q >= 0 ? (q + 2) / 4 : -((-q + 2) / 4); }

struct Bounds {
    int xmin, xmax, ymin, ymax;
};
// motion_search_bounds_px / _qpel, src/encoder_motion.cpp:1238-1255
__device__ __forceinline__ Bounds bounds_px(int requested) {
    const int r = clampi(requested, 0, 1024);
    return {-imin(r, 1024), imin(r, 1023), -imin(r, 256), imin(r, 255)};
}
__device__ __forceinline__ Bounds bounds_qpel(int requested) {
    const int r = clampi(requested, 0, 1024);
    Bounds b;
    b.xmin = -imin(r * 4, 4096);
    b.xmax = (r >= 1024) ? 4095 : imin(r * 4, 4095);
    b.ymin = -imin(r * 4, 1024);
    b.ymax = (r >= 256) ? 1023 : imin(r * 4, 1023);
    return b;
}

// ------------------------------------------------------------------ predictors

struct Pred {
    int px, py, ax, ay, cx, cy;
    bool a_valid, c_valid, hybrid;
};

__device__ __forceinline__ int median3(int a, int b, int c) {
    return a + b + c - imin(a, imin(b, c)) - imax(a, imax(b, c));
}

__device__ __forceinline__ void pullback(int qx, int qy, int X, int Y, int margin, int& px, int& py) {
    if (qx + px < margin) px = margin - qx;
    if (qx + px > X) px = X - qx;
    if (qy + py < margin) py = margin - qy;
    if (qy + py > Y) py = Y - qy;
}

// The neighbours' state, read once per analysis into LDS (every predictor
// and the propagated search read this snapshot): lane l loads dword l % 12 of
// neighbour l / 12; neighbours outside the frame read as zero.
__device__ __forceinline__ void nb_load(Ctx& C) {
    const JobDesc* J = C.J;
    if (C.lane < 4 * kNbWords) {
        const int n = C.lane / kNbWords, w = C.lane - n * kNbWords;
        const int nx = C.mx + (n == NB_LEFT || n == NB_TOPLEFT ? -1 : n == NB_TOPRIGHT ? 1 : 0);
        const int ny = C.my + (n == NB_LEFT ? 0 : -1);
        uint32_t v = 0;
        if (nx >= 0 && nx < C.g.mbw && ny >= 0) {
            const int pos = ny * C.g.mbw + nx;
            if (w < 2) v = reinterpret_cast<const uint32_t*>(J->mvs)[2 * pos + w];
            else if (w < 10) v = reinterpret_cast<const uint32_t*>(J->bmv)[8 * pos + (w - 2)];
            else if (w == 10) v = J->flags[pos];
            else v = J->prop[pos];
        }
        C.L.nb[C.lane] = v;
    }
}
__device__ __forceinline__ int2 nb_mv(const Ctx& C, int n) {
    return make_int2(static_cast<int>(C.L.nb[n * kNbWords]), static_cast<int>(C.L.nb[n * kNbWords + 1]));
}
__device__ __forceinline__ int2 nb_bmv(const Ctx& C, int n, int k) {
    return make_int2(static_cast<int>(C.L.nb[n * kNbWords + 2 + 2 * k]), static_cast<int>(C.L.nb[n * kNbWords + 3 + 2 * k]));
}
__device__ __forceinline__ uint32_t nb_flags(const Ctx& C, int n) { return C.L.nb[n * kNbWords + 10]; }
__device__ __forceinline__ uint32_t nb_prop(const Ctx& C, int n) { return C.L.nb[n * kNbWords + 11]; }
// the snapshot slot of macroblock (mx + dx, my + dy): left, top, top-left or top-right
__device__ __forceinline__ int nb_slot(int dx, int dy) {
    return dy == 0 ? NB_LEFT : dx < 0 ? NB_TOPLEFT : dx > 0 ? NB_TOPRIGHT : NB_TOP;
}

// predictor_info, src/encoder_motion.cpp:1312-1352 (reads the final 1-MV field)
__device__ __forceinline__ Pred predictor_info(const Ctx& C) {
    const int mbw = C.g.mbw, mbh = C.g.mbh, mx = C.mx, my = C.my;
    const bool av = my > 0, cv = mx > 0, bv = my > 0 && mbw > 1;
    int2 A = {0, 0}, B = {0, 0}, Cc = {0, 0};
    if (av) A = nb_mv(C, NB_TOP);
    if (cv) Cc = nb_mv(C, NB_LEFT);
    if (bv) B = nb_mv(C, mx == mbw - 1 ? NB_TOPLEFT : NB_TOPRIGHT);
    int px = 0, py = 0;
    if (av) { px = A.x; py = A.y; } else if (cv) { px = Cc.x; py = Cc.y; } else if (bv) { px = B.x; py = B.y; }
    if ((av ? 1 : 0) + (bv ? 1 : 0) + (cv ? 1 : 0) > 1) { px = median3(A.x, B.x, Cc.x); py = median3(A.y, B.y, Cc.y); }
    pullback(mx * 64, my * 64, mbw * 64 - 4, mbh * 64 - 4, -60, px, py);
    bool hyb = false;
    if (av && cv) {
        if (iabs(px - A.x) + iabs(py - A.y) > 32) hyb = true;
        else if (iabs(px - Cc.x) + iabs(py - Cc.y) > 32) hyb = true;
    }
    return {px, py, A.x, A.y, Cc.x, Cc.y, av, cv, hyb};
}

// luma block-vector field: the current macroblock's blocks from LDS own[],
// its neighbours' from the snapshot
__device__ __forceinline__ int2 field_get(const Ctx& C, int x, int y) {
    if (x < 0 || y < 0 || x >= C.g.mbw * 2 || y >= C.g.mbh * 2) return {0, 0};
    const int mmx = x >> 1, mmy = y >> 1, k = ((y & 1) << 1) | (x & 1);
    if (mmx == C.mx && mmy == C.my) return C.L.own[k];
    return nb_bmv(C, nb_slot(mmx - C.mx, mmy - C.my), k);
}
__device__ __forceinline__ bool intra_at(const Ctx& C, int x, int y) {
    if (x < 0 || y < 0 || x >= C.g.mbw * 2 || y >= C.g.mbh * 2) return false;
    const int mmx = x >> 1, mmy = y >> 1;
    if (mmx == C.mx && mmy == C.my) return false;
    return (nb_flags(C, nb_slot(mmx - C.mx, mmy - C.my)) & kFlagIntra) != 0;
}

__device__ __forceinline__ bool hybrid_flag(const Ctx& C, int px, int py, int bx, int by, int2 A, int2 Cc) {
    const int suma = intra_at(C, bx, by - 1) ? iabs(px) + iabs(py) : iabs(px - A.x) + iabs(py - A.y);
    if (suma > 32) return true;
    const int sumc = intra_at(C, bx - 1, by) ? iabs(px) + iabs(py) : iabs(px - Cc.x) + iabs(py - Cc.y);
    return sumc > 32;
}

// predictor_info_mixed_1mv, src/encoder_motion.cpp:1354-1406
__device__ __forceinline__ Pred predictor_mixed_1mv(const Ctx& C) {
    const int mbw = C.g.mbw, mbh = C.g.mbh, mx = C.mx, my = C.my;
    const int bx = mx * 2, by = my * 2;
    const bool av = my > 0, cv = mx > 0, bv = my > 0 && mbw > 1;
    const int2 A = av ? field_get(C, bx, by - 1) : make_int2(0, 0);
    const int2 Cc = cv ? field_get(C, bx - 1, by) : make_int2(0, 0);
    const int2 B = bv ? field_get(C, bx + ((mx == mbw - 1) ? -1 : 2), by - 1) : make_int2(0, 0);
    int px = 0, py = 0;
    if (av) { px = A.x; py = A.y; } else if (cv) { px = Cc.x; py = Cc.y; } else if (bv) { px = B.x; py = B.y; }
    if ((av ? 1 : 0) + (bv ? 1 : 0) + (cv ? 1 : 0) > 1) { px = median3(A.x, B.x, Cc.x); py = median3(A.y, B.y, Cc.y); }
    pullback(mx * 64, my * 64, mbw * 64 - 4, mbh * 64 - 4, -60, px, py);
    const bool hyb = av && cv && hybrid_flag(C, px, py, bx, by, A, Cc);
    return {px, py, A.x, A.y, Cc.x, Cc.y, av, cv, hyb};
}

// predictor_info_4mv, src/encoder_motion.cpp:1408-1466
__device__ __forceinline__ Pred predictor_4mv(const Ctx& C, int n) {
    const int mbw = C.g.mbw, mbh = C.g.mbh, mx = C.mx, my = C.my;
    const int bx = mx * 2 + (n & 1), by = my * 2 + ((n >> 1) & 1);
    const bool av = my > 0 || n == 2 || n == 3;
    const bool cv = mx > 0 || n == 1 || n == 3;
    const bool bv = av;
    const int2 A = av ? field_get(C, bx, by - 1) : make_int2(0, 0);
    const int2 Cc = cv ? field_get(C, bx - 1, by) : make_int2(0, 0);
    int2 B = {0, 0};
    if (bv) {
        int bxx = bx;
        if (n == 0) bxx += mx ? -1 : 1;
        else if (n == 1) bxx += (mx == mbw - 1) ? -1 : 1;
        else if (n == 2) bxx += 1;
        else bxx -= 1;
        B = field_get(C, bxx, by - 1);
    }
    int px = 0, py = 0;
    if (av) { px = A.x; py = A.y; } else if (cv) { px = Cc.x; py = Cc.y; } else if (bv) { px = B.x; py = B.y; }
    if ((av ? 1 : 0) + (bv ? 1 : 0) + (cv ? 1 : 0) > 1) { px = median3(A.x, B.x, Cc.x); py = median3(A.y, B.y, Cc.y); }
    pullback(mx * 64 + ((n == 1 || n == 3) ? 32 : 0), my * 64 + ((n == 2 || n == 3) ? 32 : 0), mbw * 64 - 4,
             mbh * 64 - 4, -28, px, py);
    const bool hyb = av && cv && hybrid_flag(C, px, py, bx, by, A, Cc);
    return {px, py, A.x, A.y, Cc.x, Cc.y, av, cv, hyb};
}

// ------------------------------------------------------------------ MV rate
// motion_vector_rate_bits and its syntax helpers, src/encoder_motion.cpp:1099-1197, :3214-3225

__device__ __forceinline__ int mv_cat(int v, bool hp) {
    const int a = iabs(v);
    if (!a) return 0;
    if (a <= 2) return 1;
    if (a <= 6) return 2;
    if (a <= 14) return 3;
    if (a <= 30) return 4;
    if (a <= (hp ? 94 : 158)) return 5;
    return -1;
}
__device__ __forceinline__ int mv_component_bits(int cat, bool hp) {
    const int sz = cat == 0 ? 0 : cat == 1 ? 2 : cat == 2 ? 3 : cat == 3 ? 4 : cat == 4 ? 5 : 8;
    return sz - ((hp && cat == 5) ? 1 : 0);
}
__device__ __forceinline__ int mvrange_kx(int r) { return r == 0 ? 9 : r == 1 ? 10 : r == 2 ? 12 : 13; }
__device__ __forceinline__ int mvrange_ky(int r) { return r == 0 ? 8 : r == 1 ? 9 : r == 2 ? 10 : 11; }
__device__ __forceinline__ int mvrange_xq(int r) { return r == 0 ? 256 : r == 1 ? 512 : r == 2 ? 2048 : 4096; }
__device__ __forceinline__ int mvrange_yq(int r) { return r == 0 ? 128 : r == 1 ? 256 : r == 2 ? 512 : 1024; }

// bits of a differential (dx,dy) (already in mode units), with `more`
__device__ __forceinline__ uint32_t mvdiff_bits(int mvrange, int dx, int dy, bool hp, bool more) {
    const int xc = mv_cat(dx, hp), yc = mv_cat(dy, hp);
    int sym;
    if (xc < 0 || yc < 0) sym = more ? 71 : 34;
    else {
        const int idx = yc * 6 + xc;
        if (!more && idx == 0) sym = 34;
        else sym = more ? idx + 36 : idx - 1;
    }
    int suffix;
    if (sym == 34 || sym == 71 || xc < 0 || yc < 0) {
        const int h = hp ? 1 : 0;
        suffix = (mvrange_kx(mvrange) - h) + (mvrange_ky(mvrange) - h);
    } else {
        suffix = mv_component_bits(xc, hp) + mv_component_bits(yc, hp);
    }
    return mv_bits0(sym) + static_cast<uint32_t>(suffix);
}
__device__ __forceinline__ uint32_t mv_rate_bits(int mvrange, int xq, int yq, int pxq, int pyq, int mode, bool more) {
    const bool hp = mode_halfpel(mode);
    int dx = xq - pxq, dy = yq - pyq;
    if (hp) { dx /= 2; dy /= 2; }
    return mvdiff_bits(mvrange, dx, dy, hp, more);
}
// modular_mv_distance, src/encoder_motion.cpp:1299-1310
__device__ __forceinline__ int mod_dist(int a, int b, int range) {
    // ((d + range) % span + span) % span is the residue of t = d + range in
    // [0, span): t itself when it already lies there (the vectors of one
    // picture differ by less than the span), the general form otherwise
    const int span = range * 2;
    int t = a - b + range;
    if (t < 0 || t >= span) {
        t %= span;
        if (t < 0) t += span;
    }
    return iabs(t - range);
}
__device__ __forceinline__ int modular_mv_distance(int mvrange, int ax, int ay, int bx, int by) {
    return mod_dist(ax, bx, mvrange_xq(mvrange)) + mod_dist(ay, by, mvrange_yq(mvrange));
}

// ------------------------------------------------------------------ interpolation
// luma_mc_sample / luma_mc_sample_mode, src/encoder_motion_compensation.cpp:360-415,
// reading the LDS window (which already holds clamped samples).

__device__ __forceinline__ int mspel(int a, int b, int c, int d, int mode) {
    if (mode == 1) return -4 * a + 53 * b + 18 * c - 3 * d;
    if (mode == 2) return -a + 9 * b + 9 * c - d;
    return -3 * a + 18 * b + 53 * c - 4 * d;
}

// ------------------------------------------------------------------ interpolated reference

// One interpolated sample (luma_mc_sample / luma_mc_sample_mode,
// src/encoder_motion_compensation.cpp:367-415) at quarter-pel phase (hm, vm)
// from its clamped 4x4 integer neighbourhood n[r][c] = ref(py - 1 + r, px - 1 + c).
__device__ __forceinline__ int mc_sample(const int n[4][4], int hm, int vm, bool bil) {
    if (bil) {
        if (hm && vm) return (n[1][1] + n[1][2] + n[2][1] + n[2][2] + 2) >> 2;
        if (hm) return (n[1][1] + n[1][2] + 1) >> 1;
        if (vm) return (n[1][1] + n[2][1] + 1) >> 1;
        return n[1][1];
    }
    if (!hm && !vm) return n[1][1];
    if (!vm) {
        const int bias = hm == 2 ? 8 : 32, sh = hm == 2 ? 4 : 6;
        return clampi(arshift(mspel(n[1][0], n[1][1], n[1][2], n[1][3], hm) + bias, sh), 0, 255);
    }
    if (!hm) {
        const int bias = vm == 2 ? 7 : 31, sh = vm == 2 ? 4 : 6;
        return clampi(arshift(mspel(n[0][1], n[1][1], n[2][1], n[3][1], vm) + bias, sh), 0, 255);
    }
    const int shift = ((hm == 2 ? 1 : 5) + (vm == 2 ? 1 : 5)) >> 1;
    const int r1 = (1 << (shift - 1)) - 1;
    int t[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) t[c] = arshift(mspel(n[0][c], n[1][c], n[2][c], n[3][c], vm) + r1, shift);
    return clampi(arshift(mspel(t[0], t[1], t[2], t[3], hm) + 64, 7), 0, 255);
}

// A block's candidate costs read the interpolated reference from phase
// planes: plane (hm, vm) holds the sample of that phase at every integer
// position px in [-2, W], py in [-2, H].  Every filter tap of a position
// outside that box reads the clamped frame edge, so the sample at any
// position is the plane sample at the clamped position.

// Four predicted samples of plane row `row` starting at plane coordinate x0
// (columns x0 .. x0+3, clamped into [-2, W]), packed little-endian.
__device__ __forceinline__ uint32_t plane_row4(gptr_u8 row, int x0, int W) {
    if (x0 >= -2 && x0 + 3 <= W) {
        const int a = x0 + 2;
        typedef const __attribute__((address_space(1))) uint32_t* gptr_u32;
        const gptr_u32 q = reinterpret_cast<gptr_u32>(row + (a & ~3));
        return __builtin_amdgcn_alignbyte(q[1], q[0], a & 3);
    }
    uint32_t r = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) r |= static_cast<uint32_t>(row[clampi(x0 + j, -2, W) + 2]) << (8 * j);
    return r;
}

// Stage the phase window with sample (0,0) at plane coordinates (wx, wy):
// kPwSide x kPwSide samples of every phase of the filter set, copied from the
// phase planes as 4-byte aligned dwords (a row of the window starts pwsh
// bytes into its LDS row).
__device__ __forceinline__ void pw_load(Ctx& C, int wx, int wy, bool bil) {
    const JobDesc* J = C.J;
    const int np = bil ? 4 : 16;
    const gptr_u8 planes = as_global(bil ? J->bp : J->qp);
    const size_t psz = static_cast<size_t>(J->pp) * J->ph;
    const int W = C.g.W, H = C.g.H;
    const int c0 = wx + 2;
    const int a0 = c0 & ~3;
    constexpr int kDwords = kPwPitch / 4;
    const int total = np * kPwSide * kDwords;
    wsync();
    if (a0 >= 0 && a0 + kPwPitch <= W + 3) {
        // every dword lies inside the plane columns: copy global -> LDS
        // directly, 64 dwords (256 bytes, dword i at byte 4 i) per instruction
        typedef const __attribute__((address_space(1))) void* gvoid;
        typedef __attribute__((address_space(3))) void* lvoid;
        // the lane's dword base + lane of the window as (plane, row r,
        // dword d), advanced by 64 dwords (9 rows and 1 dword) per copy:
        // no division, and the row offset a 24-bit product (rows and the
        // plane pitch are below 2^24)
        static_assert(kWave == 9 * kDwords + 1 && kPwSide > 10, "the window walk advances 9 rows and 1 dword");
        int r = C.lane / kDwords, d = C.lane - (C.lane / kDwords) * kDwords;
        gptr_u8 pl = planes;
        const uint32_t upp = static_cast<uint32_t>(J->pp);
        for (int base = 0; base < total; base += kWave) {
            const int prow = clampi(wy + r, -2, H) + 2;
            const gptr_u8 src = pl + (__umul24(static_cast<uint32_t>(prow), upp) + static_cast<uint32_t>(a0 + d * 4));
            // lanes past the window issue no copy (a copy lands at byte 4 * lane)
            if (C.lane < total - base) __builtin_amdgcn_global_load_lds((gvoid)src, (lvoid)(C.L.pw + base * 4), 4, 0, 0);
            r += 9;
            if (++d >= kDwords) { d -= kDwords; ++r; }
            if (r >= kPwSide) { r -= kPwSide; pl += psz; }
        }
    } else {
        // the same walk; the columns clamp to the plane per byte
        int r = C.lane / kDwords, d = C.lane - (C.lane / kDwords) * kDwords;
        gptr_u8 pl = planes;
        const uint32_t upp = static_cast<uint32_t>(J->pp);
        for (int base = 0; base < total; base += kWave) {
            const int prow = clampi(wy + r, -2, H) + 2;
            const gptr_u8 row = pl + __umul24(static_cast<uint32_t>(prow), upp);
            const int col = a0 + d * 4;
            if (C.lane < total - base) {
                uint32_t v = 0;
#pragma unroll
                for (int b = 0; b < 4; ++b) v |= static_cast<uint32_t>(row[clampi(col + b, 0, W + 2)]) << (8 * b);
                reinterpret_cast<uint32_t*>(C.L.pw)[base + C.lane] = v;
            }
            r += 9;
            if (++d >= kDwords) { d -= kDwords; ++r; }
            if (r >= kPwSide) { r -= kPwSide; pl += psz; }
        }
    }
    C.pwx = wx; C.pwy = wy; C.pwsh = c0 - a0; C.pw_ok = true; C.pw_bil = bil;
    C.wx0 = C.wy0 = -(1 << 30); C.ww = C.wh = C.wp = 0;   // the integer window is overwritten
    wsync();
}

// Costs of the candidate vectors (vx, vy) held by lanes first .. first+cnt-1
// on block (bx0,by0,bw,bh); `want` (per lane) selects the candidates to score,
// the others cost kBadSad.  metric 0 = SAD with the hidden-padding veto of
// block_sad_qpel / block_sad_motion (src/encoder_motion.cpp:2286-2293,
// :2670-2676); metric 1 = SATD of block_satd_motion
// (src/encoder_motion.cpp:2962-3042: rendered prediction, (|H D H^T| + 1) >> 1
// per 4x4 tile, zero residual outside the active area).  Each lane owns one
// 4x4 tile of one candidate: a 16x16 block takes 16 lanes (four candidates per
// pass), an 8x8 block 4 lanes (sixteen per pass).  Returns the cost of its own
// candidate in each lane first+c, `keep` in the other lanes.
__device__ __forceinline__ uint64_t qpel_eval(Ctx& C, int bx0, int by0, int bw, int bh, int vx, int vy, bool want,
                                              int first, int cnt, int mode, int metric, uint64_t keep,
                                              bool count = true) {
    uint64_t res = keep;
    if (cnt <= 0) return res;
    const JobDesc* J = C.J;
    const bool bil = mode_bilinear(mode);
    const int W = C.g.W, H = C.g.H;
    const int aw = imax(0, imin(bw, W - bx0)), ah = imax(0, imin(bh, H - by0));
    const bool chroma = chroma_block(bx0, by0, bw, bh);
    const int lshift = bw == 16 ? 4 : 2;             // log2(lanes per candidate)
    const int g = C.lane >> lshift, t = C.lane & ((1 << lshift) - 1);
    const int tx = (bw == 16 ? (t & 3) : (t & 1)) * 4, ty = (bw == 16 ? (t >> 2) : (t >> 1)) * 4;
    const int cpp = kWave >> lshift;                 // candidates per pass
    uint32_t cur[4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
        cur[r] = *reinterpret_cast<const uint32_t*>(C.L.cur + (by0 - C.y0 + ty + r) * 16 + (bx0 - C.x0 + tx));
    uint32_t cm = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (tx + j < aw) cm |= 0xffu << (8 * j);
    const bool tvalid = tx < aw && ty < ah;
    const gptr_u8 planes = as_global(bil ? J->bp : J->qp);
    const size_t psz = static_cast<size_t>(J->pp) * J->ph;
    const int wanted = want ? 1 : 0;
    for (int base = 0; base < cnt; base += cpp) {
        const int c = base + g;
        const bool mine = c < cnt;
        const int src = first + (mine ? c : 0);
        const int2 v = make_int2(__shfl(vx, src), __shfl(vy, src));
        const bool w = __shfl(wanted, src) != 0;
        bool skip = !mine || !w;
        uint64_t skip_cost = kBadSad;
        if (mine && w && metric == 0) {
            if (hidden_padding(C.g, bx0, by0, bw, bh, v.x, v.y, bil ? mode : 0, chroma)) skip = true;
            else if (bil && ((v.x & 1) || (v.y & 1))) skip = true;
        }
        if (mine && w && metric == 1) {
            if (!aw || !ah) { skip = true; skip_cost = 0; }
            else if (mode_halfpel(mode) && ((v.x & 1) || (v.y & 1))) skip = true;
        }
        const int ix = floor_div4(v.x), iy = floor_div4(v.y);
        const int hm = v.x - ix * 4, vm = v.y - iy * 4;
        uint32_t part = 0;
        if (!skip) {
            const int phase = bil ? (vm >> 1) * 2 + (hm >> 1) : vm * 4 + hm;
            const int x0 = bx0 + tx + ix;
            uint32_t pr[4];
            const bool in_pw = C.pw_ok && C.pw_bil == bil && bx0 + ix >= C.pwx && bx0 + ix + bw <= C.pwx + kPwSide &&
                               by0 + iy >= C.pwy && by0 + iy + bh <= C.pwy + kPwSide;
            if (in_pw) {
                const int a = x0 - C.pwx + C.pwsh;
                const uint8_t* base_row =
                    C.L.pw + (__umul24(static_cast<uint32_t>(phase * kPwSide + (by0 + ty + iy - C.pwy)), static_cast<uint32_t>(kPwPitch)) +
                              static_cast<uint32_t>(a & ~3));
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const uint32_t* q = reinterpret_cast<const uint32_t*>(base_row + r * kPwPitch);
                    pr[r] = __builtin_amdgcn_alignbyte(q[1], q[0], a & 3);
                }
            } else {
                const gptr_u8 pl = planes + static_cast<size_t>(phase) * psz;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int yc = clampi(by0 + ty + r + iy, -2, H) + 2;
                    pr[r] = plane_row4(pl + static_cast<size_t>(yc) * J->pp, x0, W);
                }
            }
            if (metric == 0) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (ty + r < ah) part = __builtin_amdgcn_sad_u8(pr[r] & cm, cur[r] & cm, part);
            } else {
                int h[4][4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    int d[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        d[j] = (ty + r < ah && tx + j < aw)
                                   ? static_cast<int>((cur[r] >> (8 * j)) & 255u) - static_cast<int>((pr[r] >> (8 * j)) & 255u)
                                   : 0;
                    const int a0 = d[0] + d[3], a1 = d[1] + d[2], a2 = d[1] - d[2], a3 = d[0] - d[3];
                    h[r][0] = a0 + a1; h[r][1] = a3 + a2; h[r][2] = a0 - a1; h[r][3] = a3 - a2;
                }
                uint32_t sum = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int b0 = h[0][j] + h[3][j], b1 = h[1][j] + h[2][j];
                    const int b2 = h[1][j] - h[2][j], b3 = h[0][j] - h[3][j];
                    sum += static_cast<uint32_t>(iabs(b0 + b1) + iabs(b3 + b2) + iabs(b0 - b1) + iabs(b3 - b2));
                }
                part = tvalid ? ((sum + 1) >> 1) : 0u;
            }
        }
        const uint32_t total = group_sum32(part, 1 << lshift);
        // op counts: the lead lane of each scored candidate counts it
        if (count && !skip && t == 0) {
            const uint32_t px = static_cast<uint32_t>(aw * ah);
            if (metric == 0) C.ln_sad += px;
            C.ln_taps += px * static_cast<uint32_t>(taps_per_sample(hm, vm, bil));
            if (metric == 1) C.ln_satd += 112u * static_cast<uint32_t>(((aw + 3) / 4) * ((ah + 3) / 4));
        }
        // hand each cost to the lane that holds its vector
        const uint64_t val = skip ? skip_cost : static_cast<uint64_t>(total);
        const int rel = C.lane - first - base;
        const bool here = rel >= 0 && rel < cpp && base + rel < cnt;
        const uint64_t got = lane_shfl64(val, (here ? rel : 0) << lshift);
        if (here) res = got;
    }
    return res;
}

// Op counts of one more SAD evaluation of v on the block, as qpel_eval counts
// it: for evaluations the reference repeats and this code serves from a table.
__device__ __forceinline__ void count_sad_eval(Ctx& C, int bx0, int by0, int bw, int bh, int2 v, int mode) {
    const bool bil = mode_bilinear(mode);
    const int aw = imax(0, imin(bw, C.g.W - bx0)), ah = imax(0, imin(bh, C.g.H - by0));
    if (hidden_padding(C.g, bx0, by0, bw, bh, v.x, v.y, bil ? mode : 0, chroma_block(bx0, by0, bw, bh)) ||
        (bil && ((v.x & 1) || (v.y & 1))))
        return;
    C.n_sad += static_cast<uint64_t>(aw) * ah;
    C.n_taps += static_cast<uint64_t>(aw) * ah * taps_per_sample(v.x & 3, v.y & 3, bil);
}

// The same op counts added to this lane's partial counters (lane-parallel callers).
__device__ __forceinline__ void count_sad_eval_lane(Ctx& C, int bx0, int by0, int bw, int bh, int2 v, int mode) {
    const bool bil = mode_bilinear(mode);
    const int aw = imax(0, imin(bw, C.g.W - bx0)), ah = imax(0, imin(bh, C.g.H - by0));
    if (hidden_padding(C.g, bx0, by0, bw, bh, v.x, v.y, bil ? mode : 0, chroma_block(bx0, by0, bw, bh)) ||
        (bil && ((v.x & 1) || (v.y & 1))))
        return;
    C.ln_sad += static_cast<uint32_t>(aw * ah);
    C.ln_taps += static_cast<uint32_t>(aw * ah * taps_per_sample(v.x & 3, v.y & 3, bil));
}

// ------------------------------------------------------------------ integer searches

struct IRes {
    int dx, dy;
    uint64_t sad;
};

// Division by a divisor fixed for a batch: q = floor(i / d) as the high word
// of i * ceil(2^32 / d), exact whenever i * d < 2^32 (the grid indices here
// stay below 2^17 and the divisors below 2^11).
struct FastDiv {
    uint32_t d, m;
};
__device__ __forceinline__ FastDiv fast_div(int d) {
    return {static_cast<uint32_t>(d), d > 1 ? static_cast<uint32_t>((0x100000000ull + d - 1) / d) : 0u};
}
__device__ __forceinline__ int fdiv(int i, FastDiv f) {
    return f.d > 1 ? static_cast<int>(__umulhi(static_cast<uint32_t>(i), f.m)) : i;
}

// integer_motion_umh, src/encoder_motion.cpp:1781-2023.  The reference keeps
// a running best and a top-8 list; with distinct points both are functions of
// the set of points considered so far, so each stage scores its points
// together and only the stage minimum updates `best`.  Every considered key
// is logged, and the finalists (the best four of the top-8) are extracted
// from the log once, at the end.
__device__ __forceinline__ IRes umh_search(Ctx& C, const Pred pi, int gdx, int gdy, int R) {
    // the whole search stays inside [-R,R]^2: one window serves every stage
    // (issued at the start of the analysis when it could be)
    if (!win_is(C, C.x0 - R, C.y0 - R, 16 + 2 * R, 16 + 2 * R) &&
        !win_load(C, C.x0 - R, C.y0 - R, 16 + 2 * R, 16 + 2 * R)) {
        set_err(C, 5);
        return {0, 0, kBadSad};
    }
    for (int i = C.lane; i < seen_words(C.J->local_range); i += kWave) { C.L.seen_s[i] = 0; C.L.seen_f[i] = 0; }
    wsync();
    uint64_t best = kNone;
    Top4 top{kNone, kNone, kNone, kNone};
    auto cl = [R](int v) VC1ME_INLINE { return clampi(v, -R, R); };
    auto stage = [&](int n, auto pt) VC1ME_INLINE {
        const uint64_t m = int_batch(C, n, 2, C.L.seen_s, KEY_TOTAL, 0, pt, &top);
        best = m < best ? m : best;
    };

    // predictor-driven seeds
    {
        // predictor, zero, global, then A and C where valid (seed i of lane i;
        // registers and selects, no indexed array)
        const int px = cl(round_qpel_to_pixel(pi.px)), py = cl(round_qpel_to_pixel(pi.py));
        const int ax = cl(round_qpel_to_pixel(pi.ax)), ay = cl(round_qpel_to_pixel(pi.ay));
        const int qx = cl(round_qpel_to_pixel(pi.cx)), qy = cl(round_qpel_to_pixel(pi.cy));
        const int gx = cl(gdx), gy = cl(gdy);
        const int n = 3 + (pi.a_valid ? 1 : 0) + (pi.c_valid ? 1 : 0);
        stage(n, [=](int i, int& dx, int& dy) VC1ME_INLINE {
            const bool use_a = i == 3 && pi.a_valid;
            dx = i == 0 ? px : i == 1 ? 0 : i == 2 ? gx : use_a ? ax : qx;
            dy = i == 0 ? py : i == 1 ? 0 : i == 2 ? gy : use_a ? ay : qy;
        });
    }
    // local 5x5 around the seeds' best, and the sparse lattice (R >= 8) in
    // the same batch: the lattice's points do not depend on the best so far,
    // and the best and the top list the cross starts from are functions of
    // the set of points considered, whatever their order
    {
        const int cx = key_dx(best), cy = key_dy(best);
        const int side = R >= 8 ? (2 * R) / 8 + 1 : 0;
        const FastDiv fs = fast_div(side > 0 ? side : 1);
        stage(25 + side * side + 2 * side, [=](int i, int& dx, int& dy) VC1ME_INLINE {
            if (i < 25) { dx = cx + (i % 5) - 2; dy = cy + (i / 5) - 2; }
            else if (i < 25 + side * side) {
                const int q = fdiv(i - 25, fs);
                dx = -R + 8 * (i - 25 - q * side); dy = -R + 8 * q;
            }
            else {
                const int j = i - 25 - side * side, d = -R + 8 * (j >> 1);
                if (j & 1) { dx = d; dy = R; } else { dx = R; dy = d; }
            }
            dx = cl(dx); dy = cl(dy);
        });
    }
    // uneven cross
    {
        const int cs = R >= 96 ? 8 : (R >= 32 ? 4 : 2);
        const int cx = key_dx(best), cy = key_dy(best);
        const int vlimit = imin(R, imax(8, R / 2));
        const int nh = R / cs, nv = vlimit / cs;
        stage(2 * (nh + nv), [=](int i, int& dx, int& dy) VC1ME_INLINE {
            if (i < 2 * nh) {
                const int d = cs * (i / 2 + 1);
                dx = (i & 1) ? cx + d : cx - d; dy = cy;
            } else {
                const int j = i - 2 * nh, d = cs * (j / 2 + 1);
                dx = cx; dy = (j & 1) ? cy + d : cy - d;
            }
            dx = cl(dx); dy = cl(dy);
        });
    }
    // multi-hexagon rings
    {
        const int hs = R >= 96 ? 8 : (R >= 24 ? 4 : 2);
        const int hx = key_dx(best), hy = key_dy(best);
        stage(6 * (R / hs), [=](int i, int& dx, int& dy) VC1ME_INLINE {
            const int r = hs * (i / 6 + 1), h = (r + 1) / 2, k = i % 6;
            switch (k) {
                case 0: dx = hx - r; dy = hy; break;
                case 1: dx = hx + r; dy = hy; break;
                case 2: dx = hx - h; dy = hy - r; break;
                case 3: dx = hx + h; dy = hy - r; break;
                case 4: dx = hx - h; dy = hy + r; break;
                default: dx = hx + h; dy = hy + r; break;
            }
            dx = cl(dx); dy = cl(dy);
        });
    }
    // iterative small hexagon
    for (int iter = 0; iter < 32; ++iter) {
        const int ox = key_dx(best), oy = key_dy(best);
        stage(6, [=](int i, int& dx, int& dy) VC1ME_INLINE {
            switch (i) {
                case 0: dx = ox - 2; dy = oy; break;
                case 1: dx = ox + 2; dy = oy; break;
                case 2: dx = ox - 1; dy = oy - 2; break;
                case 3: dx = ox + 1; dy = oy - 2; break;
                case 4: dx = ox - 1; dy = oy + 2; break;
                default: dx = ox + 1; dy = oy + 2; break;
            }
            dx = cl(dx); dy = cl(dy);
        });
        if (key_dx(best) == ox && key_dy(best) == oy) break;
    }
    // closing diamond
    {
        const int ox = key_dx(best), oy = key_dy(best);
        stage(4, [=](int i, int& dx, int& dy) VC1ME_INLINE {
            dx = cl(ox + (i == 0 ? -1 : i == 1 ? 1 : 0));
            dy = cl(oy + (i == 2 ? -1 : i == 3 ? 1 : 0));
        });
    }
    // the best four considered points (top[0..3] of the reference's list)
    const Top4 wt = top4_wave(top);
    const uint64_t tk[4] = {wt.k0, wt.k1, wt.k2, wt.k3};
    const int fin = (wt.k0 != kNone) + (wt.k1 != kNone) + (wt.k2 != kNone) + (wt.k3 != kNone);
    // full-resolution finalists: 3x3 around the best four
    uint64_t full = int_batch(C, fin * 9, 1, C.L.seen_f, KEY_TOTAL, 0, [=](int i, int& dx, int& dy) VC1ME_INLINE {
        const int f = i / 9, o = i % 9;
        const uint64_t k = f == 0 ? tk[0] : f == 1 ? tk[1] : f == 2 ? tk[2] : tk[3];
        dx = cl(key_dx(k) + (o % 3) - 1); dy = cl(key_dy(k) + (o / 3) - 1);
    });
    if (full == kNone) {
        const int bx = key_dx(best), by = key_dy(best);
        full = int_batch(C, 1, 1, C.L.seen_f, KEY_TOTAL, 0, [=](int, int& dx, int& dy) VC1ME_INLINE { dx = bx; dy = by; });
    }
    // rescue of genuinely poor matches
    if (R > 0 && !match_good(C, key_cost(full))) {
        const int step = R >= 8 ? 4 : (R >= 3 ? 2 : 1);
        const int side = (2 * R) / step + 1;
        const FastDiv fs = fast_div(side);
        const uint64_t coarse = int_batch(C, side * side + 2 * side, 4, nullptr, KEY_TOTAL, 0,
                                          [=](int i, int& dx, int& dy) VC1ME_INLINE {
            if (i < side * side) { const int q = fdiv(i, fs); dx = -R + step * (i - q * side); dy = -R + step * q; }
            else {
                const int j = i - side * side, d = -R + step * (j >> 1);
                if (j & 1) { dx = d; dy = R; } else { dx = R; dy = d; }
            }
        });
        const int cdx = key_dx(coarse), cdy = key_dy(coarse);
        const int rr = imax(1, step - 1);
        const int xa = imax(-R, cdx - rr), xb = imin(R, cdx + rr), ya = imax(-R, cdy - rr), yb = imin(R, cdy + rr);
        const int rw = xb - xa + 1, rh = yb - ya + 1;
        const FastDiv frw = fast_div(rw);
        const uint64_t rescue = int_batch(C, rw * rh, 1, C.L.seen_f, KEY_TOTAL, 0, [=](int i, int& dx, int& dy) VC1ME_INLINE {
            const int q = fdiv(i, frw);
            dx = xa + i - q * rw; dy = ya + q;
        });
        if (rescue < full) full = rescue;
        const uint64_t px = active_pixels(C);
        if (px && key_cost(full) > px * 24u && R <= 32) {
            const int ds = (2 * R) / 2 + 1;
            const FastDiv fds = fast_div(ds);
            const uint64_t dense = int_batch(C, ds * ds, 2, nullptr, KEY_TOTAL, 0, [=](int i, int& dx, int& dy) VC1ME_INLINE {
                const int q = fdiv(i, fds);
                dx = -R + 2 * (i - q * ds); dy = -R + 2 * q;
            });
            const int ddx = key_dx(dense), ddy = key_dy(dense);
            const int xa2 = imax(-R, ddx - 2), xb2 = imin(R, ddx + 2), ya2 = imax(-R, ddy - 2), yb2 = imin(R, ddy + 2);
            const int w2 = xb2 - xa2 + 1, h2 = yb2 - ya2 + 1;
            const FastDiv fw2 = fast_div(w2);
            const uint64_t dfull = int_batch(C, w2 * h2, 1, C.L.seen_f, KEY_TOTAL, 0, [=](int i, int& dx, int& dy) VC1ME_INLINE {
                const int q = fdiv(i, fw2);
                dx = xa2 + i - q * w2; dy = ya2 + q;
            });
            if (dfull < full) full = dfull;
        }
    }
    return {key_dx(full), key_dy(full), key_cost(full)};
}

// integer_motion_diamond, src/encoder_motion_diamond.cpp:13-73
__device__ __forceinline__ IRes diamond_search(Ctx& C, const Pred pi, int gdx, int gdy, int R) {
    if (!win_is(C, C.x0 - R, C.y0 - R, 16 + 2 * R, 16 + 2 * R) &&
        !win_load(C, C.x0 - R, C.y0 - R, 16 + 2 * R, 16 + 2 * R)) {
        set_err(C, 5);
        return {0, 0, kBadSad};
    }
    uint64_t best = int_batch(C, 1, 1, nullptr, KEY_TOTAL, 0, [](int, int& dx, int& dy) VC1ME_INLINE { dx = 0; dy = 0; });
    if (key_cost(best) == 0) return {0, 0, 0};
    // seed 0 is (0,0); seeds 1..4 (predictor, global, A, C) are kept in
    // order unless they repeat an earlier seed: seed k is slot k's vector if
    // bit k of `keep` is set
    const int p1x = clampi(round_qpel_to_pixel(pi.px), -R, R), p1y = clampi(round_qpel_to_pixel(pi.py), -R, R);
    const int p2x = clampi(gdx, -R, R), p2y = clampi(gdy, -R, R);
    const int p3x = clampi(round_qpel_to_pixel(pi.ax), -R, R), p3y = clampi(round_qpel_to_pixel(pi.ay), -R, R);
    const int p4x = clampi(round_qpel_to_pixel(pi.cx), -R, R), p4y = clampi(round_qpel_to_pixel(pi.cy), -R, R);
    unsigned keep = 0;
    {
        const bool k1 = !(p1x == 0 && p1y == 0);
        const bool k2 = !(p2x == 0 && p2y == 0) && !(k1 && p2x == p1x && p2y == p1y);
        const bool k3 = pi.a_valid && !(p3x == 0 && p3y == 0) && !(k1 && p3x == p1x && p3y == p1y) &&
                        !(k2 && p3x == p2x && p3y == p2y);
        const bool k4 = pi.c_valid && !(p4x == 0 && p4y == 0) && !(k1 && p4x == p1x && p4y == p1y) &&
                        !(k2 && p4x == p2x && p4y == p2y) && !(k3 && p4x == p3x && p4y == p3y);
        keep = (k1 ? 1u : 0u) | (k2 ? 2u : 0u) | (k3 ? 4u : 0u) | (k4 ? 8u : 0u);
    }
    const int sn = 1 + __builtin_popcount(keep);
    // the kept seeds, compacted in order into the LDS seed list and read
    // there by lane index (held in registers and picked by a lane's index,
    // the list went to scratch memory)
    if (sn > 1) {
        if (C.lane == 0) {
            int n = 0;
            if (keep & 1u) { C.L.sd[2 * n] = p1x; C.L.sd[2 * n + 1] = p1y; ++n; }
            if (keep & 2u) { C.L.sd[2 * n] = p2x; C.L.sd[2 * n + 1] = p2y; ++n; }
            if (keep & 4u) { C.L.sd[2 * n] = p3x; C.L.sd[2 * n + 1] = p3y; ++n; }
            if (keep & 8u) { C.L.sd[2 * n] = p4x; C.L.sd[2 * n + 1] = p4y; }
        }
        wsync();
        const int* sd = C.L.sd;
        const uint64_t s = int_batch(C, sn - 1, 1, nullptr, KEY_TOTAL, 0, [=](int i, int& dx, int& dy) VC1ME_INLINE {
            dx = sd[2 * i];
            dy = sd[2 * i + 1];
        });
        if (s < best) best = s;
    }
    int prevx = key_dx(best), prevy = key_dy(best);
    bool have_prev = false;
    for (int step = 0; step <= 4 * R + 1; ++step) {
        const uint64_t center = best;
        const int cx = key_dx(center), cy = key_dy(center);
        // the four neighbours left, right, up, down that lie in range and are
        // not the previous centre, in that order: bit k of `keep`
        unsigned keep = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int dx = cx + (k == 0 ? -1 : k == 1 ? 1 : 0), dy = cy + (k == 2 ? -1 : k == 3 ? 1 : 0);
            if (dx < -R || dx > R || dy < -R || dy > R) continue;
            if (have_prev && dx == prevx && dy == prevy) continue;
            keep |= 1u << k;
        }
        const int n = __builtin_popcount(keep);
        uint64_t win = center;
        if (n) {
            const uint64_t s = int_batch(C, n, 1, nullptr, KEY_TOTAL, 0, [=](int i, int& dx, int& dy) VC1ME_INLINE {
                // the i-th set bit of keep
                unsigned m = keep;
                for (int j = 0; j < i; ++j) m &= m - 1;
                const int k = __builtin_ctz(m);
                dx = cx + (k == 0 ? -1 : k == 1 ? 1 : 0);
                dy = cy + (k == 2 ? -1 : k == 3 ? 1 : 0);
            });
            if (s < win) win = s;
        }
        if (win == center) break;
        prevx = cx; prevy = cy; have_prev = true;
        best = win;
        if (key_cost(best) == 0) break;
    }
    return {key_dx(best), key_dy(best), key_cost(best)};
}

// ------------------------------------------------------------------ long-range searches

// Integer SADs of a cluster of far candidates: window over the bounding box.
template <class PtFn>
__device__ __forceinline__ uint64_t far_batch(Ctx& C, int n, int stride, int keymode, int seq0, PtFn pt) {
    int xlo = 1 << 30, xhi = -(1 << 30), ylo = 1 << 30, yhi = -(1 << 30);
    for (int i = 0; i < n; ++i) {
        int dx, dy;
        pt(i, dx, dy);
        xlo = imin(xlo, dx); xhi = imax(xhi, dx); ylo = imin(ylo, dy); yhi = imax(yhi, dy);
    }
    if (!win_load(C, C.x0 + xlo, C.y0 + ylo, 16 + xhi - xlo, 16 + yhi - ylo)) {
        set_err(C, 6);
        return kNone;
    }
    return int_batch(C, n, stride, nullptr, keymode, seq0, pt);
}

// ------------------------------------------------------------------ learned seeding

// Full SADs of up to four integer vectors at once, read straight from the
// reference frame: lane = (candidate, row).  Hidden-padding vectors cost
// kBadSad.  Returns the minimum key (cost, |d|, dy, dx) over the vectors i
// < n held in lanes i of (vx, vy) whose bit i of `which` is set.
__device__ __forceinline__ uint64_t global_sad_min(Ctx& C, int vx, int vy, int n, uint32_t which = ~0u) {
    const int W = C.g.W, H = C.g.H;
    const gptr_u8 ry = as_global(C.ry);
    uint64_t best = kNone;
    for (int base = 0; base < n; base += 4) {
        if (!((which >> base) & 15u)) continue;
        const int c = base + (C.lane >> 4), r = C.lane & 15;
        const bool mine = c < n && ((which >> c) & 1u);
        const int dx = __shfl(vx, mine ? c : 0), dy = __shfl(vy, mine ? c : 0);
        const bool hid = hidden_padding(C.g, C.x0, C.y0, 16, 16, dx * 4, dy * 4, 0, true);
        uint32_t sum = 0;
        if (mine && !hid && r < C.ah) {
            const gptr_u8 row = ry + static_cast<size_t>(clampi(C.y0 + r + dy, 0, H - 1)) * W;
            const int x = C.x0 + dx;
            const uint32_t* cr = reinterpret_cast<const uint32_t*>(C.L.cur + r * 16);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                uint32_t v = 0;
#pragma unroll
                for (int k = 0; k < 4; ++k) v |= static_cast<uint32_t>(row[clampi(x + q * 4 + k, 0, W - 1)]) << (8 * k);
                uint32_t m = 0;
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (q * 4 + k < C.aw) m |= 0xffu << (8 * k);
                sum = __builtin_amdgcn_sad_u8(v & m, cr[q] & m, sum);
            }
        }
        sum = group_sum32(sum, 16);
        if (mine && !hid && r == 0) C.ln_sad += static_cast<uint32_t>(C.aw * C.ah);
        const uint64_t key = (mine && r == 0) ? key_int(hid ? kBadSad : static_cast<uint64_t>(sum), dx, dy) : kNone;
        const uint64_t mk = wave_min64(key);
        best = mk < best ? mk : best;
    }
    return best;
}

// A seeded result is kept when its SAD is at most kSeedKeepPerPx per active
// pixel (a missed match of textured content costs several times more; the
// reference's own "good match" bound, 6 per pixel, rejects most correct
// integer matches of quarter-pel motion in noisy content).
constexpr uint64_t kSeedKeepPerPx = 16;
__device__ __forceinline__ bool seed_kept(const Ctx& C, uint64_t sad) {
    const uint64_t px = active_pixels(C);
    return px && sad < UINT64_MAX / 8 && sad <= px * kSeedKeepPerPx;
}

// This lane's seed of the learned search: lanes 0 and 1 the two proposals
// (residuals from the global vector in units of 64 px), lane 2 the predictor
// (when known, else the global vector), lane 3 zero, lane 4 the global
// vector; clamped to the search box.
__device__ __forceinline__ void seed_lane(const Ctx& C, const float* o, int gdx, int gdy, const Bounds& b, int ppx,
                                          int ppy, int& vx, int& vy) {
    const int k = C.lane;
    vx = 0; vy = 0;
    float fx = 0.f, fy = 0.f;
    // gate-only mode: both proposals sit on the global vector (selected, not
    // multiplied by zero, so a non-finite output cannot leak into the seed)
    const bool use = C.J->nn_candidates != 0;
    if (k == 0) {
        fx = (use ? o[0] * 64.f : 0.f) + static_cast<float>(gdx);
        fy = (use ? o[1] * 64.f : 0.f) + static_cast<float>(gdy);
    }
    if (k == 1) {
        fx = (use ? o[2] * 64.f : 0.f) + static_cast<float>(gdx);
        fy = (use ? o[3] * 64.f : 0.f) + static_cast<float>(gdy);
    }
    if (k < 2) {
        vx = static_cast<int>(rintf(fminf(fmaxf(fx, -4096.f), 4096.f)));
        vy = static_cast<int>(rintf(fminf(fmaxf(fy, -4096.f), 4096.f)));
    }
    if (k == 2) { vx = ppx; vy = ppy; }
    if (k == 4) { vx = gdx; vy = gdy; }
    vx = clampi(vx, b.xmin, b.xmax);
    vy = clampi(vy, b.ymin, b.ymax);
}

// Integer search seeded by the learned proposals (no reference counterpart):
// full SADs of the two proposals, the predictor, zero and the global vector,
// then a unit diamond from the best until the centre wins (or 32 steps).
// The result's SAD is never above the best seed's.
__device__ __forceinline__ IRes seeded_search(Ctx& C, const Pred pi, int gdx, int gdy, int range, const float* o) {
    const Bounds b = bounds_px(range);
    int vx, vy;
    seed_lane(C, o, gdx, gdy, b, round_qpel_to_pixel(pi.px), round_qpel_to_pixel(pi.py), vx, vy);
    // the five seeds, uniform; those inside the staged reference window (the
    // local search's, issued at the start of the analysis) are scored from
    // LDS, the others from the frame
    const int s0x = __builtin_amdgcn_readlane(vx, 0), s0y = __builtin_amdgcn_readlane(vy, 0);
    const int s1x = __builtin_amdgcn_readlane(vx, 1), s1y = __builtin_amdgcn_readlane(vy, 1);
    const int s2x = __builtin_amdgcn_readlane(vx, 2), s2y = __builtin_amdgcn_readlane(vy, 2);
    const int s3x = __builtin_amdgcn_readlane(vx, 3), s3y = __builtin_amdgcn_readlane(vy, 3);
    const int s4x = __builtin_amdgcn_readlane(vx, 4), s4y = __builtin_amdgcn_readlane(vy, 4);
    auto seed = [=](int i, int& dx, int& dy) VC1ME_INLINE {
        dx = i == 0 ? s0x : i == 1 ? s1x : i == 2 ? s2x : i == 3 ? s3x : s4x;
        dy = i == 0 ? s0y : i == 1 ? s1y : i == 2 ? s2y : i == 3 ? s3y : s4y;
    };
    uint32_t inw = 0;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        int dx, dy;
        seed(i, dx, dy);
        if (in_window(C, dx, dy)) inw |= 1u << i;
    }
    uint64_t best = kNone;
    if (inw) {
        // a seed outside the window takes the place of the first inside it
        // (scored twice, the same key)
        const int first = __builtin_ctz(inw);
        best = int_batch(C, 5, 1, nullptr, KEY_TOTAL, 0, [=](int i, int& dx, int& dy) VC1ME_INLINE {
            seed(((inw >> i) & 1u) ? i : first, dx, dy);
        });
    }
    if (inw != 31u) {
        const uint64_t g = global_sad_min(C, vx, vy, 5, ~inw & 31u);
        best = g < best ? g : best;
    }
    for (int it = 0; it < 32; ++it) {
        const int cx = key_dx(best), cy = key_dy(best);
        // one 28x28 window serves the descent until it walks out of it
        if (!(in_window(C, cx - 1, cy - 1) && in_window(C, cx + 1, cy + 1)) &&
            !win_load(C, C.x0 + cx - 6, C.y0 + cy - 6, 28, 28)) { set_err(C, 11); break; }
        const uint64_t m = int_batch(C, 4, 1, nullptr, KEY_TOTAL, 0, [=](int i, int& dx, int& dy) VC1ME_INLINE {
            dx = clampi(cx + (i == 0 ? -1 : i == 1 ? 1 : 0), b.xmin, b.xmax);
            dy = clampi(cy + (i == 2 ? -1 : i == 3 ? 1 : 0), b.ymin, b.ymax);
        });
        if (m >= best) break;
        best = m;
    }
    return {key_dx(best), key_dy(best), key_cost(best)};
}

// propagated_long_range_motion, src/encoder_motion.cpp:1658-1694
__device__ __forceinline__ IRes propagated(Ctx& C, int range) {
    const JobDesc* J = C.J;
    const Bounds b = bounds_px(range);
    uint64_t best = kNone;
    // neighbours left, top, top-left, top-right
    auto nbx = [](int k) VC1ME_INLINE { return k == 0 ? -1 : k == 1 ? 0 : k == 2 ? -1 : 1; };
    auto nby = [](int k) VC1ME_INLINE { return k == 0 ? 0 : -1; };
    for (int k = 0; k < 4; ++k) {
        const int nx = C.mx + nbx(k), ny = C.my + nby(k);
        if (nx < 0 || ny < 0 || nx >= C.g.mbw) continue;
        if (!nb_prop(C, k)) continue;           // k: left, top, top-left, top-right
        const int2 f = nb_mv(C, k);
        const int cx = round_qpel_to_pixel(f.x), cy = round_qpel_to_pixel(f.y);
        const uint64_t s = far_batch(C, 13, 1, KEY_SEQ_MAN, k * 13, [=](int i, int& dx, int& dy) VC1ME_INLINE {
            const int ox = i == 1 || i == 9 || i == 11 ? -1 : i == 2 || i == 10 || i == 12 ? 1 : i == 5 ? -2 : i == 6 ? 2 : 0;
            const int oy = i == 3 || i == 9 || i == 10 ? -1 : i == 4 || i == 11 || i == 12 ? 1 : i == 7 ? -2 : i == 8 ? 2 : 0;
            dx = clampi(cx + ox, b.xmin, b.xmax);
            dy = clampi(cy + oy, b.ymin, b.ymax);
        });
        // decode the vector of the winner: recompute its point from the sequence number
        if (s < best) best = s;
        if (best != kNone && match_good(C, key_cost(best))) break;
    }
    if (best == kNone) return {0, 0, UINT64_MAX};
    // map the sequence number back to the vector
    const int seq = static_cast<int>(best & 0xffff);
    const int k = seq / 13, i = seq % 13;
    const int2 f = nb_mv(C, k);
    const int cx = round_qpel_to_pixel(f.x), cy = round_qpel_to_pixel(f.y);
    const int ox = i == 1 || i == 9 || i == 11 ? -1 : i == 2 || i == 10 || i == 12 ? 1 : i == 5 ? -2 : i == 6 ? 2 : 0;
    const int oy = i == 3 || i == 9 || i == 10 ? -1 : i == 4 || i == 11 || i == 12 ? 1 : i == 7 ? -2 : i == 8 ? 2 : 0;
    return {clampi(cx + ox, b.xmin, b.xmax), clampi(cy + oy, b.ymin, b.ymax), key_cost(best)};
}

// long_range_signature of the current block (src/encoder_motion.cpp:1591-1601)
__device__ __forceinline__ uint64_t cur_signature(Ctx& C) {
    const JobDesc* J = C.J;
    const int W = C.g.W, H = C.g.H, cw = W / 2, ch = H / 2;
    uint32_t v[6], n[6];
    const int lx = C.lane & 7, ly = C.lane >> 3;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        int px, py, pw, ph;
        const uint8_t* p;
        if (b < 4) { px = C.x0 + (b & 1) * 8 + lx; py = C.y0 + (b >> 1) * 8 + ly; pw = W; ph = H; p = J->cy; }
        else { px = C.x0 / 2 + lx; py = C.y0 / 2 + ly; pw = cw; ph = ch; p = b == 4 ? J->cu : J->cv; }
        const bool in = px < pw && py < ph && px >= 0 && py >= 0;
        // luma from the current block in LDS (every in-frame pixel of the macroblock)
        const uint32_t pix = !in ? 0u : b < 4 ? C.L.cur[(py - C.y0) * 16 + (px - C.x0)] : p[static_cast<size_t>(py) * pw + px];
        v[b] = wave_sum32(pix);
        n[b] = wave_sum32(in ? 1u : 0u);
    }
    uint64_t s = 0;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        const uint32_t m = n[b] ? (v[b] + n[b] / 2) / n[b] : 128u;
        s |= static_cast<uint64_t>(m & 255u) << (8 * b);
    }
    return s;
}

// long_range_signature_distance, src/encoder_motion.cpp:1603-1613
// = 2 * (sum of |a_i - b_i| over the four luma bytes) + 3 * (the same over the
// two chroma bytes): two byte-SAD instructions.
__device__ __forceinline__ uint32_t sig_distance(uint64_t a, uint64_t b) {
    const uint32_t luma = __builtin_amdgcn_sad_u8(static_cast<uint32_t>(a), static_cast<uint32_t>(b), 0u);
    const uint32_t chroma = __builtin_amdgcn_sad_u8(static_cast<uint32_t>(a >> 32) & 0xffffu,
                                                    static_cast<uint32_t>(b >> 32) & 0xffffu, 0u);
    return 2u * luma + 3u * chroma;
}

// Windows of the extended search: 36 x 36 frame samples around a candidate
// (its 5x5 grid at step 4 and the 5x5 refinement around the grid's best stay
// within +-10 px), rows of 40 bytes: the dword-aligned columns that cover the
// window (its first sample at byte 0..3) and the five-dword read of the row
// SAD at an offset of up to 20 px.
constexpr int kXwSide = 36;
constexpr int kXwPitch = 40;
constexpr int kXwBytes = kXwSide * kXwPitch;
// candidates staged per round: their windows and 25 keys each must fit the
// shared LDS region, which is at least the phase window's size; five fill
// the two wave passes of a round's 125 points
constexpr int kXwPerRound = 5;
constexpr int kXwRoundBytes = kXwPerRound * (kXwBytes + 25 * 8);   // windows, then 25 keys per candidate

// The shared LDS region: the largest of its three uses.
__host__ __device__ inline int shared_bytes(int local_range) {
    const int a = int_scratch_bytes(local_range);
    const int b = kPwBytes > kXwRoundBytes ? kPwBytes : kXwRoundBytes;
    return a > b ? a : b;
}
static_assert(kXwPerRound <= 8 && kXwPerRound * 25 <= 4 * kWave, "one round's points are scored by 4 passes of the wave");

// Stage the windows of candidates t0 .. t0+cnt-1 (offsets in `cand` lanes
// t0..) into slots 0..cnt-1 of the shared LDS region.
__device__ __forceinline__ void xw_load(Ctx& C, uint64_t ckeys, int t0, int cnt) {
    const int W = C.g.W, H = C.g.H;
    uint8_t* const xw = C.L.pw;
    constexpr int kDw = kXwPitch / 4;    // dwords per window row
    wsync();
    for (int k = 0; k < cnt; ++k) {
        const uint64_t ck = lane_read64(ckeys, t0 + k);
        const int wx = C.x0 + key_dx(ck) - 10, wy = C.y0 + key_dy(ck) - 10;
        const int a0 = wx & ~3;
        const gptr_u8 ry = as_global(C.ry);
        uint8_t* const dst = xw + k * kXwBytes;
        if (a0 >= 0 && a0 + kXwPitch <= W && (W & 3) == 0) {
            typedef const __attribute__((address_space(1))) void* gvoid;
            typedef __attribute__((address_space(3))) void* lvoid;
            // the lane's dword base + lane as (row r, dword d), advanced by
            // 64 dwords (6 rows and 4 dwords) per copy; a row's offset is a
            // 24-bit product (rows and the picture width are below 2^24)
            static_assert(kWave == 6 * kDw + 4, "the window walk advances 6 rows and 4 dwords");
            int r = C.lane / kDw, d = C.lane - (C.lane / kDw) * kDw;
#pragma unroll
            for (int base = 0; base < kXwSide * kDw; base += kWave) {
                const int sy = clampi(wy + r, 0, H - 1);
                // lanes past the window issue no copy (a copy lands at byte 4 * lane)
                if (C.lane < kXwSide * kDw - base)
                    __builtin_amdgcn_global_load_lds(
                        (gvoid)(ry + (__umul24(static_cast<uint32_t>(sy), static_cast<uint32_t>(W)) + static_cast<uint32_t>(a0 + d * 4))),
                        (lvoid)(dst + base * 4), 4, 0, 0);
                r += 6;
                d += 4;
                if (d >= kDw) { d -= kDw; ++r; }
            }
        } else {
            for (int i = C.lane; i < kXwSide * kDw; i += kWave) {
                const int r = i / kDw, d = i - r * kDw;
                const int sy = clampi(wy + r, 0, H - 1);
                const gptr_u8 row = ry + static_cast<size_t>(sy) * W;
                uint32_t acc = 0;
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    acc |= static_cast<uint32_t>(row[clampi(a0 + d * 4 + b, 0, W - 1)]) << (8 * b);
                *reinterpret_cast<uint32_t*>(dst + r * kXwPitch + d * 4) = acc;
            }
        }
    }
    C.wx0 = C.wy0 = -(1 << 30); C.ww = C.wh = C.wp = 0;   // the integer window is overwritten
    wsync();
}

// SAD part of rows part, part+step, ... (sampled at `stride`) of the 16x16 block
// at offset (dx,dy) read from an extended-search window whose sample (0,0) is
// frame position (wx, wy) and whose first dword starts at column wx - sh.
__device__ __forceinline__ uint32_t xw_sad_part(const Ctx& C, const uint8_t* win, int wx, int wy, int dx, int dy,
                                                int stride, int part, int step) {
    const int sh = wx & 3;
    const int ox = C.x0 + dx - wx + sh, oy = C.y0 + dy - wy;
    const int nrows = (C.ah + stride - 1) / stride;
    const uint32_t smask = stride == 1 ? 0xffffffffu : 0x00ff00ffu;
    uint32_t m[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        uint32_t cm = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (j * 4 + b < C.aw) cm |= 0xffu << (8 * b);
        m[j] = cm & smask;
    }
    const int base = ox & ~3, al = ox & 3;
    uint32_t acc = 0;
    for (int r = part; r < nrows; r += step) {
        const int y = r * stride;
        const uint32_t* wr = reinterpret_cast<const uint32_t*>(win + (__umul24(static_cast<uint32_t>(oy + y), static_cast<uint32_t>(kXwPitch)) +
                                                                   static_cast<uint32_t>(base)));
        const uint32_t* cr = reinterpret_cast<const uint32_t*>(C.L.cur + y * 16);
        const uint32_t d0 = wr[0], d1 = wr[1], d2 = wr[2], d3 = wr[3], d4 = wr[4];
        acc = __builtin_amdgcn_sad_u8(__builtin_amdgcn_alignbyte(d1, d0, al) & m[0], cr[0] & m[0], acc);
        acc = __builtin_amdgcn_sad_u8(__builtin_amdgcn_alignbyte(d2, d1, al) & m[1], cr[1] & m[1], acc);
        acc = __builtin_amdgcn_sad_u8(__builtin_amdgcn_alignbyte(d3, d2, al) & m[2], cr[2] & m[2], acc);
        acc = __builtin_amdgcn_sad_u8(__builtin_amdgcn_alignbyte(d4, d3, al) & m[3], cr[3] & m[3], acc);
    }
    return acc;
}

// The search of extended_content_motion (src/encoder_motion.cpp:1696-1779):
// the signature scan for the top-16 cells, then per cell a 5x5 grid at step 4
// and a 5x5 refinement, stopping at the first good match.  It depends only on
// the current block, the reference and the bounds, never on other
// macroblocks' decisions, so its result is computed once per macroblock and
// step (extended_content) however often the macroblock is analysed.  The
// candidates are independent of each other, so all of
// them are searched together (kXwPerRound per round from windows staged in LDS) and
// the reference's sequential loop, with its early exit, is replayed on the
// results; the op counters count only the candidates the loop reaches.
__device__ __forceinline__ void extended_search(Ctx& C, int range, uint64_t* out_sad, int* out_dx, int* out_dy) {
    const JobDesc* J = C.J;
    const uint64_t want = cur_signature(C);
    const Bounds b = bounds_px(range);
    // cells whose offset lies inside the bounds
    const int cx_lo = imax(0, (C.x0 + b.xmin + 7) >> 3), cx_hi = imin(J->ncx - 1, (C.x0 + b.xmax) >> 3);
    const int cy_lo = imax(0, (C.y0 + b.ymin + 7) >> 3), cy_hi = imin(J->ncy - 1, (C.y0 + b.ymax) >> 3);
    // Scan: lane l keys cells l, l + 64, l + 128, ... of the cell rectangle in
    // row-major order, four per block, read straight from the lattice (it
    // stays in L2); the next block's signatures are loaded while the current
    // block is keyed.  Keys below the running 16th best are inserted into the
    // top list that lanes 0..15 hold in order.  The top sixteen are a set of
    // distinct keys, so the order in which cells are visited does not matter.
    uint64_t top = kNone;      // lane r < 16: r-th best key so far
    uint64_t thr = kNone;      // the 16th best (uniform)
    bool empty = true;         // no key has been listed yet
    uint32_t scanned = 0;
    if (cx_lo <= cx_hi && cy_lo <= cy_hi) {
        constexpr int kU = 4;
        const int nxc = cx_hi - cx_lo + 1;
        const int total = nxc * (cy_hi - cy_lo + 1);
        const int ncx = J->ncx;
        typedef const __attribute__((address_space(1))) uint64_t* gptr_u64;
        const gptr_u64 sig = (gptr_u64)(J->sig) + static_cast<size_t>(cy_lo) * ncx + cx_lo;
        const int dr = kWave / nxc, dc = kWave - dr * nxc;
        int r = C.lane / nxc, c = C.lane - (C.lane / nxc) * nxc;   // the lane's next cell
        // loads without branches (a lane past the rectangle reads its first
        // cell and keys nothing), so that the wait before keying a block
        // leaves the next block's loads in flight
        // (cell rows and columns, the lattice width and the rectangle's size
        // are below 2^24: row * width is a 24-bit product and the cell's
        // index a 32-bit offset; the lattice's top two signature bytes are
        // zero, so the chroma SAD needs no mask)
        const uint32_t uncx = static_cast<uint32_t>(ncx);
        auto fetch = [&](uint64_t (&fv)[kU], int (&fr)[kU], int (&fc)[kU], int base) VC1ME_INLINE {
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                fr[u] = r;
                fc[u] = c;
                const bool in = C.lane < total - base - u * kWave;
                const uint32_t at = __umul24(static_cast<uint32_t>(r), uncx) + static_cast<uint32_t>(c);
                fv[u] = sig[in ? at : 0u];
                c += dc;
                r += dr;
                if (c >= nxc) { c -= nxc; ++r; }
            }
        };
        const uint32_t want_lo = static_cast<uint32_t>(want), want_hi = static_cast<uint32_t>(want >> 32);
        uint32_t thr_d = 0xffffffffu >> 2;   // the distance field of thr (uniform)
        auto key_block = [&](const uint64_t (&v)[kU], const int (&vr)[kU], const int (&vc)[kU], int base) VC1ME_INLINE {
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                // the distance is the key's most significant field: a cell
                // farther than the 16th best cannot enter the list, so the
                // whole key is built only when some lane's cell might
                const uint32_t dist = 2u * __builtin_amdgcn_sad_u8(want_lo, static_cast<uint32_t>(v[u]), 0u) +
                                      3u * __builtin_amdgcn_sad_u8(want_hi, static_cast<uint32_t>(v[u] >> 32), 0u);
                const bool near = C.lane < total - base - u * kWave && dist <= thr_d;
                if (!__ballot(near)) continue;
                const int dx = (cx_lo + vc[u]) * 8 - C.x0, dy = (cy_lo + vr[u]) * 8 - C.y0;
                const bool ok = near && !(iabs(dy) <= 32 && iabs(dx) <= 32);
                const uint64_t k = ok ? key_int(dist, dx, dy) : kNone;
                uint64_t hits = __ballot(k < thr);
                if (empty && hits) {
                    // the list is empty: it becomes the (up to) sixteen least
                    // of these keys, each moved to the lane of its rank (the
                    // keys are distinct) by a forward permute
                    empty = false;
                    const int nk = __popcll(hits);
                    int rank = 0;
                    for (int j = 0; j < kWave; ++j) rank += lane_read64(k, j) < k ? 1 : 0;
                    const int to = ((hits >> C.lane) & 1) && rank < 16 ? rank * 4 : (kWave - 1) * 4;
                    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_ds_permute(to, static_cast<int>(k)));
                    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_ds_permute(to, static_cast<int>(k >> 32)));
                    top = C.lane < imin(nk, 16) ? (static_cast<uint64_t>(hi) << 32) | lo : kNone;
                    thr = lane_read64(top, 15);
                    thr_d = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(thr >> 34)));
                    continue;
                }
                while (hits) {
                    const int l = __ffsll(static_cast<long long>(hits)) - 1;
                    hits &= hits - 1;
                    const uint64_t kk = lane_read64(k, l);
                    if (kk >= thr) continue;
                    // insert kk into the ordered list of lanes 0..15
                    const int pos = __popcll(__ballot(C.lane < 16 && top < kk));
                    const uint64_t up = lane_shfl64(top, C.lane > 0 ? C.lane - 1 : 0);
                    if (C.lane < 16 && C.lane > pos) top = up;
                    if (C.lane == pos) top = kk;
                    thr = lane_read64(top, 15);
                    thr_d = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(thr >> 34)));
                }
            }
        };
        // two blocks in turn: one is keyed while the other loads
        uint64_t va[kU], vb[kU];
        int ra[kU], ca[kU], rb[kU], cb[kU];
        fetch(va, ra, ca, 0);
        for (int base = 0; base < total; base += 2 * kU * kWave) {
            fetch(vb, rb, cb, base + kU * kWave);
            key_block(va, ra, ca, base);
            if (base + kU * kWave >= total) break;
            fetch(va, ra, ca, base + 2 * kU * kWave);
            key_block(vb, rb, cb, base + kU * kWave);
        }
    }
    // the cells scanned: the rectangle less those whose offset lies within
    // +-32 px in both axes (macroblock origins are multiples of 16, so those
    // are the cell columns x0 / 8 - 4 .. x0 / 8 + 4 and rows likewise)
    if (cx_lo <= cx_hi && cy_lo <= cy_hi) {
        const int ex = imax(0, imin(cx_hi, (C.x0 >> 3) + 4) - imax(cx_lo, (C.x0 >> 3) - 4) + 1);
        const int ey = imax(0, imin(cy_hi, (C.y0 >> 3) + 4) - imax(cy_lo, (C.y0 >> 3) - 4) + 1);
        scanned = static_cast<uint32_t>((cx_hi - cx_lo + 1) * (cy_hi - cy_lo + 1) - ex * ey);
    }
    C.n_sig += scanned;
    // the top 16 by (distance, |d|, dy, dx): lane r holds the r-th
    const uint64_t ckeys = top;
    const int tn = __popcll(__ballot(C.lane < 16 && top != kNone));
    // Per candidate: the best key of its 5x5 grid at step 4 (stride-2 SAD)
    // and of the 5x5 full-SAD refinement around the grid's best, with the
    // counts of scored (not vetoed) points.  kXwPerRound candidates per
    // round; each lane scores whole points, the 25 keys of a candidate meet in
    // LDS.
    uint64_t key1 = kNone, key2 = kNone;   // lane t: results of candidate t
    uint32_t cnt1 = 0, cnt2 = 0;
    uint64_t* const kbuf = reinterpret_cast<uint64_t*>(C.L.pw + kXwPerRound * kXwBytes);
    for (int t0 = 0; t0 < tn; t0 += kXwPerRound) {
        const int cnt = imin(kXwPerRound, tn - t0);
        const int m = cnt * 25;
        xw_load(C, ckeys, t0, cnt);
        // the round's candidate offsets, lane g: candidate t0 + g
        const uint64_t myck = lane_shfl64(ckeys, imin(t0 + (C.lane & 7), 63));
        const int mydx = key_dx(myck), mydy = key_dy(myck);
        int sxv = 0, syv = 0;     // lane g < cnt: the grid winner of candidate t0 + g
        for (int stage = 0; stage < 2; ++stage) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int c = C.lane + j * kWave;
                const bool ok = c < m;
                // c / 25 = (1311 c) >> 15 for c < 4699, i / 5 = (13 i) >> 6
                // for i < 64: 24-bit products
                const int g = ok ? static_cast<int>(__umul24(static_cast<uint32_t>(c), 1311u) >> 15) : 0, i = c - g * 25;
                const int iy = static_cast<int>(__umul24(static_cast<uint32_t>(i), 13u) >> 6), ix = i - 5 * iy;
                const int cdx = __shfl(mydx, g), cdy = __shfl(mydy, g);
                const int ox = stage ? __shfl(sxv, g) : 0, oy = stage ? __shfl(syv, g) : 0;
                int dx, dy;
                if (stage == 0) {
                    dx = clampi(cdx + 4 * ix - 8, b.xmin, b.xmax);
                    dy = clampi(cdy + 4 * iy - 8, b.ymin, b.ymax);
                } else {
                    dx = clampi(ox + ix - 2, b.xmin, b.xmax);
                    dy = clampi(oy + iy - 2, b.ymin, b.ymax);
                }
                if (ok) {
                    // every point of the candidate lies within +-10 px of it:
                    // its block ends before column x0 + cdx + 26 and row
                    // y0 + cdy + 26, inside the visible picture or not (the
                    // bound of int_batch)
                    const bool may_hide = C.x0 + cdx + 26 > C.g.vw || C.y0 + cdy + 26 > C.g.vh;
                    const bool hid = may_hide && hidden_padding(C.g, C.x0, C.y0, 16, 16, dx * 4, dy * 4, 0, true);
                    uint64_t cost = kBadSad;
                    if (!hid) {
                        const int stride = stage ? 1 : 2;
                        const uint32_t sum = xw_sad_part(C, C.L.pw + g * kXwBytes, C.x0 + cdx - 10, C.y0 + cdy - 10, dx, dy,
                                                         stride, 0, 1);
                        cost = static_cast<uint64_t>(sum) * static_cast<uint64_t>(stride * stride);
                    }
                    kbuf[c] = stage ? key_seq(cost, iabs(dx) + iabs(dy), i) : key_seq(cost, 0, i);
                }
            }
            wsync();
            // lane g < cnt: reduce candidate t0 + g
            uint64_t bk = kNone;
            uint32_t nc = 0;
            if (C.lane < cnt) {
                for (int i = 0; i < 25; ++i) {
                    const uint64_t k = kbuf[C.lane * 25 + i];
                    bk = k < bk ? k : bk;
                    nc += key_cost(k) == kBadSad ? 0u : 1u;
                }
            }
            wsync();
            if (stage == 0) {
                const int i1 = static_cast<int>(bk & 0xffff);
                sxv = clampi(mydx + 4 * (i1 % 5) - 8, b.xmin, b.xmax);
                syv = clampi(mydy + 4 * (i1 / 5) - 8, b.ymin, b.ymax);
            }
            // results of candidate t0 + g to lane t0 + g
            const int rel = C.lane - t0;
            const bool mine = rel >= 0 && rel < cnt;
            const uint64_t rk = lane_shfl64(bk, mine ? rel : 0);
            const uint32_t rc = static_cast<uint32_t>(__shfl(static_cast<int>(nc), mine ? rel : 0));
            if (mine) {
                if (stage == 0) { key1 = rk; cnt1 = rc; }
                else { key2 = rk; cnt2 = rc; }
            }
        }
        // the reference loop stops at the first good candidate: skip the rest
        bool stop = false;
        {
            uint64_t best = UINT64_MAX;
            for (int u = 0; u < t0 + cnt; ++u) {
                const uint64_t fs = key_cost(lane_read64(key2, u));
                if (fs < best) best = fs;
                if (match_good(C, best)) { stop = true; break; }
            }
        }
        if (stop) break;
    }
    // replay the candidate loop in order
    uint64_t best_sad = UINT64_MAX;
    int bdx = 0, bdy = 0;
    const uint64_t npx2 = static_cast<uint64_t>((C.aw + 1) / 2) * ((C.ah + 1) / 2);
    const uint64_t npx1 = static_cast<uint64_t>(C.aw) * C.ah;
    for (int t = 0; t < tn; ++t) {
        const uint64_t ck = lane_read64(ckeys, t);
        const uint64_t s1 = lane_read64(key1, t), s2 = lane_read64(key2, t);
        C.n_sad += static_cast<uint64_t>(__builtin_amdgcn_readlane(static_cast<int>(cnt1), t)) * npx2 +
                   static_cast<uint64_t>(__builtin_amdgcn_readlane(static_cast<int>(cnt2), t)) * npx1;
        const int cdx = key_dx(ck), cdy = key_dy(ck);
        const int i1 = static_cast<int>(s1 & 0xffff);
        const int sx = clampi(cdx + 4 * (i1 % 5) - 8, b.xmin, b.xmax);
        const int sy = clampi(cdy + 4 * (i1 / 5) - 8, b.ymin, b.ymax);
        const uint64_t fs = key_cost(s2);
        if (fs < best_sad) {
            const int i = static_cast<int>(s2 & 0xffff);
            best_sad = fs;
            bdx = clampi(sx + (i % 5) - 2, b.xmin, b.xmax);
            bdy = clampi(sy + (i / 5) - 2, b.ymin, b.ymax);
        }
        if (match_good(C, best_sad)) break;
    }
    *out_sad = best_sad;
    *out_dx = bdx;
    *out_dy = bdy;
}

// extended_content_motion (src/encoder_motion.cpp:1696-1779): the search
// depends only on the macroblock, the reference and the bounds, so it runs
// the first time an analysis of this macroblock in the step needs it and is
// kept (result and op counts) for any later analysis of the macroblock.
__device__ __forceinline__ IRes extended_content(Ctx& C, IRes local, uint64_t* cand_sad, int* cand_decision) {
    const JobDesc* J = C.J;
    uint64_t best_sad;
    int2 bv;
    if (!J->ext_done[C.pos]) {
        const uint64_t sig0 = C.n_sig, sad0 = C.n_sad;
        int dx = 0, dy = 0;
        extended_search(C, J->range, &best_sad, &dx, &dy);
        bv = make_int2(dx, dy);
        C.pw_ok = false;   // the staging shared the phase-window bytes
        if (C.lane == 0) {
            J->ext_sad[C.pos] = best_sad;
            J->ext_mv[C.pos] = bv;
            J->ext_ops[2 * C.pos] = static_cast<uint32_t>(C.n_sig - sig0);
            J->ext_ops[2 * C.pos + 1] = static_cast<uint32_t>(C.n_sad - sad0);
            J->ext_done[C.pos] = 1;
        }
    } else {
        best_sad = J->ext_sad[C.pos];
        bv = J->ext_mv[C.pos];
        C.n_sig += J->ext_ops[2 * C.pos];
        C.n_sad += J->ext_ops[2 * C.pos + 1];
    }
    *cand_sad = best_sad;
    const int decision = best_sad < UINT64_MAX / 8 ? distant_decision(C, best_sad, local.sad) : 0;
    *cand_decision = decision;
    if (best_sad >= local.sad || decision < 2) return local;
    return {bv.x, bv.y, best_sad};
}

// ------------------------------------------------------------------ staged sub-pel refinement

struct SubParams {
    int bx0, by0, bw, bh, mode, range;
    bool more;
    // the guess pass's refinement: the fast stencils and the SAD+lambda*R
    // winner only (no SATD finalists), whatever the job's settings
    bool cheap = false;
};

__device__ __forceinline__ int make_even(int v) { return (v & 1) ? (v > 0 ? v - 1 : v + 1) : v; }

// (cost, |x|+|y|, y, x) key of the sub-pel stages (tie order of
// src/encoder_motion.cpp:3320-3325)
__device__ __forceinline__ uint64_t key_q(uint64_t cost, int xq, int yq) {
    return (cost21(cost) << 37) | (static_cast<uint64_t>(iabs(xq) + iabs(yq)) << 24) |
           (static_cast<uint64_t>(yq + 1024) << 13) | static_cast<uint64_t>(xq + 4096);
}
__device__ __forceinline__ int keyq_x(uint64_t k) { return static_cast<int>(k & 8191) - 4096; }
__device__ __forceinline__ int keyq_y(uint64_t k) { return static_cast<int>((k >> 13) & 2047) - 1024; }

// The candidate list of refine_motion_block_staged, one entry per lane:
// lane k < n holds entry k (vector and SAD).  n is uniform.
struct SubList {
    int n;
    int xq, yq;
    uint64_t sad;
};

// add(mv) (src/encoder_motion.cpp:3330-3343) for the m vectors (rx, ry) held
// by lanes 0..m-1: a vector already listed is found, a new one is appended
// in request order, and the new entries are scored together.  Returns the
// mask of list slots the batch refers to.
__device__ __forceinline__ uint64_t sub_add(Ctx& C, const SubParams& sp, SubList& L, int rx, int ry, int m) {
    uint64_t slots = 0;
    const int n0 = L.n;
    int n = n0;
    for (int j = 0; j < m; ++j) {
        const int vx = __builtin_amdgcn_readlane(rx, j), vy = __builtin_amdgcn_readlane(ry, j);
        const uint64_t hit = __ballot(C.lane < n && L.xq == vx && L.yq == vy);
        if (hit) {
            slots |= 1ull << (__ffsll(static_cast<long long>(hit)) - 1);
            continue;
        }
        if (C.lane == n) { L.xq = vx; L.yq = vy; }
        slots |= 1ull << n;
        ++n;
    }
    L.sad = qpel_eval(C, sp.bx0, sp.by0, sp.bw, sp.bh, L.xq, L.yq, true, n0, n - n0, sp.mode, 0, L.sad);
    L.n = n;
    return slots;
}

// minimum (SAD, tie order) key over the list slots of a batch
__device__ __forceinline__ uint64_t batch_min_key(const Ctx& C, const SubList& L, uint64_t slots) {
    return wave_min64(((slots >> C.lane) & 1) ? key_q(L.sad, L.xq, L.yq) : kNone);
}

// refine_motion_integer_only, src/encoder_motion.cpp:3247-3269
__device__ __forceinline__ int2 refine_integer_only(Ctx& C, const SubParams& sp, int sxq, int syq, uint64_t* best_sad,
                                                    uint64_t known) {
    const Bounds qb = bounds_qpel(sp.range);
    auto snap = [](int q, int lo, int hi) VC1ME_INLINE {
        q = clampi(q, lo, hi);
        const int nearest = q >= 0 ? (q + 1) / 4 : (q - 1) / 4;
        return clampi(nearest * 4, lo, hi / 4 * 4);
    };
    int ix = snap(sxq, qb.xmin, qb.xmax), iy = snap(syq, qb.ymin, qb.ymax);
    if (hidden_padding(C.g, sp.bx0, sp.by0, sp.bw, sp.bh, ix, iy, sp.mode, chroma_block(sp.bx0, sp.by0, sp.bw, sp.bh))) {
        ix = 0; iy = 0;
    }
    if (ix == sxq && iy == syq && known != UINT64_MAX) *best_sad = known;
    else *best_sad = lane_read64(qpel_eval(C, sp.bx0, sp.by0, sp.bw, sp.bh, ix, iy, true, 0, 1, sp.mode, 0, 0), 0);
    return make_int2(ix, iy);
}

// refine_qpel_block, src/encoder_motion.cpp:3596-3685: the SAD-only
// quarter-pel refinement of the motion-failure check.  Each stage scores its
// candidates at once (one per lane group) and takes the least in the
// (SAD, |mv|, y, x) order, which is the reference's sequential scan; a
// candidate equal to the seed is skipped, so the seed keeps its given cost as
// in the scan's visited list (other repeats score equal to their first visit
// and cannot win).
__device__ __forceinline__ bool qpel_key_less(uint64_t s, int x, int y, uint64_t bs, int bx, int by) {
    if (s != bs) return s < bs;
    const int m = iabs(x) + iabs(y), bm = iabs(bx) + iabs(by);
    if (m != bm) return m < bm;
    if (y != by) return y < by;
    return x < bx;
}
template <typename F>
__device__ __forceinline__ void qpel_sad_stage(Ctx& C, int n, F cand, int xmin, int xmax, int ymin, int ymax, int2 seed,
                                               uint64_t& cost, int2& best) {
    int cx = 0, cy = 0;
    if (C.lane < n) {
        cand(C.lane, cx, cy);
        cx = clampi(cx, xmin, xmax);
        cy = clampi(cy, ymin, ymax);
    }
    const bool skip = C.lane >= n || (cx == seed.x && cy == seed.y);
    const uint64_t sad = qpel_eval(C, C.x0, C.y0, 16, 16, cx, cy, !skip, 0, n, 0, 0, 0);
    for (int k = 0; k < n; ++k) {
        const int kx = __shfl(cx, k), ky = __shfl(cy, k);
        const bool ks = __shfl(skip ? 1 : 0, k) != 0;
        const uint64_t v = lane_read64(sad, k);
        if (!ks && qpel_key_less(v, kx, ky, cost, best.x, best.y)) { cost = v; best = make_int2(kx, ky); }
    }
}
__device__ __forceinline__ int2 refine_qpel_sad(Ctx& C, int range, int sxq, int syq, uint64_t known,
                                                uint64_t* best_sad) {
    const JobDesc* J = C.J;
    const int limit_q = 4 * J->subpel_range;
    if (!J->subpel || limit_q == 0 || imax(iabs(sxq), iabs(syq)) > limit_q) {
        const SubParams sp{C.x0, C.y0, 16, 16, 0, range, true};
        return refine_integer_only(C, sp, sxq, syq, best_sad, known);
    }
    const Bounds qb = bounds_qpel(range);
    const int xmin = imax(qb.xmin, -limit_q), xmax = imin(qb.xmax, limit_q);
    const int ymin = imax(qb.ymin, -limit_q), ymax = imin(qb.ymax, limit_q);
    int2 seed = make_int2(clampi(sxq, xmin, xmax), clampi(syq, ymin, ymax));
    if (hidden_padding(C.g, C.x0, C.y0, 16, 16, seed.x, seed.y, 0, true)) {
        seed = make_int2(0, 0);
        known = UINT64_MAX;
    }
    int2 best = seed;
    uint64_t cost = known != UINT64_MAX
                        ? known
                        : lane_read64(qpel_eval(C, C.x0, C.y0, 16, 16, seed.x, seed.y, true, 0, 1, 0, 0, 0), 0);
    if (J->fast) {
        if (cost == 0 && best.x == 0 && best.y == 0) { *best_sad = 0; return best; }
        qpel_sad_stage(C, 4, [&](int k, int& x, int& y) VC1ME_INLINE {
            x = seed.x + (k == 0 ? -2 : k == 1 ? 2 : 0); y = seed.y + (k == 2 ? -2 : k == 3 ? 2 : 0); },
            xmin, xmax, ymin, ymax, seed, cost, best);
        const int2 axis = best;
        const int dx = axis.x - seed.x, dy = axis.y - seed.y;
        if (dx && !dy)
            qpel_sad_stage(C, 2, [&](int k, int& x, int& y) VC1ME_INLINE { x = axis.x; y = axis.y + (k ? 2 : -2); },
                           xmin, xmax, ymin, ymax, seed, cost, best);
        else if (dy && !dx)
            qpel_sad_stage(C, 2, [&](int k, int& x, int& y) VC1ME_INLINE { x = axis.x + (k ? 2 : -2); y = axis.y; },
                           xmin, xmax, ymin, ymax, seed, cost, best);
        const int2 half = best;
        qpel_sad_stage(C, 4, [&](int k, int& x, int& y) VC1ME_INLINE {
            x = half.x + (k == 0 ? -1 : k == 1 ? 1 : 0); y = half.y + (k == 2 ? -1 : k == 3 ? 1 : 0); },
            xmin, xmax, ymin, ymax, seed, cost, best);
        const int2 qaxis = best;
        const int qdx = qaxis.x - half.x, qdy = qaxis.y - half.y;
        if (qdx && !qdy)
            qpel_sad_stage(C, 2, [&](int k, int& x, int& y) VC1ME_INLINE { x = qaxis.x; y = qaxis.y + (k ? 1 : -1); },
                           xmin, xmax, ymin, ymax, seed, cost, best);
        else if (qdy && !qdx)
            qpel_sad_stage(C, 2, [&](int k, int& x, int& y) VC1ME_INLINE { x = qaxis.x + (k ? 1 : -1); y = qaxis.y; },
                           xmin, xmax, ymin, ymax, seed, cost, best);
    } else {
        qpel_sad_stage(C, 9, [&](int k, int& x, int& y) VC1ME_INLINE {
            x = seed.x + (k % 3 - 1) * 2; y = seed.y + (k / 3 - 1) * 2; }, xmin, xmax, ymin, ymax, seed, cost, best);
        const int2 half = best;
        qpel_sad_stage(C, 9, [&](int k, int& x, int& y) VC1ME_INLINE {
            x = half.x + (k % 3 - 1); y = half.y + (k / 3 - 1); }, xmin, xmax, ymin, ymax, seed, cost, best);
    }
    *best_sad = cost;
    return best;
}

// refine_motion_block_staged, src/encoder_motion.cpp:3271-3594 (quality RATE / SATD)
__device__ __forceinline__ int2 refine_staged(Ctx& C, const SubParams& sp, int sxq, int syq, const Pred pi,
                                              uint64_t* best_sad, uint64_t known_seed) {
    const JobDesc* J = C.J;
    const int limit_q = 4 * J->subpel_range;
    if (!J->subpel || limit_q == 0 || imax(iabs(sxq), iabs(syq)) > limit_q)
        return refine_integer_only(C, sp, sxq, syq, best_sad, known_seed);
    const Bounds qb = bounds_qpel(sp.range);
    const int xmin = imax(qb.xmin, -limit_q), xmax = imin(qb.xmax, limit_q);
    const int ymin = imax(qb.ymin, -limit_q), ymax = imin(qb.ymax, limit_q);
    const bool hp = mode_halfpel(sp.mode);
    const bool chroma = chroma_block(sp.bx0, sp.by0, sp.bw, sp.bh);
    const int lane = C.lane;
    sxq = clampi(sxq, xmin, xmax);
    syq = clampi(syq, ymin, ymax);
    if (hp) { sxq = make_even(sxq); syq = make_even(syq); }
    SubList L;
    L.n = 0; L.xq = 0; L.yq = 0; L.sad = kNone;
    const bool hid = hidden_padding(C.g, sp.bx0, sp.by0, sp.bw, sp.bh, sxq, syq, sp.mode, chroma);
    if (hid || known_seed != UINT64_MAX) {
        if (lane == 0) { L.xq = sxq; L.yq = syq; L.sad = hid ? kBadSad : known_seed; }
        L.n = 1;
    } else {
        sub_add(C, sp, L, sxq, syq, 1);
    }
    if (hid) sub_add(C, sp, L, 0, 0, 1);
    uint64_t half = key_q(lane_read64(L.sad, 0), sxq, syq);
    const bool fast = J->fast != 0 || sp.cheap;
    // half-pel stencil around the seed
    {
        int rx = sxq, ry = syq, nb;
        if (fast) {
            nb = 5;
            if (lane == 1) rx = sxq - 2; else if (lane == 2) rx = sxq + 2;
            else if (lane == 3) ry = syq - 2; else if (lane == 4) ry = syq + 2;
            rx = clampi(rx, xmin, xmax); ry = clampi(ry, ymin, ymax);
        } else {
            nb = 9;
            rx = clampi(sxq + 2 * (lane % 3) - 2, xmin, xmax);
            ry = clampi(syq + 2 * ((lane / 3) % 3) - 2, ymin, ymax);
        }
        if (hp) { rx = make_even(rx); ry = make_even(ry); }
        const uint64_t m = batch_min_key(C, L, sub_add(C, sp, L, rx, ry, nb));
        half = m < half ? m : half;
    }
    int hbx = keyq_x(half), hby = keyq_y(half);
    if (fast && !(hbx == sxq && hby == syq)) {
        const int dx = hbx - sxq, dy = hby - syq;
        int nd = 0, rx = 0, ry = 0;
        if (dx && !dy) {
            nd = 2;
            rx = hbx; ry = clampi(syq + ((lane & 1) ? 2 : -2), ymin, ymax);
        } else if (dy && !dx) {
            nd = 2;
            rx = clampi(sxq + ((lane & 1) ? 2 : -2), xmin, xmax); ry = hby;
        }
        if (nd) {
            if (hp) { rx = make_even(rx); ry = make_even(ry); }
            const uint64_t m = batch_min_key(C, L, sub_add(C, sp, L, rx, ry, nd));
            half = m < half ? m : half;
            hbx = keyq_x(half); hby = keyq_y(half);
        }
    }
    if (!hp) {
        if (fast) {
            const int qcx = hbx, qcy = hby;
            int rx = qcx, ry = qcy;
            if (lane == 1) rx = qcx - 1; else if (lane == 2) rx = qcx + 1;
            else if (lane == 3) ry = qcy - 1; else if (lane == 4) ry = qcy + 1;
            rx = clampi(rx, xmin, xmax); ry = clampi(ry, ymin, ymax);
            const uint64_t sl = sub_add(C, sp, L, rx, ry, 5);
            // the quarter centre's cost, then the best of the five
            uint64_t qk = kNone;
            const uint64_t at = __ballot(lane < L.n && L.xq == qcx && L.yq == qcy);
            if (at) qk = key_q(lane_read64(L.sad, __ffsll(static_cast<long long>(at)) - 1), qcx, qcy);
            const uint64_t m = batch_min_key(C, L, sl);
            qk = m < qk ? m : qk;
            const int qbx = keyq_x(qk), qby = keyq_y(qk);
            const int dx = qbx - qcx, dy = qby - qcy;
            int n2 = 0;
            if (dx && !dy) {
                n2 = 2;
                rx = qbx; ry = clampi(qcy + ((lane & 1) ? 1 : -1), ymin, ymax);
            } else if (dy && !dx) {
                n2 = 2;
                rx = clampi(qcx + ((lane & 1) ? 1 : -1), xmin, xmax); ry = qby;
            }
            if (n2) sub_add(C, sp, L, rx, ry, n2);
        } else {
            const int rx = clampi(hbx + (lane % 3) - 1, xmin, xmax), ry = clampi(hby + ((lane / 3) % 3) - 1, ymin, ymax);
            sub_add(C, sp, L, rx, ry, 9);
        }
    }
    // sort by (SAD, tie order) and keep the shortlist: the entry of rank r
    // moves to lane r
    const int keep = fast ? 4 : 8;
    int n = L.n;
    int xq, yq;
    uint64_t sad;
    {
        const uint64_t myk = lane < n ? key_q(L.sad, L.xq, L.yq) : kNone;
        int rank = 0;
        for (int k = 0; k < n; ++k) rank += lane_read64(myk, k) < myk ? 1 : 0;
        if (n > keep) n = keep;
        // every listed entry moves to the lane of its rank (the keys, and so
        // the ranks, are distinct) by a forward permute; the lanes past the
        // list all go to lane 63, which is never a shortlist lane
        const int to = (lane < L.n ? rank : kWave - 1) * 4;
        xq = __builtin_amdgcn_ds_permute(to, L.xq);
        yq = __builtin_amdgcn_ds_permute(to, L.yq);
        sad = (static_cast<uint64_t>(static_cast<uint32_t>(__builtin_amdgcn_ds_permute(to, static_cast<int>(L.sad >> 32)))) << 32) |
              static_cast<uint32_t>(__builtin_amdgcn_ds_permute(to, static_cast<int>(L.sad)));
    }
    const uint64_t screen = lane_read64(sad, 0);
    // rates and x87 scores, one candidate per lane
    const uint64_t lm = J->lam_m;
    const int lk = J->lam_k;
    const bool in = lane < n;
    uint32_t rate = 0;
    unsigned __int128 score = 0;
    if (in) {
        int prx = pi.px, pry = pi.py;
        if (pi.hybrid) {
            const bool use_a = modular_mv_distance(J->mvrange, xq, yq, pi.ax, pi.ay) <=
                               modular_mv_distance(J->mvrange, xq, yq, pi.cx, pi.cy);
            prx = use_a ? pi.ax : pi.cx;
            pry = use_a ? pi.ay : pi.cy;
            ++rate;
        }
        rate += mv_rate_bits(J->mvrange, xq, yq, prx, pry, sp.mode, sp.more);
        score = x87_score(sad, rate, lm, lk);
    }
    // the SAD+lambda*R winner: the least (score, SAD, tie order) entry of the
    // shortlist (at most eight, lanes 0..7; lanes past it rank last).  key_q
    // orders (SAD, tie order) exactly: SADs are below 2^21 except the
    // hidden-padding veto, and the vector fields cover the quarter-pel bounds.
    const WinKey w1 = lanes8_min(WinKey{in ? 0u : 1u, score, key_q(sad, xq, yq), 0});
    const uint64_t k1 = lane_read64(w1.a, 0);
    int2 result = make_int2(keyq_x(k1), keyq_y(k1));
    if (J->quality == 2 && !sp.cheap) {
        const int rw = __ffsll(static_cast<long long>(__ballot(in && xq == result.x && yq == result.y))) - 1;
        const uint64_t rws = lane_read64(sad, rw);
        // (rws / 100) * 3 + ceil((rws % 100) * 3 / 100) = ceil(3 rws / 100),
        // in 32 bits for the SADs that fit
        const uint64_t pct3 = rws < (1ull << 30)
                                  ? static_cast<uint64_t>((static_cast<uint32_t>(rws) * 3u + 99u) / 100u)
                                  : (rws / 100) * 3 + (((rws % 100) * 3 + 99) / 100);
        const uint64_t slack = pct3 > 4 ? pct3 : 4;
        const uint64_t limit = rws > UINT64_MAX - slack ? UINT64_MAX : rws + slack;
        const bool el = in && sad < UINT64_MAX / 8 && sad <= limit;
        const uint64_t satd = qpel_eval(C, sp.bx0, sp.by0, sp.bw, sp.bh, xq, yq, el, 0, n, sp.mode, 1, kBadSad);
        const unsigned __int128 s2 = x87_score(satd, rate, lm, lk);
        // the SATD winner: eligible entries first, then (score, SATD, SAD,
        // tie order)
        const WinKey w2 = lanes8_min(WinKey{in ? (el ? 0u : 1u) : 2u, s2, satd, key_q(sad, xq, yq)});
        const uint64_t k2 = lane_read64(w2.b, 0);
        result = make_int2(keyq_x(k2), keyq_y(k2));
    }
    *best_sad = screen;
    return result;
}

// ------------------------------------------------------------------ the macroblock kernel

// median4, src/encoder_motion_compensation.cpp:522-529
__device__ __forceinline__ int median4(int a, int b, int c, int d) {
    if (a < b) {
        if (c < d) return (imin(b, d) + imax(a, c)) / 2;
        return (imin(b, c) + imax(a, d)) / 2;
    }
    if (c < d) return (imin(a, d) + imax(b, c)) / 2;
    return (imin(a, c) + imax(b, d)) / 2;
}

// The part of a macroblock's result that its right, bottom-left, bottom and
// bottom-right neighbours read: 1-MV and block vectors, intra flag and the
// "propagated" flag of the long-range stage.
struct NbState {
    int2 mv;
    int2 b[4];
    uint32_t fp;        // intra flag | propagated flag << 8
    uint32_t changed;   // (lane 0) mask of the dependents that read a changed field (requeue_mask)
};

__device__ __forceinline__ NbState nb_state(const JobDesc* J, int pos) {
    NbState s;
    s.mv = J->mvs[pos];
#pragma unroll
    for (int k = 0; k < 4; ++k) s.b[k] = J->bmv[pos * 4 + k];
    s.fp = static_cast<uint32_t>(J->flags[pos] & kFlagIntra) | (static_cast<uint32_t>(J->prop[pos]) << 8);
    s.changed = 0;
    return s;
}

// The dependents of macroblock (mx, my) whose analysis reads a field that
// differs between its states o and n: bit k for dependent k of
// queue_dependents (right, bottom-left, bottom, bottom-right).  What each
// reads of this macroblock through its neighbour snapshot:
//   right (this is its left):  the 1-MV vector (predictor_info), block 1
//     (predictor_info_mixed_1mv, the 4-MV predictor of its block 0), block 3
//     (that of its block 2), the intra flag (hybrid tests);
//   bottom (this is its top):  the vector, blocks 2 and 3 (the mixed 1-MV
//     and 4-MV predictors of its blocks 0 and 1), the intra flag;
//   bottom-left (top-right):   the vector, block 2;
//   bottom-right (top-left):   block 3 (4-MV predictor of its block 0); the
//     vector and block 3 in predictor_info and the mixed 1-MV predictor only
//     when it sits in the last column;
// and every one the propagated flag, and the vector when that flag is set
// (propagated_long_range_motion).  Block 0 is read by none of them.  A
// dependent whose read fields are all unchanged would repeat its last
// analysis exactly, so it is not queued.
__device__ __forceinline__ uint32_t requeue_mask(const NbState& o, const NbState& n, int mx, int mbw) {
    auto ne = [](int2 a, int2 b) VC1ME_INLINE { return a.x != b.x || a.y != b.y; };
    const bool dmv = ne(o.mv, n.mv), db1 = ne(o.b[1], n.b[1]), db2 = ne(o.b[2], n.b[2]), db3 = ne(o.b[3], n.b[3]);
    const bool dintra = ((o.fp ^ n.fp) & 0xffu) != 0, dprop = ((o.fp ^ n.fp) >> 8) != 0;
    const bool prop = (n.fp >> 8) != 0;
    uint32_t m = 0;
    if (dmv || db1 || db3 || dintra || dprop) m |= 1u;                      // right
    if (dmv || db2 || dprop) m |= 2u;                                        // bottom-left
    if (dmv || db2 || db3 || dintra || dprop) m |= 4u;                       // bottom
    if (db3 || dprop || (dmv && (prop || mx + 1 == mbw - 1))) m |= 8u;       // bottom-right
    return m;
}

// Lane 0 writes the macroblock's neighbour-visible state (vectors, flags,
// propagated flag) and notes in st.changed which dependents read a field
// that differs from the state the macroblock held (requeue_mask); the
// relaxation queues those.  No other workgroup writes this macroblock's
// state while it is analysed, so the state read here is the one the analysis
// started from (and no register holds it through the analysis).
__device__ __forceinline__ void commit_state(const Ctx& C, NbState& st, uint8_t flags, uint8_t prop) {
    st.changed = 0;
    if (C.lane == 0) {
        const JobDesc* J = C.J;
        const int pos = C.pos;
        st.changed = requeue_mask(nb_state(J, pos), st, C.mx, C.g.mbw);
        J->mvs[pos] = st.mv;
#pragma unroll
        for (int k = 0; k < 4; ++k) J->bmv[pos * 4 + k] = st.b[k];
        J->flags[pos] = flags;
        J->prop[pos] = prop;
    }
}

// kGuess: the relaxation's starting state only (never a result): the
// analysis without the 4-MV stage, the costliest part, whose outcome rarely
// changes a neighbour's predictors, with the diamond integer search in place
// of UMH and the fast sub-pel stencils without SATD finalists.  The extended long-range search does run:
// its result depends on no neighbour and is kept for the step (pass 0 reads
// it back), and the long-range vectors it finds reach the neighbours'
// predictors and propagated searches in pass 0 already, instead of spreading
// one macroblock per pass.
template <bool kGuess>
__device__ __forceinline__ void analyze_mb(Ctx& C, NbState& st) {
    const JobDesc* J = C.J;
    const int mode = J->mode;
    const bool allow_4mv = mode == 3 && !kGuess;
    const int one_mode = allow_4mv ? 0 : mode;
    const int range = J->range;
    const int local_range = J->local_range;
    const int pos = C.pos;

    // the local integer search's window, issued first when the analysis will
    // start with that search or the seeded one (its load overlaps the ones
    // below; the wsync after them waits for all)
    if (range > 0 && !J->seed_mvs && J->lr_mode != 2)
        win_issue(C, C.x0 - local_range, C.y0 - local_range, 16 + 2 * local_range, 16 + 2 * local_range);
    // current block to LDS, zero outside the active area (a whole block:
    // one direct global-to-LDS dword copy, lane i the dword of row i / 4 at
    // column 4 (i % 4)); the neighbours' state snapshot
    if (C.aw == 16 && C.ah == 16 && (C.g.W & 3) == 0) {
        typedef const __attribute__((address_space(1))) void* gvoid;
        typedef __attribute__((address_space(3))) void* lvoid;
        const gptr_u8 src = as_global(J->cy) + static_cast<size_t>(C.y0 + (C.lane >> 2)) * C.g.W + C.x0 + (C.lane & 3) * 4;
        __builtin_amdgcn_global_load_lds((gvoid)src, (lvoid)C.L.cur, 4, 0, 0);
    } else {
        for (int i = C.lane; i < 256; i += kWave) {
            const int x = i & 15, y = i >> 4;
            C.L.cur[i] = (x < C.aw && y < C.ah) ? J->cy[static_cast<size_t>(C.y0 + y) * C.g.W + C.x0 + x] : 0;
        }
    }
    nb_load(C);
    wsync();
    pack_sampled_rows(C);
    wsync();

    uint64_t one_sad = 0;
    int2 one = {0, 0};
    double dl_mae = -1.0, dc_mae = -1.0;
    int decision_out = 0;
    if (range == 0) {
        // zero-range mode: zero vectors with their SAD (src/encoder_motion.cpp:73-86)
        win_load(C, C.x0, C.y0, 16, 16);
        const uint64_t z = int_batch(C, 1, 1, nullptr, KEY_TOTAL, 0, [](int, int& dx, int& dy) VC1ME_INLINE { dx = 0; dy = 0; });
        // the motion-failure check at range 0: the refined and the zero-vector
        // SAD are both this SAD (src/encoder_motion.cpp:741-748)
        const uint64_t zs = key_cost(z), zpx = active_pixels(C);
        const bool zgood = J->motion_check && zpx && (zs <= zpx * 12u || (zs > 0 && zs * 10u <= zs * 9u));
        st.mv = make_int2(0, 0);
        for (int k = 0; k < 4; ++k) st.b[k] = make_int2(0, 0);
        st.fp = 0;
        commit_state(C, st, zgood ? kFlagGoodMatch : 0, 0);
        if (C.lane == 0) {
            J->screen[pos] = key_cost(z);
            J->chosen[pos] = key_cost(z);
            J->temporal[pos] = key_cost(z);
            J->ratio[pos] = -1.0;
            J->dl_mae[pos] = -1.0;
            J->dc_mae[pos] = -1.0;
            J->decision[pos] = 0;
            uint64_t* o = J->mbops + static_cast<size_t>(pos) * 5;
            o[0] = C.n_sad; o[1] = 0; o[2] = 0; o[3] = 0; o[4] = 0;
        }
        return;
    }

    if (C.lane < 4) C.L.own[C.lane] = make_int2(0, 0);
    wsync();
    const Pred rate_pi = predictor_mixed_1mv(C);
    bool was_seeded = false;
    uint8_t prop_out = 0;
    if (J->seed_mvs) {
        // the seeded core (src/encoder_motion.cpp:120-130): the seed, clamped
        // to the search bounds, starts the staged refinement; no integer or
        // long-range search, nothing propagated
        const Bounds qb = bounds_qpel(range);
        const int2 sv = J->seed_mvs[pos];
        const int sx = clampi(sv.x, qb.xmin, qb.xmax), sy = clampi(sv.y, qb.ymin, qb.ymax);
        if (J->subpel) pw_load(C, C.x0 + floor_div4(sx) - kPwMargin, C.y0 + floor_div4(sy) - kPwMargin,
                               mode_bilinear(one_mode));
        SubParams sp{C.x0, C.y0, 16, 16, one_mode, range, true};
        one = refine_staged(C, sp, sx, sy, rate_pi, &one_sad, UINT64_MAX);
    } else {
        const Pred pi = predictor_info(C);
        const int gdx = J->g->dx, gdy = J->g->dy;
        IRes im;
        bool selected_distant = false;
        // learned seeding: this macroblock's proposals, used when confident
        const float* nn = J->nn_out ? J->nn_out + static_cast<size_t>(pos) * 5 : nullptr;
        // a confident macroblock keeps the seeded result if it matches well
        // enough (seed_kept); otherwise it runs the plain search as well
        bool seeded_ok = false;
        if (nn && nn[4] >= J->nn_logit) {
            im = seeded_search(C, pi, gdx, gdy, range, nn);
            seeded_ok = seed_kept(C, im.sad);
        }
        was_seeded = seeded_ok;
        const uint64_t px = active_pixels(C);
        // long-range mode 2 tries the propagated neighbours first and skips
        // the local search on a good match; the other modes search locally
        // first (one call site per search keeps the predictor in registers)
        bool sc = seeded_ok;
        if (J->lr_mode == 2 && range > local_range && !sc) {
            const IRes pr = propagated(C, range);
            if (match_good(C, pr.sad)) { im = pr; sc = true; selected_distant = true; }
        }
        if (!sc) {
            // (the guess pass takes the cheaper diamond search: its result is
            // only the relaxation's starting state, and the diamond lands on
            // the UMH result for most macroblocks)
            if (J->algo == 1 || kGuess) im = diamond_search(C, pi, gdx, gdy, local_range);
            else im = umh_search(C, pi, gdx, gdy, local_range);
        }
        if (J->lr_mode != 2) {
            const uint64_t local_sad = im.sad;
            if (px) dl_mae = static_cast<double>(local_sad) / static_cast<double>(px);
            const bool skip = J->lr_mode == 1 && match_good(C, im.sad);
            if (range > local_range && !skip && !seeded_ok) {
                const IRes pr = propagated(C, range);
                const int dec = distant_decision(C, pr.sad, local_sad);
                if (px && pr.sad < UINT64_MAX / 8) {
                    const double mae = static_cast<double>(pr.sad) / static_cast<double>(px);
                    if (dc_mae < 0.0 || mae < dc_mae) { dc_mae = mae; decision_out = dec; }
                }
                if (key_int(pr.sad == UINT64_MAX ? kBadSad + 1 : pr.sad, pr.dx, pr.dy) < key_int(im.sad, im.dx, im.dy) &&
                    pr.sad != UINT64_MAX && dec >= 2) {
                    im = pr;
                    selected_distant = true;
                }
            }
        }
        if (range > local_range && !seeded_ok && !match_good(C, im.sad)) {
            uint64_t ext_sad = UINT64_MAX;
            int ext_dec = 0;
            const IRes ext = extended_content(C, im, &ext_sad, &ext_dec);
            if (J->lr_mode != 2 && px && ext_sad < UINT64_MAX / 8) {
                const double mae = static_cast<double>(ext_sad) / static_cast<double>(px);
                if (dc_mae < 0.0 || mae < dc_mae) { dc_mae = mae; decision_out = ext_dec; }
            }
            if (ext.dx != im.dx || ext.dy != im.dy) { im = ext; selected_distant = true; }
        }
        prop_out = selected_distant ? 1 : 0;
        if (J->motion_check) {
            // the motion-failure check (src/encoder_motion.cpp:731-748): the
            // SAD-only refinement, then the usable-match test against 12 per
            // pixel and 90% of the zero-vector SAD (u64 arithmetic as written)
            uint64_t refined = im.sad;
            if (J->subpel) pw_load(C, C.x0 + im.dx - kPwMargin, C.y0 + im.dy - kPwMargin, false);
            const int2 r = refine_qpel_sad(C, range, im.dx * 4, im.dy * 4, im.sad, &refined);
            const uint64_t zero = lane_read64(qpel_eval(C, C.x0, C.y0, 16, 16, 0, 0, true, 0, 1, 0, 0, 0), 0);
            const bool good = px && (refined <= px * 12u || (zero > 0 && refined * 10u <= zero * 9u));
            C.n_sad += wave_sum32(C.ln_sad);
            C.n_taps += wave_sum32(C.ln_taps);
            st.mv = r;
            for (int k = 0; k < 4; ++k) st.b[k] = r;
            st.fp = static_cast<uint32_t>(prop_out) << 8;
            commit_state(C, st, good ? kFlagGoodMatch : 0, prop_out);
            if (C.lane == 0) {
                J->screen[pos] = refined;
                J->chosen[pos] = refined;
                J->temporal[pos] = refined;
                J->ratio[pos] = -1.0;
                J->dl_mae[pos] = dl_mae;
                J->dc_mae[pos] = dc_mae;
                J->decision[pos] = static_cast<uint8_t>(decision_out);
                uint64_t* o = J->mbops + static_cast<size_t>(pos) * 5;
                o[0] = C.n_sad; o[1] = C.n_taps; o[2] = 0; o[3] = C.n_sig; o[4] = 0;
            }
            return;
        }
        if (J->subpel) pw_load(C, C.x0 + im.dx - kPwMargin, C.y0 + im.dy - kPwMargin, mode_bilinear(one_mode));
        SubParams sp{C.x0, C.y0, 16, 16, one_mode, range, true, kGuess};
        one = refine_staged(C, sp, im.dx * 4, im.dy * 4, rate_pi, &one_sad, im.sad);
    }
    if (hidden_padding(C.g, C.x0, C.y0, 16, 16, one.x, one.y, one_mode, true)) {
        one = make_int2(0, 0);
        one_sad = lane_read64(qpel_eval(C, C.x0, C.y0, 16, 16, 0, 0, true, 0, 1, one_mode, 0, 0), 0);
    }
    // the cost report's vector: the final 1-MV (16x16) vector, costed by
    // k_cost_final once the field is final
    if (J->cost20 && C.lane == 0) J->one_mv[pos] = one;
    const uint64_t one_screen = one_sad;
    int2 outmv = one;
    // block vectors of the decision: one (1-MV), the four 8x8 results (4-MV,
    // kept in LDS own[]) or zero (intra), held as four registers
    int2 ob0 = one, ob1 = one, ob2 = one, ob3 = one;
    uint8_t flags = 0;
    wsync();
    if (C.lane < 4) C.L.own[C.lane] = one;
    wsync();
    const double lambda = J->lambda;
    uint64_t chosen_sad = one_sad;
    uint64_t temporal_screen = one_screen;
    if (allow_4mv) {
        uint64_t four_sad = 0, four_screen = 0, extra_bits = 1;
        const Bounds qb = bounds_qpel(range);
        const int limit = 4 * J->subpel_range;
        for (int k = 0; k < 4; ++k) {
            const Pred pi4 = predictor_4mv(C, k);
            const int bx = C.x0 + (k & 1) * 8, by = C.y0 + ((k >> 1) & 1) * 8;
            auto within = [limit](int2 m) VC1ME_INLINE {
                if (imax(iabs(m.x), iabs(m.y)) > limit) { m.x = (m.x / 4) * 4; m.y = (m.y / 4) * 4; }
                return m;
            };
            // the eleven seed vectors, one per lane
            int sx = one.x, sy = one.y;
            {
                const int s = C.lane - 1;
                int x = one.x, y = one.y;
                switch (s) {
                    case 1: x = pi4.px; y = pi4.py; break;
                    case 2: x -= 8; break;
                    case 3: x += 8; break;
                    case 4: y -= 8; break;
                    case 5: y += 8; break;
                    case 6: x -= 4; break;
                    case 7: x += 4; break;
                    case 8: y -= 4; break;
                    case 9: y += 4; break;
                    default: break;
                }
                if (C.lane > 0) { sx = clampi(x, qb.xmin, qb.xmax); sy = clampi(y, qb.ymin, qb.ymax); }
                const int2 w = within(make_int2(sx, sy));
                sx = w.x; sy = w.y;
            }
            // all eleven seeds scored at once (lanes 0..10).  The reference's
            // scan keeps the first seed of least (cost, |x|+|y|): the least
            // (cost, |x|+|y|, seed) key.  It scores seed 0 and every later seed
            // that differs from the best of the seeds before it (the prefix
            // minimum), and those scorings are what the op counts record.
            const bool is_seed = C.lane < 11;
            const uint64_t scost = qpel_eval(C, bx, by, 8, 8, sx, sy, is_seed, 0, 11, 0, 0, kNone, false);
            const uint64_t skey = is_seed ? (cost21(scost) << 24) | (static_cast<uint64_t>(iabs(sx) + iabs(sy)) << 4) |
                                                static_cast<uint64_t>(C.lane)
                                          : kNone;
            const uint64_t upto = row_prefix_min64(skey);
            const uint64_t before = row_prev64(upto);
            const int prev_best = before == kNone ? 0 : static_cast<int>(before & 15);
            const int pbx = __shfl(sx, prev_best), pby = __shfl(sy, prev_best);
            if (is_seed && (C.lane == 0 || sx != pbx || sy != pby))
                count_sad_eval_lane(C, bx, by, 8, 8, make_int2(sx, sy), 0);
            const int win = static_cast<int>(lane_read64(upto, 10) & 15);
            const int2 best_seed = make_int2(__builtin_amdgcn_readlane(sx, win), __builtin_amdgcn_readlane(sy, win));
            const uint64_t seed_cost = lane_read64(scost, win);
            uint64_t bsad = 0;
            SubParams sp{bx, by, 8, 8, 3, range, true};
            const int2 mv = refine_staged(C, sp, best_seed.x, best_seed.y, pi4, &bsad, seed_cost);
            four_screen += bsad;
            wsync();
            if (C.lane == 0) C.L.own[k] = mv;
            wsync();
            four_sad += bsad;
            int apx = pi4.px, apy = pi4.py;
            if (pi4.hybrid) {
                const int da = modular_mv_distance(J->mvrange, mv.x, mv.y, pi4.ax, pi4.ay);
                const int dc = modular_mv_distance(J->mvrange, mv.x, mv.y, pi4.cx, pi4.cy);
                apx = da <= dc ? pi4.ax : pi4.cx;
                apy = da <= dc ? pi4.ay : pi4.cy;
                ++extra_bits;
            }
            if (!(mv.x == apx && mv.y == apy)) extra_bits += mvdiff_bits(J->mvrange, mv.x - apx, mv.y - apy, false, false);
        }
        // the four block results (lane 0 stored them in own[] after each block)
        const int2 f0 = C.L.own[0], f1 = C.L.own[1], f2 = C.L.own[2], f3 = C.L.own[3];
        const int cmx = median4(f0.x, f1.x, f2.x, f3.x);
        const int cmy = median4(f0.y, f1.y, f2.y, f3.y);
        const bool chroma_safe = !hidden_chroma(C.g, C.x0, C.y0, 16, 16, cmx, cmy);
        bool choose4 = false;
        if (chroma_safe) {
            const double lb = __dmul_rn(lambda, static_cast<double>(extra_bits));
            const double os = __dmul_rn(static_cast<double>(one_sad), 0.985);
            choose4 = four_sad + static_cast<uint64_t>(llround_exact(lb)) + 8 < static_cast<uint64_t>(llround_exact(os));
        }
        if (choose4) {
            ob0 = f0; ob1 = f1; ob2 = f2; ob3 = f3;
            flags |= kFlag4mvChosen;
            outmv = f0;
            chosen_sad = four_sad;
            temporal_screen = four_screen;
        } else {
            ob0 = ob1 = ob2 = ob3 = one;
            outmv = one;
            temporal_screen = one_screen;
        }
    }
    const uint64_t temporal_sad = chosen_sad;
    double ratio = -1.0;
    if (J->allow_intra) {
        // visible_block_mean (src/encoder_entropy.cpp:496-509) per 8x8 block and the DC-activity proxy
        uint64_t intra_sad = 0;
        const int lx = C.lane & 7, ly = C.lane >> 3;
        for (int k = 0; k < 4; ++k) {
            const int bx = C.x0 + (k & 1) * 8, by = C.y0 + ((k >> 1) & 1) * 8;
            const bool in = bx + lx < C.g.W && by + ly < C.g.H;
            // the current block in LDS holds every in-frame pixel of the macroblock
            const uint32_t v = in ? C.L.cur[(by - C.y0 + ly) * 16 + (bx - C.x0 + lx)] : 0u;
            const uint32_t sum = wave_sum32(v);
            const uint32_t n = wave_sum32(in ? 1u : 0u);
            const int mean = n ? static_cast<int>((sum + n / 2) / n) : 128;
            intra_sad += wave_sum32(in ? static_cast<uint32_t>(iabs(static_cast<int>(v) - mean)) : 0u);
        }
        const uint64_t intra_rate = static_cast<uint64_t>(llround_exact(__dmul_rn(lambda, 72.0)));
        ratio = temporal_screen > 0 ? __ddiv_rn(static_cast<double>(intra_sad + intra_rate), static_cast<double>(temporal_screen))
                                    : __builtin_huge_val();
        const double thr = __dmul_rn(static_cast<double>(temporal_screen), J->intra_thr);
        if (intra_sad + intra_rate < static_cast<uint64_t>(llround_exact(thr))) {
            flags |= kFlagIntra;
            outmv = make_int2(0, 0);
            ob0 = ob1 = ob2 = ob3 = make_int2(0, 0);
            chosen_sad = intra_sad;
        }
    }
    C.n_sad += wave_sum32(C.ln_sad);
    C.n_taps += wave_sum32(C.ln_taps);
    C.n_satd += wave_sum32(C.ln_satd);
    st.mv = outmv;
    st.b[0] = ob0; st.b[1] = ob1; st.b[2] = ob2; st.b[3] = ob3;
    st.fp = static_cast<uint32_t>(flags & kFlagIntra) | (static_cast<uint32_t>(prop_out) << 8);
    commit_state(C, st, flags, prop_out);
    if (C.lane == 0) {
        J->screen[pos] = temporal_screen;
        J->ratio[pos] = ratio;
        J->dl_mae[pos] = dl_mae;
        J->dc_mae[pos] = dc_mae;
        J->decision[pos] = static_cast<uint8_t>(decision_out);
        J->chosen[pos] = chosen_sad;
        J->temporal[pos] = temporal_sad;
        // this analysis's op counts replace any earlier analysis of the
        // macroblock (the relaxation schedule may analyse it more than once)
        uint64_t* o = J->mbops + static_cast<size_t>(pos) * 5;
        o[0] = C.n_sad; o[1] = C.n_taps; o[2] = C.n_satd; o[3] = C.n_sig; o[4] = was_seeded ? 1 : 0;
    }
}

__device__ __forceinline__ void init_ctx(Ctx& C, const JobDesc* J, int mx, int my, uint8_t* smem) {
    C.J = J;
    C.g = {J->W, J->H, J->vw, J->vh, J->mbw, J->mbh};
    C.ry = J->ry;
    C.mx = mx; C.my = my; C.x0 = mx * 16; C.y0 = my * 16;
    C.aw = imax(0, imin(16, J->W - C.x0));
    C.ah = imax(0, imin(16, J->H - C.y0));
    C.pos = my * J->mbw + mx;
    // the lane index passes through an empty asm: in the work-queue pass the
    // analysis runs in a loop, and values derived from the lane alone would
    // otherwise be hoisted out of it and held across the whole analysis
    // (the kernel spilled)
    int lane = threadIdx.x;
    __asm__ volatile("" : "+v"(lane));
    C.lane = lane;
    C.L = carve(smem, J->local_range);
    C.winCap = win_capacity(J->local_range);
    C.wx0 = C.wy0 = -(1 << 30); C.ww = C.wh = C.wp = 0;
    C.pw_ok = false; C.pw_bil = false; C.pwx = C.pwy = C.pwsh = 0;
    C.n_sad = C.n_taps = C.n_satd = C.n_sig = 0;
    C.ln_sad = C.ln_taps = C.ln_satd = 0;
}

template <bool kGuess = false>
__device__ __forceinline__ NbState run_mb(const JobDesc* J, int mx, int my, uint8_t* smem) {
    Ctx C;
    init_ctx(C, J, mx, my, smem);
    NbState st;
    analyze_mb<kGuess>(C, st);
    return st;
}

// The extended long-range search of every macroblock of plain jobs with a
// search range beyond the local range, one 64-lane workgroup per macroblock
// at full occupancy: used ahead of the anti-diagonal sweep, whose launches
// are latency-bound, so that the sweep only reads the kept results.
__global__ void __launch_bounds__(64) k_extended(const JobDesc* __restrict__ jobs) {
    extern __shared__ __align__(16) uint8_t smem[];
    const JobDesc* J = &jobs[blockIdx.y];
    if (J->range <= J->local_range || J->nn_out || J->seed_mvs) return;
    const int pos = static_cast<int>(blockIdx.x);
    if (pos >= J->mbw * J->mbh) return;
    Ctx C;
    init_ctx(C, J, pos % J->mbw, pos / J->mbw, smem);
    for (int i = C.lane; i < 256; i += kWave) {
        const int x = i & 15, y = i >> 4;
        C.L.cur[i] = (x < C.aw && y < C.ah) ? J->cy[static_cast<size_t>(C.y0 + y) * C.g.W + C.x0 + x] : 0;
    }
    wsync();
    uint64_t sad = UINT64_MAX;
    int dx = 0, dy = 0;
    extended_search(C, J->range, &sad, &dx, &dy);
    if (C.lane == 0) {
        J->ext_sad[pos] = sad;
        J->ext_mv[pos] = make_int2(dx, dy);
        J->ext_ops[2 * pos] = static_cast<uint32_t>(C.n_sig);
        J->ext_ops[2 * pos + 1] = static_cast<uint32_t>(C.n_sad);
        J->ext_done[pos] = 1;
    }
}

// Block costs under the search's own laws (block_sad_motion / block_satd_motion,
// hidden-padding veto included) for an arbitrary candidate list: the device
// primitive of the sub-pel stages, exposed for direct checking.
__global__ void __launch_bounds__(64) k_block_costs(const JobDesc* __restrict__ J, int x0, int y0, int bw, int bh,
                                                    int mode, int metric, const int2* __restrict__ cands, int n,
                                                    uint64_t* __restrict__ out) {
    extern __shared__ __align__(16) uint8_t smem[];
    Ctx C;
    C.J = J;
    C.g = {J->W, J->H, J->vw, J->vh, J->mbw, J->mbh};
    C.ry = J->ry;
    C.mx = x0 >> 4; C.my = y0 >> 4; C.x0 = x0; C.y0 = y0;
    C.aw = imax(0, imin(16, J->W - x0));
    C.ah = imax(0, imin(16, J->H - y0));
    C.pos = 0;
    C.lane = threadIdx.x;
    C.L = carve(smem, J->local_range);
    C.winCap = win_capacity(J->local_range);
    C.wx0 = C.wy0 = -(1 << 30); C.ww = C.wh = C.wp = 0;
    C.pw_ok = false; C.pw_bil = false; C.pwx = C.pwy = C.pwsh = 0;
    C.n_sad = C.n_taps = C.n_satd = C.n_sig = 0;
    C.ln_sad = C.ln_taps = C.ln_satd = 0;
    for (int i = C.lane; i < 256; i += kWave) {
        const int x = i & 15, y = i >> 4;
        C.L.cur[i] = (x0 + x < J->W && y0 + y < J->H) ? J->cy[static_cast<size_t>(y0 + y) * J->W + x0 + x] : 0;
    }
    wsync();
    for (int base = 0; base < n; base += kWave) {
        const int m = imin(kWave, n - base);
        int2 v = make_int2(0, 0);
        if (C.lane < m) v = cands[base + C.lane];
        const uint64_t cost = qpel_eval(C, x0, y0, bw, bh, v.x, v.y, true, 0, m, mode, metric, 0);
        if (C.lane < m) out[base + C.lane] = cost;
    }
}

__global__ void __launch_bounds__(64) k_mb_diagonal(const JobDesc* __restrict__ jobs, int diag) {
    extern __shared__ __align__(16) uint8_t smem[];
    const JobDesc* J = &jobs[blockIdx.y];
    const int mbw = J->mbw, mbh = J->mbh;
    if (diag > J->max_diag) return;
    // macroblocks on the anti-diagonal mx + 2*my == diag
    const int my_lo = imax(0, (diag - (mbw - 1) + 1) / 2);
    const int my = my_lo + static_cast<int>(blockIdx.x);
    if (my > imin(mbh - 1, diag / 2)) return;
    const int mx = diag - 2 * my;
    if (mx < 0 || mx >= mbw) return;
    run_mb(J, mx, my, smem);
}

// ------------------------------------------------------------------ relaxation schedule
//
// A macroblock's analysis is a deterministic function of the frames, the
// frame-level results (global vector, signature index, extended-search
// results, proposals) and the state of its left, top-left, top and top-right
// neighbours: their 1-MV and block vectors, intra flag and "propagated"
// flag.  Those neighbours precede it in raster order, so the dependency
// graph is acyclic and the raster-order result S is the only state F with
// F[m] = analysis(m, F[neighbours of m]) for every m.
//
// The relaxation schedule starts from a guessed state (every macroblock at
// the global vector), analyses every macroblock at once, then re-analyses,
// pass after pass, exactly the macroblocks with a neighbour whose state
// changed in the previous pass.  When a pass changes nothing, every
// macroblock's last analysis read its neighbours' final state (a neighbour
// that changed later would have queued it again), so the state is that
// fixed point: S, bit for bit.  Reads of a neighbour that is being
// rewritten in the same pass may see either value; the rewrite queues the
// reader again, so the race cannot leak into the result.


// starting state: every macroblock at the frame's global vector, no intra,
// nothing propagated; no macroblock queued yet
__global__ void __launch_bounds__(256) k_relax_init(const JobDesc* __restrict__ jobs) {
    const JobDesc* J = &jobs[blockIdx.y];
    const int pos = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (pos >= J->mbw * J->mbh) return;
    const int2 g = J->range == 0 ? make_int2(0, 0)
                   : J->seed_mvs ? J->seed_mvs[pos] : make_int2(J->g->dx * 4, J->g->dy * 4);
    J->mvs[pos] = g;
#pragma unroll
    for (int k = 0; k < 4; ++k) J->bmv[pos * 4 + k] = g;
    J->flags[pos] = 0;
    J->prop[pos] = 0;
    J->mark[pos] = 0;
}

// The starting state of the relaxation: every macroblock's guess analysis at
// once (its outputs are all overwritten by pass 0, which analyses every
// macroblock for real).
//
// Work of a (macroblocks x jobs) grid, XCD-aware: the dispatcher deals
// workgroups to the eight XCDs in turn (linear id L runs on XCD L % 8), and
// neighbouring macroblocks read mostly the same reference windows, phase
// planes and signature index.  Workgroup L therefore takes item
// band(L % 8) + L / 8 of the job-major item order, so that each XCD works
// through one contiguous band of macroblocks (and of jobs) and its L2 holds
// what its neighbours read (xcd_band_start: a bijection of [0, T)).
struct GridItem { int job, pos; };
// Waves per SIMD the macroblock kernels are compiled for (their register
// budget: 512 / waves VGPRs per lane); their LDS allows four.
#ifndef VC1ME_MB_WAVES
#define VC1ME_MB_WAVES 4
#endif

__device__ __forceinline__ GridItem xcd_banded_item() {
    const int T = static_cast<int>(gridDim.x * gridDim.y);
    const int L = static_cast<int>(blockIdx.y * gridDim.x + blockIdx.x);
    const int item = xcd_band_start(T, L % kXcds) + L / kXcds;
    const int per = static_cast<int>(gridDim.x);
    return {item / per, item - (item / per) * per};
}

// In `phases` launches by rows like pass 0 (launch `phase` takes the rows y
// with y % phases == phase), so that the guesses of later rows start from
// the guesses of the rows above them.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(VC1ME_MB_WAVES))) k_mb_guess(const JobDesc* __restrict__ jobs,
                                                                                      int phase, int phases) {
    extern __shared__ __align__(16) uint8_t smem[];
    const GridItem it = xcd_banded_item();
    const JobDesc* J = &jobs[it.job];
    const int band = it.pos / J->mbw;
    const int mx = it.pos - band * J->mbw;
    const int my = phase + phases * band;
    if (my >= J->mbh) return;
    run_mb<true>(J, mx, my, smem);
}

// Appends the dependents in `mask` (bit k: right, bottom-left, bottom,
// bottom-right) of a macroblock whose state changed to the list of the next
// pass (each once per pass: the queue stamp), as relax_entry(job,
// macroblock) entries.
__device__ __forceinline__ void queue_dependents(const JobDesc* J, int job, int pos, int next_pass, RelaxQueue Q,
                                                 uint32_t mask) {
    const int mbw = J->mbw, mbh = J->mbh;
    int32_t* next = (next_pass & 1) ? Q.wl[1] : Q.wl[0];     // a select: indexed, the pair went to scratch
    const int mx = pos % mbw, my = pos / mbw;
    const int dxs[4] = {1, -1, 0, 1}, dys[4] = {0, 1, 1, 1};
    for (int k = 0; k < 4; ++k) {
        if (!((mask >> k) & 1u)) continue;
        const int x = mx + dxs[k], y = my + dys[k];
        if (x < 0 || x >= mbw || y >= mbh) continue;
        const int d = y * mbw + x;
        if (atomicExch(&J->mark[d], static_cast<uint32_t>(next_pass)) != static_cast<uint32_t>(next_pass)) {
            const int at = atomicAdd(&Q.cnt[next_pass], 1);
            if (at < Q.cap) next[at] = static_cast<int32_t>(relax_entry(job, d));
            else atomicMax(J->err, 12);      // cannot happen: a pass lists each macroblock at most once
        }
    }
}

// Pass 0: every macroblock of every job (one workgroup each, full
// occupancy), in `phases` launches by rows: launch `phase` analyses the rows
// y with y % phases == phase of every job.  A macroblock of a later launch
// reads the state its top-left, top and top-right neighbours committed in
// this pass (the launches are ordered on the stream), so a macroblock whose
// state changed queues its bottom-left, bottom and bottom-right dependents
// only when their row came first (its own row is of the last phase, the row
// below of phase 0); its right neighbour is analysed in the same launch, in
// no set order, and is queued as in a single launch.  The dependents that
// read a changed field and were analysed before it form the list of pass 1.
// Grid: (max over jobs of mbw * ceil(mbh / phases), jobs).
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(VC1ME_MB_WAVES))) k_mb_relax0(const JobDesc* __restrict__ jobs,
                                                                                       RelaxQueue Q, int phase,
                                                                                       int phases) {
    extern __shared__ __align__(16) uint8_t smem[];
    const GridItem it = xcd_banded_item();
    const int job = it.job;
    const JobDesc* J = &jobs[job];
    const int mbw = J->mbw;
    const int band = it.pos / mbw;
    const int mx = it.pos - band * mbw;
    const int my = phase + phases * band;
    if (my >= J->mbh) return;
    const int pos = my * mbw + mx;
    const NbState after = run_mb(J, mx, my, smem);
    const uint32_t mask = phase + 1 < phases ? (after.changed & 1u) : after.changed;
    if (threadIdx.x == 0 && mask) queue_dependents(J, job, pos, 1, Q, mask);
}

// Pass p > 0 as a work queue: a grid of resident workgroups takes entries of
// the pass's list by ticket until the list is exhausted (no workgroup ever
// waits for another), and queues the dependents of the macroblocks that
// changed for pass p + 1.  The list length was final when the previous pass
// ended; an empty list ends every workgroup at once, so passes launched
// after the state settled cost only their launch.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(VC1ME_MB_WAVES))) k_mb_relaxq(const JobDesc* __restrict__ jobs,
                                                                                       RelaxQueue Q, int pass) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int n = imin(Q.cnt[pass], Q.cap);
    if (n == 0) return;
    const int32_t* list = (pass & 1) ? Q.wl[1] : Q.wl[0];
    // one wave per workgroup; the barrier at the end of each entry keeps the
    // compiler from moving one macroblock's LDS accesses past the next one's
    for (int t = __builtin_amdgcn_readfirstlane(threadIdx.x == 0 ? atomicAdd(&Q.tk[pass], 1) : 0); t < n;) {
        // the entry is the same for every lane: read it into a scalar register,
        // so the job and macroblock of the analysis are known to be uniform
        // (loaded into vector registers, every job field the analysis reads
        // became per-lane work and the kernel spilled)
        const uint32_t code = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(list[t]));
        const int job = relax_entry_job(code), pos = relax_entry_pos(code);
        const JobDesc* J = &jobs[job];
        const NbState after = run_mb(J, pos % J->mbw, pos / J->mbw, smem);
        if (threadIdx.x == 0 && after.changed) queue_dependents(J, job, pos, pass + 1, Q, after.changed);
        wsync();
        int nt = 0;
        if (threadIdx.x == 0) nt = atomicAdd(&Q.tk[pass], 1);
        t = __builtin_amdgcn_readfirstlane(__shfl(nt, 0));
    }
}

// The cost report of a finished step (judge of the learned form): per
// macroblock C20 = 20 * (SAD + lambda * R) of its final 1-MV vector, SAD by
// block_sad_motion at that vector and R against predictor_info_mixed_1mv of
// the final field with the HYBRIDPRED bit (the rate-winner score of
// src/encoder_motion.cpp:3459-3471, 20 * lambda = lam20).  Evaluated on the
// final field, so that a schedule that stops before the raster-order fixed
// point is judged by the field it returns; at the fixed point it is the score
// the last analysis saw.  One workgroup per macroblock.
__global__ void __launch_bounds__(64) k_cost_final(const JobDesc* __restrict__ jobs) {
    extern __shared__ __align__(16) uint8_t smem[];
    const JobDesc* J = &jobs[blockIdx.y];
    const int pos = static_cast<int>(blockIdx.x);
    if (!J->cost20 || J->range == 0 || J->motion_check || pos >= J->mbw * J->mbh) return;
    Ctx C;
    init_ctx(C, J, pos % J->mbw, pos / J->mbw, smem);
    for (int i = C.lane; i < 256; i += kWave) {
        const int x = i & 15, y = i >> 4;
        C.L.cur[i] = (x < C.aw && y < C.ah) ? J->cy[static_cast<size_t>(C.y0 + y) * C.g.W + C.x0 + x] : 0;
    }
    nb_load(C);
    wsync();
    const Pred rate_pi = predictor_mixed_1mv(C);
    const int one_mode = J->mode == 3 ? 0 : J->mode;
    const int2 one = J->one_mv[pos];
    const uint64_t s = lane_read64(qpel_eval(C, C.x0, C.y0, 16, 16, one.x, one.y, true, 0, 1, one_mode, 0, 0), 0);
    uint32_t r = 0;
    int prx = rate_pi.px, pry = rate_pi.py;
    if (rate_pi.hybrid) {
        const bool use_a = modular_mv_distance(J->mvrange, one.x, one.y, rate_pi.ax, rate_pi.ay) <=
                           modular_mv_distance(J->mvrange, one.x, one.y, rate_pi.cx, rate_pi.cy);
        prx = use_a ? rate_pi.ax : rate_pi.cx;
        pry = use_a ? rate_pi.ay : rate_pi.cy;
        ++r;
    }
    r += mv_rate_bits(J->mvrange, one.x, one.y, prx, pry, one_mode, true);
    if (C.lane == 0) J->cost20[pos] = 20u * s + static_cast<uint64_t>(J->lam20) * r;
}

// Start of a step: clears the per-job error flag, op counters, extended
// search marks and cost report, and the relaxation's list lengths and
// ticket counters (a fill kernel, so a captured step has no memset nodes).
__global__ void __launch_bounds__(256) k_step_begin(const JobDesc* __restrict__ jobs, int32_t* __restrict__ counters,
                                                    int ncounters) {
    const JobDesc* J = &jobs[blockIdx.y];
    const int pos = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (blockIdx.x == 0 && blockIdx.y == 0)
        for (int i = static_cast<int>(threadIdx.x); i < ncounters; i += 256) counters[i] = 0;
    if (blockIdx.x == 0) {
        if (threadIdx.x < 8) J->ops[threadIdx.x] = 0;
        if (threadIdx.x == 8) *J->err = 0;
    }
    if (pos >= J->mbw * J->mbh) return;
    J->ext_done[pos] = 0;
    if (J->cost20) {
        J->cost20[pos] = 0;
        J->one_mv[pos] = make_int2(0, 0);
    }
}

// End of a step, one workgroup per job: the totals of the final macroblock
// analyses' op counts, the frame aggregates of analyze_p_picture_core
// (src/encoder_motion.cpp:416-441: selected SAD, temporal SAD, active
// pixels, moved / fractional-chroma / 4-MV / intra macroblocks) as exact
// integers, the global vector and the error flag, into the job's summary;
// with J->pack set, every per-macroblock output field in the packed layout
// the host copies back in one transfer.  Reads the state only, so it can run
// again after more passes.
__global__ void __launch_bounds__(256) k_step_end(const JobDesc* __restrict__ jobs, StepSummary* __restrict__ out) {
    const int job = static_cast<int>(blockIdx.x);
    const JobDesc* J = &jobs[job];
    const int mbw = J->mbw, mbs = mbw * J->mbh;
    constexpr int kN = 12;   // 5 op sums, then sad, rate, px, moved, frac, four, intra
    unsigned long long s[kN] = {};
    PackedOut P = pack_layout(J->pack, mbs);
    for (int pos = static_cast<int>(threadIdx.x); pos < mbs; pos += 256) {
#pragma unroll
        for (int k = 0; k < 5; ++k) s[k] += J->mbops[static_cast<size_t>(pos) * 5 + k];
        const int mx = pos % mbw, my = pos / mbw;
        const uint64_t px = static_cast<uint64_t>(imax(0, imin(16, J->W - mx * 16))) *
                            static_cast<uint64_t>(imax(0, imin(16, J->H - my * 16)));
        const uint8_t f = J->flags[pos];
        bool moved = false, frac = false;
        int2 b[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            b[k] = J->bmv[pos * 4 + k];
            moved |= b[k].x != 0 || b[k].y != 0;
            frac |= (b[k].x % 8) != 0 || (b[k].y % 8) != 0;
        }
        s[5] += J->chosen[pos];
        s[6] += J->temporal[pos];
        s[7] += px;
        s[8] += moved;
        s[9] += frac;
        s[10] += (f & kFlag4mvChosen) ? 1 : 0;
        s[11] += (f & kFlagIntra) ? 1 : 0;
        if (J->pack) {
            P.mvs[pos] = J->mvs[pos];
#pragma unroll
            for (int k = 0; k < 4; ++k) P.bmv[pos * 4 + k] = b[k];
            P.screen[pos] = J->screen[pos];
            P.ratio[pos] = J->ratio[pos];
            P.dl[pos] = J->dl_mae[pos];
            P.dc[pos] = J->dc_mae[pos];
            P.four[pos] = ((f & kFlag4mvChosen) && !(f & kFlagIntra)) ? 1 : 0;
            P.intra[pos] = (f & kFlagIntra) ? 1 : 0;
            P.decision[pos] = J->decision[pos];
        }
    }
    __shared__ unsigned long long part[kN][256 / kWave];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < kN; ++k) {
        unsigned long long v = s[k];
        for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
        if (lane == 0) part[k][w] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t[kN] = {};
        for (int k = 0; k < kN; ++k)
            for (int j = 0; j < 256 / kWave; ++j) t[k] += part[k][j];
        StepSummary& o = out[job];
        o.err = *J->err;
        o.gdx = J->g->dx;
        o.gdy = J->g->dy;
        o.global_sad = J->g->full_sad;
        // ops[4] (frame SAD pixels) comes from the global search; the others
        // are this step's final analyses
        for (int k = 0; k < 8; ++k) o.ops[k] = J->ops[k];
        for (int k = 0; k < 4; ++k) o.ops[k] += t[k];
        o.ops[5] += t[4];
        o.total_sad = t[5];
        o.rate_total = t[6];
        o.total_px = t[7];
        o.moved = t[8];
        o.frac = t[9];
        o.four = t[10];
        o.intra = t[11];
    }
}

// ------------------------------------------------------------------ global search

// Frame SAD partial sums (frame_sad, src/encoder_motion.cpp:1565-1575):
// block = (row chunk, point group, job), 256 threads, 8 sampled rows per
// block and up to kFsPoints points per group.  Each thread reads its current
// samples once and the reference at every point of the group; unsampled
// (stride 1) rows of a width divisible by 4 are read as dwords, four pixels
// per v_sad_u8, with the clamped byte reads only where the reference dword
// would leave the row.
constexpr int kFsPoints = 8;
__global__ void __launch_bounds__(256) k_frame_sads(const JobDesc* __restrict__ jobs, int per) {
    const JobDesc* J = &jobs[blockIdx.z];
    GlobalState* g = J->g;
    const int p0 = static_cast<int>(blockIdx.y) * per;
    const int np = imin(imin(per, kFsPoints), g->n - p0);
    if (np <= 0) return;
    const int s = g->stride;
    const int W = J->W, H = J->H;
    const int rows = (H + s - 1) / s, cols = (W + s - 1) / s;
    const int r0 = static_cast<int>(blockIdx.x) * 8;
    if (r0 >= rows) return;
    const int nr = imin(rows, r0 + 8) - r0;
    int pdx[kFsPoints], pdy[kFsPoints];
#pragma unroll
    for (int k = 0; k < kFsPoints; ++k) {
        pdx[k] = k < np ? g->pts[2 * (p0 + k)] : 0;
        pdy[k] = k < np ? g->pts[2 * (p0 + k) + 1] : 0;
    }
    uint32_t acc[kFsPoints];
#pragma unroll
    for (int k = 0; k < kFsPoints; ++k) acc[k] = 0;
    const gptr_u8 cy = as_global(J->cy), ry = as_global(J->ry);
    typedef const __attribute__((address_space(1))) uint32_t* gptr_u32;
    const bool dwords = s == 1 && (W & 3) == 0;
    const int units = dwords ? W >> 2 : cols;     // per row: dwords or samples
    const int total = nr * units;
    int rr = static_cast<int>(threadIdx.x) / units, u = static_cast<int>(threadIdx.x) - rr * units;
    for (int i = static_cast<int>(threadIdx.x); i < total; i += 256) {
        const int y = (r0 + rr) * s;
        if (dwords) {
            const int x = 4 * u;
            const uint32_t c = *reinterpret_cast<gptr_u32>(cy + static_cast<size_t>(y) * W + x);
#pragma unroll
            for (int k = 0; k < kFsPoints; ++k) {
                if (k >= np) break;
                const gptr_u8 row = ry + static_cast<size_t>(clampi(y + pdy[k], 0, H - 1)) * W;
                const int a = x + pdx[k];
                uint32_t r;
                if (a >= 0 && (a & ~3) + 8 <= W) {
                    const gptr_u32 q = reinterpret_cast<gptr_u32>(row + (a & ~3));
                    r = __builtin_amdgcn_alignbyte(q[1], q[0], a & 3);
                } else {
                    r = 0;
#pragma unroll
                    for (int j = 0; j < 4; ++j) r |= static_cast<uint32_t>(row[clampi(a + j, 0, W - 1)]) << (8 * j);
                }
                acc[k] = __builtin_amdgcn_sad_u8(r, c, acc[k]);
            }
        } else {
            const int x = u * s;
            const int c = cy[static_cast<size_t>(y) * W + x];
#pragma unroll
            for (int k = 0; k < kFsPoints; ++k) {
                if (k >= np) break;
                const int sy = clampi(y + pdy[k], 0, H - 1);
                acc[k] += static_cast<uint32_t>(iabs(c - static_cast<int>(ry[static_cast<size_t>(sy) * W +
                                                                               clampi(x + pdx[k], 0, W - 1)])));
            }
        }
        u += 256;
        while (u >= units) { u -= units; ++rr; }
    }
    // block sums per point
    __shared__ uint32_t red[kFsPoints][4];
    const int wave = static_cast<int>(threadIdx.x) >> 6;
#pragma unroll
    for (int k = 0; k < kFsPoints; ++k) {
        if (k >= np) break;
        const uint32_t t = wave_sum32(acc[k]);
        if ((threadIdx.x & 63) == 0) red[k][wave] = t;
    }
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < np) {
        const int k = static_cast<int>(threadIdx.x);
        const uint64_t t = static_cast<uint64_t>(red[k][0]) + red[k][1] + red[k][2] + red[k][3];
        atomicAdd(reinterpret_cast<unsigned long long*>(&g->cost[p0 + k]), static_cast<unsigned long long>(t));
    }
    if (threadIdx.x == 0)
        atomicAdd(&J->ops[4], static_cast<unsigned long long>(np) * static_cast<unsigned long long>(nr) *
                                  static_cast<unsigned long long>(cols));
}

__device__ __forceinline__ bool g_better(uint64_t ac, int ax, int ay, uint64_t bc, int bx, int by) {
    if (ac != bc) return ac < bc;
    const int am = iabs(ax) + iabs(ay), bm = iabs(bx) + iabs(by);
    if (am != bm) return am < bm;
    if (ay != by) return ay < by;
    return ax < bx;
}

// The extended search's sorted candidate list (src/encoder_motion.cpp:2036-2041)
// held in registers while a stage is replayed: consider() keeps the lower
// cost of a repeated point (without re-sorting, as the reference does); a new
// point is appended, the list sorted by the strict order of g_better (points
// are distinct, so the order is unique) and truncated to 8.  `sorted` tracks
// whether the list is in that order, so that an append to a sorted list is a
// single insertion; only a lowered repeat can leave it unsorted.
struct GTop {
    uint64_t c[9];
    int x[9], y[9];
    int n;
    bool sorted;
};

__device__ __forceinline__ void gtop_load(GTop& t, const GlobalState* g) {
    t.n = g->tn;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        t.c[i] = i < t.n ? g->tcost[i] : UINT64_MAX;
        t.x[i] = i < t.n ? g->tdx[i] : 0;
        t.y[i] = i < t.n ? g->tdy[i] : 0;
    }
    t.sorted = true;
#pragma unroll
    for (int i = 1; i < 9; ++i)
        if (i < t.n && g_better(t.c[i], t.x[i], t.y[i], t.c[i - 1], t.x[i - 1], t.y[i - 1])) t.sorted = false;
}

__device__ __forceinline__ void gtop_store(const GTop& t, GlobalState* g) {
    g->tn = t.n;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        if (i < t.n) { g->tcost[i] = t.c[i]; g->tdx[i] = t.x[i]; g->tdy[i] = t.y[i]; }
    }
}

__device__ __forceinline__ void g_consider(GTop& t, uint64_t cost, int dx, int dy) {
    bool found = false;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        if (i < t.n && t.x[i] == dx && t.y[i] == dy) {
            if (cost < t.c[i]) {
                t.c[i] = cost;
                if (i > 0 && g_better(t.c[i], t.x[i], t.y[i], t.c[i - 1], t.x[i - 1], t.y[i - 1])) t.sorted = false;
            }
            found = true;
        }
    }
    if (found) return;
    if (t.sorted) {
        // insertion at the new point's rank, shifting the worse entries up
        int p = 0;
#pragma unroll
        for (int i = 0; i < 9; ++i)
            if (i < t.n && g_better(t.c[i], t.x[i], t.y[i], cost, dx, dy)) ++p;
#pragma unroll
        for (int i = 8; i >= 1; --i)
            if (i > p && i <= t.n) { t.c[i] = t.c[i - 1]; t.x[i] = t.x[i - 1]; t.y[i] = t.y[i - 1]; }
#pragma unroll
        for (int i = 0; i < 9; ++i)
            if (i == p) { t.c[i] = cost; t.x[i] = dx; t.y[i] = dy; }
        t.n = t.n + 1 > 8 ? 8 : t.n + 1;
        return;
    }
    // append at n (n <= 8), then sort the n + 1 entries (odd-even
    // transposition: 9 rounds sort any 9 entries)
#pragma unroll
    for (int i = 0; i < 9; ++i)
        if (i == t.n) { t.c[i] = cost; t.x[i] = dx; t.y[i] = dy; }
    const int n = t.n + 1;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
#pragma unroll
        for (int i = r & 1; i + 1 < 9; i += 2) {
            if (i + 1 < n && g_better(t.c[i + 1], t.x[i + 1], t.y[i + 1], t.c[i], t.x[i], t.y[i])) {
                const uint64_t tc = t.c[i]; t.c[i] = t.c[i + 1]; t.c[i + 1] = tc;
                const int tx = t.x[i]; t.x[i] = t.x[i + 1]; t.x[i + 1] = tx;
                const int ty = t.y[i]; t.y[i] = t.y[i + 1]; t.y[i + 1] = ty;
            }
        }
    }
    t.n = n > 8 ? 8 : n;
    t.sorted = true;
}

__device__ __forceinline__ void g_set_stage(GlobalState* g, int stride) {
    g->stride = stride;
    for (int i = 0; i < g->n; ++i) g->cost[i] = 0;
}

__device__ __forceinline__ int g_range(const JobDesc* J) { return imin(J->local_range, 32); }

// Stage 0: the first point list.  Extended search (range 32): cross and
// multi-hex at stride 16; grid search (range 1..31): coarse grid at stride 4.
__global__ void k_global_begin(const JobDesc* __restrict__ jobs, int njobs) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= njobs) return;
    const JobDesc* J = &jobs[j];
    GlobalState* g = J->g;
    g->tn = 0;
    g->n = 0;
    g->dx = 0; g->dy = 0; g->full_sad = 0;
    if (J->range == 0 || J->seed_mvs) return;   // the seeded core has no global search
    const int R = g_range(J);
    if (R <= 0) {
        g->pts[0] = 0; g->pts[1] = 0; g->n = 1;
        g_set_stage(g, 1);
        return;
    }
    int n = 0;
    auto push = [&](int dx, int dy) VC1ME_INLINE { g->pts[2 * n] = dx; g->pts[2 * n + 1] = dy; ++n; };
    if (R > 31) {
        push(0, 0);
        const int step = R >= 160 ? 16 : 8;
        for (int d = step; d <= R; d += step) { push(-d, 0); push(d, 0); push(0, -d); push(0, d); }
        for (int r = step; r <= R; r += step) {
            const int h = (r + 1) / 2;
            push(-r, -h); push(-r, h); push(r, -h); push(r, h);
            push(-h, -r); push(h, -r); push(-h, r); push(h, r);
        }
        for (int i = 0; i < n; ++i) {
            g->pts[2 * i] = clampi(g->pts[2 * i], -R, R);
            g->pts[2 * i + 1] = clampi(g->pts[2 * i + 1], -R, R);
        }
        g->n = n;
        g_set_stage(g, 16);
    } else {
        const int cs = R >= 4 ? 4 : (R >= 2 ? 2 : 1);
        for (int dy = -R; dy <= R; dy += cs)
            for (int dx = -R; dx <= R; dx += cs) push(dx, dy);
        push(0, 0);
        for (int d = -R; d <= R; d += cs) { push(R, d); push(d, R); }
        g->n = n;
        g_set_stage(g, 4);
    }
}

// Replay the sequential bookkeeping of one stage on the computed costs and
// emit the next stage (src/encoder_motion.cpp:2043-2073 and :2094-2140).
// One 64-thread block per job: the stage's points and costs are staged in LDS
// by the block, then one thread replays the stage (its loads are then LDS
// hits instead of a chain of dependent global loads).
__global__ void __launch_bounds__(64) k_global_step(const JobDesc* __restrict__ jobs, int njobs, int stage) {
    __shared__ uint64_t sc[kGlobalMaxPts];
    __shared__ int32_t sp[2 * kGlobalMaxPts];
    const int j = static_cast<int>(blockIdx.x);
    if (j >= njobs) return;
    const JobDesc* J = &jobs[j];
    GlobalState* g = J->g;
    const int np_in = J->range == 0 ? 0 : imin(g->n, kGlobalMaxPts);
    for (int i = static_cast<int>(threadIdx.x); i < np_in; i += static_cast<int>(blockDim.x)) {
        sc[i] = g->cost[i];
        sp[2 * i] = g->pts[2 * i];
        sp[2 * i + 1] = g->pts[2 * i + 1];
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (J->range == 0 || np_in == 0) return;
    const int R = g_range(J);
    const int s = g->stride;
    const uint64_t mul = static_cast<uint64_t>(s) * s;
    if (R <= 0) {
        g->dx = 0; g->dy = 0; g->full_sad = g->cost[0];
        g->n = 0;
        return;
    }
    if (R > 31) {
        const int step = R >= 160 ? 16 : 8;
        if (stage == 0 || stage == 1 || stage == 2) {
            GTop t;
            gtop_load(t, g);
            const int np = np_in;
            for (int i = 0; i < np; ++i) g_consider(t, sc[i] * (s > 1 ? mul : 1), sp[2 * i], sp[2 * i + 1]);
            gtop_store(t, g);
            int n = 0;
            if (stage == 0) {
                // densify the best four of the coarse list at stride 8
                const int tn = g->tn;
                int cx[4], cy[4];
                for (int i = 0; i < imin(4, tn); ++i) { cx[i] = g->tdx[i]; cy[i] = g->tdy[i]; }
                for (int i = 0; i < imin(4, tn); ++i)
                    for (int oy = -step; oy <= step; oy += 4)
                        for (int ox = -step; ox <= step; ox += 4) {
                            g->pts[2 * n] = clampi(cx[i] + ox, -R, R);
                            g->pts[2 * n + 1] = clampi(cy[i] + oy, -R, R);
                            ++n;
                        }
                g->n = n;
                g_set_stage(g, 8);
            } else if (stage == 1) {
                const int mdx = g->tdx[0], mdy = g->tdy[0];
                for (int oy = -4; oy <= 4; ++oy)
                    for (int ox = -4; ox <= 4; ++ox) {
                        g->pts[2 * n] = clampi(mdx + ox, -R, R);
                        g->pts[2 * n + 1] = clampi(mdy + oy, -R, R);
                        ++n;
                    }
                g->n = n;
                g_set_stage(g, 4);
            } else {
                const int fdx = g->tdx[0], fdy = g->tdy[0];
                for (int oy = -1; oy <= 1; ++oy)
                    for (int ox = -1; ox <= 1; ++ox) {
                        g->pts[2 * n] = clampi(fdx + ox, -R, R);
                        g->pts[2 * n + 1] = clampi(fdy + oy, -R, R);
                        ++n;
                    }
                g->n = n;
                g_set_stage(g, 1);
            }
        } else {
            uint64_t bc = UINT64_MAX;
            int bx = 0, by = 0;
            for (int i = 0; i < np_in; ++i)
                if (g_better(sc[i], sp[2 * i], sp[2 * i + 1], bc, bx, by)) {
                    bc = sc[i]; bx = sp[2 * i]; by = sp[2 * i + 1];
                }
            g->dx = bx; g->dy = by; g->full_sad = bc;
            g->n = 0;
        }
        return;
    }
    // grid path (range <= 31)
    const int cs = R >= 4 ? 4 : (R >= 2 ? 2 : 1);
    if (stage == 0) {
        // sorted top-8 of every coarse point (push, sort, truncate)
        g->tn = 0;
        for (int i = 0; i < np_in; ++i) {
            const uint64_t c = sc[i] * mul;
            const int x = sp[2 * i], y = sp[2 * i + 1];
            g->tcost[g->tn] = c; g->tdx[g->tn] = x; g->tdy[g->tn] = y; g->tn++;
            for (int a = g->tn - 1; a > 0 && g_better(g->tcost[a], g->tdx[a], g->tdy[a], g->tcost[a - 1], g->tdx[a - 1], g->tdy[a - 1]); --a) {
                const uint64_t tc = g->tcost[a]; g->tcost[a] = g->tcost[a - 1]; g->tcost[a - 1] = tc;
                const int tx = g->tdx[a]; g->tdx[a] = g->tdx[a - 1]; g->tdx[a - 1] = tx;
                const int ty = g->tdy[a]; g->tdy[a] = g->tdy[a - 1]; g->tdy[a - 1] = ty;
            }
            if (g->tn > 8) g->tn = 8;
        }
        // neighbourhoods at stride 2 in first-visit order (63x63 seen table)
        uint64_t seen[63 * 63 / 64 + 1];
        for (int i = 0; i < 63 * 63 / 64 + 1; ++i) seen[i] = 0;
        int n = 0;
        for (int t = 0; t < g->tn; ++t) {
            const int cx = g->tdx[t], cy = g->tdy[t];
            for (int dy = imax(-R, cy - cs + 1); dy <= imin(R, cy + cs - 1); ++dy)
                for (int dx = imax(-R, cx - cs + 1); dx <= imin(R, cx + cs - 1); ++dx) {
                    const int key = (dy + 31) * 63 + (dx + 31);
                    if ((seen[key >> 6] >> (key & 63)) & 1) continue;
                    seen[key >> 6] |= 1ull << (key & 63);
                    if (n < kGlobalMaxPts) {
                        g->pts[2 * n] = dx; g->pts[2 * n + 1] = dy;
                    }
                    ++n;
                }
        }
        if (n > kGlobalMaxPts) { atomicMax(J->err, 7); n = kGlobalMaxPts; }
        g->n = n;
        g_set_stage(g, 2);
    } else if (stage == 1) {
        uint64_t bc = UINT64_MAX;
        int bx = 0, by = 0;
        for (int i = 0; i < np_in; ++i) {
            const uint64_t c = sc[i] * mul;
            const int x = sp[2 * i], y = sp[2 * i + 1];
            const int cm = iabs(x) + iabs(y), bm = iabs(bx) + iabs(by);
            if (c < bc || (c == bc && cm < bm)) { bc = c; bx = x; by = y; }
        }
        g->pts[0] = bx; g->pts[1] = by;
        g->n = 1;
        g_set_stage(g, 1);
    } else {
        g->dx = g->pts[0]; g->dy = g->pts[1]; g->full_sad = g->cost[0];
        g->n = 0;
    }
}

// ------------------------------------------------------------------ long-range index

// long_range_signature of every 8-px lattice cell of the reference
// (src/encoder_motion.cpp:1578-1625), packed 6 bytes per cell: the rounded
// means of the four 8x8 luma blocks of the cell's 16x16 area and of the 8x8
// chroma blocks at half its position, each over the pixels inside the plane
// (128 when none is).  A workgroup takes a tile of 64 x 4 cells: it first sums
// the aligned 8x8 luma blocks and the 4x4 chroma sub-blocks the tile's cells
// cover (65 x 5 of each, every pixel read once, rows as dwords where the
// plane allows), then each thread forms one cell's means from those sums.
constexpr int kSigTileX = 64, kSigTileY = 4;
constexpr int kSigBx = kSigTileX + 1, kSigBy = kSigTileY + 1;

// sum of the pixels of p (pitch pw, pw x ph) in [x0, x0 + n) x [y0, y0 + n)
__device__ __forceinline__ uint32_t block_sum(gptr_u8 p, int pw, int ph, int x0, int y0, int n, bool dwords) {
    typedef const __attribute__((address_space(1))) uint32_t* gptr_u32;
    uint32_t sum = 0;
    const int rows = imin(n, ph - y0);
    if (dwords && x0 + n <= pw) {
        for (int y = 0; y < rows; ++y) {
            const gptr_u32 q = reinterpret_cast<gptr_u32>(p + static_cast<size_t>(y0 + y) * pw + x0);
            for (int k = 0; k < n / 4; ++k) sum = __builtin_amdgcn_sad_u8(q[k], 0u, sum);
        }
    } else {
        const int cols = imin(n, pw - x0);
        for (int y = 0; y < rows; ++y)
            for (int x = 0; x < cols; ++x) sum += p[static_cast<size_t>(y0 + y) * pw + x0 + x];
    }
    return sum;
}
__device__ __forceinline__ uint32_t rounded_mean(uint32_t sum, int n) {
    return n > 0 ? (sum + static_cast<uint32_t>(n) / 2) / static_cast<uint32_t>(n) : 128u;
}

__global__ void __launch_bounds__(256) k_signatures(const JobDesc* __restrict__ jobs) {
    const JobDesc* J = &jobs[blockIdx.z];
    if (J->range <= J->local_range || J->range == 0) return;
    const int ncx = J->ncx, ncy = J->ncy;
    const int cx0 = static_cast<int>(blockIdx.x) * kSigTileX, cy0 = static_cast<int>(blockIdx.y) * kSigTileY;
    if (cx0 >= ncx || cy0 >= ncy) return;
    const int W = J->W, H = J->H, cw = W / 2, ch = H / 2;
    __shared__ uint32_t ls[kSigBy][kSigBx], us[kSigBy][kSigBx], vs[kSigBy][kSigBx];
    const gptr_u8 py = as_global(J->ry), pu = as_global(J->ru), pv = as_global(J->rv);
    const bool ydw = (W & 3) == 0 && (reinterpret_cast<uintptr_t>(J->ry) & 3) == 0;
    const bool udw = (cw & 3) == 0 && (reinterpret_cast<uintptr_t>(J->ru) & 3) == 0;
    const bool vdw = (cw & 3) == 0 && (reinterpret_cast<uintptr_t>(J->rv) & 3) == 0;
    for (int i = static_cast<int>(threadIdx.x); i < kSigBx * kSigBy; i += 256) {
        const int by = i / kSigBx, bx = i - by * kSigBx;
        const int gx = cx0 + bx, gy = cy0 + by;        // luma block / chroma sub-block index
        uint32_t l = 0, u = 0, v = 0;
        if (8 * gx < W && 8 * gy < H) l = block_sum(py, W, H, 8 * gx, 8 * gy, 8, ydw);
        if (4 * gx < cw && 4 * gy < ch) {
            u = block_sum(pu, cw, ch, 4 * gx, 4 * gy, 4, udw);
            v = block_sum(pv, cw, ch, 4 * gx, 4 * gy, 4, vdw);
        }
        ls[by][bx] = l; us[by][bx] = u; vs[by][bx] = v;
    }
    __syncthreads();
    const int tx = static_cast<int>(threadIdx.x) & (kSigTileX - 1), ty = static_cast<int>(threadIdx.x) / kSigTileX;
    const int cx = cx0 + tx, cy = cy0 + ty;
    if (cx >= ncx || cy >= ncy) return;
    const int x0 = cx * 8, y0 = cy * 8;
    auto lum_n = [&](int bx, int by) VC1ME_INLINE { return imax(0, imin(8, W - bx)) * imax(0, imin(8, H - by)); };
    const int cn = imax(0, imin(8, cw - x0 / 2)) * imax(0, imin(8, ch - y0 / 2));
    uint64_t s = 0;
    s |= static_cast<uint64_t>(rounded_mean(ls[ty][tx], lum_n(x0, y0)));
    s |= static_cast<uint64_t>(rounded_mean(ls[ty][tx + 1], lum_n(x0 + 8, y0))) << 8;
    s |= static_cast<uint64_t>(rounded_mean(ls[ty + 1][tx], lum_n(x0, y0 + 8))) << 16;
    s |= static_cast<uint64_t>(rounded_mean(ls[ty + 1][tx + 1], lum_n(x0 + 8, y0 + 8))) << 24;
    s |= static_cast<uint64_t>(rounded_mean(us[ty][tx] + us[ty][tx + 1] + us[ty + 1][tx] + us[ty + 1][tx + 1], cn)) << 32;
    s |= static_cast<uint64_t>(rounded_mean(vs[ty][tx] + vs[ty][tx + 1] + vs[ty + 1][tx] + vs[ty + 1][tx + 1], cn)) << 40;
    J->sig[static_cast<size_t>(cy) * ncx + cx] = s;
}

// ------------------------------------------------------------------ batch scorer

// One lane per candidate (src/shaders/vulkan_motion_cost.comp:178-183):
// sampled SAD or SATD4x4 of one block, clamped fetch, 32-bit accumulation.
__global__ void __launch_bounds__(64) k_score_batch(const uint8_t* __restrict__ cur, const uint8_t* __restrict__ ref,
                                                    int width, int height, int x0, int y0, int bw, int bh, int ss,
                                                    int bilinear, int rnd, int metric, const int2* __restrict__ cands,
                                                    int n, uint32_t* __restrict__ costs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int2 mv = cands[i];
    const int ix = floor_div4(mv.x), iy = floor_div4(mv.y);
    const int hm = mv.x - ix * 4, vm = mv.y - iy * 4;
    const int aw = imax(0, imin(bw, width - x0)), ah = imax(0, imin(bh, height - y0));
    auto S = [&](int x, int y) VC1ME_INLINE {
        return static_cast<int>(ref[static_cast<size_t>(clampi(y, 0, height - 1)) * width + clampi(x, 0, width - 1)]);
    };
    auto pred = [&](int x, int y) VC1ME_INLINE -> int {
        const int sx = x + ix, sy = y + iy;
        if (bilinear) {
            const int a = S(sx, sy);
            if (!hm && !vm) return a;
            const int b = S(sx + 1, sy), c = S(sx, sy + 1);
            if (hm && vm) return (a + b + c + S(sx + 1, sy + 1) + (rnd ? 1 : 2)) >> 2;
            if (hm) return (a + b + (rnd ? 0 : 1)) >> 1;
            return (a + c + (rnd ? 0 : 1)) >> 1;
        }
        if (!hm && !vm) return S(sx, sy);
        int v;
        if (hm && vm) {
            const int shift = ((hm == 2 ? 1 : 5) + (vm == 2 ? 1 : 5)) >> 1;
            const int r1 = (1 << (shift - 1)) + (rnd ? 1 : 0) - 1;
            int t[4];
            for (int k = -1; k <= 2; ++k)
                t[k + 1] = arshift(mspel(S(sx + k, sy - 1), S(sx + k, sy), S(sx + k, sy + 1), S(sx + k, sy + 2), vm) + r1, shift);
            v = arshift(mspel(t[0], t[1], t[2], t[3], hm) + 64 - (rnd ? 1 : 0), 7);
        } else if (vm) {
            v = arshift(mspel(S(sx, sy - 1), S(sx, sy), S(sx, sy + 1), S(sx, sy + 2), vm) + (vm == 2 ? 7 + (rnd ? 1 : 0) : 31 + (rnd ? 1 : 0)),
                        vm == 2 ? 4 : 6);
        } else {
            v = arshift(mspel(S(sx - 1, sy), S(sx, sy), S(sx + 1, sy), S(sx + 2, sy), hm) + (hm == 2 ? 8 - (rnd ? 1 : 0) : 32 - (rnd ? 1 : 0)),
                        hm == 2 ? 4 : 6);
        }
        return clampi(v, 0, 255);
    };
    uint32_t total = 0;
    if (metric == 0) {
        const int s = imax(ss, 1);
        for (int y = 0; y < ah; y += s)
            for (int x = 0; x < aw; x += s)
                total += static_cast<uint32_t>(iabs(static_cast<int>(cur[static_cast<size_t>(y0 + y) * width + x0 + x]) - pred(x0 + x, y0 + y)));
        if (s > 1) total *= static_cast<uint32_t>(s * s);
    } else {
        for (int by = 0; by < bh; by += 4)
            for (int bx = 0; bx < bw; bx += 4) {
                if (!(bx < aw && by < ah)) continue;
                int d[16];
                for (int yy = 0; yy < 4; ++yy)
                    for (int xx = 0; xx < 4; ++xx) {
                        const int px = bx + xx, py = by + yy;
                        d[yy * 4 + xx] = (px < aw && py < ah)
                                             ? static_cast<int>(cur[static_cast<size_t>(y0 + py) * width + x0 + px]) - pred(x0 + px, y0 + py)
                                             : 0;
                    }
                int t[16];
                for (int y = 0; y < 4; ++y) {
                    const int a0 = d[y * 4] + d[y * 4 + 3], a1 = d[y * 4 + 1] + d[y * 4 + 2];
                    const int a2 = d[y * 4 + 1] - d[y * 4 + 2], a3 = d[y * 4] - d[y * 4 + 3];
                    t[y * 4] = a0 + a1; t[y * 4 + 1] = a3 + a2; t[y * 4 + 2] = a0 - a1; t[y * 4 + 3] = a3 - a2;
                }
                uint32_t tile = 0;
                for (int x = 0; x < 4; ++x) {
                    const int a0 = t[x] + t[12 + x], a1 = t[4 + x] + t[8 + x];
                    const int a2 = t[4 + x] - t[8 + x], a3 = t[x] - t[12 + x];
                    tile += static_cast<uint32_t>(iabs(a0 + a1) + iabs(a3 + a2) + iabs(a0 - a1) + iabs(a3 - a2));
                }
                total += (tile + 1) >> 1;
            }
    }
    costs[i] = total;
}


// ------------------------------------------------------------------ B pictures
// Vc1Encoder::analyze_b_picture_mode (src/encoder_motion.cpp:753-1004) after
// its two core passes: job 0 = forward (cur vs past), job 1 = backward (cur
// vs future).

// motion_uses_hidden_chroma with the B future-reference rounding (future:
// odd chroma components round away from zero, src/encoder_motion.cpp:2176-2184)
__device__ __forceinline__ bool hidden_chroma_future(const Geo& g, int x0, int y0, int bw, int bh, int xq, int yq) {
    const int cw = (g.vw + 1) / 2, ch = (g.vh + 1) / 2;
    const bool rp = cw < g.mbw * 8, bp = ch < g.mbh * 8;
    if (!rp && !bp) return false;
    const int cx = x0 / 2, cy = y0 / 2;
    const int aw = imax(0, imin(bw / 2, cw - cx)), ah = imax(0, imin(bh / 2, ch - cy));
    if (!aw || !ah) return false;
    int uvx = arshift(xq + ((xq & 3) == 3), 1), uvy = arshift(yq + ((yq & 3) == 3), 1);
    uvx = uvx < 0 ? uvx - (uvx & 1) : uvx + (uvx & 1);
    uvy = uvy < 0 ? uvy - (uvy & 1) : uvy + (uvy & 1);
    const int ix = floor_div4(uvx), iy = floor_div4(uvy);
    const bool fx = (uvx - ix * 4) != 0, fy = (uvy - iy * 4) != 0;
    return (rp && cx + aw - 1 + ix + (fx ? 1 : 0) >= cw) || (bp && cy + ah - 1 + iy + (fy ? 1 : 0) >= ch);
}
__device__ __forceinline__ bool b_unsafe(const Geo& g, int x0, int y0, int2 v, int mode, bool future) {
    return hidden_luma(g, x0, y0, 16, 16, v.x, v.y, mode) ||
           (future ? hidden_chroma_future(g, x0, y0, 16, 16, v.x, v.y) : hidden_chroma(g, x0, y0, 16, 16, v.x, v.y));
}

// the direct-mode vector: the collocated anchor vector scaled by bfrac / 256
// (or bfrac - 256), half-pel rounded in half-pel modes, clamped to the
// picture's MV bounds (src/encoder_motion.cpp:815-833)
__device__ __forceinline__ int2 b_scale_direct(int2 col, bool inverse, int bfrac, int mode, int mx, int my, int mbw,
                                               int mbh) {
    const int n = inverse ? bfrac - 256 : bfrac;
    int2 mv;
    if (mode_halfpel(mode)) {
        mv.x = 2 * arshift(col.x * n + 255, 9);
        mv.y = 2 * arshift(col.y * n + 255, 9);
    } else {
        mv.x = arshift(col.x * n + 128, 8);
        mv.y = arshift(col.y * n + 128, 8);
    }
    mv.x = clampi(mv.x, -60 - mx * 64, mbw * 64 - 4 - mx * 64);
    mv.y = clampi(mv.y, -60 - my * 64, mbh * 64 - 4 - my * 64);
    return mv;
}

// Four predicted samples (row r of the macroblock, columns c0..c0+3) of
// reference job R at vector v: the job's phase plane of v's phase.
__device__ __forceinline__ uint32_t b_pred4(const JobDesc* R, int mode, int x, int y, int2 v) {
    const bool bil = mode_bilinear(mode);
    const int ix = floor_div4(v.x), iy = floor_div4(v.y);
    const int hm = v.x - ix * 4, vm = v.y - iy * 4;
    const int phase = bil ? (vm >> 1) * 2 + (hm >> 1) : vm * 4 + hm;
    const gptr_u8 pl = as_global(bil ? R->bp : R->qp) + static_cast<size_t>(phase) * R->pp * R->ph;
    const int yc = clampi(y + iy, -2, R->H) + 2;
    return plane_row4(pl + static_cast<size_t>(yc) * R->pp, x + ix, R->W);
}

// rounding average of packed bytes, (a + b + 1) >> 1 per byte
__device__ __forceinline__ uint32_t avg_u8x4(uint32_t a, uint32_t b) { return (a | b) - (((a ^ b) >> 1) & 0x7f7f7f7fu); }

// One 64-lane wave per macroblock: the direct vectors, the four prediction
// SADs over the visible pixels (each lane four pixels of one row), their
// hidden-padding screens, and the B-intra proxy (the four 8x8 blocks' SAD
// about their visible means).
__global__ void __launch_bounds__(64) k_b_costs(const JobDesc* __restrict__ jobs, const BDesc* __restrict__ Bs) {
    const BDesc* B = &Bs[blockIdx.y];
    const JobDesc* F = &jobs[2 * blockIdx.y];
    const JobDesc* K = &jobs[2 * blockIdx.y + 1];
    const int pos = static_cast<int>(blockIdx.x);
    if (pos >= F->mbw * F->mbh) return;
    const int mbw = F->mbw, mbh = F->mbh, W = F->W, H = F->H, mode = B->mode;
    const int mx = pos % mbw, my = pos / mbw, x0 = mx * 16, y0 = my * 16;
    const int lane = static_cast<int>(threadIdx.x);
    const Geo g = {F->W, F->H, F->vw, F->vh, F->mbw, F->mbh};
    const int2 col = B->anchor_mvs[pos];
    const int2 df = b_scale_direct(col, false, B->bfrac, mode, mx, my, mbw, mbh);
    const int2 db = b_scale_direct(col, true, B->bfrac, mode, mx, my, mbw, mbh);
    const int2 fm = F->mvs[pos], bm = K->mvs[pos];
    const int vw = imin(16, W - x0), vh = imin(16, H - y0);
    const int r = lane >> 2, c0 = (lane & 3) * 4;
    uint32_t cur = 0, cm = 0;
    if (r < vh) {
        const gptr_u8 row = as_global(F->cy) + static_cast<size_t>(y0 + r) * W + x0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (c0 + j < vw) { cur |= static_cast<uint32_t>(row[c0 + j]) << (8 * j); cm |= 0xffu << (8 * j); }
    }
    uint32_t sf = 0, sb = 0, si = 0, sd = 0;
    if (r < vh) {
        const uint32_t pf = b_pred4(F, mode, x0 + c0, y0 + r, fm) & cm;
        const uint32_t pb = b_pred4(K, mode, x0 + c0, y0 + r, bm) & cm;
        const uint32_t pdf = b_pred4(F, mode, x0 + c0, y0 + r, df) & cm;
        const uint32_t pdb = b_pred4(K, mode, x0 + c0, y0 + r, db) & cm;
        sf = __builtin_amdgcn_sad_u8(pf, cur, 0u);
        sb = __builtin_amdgcn_sad_u8(pb, cur, 0u);
        si = __builtin_amdgcn_sad_u8(avg_u8x4(pf, pb), cur, 0u);
        sd = __builtin_amdgcn_sad_u8(avg_u8x4(pdf, pdb), cur, 0u);
    }
    sf = wave_sum32(sf); sb = wave_sum32(sb); si = wave_sum32(si); sd = wave_sum32(sd);
    // visible_block_mean (src/encoder_entropy.cpp:496-509) of each 8x8 block and the SAD about it
    uint64_t intra = 0;
    const int lx = lane & 7, ly = lane >> 3;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int bx = x0 + (k & 1) * 8, by = y0 + ((k >> 1) & 1) * 8;
        const bool in = bx + lx < W && by + ly < H;
        const uint32_t v = in ? F->cy[static_cast<size_t>(by + ly) * W + bx + lx] : 0u;
        const uint32_t sum = wave_sum32(v), n = wave_sum32(in ? 1u : 0u);
        const int mean = n ? static_cast<int>((sum + n / 2) / n) : 128;
        intra += wave_sum32(in ? static_cast<uint32_t>(iabs(static_cast<int>(v) - mean)) : 0u);
    }
    if (lane == 0) {
        constexpr uint64_t kInvalid = uint64_t{1} << 54;
        const bool ok_fm = !b_unsafe(g, x0, y0, fm, mode, false), ok_bm = !b_unsafe(g, x0, y0, bm, mode, true);
        const bool ok_df = !b_unsafe(g, x0, y0, df, mode, false), ok_db = !b_unsafe(g, x0, y0, db, mode, true);
        const uint64_t pen = B->anchor_4mv[pos] ? 1000000000ull : 0ull;
        uint64_t* c = B->cost + 4 * static_cast<size_t>(pos);
        c[0] = (ok_df && ok_db ? sd : kInvalid) + pen;
        c[1] = (ok_fm ? sf : kInvalid) + pen;
        c[2] = (ok_bm ? sb : kInvalid) + pen;
        c[3] = ok_fm && ok_bm ? si : kInvalid;
        B->intra_sad[pos] = intra;
        B->direct[2 * pos] = df;
        B->direct[2 * pos + 1] = db;
    }
}

// predictor_info(...).pre of a final B field (src/encoder_motion.cpp:1312-1352)
// from its top (A), top-right (B; top-left in the last column) and left (C)
// vectors: their median (or the one valid of them), pulled inside the
// picture's MV bounds
__device__ __forceinline__ int2 b_predictor(int2 A, bool a_valid, int2 Bv, bool b_valid, int2 Cv, bool c_valid, int mx,
                                            int my, int mbw, int mbh) {
    const int2 z = make_int2(0, 0);
    if (!a_valid) A = z;
    if (!b_valid) Bv = z;
    if (!c_valid) Cv = z;
    int2 pre = a_valid ? A : (c_valid ? Cv : (b_valid ? Bv : z));
    if ((a_valid ? 1 : 0) + (b_valid ? 1 : 0) + (c_valid ? 1 : 0) > 1) {
        pre.x = median3(A.x, Bv.x, Cv.x);
        pre.y = median3(A.y, Bv.y, Cv.y);
    }
    const int qx = mx * 64, qy = my * 64, X = mbw * 64 - 4, Y = mbh * 64 - 4;
    if (qx + pre.x < -60) pre.x = -60 - qx;
    if (qx + pre.x > X) pre.x = X - qx;
    if (qy + pre.y < -60) pre.y = -60 - qy;
    if (qy + pre.y > Y) pre.y = Y - qy;
    return pre;
}

// (long double)cost + (double)(lambda * bits), rounded to a 64-bit
// significand, as an integer multiple of 2^-52 (lambda >= 1, so every
// nonzero double product is a multiple of 2^-52 below 2^11)
__device__ __forceinline__ unsigned __int128 b_score(uint64_t cost, uint64_t bits, double lambda) {
    const double prod = __dmul_rn(lambda, static_cast<double>(bits));
    const uint64_t p52 = static_cast<uint64_t>(__dmul_rn(prod, 4503599627370496.0));
    return x87_round((static_cast<unsigned __int128>(cost) << 52) + p52);
}

// The raster-order decision: macroblock (mx, my) needs the final fields of
// its left, top-left, top and top-right neighbours, so the macroblocks of
// anti-diagonal t = mx + 2 my are independent once t - 1 is done.  One
// workgroup walks the diagonals with one thread per macroblock row (row my
// decides macroblock t - 2 my on diagonal t).  The final fields of the last
// three diagonals live in an LDS ring (entry [t & 3][my]: forward and
// backward vectors), which holds every neighbour a decision reads; each
// thread loads its next macroblock's costs and vectors while the current one
// is decided.
struct BMb {
    int2 fm, bm, df, db;
    uint64_t c0, c1, c2, c3, isad, sf, sb;
};
__device__ __forceinline__ BMb b_load(const JobDesc* F, const JobDesc* K, const BDesc* B, int pos) {
    BMb m;
    m.fm = F->mvs[pos];
    m.bm = K->mvs[pos];
    m.df = B->direct[2 * pos];
    m.db = B->direct[2 * pos + 1];
    const uint64_t* c = B->cost + 4 * static_cast<size_t>(pos);
    m.c0 = c[0]; m.c1 = c[1]; m.c2 = c[2]; m.c3 = c[3];
    m.isad = B->intra_sad[pos];
    m.sf = F->screen[pos];
    m.sb = K->screen[pos];
    return m;
}

__global__ void __launch_bounds__(1024) k_b_decide(const JobDesc* __restrict__ jobs, const BDesc* __restrict__ Bs) {
    extern __shared__ int4 ring[];   // [4][mbh]: (forward x, y, backward x, y)
    const BDesc* B = &Bs[blockIdx.x];
    const JobDesc* F = &jobs[2 * blockIdx.x];
    const JobDesc* K = &jobs[2 * blockIdx.x + 1];
    const int mbw = F->mbw, mbh = F->mbh, mode = B->mode;
    const int tmax = (mbw - 1) + 2 * (mbh - 1);
    const double lambda = B->lambda;
    const uint32_t marker = imax(static_cast<int>(mv_bits0(35)), static_cast<int>(mv_bits0(72))) + 1;
    const uint32_t ftype = B->below_half ? 1u : 2u, btype = B->below_half ? 2u : 1u;
    const int my = static_cast<int>(threadIdx.x);
    const bool row = my < mbh;
    BMb cur{};
    if (row) cur = b_load(F, K, B, my * mbw);
    for (int t = 0; t <= tmax; ++t) {
        const int mx = t - 2 * my;
        if (row && mx >= 0 && mx < mbw) {
            const int pos = my * mbw + mx;
            const BMb m = cur;
            if (mx + 1 < mbw) cur = b_load(F, K, B, pos + 1);
            // neighbours: left (t-1, my), top (t-2, my-1), top-right (t-1, my-1),
            // top-left in the last column (t-3, my-1)
            const bool a_valid = my > 0, c_valid = mx > 0, b_valid = my > 0 && mbw > 1;
            const int4 L = c_valid ? ring[((t - 1) & 3) * mbh + my] : make_int4(0, 0, 0, 0);
            const int4 T = a_valid ? ring[((t - 2) & 3) * mbh + my - 1] : make_int4(0, 0, 0, 0);
            const int4 D = !b_valid ? make_int4(0, 0, 0, 0)
                                    : (mx == mbw - 1 ? ring[((t - 3) & 3) * mbh + my - 1] : ring[((t - 1) & 3) * mbh + my - 1]);
            const int2 fp = b_predictor(make_int2(T.x, T.y), a_valid, make_int2(D.x, D.y), b_valid, make_int2(L.x, L.y),
                                        c_valid, mx, my, mbw, mbh);
            const int2 bp = b_predictor(make_int2(T.z, T.w), a_valid, make_int2(D.z, D.w), b_valid, make_int2(L.z, L.w),
                                        c_valid, mx, my, mbw, mbh);
            const uint32_t fbits = mv_rate_bits(F->mvrange, m.fm.x, m.fm.y, fp.x, fp.y, mode, false);
            const uint32_t bbits = mv_rate_bits(F->mvrange, m.bm.x, m.bm.y, bp.x, bp.y, mode, false);
            const uint32_t ibits = mv_rate_bits(F->mvrange, m.bm.x, m.bm.y, bp.x, bp.y, mode, true) + fbits + 2;
            // candidates in the reference's order: direct, forward, backward, interpolated
            const uint64_t cost[4] = {m.c0, m.c1, m.c2, m.c3};
            const uint32_t rate[4] = {0u, fbits + ftype, bbits + btype, ibits};
            int best = 0;
            uint64_t best_cost = cost[0];
            uint32_t best_rate = rate[0];
            unsigned __int128 bs = b_score(cost[0], rate[0], lambda);
#pragma unroll
            for (int k = 1; k < 4; ++k) {
                const unsigned __int128 sc = b_score(cost[k], rate[k], lambda);
                if (sc < bs) { bs = sc; best = k; best_cost = cost[k]; best_rate = rate[k]; }
            }
            uint64_t gate = best_cost;
            if (m.sf < gate) gate = m.sf;
            if (m.sb < gate) gate = m.sb;
            const uint64_t guarded = m.isad + B->intra_guard;
            B->ratio[pos] = gate > 0 ? __ddiv_rn(static_cast<double>(guarded), static_cast<double>(gate))
                                     : __builtin_huge_val();
            const bool intra = B->allow_intra &&
                               guarded < static_cast<uint64_t>(llround_exact(__dmul_rn(static_cast<double>(gate),
                                                                                       B->intra_thr)));
            int2 fo = make_int2(0, 0), bo = make_int2(0, 0);
            uint8_t mm = 4;
            if (!intra) {
                mm = best == 0 ? 3 : static_cast<uint8_t>(best - 1);   // direct, forward, backward, interpolated
                fo = (best == 1 || best == 3) ? m.fm : m.df;
                bo = (best == 2 || best == 3) ? m.bm : m.db;
            }
            ring[(t & 3) * mbh + my] = make_int4(fo.x, fo.y, bo.x, bo.y);
            B->modes[pos] = mm;
            B->fmv[pos] = fo;
            B->bmv[pos] = bo;
            B->sel_sad[pos] = intra ? m.isad : best_cost;
            B->rate_sad[pos] = best_cost;
            B->sel_rate[pos] = intra ? marker : best_rate;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ motion compensation
// The P prediction picture of an analysis: motion_compensate_mb /
// motion_compensate_4mv (src/encoder_motion_compensation.cpp:474-571), one
// thread per output sample.  The same kernels write the padded pictures of
// the loop reconstruction, motion_compensate_mb_padded / _4mv_padded
// (:573-657): those cover the whole macroblock raster (DW x DH = mbw*16 x
// mbh*16) while the samples clamp to the reference's own raster (RW x RH),
// which is the coded raster or the padded one.  The plain picture is the
// case DW = RW = W, DH = RH = H.

// luma: the picture's mode for 1-MV macroblocks, bicubic for the four
// vectors of a 4-MV macroblock (luma_mc_sample_mode / luma_mc_sample)
__global__ void __launch_bounds__(256) k_predict_luma(const uint8_t* __restrict__ ref, int RW, int RH, int DW, int DH,
                                                      int mbw, int mode, const int2* __restrict__ mvs,
                                                      const int2* __restrict__ bmv, const uint8_t* __restrict__ four,
                                                      uint8_t* __restrict__ dst) {
    const int px = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x), py = static_cast<int>(blockIdx.y);
    if (px >= DW || py >= DH) return;
    const int pos = (py >> 4) * mbw + (px >> 4);
    int2 v;
    bool bil = mode == 2;
    if (four[pos]) {
        v = bmv[pos * 4 + ((py >> 3) & 1) * 2 + ((px >> 3) & 1)];
        bil = false;
    } else {
        v = mvs[pos];
    }
    const int ix = floor_div4(v.x), iy = floor_div4(v.y);
    const int hm = v.x - ix * 4, vm = v.y - iy * 4;
    int n[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int yy = clampi(py + iy - 1 + r, 0, RH - 1);
#pragma unroll
        for (int c = 0; c < 4; ++c) n[r][c] = ref[static_cast<size_t>(yy) * RW + clampi(px + ix - 1 + c, 0, RW - 1)];
    }
    dst[static_cast<size_t>(py) * DW + px] = static_cast<uint8_t>(mc_sample(n, hm, vm, bil));
}

// chroma_mc_sample (src/encoder_motion_compensation.cpp:417-447): `future`
// rounds an odd half-chroma component away from the even value (the
// future reference of a bi-directional prediction), otherwise toward it
__device__ __forceinline__ int chroma_sample(const uint8_t* p, int cw, int ch, int px, int py, int2 v, bool future) {
    int uvx = arshift(v.x + ((v.x & 3) == 3), 1), uvy = arshift(v.y + ((v.y & 3) == 3), 1);
    if (future) {
        uvx = uvx < 0 ? uvx - (uvx & 1) : uvx + (uvx & 1);
        uvy = uvy < 0 ? uvy - (uvy & 1) : uvy + (uvy & 1);
    } else {
        uvx = uvx < 0 ? uvx + (uvx & 1) : uvx - (uvx & 1);
        uvy = uvy < 0 ? uvy + (uvy & 1) : uvy - (uvy & 1);
    }
    const int ix = floor_div4(uvx), iy = floor_div4(uvy);
    const int fx = (uvx - ix * 4) * 2, fy = (uvy - iy * 4) * 2;
    const int x0 = clampi(px + ix, 0, cw - 1), x1 = clampi(px + ix + 1, 0, cw - 1);
    const int y0 = clampi(py + iy, 0, ch - 1), y1 = clampi(py + iy + 1, 0, ch - 1);
    const int a = p[static_cast<size_t>(y0) * cw + x0];
    if (!fx && !fy) return a;
    const int b = p[static_cast<size_t>(y0) * cw + x1], c = p[static_cast<size_t>(y1) * cw + x0];
    const int d = p[static_cast<size_t>(y1) * cw + x1];
    return (a * (8 - fx) * (8 - fy) + b * fx * (8 - fy) + c * (8 - fx) * fy + d * fx * fy + 32) >> 6;
}

// chroma, P form, at the macroblock's vector or, for 4-MV, the median-4
// vector (:522-535); reference chroma RCW x RCH, destination DCW x DCH
__global__ void __launch_bounds__(256) k_predict_chroma(const uint8_t* __restrict__ ru, const uint8_t* __restrict__ rv,
                                                        int RCW, int RCH, int DCW, int DCH, int mbw,
                                                        const int2* __restrict__ mvs, const int2* __restrict__ bmv,
                                                        const uint8_t* __restrict__ four, uint8_t* __restrict__ du,
                                                        uint8_t* __restrict__ dv) {
    const int px = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x), py = static_cast<int>(blockIdx.y);
    if (px >= DCW || py >= DCH) return;
    const int pos = (py >> 3) * mbw + (px >> 3);
    int2 v;
    if (four[pos]) {
        const int2* b = bmv + pos * 4;
        v = make_int2(median4(b[0].x, b[1].x, b[2].x, b[3].x), median4(b[0].y, b[1].y, b[2].y, b[3].y));
    } else {
        v = mvs[pos];
    }
    const size_t o = static_cast<size_t>(py) * DCW + px;
    du[o] = static_cast<uint8_t>(chroma_sample(ru, RCW, RCH, px, py, v, false));
    dv[o] = static_cast<uint8_t>(chroma_sample(rv, RCW, RCH, px, py, v, false));
}

// The bi-directional prediction picture: motion_compensate_bi_mb
// (src/encoder_motion_compensation.cpp:659-711) or, over the macroblock
// raster, motion_compensate_bi_mb_padded (:712-763) for every macroblock:
// the rounding average of the past prediction at fmvs and the future
// prediction at bmvs, each reference clamped to its own raster (PW x PH,
// FW x FH; planes contiguous, chroma after luma).
__device__ __forceinline__ int luma_sample(const uint8_t* ref, int W, int H, int px, int py, int2 v, bool bil) {
    const int ix = floor_div4(v.x), iy = floor_div4(v.y);
    int n[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int yy = clampi(py + iy - 1 + r, 0, H - 1);
#pragma unroll
        for (int c = 0; c < 4; ++c) n[r][c] = ref[static_cast<size_t>(yy) * W + clampi(px + ix - 1 + c, 0, W - 1)];
    }
    return mc_sample(n, v.x - ix * 4, v.y - iy * 4, bil);
}
__global__ void __launch_bounds__(256) k_predict_bi(const uint8_t* __restrict__ past, int PW, int PH,
                                                    const uint8_t* __restrict__ fut, int FW, int FH, int DW, int DH,
                                                    int mbw, int mode, const int2* __restrict__ fmv,
                                                    const int2* __restrict__ bmv, uint8_t* __restrict__ dy,
                                                    uint8_t* __restrict__ du, uint8_t* __restrict__ dv) {
    const int px = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x), py = static_cast<int>(blockIdx.y);
    if (blockIdx.z == 0) {
        if (px >= DW || py >= DH) return;
        const int pos = (py >> 4) * mbw + (px >> 4);
        const int a = luma_sample(past, PW, PH, px, py, fmv[pos], mode == 2);
        const int z = luma_sample(fut, FW, FH, px, py, bmv[pos], mode == 2);
        dy[static_cast<size_t>(py) * DW + px] = static_cast<uint8_t>((a + z + 1) >> 1);
        return;
    }
    const int dcw = DW / 2, dch = DH / 2;
    if (px >= dcw || py >= dch) return;
    const int pos = (py >> 3) * mbw + (px >> 3);
    const int pcw = PW / 2, pch = PH / 2, fcw = FW / 2, fch = FH / 2;
    const uint8_t* pp = past + static_cast<size_t>(PW) * PH + (blockIdx.z == 2 ? static_cast<size_t>(pcw) * pch : 0);
    const uint8_t* fp = fut + static_cast<size_t>(FW) * FH + (blockIdx.z == 2 ? static_cast<size_t>(fcw) * fch : 0);
    uint8_t* out = blockIdx.z == 2 ? dv : du;
    const int a = chroma_sample(pp, pcw, pch, px, py, fmv[pos], false);
    const int b = chroma_sample(fp, fcw, fch, px, py, bmv[pos], true);
    out[static_cast<size_t>(py) * dcw + px] = static_cast<uint8_t>((a + b + 1) >> 1);
}
}  // namespace

// ------------------------------------------------------------------ launchers

hipError_t launch_predict_bi(const uint8_t* past, int PW, int PH, const uint8_t* fut, int FW, int FH, int DW, int DH,
                             int mode, const int2* fmv, const int2* bmv, uint8_t* dy, uint8_t* du, uint8_t* dv,
                             hipStream_t s) {
    if (DW <= 0 || DH <= 0 || PW <= 0 || PH <= 0 || FW <= 0 || FH <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_predict_bi, dim3((DW + 255) / 256, DH, 3), dim3(256), 0, s, past, PW, PH, fut, FW, FH, DW, DH,
                       (DW + 15) / 16, mode, fmv, bmv, dy, du, dv);
    return hipGetLastError();
}

hipError_t launch_predict_p(const uint8_t* ry, const uint8_t* ru, const uint8_t* rv, int RW, int RH, int DW, int DH,
                            int mode, const int2* mvs, const int2* bmv, const uint8_t* four, uint8_t* dy, uint8_t* du,
                            uint8_t* dv, hipStream_t s) {
    if (DW <= 0 || DH <= 0 || RW <= 0 || RH <= 0) return hipSuccess;
    const int mbw = (DW + 15) / 16;
    hipLaunchKernelGGL(k_predict_luma, dim3((DW + 255) / 256, DH), dim3(256), 0, s, ry, RW, RH, DW, DH, mbw, mode, mvs,
                       bmv, four, dy);
    if (DW / 2 > 0 && DH / 2 > 0 && RW / 2 > 0 && RH / 2 > 0)
        hipLaunchKernelGGL(k_predict_chroma, dim3((DW / 2 + 255) / 256, DH / 2), dim3(256), 0, s, ru, rv, RW / 2,
                           RH / 2, DW / 2, DH / 2, mbw, mvs, bmv, four, du, dv);
    return hipGetLastError();
}

hipError_t launch_b_costs(const JobDesc* d_jobs, const BDesc* d_b, int nb, int max_mbs, hipStream_t s) {
    if (max_mbs <= 0 || nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_b_costs, dim3(max_mbs, nb), dim3(64), 0, s, d_jobs, d_b);
    return hipGetLastError();
}

hipError_t launch_b_decide(const JobDesc* d_jobs, const BDesc* d_b, int nb, int max_mbh, hipStream_t s) {
    if (nb <= 0) return hipSuccess;
    if (max_mbh <= 0 || max_mbh > kBMaxRows) return hipErrorInvalidValue;
    const int threads = (max_mbh + kWave - 1) / kWave * kWave;
    hipLaunchKernelGGL(k_b_decide, dim3(nb), dim3(threads), 4 * max_mbh * sizeof(int4), s, d_jobs, d_b);
    return hipGetLastError();
}

int mb_kernel_lds_bytes(int local_range) { return lds_bytes_for(local_range); }

hipError_t launch_global_search(const JobDesc* d_jobs, int njobs, const JobDesc* h_jobs, hipStream_t s,
                                hipEvent_t ev0, hipEvent_t ev1) {
    // per-stage upper bounds of point count and sampled-row chunks over the batch
    int pts[4] = {0, 0, 0, 0}, chunks[4] = {0, 0, 0, 0};
    for (int i = 0; i < njobs; ++i) {
        const JobDesc& J = h_jobs[i];
        if (J.range == 0 || J.seed_mvs) continue;
        const int R = J.local_range < 32 ? J.local_range : 32;
        int np[4] = {0, 0, 0, 0}, st[4] = {1, 1, 1, 1};
        if (R <= 0) {
            np[0] = 1;
        } else if (R > 31) {
            const int step = R >= 160 ? 16 : 8;
            np[0] = 1 + 12 * (R / step); st[0] = 16;
            np[1] = 4 * (2 * step / 4 + 1) * (2 * step / 4 + 1); st[1] = 8;
            np[2] = 81; st[2] = 4;
            np[3] = 9; st[3] = 1;
        } else {
            const int cs = R >= 4 ? 4 : (R >= 2 ? 2 : 1);
            const int side = 2 * R / cs + 1;
            np[0] = side * side + 1 + 2 * side; st[0] = 4;
            np[1] = 8 * (2 * cs - 1) * (2 * cs - 1); st[1] = 2;
            np[2] = 1; st[2] = 1;
        }
        for (int k = 0; k < 4; ++k) {
            if (np[k] > kGlobalMaxPts) return hipErrorInvalidValue;
            if (!np[k]) continue;
            pts[k] = np[k] > pts[k] ? np[k] : pts[k];
            const int rows = (J.H + st[k] - 1) / st[k];
            const int c = (rows + 7) / 8;
            chunks[k] = c > chunks[k] ? c : chunks[k];
        }
    }
    const int tb = 64, gb = (njobs + tb - 1) / tb;
    if (ev0) (void)hipEventRecord(ev0, s);
    hipLaunchKernelGGL(k_global_begin, dim3(gb), dim3(tb), 0, s, d_jobs, njobs);
    // up to four stages; each: frame SADs of the stage's points, then the replay
    for (int stage = 0; stage < 4; ++stage) {
        if (!pts[stage]) continue;
        // point groups of at most kFsPoints, as even as the count allows
        const int groups = (pts[stage] + kFsPoints - 1) / kFsPoints;
        const int per = (pts[stage] + groups - 1) / groups;
        hipLaunchKernelGGL(k_frame_sads, dim3(chunks[stage], groups, njobs), dim3(256), 0, s, d_jobs, per);
        hipLaunchKernelGGL(k_global_step, dim3(njobs), dim3(64), 0, s, d_jobs, njobs, stage);
    }
    if (ev1) (void)hipEventRecord(ev1, s);
    return hipGetLastError();
}

hipError_t launch_signatures(const JobDesc* d_jobs, int njobs, int max_ncx, int max_ncy, hipStream_t s) {
    if (max_ncx <= 0 || max_ncy <= 0) return hipSuccess;
    const dim3 grid((max_ncx + kSigTileX - 1) / kSigTileX, (max_ncy + kSigTileY - 1) / kSigTileY, njobs);
    hipLaunchKernelGGL(k_signatures, grid, dim3(256), 0, s, d_jobs);
    return hipGetLastError();
}

// Phase planes of one reference (see qpel_eval).  Each thread produces four
// consecutive plane positions px0 .. px0+3 of row py, px0 = 4 t - 2, which
// fill one aligned dword of every plane (plane column px + 2).  Their
// clamped 4x7 neighbourhood v[r][k] = ref(py - 1 + r, px0 - 1 + k) serves
// every phase: position j reads n[r][c] = v[r][j + c] of luma_mc_sample
// (mc_sample).  A vertical filter sum depends only on the column and vm, so
// it is formed once per column and rounded for both horizontal classes
// (shift 1 or 3 at vm 2, 3 or 5 at vm 1 and 3: hm 2 against hm 1 and 3).
// Four byte values (each in [0, 255]) packed little-endian through v_perm_b32.
// Packed with shifts and ors, the clamped filter outputs are fused by the
// compiler into v_ashr_pk_u8_i32, which writes only the low half of its
// destination (still holding the first unshifted sum, all ones above bit 15
// when negative), and the two bytes or-ed in above came out as 255.
__device__ __forceinline__ uint32_t pack4(int a, int b, int c, int d) {
    const uint32_t lo = __builtin_amdgcn_perm(static_cast<uint32_t>(b), static_cast<uint32_t>(a), 0x0c0c0400u);
    const uint32_t hi = __builtin_amdgcn_perm(static_cast<uint32_t>(d), static_cast<uint32_t>(c), 0x0c0c0400u);
    return __builtin_amdgcn_perm(hi, lo, 0x05040100u);
}
__device__ __forceinline__ void phase_quad(const uint8_t* __restrict__ ref, int W, int H, uint8_t* __restrict__ out,
                                           int pp, int ph, int bil, int t, int py) {
    const int px0 = 4 * t - 2;
    if (px0 > W || py > H) return;
    const gptr_u8 r = as_global(ref);
    typedef const __attribute__((address_space(1))) uint32_t* gptr_u32;
    int v[4][7];
    const int x0 = px0 - 1;              // = 4 t - 3
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const size_t row = static_cast<size_t>(clampi(py - 1 + a, 0, H - 1)) * W;
        if ((W & 3) == 0 && x0 - 1 >= 0 && x0 + 6 <= W - 1) {
            // bytes 4t-4 .. 4t+3 of the row: two aligned dwords
            const gptr_u32 q = reinterpret_cast<gptr_u32>(r + row + (x0 - 1));
            const uint32_t d0 = q[0], d1 = q[1];
#pragma unroll
            for (int k = 0; k < 7; ++k) v[a][k] = static_cast<int>(((k < 3 ? d0 >> (8 * (k + 1)) : d1 >> (8 * (k - 3)))) & 255u);
        } else {
#pragma unroll
            for (int k = 0; k < 7; ++k) v[a][k] = r[row + clampi(x0 + k, 0, W - 1)];
        }
    }
    const size_t psz = static_cast<size_t>(pp) * ph;
    uint8_t* o = out + static_cast<size_t>(py + 2) * pp + (px0 + 2);
    uint32_t w[16];
    int np;
    if (bil) {
        np = 4;
        int s1[4], s2[4], s3[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int n11 = v[1][j + 1], n12 = v[1][j + 2], n21 = v[2][j + 1], n22 = v[2][j + 2];
            s1[j] = (n11 + n12 + 1) >> 1;
            s2[j] = (n11 + n21 + 1) >> 1;
            s3[j] = (n11 + n12 + n21 + n22 + 2) >> 2;
        }
        w[0] = pack4(v[1][1], v[1][2], v[1][3], v[1][4]);
        w[1] = pack4(s1[0], s1[1], s1[2], s1[3]);
        w[2] = pack4(s2[0], s2[1], s2[2], s2[3]);
        w[3] = pack4(s3[0], s3[1], s3[2], s3[3]);
    } else {
        np = 16;
        w[0] = pack4(v[1][1], v[1][2], v[1][3], v[1][4]);
        // vm = 0: horizontal only
#pragma unroll
        for (int hm = 1; hm < 4; ++hm) {
            const int bias = hm == 2 ? 8 : 32, sh = hm == 2 ? 4 : 6;
            int q[4];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                q[j] = clampi(arshift(mspel(v[1][j], v[1][j + 1], v[1][j + 2], v[1][j + 3], hm) + bias, sh), 0, 255);
            w[hm] = pack4(q[0], q[1], q[2], q[3]);
        }
#pragma unroll
        for (int vm = 1; vm < 4; ++vm) {
            // hm = 0: vertical only
            {
                const int bias = vm == 2 ? 7 : 31, sh = vm == 2 ? 4 : 6;
                int q[4];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    q[j] = clampi(arshift(mspel(v[0][j + 1], v[1][j + 1], v[2][j + 1], v[3][j + 1], vm) + bias, sh), 0, 255);
                w[vm * 4] = pack4(q[0], q[1], q[2], q[3]);
            }
            // the vertical sums of the seven columns, rounded per class
            int sv[7];
#pragma unroll
            for (int k = 0; k < 7; ++k) sv[k] = mspel(v[0][k], v[1][k], v[2][k], v[3][k], vm);
#pragma unroll
            for (int hm = 1; hm < 4; ++hm) {
                const int shift = ((hm == 2 ? 1 : 5) + (vm == 2 ? 1 : 5)) >> 1;
                const int r1 = (1 << (shift - 1)) - 1;
                int tv[7];
#pragma unroll
                for (int k = 0; k < 7; ++k) tv[k] = arshift(sv[k] + r1, shift);
                int q[4];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    q[j] = clampi(arshift(mspel(tv[j], tv[j + 1], tv[j + 2], tv[j + 3], hm) + 64, 7), 0, 255);
                w[vm * 4 + hm] = pack4(q[0], q[1], q[2], q[3]);
            }
        }
    }
    // the dword lies inside the row's pitch (pitch >= W + 11 and a multiple
    // of 16; positions past W only fill columns no reader uses)
    typedef uint32_t* u32p;
#pragma unroll
    for (int k = 0; k < 16; ++k)
        if (k < np) *reinterpret_cast<u32p>(o + k * psz) = w[k];
}

// grid: x = groups of 64 quads of a row, y = groups of 4 rows (positions
// px in [-2, W], py in [-2, H])
__device__ __forceinline__ int phase_quad_t() { return static_cast<int>(blockIdx.x) * 64 + static_cast<int>(threadIdx.x & 63); }
__device__ __forceinline__ int phase_quad_y() { return static_cast<int>(blockIdx.y) * 4 + static_cast<int>(threadIdx.x >> 6) - 2; }

__global__ void __launch_bounds__(256) k_phase_planes(const uint8_t* __restrict__ ref, int W, int H,
                                                      uint8_t* __restrict__ out, int pp, int ph, int bil) {
    phase_quad(ref, W, H, out, pp, ph, bil, phase_quad_t(), phase_quad_y());
}

// The phase planes of every job that owns its reference's planes in the step
// (one job per distinct reference and filter); grid z = job.
__global__ void __launch_bounds__(256) k_phase_planes_jobs(const JobDesc* __restrict__ jobs) {
    const JobDesc* J = &jobs[blockIdx.z];
    if (!J->phase_owner) return;
    const bool bil = J->mode == 2;
    phase_quad(J->ry, J->W, J->H, const_cast<uint8_t*>(bil ? J->bp : J->qp), J->pp, J->ph, bil ? 1 : 0,
               phase_quad_t(), phase_quad_y());
}

// quads of a row: positions -2 .. W, four per quad
__host__ __device__ inline int phase_quads(int width) { return (width + 2) / 4 + 1; }

hipError_t launch_phase_planes(const uint8_t* ref, int width, int height, uint8_t* out, int pitch, int rows,
                               int bilinear, hipStream_t s) {
    if (rows < height + 3 || pitch < width + 3 + 8 || (pitch & 15)) return hipErrorInvalidValue;
    const dim3 grid((phase_quads(width) + 63) / 64, (height + 3 + 3) / 4);
    hipLaunchKernelGGL(k_phase_planes, grid, dim3(256), 0, s, ref, width, height, out, pitch, rows, bilinear);
    return hipGetLastError();
}

hipError_t launch_phase_planes_jobs(const JobDesc* d_jobs, int njobs, int max_w, int max_h, hipStream_t s) {
    if (njobs <= 0 || max_w <= 0) return hipSuccess;
    const dim3 grid((phase_quads(max_w) + 63) / 64, (max_h + 3 + 3) / 4, njobs);
    hipLaunchKernelGGL(k_phase_planes_jobs, grid, dim3(256), 0, s, d_jobs);
    return hipGetLastError();
}

hipError_t launch_mb_diagonal(const JobDesc* d_jobs, int njobs, int diag, int max_per_diag, int lds_bytes,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_mb_diagonal, dim3(max_per_diag, njobs), dim3(64), lds_bytes, s, d_jobs, diag);
    return hipGetLastError();
}

hipError_t launch_extended(const JobDesc* d_jobs, int njobs, int max_mbs, int lds_bytes, hipStream_t s) {
    if (njobs <= 0 || max_mbs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_extended, dim3(max_mbs, njobs), dim3(64), lds_bytes, s, d_jobs);
    return hipGetLastError();
}

hipError_t launch_relax_init(const JobDesc* d_jobs, int njobs, int max_mbs, hipStream_t s) {
    if (njobs <= 0 || max_mbs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_relax_init, dim3((max_mbs + 255) / 256, njobs), dim3(256), 0, s, d_jobs);
    return hipGetLastError();
}

hipError_t launch_mb_guess(const JobDesc* d_jobs, int njobs, int max_items, int phase, int phases, int lds_bytes,
                           hipStream_t s) {
    if (njobs <= 0 || max_items <= 0) return hipSuccess;
    if (phases < 1 || phase < 0 || phase >= phases) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mb_guess, dim3(max_items, njobs), dim3(64), lds_bytes, s, d_jobs, phase, phases);
    return hipGetLastError();
}

hipError_t launch_mb_relax0(const JobDesc* d_jobs, int njobs, int max_items, int phase, int phases, int lds_bytes,
                            const RelaxQueue& q, hipStream_t s) {
    if (njobs <= 0 || max_items <= 0) return hipSuccess;
    if (phases < 1 || phase < 0 || phase >= phases) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mb_relax0, dim3(max_items, njobs), dim3(64), lds_bytes, s, d_jobs, q, phase, phases);
    return hipGetLastError();
}

int relax_queue_grid(int lds_bytes, int cu_count) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mb_relaxq, 64, lds_bytes) != hipSuccess || per_cu <= 0)
        per_cu = 8;
    return per_cu * (cu_count > 0 ? cu_count : 256);
}

hipError_t launch_mb_relaxq(const JobDesc* d_jobs, int pass, int grid, int lds_bytes, const RelaxQueue& q,
                            hipStream_t s) {
    if (grid <= 0) return hipSuccess;
    if (pass <= 0 || pass >= kMaxRelaxPasses) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mb_relaxq, dim3(grid), dim3(64), lds_bytes, s, d_jobs, q, pass);
    return hipGetLastError();
}

hipError_t launch_step_begin(const JobDesc* d_jobs, int njobs, int max_mbs, int32_t* counters, int ncounters,
                             hipStream_t s) {
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_step_begin, dim3(((max_mbs > 0 ? max_mbs : 1) + 255) / 256, njobs), dim3(256), 0, s, d_jobs,
                       counters, ncounters);
    return hipGetLastError();
}

hipError_t launch_cost_final(const JobDesc* d_jobs, int njobs, int max_mbs, int lds_bytes, hipStream_t s) {
    if (njobs <= 0 || max_mbs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cost_final, dim3(max_mbs, njobs), dim3(64), lds_bytes, s, d_jobs);
    return hipGetLastError();
}

hipError_t launch_step_end(const JobDesc* d_jobs, int njobs, StepSummary* out, hipStream_t s) {
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_step_end, dim3(njobs), dim3(256), 0, s, d_jobs, out);
    return hipGetLastError();
}

hipError_t launch_block_costs(const JobDesc* d_job, int x0, int y0, int bw, int bh, int mode, int metric,
                              const int2* cands, int n, uint64_t* out, int lds_bytes, hipStream_t s) {
    hipLaunchKernelGGL(k_block_costs, dim3(1), dim3(64), lds_bytes, s, d_job, x0, y0, bw, bh, mode, metric, cands, n,
                       out);
    return hipGetLastError();
}

hipError_t launch_score_batch(const uint8_t* cur, const uint8_t* ref, int width, int height, int x0, int y0, int bw,
                              int bh, int sample_stride, int bilinear, int rnd, int metric, const int2* cands, int n,
                              uint32_t* costs, hipStream_t s) {
    hipLaunchKernelGGL(k_score_batch, dim3((n + 63) / 64), dim3(64), 0, s, cur, ref, width, height, x0, y0, bw, bh,
                       sample_stride, bilinear, rnd, metric, cands, n, costs);
    return hipGetLastError();
}

}  // namespace vc1me
