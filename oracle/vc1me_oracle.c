/*
 * vc1me_oracle.c — CPU restatement of libvc1's P-frame motion estimation.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP backend.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Written from the behaviour of the reference (libvc1 0.2.69); each function
 * names the file:line it restates.  Sequential, scalar and deliberately
 * literal: evaluation order, tie-breaks, dedupe sets and the x87
 * long-double scoring are kept as the reference has them.  The scores of the
 * staged sub-pel refinement use the platform `long double`; on x86-64 Linux
 * that is the same 80-bit x87 format the reference computes in.
 *
 * Pinned by tests/test_oracle.py against the compiled reference
 * (oracle/_ref), against tests/golden/ fixtures written from it, and against
 * the reference's own known-answer tests restated there.
 */
#include "vc1me_oracle.h"

#include <math.h>
#include <stdbool.h>
#include <stdlib.h>
#include <string.h>

#define U64MAX UINT64_MAX
#define BAD_SAD (UINT64_MAX / 4)

typedef struct { int xq, yq; } mv_t;

typedef struct {
    const vc1me_params* p;
    int W, H;      /* coded raster */
    int vw, vh;    /* visible raster */
    int mbw, mbh;
    const uint8_t *cy, *cu, *cv;
    const uint8_t *ry, *ru, *rv;
    oracle_ops* ops;
} ctx_t;

static oracle_ops g_dummy_ops;

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline int iabs(int v) { return v < 0 ? -v : v; }
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

/* ---------------------------------------------------------------- laws */

/* src/encoder_motion_compensation.cpp:350-353 */
static inline int sample_clamped(const uint8_t* p, int w, int h, int x, int y) {
    x = clampi(x, 0, w - 1);
    y = clampi(y, 0, h - 1);
    return p[(size_t)y * w + x];
}
/* src/encoder_motion_compensation.cpp:355-358 */
static inline int floor_div4(int v) { return v >= 0 ? v / 4 : -((-v + 3) / 4); }
/* src/encoder_entropy.cpp:377-381 */
static inline int arshift(int v, int n) {
    if (v >= 0) return v >> n;
    const int d = 1 << n;
    return -((-v + d - 1) / d);
}
/* src/encoder_motion_compensation.cpp:360-365 */
static inline int mspel_filter4(int a, int b, int c, int d, int mode) {
    if (mode == 1) return -4 * a + 53 * b + 18 * c - 3 * d;
    if (mode == 2) return -a + 9 * b + 9 * c - d;
    if (mode == 3) return -3 * a + 18 * b + 53 * c - 4 * d;
    return b;
}
static inline bool mode_halfpel(int m) { return m == VC1ME_MODE_ONE_MV_HPEL || m == VC1ME_MODE_ONE_MV_HPEL_BILINEAR; }
static inline bool mode_bilinear(int m) { return m == VC1ME_MODE_ONE_MV_HPEL_BILINEAR; }

/* luma_mc_sample, src/encoder_motion_compensation.cpp:367-396 */
static int luma_mc_sample(const uint8_t* p, int w, int h, int x, int y, int mvx, int mvy, int rnd) {
    const int ix = floor_div4(mvx), iy = floor_div4(mvy);
    const int hm = mvx - ix * 4, vm = mvy - iy * 4;
    const int sx = x + ix, sy = y + iy;
    if (!hm && !vm) return sample_clamped(p, w, h, sx, sy);
    int v;
    if (vm && hm) {
        static const int shift_value[4] = {0, 5, 1, 5};
        const int shift = (shift_value[hm] + shift_value[vm]) >> 1;
        const int r1 = (1 << (shift - 1)) + (rnd ? 1 : 0) - 1;
        int t[4];
        for (int k = -1; k <= 2; ++k) {
            const int raw = mspel_filter4(sample_clamped(p, w, h, sx + k, sy - 1), sample_clamped(p, w, h, sx + k, sy),
                                          sample_clamped(p, w, h, sx + k, sy + 1), sample_clamped(p, w, h, sx + k, sy + 2), vm);
            t[k + 1] = arshift(raw + r1, shift);
        }
        const int raw = mspel_filter4(t[0], t[1], t[2], t[3], hm);
        v = arshift(raw + 64 - (rnd ? 1 : 0), 7);
    } else if (vm) {
        const int raw = mspel_filter4(sample_clamped(p, w, h, sx, sy - 1), sample_clamped(p, w, h, sx, sy),
                                      sample_clamped(p, w, h, sx, sy + 1), sample_clamped(p, w, h, sx, sy + 2), vm);
        const int bias = (vm == 2) ? (7 + (rnd ? 1 : 0)) : (31 + (rnd ? 1 : 0));
        v = arshift(raw + bias, vm == 2 ? 4 : 6);
    } else {
        const int raw = mspel_filter4(sample_clamped(p, w, h, sx - 1, sy), sample_clamped(p, w, h, sx, sy),
                                      sample_clamped(p, w, h, sx + 1, sy), sample_clamped(p, w, h, sx + 2, sy), hm);
        const int bias = (hm == 2) ? (8 - (rnd ? 1 : 0)) : (32 - (rnd ? 1 : 0));
        v = arshift(raw + bias, hm == 2 ? 4 : 6);
    }
    return clampi(v, 0, 255);
}

/* luma_mc_sample_mode, src/encoder_motion_compensation.cpp:398-415 */
static int luma_mc_sample_mode(const uint8_t* p, int w, int h, int x, int y, int mvx, int mvy, int mode, int rnd) {
    if (!mode_bilinear(mode)) return luma_mc_sample(p, w, h, x, y, mvx, mvy, rnd);
    const int ix = floor_div4(mvx), iy = floor_div4(mvy);
    const int fx = mvx - ix * 4, fy = mvy - iy * 4;
    const int sx = x + ix, sy = y + iy;
    const int a = sample_clamped(p, w, h, sx, sy);
    if (!fx && !fy) return a;
    const int b = sample_clamped(p, w, h, sx + 1, sy);
    const int c = sample_clamped(p, w, h, sx, sy + 1);
    if (fx && fy) {
        const int d = sample_clamped(p, w, h, sx + 1, sy + 1);
        return (a + b + c + d + (rnd ? 1 : 2)) >> 2;
    }
    if (fx) return (a + b + (rnd ? 0 : 1)) >> 1;
    return (a + c + (rnd ? 0 : 1)) >> 1;
}

/* Tap-op count of one interpolated sample (survey unit: 1D = 1, 2D = 5). */
static inline uint64_t taps_per_sample(int mvx, int mvy, int mode) {
    const int hm = mvx & 3, vm = mvy & 3;
    if (!hm && !vm) return 0;
    if (mode_bilinear(mode)) return 1;
    return (hm && vm) ? 5 : 1;
}

/* ---------------------------------------------------------------- hidden padding */

/* motion_uses_hidden_luma, src/encoder_motion.cpp:2143-2164 */
static bool hidden_luma(const ctx_t* c, int x0, int y0, int bw, int bh, mv_t mv, int mode) {
    const int vw = c->vw, vh = c->vh;
    const bool right_padding = vw < c->mbw * 16;
    const bool bottom_padding = vh < c->mbh * 16;
    if (!right_padding && !bottom_padding) return false;
    const int aw = imax(0, imin(bw, vw - x0));
    const int ah = imax(0, imin(bh, vh - y0));
    if (!aw || !ah) return false;
    const int ix = floor_div4(mv.xq), iy = floor_div4(mv.yq);
    const int fx = mv.xq - ix * 4, fy = mv.yq - iy * 4;
    const int right = fx ? (mode_bilinear(mode) ? 1 : 2) : 0;
    const int bottom = fy ? (mode_bilinear(mode) ? 1 : 2) : 0;
    return (right_padding && x0 + aw - 1 + ix + right >= vw) ||
           (bottom_padding && y0 + ah - 1 + iy + bottom >= vh);
}

/* motion_uses_hidden_chroma, src/encoder_motion.cpp:2166-2189 */
static bool hidden_chroma(const ctx_t* c, int x0, int y0, int bw, int bh, mv_t mv, bool future) {
    const int cw = (c->vw + 1) / 2, ch = (c->vh + 1) / 2;
    const bool right_padding = cw < c->mbw * 8;
    const bool bottom_padding = ch < c->mbh * 8;
    if (!right_padding && !bottom_padding) return false;
    const int cx = x0 / 2, cy = y0 / 2;
    const int aw = imax(0, imin(bw / 2, cw - cx));
    const int ah = imax(0, imin(bh / 2, ch - cy));
    if (!aw || !ah) return false;
    int uvx = arshift(mv.xq + ((mv.xq & 3) == 3), 1);
    int uvy = arshift(mv.yq + ((mv.yq & 3) == 3), 1);
    if (future) {
        uvx = uvx < 0 ? uvx - (uvx & 1) : uvx + (uvx & 1);
        uvy = uvy < 0 ? uvy - (uvy & 1) : uvy + (uvy & 1);
    } else {
        uvx = uvx < 0 ? uvx + (uvx & 1) : uvx - (uvx & 1);
        uvy = uvy < 0 ? uvy + (uvy & 1) : uvy - (uvy & 1);
    }
    const int ix = floor_div4(uvx), iy = floor_div4(uvy);
    const bool fx = (uvx - ix * 4) != 0, fy = (uvy - iy * 4) != 0;
    return (right_padding && cx + aw - 1 + ix + (fx ? 1 : 0) >= cw) ||
           (bottom_padding && cy + ah - 1 + iy + (fy ? 1 : 0) >= ch);
}

/* motion_uses_hidden_padding, src/encoder_motion.cpp:2191-2195 */
static bool hidden_padding(const ctx_t* c, int x0, int y0, int bw, int bh, mv_t mv, int mode, bool chroma) {
    return hidden_luma(c, x0, y0, bw, bh, mv, mode) || (chroma && hidden_chroma(c, x0, y0, bw, bh, mv, false));
}

static inline bool chroma_block(int x0, int y0, int bw, int bh) { return bw == 16 && bh == 16 && !(x0 & 15) && !(y0 & 15); }

/* ---------------------------------------------------------------- costs */

/* mb_sad, src/encoder_motion.cpp:2197-2284 (Advanced profile; the SIMD
 * branches compute the same sums). */
static uint64_t mb_sad(const ctx_t* c, int mx, int my, int dx, int dy, int stride) {
    const int x0 = mx * 16, y0 = my * 16;
    const mv_t m = {dx * 4, dy * 4};
    if (hidden_padding(c, x0, y0, 16, 16, m, VC1ME_MODE_ONE_MV_QPEL, true)) return BAD_SAD;
    uint64_t sad = 0, n = 0;
    for (int y = 0; y < 16 && y0 + y < c->H; y += stride) {
        const int sy = clampi(y0 + y + dy, 0, c->H - 1);
        for (int x = 0; x < 16 && x0 + x < c->W; x += stride) {
            const int sx = clampi(x0 + x + dx, 0, c->W - 1);
            sad += (uint64_t)iabs((int)c->cy[(size_t)(y0 + y) * c->W + x0 + x] - (int)c->ry[(size_t)sy * c->W + sx]);
            ++n;
        }
    }
    c->ops->sad_px += n;
    if (stride > 1) sad *= (uint64_t)(stride * stride);
    return sad;
}

/* frame_sad, src/encoder_motion.cpp:1565-1575 */
static uint64_t frame_sad(const ctx_t* c, int dx, int dy, int stride) {
    uint64_t sad = 0, n = 0;
    for (int y = 0; y < c->H; y += stride) {
        const int sy = clampi(y + dy, 0, c->H - 1);
        for (int x = 0; x < c->W; x += stride) {
            const int sx = clampi(x + dx, 0, c->W - 1);
            sad += (uint64_t)iabs((int)c->cy[(size_t)y * c->W + x] - (int)c->ry[(size_t)sy * c->W + sx]);
            ++n;
        }
    }
    c->ops->frame_px += n;
    if (stride > 1) sad *= (uint64_t)(stride * stride);
    return sad;
}

/* block_sad_qpel / block_sad_motion, src/encoder_motion.cpp:2286-2668 and
 * :2670-2933 (scalar law; Advanced profile). */
static uint64_t block_sad_motion(const ctx_t* c, int x0, int y0, int bw, int bh, mv_t mv, int mode) {
    if (hidden_padding(c, x0, y0, bw, bh, mv, mode_bilinear(mode) ? mode : VC1ME_MODE_ONE_MV_QPEL,
                       chroma_block(x0, y0, bw, bh)))
        return BAD_SAD;
    if (mode_bilinear(mode) && ((mv.xq & 1) || (mv.yq & 1))) return BAD_SAD;
    const int aw = imax(0, imin(bw, c->W - x0));
    const int ah = imax(0, imin(bh, c->H - y0));
    uint64_t sad = 0;
    for (int y = 0; y < ah; ++y)
        for (int x = 0; x < aw; ++x) {
            const int pred = luma_mc_sample_mode(c->ry, c->W, c->H, x0 + x, y0 + y, mv.xq, mv.yq, mode, 0);
            sad += (uint64_t)iabs((int)c->cy[(size_t)(y0 + y) * c->W + x0 + x] - pred);
        }
    c->ops->sad_px += (uint64_t)aw * ah;
    c->ops->interp_taps += (uint64_t)aw * ah * taps_per_sample(mv.xq, mv.yq, mode);
    return sad;
}

/* satd4x4_scalar, src/encoder_motion.cpp:2936-2958 */
static uint64_t satd4x4(const uint8_t* cur, int cs, const uint8_t* pred, int ps) {
    int d[4][4], t[4][4];
    for (int y = 0; y < 4; ++y)
        for (int x = 0; x < 4; ++x) d[y][x] = (int)cur[y * cs + x] - (int)pred[y * ps + x];
    for (int y = 0; y < 4; ++y) {
        const int a0 = d[y][0] + d[y][3], a1 = d[y][1] + d[y][2];
        const int a2 = d[y][1] - d[y][2], a3 = d[y][0] - d[y][3];
        t[y][0] = a0 + a1; t[y][1] = a3 + a2; t[y][2] = a0 - a1; t[y][3] = a3 - a2;
    }
    uint64_t tile = 0;
    for (int x = 0; x < 4; ++x) {
        const int a0 = t[0][x] + t[3][x], a1 = t[1][x] + t[2][x];
        const int a2 = t[1][x] - t[2][x], a3 = t[0][x] - t[3][x];
        tile += (uint64_t)iabs(a0 + a1) + (uint64_t)iabs(a3 + a2) + (uint64_t)iabs(a0 - a1) + (uint64_t)iabs(a3 - a2);
    }
    return (tile + 1) >> 1;
}

/* block_satd_motion, src/encoder_motion.cpp:2962-3042, with the prediction
 * rendered by predict_luma_motion_block (src/encoder_motion_compensation.cpp:449-472). */
static uint64_t block_satd_motion(const ctx_t* c, int x0, int y0, int bw, int bh, mv_t mv, int mode) {
    const int aw = imax(0, imin(bw, c->W - x0));
    const int ah = imax(0, imin(bh, c->H - y0));
    if (!aw || !ah) return 0;
    if (mode_halfpel(mode) && ((mv.xq & 1) || (mv.yq & 1))) return BAD_SAD;
    uint8_t pred[16 * 16];
    memset(pred, 0, sizeof pred);
    for (int y = 0; y < ah; ++y)
        for (int x = 0; x < aw; ++x)
            pred[y * 16 + x] = (uint8_t)luma_mc_sample_mode(c->ry, c->W, c->H, x0 + x, y0 + y, mv.xq, mv.yq, mode, 0);
    c->ops->interp_taps += (uint64_t)aw * ah * taps_per_sample(mv.xq, mv.yq, mode);
    uint64_t total = 0;
    for (int by = 0; by < bh; by += 4)
        for (int bx = 0; bx < bw; bx += 4) {
            const int tw = imax(0, imin(4, aw - bx));
            const int th = imax(0, imin(4, ah - by));
            if (!tw || !th) continue;
            uint8_t ct[16], pt[16];
            memset(ct, 0, sizeof ct);
            memset(pt, 0, sizeof pt);
            for (int yy = 0; yy < th; ++yy)
                for (int xx = 0; xx < tw; ++xx) {
                    ct[yy * 4 + xx] = c->cy[(size_t)(y0 + by + yy) * c->W + x0 + bx + xx];
                    pt[yy * 4 + xx] = pred[(by + yy) * 16 + bx + xx];
                }
            total += satd4x4(ct, 4, pt, 4);
            c->ops->satd_ops += 112;
        }
    return total;
}

/* ---------------------------------------------------------------- MV syntax and rate */

/* motion_mvrange, src/encoder_motion.cpp:1257-1263 */
static int motion_mvrange(const ctx_t* c) {
    const int r = c->p->motion_search_range;
    if (r <= 31) return 0;
    if (r <= 63) return 1;
    if (r <= 127) return 2;
    return 3;
}
static int mvrange_x_qpel(int r) { static const int v[4] = {256, 512, 2048, 4096}; return v[clampi(r, 0, 3)]; }
static int mvrange_y_qpel(int r) { static const int v[4] = {128, 256, 512, 1024}; return v[clampi(r, 0, 3)]; }
static int mvrange_kx(int r) { static const int v[4] = {9, 10, 12, 13}; return v[clampi(r, 0, 3)]; }
static int mvrange_ky(int r) { static const int v[4] = {8, 9, 10, 11}; return v[clampi(r, 0, 3)]; }

/* MVDATA code lengths of MV table 0 (ST 421 normative data, the table the
 * reference's rate proxy reads: src/vc1_groupa_tables.h:26-27 row 0). */
static const uint8_t kMvBits0[73] = {
    6, 7, 7, 8, 14, 6, 5, 6, 7, 14, 14, 6, 6, 6, 8, 9, 10, 9, 7, 7, 7, 14, 14, 10, 9,
    10, 8, 10, 10, 14, 14, 14, 13, 13, 13, 6, 3, 5, 6, 8, 9, 13, 5, 4, 4, 5, 7, 9, 6, 5,
    5, 5, 6, 9, 8, 5, 5, 5, 7, 10, 7, 7, 7, 7, 7, 8, 10, 9, 8, 9, 9, 9, 3};

typedef struct { int dx, dy; bool more, intra, halfpel; } mvdata_t;

/* mv_cat, src/encoder_motion.cpp:1150-1162 */
static int mv_cat(int v, bool halfpel) {
    const int a = iabs(v);
    if (!a) return 0;
    if (a <= 2) return 1;
    if (a <= 6) return 2;
    if (a <= 14) return 3;
    if (a <= 30) return 4;
    if (a <= (halfpel ? 94 : 158)) return 5;
    return -1;
}
/* mv_symbol, src/encoder_motion.cpp:1164-1171 */
static int mv_symbol(const mvdata_t* m) {
    if (m->intra) return m->more ? 72 : 35;
    const int xc = mv_cat(m->dx, m->halfpel), yc = mv_cat(m->dy, m->halfpel);
    if (xc < 0 || yc < 0) return m->more ? 71 : 34;
    const int idx = yc * 6 + xc;
    if (!m->more && idx == 0) return 34;
    return m->more ? idx + 36 : idx - 1;
}
/* mv_component_bits, src/encoder_motion.cpp:1173-1176 */
static int mv_component_bits(int cat, bool halfpel) {
    static const int sz[6] = {0, 2, 3, 4, 5, 8};
    return sz[cat] - ((halfpel && cat == 5) ? 1 : 0);
}
/* mv_suffix_bits, src/encoder_motion.cpp:1178-1197 */
static int mv_suffix_bits(const ctx_t* c, const mvdata_t* m) {
    if (m->intra) return 0;
    const int sym = mv_symbol(m);
    const int xc = mv_cat(m->dx, m->halfpel), yc = mv_cat(m->dy, m->halfpel);
    if (sym == 34 || sym == 71 || xc < 0 || yc < 0) {
        const int r = motion_mvrange(c);
        const int hp = m->halfpel ? 1 : 0;
        return (mvrange_kx(r) - hp) + (mvrange_ky(r) - hp);
    }
    return mv_component_bits(xc, m->halfpel) + mv_component_bits(yc, m->halfpel);
}
/* mvdata_for_mode, src/encoder_motion.cpp:1099-1114 (the odd-differential
 * throw cannot trigger: half-pel candidates and predictors are even). */
static mvdata_t mvdata_for_mode(mv_t d, mv_t pred, int mode, bool more) {
    mvdata_t m;
    m.halfpel = mode_halfpel(mode);
    m.more = more;
    m.intra = false;
    m.dx = d.xq - pred.xq;
    m.dy = d.yq - pred.yq;
    if (m.halfpel) { m.dx /= 2; m.dy /= 2; }
    return m;
}
/* motion_vector_rate_bits, src/encoder_motion.cpp:3214-3225 */
static uint64_t mv_rate_bits(const ctx_t* c, mv_t d, mv_t pred, int mode, bool more) {
    const mvdata_t m = mvdata_for_mode(d, pred, mode, more);
    return (uint64_t)kMvBits0[mv_symbol(&m)] + (uint64_t)mv_suffix_bits(c, &m);
}

/* ---------------------------------------------------------------- predictors */

static int median3(int a, int b, int c) {
    const int mn = imin(a, imin(b, c)), mx = imax(a, imax(b, c));
    return a + b + c - mn - mx;
}
/* src/encoder_motion.cpp:1295-1297 */
static int round_qpel_to_pixel(int q) { return q >= 0 ? (q + 2) / 4 : -((-q + 2) / 4); }
/* src/encoder_motion.cpp:1299-1310 */
static int modular_component_distance(int a, int b, int range) {
    int d = a - b;
    const int span = range * 2;
    d = ((d + range) % span + span) % span - range;
    return iabs(d);
}
static int modular_mv_distance(const ctx_t* c, mv_t a, mv_t b) {
    const int r = motion_mvrange(c);
    return modular_component_distance(a.xq, b.xq, mvrange_x_qpel(r)) +
           modular_component_distance(a.yq, b.yq, mvrange_y_qpel(r));
}
static inline bool same_mv(mv_t a, mv_t b) { return a.xq == b.xq && a.yq == b.yq; }

typedef struct { mv_t pre, a, c; bool a_valid, c_valid, hybrid; } pred_t;

/* predictor_info, src/encoder_motion.cpp:1312-1352 */
static pred_t predictor_info(const ctx_t* cx, const mv_t* mv, int mx, int my) {
    const int mbw = cx->mbw, mbh = cx->mbh;
    pred_t o;
    mv_t A = {0, 0}, B = {0, 0}, C = {0, 0};
    const bool a_valid = my > 0, c_valid = mx > 0, b_valid = my > 0 && mbw > 1;
    if (a_valid) A = mv[(size_t)(my - 1) * mbw + mx];
    if (c_valid) C = mv[(size_t)my * mbw + mx - 1];
    if (b_valid) B = mv[(size_t)(my - 1) * mbw + ((mx == mbw - 1) ? mx - 1 : mx + 1)];
    mv_t pre = {0, 0};
    if (a_valid) pre = A; else if (c_valid) pre = C; else if (b_valid) pre = B;
    if ((a_valid ? 1 : 0) + (b_valid ? 1 : 0) + (c_valid ? 1 : 0) > 1) {
        pre.xq = median3(A.xq, B.xq, C.xq);
        pre.yq = median3(A.yq, B.yq, C.yq);
    }
    const int qx = mx * 64, qy = my * 64, X = mbw * 64 - 4, Y = mbh * 64 - 4;
    if (qx + pre.xq < -60) pre.xq = -60 - qx;
    if (qx + pre.xq > X) pre.xq = X - qx;
    if (qy + pre.yq < -60) pre.yq = -60 - qy;
    if (qy + pre.yq > Y) pre.yq = Y - qy;
    bool hybrid = false;
    if (a_valid && c_valid) {
        if (iabs(pre.xq - A.xq) + iabs(pre.yq - A.yq) > 32) hybrid = true;
        else if (iabs(pre.xq - C.xq) + iabs(pre.yq - C.yq) > 32) hybrid = true;
    }
    o.pre = pre; o.a = A; o.c = C; o.a_valid = a_valid; o.c_valid = c_valid; o.hybrid = hybrid;
    return o;
}

static mv_t field_get(const ctx_t* cx, const mv_t* field, int x, int y) {
    mv_t z = {0, 0};
    if (x < 0 || y < 0 || x >= cx->mbw * 2 || y >= cx->mbh * 2) return z;
    return field[((size_t)(y >> 1) * cx->mbw + (x >> 1)) * 4 + (((y & 1) << 1) | (x & 1))];
}
static bool intra_at(const ctx_t* cx, const uint8_t* intra, int x, int y) {
    if (x < 0 || y < 0 || x >= cx->mbw * 2 || y >= cx->mbh * 2) return false;
    return intra[(size_t)(y >> 1) * cx->mbw + (x >> 1)] != 0;
}

/* predictor_info_mixed_1mv, src/encoder_motion.cpp:1354-1406 */
static pred_t predictor_info_mixed_1mv(const ctx_t* cx, const mv_t* field, const uint8_t* intra, int mx, int my) {
    const int mbw = cx->mbw, mbh = cx->mbh;
    pred_t o;
    const int bx = mx * 2, by = my * 2;
    const bool a_valid = my > 0, c_valid = mx > 0, b_valid = my > 0 && mbw > 1;
    mv_t z = {0, 0};
    mv_t A = a_valid ? field_get(cx, field, bx, by - 1) : z;
    mv_t C = c_valid ? field_get(cx, field, bx - 1, by) : z;
    mv_t B = z;
    if (b_valid) B = field_get(cx, field, bx + ((mx == mbw - 1) ? -1 : 2), by - 1);
    mv_t pre = z;
    if (a_valid) pre = A; else if (c_valid) pre = C; else if (b_valid) pre = B;
    if ((a_valid ? 1 : 0) + (b_valid ? 1 : 0) + (c_valid ? 1 : 0) > 1) {
        pre.xq = median3(A.xq, B.xq, C.xq);
        pre.yq = median3(A.yq, B.yq, C.yq);
    }
    const int qx = mx * 64, qy = my * 64, X = mbw * 64 - 4, Y = mbh * 64 - 4;
    if (qx + pre.xq < -60) pre.xq = -60 - qx;
    if (qx + pre.xq > X) pre.xq = X - qx;
    if (qy + pre.yq < -60) pre.yq = -60 - qy;
    if (qy + pre.yq > Y) pre.yq = Y - qy;
    bool hybrid = false;
    if (a_valid && c_valid) {
        const int suma = intra_at(cx, intra, bx, by - 1) ? iabs(pre.xq) + iabs(pre.yq)
                                                         : iabs(pre.xq - A.xq) + iabs(pre.yq - A.yq);
        if (suma > 32) hybrid = true;
        else {
            const int sumc = intra_at(cx, intra, bx - 1, by) ? iabs(pre.xq) + iabs(pre.yq)
                                                             : iabs(pre.xq - C.xq) + iabs(pre.yq - C.yq);
            if (sumc > 32) hybrid = true;
        }
    }
    o.pre = pre; o.a = A; o.c = C; o.a_valid = a_valid; o.c_valid = c_valid; o.hybrid = hybrid;
    return o;
}

/* predictor_info_4mv, src/encoder_motion.cpp:1408-1466 */
static pred_t predictor_info_4mv(const ctx_t* cx, const mv_t* field, const uint8_t* intra, int mx, int my, int n) {
    const int mbw = cx->mbw, mbh = cx->mbh;
    pred_t o;
    const int bx = mx * 2 + (n & 1), by = my * 2 + ((n >> 1) & 1);
    const bool a_valid = my > 0 || n == 2 || n == 3;
    const bool c_valid = mx > 0 || n == 1 || n == 3;
    const bool b_valid = a_valid;
    mv_t z = {0, 0};
    mv_t A = a_valid ? field_get(cx, field, bx, by - 1) : z;
    mv_t C = c_valid ? field_get(cx, field, bx - 1, by) : z;
    mv_t B = z;
    if (b_valid) {
        int bxx = bx;
        if (n == 0) bxx += mx ? -1 : 1;
        else if (n == 1) bxx += (mx == mbw - 1) ? -1 : 1;
        else if (n == 2) bxx += 1;
        else bxx -= 1;
        B = field_get(cx, field, bxx, by - 1);
    }
    mv_t pre = z;
    if (a_valid) pre = A; else if (c_valid) pre = C; else if (b_valid) pre = B;
    if ((a_valid ? 1 : 0) + (b_valid ? 1 : 0) + (c_valid ? 1 : 0) > 1) {
        pre.xq = median3(A.xq, B.xq, C.xq);
        pre.yq = median3(A.yq, B.yq, C.yq);
    }
    const int qx = mx * 64 + ((n == 1 || n == 3) ? 32 : 0);
    const int qy = my * 64 + ((n == 2 || n == 3) ? 32 : 0);
    const int X = mbw * 64 - 4, Y = mbh * 64 - 4;
    if (qx + pre.xq < -28) pre.xq = -28 - qx;
    if (qx + pre.xq > X) pre.xq = X - qx;
    if (qy + pre.yq < -28) pre.yq = -28 - qy;
    if (qy + pre.yq > Y) pre.yq = Y - qy;
    bool hybrid = false;
    if (a_valid && c_valid) {
        const int suma = intra_at(cx, intra, bx, by - 1) ? iabs(pre.xq) + iabs(pre.yq)
                                                         : iabs(pre.xq - A.xq) + iabs(pre.yq - A.yq);
        if (suma > 32) hybrid = true;
        else {
            const int sumc = intra_at(cx, intra, bx - 1, by) ? iabs(pre.xq) + iabs(pre.yq)
                                                             : iabs(pre.xq - C.xq) + iabs(pre.yq - C.yq);
            if (sumc > 32) hybrid = true;
        }
    }
    o.pre = pre; o.a = A; o.c = C; o.a_valid = a_valid; o.c_valid = c_valid; o.hybrid = hybrid;
    return o;
}

/* ---------------------------------------------------------------- bounds */

typedef struct { int xmin, xmax, ymin, ymax; } bounds_t;
/* motion_search_bounds_px, src/encoder_motion.cpp:1238-1246 */
static bounds_t bounds_px(int requested) {
    const int r = clampi(requested, 0, 1024);
    bounds_t b = {-imin(r, 1024), imin(r, 1023), -imin(r, 256), imin(r, 255)};
    return b;
}
/* motion_search_bounds_qpel, src/encoder_motion.cpp:1248-1255 */
static bounds_t bounds_qpel(int requested) {
    const int r = clampi(requested, 0, 1024);
    bounds_t b;
    b.xmin = -imin(r * 4, 4096);
    b.xmax = (r >= 1024) ? 4095 : imin(r * 4, 4095);
    b.ymin = -imin(r * 4, 1024);
    b.ymax = (r >= 256) ? 1023 : imin(r * 4, 1023);
    return b;
}

/* ---------------------------------------------------------------- global search */

typedef struct { uint64_t cost; int dx, dy; } cand_t;

/* the (cost, |d|, dy, dx) total order shared by the integer searches */
static bool better4(uint64_t s, int dx, int dy, uint64_t bs, int bx, int by) {
    if (s != bs) return s < bs;
    const int m = iabs(dx) + iabs(dy), bm = iabs(bx) + iabs(by);
    if (m != bm) return m < bm;
    if (dy != by) return dy < by;
    return dx < bx;
}
static bool cand_better(const cand_t* a, const cand_t* b) { return better4(a->cost, a->dx, a->dy, b->cost, b->dx, b->dy); }

static void sort_cands(cand_t* v, int n) {
    for (int i = 1; i < n; ++i) {
        cand_t x = v[i];
        int j = i - 1;
        while (j >= 0 && cand_better(&x, &v[j])) { v[j + 1] = v[j]; --j; }
        v[j + 1] = x;
    }
}

/* global_motion_search_extended, src/encoder_motion.cpp:2025-2074 */
typedef struct { cand_t v[9]; int n; } topg_t;
static void gconsider(const ctx_t* c, topg_t* top, int range, int dx, int dy, int stride) {
    dx = clampi(dx, -range, range);
    dy = clampi(dy, -range, range);
    cand_t cd = {frame_sad(c, dx, dy, stride), dx, dy};
    for (int i = 0; i < top->n; ++i)
        if (top->v[i].dx == dx && top->v[i].dy == dy) {
            if (cd.cost < top->v[i].cost) top->v[i] = cd;
            return;
        }
    top->v[top->n++] = cd;
    sort_cands(top->v, top->n);
    if (top->n > 8) top->n = 8;
}
static cand_t global_extended(const ctx_t* c, int range) {
    topg_t top;
    top.n = 0;
    gconsider(c, &top, range, 0, 0, 16);
    const int step = range >= 160 ? 16 : 8;
    for (int d = step; d <= range; d += step) {
        gconsider(c, &top, range, -d, 0, 16); gconsider(c, &top, range, d, 0, 16);
        gconsider(c, &top, range, 0, -d, 16); gconsider(c, &top, range, 0, d, 16);
    }
    for (int r = step; r <= range; r += step) {
        const int h = (r + 1) / 2;
        gconsider(c, &top, range, -r, -h, 16); gconsider(c, &top, range, -r, h, 16);
        gconsider(c, &top, range, r, -h, 16); gconsider(c, &top, range, r, h, 16);
        gconsider(c, &top, range, -h, -r, 16); gconsider(c, &top, range, h, -r, 16);
        gconsider(c, &top, range, -h, r, 16); gconsider(c, &top, range, h, r, 16);
    }
    topg_t coarse = top;
    for (int i = 0; i < imin(4, coarse.n); ++i) {
        const cand_t cc = coarse.v[i];
        for (int oy = -step; oy <= step; oy += 4)
            for (int ox = -step; ox <= step; ox += 4) gconsider(c, &top, range, cc.dx + ox, cc.dy + oy, 8);
    }
    if (top.n == 0) gconsider(c, &top, range, 0, 0, 8);
    const cand_t mid = top.v[0];
    for (int oy = -4; oy <= 4; ++oy)
        for (int ox = -4; ox <= 4; ++ox) gconsider(c, &top, range, mid.dx + ox, mid.dy + oy, 4);
    const cand_t fine = top.v[0];
    cand_t full = {U64MAX, 0, 0};
    for (int oy = -1; oy <= 1; ++oy)
        for (int ox = -1; ox <= 1; ++ox) {
            const int dx = clampi(fine.dx + ox, -range, range), dy = clampi(fine.dy + oy, -range, range);
            cand_t cd = {frame_sad(c, dx, dy, 1), dx, dy};
            if (cand_better(&cd, &full)) full = cd;
        }
    return full;
}

/* global_motion_search, src/encoder_motion.cpp:2076-2141 */
static cand_t global_search(const ctx_t* c, int range) {
    range = imin(range, 32);
    if (range <= 0) { cand_t r = {frame_sad(c, 0, 0, 1), 0, 0}; return r; }
    if (range > 31) return global_extended(c, range);
    const int coarse_step = range >= 4 ? 4 : (range >= 2 ? 2 : 1);
    cand_t top[9];
    int n = 0;
#define GADD(cost_, dx_, dy_)                           \
    do {                                                \
        cand_t cd_ = {(cost_), (dx_), (dy_)};           \
        top[n++] = cd_;                                 \
        sort_cands(top, n);                             \
        if (n > 8) n = 8;                               \
    } while (0)
    for (int dy = -range; dy <= range; dy += coarse_step)
        for (int dx = -range; dx <= range; dx += coarse_step) GADD(frame_sad(c, dx, dy, 4), dx, dy);
    GADD(frame_sad(c, 0, 0, 4), 0, 0);
    for (int d = -range; d <= range; d += coarse_step) {
        GADD(frame_sad(c, range, d, 4), range, d);
        GADD(frame_sad(c, d, range, 4), d, range);
    }
#undef GADD
    cand_t best = {U64MAX, 0, 0};
    uint8_t seen[63 * 63];
    memset(seen, 0, sizeof seen);
    for (int i = 0; i < n; ++i) {
        const cand_t cc = top[i];
        for (int dy = imax(-range, cc.dy - coarse_step + 1); dy <= imin(range, cc.dy + coarse_step - 1); ++dy)
            for (int dx = imax(-range, cc.dx - coarse_step + 1); dx <= imin(range, cc.dx + coarse_step - 1); ++dx) {
                const size_t key = (size_t)(dy + 31) * 63 + (size_t)(dx + 31);
                if (seen[key]) continue;
                seen[key] = 1;
                const uint64_t cost = frame_sad(c, dx, dy, 2);
                const int cm = iabs(dx) + iabs(dy), bm = iabs(best.dx) + iabs(best.dy);
                if (cost < best.cost || (cost == best.cost && cm < bm)) { best.cost = cost; best.dx = dx; best.dy = dy; }
            }
    }
    cand_t r = {frame_sad(c, best.dx, best.dy, 1), best.dx, best.dy};
    return r;
}

/* ---------------------------------------------------------------- integer searches */

static uint64_t active_mb_pixels(const ctx_t* c, int mx, int my) {
    const int x0 = mx * 16, y0 = my * 16;
    return (uint64_t)imax(0, imin(16, c->W - x0)) * (uint64_t)imax(0, imin(16, c->H - y0));
}
/* motion_match_good, src/encoder_motion.cpp:1633-1640 */
static bool match_good(const ctx_t* c, uint64_t sad, int mx, int my) {
    const uint64_t px = active_mb_pixels(c, mx, my);
    if (!px || sad >= U64MAX / 8) return false;
    return sad <= px * 6u;
}
/* distant_match_decision, src/encoder_motion.cpp:1642-1656 */
static int distant_decision(const ctx_t* c, uint64_t cand, uint64_t local, int mx, int my) {
    const uint64_t px = active_mb_pixels(c, mx, my);
    if (!px || cand >= U64MAX / 8) return 0;
    if (cand <= px * 6u) return 2;
    const long double max_sad = (long double)px * (long double)c->p->distant_match_max_mae;
    if ((long double)cand > max_sad) return 1;
    if (local >= U64MAX / 8) return 3;
    return cand * 4u <= local * 3u ? 3 : 1;
}

typedef struct { int dx, dy; uint64_t sad; } imr_t;

#define SEEN_R 64
#define SEEN_SIDE (2 * SEEN_R + 1)
typedef struct { uint64_t w[(SEEN_SIDE * SEEN_SIDE + 63) / 64]; } seen_t;
static bool seen_test_set(seen_t* s, int range, int dx, int dy) {
    if (range > SEEN_R) return false;
    const size_t bit = (size_t)(dy + SEEN_R) * SEEN_SIDE + (size_t)(dx + SEEN_R);
    uint64_t* word = &s->w[bit >> 6];
    const uint64_t mask = (uint64_t)1 << (bit & 63);
    if (*word & mask) return true;
    *word |= mask;
    return false;
}

typedef struct {
    const ctx_t* c;
    int mx, my, range;
    cand_t top[8];
    int top_count;
    cand_t best;
    seen_t seen;
} umh_t;

/* the consider lambda of integer_motion_umh, src/encoder_motion.cpp:1820-1838 */
static void umh_consider(umh_t* u, int dx, int dy) {
    dx = clampi(dx, -u->range, u->range);
    dy = clampi(dy, -u->range, u->range);
    if (seen_test_set(&u->seen, u->range, dx, dy)) return;
    const uint64_t cost = mb_sad(u->c, u->mx, u->my, dx, dy, 2);
    if (better4(cost, dx, dy, u->best.cost, u->best.dx, u->best.dy)) { u->best.cost = cost; u->best.dx = dx; u->best.dy = dy; }
    for (int i = 0; i < u->top_count; ++i)
        if (u->top[i].dx == dx && u->top[i].dy == dy) {
            if (cost < u->top[i].cost) u->top[i].cost = cost;
            return;
        }
    int at = 0;
    while (at < u->top_count && !better4(cost, dx, dy, u->top[at].cost, u->top[at].dx, u->top[at].dy)) ++at;
    if (at < 8) {
        const int new_count = imin(u->top_count + 1, 8);
        for (int j = new_count - 1; j > at; --j) u->top[j] = u->top[j - 1];
        u->top[at].cost = cost; u->top[at].dx = dx; u->top[at].dy = dy;
        u->top_count = new_count;
    }
}

/* integer_motion_umh, src/encoder_motion.cpp:1781-2023 (CPU path: every
 * score_integer_batch falls back to mb_sad). */
static imr_t integer_umh(const ctx_t* c, int mx, int my, const pred_t* pi, cand_t global, int range) {
    umh_t* u = (umh_t*)calloc(1, sizeof(umh_t));
    u->c = c; u->mx = mx; u->my = my; u->range = range;
    u->best.cost = U64MAX;
    const int pdx = clampi(round_qpel_to_pixel(pi->pre.xq), -range, range);
    const int pdy = clampi(round_qpel_to_pixel(pi->pre.yq), -range, range);
    umh_consider(u, pdx, pdy);
    umh_consider(u, 0, 0);
    umh_consider(u, global.dx, global.dy);
    if (pi->a_valid) umh_consider(u, round_qpel_to_pixel(pi->a.xq), round_qpel_to_pixel(pi->a.yq));
    if (pi->c_valid) umh_consider(u, round_qpel_to_pixel(pi->c.xq), round_qpel_to_pixel(pi->c.yq));
    {
        const int cx = u->best.dx, cy = u->best.dy;
        for (int oy = -2; oy <= 2; ++oy)
            for (int ox = -2; ox <= 2; ++ox) umh_consider(u, cx + ox, cy + oy);
    }
    if (range >= 8) {
        const int lattice = 8;
        for (int dy = -range; dy <= range; dy += lattice)
            for (int dx = -range; dx <= range; dx += lattice) umh_consider(u, dx, dy);
        for (int d = -range; d <= range; d += lattice) { umh_consider(u, range, d); umh_consider(u, d, range); }
    }
    {
        const int cross_step = range >= 96 ? 8 : (range >= 32 ? 4 : 2);
        const int cx = u->best.dx, cy = u->best.dy;
        const int vlimit = imin(range, imax(8, range / 2));
        for (int d = cross_step; d <= range; d += cross_step) { umh_consider(u, cx - d, cy); umh_consider(u, cx + d, cy); }
        for (int d = cross_step; d <= vlimit; d += cross_step) { umh_consider(u, cx, cy - d); umh_consider(u, cx, cy + d); }
    }
    {
        const int hex_step = range >= 96 ? 8 : (range >= 24 ? 4 : 2);
        const int hx = u->best.dx, hy = u->best.dy;
        for (int r = hex_step; r <= range; r += hex_step) {
            const int h = (r + 1) / 2;
            umh_consider(u, hx - r, hy); umh_consider(u, hx + r, hy);
            umh_consider(u, hx - h, hy - r); umh_consider(u, hx + h, hy - r);
            umh_consider(u, hx - h, hy + r); umh_consider(u, hx + h, hy + r);
        }
    }
    for (int iter = 0; iter < 32; ++iter) {
        const int ox = u->best.dx, oy = u->best.dy;
        umh_consider(u, ox - 2, oy); umh_consider(u, ox + 2, oy);
        umh_consider(u, ox - 1, oy - 2); umh_consider(u, ox + 1, oy - 2);
        umh_consider(u, ox - 1, oy + 2); umh_consider(u, ox + 1, oy + 2);
        if (u->best.dx == ox && u->best.dy == oy) break;
    }
    {
        const int ox = u->best.dx, oy = u->best.dy;
        umh_consider(u, ox - 1, oy); umh_consider(u, ox + 1, oy);
        umh_consider(u, ox, oy - 1); umh_consider(u, ox, oy + 1);
    }
    /* finalists at full resolution */
    cand_t full = {U64MAX, 0, 0};
    seen_t* sf = (seen_t*)calloc(1, sizeof(seen_t));
    const int finalists = imin(4, u->top_count);
    for (int i = 0; i < finalists; ++i) {
        const cand_t cc = u->top[i];
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) {
                const int dx = clampi(cc.dx + ox, -range, range), dy = clampi(cc.dy + oy, -range, range);
                if (seen_test_set(sf, range, dx, dy)) continue;
                const uint64_t sad = mb_sad(c, mx, my, dx, dy, 1);
                if (better4(sad, dx, dy, full.cost, full.dx, full.dy)) { full.cost = sad; full.dx = dx; full.dy = dy; }
            }
    }
    if (full.cost == U64MAX) {
        (void)seen_test_set(sf, range, u->best.dx, u->best.dy);
        full.cost = mb_sad(c, mx, my, u->best.dx, u->best.dy, 1);
        full.dx = u->best.dx; full.dy = u->best.dy;
    }
    if (range > 0 && !match_good(c, full.cost, mx, my)) {
        cand_t coarse = {U64MAX, 0, 0};
        const int step = range >= 8 ? 4 : (range >= 3 ? 2 : 1);
        for (int dy = -range; dy <= range; dy += step)
            for (int dx = -range; dx <= range; dx += step) {
                const uint64_t sad = mb_sad(c, mx, my, dx, dy, 4);
                if (better4(sad, dx, dy, coarse.cost, coarse.dx, coarse.dy)) { coarse.cost = sad; coarse.dx = dx; coarse.dy = dy; }
            }
        for (int d = -range; d <= range; d += step) {
            uint64_t sad = mb_sad(c, mx, my, range, d, 4);
            if (better4(sad, range, d, coarse.cost, coarse.dx, coarse.dy)) { coarse.cost = sad; coarse.dx = range; coarse.dy = d; }
            sad = mb_sad(c, mx, my, d, range, 4);
            if (better4(sad, d, range, coarse.cost, coarse.dx, coarse.dy)) { coarse.cost = sad; coarse.dx = d; coarse.dy = range; }
        }
        cand_t rescue = {U64MAX, coarse.dx, coarse.dy};
        const int rr = imax(1, step - 1);
        for (int dy = imax(-range, coarse.dy - rr); dy <= imin(range, coarse.dy + rr); ++dy)
            for (int dx = imax(-range, coarse.dx - rr); dx <= imin(range, coarse.dx + rr); ++dx) {
                if (seen_test_set(sf, range, dx, dy)) continue;
                const uint64_t sad = mb_sad(c, mx, my, dx, dy, 1);
                if (better4(sad, dx, dy, rescue.cost, rescue.dx, rescue.dy)) { rescue.cost = sad; rescue.dx = dx; rescue.dy = dy; }
            }
        if (better4(rescue.cost, rescue.dx, rescue.dy, full.cost, full.dx, full.dy)) full = rescue;
        const uint64_t pixels = active_mb_pixels(c, mx, my);
        if (pixels && full.cost > pixels * 24u && range <= 32) {
            cand_t dense = {U64MAX, 0, 0};
            for (int dy = -range; dy <= range; dy += 2)
                for (int dx = -range; dx <= range; dx += 2) {
                    const uint64_t sad = mb_sad(c, mx, my, dx, dy, 2);
                    if (better4(sad, dx, dy, dense.cost, dense.dx, dense.dy)) { dense.cost = sad; dense.dx = dx; dense.dy = dy; }
                }
            cand_t dfull = {U64MAX, dense.dx, dense.dy};
            for (int dy = imax(-range, dense.dy - 2); dy <= imin(range, dense.dy + 2); ++dy)
                for (int dx = imax(-range, dense.dx - 2); dx <= imin(range, dense.dx + 2); ++dx) {
                    if (seen_test_set(sf, range, dx, dy)) continue;
                    const uint64_t sad = mb_sad(c, mx, my, dx, dy, 1);
                    if (better4(sad, dx, dy, dfull.cost, dfull.dx, dfull.dy)) { dfull.cost = sad; dfull.dx = dx; dfull.dy = dy; }
                }
            if (better4(dfull.cost, dfull.dx, dfull.dy, full.cost, full.dx, full.dy)) full = dfull;
        }
    }
    imr_t r = {full.dx, full.dy, full.cost};
    free(sf);
    free(u);
    return r;
}

/* integer_motion_diamond, src/encoder_motion_diamond.cpp:13-73 */
static imr_t integer_diamond(const ctx_t* c, int mx, int my, const pred_t* pi, cand_t global, int range) {
    imr_t best = {0, 0, mb_sad(c, mx, my, 0, 0, 1)};
    if (best.sad == 0) { imr_t z = {0, 0, 0}; return z; }
    int sx[5] = {0}, sy[5] = {0}, sn = 1;
#define DSEED(dx_, dy_)                                                          \
    do {                                                                         \
        int dx = clampi((dx_), -range, range), dy = clampi((dy_), -range, range); \
        bool dup = false;                                                        \
        for (int i = 0; i < sn; ++i) if (dx == sx[i] && dy == sy[i]) dup = true; \
        if (!dup) {                                                              \
            sx[sn] = dx; sy[sn] = dy; ++sn;                                      \
            const uint64_t s = mb_sad(c, mx, my, dx, dy, 1);                     \
            if (better4(s, dx, dy, best.sad, best.dx, best.dy)) {                \
                best.dx = dx; best.dy = dy; best.sad = s;                        \
            }                                                                    \
        }                                                                        \
    } while (0)
    DSEED(round_qpel_to_pixel(pi->pre.xq), round_qpel_to_pixel(pi->pre.yq));
    DSEED(global.dx, global.dy);
    if (pi->a_valid) DSEED(round_qpel_to_pixel(pi->a.xq), round_qpel_to_pixel(pi->a.yq));
    if (pi->c_valid) DSEED(round_qpel_to_pixel(pi->c.xq), round_qpel_to_pixel(pi->c.yq));
#undef DSEED
    int prevx = best.dx, prevy = best.dy;
    bool have_prev = false;
    static const int off[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
    for (int step = 0; step <= 4 * range + 1; ++step) {
        const imr_t center = best;
        imr_t win = center;
        for (int k = 0; k < 4; ++k) {
            const int dx = center.dx + off[k][0], dy = center.dy + off[k][1];
            if (dx < -range || dx > range || dy < -range || dy > range) continue;
            if (have_prev && dx == prevx && dy == prevy) continue;
            const uint64_t s = mb_sad(c, mx, my, dx, dy, 1);
            if (better4(s, dx, dy, win.sad, win.dx, win.dy)) { win.dx = dx; win.dy = dy; win.sad = s; }
        }
        if (win.dx == center.dx && win.dy == center.dy) break;
        prevx = center.dx; prevy = center.dy; have_prev = true;
        best = win;
        if (best.sad == 0) break;
    }
    return best;
}

/* integer_motion_local, src/encoder_motion_diamond.cpp:5-11 */
static imr_t integer_local(const ctx_t* c, int mx, int my, const pred_t* pi, cand_t global, int range) {
    if (c->p->me_algorithm == 1) return integer_diamond(c, mx, my, pi, global, range);
    return integer_umh(c, mx, my, pi, global, range);
}

/* ---------------------------------------------------------------- long range */

typedef struct { uint8_t ydc[4], udc, vdc; } sig_t;
typedef struct { int x, y; sig_t s; } cell_t;

/* mean_region, src/encoder_motion.cpp:1578-1589 */
static uint8_t mean_region(const uint8_t* p, int pw, int ph, int x0, int y0, int bw, int bh) {
    uint64_t sum = 0;
    int n = 0;
    for (int y = 0; y < bh && y0 + y < ph; ++y) {
        if (y0 + y < 0) continue;
        for (int x = 0; x < bw && x0 + x < pw; ++x) {
            if (x0 + x < 0) continue;
            sum += p[(size_t)(y0 + y) * pw + x0 + x];
            ++n;
        }
    }
    return n ? (uint8_t)((sum + (uint64_t)(n / 2)) / (uint64_t)n) : 128;
}
/* long_range_signature, src/encoder_motion.cpp:1591-1601 */
static sig_t signature(const uint8_t* y, const uint8_t* u, const uint8_t* v, int W, int H, int x0, int y0) {
    sig_t s;
    s.ydc[0] = mean_region(y, W, H, x0, y0, 8, 8);
    s.ydc[1] = mean_region(y, W, H, x0 + 8, y0, 8, 8);
    s.ydc[2] = mean_region(y, W, H, x0, y0 + 8, 8, 8);
    s.ydc[3] = mean_region(y, W, H, x0 + 8, y0 + 8, 8, 8);
    const int cw = W / 2, ch = H / 2;
    s.udc = mean_region(u, cw, ch, x0 / 2, y0 / 2, 8, 8);
    s.vdc = mean_region(v, cw, ch, x0 / 2, y0 / 2, 8, 8);
    return s;
}
/* long_range_signature_distance, src/encoder_motion.cpp:1603-1613 */
static unsigned sig_distance(const sig_t* a, const sig_t* b) {
    unsigned d = 0;
    for (int i = 0; i < 4; ++i) d += (unsigned)iabs((int)a->ydc[i] - (int)b->ydc[i]) * 2u;
    d += (unsigned)iabs((int)a->udc - (int)b->udc) * 3u;
    d += (unsigned)iabs((int)a->vdc - (int)b->vdc) * 3u;
    return d;
}

/* propagated_long_range_motion, src/encoder_motion.cpp:1658-1694 */
static imr_t propagated_long_range(const ctx_t* c, int mx, int my, const mv_t* field, const uint8_t* prop, int requested) {
    imr_t best = {0, 0, U64MAX};
    const bounds_t b = bounds_px(requested);
    static const int offs[13][2] = {{0, 0}, {-1, 0}, {1, 0}, {0, -1}, {0, 1}, {-2, 0}, {2, 0},
                                    {0, -2}, {0, 2}, {-1, -1}, {1, -1}, {-1, 1}, {1, 1}};
    static const int nb[4][2] = {{-1, 0}, {0, -1}, {-1, -1}, {1, -1}};
    for (int k = 0; k < 4; ++k) {
        const int nx = mx + nb[k][0], ny = my + nb[k][1];
        if (nx < 0 || ny < 0 || nx >= c->mbw) continue;
        const size_t pos = (size_t)ny * c->mbw + nx;
        if (pos >= (size_t)c->mbw * c->mbh || !prop[pos]) continue;
        const int cx = round_qpel_to_pixel(field[pos].xq), cy = round_qpel_to_pixel(field[pos].yq);
        for (int i = 0; i < 13; ++i) {
            const int dx = clampi(cx + offs[i][0], b.xmin, b.xmax);
            const int dy = clampi(cy + offs[i][1], b.ymin, b.ymax);
            const uint64_t sad = mb_sad(c, mx, my, dx, dy, 1);
            if (sad < best.sad || (sad == best.sad && iabs(dx) + iabs(dy) < iabs(best.dx) + iabs(best.dy))) {
                best.dx = dx; best.dy = dy; best.sad = sad;
            }
        }
        if (match_good(c, best.sad, mx, my)) break;
    }
    return best;
}

typedef struct { unsigned score; int dx, dy; } sigc_t;
static bool sig_less(const sigc_t* a, const sigc_t* b) {
    if (a->score != b->score) return a->score < b->score;
    const int am = iabs(a->dx) + iabs(a->dy), bm = iabs(b->dx) + iabs(b->dy);
    if (am != bm) return am < bm;
    if (a->dy != b->dy) return a->dy < b->dy;
    return a->dx < b->dx;
}

/* extended_content_motion, src/encoder_motion.cpp:1696-1779 */
static imr_t extended_content(const ctx_t* c, int mx, int my, int requested, const cell_t* cells, int ncells,
                              imr_t local, uint64_t* cand_sad, int* cand_decision) {
    const int x0 = mx * 16, y0 = my * 16;
    const sig_t want = signature(c->cy, c->cu, c->cv, c->W, c->H, x0, y0);
    const bounds_t b = bounds_px(requested);
    sigc_t top[17];
    int tn = 0;
    for (int i = 0; i < ncells; ++i) {
        const int dx = cells[i].x - x0, dy = cells[i].y - y0;
        if (dx < b.xmin || dx > b.xmax || dy < b.ymin || dy > b.ymax) continue;
        if (iabs(dx) <= 32 && iabs(dy) <= 32) continue;
        c->ops->sig_ops++;
        sigc_t cd = {sig_distance(&want, &cells[i].s), dx, dy};
        if (tn == 16 && !sig_less(&cd, &top[15])) continue;
        int at = 0;
        while (at < tn && sig_less(&top[at], &cd)) ++at; /* lower_bound */
        for (int j = tn; j > at; --j) top[j] = top[j - 1];
        top[at] = cd;
        ++tn;
        if (tn > 16) tn = 16;
    }
    imr_t best = {0, 0, U64MAX};
    for (int t = 0; t < tn; ++t) {
        const sigc_t cc = top[t];
        int sx = cc.dx, sy = cc.dy;
        uint64_t sampled = U64MAX;
        for (int oy = -8; oy <= 8; oy += 4)
            for (int ox = -8; ox <= 8; ox += 4) {
                const int dx = clampi(cc.dx + ox, b.xmin, b.xmax), dy = clampi(cc.dy + oy, b.ymin, b.ymax);
                const uint64_t sad = mb_sad(c, mx, my, dx, dy, 2);
                if (sad < sampled) { sampled = sad; sx = dx; sy = dy; }
            }
        imr_t full = {sx, sy, U64MAX};
        for (int oy = -2; oy <= 2; ++oy)
            for (int ox = -2; ox <= 2; ++ox) {
                const int dx = clampi(sx + ox, b.xmin, b.xmax), dy = clampi(sy + oy, b.ymin, b.ymax);
                const uint64_t sad = mb_sad(c, mx, my, dx, dy, 1);
                if (sad < full.sad || (sad == full.sad && iabs(dx) + iabs(dy) < iabs(full.dx) + iabs(full.dy))) {
                    full.dx = dx; full.dy = dy; full.sad = sad;
                }
            }
        if (full.sad < best.sad) best = full;
        if (match_good(c, best.sad, mx, my)) break;
    }
    if (cand_sad) *cand_sad = best.sad;
    const int decision = best.sad < U64MAX / 8 ? distant_decision(c, best.sad, local.sad, mx, my) : 0;
    if (cand_decision) *cand_decision = decision;
    if (best.sad >= local.sad || decision < 2) return local;
    return best;
}

/* ---------------------------------------------------------------- sub-pel refinement */

/* refine_motion_integer_only, src/encoder_motion.cpp:3247-3269 */
static mv_t refine_integer_only(const ctx_t* c, int x0, int y0, int bw, int bh, mv_t seed, int range, int mode,
                                uint64_t* best_sad, uint64_t known_seed_sad) {
    const bounds_t qb = bounds_qpel(range);
#define SNAP(q, lo, hi) clampi(((clampi((q), (lo), (hi)) >= 0) ? (clampi((q), (lo), (hi)) + 1) / 4 \
                                                               : (clampi((q), (lo), (hi)) - 1) / 4) * 4, \
                               (lo), (hi) / 4 * 4)
    mv_t integer = {SNAP(seed.xq, qb.xmin, qb.xmax), SNAP(seed.yq, qb.ymin, qb.ymax)};
#undef SNAP
    if (hidden_padding(c, x0, y0, bw, bh, integer, mode, chroma_block(x0, y0, bw, bh))) integer.xq = integer.yq = 0;
    if (best_sad)
        *best_sad = (integer.xq == seed.xq && integer.yq == seed.yq && known_seed_sad != U64MAX)
                        ? known_seed_sad
                        : block_sad_motion(c, x0, y0, bw, bh, integer, mode);
    return integer;
}

typedef struct {
    mv_t mv;
    uint64_t sad, rate, satd;
    long double score;
    bool eligible;
} sub_t;

static bool tie_less(mv_t a, mv_t b) {
    const int am = iabs(a.xq) + iabs(a.yq), bm = iabs(b.xq) + iabs(b.yq);
    if (am != bm) return am < bm;
    if (a.yq != b.yq) return a.yq < b.yq;
    return a.xq < b.xq;
}

typedef struct {
    const ctx_t* c;
    int x0, y0, bw, bh, mode;
    sub_t all[32];
    int n;
} staged_t;

static sub_t* st_find(staged_t* s, mv_t mv) {
    for (int i = 0; i < s->n; ++i)
        if (same_mv(s->all[i].mv, mv)) return &s->all[i];
    return NULL;
}
static void st_add(staged_t* s, mv_t mv, uint64_t known) {
    if (st_find(s, mv)) return;
    sub_t cd;
    memset(&cd, 0, sizeof cd);
    cd.mv = mv;
    cd.sad = hidden_padding(s->c, s->x0, s->y0, s->bw, s->bh, mv, s->mode, chroma_block(s->x0, s->y0, s->bw, s->bh))
                 ? BAD_SAD
                 : (known != U64MAX ? known : block_sad_motion(s->c, s->x0, s->y0, s->bw, s->bh, mv, s->mode));
    cd.eligible = true;
    s->all[s->n++] = cd;
}
static int make_even(int v) { return (v & 1) ? (v > 0 ? v - 1 : v + 1) : v; }

static bool sad_order(const sub_t* a, const sub_t* b) {
    if (a->sad != b->sad) return a->sad < b->sad;
    return tie_less(a->mv, b->mv);
}
static bool score_order(const sub_t* a, const sub_t* b) {
    if (a->score != b->score) return a->score < b->score;
    if (a->sad != b->sad) return a->sad < b->sad;
    return tie_less(a->mv, b->mv);
}
static bool final_order(const sub_t* a, const sub_t* b) {
    if (a->eligible != b->eligible) return a->eligible > b->eligible;
    if (a->score != b->score) return a->score < b->score;
    if (a->satd != b->satd) return a->satd < b->satd;
    if (a->sad != b->sad) return a->sad < b->sad;
    return tie_less(a->mv, b->mv);
}
static void sort_sub(sub_t* v, int n, bool (*less)(const sub_t*, const sub_t*)) {
    for (int i = 1; i < n; ++i) {
        sub_t x = v[i];
        int j = i - 1;
        while (j >= 0 && less(&x, &v[j])) { v[j + 1] = v[j]; --j; }
        v[j + 1] = x;
    }
}

/* lambda_me = max(1, 0.35*pq) as a double, widened to long double
 * (src/encoder_motion.cpp:3458) */
static long double lambda_me(int pq) {
    const double l = 0.35 * (double)pq;
    return (long double)(l > 1.0 ? l : 1.0);
}

/* refine_motion_block_staged, src/encoder_motion.cpp:3271-3594, for
 * me_quality RATE and SATD (RD needs the transform coder and is not a
 * motion-estimation configuration this backend offers). */
static mv_t refine_staged(const ctx_t* c, int x0, int y0, int bw, int bh, mv_t seed, int range, int mode,
                          const pred_t* pi, bool more, uint64_t* best_sad, uint64_t known_seed_sad) {
    const vc1me_params* P = c->p;
    const int limit_q = 4 * P->subpixel_refinement_range;
    if (!P->subpixel_refinement || limit_q == 0 || imax(iabs(seed.xq), iabs(seed.yq)) > limit_q)
        return refine_integer_only(c, x0, y0, bw, bh, seed, range, mode, best_sad, known_seed_sad);
    staged_t* s = (staged_t*)calloc(1, sizeof(staged_t));
    s->c = c; s->x0 = x0; s->y0 = y0; s->bw = bw; s->bh = bh; s->mode = mode;
    const bounds_t qb = bounds_qpel(range);
    const int xmin = imax(qb.xmin, -limit_q), xmax = imin(qb.xmax, limit_q);
    const int ymin = imax(qb.ymin, -limit_q), ymax = imin(qb.ymax, limit_q);
    const bool hp = mode_halfpel(mode);
    seed.xq = clampi(seed.xq, xmin, xmax);
    seed.yq = clampi(seed.yq, ymin, ymax);
    if (hp) { seed.xq = make_even(seed.xq); seed.yq = make_even(seed.yq); }
    st_add(s, seed, known_seed_sad);
    if (hidden_padding(c, x0, y0, bw, bh, seed, mode, chroma_block(x0, y0, bw, bh))) {
        mv_t z = {0, 0};
        st_add(s, z, U64MAX);
    }
    mv_t half_best = seed;
    uint64_t half_sad = s->all[0].sad;
    const bool fast = P->fast_subpixel_refinement != 0;
    mv_t batch[9];
    int nb = 0;
    if (fast) {
        batch[nb++] = seed;
        static const int o4[4][2] = {{-2, 0}, {2, 0}, {0, -2}, {0, 2}};
        for (int k = 0; k < 4; ++k) {
            mv_t mv = {clampi(seed.xq + o4[k][0], xmin, xmax), clampi(seed.yq + o4[k][1], ymin, ymax)};
            if (hp) { mv.xq = make_even(mv.xq); mv.yq = make_even(mv.yq); }
            batch[nb++] = mv;
        }
    } else {
        for (int oy = -2; oy <= 2; oy += 2)
            for (int ox = -2; ox <= 2; ox += 2) {
                mv_t mv = {clampi(seed.xq + ox, xmin, xmax), clampi(seed.yq + oy, ymin, ymax)};
                if (hp) { mv.xq = make_even(mv.xq); mv.yq = make_even(mv.yq); }
                batch[nb++] = mv;
            }
    }
    for (int i = 0; i < nb; ++i) st_add(s, batch[i], U64MAX);
    for (int i = 0; i < nb; ++i) {
        sub_t* cd = st_find(s, batch[i]);
        if (cd && (cd->sad < half_sad || (cd->sad == half_sad && tie_less(cd->mv, half_best)))) {
            half_sad = cd->sad;
            half_best = cd->mv;
        }
    }
    if (fast && !same_mv(half_best, seed)) {
        mv_t dg[2];
        int nd = 0;
        const int dx = half_best.xq - seed.xq, dy = half_best.yq - seed.yq;
        if (dx && !dy) {
            dg[nd].xq = half_best.xq; dg[nd].yq = clampi(seed.yq - 2, ymin, ymax); ++nd;
            dg[nd].xq = half_best.xq; dg[nd].yq = clampi(seed.yq + 2, ymin, ymax); ++nd;
        } else if (dy && !dx) {
            dg[nd].xq = clampi(seed.xq - 2, xmin, xmax); dg[nd].yq = half_best.yq; ++nd;
            dg[nd].xq = clampi(seed.xq + 2, xmin, xmax); dg[nd].yq = half_best.yq; ++nd;
        }
        if (hp)
            for (int i = 0; i < nd; ++i) { dg[i].xq = make_even(dg[i].xq); dg[i].yq = make_even(dg[i].yq); }
        for (int i = 0; i < nd; ++i) st_add(s, dg[i], U64MAX);
        for (int i = 0; i < nd; ++i) {
            sub_t* cd = st_find(s, dg[i]);
            if (cd && (cd->sad < half_sad || (cd->sad == half_sad && tie_less(cd->mv, half_best)))) {
                half_sad = cd->sad;
                half_best = cd->mv;
            }
        }
    }
    if (!hp) {
        if (fast) {
            const mv_t qc = half_best;
            mv_t qb5[5];
            int nq = 0;
            qb5[nq++] = qc;
            static const int o4[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
            for (int k = 0; k < 4; ++k) {
                mv_t mv = {clampi(qc.xq + o4[k][0], xmin, xmax), clampi(qc.yq + o4[k][1], ymin, ymax)};
                qb5[nq++] = mv;
            }
            for (int i = 0; i < nq; ++i) st_add(s, qb5[i], U64MAX);
            mv_t qbmv = qc;
            uint64_t qcost = st_find(s, qc)->sad;
            for (int i = 0; i < nq; ++i) {
                sub_t* cd = st_find(s, qb5[i]);
                if (cd && (cd->sad < qcost || (cd->sad == qcost && tie_less(cd->mv, qbmv)))) {
                    qbmv = cd->mv;
                    qcost = cd->sad;
                }
            }
            mv_t q2[2];
            int n2 = 0;
            const int dx = qbmv.xq - qc.xq, dy = qbmv.yq - qc.yq;
            if (dx && !dy) {
                q2[n2].xq = qbmv.xq; q2[n2].yq = clampi(qc.yq - 1, ymin, ymax); ++n2;
                q2[n2].xq = qbmv.xq; q2[n2].yq = clampi(qc.yq + 1, ymin, ymax); ++n2;
            } else if (dy && !dx) {
                q2[n2].xq = clampi(qc.xq - 1, xmin, xmax); q2[n2].yq = qbmv.yq; ++n2;
                q2[n2].xq = clampi(qc.xq + 1, xmin, xmax); q2[n2].yq = qbmv.yq; ++n2;
            }
            for (int i = 0; i < n2; ++i) st_add(s, q2[i], U64MAX);
        } else {
            for (int oy = -1; oy <= 1; ++oy)
                for (int ox = -1; ox <= 1; ++ox) {
                    mv_t mv = {clampi(half_best.xq + ox, xmin, xmax), clampi(half_best.yq + oy, ymin, ymax)};
                    st_add(s, mv, U64MAX);
                }
        }
    }
    sort_sub(s->all, s->n, sad_order);
    const uint64_t screen_sad = s->all[0].sad;
    const int keep = fast ? 4 : 8;
    if (s->n > keep) s->n = keep;
    const long double lam = lambda_me(P->pqindex);
    for (int i = 0; i < s->n; ++i) {
        sub_t* cd = &s->all[i];
        mv_t pred = pi->pre;
        cd->rate = 0;
        if (pi->hybrid) {
            pred = modular_mv_distance(c, cd->mv, pi->a) <= modular_mv_distance(c, cd->mv, pi->c) ? pi->a : pi->c;
            ++cd->rate;
        }
        cd->rate += mv_rate_bits(c, cd->mv, pred, mode, more);
        cd->score = (long double)cd->sad + lam * (long double)cd->rate;
    }
    int rw = 0;
    for (int i = 1; i < s->n; ++i)
        if (score_order(&s->all[i], &s->all[rw])) rw = i;
    const sub_t rate_winner = s->all[rw];
    mv_t result;
    if (P->me_quality == 1) {
        result = rate_winner.mv;
    } else {
        const uint64_t pct3 = (rate_winner.sad / 100) * 3 + (((rate_winner.sad % 100) * 3 + 99) / 100);
        const uint64_t slack = pct3 > 4 ? pct3 : 4;
        const uint64_t limit = rate_winner.sad > U64MAX - slack ? U64MAX : rate_winner.sad + slack;
        for (int i = 0; i < s->n; ++i) {
            sub_t* cd = &s->all[i];
            cd->eligible = cd->sad < U64MAX / 8 && cd->sad <= limit;
            cd->satd = cd->eligible ? block_satd_motion(c, x0, y0, bw, bh, cd->mv, mode) : BAD_SAD;
            cd->score = (long double)cd->satd + lam * (long double)cd->rate;
        }
        sort_sub(s->all, s->n, final_order);
        result = s->all[0].mv;
    }
    if (best_sad) *best_sad = screen_sad;
    free(s);
    return result;
}

/* ---------------------------------------------------------------- chroma MV of 4-MV */

/* median4, src/encoder_motion_compensation.cpp:522-529 */
static int median4(int a, int b, int c, int d) {
    if (a < b) {
        if (c < d) return (imin(b, d) + imax(a, c)) / 2;
        return (imin(b, c) + imax(a, d)) / 2;
    }
    if (c < d) return (imin(a, d) + imax(b, c)) / 2;
    return (imin(a, c) + imax(b, d)) / 2;
}

/* visible_block_mean, src/encoder_entropy.cpp:496-509 */
static int visible_block_mean(const uint8_t* p, int w, int h, int x0, int y0) {
    const int vw = imax(0, imin(8, w - x0)), vh = imax(0, imin(8, h - y0));
    if (!vw || !vh) return 128;
    uint64_t sum = 0;
    for (int y = 0; y < vh; ++y)
        for (int x = 0; x < vw; ++x) sum += p[(size_t)(y0 + y) * w + x0 + x];
    const uint64_t n = (uint64_t)vw * vh;
    return (int)((sum + n / 2) / n);
}

/* ---------------------------------------------------------------- frame driver */

static void init_ctx(ctx_t* c, const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref, oracle_ops* ops) {
    memset(c, 0, sizeof *c);
    c->p = p;
    c->W = p->width;
    c->H = p->height;
    c->vw = p->display_width > 0 ? p->display_width : p->width;
    c->vh = p->display_height > 0 ? p->display_height : p->height;
    c->mbw = (p->width + 15) / 16;
    c->mbh = (p->height + 15) / 16;
    c->cy = cur->y; c->cu = cur->u; c->cv = cur->v;
    c->ry = ref->y; c->ru = ref->u; c->rv = ref->v;
    c->ops = ops ? ops : &g_dummy_ops;
}

/* Vc1Encoder::analyze_p_picture_core with shared_seed == nullptr and
 * b_reuse == nullptr, src/encoder_motion.cpp:47-442 (Advanced profile). */
/* analyze_p_picture_core with shared_seed == seed (NULL: the full search).
 * A seed supplies each macroblock's 1-MV start vector (clamped to the search
 * bounds) for the staged refinement instead of the integer search, and its
 * distant-match fields are the result's (src/encoder_motion.cpp:97-101,
 * :120-130); there is no global search, so global_mad is the local MAD
 * (:432-434). */
static int analyze_core(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                        const vc1me_p_result* seed, vc1me_p_result* out, oracle_ops* ops) {
    if (p->me_quality != 1 && p->me_quality != 2) return -1;
    ctx_t cx;
    init_ctx(&cx, p, cur, ref, ops);
    const ctx_t* c = &cx;
    const int mbw = c->mbw, mbh = c->mbh;
    const size_t mbs = (size_t)mbw * mbh;
    const bool allow_4mv = p->mode == VC1ME_MODE_MIXED_MV;
    const int one_mode = allow_4mv ? VC1ME_MODE_ONE_MV_QPEL : p->mode;

    mv_t* mvs = (mv_t*)calloc(mbs, sizeof(mv_t));
    mv_t* bmv = (mv_t*)calloc(mbs * 4, sizeof(mv_t));
    uint8_t* use_4mv = (uint8_t*)calloc(mbs, 1);
    uint8_t* use_intra = (uint8_t*)calloc(mbs, 1);
    uint8_t* prop = (uint8_t*)calloc(mbs, 1);
    for (size_t i = 0; i < mbs; ++i) {
        out->screen_sads[i] = 0;
        out->inter_intra_cost_ratio[i] = -1.0;
        out->distant_local_mae[i] = -1.0;
        out->distant_candidate_mae[i] = -1.0;
        out->distant_match_decision[i] = 0;
    }
    out->moved_macroblocks = out->fractional_chroma_macroblocks = out->four_mv_macroblocks = out->intra_macroblocks = 0;
    uint64_t total_sad = 0, rate_sad_total = 0, total_pixels = 0;
    cand_t global = {0, 0, 0};
    out->global_dx = 0; out->global_dy = 0; out->global_sad = 0;

    if (p->motion_search_range == 0) {
        for (int my = 0; my < mbh; ++my)
            for (int mx = 0; mx < mbw; ++mx) {
                const uint64_t z = mb_sad(c, mx, my, 0, 0, 1);
                total_sad += z;
                out->screen_sads[(size_t)my * mbw + mx] = z;
                total_pixels += active_mb_pixels(c, mx, my);
            }
        out->mean_abs_residual = total_pixels ? (double)total_sad / (double)total_pixels : 0.0;
        out->rate_complexity = out->mean_abs_residual;
        out->motion_sad = 0;
        goto done;
    }
    {
        const int range = p->motion_search_range;
        const int local_range = imin(range, imax(0, p->motion_local_search_range));
        if (!seed) {
            global = global_search(c, local_range);
            out->global_dx = global.dx; out->global_dy = global.dy; out->global_sad = global.cost;
        } else {
            for (size_t i = 0; i < mbs; ++i) {
                out->distant_local_mae[i] = seed->distant_local_mae[i];
                out->distant_candidate_mae[i] = seed->distant_candidate_mae[i];
                out->distant_match_decision[i] = seed->distant_match_decision[i];
            }
        }
        cell_t* cells = NULL;
        int ncells = 0;
        const double lambda = fmax(1.0, 0.35 * (double)p->pqindex);
        for (int my = 0; my < mbh; ++my)
            for (int mx = 0; mx < mbw; ++mx) {
                const size_t pos = (size_t)my * mbw + mx;
                const int x0 = mx * 16, y0 = my * 16;
                const pred_t rate_pi = predictor_info_mixed_1mv(c, bmv, use_intra, mx, my);
                uint64_t one_sad = 0;
                mv_t one = {0, 0};
                if (seed) {
                    const bounds_t qb = bounds_qpel(range);
                    mv_t sv = {clampi(seed->mvs[2 * pos], qb.xmin, qb.xmax), clampi(seed->mvs[2 * pos + 1], qb.ymin, qb.ymax)};
                    one = refine_staged(c, x0, y0, 16, 16, sv, range, one_mode, &rate_pi, true, &one_sad, U64MAX);
                } else {
                    const pred_t pi = predictor_info(c, mvs, mx, my);
                    imr_t im;
                    bool selected_distant = false;
                    if (p->long_range_search_mode == 2) {
                        bool short_circuit = false;
                        if (range > local_range) {
                            const imr_t pr = propagated_long_range(c, mx, my, mvs, prop, range);
                            if (match_good(c, pr.sad, mx, my)) { im = pr; short_circuit = true; selected_distant = true; }
                        }
                        if (!short_circuit) im = integer_local(c, mx, my, &pi, global, local_range);
                    } else {
                        im = integer_local(c, mx, my, &pi, global, local_range);
                        const uint64_t local_sad = im.sad;
                        const uint64_t pixels = active_mb_pixels(c, mx, my);
                        if (pixels) out->distant_local_mae[pos] = (double)local_sad / (double)pixels;
                        const bool skip = p->long_range_search_mode == 1 && match_good(c, im.sad, mx, my);
                        if (range > local_range && !skip) {
                            const imr_t pr = propagated_long_range(c, mx, my, mvs, prop, range);
                            const int decision = distant_decision(c, pr.sad, local_sad, mx, my);
                            if (pixels && pr.sad < U64MAX / 8) {
                                const double mae = (double)pr.sad / (double)pixels;
                                if (out->distant_candidate_mae[pos] < 0.0 || mae < out->distant_candidate_mae[pos]) {
                                    out->distant_candidate_mae[pos] = mae;
                                    out->distant_match_decision[pos] = (uint8_t)decision;
                                }
                            }
                            if (better4(pr.sad, pr.dx, pr.dy, im.sad, im.dx, im.dy) && decision >= 2) {
                                im = pr;
                                selected_distant = true;
                            }
                        }
                    }
                    if (range > local_range && !match_good(c, im.sad, mx, my)) {
                        if (!cells) {
                            const int nx = (c->W + 7) / 8, ny = (c->H + 7) / 8;
                            cells = (cell_t*)malloc(sizeof(cell_t) * (size_t)nx * ny);
                            for (int y = 0; y < c->H; y += 8)
                                for (int x = 0; x < c->W; x += 8) {
                                    cells[ncells].x = x;
                                    cells[ncells].y = y;
                                    cells[ncells].s = signature(c->ry, c->ru, c->rv, c->W, c->H, x, y);
                                    ++ncells;
                                }
                        }
                        uint64_t ext_sad = U64MAX;
                        int ext_dec = 0;
                        const imr_t ext = extended_content(c, mx, my, range, cells, ncells, im, &ext_sad, &ext_dec);
                        const uint64_t pixels = active_mb_pixels(c, mx, my);
                        if (p->long_range_search_mode != 2 && pixels && ext_sad < U64MAX / 8) {
                            const double mae = (double)ext_sad / (double)pixels;
                            if (out->distant_candidate_mae[pos] < 0.0 || mae < out->distant_candidate_mae[pos]) {
                                out->distant_candidate_mae[pos] = mae;
                                out->distant_match_decision[pos] = (uint8_t)ext_dec;
                            }
                        }
                        if (ext.dx != im.dx || ext.dy != im.dy) {
                            im = ext;
                            selected_distant = true;
                        }
                    }
                    if (selected_distant) prop[pos] = 1;
                    mv_t seed = {im.dx * 4, im.dy * 4};
                    one = refine_staged(c, x0, y0, 16, 16, seed, range, one_mode, &rate_pi, true, &one_sad, im.sad);
                }
                if (hidden_padding(c, x0, y0, 16, 16, one, one_mode, true)) {
                    one.xq = one.yq = 0;
                    one_sad = block_sad_motion(c, x0, y0, 16, 16, one, one_mode);
                }
                const uint64_t one_screen_sad = one_sad;
                mvs[pos] = one;
                for (int k = 0; k < 4; ++k) bmv[pos * 4 + k] = one;
                uint64_t chosen_sad = one_sad;
                uint64_t temporal_screen_sad = one_screen_sad;
                if (allow_4mv) {
                    mv_t four[4] = {one, one, one, one};
                    uint64_t four_sad = 0, four_screen_sad = 0, extra_bits = 1;
                    const bounds_t qb = bounds_qpel(range);
                    const int limit = 4 * p->subpixel_refinement_range;
                    for (int k = 0; k < 4; ++k) {
                        const pred_t pi4 = predictor_info_4mv(c, bmv, use_intra, mx, my, k);
                        const int bx = x0 + (k & 1) * 8, by = y0 + ((k >> 1) & 1) * 8;
#define WITHIN(m_)                                                      \
    do {                                                                \
        if (imax(iabs((m_).xq), iabs((m_).yq)) > limit) {               \
            (m_).xq = ((m_).xq / 4) * 4;                                \
            (m_).yq = ((m_).yq / 4) * 4;                                \
        }                                                               \
    } while (0)
                        mv_t best_seed = one;
                        WITHIN(best_seed);
                        uint64_t seed_cost = block_sad_motion(c, bx, by, 8, 8, best_seed, VC1ME_MODE_ONE_MV_QPEL);
                        const mv_t seeds[10] = {one, pi4.pre,
                                                {one.xq - 8, one.yq}, {one.xq + 8, one.yq},
                                                {one.xq, one.yq - 8}, {one.xq, one.yq + 8},
                                                {one.xq - 4, one.yq}, {one.xq + 4, one.yq},
                                                {one.xq, one.yq - 4}, {one.xq, one.yq + 4}};
                        for (int si = 0; si < 10; ++si) {
                            mv_t sd = seeds[si];
                            sd.xq = clampi(sd.xq, qb.xmin, qb.xmax);
                            sd.yq = clampi(sd.yq, qb.ymin, qb.ymax);
                            WITHIN(sd);
                            if (same_mv(sd, best_seed)) continue;
                            const uint64_t cc = block_sad_motion(c, bx, by, 8, 8, sd, VC1ME_MODE_ONE_MV_QPEL);
                            const int cm = iabs(sd.xq) + iabs(sd.yq), bm = iabs(best_seed.xq) + iabs(best_seed.yq);
                            if (cc < seed_cost || (cc == seed_cost && cm < bm)) { seed_cost = cc; best_seed = sd; }
                        }
#undef WITHIN
                        uint64_t bsad = 0;
                        const mv_t mv = refine_staged(c, bx, by, 8, 8, best_seed, range, VC1ME_MODE_MIXED_MV, &pi4, true,
                                                      &bsad, seed_cost);
                        four_screen_sad += bsad;
                        four[k] = mv;
                        bmv[pos * 4 + k] = mv;
                        four_sad += bsad;
                        mv_t actual = pi4.pre;
                        if (pi4.hybrid) {
                            const int da = modular_mv_distance(c, mv, pi4.a), dc = modular_mv_distance(c, mv, pi4.c);
                            actual = (da <= dc) ? pi4.a : pi4.c;
                            ++extra_bits;
                        }
                        if (!same_mv(mv, actual)) {
                            mvdata_t est;
                            est.dx = mv.xq - actual.xq; est.dy = mv.yq - actual.yq;
                            est.more = false; est.intra = false; est.halfpel = false;
                            extra_bits += kMvBits0[mv_symbol(&est)] + (uint64_t)mv_suffix_bits(c, &est);
                        }
                    }
                    const mv_t cmv = {median4(four[0].xq, four[1].xq, four[2].xq, four[3].xq),
                                      median4(four[0].yq, four[1].yq, four[2].yq, four[3].yq)};
                    const bool chroma_safe = !hidden_chroma(c, x0, y0, 16, 16, cmv, false);
                    bool choose4 = false;
                    if (chroma_safe) {
                        const double lb = lambda * (double)extra_bits;
                        const double os = (double)one_sad * 0.985;
                        choose4 = four_sad + (uint64_t)llround(lb) + 8 < (uint64_t)llround(os);
                    }
                    if (choose4) {
                        for (int k = 0; k < 4; ++k) bmv[pos * 4 + k] = four[k];
                        use_4mv[pos] = 1;
                        mvs[pos] = four[0];
                        out->four_mv_macroblocks++;
                        chosen_sad = four_sad;
                        temporal_screen_sad = four_screen_sad;
                    } else {
                        for (int k = 0; k < 4; ++k) bmv[pos * 4 + k] = one;
                        mvs[pos] = one;
                        temporal_screen_sad = one_screen_sad;
                    }
                }
                const uint64_t temporal_sad = chosen_sad;
                out->screen_sads[pos] = temporal_screen_sad;
                if (p->allow_intra && !p->debug_disable_p_intra) {
                    uint64_t intra_sad = 0;
                    for (int k = 0; k < 4; ++k) {
                        const int bx = x0 + (k & 1) * 8, by = y0 + ((k >> 1) & 1) * 8;
                        const int mean = visible_block_mean(c->cy, c->W, c->H, bx, by);
                        for (int yy = 0; yy < 8 && by + yy < c->H; ++yy)
                            for (int xx = 0; xx < 8 && bx + xx < c->W; ++xx)
                                intra_sad += (uint64_t)iabs((int)c->cy[(size_t)(by + yy) * c->W + bx + xx] - mean);
                    }
                    const uint64_t intra_rate = (uint64_t)llround(lambda * 72.0);
                    const double ratio = temporal_screen_sad > 0
                                             ? (double)(intra_sad + intra_rate) / (double)temporal_screen_sad
                                             : INFINITY;
                    out->inter_intra_cost_ratio[pos] = ratio;
                    const double thr = (double)temporal_screen_sad * p->inter_intra_threshold;
                    if (intra_sad + intra_rate < (uint64_t)llround(thr)) {
                        use_intra[pos] = 1;
                        use_4mv[pos] = 0;
                        mvs[pos].xq = mvs[pos].yq = 0;
                        for (int k = 0; k < 4; ++k) bmv[pos * 4 + k].xq = bmv[pos * 4 + k].yq = 0;
                        out->intra_macroblocks++;
                        chosen_sad = intra_sad;
                    }
                }
                bool moved = false, frac = false;
                for (int k = 0; k < 4; ++k) {
                    const mv_t m = bmv[pos * 4 + k];
                    moved |= m.xq != 0 || m.yq != 0;
                    frac |= (m.xq % 8) != 0 || (m.yq % 8) != 0;
                }
                if (moved) out->moved_macroblocks++;
                if (frac) out->fractional_chroma_macroblocks++;
                total_sad += chosen_sad;
                rate_sad_total += temporal_sad;
                total_pixels += active_mb_pixels(c, mx, my);
            }
        free(cells);
        const double local_mad = total_pixels ? (double)total_sad / (double)total_pixels : 0.0;
        const double global_mad = seed ? local_mad
                                       : (c->W > 0 && c->H > 0) ? (double)global.cost / (double)((uint64_t)c->W * c->H) : 0.0;
        out->mean_abs_residual = fmin(local_mad, global_mad);
        const double rate_local = total_pixels ? (double)rate_sad_total / (double)total_pixels : 0.0;
        out->rate_complexity = fmin(rate_local, global_mad);
        out->motion_sad = total_sad;
    }
done:
    for (size_t i = 0; i < mbs; ++i) {
        out->mvs[2 * i] = mvs[i].xq;
        out->mvs[2 * i + 1] = mvs[i].yq;
        for (int k = 0; k < 4; ++k) {
            out->block_mvs[8 * i + 2 * k] = bmv[i * 4 + k].xq;
            out->block_mvs[8 * i + 2 * k + 1] = bmv[i * 4 + k].yq;
        }
        out->use_4mv[i] = use_4mv[i];
        out->use_intra[i] = use_intra[i];
    }
    free(mvs); free(bmv); free(use_4mv); free(use_intra); free(prop);
    return 0;
}

int oracle_analyze_p_frame(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                           vc1me_p_result* out, oracle_ops* ops) {
    return analyze_core(p, cur, ref, NULL, out, ops);
}

int oracle_analyze_p_frame_seeded(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                                  const vc1me_p_result* seed, vc1me_p_result* out, oracle_ops* ops) {
    if (!seed || !seed->mvs || !seed->distant_local_mae || !seed->distant_candidate_mae || !seed->distant_match_decision)
        return -1;
    return analyze_core(p, cur, ref, seed, out, ops);
}

/* ---------------------------------------------------------------- batch scorer */

/* GLSL predict_luma / sad_cost / satd_cost,
 * src/shaders/vulkan_motion_cost.comp:58-176 */
void oracle_score_batch(const uint8_t* cur, const uint8_t* ref, int width, int height, int x0, int y0, int bw,
                        int bh, int sample_stride, int bilinear, int rnd, int metric, const vc1me_candidate* cd,
                        size_t n, uint64_t* costs) {
    const int mode = bilinear ? VC1ME_MODE_ONE_MV_HPEL_BILINEAR : VC1ME_MODE_ONE_MV_QPEL;
    const int aw = imax(0, imin(bw, width - x0)), ah = imax(0, imin(bh, height - y0));
    const int ss = imax(sample_stride, 1);
    for (size_t i = 0; i < n; ++i) {
        const int xq = cd[i].xq, yq = cd[i].yq;
        uint64_t total = 0;
        if (metric == VC1ME_METRIC_SAD) {
            for (int y = 0; y < ah; y += ss)
                for (int x = 0; x < aw; x += ss) {
                    const int cv = cur[(size_t)(y0 + y) * width + x0 + x];
                    const int pv = luma_mc_sample_mode(ref, width, height, x0 + x, y0 + y, xq, yq, mode, rnd);
                    total += (uint64_t)iabs(cv - pv);
                }
            if (ss > 1) total *= (uint64_t)(ss * ss);
            total &= 0xffffffffu; /* the kernel accumulates in a 32-bit uint */
        } else {
            for (int by = 0; by < bh; by += 4)
                for (int bx = 0; bx < bw; bx += 4) {
                    if (!(bx < aw && by < ah)) continue;
                    uint8_t ct[16], pt[16];
                    for (int yy = 0; yy < 4; ++yy)
                        for (int xx = 0; xx < 4; ++xx) {
                            const int px = bx + xx, py = by + yy;
                            if (px < aw && py < ah) {
                                ct[yy * 4 + xx] = cur[(size_t)(y0 + py) * width + x0 + px];
                                pt[yy * 4 + xx] = (uint8_t)luma_mc_sample_mode(ref, width, height, x0 + px, y0 + py, xq,
                                                                               yq, mode, rnd);
                            } else {
                                ct[yy * 4 + xx] = 0;
                                pt[yy * 4 + xx] = 0;
                            }
                        }
                    total += satd4x4(ct, 4, pt, 4);
                }
        }
        costs[i] = total;
    }
}

/* ---------------------------------------------------------------- test hooks */

int oracle_luma_mc_sample(const uint8_t* plane, int w, int h, int x, int y, int xq, int yq, int mode, int rnd) {
    return luma_mc_sample_mode(plane, w, h, x, y, xq, yq, mode, rnd);
}
uint64_t oracle_mb_sad(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref, int mx, int my,
                       int dx, int dy, int stride) {
    ctx_t c;
    init_ctx(&c, p, cur, ref, NULL);
    return mb_sad(&c, mx, my, dx, dy, stride);
}
uint64_t oracle_block_cost(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref, int x0, int y0,
                           int bw, int bh, int xq, int yq, int mode, int metric) {
    ctx_t c;
    init_ctx(&c, p, cur, ref, NULL);
    const mv_t mv = {xq, yq};
    return metric ? block_satd_motion(&c, x0, y0, bw, bh, mv, mode) : block_sad_motion(&c, x0, y0, bw, bh, mv, mode);
}
uint64_t oracle_frame_sad(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref, int dx, int dy,
                          int stride) {
    ctx_t c;
    init_ctx(&c, p, cur, ref, NULL);
    return frame_sad(&c, dx, dy, stride);
}
void oracle_global_motion(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref, int range,
                          int32_t* dx, int32_t* dy, uint64_t* sad) {
    ctx_t c;
    init_ctx(&c, p, cur, ref, NULL);
    const cand_t g = global_search(&c, range);
    *dx = g.dx; *dy = g.dy; *sad = g.cost;
}
int oracle_mv_rate_bits(const vc1me_params* p, int dxq, int dyq, int pxq, int pyq, int mode, int more) {
    ctx_t c;
    memset(&c, 0, sizeof c);
    c.p = p;
    const mv_t d = {dxq, dyq}, pr = {pxq, pyq};
    return (int)mv_rate_bits(&c, d, pr, mode, more != 0);
}
int oracle_score_cmp(int pqindex, uint64_t sad_a, uint64_t rate_a, uint64_t sad_b, uint64_t rate_b) {
    const long double lam = lambda_me(pqindex);
    const long double a = (long double)sad_a + lam * (long double)rate_a;
    const long double b = (long double)sad_b + lam * (long double)rate_b;
    return a < b ? -1 : (a > b ? 1 : 0);
}

int oracle_hidden_padding(const vc1me_params* p, int x0, int y0, int bw, int bh, int xq, int yq, int mode, int chroma) {
    ctx_t c;
    memset(&c, 0, sizeof c);
    c.p = p;
    c.W = p->width; c.H = p->height;
    c.vw = p->display_width > 0 ? p->display_width : p->width;
    c.vh = p->display_height > 0 ? p->display_height : p->height;
    c.mbw = (p->width + 15) / 16; c.mbh = (p->height + 15) / 16;
    const mv_t mv = {xq, yq};
    return hidden_padding(&c, x0, y0, bw, bh, mv, mode, chroma != 0) ? 1 : 0;
}
int oracle_distant_decision(const vc1me_params* p, uint64_t cand, uint64_t local, int mx, int my) {
    ctx_t c;
    memset(&c, 0, sizeof c);
    c.p = p;
    c.W = p->width; c.H = p->height;
    return distant_decision(&c, cand, local, mx, my);
}
void oracle_integer_search(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref, int mx, int my,
                           int pre_xq, int pre_yq, int gdx, int gdy, int range, int32_t* dx, int32_t* dy,
                           uint64_t* sad) {
    ctx_t c;
    init_ctx(&c, p, cur, ref, NULL);
    pred_t pi;
    memset(&pi, 0, sizeof pi);
    pi.pre.xq = pre_xq; pi.pre.yq = pre_yq;
    const cand_t g = {0, gdx, gdy};
    const imr_t r = integer_local(&c, mx, my, &pi, g, range);
    *dx = r.dx; *dy = r.dy; *sad = r.sad;
}

/* Per-macroblock judge of a finished analysis: 20 * (SAD + lambda * R) of the
 * final 1-MV vector one_mvs[pos] -- block_sad_motion at that vector, R its
 * rate against predictor_info_mixed_1mv of the final block field and intra
 * flags, HYBRIDPRED bit included, lambda = max(1, 0.35 * PQINDEX).  This is
 * the rate-winner score of src/encoder_motion.cpp:3458-3471 (the same terms
 * refine_staged above forms per candidate) evaluated once per macroblock;
 * 20 * lambda = max(20, 7 * PQINDEX) keeps it an integer. */
void oracle_mb_costs20(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                       const int32_t* one_mvs, const int32_t* block_mvs, const uint8_t* use_intra, uint64_t* out) {
    ctx_t c;
    init_ctx(&c, p, cur, ref, NULL);
    const int one_mode = p->mode == VC1ME_MODE_MIXED_MV ? VC1ME_MODE_ONE_MV_QPEL : p->mode;
    const uint64_t lam20 = (uint64_t)imax(20, 7 * p->pqindex);
    const mv_t* field = (const mv_t*)block_mvs;
    for (int my = 0; my < c.mbh; ++my)
        for (int mx = 0; mx < c.mbw; ++mx) {
            const size_t pos = (size_t)my * c.mbw + mx;
            const mv_t one = {one_mvs[2 * pos], one_mvs[2 * pos + 1]};
            const pred_t pi = predictor_info_mixed_1mv(&c, field, use_intra, mx, my);
            mv_t pred = pi.pre;
            uint64_t rate = 0;
            if (pi.hybrid) {
                pred = modular_mv_distance(&c, one, pi.a) <= modular_mv_distance(&c, one, pi.c) ? pi.a : pi.c;
                ++rate;
            }
            rate += mv_rate_bits(&c, one, pred, one_mode, true);
            out[pos] = 20u * block_sad_motion(&c, mx * 16, my * 16, 16, 16, one, one_mode) + lam20 * rate;
        }
}

/* ------------------------------------------------------------------ intensity compensation */

/* Vc1Encoder::intensity_map_y (src/encoder_motion_compensation.cpp:175-187):
 * LUMSCALE 0 is the inverting map, otherwise scale = LUMSCALE + 32 over 64
 * and LUMSHIFT a 6-bit two's-complement offset. */
static void intensity_scale_shift(const vc1me_intensity* ic, int* scale, int* shift) {
    if (ic->lumscale == 0) {
        *scale = -64;
        *shift = (255 - ic->lumshift * 2) * 64;
        if (ic->lumshift > 31) *shift += 128 * 64;
    } else {
        *scale = ic->lumscale + 32;
        *shift = (ic->lumshift > 31 ? ic->lumshift - 64 : ic->lumshift) * 64;
    }
}

uint8_t oracle_intensity_map_y(uint8_t v, const vc1me_intensity* ic) {
    if (!ic->enabled) return v;
    int scale, shift;
    intensity_scale_shift(ic, &scale, &shift);
    return (uint8_t)clampi((scale * (int)v + shift + 32) >> 6, 0, 255);
}

/* Vc1Encoder::intensity_map_uv (src/encoder_motion_compensation.cpp:189-193) */
uint8_t oracle_intensity_map_uv(uint8_t v, const vc1me_intensity* ic) {
    if (!ic->enabled) return v;
    const int scale = ic->lumscale == 0 ? -64 : ic->lumscale + 32;
    return (uint8_t)clampi((scale * ((int)v - 128) + 128 * 64 + 32) >> 6, 0, 255);
}

/* Vc1Encoder::intensity_compensated_reference
 * (src/encoder_motion_compensation.cpp:195-294): every sample through the
 * map of its plane. */
void oracle_intensity_compensate(const vc1me_planes* ref, const vc1me_intensity* ic, uint8_t* y, uint8_t* u,
                                 uint8_t* v) {
    const size_t ys = (size_t)ref->width * ref->height, cs = (size_t)(ref->width / 2) * (ref->height / 2);
    for (size_t i = 0; i < ys; ++i) y[i] = oracle_intensity_map_y(ref->y[i], ic);
    for (size_t i = 0; i < cs; ++i) {
        u[i] = oracle_intensity_map_uv(ref->u ? ref->u[i] : 128, ic);
        v[i] = oracle_intensity_map_uv(ref->v ? ref->v[i] : 128, ic);
    }
}

/* Vc1Encoder::estimate_intensity_comp (src/encoder_motion_compensation.cpp:296-348).
 * Samples: every 4th pixel in x and y of each macroblock, the reference
 * sample displaced by the macroblock's vector rounded to whole pixels
 * (round_qpel_to_pixel, src/encoder_motion.cpp:1295-1297) and clamped to
 * the raster.  The least-squares fit cur ~ a * ref + b runs in long double
 * on exact integer sums; LUMSCALE = clamp(round(64a - 32), 1, 63), LUMSHIFT =
 * clamp(round(b), -32, 31); then the (scale, shift) of smallest sampled SAD
 * within +-3 of both, ties to the smaller |scale - 32| + |shift| and then to
 * the first in (scale, shift) order.  (32, 0) means no compensation. */
int oracle_estimate_intensity_comp(const vc1me_planes* cur, const vc1me_planes* ref, const int32_t* mvs,
                                   vc1me_intensity* out) {
    const int W = cur->width, H = cur->height;
    const int mbw = (W + 15) / 16, mbh = (H + 15) / 16;
    out->enabled = 0; out->lumscale = 32; out->lumshift = 0;
    const size_t cap = (size_t)((W + 3) / 4) * ((H + 3) / 4);
    uint8_t* sx8 = (uint8_t*)malloc(cap);
    uint8_t* sy8 = (uint8_t*)malloc(cap);
    uint64_t sx = 0, sy = 0, sxx = 0, sxy = 0, n = 0;
    for (int my = 0; my < mbh; ++my)
        for (int mx = 0; mx < mbw; ++mx) {
            const int32_t xq = mvs[2 * ((size_t)my * mbw + mx)], yq = mvs[2 * ((size_t)my * mbw + mx) + 1];
            const int dx = xq >= 0 ? (xq + 2) / 4 : -((-xq + 2) / 4);
            const int dy = yq >= 0 ? (yq + 2) / 4 : -((-yq + 2) / 4);
            for (int y = 0; y < 16 && my * 16 + y < H; y += 4)
                for (int x = 0; x < 16 && mx * 16 + x < W; x += 4) {
                    const int px = mx * 16 + x, py = my * 16 + y;
                    const uint8_t rv = ref->y[(size_t)clampi(py + dy, 0, H - 1) * W + clampi(px + dx, 0, W - 1)];
                    const uint8_t cv = cur->y[(size_t)py * W + px];
                    sx8[n] = rv; sy8[n] = cv;
                    sx += rv; sy += cv; sxx += (uint64_t)rv * rv; sxy += (uint64_t)rv * cv; ++n;
                }
        }
    if (n < 16) { free(sx8); free(sy8); return 0; }
    const long double ln = (long double)n, lsx = (long double)sx, lsy = (long double)sy;
    const long double denom = ln * (long double)sxx - lsx * lsx;
    long double a = 1.0L, b = 0.0L;
    if (fabsl(denom) > 1e-9L) {
        a = (ln * (long double)sxy - lsx * lsy) / denom;
        b = (lsy - a * lsx) / ln;
    } else {
        b = (lsy - lsx) / ln;
    }
    const int ls = clampi((int)llroundl(a * 64.0L - 32.0L), 1, 63);
    const int sh = clampi((int)llroundl(b), -32, 31);
    uint64_t best = UINT64_MAX;
    int best_ls = 32, best_sh = 0;
    for (int l = imax(1, ls - 3); l <= imin(63, ls + 3); ++l)
        for (int z = imax(-32, sh - 3); z <= imin(31, sh + 3); ++z) {
            const vc1me_intensity ic = {1, l, z < 0 ? z + 64 : z};
            uint64_t sad = 0;
            for (uint64_t i = 0; i < n; ++i) sad += (uint64_t)iabs((int)sy8[i] - (int)oracle_intensity_map_y(sx8[i], &ic));
            if (sad < best || (sad == best && iabs(l - 32) + iabs(z) < iabs(best_ls - 32) + iabs(best_sh))) {
                best = sad; best_ls = l; best_sh = z;
            }
        }
    free(sx8);
    free(sy8);
    if (best_ls == 32 && best_sh == 0) return 0;
    out->enabled = 1;
    out->lumscale = best_ls;
    out->lumshift = best_sh < 0 ? best_sh + 64 : best_sh;
    return 0;
}

/* ------------------------------------------------------------------ B pictures */

static void p_result_alloc(vc1me_p_result* r, size_t mbs) {
    memset(r, 0, sizeof *r);
    r->mvs = (int32_t*)calloc(mbs * 2, sizeof(int32_t));
    r->block_mvs = (int32_t*)calloc(mbs * 8, sizeof(int32_t));
    r->use_4mv = (uint8_t*)calloc(mbs, 1);
    r->use_intra = (uint8_t*)calloc(mbs, 1);
    r->screen_sads = (uint64_t*)calloc(mbs, sizeof(uint64_t));
    r->inter_intra_cost_ratio = (double*)calloc(mbs, sizeof(double));
    r->distant_local_mae = (double*)calloc(mbs, sizeof(double));
    r->distant_candidate_mae = (double*)calloc(mbs, sizeof(double));
    r->distant_match_decision = (uint8_t*)calloc(mbs, 1);
}

static void p_result_free(vc1me_p_result* r) {
    free(r->mvs); free(r->block_mvs); free(r->use_4mv); free(r->use_intra); free(r->screen_sads);
    free(r->inter_intra_cost_ratio); free(r->distant_local_mae); free(r->distant_candidate_mae);
    free(r->distant_match_decision);
}

/* The direct-mode vector of a macroblock: the future anchor's collocated
 * vector scaled by the temporal fraction (bfrac / 256, or bfrac - 256 for
 * the backward half), rounded on the half-pel grid in half-pel modes, then
 * clamped to the picture's MV bounds (src/encoder_motion.cpp:815-833). */
static mv_t b_scale_direct(const ctx_t* c, mv_t col, bool inverse, int bfrac, int mode, int mx, int my) {
    const int n = inverse ? bfrac - 256 : bfrac;
    mv_t mv;
    if (mode_halfpel(mode)) {
        mv.xq = 2 * arshift(col.xq * n + 255, 9);
        mv.yq = 2 * arshift(col.yq * n + 255, 9);
    } else {
        mv.xq = arshift(col.xq * n + 128, 8);
        mv.yq = arshift(col.yq * n + 128, 8);
    }
    mv.xq = clampi(mv.xq, -60 - mx * 64, c->mbw * 64 - 4 - mx * 64);
    mv.yq = clampi(mv.yq, -60 - my * 64, c->mbh * 64 - 4 - my * 64);
    return mv;
}

/* Vc1Encoder::analyze_b_picture_mode (src/encoder_motion.cpp:753-1004,
 * Advanced profile): the forward and backward core passes (no intra), then
 * per macroblock in raster order the four prediction SADs, the mode choice
 * by (long double)SAD + (double)(lambda * bits) with lambda = max(1, 0.45 *
 * PQINDEX) against the predictors of the final B fields, and the B-intra
 * decision with its 20% hysteresis. */
int oracle_analyze_b_frame(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* past,
                           const vc1me_planes* future, const vc1me_b_params* b, vc1me_b_result* out) {
    const int fn = b->fraction_num, fd = b->fraction_den;
    if (!((fd == 2 && fn == 1) || (fd == 3 && (fn == 1 || fn == 2)))) return -1;
    if (p->mode == VC1ME_MODE_MIXED_MV || p->me_quality < 1 || p->me_quality > 2) return -1;
    ctx_t cx;
    init_ctx(&cx, p, cur, past, NULL);
    const ctx_t* c = &cx;
    const int mbw = c->mbw, mbh = c->mbh, mode = p->mode;
    const size_t mbs = (size_t)mbw * mbh;
    mv_t* fout = (mv_t*)calloc(mbs, sizeof(mv_t));
    mv_t* bout = (mv_t*)calloc(mbs, sizeof(mv_t));
    out->forward_macroblocks = out->backward_macroblocks = out->interpolated_macroblocks = 0;
    out->direct_macroblocks = out->intra_macroblocks = out->moved_macroblocks = 0;
    out->fractional_chroma_macroblocks = 0;
    out->mean_abs_residual = out->rate_complexity = 0.0;
    out->motion_sad = out->motion_bits = 0;
    if (b->zero_motion_shortcut && p->motion_search_range == 0) {
        for (size_t i = 0; i < mbs; ++i) {
            if (out->modes) out->modes[i] = VC1ME_B_FORWARD;
            if (out->forward_mvs) { out->forward_mvs[2 * i] = 0; out->forward_mvs[2 * i + 1] = 0; }
            if (out->backward_mvs) { out->backward_mvs[2 * i] = 0; out->backward_mvs[2 * i + 1] = 0; }
            if (out->inter_intra_cost_ratio) out->inter_intra_cost_ratio[i] = -1.0;
            if (out->forward_distant_local_mae) out->forward_distant_local_mae[i] = -1.0;
            if (out->forward_distant_candidate_mae) out->forward_distant_candidate_mae[i] = -1.0;
            if (out->forward_distant_match_decision) out->forward_distant_match_decision[i] = 0;
            if (out->backward_distant_local_mae) out->backward_distant_local_mae[i] = -1.0;
            if (out->backward_distant_candidate_mae) out->backward_distant_candidate_mae[i] = -1.0;
            if (out->backward_distant_match_decision) out->backward_distant_match_decision[i] = 0;
        }
        out->forward_macroblocks = mbs;
        free(fout); free(bout);
        return 0;
    }
    /* the forward reference carries the anchor's intensity compensation */
    vc1me_planes pw = *past;
    uint8_t *wy = NULL, *wu = NULL, *wv = NULL;
    if (b->forward_intensity.enabled) {
        const size_t ys = (size_t)past->width * past->height, cs = (size_t)(past->width / 2) * (past->height / 2);
        wy = (uint8_t*)malloc(ys); wu = (uint8_t*)malloc(cs); wv = (uint8_t*)malloc(cs);
        oracle_intensity_compensate(past, &b->forward_intensity, wy, wu, wv);
        pw.y = wy; pw.u = wu; pw.v = wv;
    }
    vc1me_params q = *p;
    q.allow_intra = 0;
    vc1me_p_result fwd, bwd;
    p_result_alloc(&fwd, mbs);
    p_result_alloc(&bwd, mbs);
    int rc = analyze_core(&q, cur, &pw, NULL, &fwd, NULL);
    if (!rc) rc = analyze_core(&q, cur, future, NULL, &bwd, NULL);
    if (rc) goto cleanup;
    for (size_t i = 0; i < mbs; ++i) {
        if (out->forward_distant_local_mae) out->forward_distant_local_mae[i] = fwd.distant_local_mae[i];
        if (out->forward_distant_candidate_mae) out->forward_distant_candidate_mae[i] = fwd.distant_candidate_mae[i];
        if (out->forward_distant_match_decision) out->forward_distant_match_decision[i] = fwd.distant_match_decision[i];
        if (out->backward_distant_local_mae) out->backward_distant_local_mae[i] = bwd.distant_local_mae[i];
        if (out->backward_distant_candidate_mae) out->backward_distant_candidate_mae[i] = bwd.distant_candidate_mae[i];
        if (out->backward_distant_match_decision) out->backward_distant_match_decision[i] = bwd.distant_match_decision[i];
    }
    {
        const int bfrac = fd == 2 ? 128 : (fn == 1 ? 85 : 170);
        const bool below_half = 2 * fn < fd;
        const double lambda = fmax(1.0, 0.45 * (double)p->pqindex);
        const uint64_t intra_guard = (uint64_t)llround(lambda * 72.0);
        const uint64_t intra_marker_bits = (uint64_t)imax(kMvBits0[35], kMvBits0[72]) + 1u;
        const uint64_t invalid = (uint64_t)1 << 54;
        uint64_t sel_sad = 0, rate_sad = 0, sel_rate = 1, sel_px = 0;
        const int W = c->W, H = c->H;
        for (int my = 0; my < mbh; ++my)
            for (int mx = 0; mx < mbw; ++mx) {
                const size_t pos = (size_t)my * mbw + mx;
                const mv_t col = {b->future_anchor_mvs[2 * pos], b->future_anchor_mvs[2 * pos + 1]};
                const mv_t df = b_scale_direct(c, col, false, bfrac, mode, mx, my);
                const mv_t db = b_scale_direct(c, col, true, bfrac, mode, mx, my);
                const mv_t fm = {fwd.mvs[2 * pos], fwd.mvs[2 * pos + 1]};
                const mv_t bm = {bwd.mvs[2 * pos], bwd.mvs[2 * pos + 1]};
                const int x0 = mx * 16, y0 = my * 16;
                const int vw = imin(16, W - x0), vh = imin(16, H - y0);
                const bool safe_fm = !(hidden_luma(c, x0, y0, 16, 16, fm, mode) || hidden_chroma(c, x0, y0, 16, 16, fm, false));
                const bool safe_bm = !(hidden_luma(c, x0, y0, 16, 16, bm, mode) || hidden_chroma(c, x0, y0, 16, 16, bm, true));
                const bool safe_df = !(hidden_luma(c, x0, y0, 16, 16, df, mode) || hidden_chroma(c, x0, y0, 16, 16, df, false));
                const bool safe_db = !(hidden_luma(c, x0, y0, 16, 16, db, mode) || hidden_chroma(c, x0, y0, 16, 16, db, true));
                uint64_t cf = 0, cb = 0, ci = 0, cdir = 0;
                for (int y = 0; y < vh; ++y)
                    for (int x = 0; x < vw; ++x) {
                        const int src = cur->y[(size_t)(y0 + y) * W + x0 + x];
                        const int pf = luma_mc_sample_mode(pw.y, W, H, x0 + x, y0 + y, fm.xq, fm.yq, mode, 0);
                        const int pb = luma_mc_sample_mode(future->y, W, H, x0 + x, y0 + y, bm.xq, bm.yq, mode, 0);
                        const int pdf = luma_mc_sample_mode(pw.y, W, H, x0 + x, y0 + y, df.xq, df.yq, mode, 0);
                        const int pdb = luma_mc_sample_mode(future->y, W, H, x0 + x, y0 + y, db.xq, db.yq, mode, 0);
                        cf += (uint64_t)iabs(src - pf);
                        cb += (uint64_t)iabs(src - pb);
                        ci += (uint64_t)iabs(src - ((pf + pb + 1) >> 1));
                        cdir += (uint64_t)iabs(src - ((pdf + pdb + 1) >> 1));
                    }
                if (!safe_fm) cf = invalid;
                if (!safe_bm) cb = invalid;
                if (!safe_fm || !safe_bm) ci = invalid;
                if (!safe_df || !safe_db) cdir = invalid;
                const mv_t fpred = predictor_info(c, fout, mx, my).pre;
                const mv_t bpred = predictor_info(c, bout, mx, my).pre;
                const uint64_t fbits = mv_rate_bits(c, fm, fpred, mode, false);
                const uint64_t bbits = mv_rate_bits(c, bm, bpred, mode, false);
                const uint64_t ibits = mv_rate_bits(c, bm, bpred, mode, true) + mv_rate_bits(c, fm, fpred, mode, false) + 2;
                const uint64_t ftype = below_half ? 1u : 2u, btype = below_half ? 2u : 1u;
                const uint64_t pen = b->future_anchor_4mv[pos] ? 1000000000ull : 0ull;
                const int cmode[4] = {VC1ME_B_DIRECT, VC1ME_B_FORWARD, VC1ME_B_BACKWARD, VC1ME_B_INTERPOLATED};
                const uint64_t ccost[4] = {cdir + pen, cf + pen, cb + pen, ci};
                const uint64_t crate[4] = {0, fbits + ftype, bbits + btype, ibits};
                int best = 0;
                long double best_score = (long double)ccost[0] + lambda * (double)crate[0];
                for (int k = 0; k < 4; ++k) {
                    const long double s = (long double)ccost[k] + lambda * (double)crate[k];
                    if (s < best_score) { best = k; best_score = s; }
                }
                uint64_t intra_sad = 0;
                for (int k = 0; k < 4; ++k) {
                    const int bx = x0 + (k & 1) * 8, by = y0 + ((k >> 1) & 1) * 8;
                    const int mean = visible_block_mean(cur->y, W, H, bx, by);
                    for (int yy = 0; yy < 8 && by + yy < H; ++yy)
                        for (int xx = 0; xx < 8 && bx + xx < W; ++xx)
                            intra_sad += (uint64_t)iabs((int)cur->y[(size_t)(by + yy) * W + bx + xx] - mean);
                }
                uint64_t gate = ccost[best];
                if (fwd.screen_sads[pos] < gate) gate = fwd.screen_sads[pos];
                if (bwd.screen_sads[pos] < gate) gate = bwd.screen_sads[pos];
                const double ratio = gate > 0 ? (double)(intra_sad + intra_guard) / (double)gate : INFINITY;
                if (out->inter_intra_cost_ratio) out->inter_intra_cost_ratio[pos] = ratio;
                const bool intra = b->allow_intra &&
                                   intra_sad + intra_guard < (uint64_t)llround((double)gate * p->inter_intra_threshold);
                sel_sad += intra ? intra_sad : ccost[best];
                rate_sad += ccost[best];
                sel_rate += intra ? intra_marker_bits : crate[best];
                sel_px += (uint64_t)imax(0, imin(16, W - x0)) * (uint64_t)imax(0, imin(16, H - y0));
                mv_t fo = {0, 0}, bo = {0, 0};
                int m = VC1ME_B_INTRA;
                if (intra) {
                    ++out->intra_macroblocks;
                } else {
                    m = cmode[best];
                    if (m == VC1ME_B_DIRECT) { fo = df; bo = db; ++out->direct_macroblocks; }
                    else if (m == VC1ME_B_FORWARD) { fo = fm; bo = db; ++out->forward_macroblocks; }
                    else if (m == VC1ME_B_BACKWARD) { fo = df; bo = bm; ++out->backward_macroblocks; }
                    else { fo = fm; bo = bm; ++out->interpolated_macroblocks; }
                }
                fout[pos] = fo;
                bout[pos] = bo;
                if (out->modes) out->modes[pos] = (uint8_t)m;
                if (out->forward_mvs) { out->forward_mvs[2 * pos] = fo.xq; out->forward_mvs[2 * pos + 1] = fo.yq; }
                if (out->backward_mvs) { out->backward_mvs[2 * pos] = bo.xq; out->backward_mvs[2 * pos + 1] = bo.yq; }
                if (fo.xq || fo.yq || bo.xq || bo.yq) ++out->moved_macroblocks;
#define FRAC_CHROMA(v) (((v).xq & 3) || ((v).yq & 3) || (((v).xq / 4) & 1) || (((v).yq / 4) & 1))
                if (FRAC_CHROMA(fo) || FRAC_CHROMA(bo)) ++out->fractional_chroma_macroblocks;
#undef FRAC_CHROMA
            }
        out->mean_abs_residual = sel_px ? (double)sel_sad / (double)sel_px : 0.0;
        out->rate_complexity = sel_px ? (double)rate_sad / (double)sel_px : 0.0;
        out->motion_sad = sel_sad;
        out->motion_bits = sel_rate;
    }
cleanup:
    p_result_free(&fwd);
    p_result_free(&bwd);
    free(wy); free(wu); free(wv);
    free(fout); free(bout);
    return rc;
}

/* ------------------------------------------------------------------ scene cuts */

/* global_motion_search(cur, ref, range) as a frame-level call (the search
 * range is clamped to 32, src/encoder_motion.cpp:2076-2141) */
static cand_t global_for(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref, int range) {
    ctx_t cx;
    init_ctx(&cx, p, cur, ref, NULL);
    return global_search(&cx, range);
}

/* Vc1Encoder::scene_change_score (src/encoder_motion.cpp:599-676): the
 * global-translation MAD, the intensity-compensated one when fade
 * compensation is on, and, for a score at or above the threshold, the 1-MV
 * quarter-pel core's MAD if the interior vector field fits translation plus
 * isotropic zoom (>= 60% of vectors within 2.5 px, 0.005 <= |zoom| <= 0.08)
 * and that MAD is clearly lower. */
int oracle_scene_change_score(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                              int fade_compensation, double scene_threshold, double* out) {
    const int W = p->width, H = p->height;
    if (cur->width != W || cur->height != H || ref->width != W || ref->height != H) return -1;
    const cand_t g = global_for(p, cur, ref, p->motion_search_range);
    const uint64_t pixels = (uint64_t)W * H;
    double best = pixels ? (double)g.cost / (double)pixels : 0.0;
    const int mbw = (W + 15) / 16, mbh = (H + 15) / 16;
    const size_t mbs = (size_t)mbw * mbh;
    if (fade_compensation) {
        int32_t* uni = (int32_t*)malloc(mbs * 2 * sizeof(int32_t));
        for (size_t i = 0; i < mbs; ++i) { uni[2 * i] = g.dx * 4; uni[2 * i + 1] = g.dy * 4; }
        vc1me_intensity ic;
        oracle_estimate_intensity_comp(cur, ref, uni, &ic);
        free(uni);
        if (ic.enabled) {
            const size_t ys = (size_t)W * H, cs = (size_t)(W / 2) * (H / 2);
            uint8_t* wy = (uint8_t*)malloc(ys);
            uint8_t* wu = (uint8_t*)malloc(cs);
            uint8_t* wv = (uint8_t*)malloc(cs);
            oracle_intensity_compensate(ref, &ic, wy, wu, wv);
            const vc1me_planes wp = {wy, wu, wv, W, H};
            const cand_t wg = global_for(p, cur, &wp, p->motion_search_range);
            const double wm = pixels ? (double)wg.cost / (double)pixels : 0.0;
            if (wm < best) best = wm;
            free(wy); free(wu); free(wv);
        }
    }
    if (best >= scene_threshold) {
        vc1me_params q = *p;
        q.mode = VC1ME_MODE_ONE_MV_QPEL;
        q.allow_intra = 0;
        vc1me_p_result local;
        p_result_alloc(&local, mbs);
        const int rc = analyze_core(&q, cur, ref, NULL, &local, NULL);
        if (rc) { p_result_free(&local); return rc; }
        const double lm = local.mean_abs_residual;
        long double mx = 0.0L, my = 0.0L, mdx = 0.0L, mdy = 0.0L;
        size_t samples = 0;
        for (int by = 1; by < mbh - 1; ++by)
            for (int bx = 1; bx < mbw - 1; ++bx) {
                const size_t pos = (size_t)by * mbw + bx;
                mx += bx * 16.0L + 8.0L; my += by * 16.0L + 8.0L;
                mdx += local.mvs[2 * pos] / 4.0L; mdy += local.mvs[2 * pos + 1] / 4.0L;
                ++samples;
            }
        bool coherent = false;
        if (samples >= 12) {
            mx /= samples; my /= samples; mdx /= samples; mdy /= samples;
            long double num = 0.0L, den = 0.0L;
            for (int by = 1; by < mbh - 1; ++by)
                for (int bx = 1; bx < mbw - 1; ++bx) {
                    const size_t pos = (size_t)by * mbw + bx;
                    const long double x = bx * 16.0L + 8.0L - mx, y = by * 16.0L + 8.0L - my;
                    const long double dx = local.mvs[2 * pos] / 4.0L - mdx, dy = local.mvs[2 * pos + 1] / 4.0L - mdy;
                    num += x * dx + y * dy;
                    den += x * x + y * y;
                }
            const long double z = den > 0.0L ? num / den : 0.0L;
            size_t inliers = 0;
            for (int by = 1; by < mbh - 1; ++by)
                for (int bx = 1; bx < mbw - 1; ++bx) {
                    const size_t pos = (size_t)by * mbw + bx;
                    const long double x = bx * 16.0L + 8.0L - mx, y = by * 16.0L + 8.0L - my;
                    const long double ex = local.mvs[2 * pos] / 4.0L - (mdx + z * x);
                    const long double ey = local.mvs[2 * pos + 1] / 4.0L - (mdy + z * y);
                    if (ex * ex + ey * ey <= 6.25L) ++inliers;
                }
            const long double az = fabsl(z);
            coherent = inliers * 5 >= samples * 3 && az >= 0.005L && az <= 0.08L;
        }
        if (coherent && lm < best * 0.72 && lm < scene_threshold * 0.86) best = lm;
        p_result_free(&local);
    }
    *out = best;
    return 0;
}

/* ------------------------------------------------------------------ motion failure */

/* the (SAD, |mv|, y, x) order of refine_qpel_block */
static bool qpel_better(uint64_t sad, mv_t mv, uint64_t best, mv_t bmv) {
    const int mag = iabs(mv.xq) + iabs(mv.yq), bmag = iabs(bmv.xq) + iabs(bmv.yq);
    if (sad != best) return sad < best;
    if (mag != bmag) return mag < bmag;
    if (mv.yq != bmv.yq) return mv.yq < bmv.yq;
    return mv.xq < bmv.xq;
}

/* Vc1Encoder::refine_qpel_block (src/encoder_motion.cpp:3596-3685): the
 * SAD-only quarter-pel refinement (half-pel 3x3 around the seed, then
 * quarter-pel 3x3 around the half-pel winner; the fast mode probes
 * cardinals, then the corners beside the winning axis, at each scale). */
static mv_t refine_qpel_sad(const ctx_t* c, int x0, int y0, int bw, int bh, mv_t seed, int range, uint64_t* best_sad,
                            uint64_t known_seed_sad) {
    const vc1me_params* P = c->p;
    const int limit_q = 4 * P->subpixel_refinement_range;
    if (!P->subpixel_refinement || limit_q == 0 || imax(iabs(seed.xq), iabs(seed.yq)) > limit_q)
        return refine_integer_only(c, x0, y0, bw, bh, seed, range, VC1ME_MODE_ONE_MV_QPEL, best_sad, known_seed_sad);
    const bounds_t qb = bounds_qpel(range);
    const int xmin = imax(qb.xmin, -limit_q), xmax = imin(qb.xmax, limit_q);
    const int ymin = imax(qb.ymin, -limit_q), ymax = imin(qb.ymax, limit_q);
    seed.xq = clampi(seed.xq, xmin, xmax);
    seed.yq = clampi(seed.yq, ymin, ymax);
    if (hidden_padding(c, x0, y0, bw, bh, seed, VC1ME_MODE_ONE_MV_QPEL, chroma_block(x0, y0, bw, bh))) {
        seed.xq = seed.yq = 0;
        known_seed_sad = U64MAX;
    }
    mv_t best = seed;
    uint64_t cost = known_seed_sad != U64MAX ? known_seed_sad
                                             : block_sad_motion(c, x0, y0, bw, bh, best, VC1ME_MODE_ONE_MV_QPEL);
    mv_t seen[20];
    int nseen = 1;
    seen[0] = seed;
#define EVAL(mx_, my_)                                                                         \
    do {                                                                                       \
        mv_t mv_ = {clampi((mx_), xmin, xmax), clampi((my_), ymin, ymax)};                     \
        bool dup_ = false;                                                                     \
        for (int s_ = 0; s_ < nseen; ++s_) dup_ |= same_mv(seen[s_], mv_);                     \
        if (!dup_) {                                                                           \
            if (nseen < 20) seen[nseen++] = mv_;                                               \
            const uint64_t sad_ = block_sad_motion(c, x0, y0, bw, bh, mv_, VC1ME_MODE_ONE_MV_QPEL); \
            if (qpel_better(sad_, mv_, cost, best)) { cost = sad_; best = mv_; }                \
        }                                                                                      \
    } while (0)
    if (P->fast_subpixel_refinement) {
        if (cost == 0 && best.xq == 0 && best.yq == 0) {
            if (best_sad) *best_sad = 0;
            return best;
        }
        static const int card[4][2] = {{-2, 0}, {2, 0}, {0, -2}, {0, 2}};
        for (int k = 0; k < 4; ++k) EVAL(seed.xq + card[k][0], seed.yq + card[k][1]);
        const mv_t axis = best;
        const int dx = axis.xq - seed.xq, dy = axis.yq - seed.yq;
        if (dx && !dy) { EVAL(axis.xq, axis.yq - 2); EVAL(axis.xq, axis.yq + 2); }
        else if (dy && !dx) { EVAL(axis.xq - 2, axis.yq); EVAL(axis.xq + 2, axis.yq); }
        const mv_t half = best;
        static const int quar[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
        for (int k = 0; k < 4; ++k) EVAL(half.xq + quar[k][0], half.yq + quar[k][1]);
        const mv_t qaxis = best;
        const int qdx = qaxis.xq - half.xq, qdy = qaxis.yq - half.yq;
        if (qdx && !qdy) { EVAL(qaxis.xq, qaxis.yq - 1); EVAL(qaxis.xq, qaxis.yq + 1); }
        else if (qdy && !qdx) { EVAL(qaxis.xq - 1, qaxis.yq); EVAL(qaxis.xq + 1, qaxis.yq); }
    } else {
        for (int oy = -2; oy <= 2; oy += 2)
            for (int ox = -2; ox <= 2; ox += 2) EVAL(seed.xq + ox, seed.yq + oy);
        const mv_t half = best;
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) EVAL(half.xq + ox, half.yq + oy);
    }
#undef EVAL
    if (best_sad) *best_sad = cost;
    return best;
}

/* Vc1Encoder::has_good_motion_match (src/encoder_motion.cpp:678-751): the
 * P search's integer and long-range stages per macroblock in raster order
 * (against the fade-compensated reference when fade compensation finds a
 * code on the uniform global field), each refined by refine_qpel_block; true
 * at the first macroblock whose refined SAD is at most 12 per pixel or at
 * most 90% of its zero-vector SAD. */
int oracle_has_good_motion_match(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref,
                                 int fade_compensation, int* out, int32_t* field_out) {
    const int W = p->width, H = p->height;
    if (cur->width != W || cur->height != H || ref->width != W || ref->height != H) return -1;
    const int mbw = (W + 15) / 16, mbh = (H + 15) / 16;
    const size_t mbs = (size_t)mbw * mbh;
    const int range = imax(0, p->motion_search_range);
    const int local_range = imin(range, imax(0, p->motion_local_search_range));
    cand_t global = global_for(p, cur, ref, local_range);
    vc1me_planes pw = *ref;
    uint8_t *wy = NULL, *wu = NULL, *wv = NULL;
    if (fade_compensation) {
        int32_t* uni = (int32_t*)malloc(mbs * 2 * sizeof(int32_t));
        for (size_t i = 0; i < mbs; ++i) { uni[2 * i] = global.dx * 4; uni[2 * i + 1] = global.dy * 4; }
        vc1me_intensity ic;
        oracle_estimate_intensity_comp(cur, ref, uni, &ic);
        free(uni);
        if (ic.enabled) {
            const size_t ys = (size_t)W * H, cs = (size_t)(W / 2) * (H / 2);
            wy = (uint8_t*)malloc(ys); wu = (uint8_t*)malloc(cs); wv = (uint8_t*)malloc(cs);
            oracle_intensity_compensate(ref, &ic, wy, wu, wv);
            pw.y = wy; pw.u = wu; pw.v = wv;
            global = global_for(p, cur, &pw, local_range);
        }
    }
    ctx_t cx;
    init_ctx(&cx, p, cur, &pw, NULL);
    const ctx_t* c = &cx;
    mv_t* field = (mv_t*)calloc(mbs, sizeof(mv_t));
    uint8_t* prop = (uint8_t*)calloc(mbs, 1);
    cell_t* cells = NULL;
    int ncells = 0;
    int good = 0;
    /* with field_out the scan continues past the first usable match and
     * reports every macroblock's refined vector */
    for (int my = 0; my < mbh && (!good || field_out); ++my)
        for (int mx = 0; mx < mbw && (!good || field_out); ++mx) {
            const size_t pos = (size_t)my * mbw + mx;
            const pred_t pi = predictor_info(c, field, mx, my);
            imr_t im;
            bool selected_distant = false;
            if (p->long_range_search_mode == 2) {
                bool done = false;
                if (range > local_range) {
                    const imr_t pr = propagated_long_range(c, mx, my, field, prop, range);
                    if (match_good(c, pr.sad, mx, my)) { im = pr; done = true; selected_distant = true; }
                }
                if (!done) im = integer_local(c, mx, my, &pi, global, local_range);
            } else {
                im = integer_local(c, mx, my, &pi, global, local_range);
                const bool skip = p->long_range_search_mode == 1 && match_good(c, im.sad, mx, my);
                if (range > local_range && !skip) {
                    const imr_t pr = propagated_long_range(c, mx, my, field, prop, range);
                    if (better4(pr.sad, pr.dx, pr.dy, im.sad, im.dx, im.dy) &&
                        distant_decision(c, pr.sad, im.sad, mx, my) >= 2) {
                        im = pr;
                        selected_distant = true;
                    }
                }
            }
            if (range > local_range && !match_good(c, im.sad, mx, my)) {
                if (!cells) {
                    const int nx = (W + 7) / 8, ny = (H + 7) / 8;
                    cells = (cell_t*)malloc(sizeof(cell_t) * (size_t)nx * ny);
                    for (int y = 0; y < H; y += 8)
                        for (int x = 0; x < W; x += 8) {
                            cells[ncells].x = x;
                            cells[ncells].y = y;
                            cells[ncells].s = signature(c->ry, c->ru, c->rv, W, H, x, y);
                            ++ncells;
                        }
                }
                uint64_t ext_sad = U64MAX;
                int ext_dec = 0;
                const imr_t ext = extended_content(c, mx, my, range, cells, ncells, im, &ext_sad, &ext_dec);
                if (ext.dx != im.dx || ext.dy != im.dy) { im = ext; selected_distant = true; }
            }
            if (selected_distant) prop[pos] = 1;
            uint64_t refined = im.sad;
            const mv_t seed = {im.dx * 4, im.dy * 4};
            field[pos] = refine_qpel_sad(c, mx * 16, my * 16, 16, 16, seed, range, &refined, im.sad);
            const uint64_t pixels = active_mb_pixels(c, mx, my);
            const uint64_t zero_sad = mb_sad(c, mx, my, 0, 0, 1);
            if (pixels && (refined <= pixels * 12u || (zero_sad > 0 && refined * 10u <= zero_sad * 9u))) good = 1;
        }
    if (field_out)
        for (size_t i = 0; i < mbs; ++i) { field_out[2 * i] = field[i].xq; field_out[2 * i + 1] = field[i].yq; }
    free(field); free(prop); free(cells);
    free(wy); free(wu); free(wv);
    *out = good;
    return 0;
}

/* refine_qpel_block on one block (for its tests) */
int oracle_refine_qpel_block(const vc1me_params* p, const vc1me_planes* cur, const vc1me_planes* ref, int x0, int y0,
                             int bw, int bh, int seed_xq, int seed_yq, int range, uint64_t known_seed_sad,
                             int32_t* xq, int32_t* yq, uint64_t* sad) {
    ctx_t cx;
    init_ctx(&cx, p, cur, ref, NULL);
    const mv_t seed = {seed_xq, seed_yq};
    const mv_t r = refine_qpel_sad(&cx, x0, y0, bw, bh, seed, range, sad, known_seed_sad);
    *xq = r.xq;
    *yq = r.yq;
    return 0;
}

/* ------------------------------------------------------------------ motion compensation */

/* Vc1Encoder::chroma_mc_sample (src/encoder_motion_compensation.cpp:417-447),
 * rounding control off; future: the B future-reference rounding of odd
 * chroma components (away from the even value) */
static int chroma_mc_sample_dir(const uint8_t* p, int w, int h, int x, int y, int mvx, int mvy, bool future) {
    int uvx = arshift(mvx + ((mvx & 3) == 3), 1), uvy = arshift(mvy + ((mvy & 3) == 3), 1);
    if (future) {
        uvx = uvx < 0 ? uvx - (uvx & 1) : uvx + (uvx & 1);
        uvy = uvy < 0 ? uvy - (uvy & 1) : uvy + (uvy & 1);
    } else {
        uvx = uvx < 0 ? uvx + (uvx & 1) : uvx - (uvx & 1);
        uvy = uvy < 0 ? uvy + (uvy & 1) : uvy - (uvy & 1);
    }
    const int ix = floor_div4(uvx), iy = floor_div4(uvy);
    const int fx = (uvx - ix * 4) * 2, fy = (uvy - iy * 4) * 2;
    const int sx = x + ix, sy = y + iy;
    const int a = sample_clamped(p, w, h, sx, sy);
    if (!fx && !fy) return a;
    const int b = sample_clamped(p, w, h, sx + 1, sy), cc = sample_clamped(p, w, h, sx, sy + 1);
    const int d = sample_clamped(p, w, h, sx + 1, sy + 1);
    return (a * (8 - fx) * (8 - fy) + b * fx * (8 - fy) + cc * (8 - fx) * fy + d * fx * fy + 32) >> 6;
}
static int chroma_mc_sample(const uint8_t* p, int w, int h, int x, int y, int mvx, int mvy) {
    return chroma_mc_sample_dir(p, w, h, x, y, mvx, mvy, false);
}

/* The P prediction picture of an analysis: motion_compensate_4mv
 * (src/encoder_motion_compensation.cpp:536-571) for 4-MV macroblocks (four
 * bicubic luma blocks, chroma at the median-4 vector), motion_compensate_mb
 * (:474-521) with the picture's mode for the others. */
void oracle_predict_p_frame(const vc1me_planes* ref, int mode, const int32_t* mvs, const int32_t* block_mvs,
                            const uint8_t* use_4mv, uint8_t* y, uint8_t* u, uint8_t* v) {
    const int W = ref->width, H = ref->height, cw = W / 2, ch = H / 2;
    const int mbw = (W + 15) / 16;
    for (int py = 0; py < H; ++py)
        for (int px = 0; px < W; ++px) {
            const size_t pos = (size_t)(py / 16) * mbw + px / 16;
            int mvx, mvy, m = mode;
            if (use_4mv[pos]) {
                const int k = ((py / 8) & 1) * 2 + ((px / 8) & 1);
                mvx = block_mvs[8 * pos + 2 * k];
                mvy = block_mvs[8 * pos + 2 * k + 1];
                m = VC1ME_MODE_ONE_MV_QPEL;
            } else {
                mvx = mvs[2 * pos];
                mvy = mvs[2 * pos + 1];
            }
            y[(size_t)py * W + px] = (uint8_t)luma_mc_sample_mode(ref->y, W, H, px, py, mvx, mvy, m, 0);
        }
    for (int py = 0; py < ch; ++py)
        for (int px = 0; px < cw; ++px) {
            const size_t pos = (size_t)(py / 8) * mbw + px / 8;
            int mvx, mvy;
            if (use_4mv[pos]) {
                const int32_t* b = block_mvs + 8 * pos;
                mvx = median4(b[0], b[2], b[4], b[6]);
                mvy = median4(b[1], b[3], b[5], b[7]);
            } else {
                mvx = mvs[2 * pos];
                mvy = mvs[2 * pos + 1];
            }
            u[(size_t)py * cw + px] = (uint8_t)chroma_mc_sample(ref->u, cw, ch, px, py, mvx, mvy);
            v[(size_t)py * cw + px] = (uint8_t)chroma_mc_sample(ref->v, cw, ch, px, py, mvx, mvy);
        }
}

/* The bi-directional prediction picture: motion_compensate_bi_mb
 * (src/encoder_motion_compensation.cpp:659-711) for every macroblock, the
 * rounding average of the past prediction at fmvs and the future prediction
 * at bmvs (chroma with the future-reference rounding). */
void oracle_predict_bi_frame(const vc1me_planes* past, const vc1me_planes* future, int mode, const int32_t* fmvs,
                             const int32_t* bmvs, uint8_t* y, uint8_t* u, uint8_t* v) {
    const int W = past->width, H = past->height, cw = W / 2, ch = H / 2;
    const int mbw = (W + 15) / 16;
    for (int py = 0; py < H; ++py)
        for (int px = 0; px < W; ++px) {
            const size_t pos = (size_t)(py / 16) * mbw + px / 16;
            const int a = luma_mc_sample_mode(past->y, W, H, px, py, fmvs[2 * pos], fmvs[2 * pos + 1], mode, 0);
            const int z = luma_mc_sample_mode(future->y, W, H, px, py, bmvs[2 * pos], bmvs[2 * pos + 1], mode, 0);
            y[(size_t)py * W + px] = (uint8_t)((a + z + 1) >> 1);
        }
    for (int py = 0; py < ch; ++py)
        for (int px = 0; px < cw; ++px) {
            const size_t pos = (size_t)(py / 8) * mbw + px / 8;
            const int fx = fmvs[2 * pos], fy = fmvs[2 * pos + 1], bx = bmvs[2 * pos], by = bmvs[2 * pos + 1];
            const size_t o = (size_t)py * cw + px;
            u[o] = (uint8_t)((chroma_mc_sample_dir(past->u, cw, ch, px, py, fx, fy, false) +
                              chroma_mc_sample_dir(future->u, cw, ch, px, py, bx, by, true) + 1) >> 1);
            v[o] = (uint8_t)((chroma_mc_sample_dir(past->v, cw, ch, px, py, fx, fy, false) +
                              chroma_mc_sample_dir(future->v, cw, ch, px, py, bx, by, true) + 1) >> 1);
        }
}

/* The padded P prediction picture: motion_compensate_mb_padded and
 * motion_compensate_4mv_padded (src/encoder_motion_compensation.cpp:573-657)
 * for every macroblock of an mbw x mbh raster.  Unlike the picture above,
 * every sample of each 16x16 luma and 8x8 chroma macroblock is written, into
 * a picture of mbw*16 x mbh*16 (chroma mbw*8 x mbh*8), and the samples clamp
 * to the reference's own raster, which may itself be the padded one. */
void oracle_predict_p_frame_padded(const vc1me_planes* ref, int mbw, int mbh, int mode, const int32_t* mvs,
                                   const int32_t* block_mvs, const uint8_t* use_4mv, uint8_t* y, uint8_t* u,
                                   uint8_t* v) {
    const int RW = ref->width, RH = ref->height, rcw = RW / 2, rch = RH / 2;
    const int DW = mbw * 16, DH = mbh * 16, dcw = DW / 2, dch = DH / 2;
    for (int py = 0; py < DH; ++py)
        for (int px = 0; px < DW; ++px) {
            const size_t pos = (size_t)(py / 16) * mbw + px / 16;
            int mvx, mvy, m = mode;
            if (use_4mv[pos]) {
                /* the four 8x8 blocks are bicubic whatever the mode (:573-595) */
                const int k = ((py / 8) & 1) * 2 + ((px / 8) & 1);
                mvx = block_mvs[8 * pos + 2 * k];
                mvy = block_mvs[8 * pos + 2 * k + 1];
                m = VC1ME_MODE_ONE_MV_QPEL;
            } else {
                mvx = mvs[2 * pos];
                mvy = mvs[2 * pos + 1];
            }
            y[(size_t)py * DW + px] = (uint8_t)luma_mc_sample_mode(ref->y, RW, RH, px, py, mvx, mvy, m, 0);
        }
    for (int py = 0; py < dch; ++py)
        for (int px = 0; px < dcw; ++px) {
            const size_t pos = (size_t)(py / 8) * mbw + px / 8;
            int mvx, mvy;
            if (use_4mv[pos]) {
                const int32_t* b = block_mvs + 8 * pos;
                mvx = median4(b[0], b[2], b[4], b[6]);
                mvy = median4(b[1], b[3], b[5], b[7]);
            } else {
                mvx = mvs[2 * pos];
                mvy = mvs[2 * pos + 1];
            }
            u[(size_t)py * dcw + px] = (uint8_t)chroma_mc_sample(ref->u, rcw, rch, px, py, mvx, mvy);
            v[(size_t)py * dcw + px] = (uint8_t)chroma_mc_sample(ref->v, rcw, rch, px, py, mvx, mvy);
        }
}

/* The padded bi-directional picture: motion_compensate_bi_mb_padded
 * (src/encoder_motion_compensation.cpp:712-763) for every macroblock of an
 * mbw x mbh raster, each reference sampled with clamping to its own raster. */
void oracle_predict_bi_frame_padded(const vc1me_planes* past, const vc1me_planes* future, int mbw, int mbh, int mode,
                                    const int32_t* fmvs, const int32_t* bmvs, uint8_t* y, uint8_t* u, uint8_t* v) {
    const int PW = past->width, PH = past->height, FW = future->width, FH = future->height;
    const int DW = mbw * 16, DH = mbh * 16, dcw = DW / 2, dch = DH / 2;
    for (int py = 0; py < DH; ++py)
        for (int px = 0; px < DW; ++px) {
            const size_t pos = (size_t)(py / 16) * mbw + px / 16;
            const int a = luma_mc_sample_mode(past->y, PW, PH, px, py, fmvs[2 * pos], fmvs[2 * pos + 1], mode, 0);
            const int z = luma_mc_sample_mode(future->y, FW, FH, px, py, bmvs[2 * pos], bmvs[2 * pos + 1], mode, 0);
            y[(size_t)py * DW + px] = (uint8_t)((a + z + 1) >> 1);
        }
    for (int py = 0; py < dch; ++py)
        for (int px = 0; px < dcw; ++px) {
            const size_t pos = (size_t)(py / 8) * mbw + px / 8;
            const int fx = fmvs[2 * pos], fy = fmvs[2 * pos + 1], bx = bmvs[2 * pos], by = bmvs[2 * pos + 1];
            const size_t o = (size_t)py * dcw + px;
            u[o] = (uint8_t)((chroma_mc_sample_dir(past->u, PW / 2, PH / 2, px, py, fx, fy, false) +
                              chroma_mc_sample_dir(future->u, FW / 2, FH / 2, px, py, bx, by, true) + 1) >> 1);
            v[o] = (uint8_t)((chroma_mc_sample_dir(past->v, PW / 2, PH / 2, px, py, fx, fy, false) +
                              chroma_mc_sample_dir(future->v, FW / 2, FH / 2, px, py, bx, by, true) + 1) >> 1);
        }
}
