// vc1me_nn.hip — the motion-proposal network of the learned-seeded form.
//
// No counterpart in the reference: the network only proposes integer-search
// seeds; the search judges them with exact costs.
//
// Network (per frame pair, 1/4-resolution luma of cur and ref as 2 channels):
//   conv3x3 2->32 (dilation 1), conv3x3 32->32 (dilation 2, 4, 8), each with
//   bias and ReLU; mean pool over each macroblock's 4x4 cell; 1x1 head 32->5:
//   two vector candidates (dx, dy in units of 64 px) and a confidence logit.
// Activations are bf16 in NHWC order; the three GEMM-shaped operations of each
// convolution (forward, data gradient, weight gradient) run on MFMA
// (v_mfma_f32_16x16x16_bf16, fp32 accumulation).  Lane layout of that
// instruction: lane l holds A[l%16][4(l/16)+j], B[4(l/16)+j][l%16] and
// D[4(l/16)+j][l%16], j = 0..3.
#include "vc1me_nn_internal.h"

#include <atomic>

namespace pnet {
namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

// four consecutive bf16 at p (8-byte aligned) or zeros
__device__ __forceinline__ v4s load4(const uint16_t* p, bool ok) {
    v4s r = {0, 0, 0, 0};
    if (ok) {
        const uint2 u = *reinterpret_cast<const uint2*>(p);
        r[0] = static_cast<short>(u.x & 0xffffu); r[1] = static_cast<short>(u.x >> 16);
        r[2] = static_cast<short>(u.y & 0xffffu); r[3] = static_cast<short>(u.y >> 16);
    }
    return r;
}

// ------------------------------------------------------------------ input

// 4x4 box mean of cur and ref luma (edge samples repeated past the frame),
// scaled to [-0.5, 0.5], bf16 [h4][w4][2] with h4 = ceil(H/4), w4 = ceil(W/4).
// The ref block is taken displaced by the global vector gxy = {dx, dy} (px,
// device memory, null = none), so the network sees ref aligned with the
// frame's dominant motion and predicts residuals from it.
__device__ __forceinline__ void downsample_cell(const uint8_t* __restrict__ cur, const uint8_t* __restrict__ ref, int W,
                                                int H, int w4, int i, int gx, int gy, uint16_t* __restrict__ x0) {
    const int y = i / w4, x = i % w4;
    uint32_t sc = 0, sr = 0;
    for (int dy = 0; dy < 4; ++dy)
        for (int dx = 0; dx < 4; ++dx) {
            const size_t o = static_cast<size_t>(min(y * 4 + dy, H - 1)) * W + min(x * 4 + dx, W - 1);
            const size_t r = static_cast<size_t>(min(max(y * 4 + dy + gy, 0), H - 1)) * W +
                             min(max(x * 4 + dx + gx, 0), W - 1);
            sc += cur[o];
            sr += ref[r];
        }
    x0[2 * i] = to_bf16(static_cast<float>(sc) / (16.0f * 255.0f) - 0.5f);
    x0[2 * i + 1] = to_bf16(static_cast<float>(sr) / (16.0f * 255.0f) - 0.5f);
}

__global__ void k_downsample(const uint8_t* __restrict__ cur, const uint8_t* __restrict__ ref, int W, int H, int h4,
                             int w4, const int32_t* __restrict__ gxy, uint16_t* __restrict__ x0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= h4 * w4) return;
    downsample_cell(cur, ref, W, H, w4, i, gxy ? gxy[0] : 0, gxy ? gxy[1] : 0, x0);
}

// The same for the pairs of an analysis step, read from its job table on the
// device (frames and global vectors), so that a captured step replays on new
// frames: image b of the batch is job idx.k[b]; grid (cells, images).
__global__ void k_downsample_jobs(const vc1me::JobDesc* __restrict__ jobs, JobIndex idx, int h4, int w4,
                                  uint16_t* __restrict__ x0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= h4 * w4) return;
    const vc1me::JobDesc* J = &jobs[idx.k[blockIdx.y]];
    downsample_cell(J->cy, J->ry, J->W, J->H, w4, i, J->g->dx, J->g->dy,
                    x0 + static_cast<size_t>(blockIdx.y) * h4 * w4 * 2);
}

// fp32 host-provided input [b][h4][w4][2] -> bf16 (gradient checks)
__global__ void k_to_bf16(const float* __restrict__ in, uint16_t* __restrict__ out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = to_bf16(in[i]);
}

// ------------------------------------------------------------------ convolution

// Forward of one 3x3 convolution with dilation dil, as D^T = W X^T: the wave
// keeps the layer's weights in registers as MFMA A operands (M = output
// channel, K = tap x input channel, tap-major) and
// walks tiles of 16 output pixels of one row (N = pixel), issuing every input
// load of a tile before its MFMAs.  D[4(l/16)+j][l%16] gives lane l four
// consecutive output channels of one pixel: one 8-byte store per 16 channels.
// Cin 32: 18 K steps (tap t = step / 2, channel half step % 2); Cin 2: taps x
// channels packed into one K of 32 (18 used), two steps.
template <int kCin>
__global__ void __launch_bounds__(64) k_conv_fwd_regw(const uint16_t* __restrict__ x, const uint16_t* __restrict__ wt,
                                                      const float* __restrict__ bias, uint16_t* __restrict__ out,
                                                      int nb, int h, int w, int dil) {
    constexpr int kSteps = kCin == 32 ? 18 : 2;
    const int l = threadIdx.x;
    const int m = l & 15, kq = 4 * (l >> 4);
    v4s wa[kSteps], wb[kSteps];
#pragma unroll
    for (int ks = 0; ks < kSteps; ++ks) {
        if (kCin == 32) {
            const int t = ks >> 1, c0 = (ks & 1) * 16 + kq;
            wa[ks] = load4(wt + (static_cast<size_t>(m) * 9 + t) * 32 + c0, true);
            wb[ks] = load4(wt + (static_cast<size_t>(16 + m) * 9 + t) * 32 + c0, true);
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = ks * 16 + kq + j;
                wa[ks][j] = k < 18 ? static_cast<short>(wt[m * 18 + k]) : 0;
                wb[ks][j] = k < 18 ? static_cast<short>(wt[(16 + m) * 18 + k]) : 0;
            }
        }
    }
    float b0[4], b1[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) { b0[j] = bias[kq + j]; b1[j] = bias[16 + kq + j]; }
    const int tiles_x = (w + 15) / 16;
    const int per = h * tiles_x, ntiles = nb * per;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int b = tile / per, rem = tile - b * per;
        const int y = rem / tiles_x, px = (rem - y * tiles_x) * 16 + m;
        const uint16_t* xb = x + static_cast<size_t>(b) * h * w * kCin;
        v4s xv[kSteps];
#pragma unroll
        for (int ks = 0; ks < kSteps; ++ks) {
            if (kCin == 32) {
                const int t = ks >> 1, c0 = (ks & 1) * 16 + kq;
                const int sy = y + (t / 3 - 1) * dil, sx = px + (t % 3 - 1) * dil;
                const bool in = sy >= 0 && sy < h && sx >= 0 && sx < w && px < w;
                xv[ks] = load4(xb + (static_cast<size_t>(sy) * w + sx) * 32 + c0, in);
            } else {
                v4s a = {0, 0, 0, 0};
#pragma unroll
                for (int pr = 0; pr < 2; ++pr) {
                    const int k = ks * 16 + kq + 2 * pr, t = k >> 1;   // channels 0 and 1 of tap t
                    const int sy = y + (t / 3 - 1) * dil, sx = px + (t % 3 - 1) * dil;
                    if (k < 18 && sy >= 0 && sy < h && sx >= 0 && sx < w && px < w) {
                        const uint32_t u = *reinterpret_cast<const uint32_t*>(xb + (static_cast<size_t>(sy) * w + sx) * 2);
                        a[2 * pr] = static_cast<short>(u & 0xffffu);
                        a[2 * pr + 1] = static_cast<short>(u >> 16);
                    }
                }
                xv[ks] = a;
            }
        }
        v4f acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < kSteps; ++ks) {
            acc0 = mfma(wa[ks], xv[ks], acc0);
            acc1 = mfma(wb[ks], xv[ks], acc1);
        }
        if (px < w) {
            uint16_t* o = out + ((static_cast<size_t>(b) * h + y) * w + px) * 32 + kq;
            uint32_t lo0 = to_bf16(fmaxf(acc0[0] + b0[0], 0.f)) | (static_cast<uint32_t>(to_bf16(fmaxf(acc0[1] + b0[1], 0.f))) << 16);
            uint32_t hi0 = to_bf16(fmaxf(acc0[2] + b0[2], 0.f)) | (static_cast<uint32_t>(to_bf16(fmaxf(acc0[3] + b0[3], 0.f))) << 16);
            uint32_t lo1 = to_bf16(fmaxf(acc1[0] + b1[0], 0.f)) | (static_cast<uint32_t>(to_bf16(fmaxf(acc1[1] + b1[1], 0.f))) << 16);
            uint32_t hi1 = to_bf16(fmaxf(acc1[2] + b1[2], 0.f)) | (static_cast<uint32_t>(to_bf16(fmaxf(acc1[3] + b1[3], 0.f))) << 16);
            *reinterpret_cast<uint2*>(o) = make_uint2(lo0, hi0);
            *reinterpret_cast<uint2*>(o + 16) = make_uint2(lo1, hi1);
        }
    }
}

// Data gradient of a 3x3 convolution (Cin = Cout = 32): dx[p][ci] =
// sum_t sum_co dz[p - off_t][co] * W[co][t][ci], masked by the ReLU of the
// layer below (x > 0).  wtt: bf16 [9][ci][co].  One wave: 16 pixels x 32 ci.
__global__ void __launch_bounds__(64) k_conv_bwd_data(const uint16_t* __restrict__ dz, const uint16_t* __restrict__ wtt,
                                                      const uint16_t* __restrict__ xact, uint16_t* __restrict__ dx,
                                                      int nb, int h, int w, int dil) {
    const int l = threadIdx.x;
    const int tiles_x = (w + 15) / 16;
    const int tile = blockIdx.x;
    const int b = tile / (h * tiles_x), rem = tile - b * h * tiles_x;
    const int y = rem / tiles_x, x0 = (rem - y * tiles_x) * 16;
    if (b >= nb) return;
    const int px = x0 + (l & 15);
    const int kq = 4 * (l >> 4);
    v4f acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
    const uint16_t* db = dz + static_cast<size_t>(b) * h * w * 32;
    for (int t = 0; t < 9; ++t) {
        const int sy = y - (t / 3 - 1) * dil, sx = px - (t % 3 - 1) * dil;
        const bool in = sy >= 0 && sy < h && sx >= 0 && sx < w && px < w;
        for (int cb = 0; cb < 2; ++cb) {
            const int c0 = cb * 16 + kq;   // output-channel slice (K)
            const v4s a = load4(db + (static_cast<size_t>(sy) * w + sx) * 32 + c0, in);
            const v4s b0 = load4(wtt + (static_cast<size_t>(t) * 32 + (l & 15)) * 32 + c0, true);
            const v4s b1 = load4(wtt + (static_cast<size_t>(t) * 32 + 16 + (l & 15)) * 32 + c0, true);
            acc0 = mfma(a, b0, acc0);
            acc1 = mfma(a, b1, acc1);
        }
    }
    const uint16_t* xb = xact + static_cast<size_t>(b) * h * w * 32;
    uint16_t* ob = dx + static_cast<size_t>(b) * h * w * 32;
    for (int j = 0; j < 4; ++j) {
        const int opx = x0 + 4 * (l >> 4) + j;
        if (opx >= w) continue;
        const int ci = l & 15;
        const size_t o = (static_cast<size_t>(y) * w + opx) * 32;
        ob[o + ci] = from_bf16(xb[o + ci]) > 0.f ? to_bf16(acc0[j]) : 0;
        ob[o + 16 + ci] = from_bf16(xb[o + 16 + ci]) > 0.f ? to_bf16(acc1[j]) : 0;
    }
}

// Weight and bias gradients: dW[co][t][ci] = sum_p dz[p][co] * x[p + off_t][ci]
// over a chunk of pixels per wave (M = co, N = (t, ci), K = pixels), added to
// fp32 accumulators.  dz: bf16 [b][h][w][32]; x: bf16 [b][h][w][cin].
constexpr int kChunk = 1024;   // pixels per wave
__global__ void __launch_bounds__(64) k_conv_bwd_weight(const uint16_t* __restrict__ dz, const uint16_t* __restrict__ x,
                                                        float* __restrict__ dw, float* __restrict__ dbias, int nb,
                                                        int h, int w, int cin, int dil) {
    const int l = threadIdx.x;
    const int npx = nb * h * w;
    const int p0 = blockIdx.x * kChunk;
    if (p0 >= npx) return;
    const int kq = 4 * (l >> 4);
    const int mrow = l & 15;
    const int ntiles = cin == 32 ? 18 : 2;   // N = 9 * cin (32: 288 = 18 x 16; 2: 18 -> 2 tiles of 16)
    for (int nt = 0; nt < ntiles; ++nt) {
        v4f acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
        // this lane's B column n = nt * 16 + (l & 15) -> (tap, ci)
        const int n = nt * 16 + (l & 15);
        const int t = cin == 32 ? n / 32 : n / 2, ci = cin == 32 ? n % 32 : n % 2;
        const bool nvalid = n < 9 * cin;
        const int oy = (t / 3 - 1) * dil, ox = (t % 3 - 1) * dil;
        for (int k0 = 0; k0 < kChunk; k0 += 16) {
            v4s a0 = {0, 0, 0, 0}, a1 = {0, 0, 0, 0}, bb = {0, 0, 0, 0};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int p = p0 + k0 + kq + j;
                if (p >= npx) continue;
                const int b = p / (h * w), r = p - b * h * w, py = r / w, pxx = r - py * w;
                a0[j] = static_cast<short>(dz[static_cast<size_t>(p) * 32 + mrow]);
                a1[j] = static_cast<short>(dz[static_cast<size_t>(p) * 32 + 16 + mrow]);
                const int sy = py + oy, sx = pxx + ox;
                if (nvalid && sy >= 0 && sy < h && sx >= 0 && sx < w)
                    bb[j] = static_cast<short>(x[((static_cast<size_t>(b) * h + sy) * w + sx) * cin + ci]);
            }
            acc0 = mfma(a0, bb, acc0);
            acc1 = mfma(a1, bb, acc1);
        }
        // D[m = co 4(l/16)+j][n]
        if (nvalid)
            for (int j = 0; j < 4; ++j) {
                const int co = 4 * (l >> 4) + j;
                atomicAdd(&dw[(static_cast<size_t>(co) * 9 + t) * cin + ci], acc0[j]);
                atomicAdd(&dw[(static_cast<size_t>(16 + co) * 9 + t) * cin + ci], acc1[j]);
            }
    }
    // bias gradient: lanes 0..31 sum their channel over the chunk
    if (l < 32) {
        const int pend = min(p0 + kChunk, npx);
        float bsum = 0.f;
        for (int p = p0; p < pend; ++p) bsum += from_bf16(dz[static_cast<size_t>(p) * 32 + l]);
        atomicAdd(&dbias[l], bsum);
    }
}

// ------------------------------------------------------------------ pool, head, loss

// Mean pool of a4 over each macroblock's 4x4 cell and the 1x1 head.
// One thread per (image, macroblock).  head: [5][32] weights then 5 biases.
__global__ void k_pool_head(const uint16_t* __restrict__ a, const float* __restrict__ head, float* __restrict__ pooled,
                            float* __restrict__ out, int nb, int h, int w, int mbh, int mbw) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb * mbh * mbw) return;
    const int b = i / (mbh * mbw), r = i - b * mbh * mbw, my = r / mbw, mx = r - my * mbw;
    float s[32];
    for (int c = 0; c < 32; ++c) s[c] = 0.f;
    int cnt = 0;
    for (int dy = 0; dy < 4; ++dy)
        for (int dx = 0; dx < 4; ++dx) {
            const int y = my * 4 + dy, x = mx * 4 + dx;
            if (y >= h || x >= w) continue;
            ++cnt;
            const uint16_t* p = a + ((static_cast<size_t>(b) * h + y) * w + x) * 32;
            for (int c = 0; c < 32; ++c) s[c] += from_bf16(p[c]);
        }
    const float inv = cnt ? 1.f / static_cast<float>(cnt) : 0.f;
    for (int c = 0; c < 32; ++c) {
        s[c] *= inv;
        pooled[static_cast<size_t>(i) * 32 + c] = s[c];
    }
    for (int o = 0; o < 5; ++o) {
        float v = head[5 * 32 + o];
        for (int c = 0; c < 32; ++c) v += head[o * 32 + c] * s[c];
        out[static_cast<size_t>(i) * 5 + o] = v;
    }
}

// Head and pool backward: dhead += dout x pooled, and d(a4) spread evenly over
// each cell, masked by a4 > 0 (the layer's ReLU).  One thread per macroblock;
// head gradients accumulate atomically.
__global__ void k_pool_head_bwd(const float* __restrict__ dout, const float* __restrict__ pooled,
                                const float* __restrict__ head, const uint16_t* __restrict__ a,
                                float* __restrict__ dhead, uint16_t* __restrict__ da, int nb, int h, int w, int mbh,
                                int mbw) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb * mbh * mbw) return;
    const int b = i / (mbh * mbw), r = i - b * mbh * mbw, my = r / mbw, mx = r - my * mbw;
    float dp[32];
    for (int c = 0; c < 32; ++c) dp[c] = 0.f;
    for (int o = 0; o < 5; ++o) {
        const float g = dout[static_cast<size_t>(i) * 5 + o];
        if (g == 0.f) continue;
        atomicAdd(&dhead[5 * 32 + o], g);
        for (int c = 0; c < 32; ++c) {
            atomicAdd(&dhead[o * 32 + c], g * pooled[static_cast<size_t>(i) * 32 + c]);
            dp[c] += g * head[o * 32 + c];
        }
    }
    int cnt = 0;
    for (int dy = 0; dy < 4; ++dy)
        for (int dx = 0; dx < 4; ++dx)
            if (my * 4 + dy < h && mx * 4 + dx < w) ++cnt;
    const float inv = cnt ? 1.f / static_cast<float>(cnt) : 0.f;
    for (int dy = 0; dy < 4; ++dy)
        for (int dx = 0; dx < 4; ++dx) {
            const int y = my * 4 + dy, x = mx * 4 + dx;
            if (y >= h || x >= w) continue;
            const size_t o = ((static_cast<size_t>(b) * h + y) * w + x) * 32;
            for (int c = 0; c < 32; ++c) da[o + c] = from_bf16(a[o + c]) > 0.f ? to_bf16(dp[c] * inv) : 0;
        }
}

// Loss of one macroblock: candidates c_k = (o[2k], o[2k+1]) against the
// target t (64-px units); regression = min_k smoothL1(c_k - t); confidence
// label = the nearer candidate lies within 8 px in both components.
// loss = mean over valid macroblocks of regression + 0.5 * BCE(conf).
__global__ void k_loss(const float* __restrict__ out, const float* __restrict__ tgt, const uint8_t* __restrict__ valid,
                       float* __restrict__ dout, float* __restrict__ loss, int n, float inv_count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float* d = dout + static_cast<size_t>(i) * 5;
    for (int o = 0; o < 5; ++o) d[o] = 0.f;
    if (!valid[i]) return;
    const float* o = out + static_cast<size_t>(i) * 5;
    const float tx = tgt[2 * i], ty = tgt[2 * i + 1];
    auto sl1 = [](float e) { const float a = fabsf(e); return a < 1.f ? 0.5f * e * e : a - 0.5f; };
    auto dsl1 = [](float e) { return fabsf(e) < 1.f ? e : (e > 0.f ? 1.f : -1.f); };
    // errors in units of 8 px so the quadratic zone is the search's reach
    const float s = 8.f;
    float e[2];
    for (int k = 0; k < 2; ++k) e[k] = sl1((o[2 * k] - tx) * s) + sl1((o[2 * k + 1] - ty) * s);
    const int kb = e[1] < e[0] ? 1 : 0;
    d[2 * kb] = dsl1((o[2 * kb] - tx) * s) * s * inv_count;
    d[2 * kb + 1] = dsl1((o[2 * kb + 1] - ty) * s) * s * inv_count;
    const float near = fmaxf(fabsf(o[2 * kb] - tx), fabsf(o[2 * kb + 1] - ty)) * 64.f <= 8.f ? 1.f : 0.f;
    const float z = o[4];
    const float sig = 1.f / (1.f + expf(-z));
    const float bce = fmaxf(z, 0.f) - z * near + log1pf(expf(-fabsf(z)));
    d[4] = 0.5f * (sig - near) * inv_count;
    atomicAdd(loss, (e[kb] + 0.5f * bce) * inv_count);
}

// ------------------------------------------------------------------ optimizer

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int n, float lr, float b1, float b2, float c1, float c2, float eps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 copies of one layer's weights: [co][9][ci] and, for Cin = 32, [9][ci][co]
__global__ void k_pack_weights(const float* __restrict__ wf, uint16_t* __restrict__ wt, uint16_t* __restrict__ wtt,
                               int cin) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 32 * 9 * cin) return;
    const int co = i / (9 * cin), r = i - co * 9 * cin, t = r / cin, ci = r - t * cin;
    const uint16_t b = to_bf16(wf[i]);
    wt[i] = b;
    if (wtt) wtt[(static_cast<size_t>(t) * cin + ci) * 32 + co] = b;
}

}  // namespace

// ------------------------------------------------------------------ launchers

hipError_t launch_downsample(const uint8_t* cur, const uint8_t* ref, int W, int H, int h4, int w4, const int32_t* gxy,
                             uint16_t* x0, hipStream_t s) {
    const int n = h4 * w4;
    hipLaunchKernelGGL(k_downsample, dim3((n + 255) / 256), dim3(256), 0, s, cur, ref, W, H, h4, w4, gxy, x0);
    return hipGetLastError();
}
hipError_t launch_downsample_jobs(const vc1me::JobDesc* jobs, const JobIndex& idx, int nb, int h4, int w4, uint16_t* x0,
                                  hipStream_t s) {
    const int n = h4 * w4;
    if (nb <= 0 || nb > 255 || n <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_downsample_jobs, dim3((n + 255) / 256, nb), dim3(256), 0, s, jobs, idx, h4, w4, x0);
    return hipGetLastError();
}
hipError_t launch_to_bf16(const float* in, uint16_t* out, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_to_bf16, dim3((n + 255) / 256), dim3(256), 0, s, in, out, n);
    return hipGetLastError();
}
hipError_t launch_conv_fwd(const uint16_t* x, const uint16_t* wt, const float* bias, uint16_t* out, int nb, int h, int w,
                           int cin, int dil, hipStream_t s) {
    const int tiles = nb * h * ((w + 15) / 16);
    if (tiles <= 0) return hipSuccess;
    // resident waves that each keep the weights in registers and walk tiles
    // (one round of the current device's occupancy, looked up once per
    // device), fewer for small inputs
    constexpr int kDevices = 64;
    static std::atomic<int> resident[kDevices][2] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kDevices) dev = -1;
    int r = dev >= 0 ? resident[dev][cin == 32 ? 1 : 0].load(std::memory_order_relaxed) : 0;
    if (!r) {
        int cus = 0, per_cu = 0;
        if (dev >= 0 &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, cin == 32 ? reinterpret_cast<const void*>(k_conv_fwd_regw<32>)
                                   : reinterpret_cast<const void*>(k_conv_fwd_regw<2>), 64, 0) == hipSuccess &&
            cus > 0 && per_cu > 0)
            r = cus * per_cu;
        else
            r = 3072;
        if (dev >= 0) resident[dev][cin == 32 ? 1 : 0].store(r, std::memory_order_relaxed);
    }
    const int grid = tiles < r ? tiles : r;
    if (cin == 32)
        hipLaunchKernelGGL(k_conv_fwd_regw<32>, dim3(grid), dim3(64), 0, s, x, wt, bias, out, nb, h, w, dil);
    else if (cin == 2)
        hipLaunchKernelGGL(k_conv_fwd_regw<2>, dim3(grid), dim3(64), 0, s, x, wt, bias, out, nb, h, w, dil);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}
hipError_t launch_conv_bwd_data(const uint16_t* dz, const uint16_t* wtt, const uint16_t* xact, uint16_t* dx, int nb,
                                int h, int w, int dil, hipStream_t s) {
    const int tiles = nb * h * ((w + 15) / 16);
    hipLaunchKernelGGL(k_conv_bwd_data, dim3(tiles), dim3(64), 0, s, dz, wtt, xact, dx, nb, h, w, dil);
    return hipGetLastError();
}
hipError_t launch_conv_bwd_weight(const uint16_t* dz, const uint16_t* x, float* dw, float* dbias, int nb, int h, int w,
                                  int cin, int dil, hipStream_t s) {
    const int npx = nb * h * w;
    hipLaunchKernelGGL(k_conv_bwd_weight, dim3((npx + kChunk - 1) / kChunk), dim3(64), 0, s, dz, x, dw, dbias, nb, h, w,
                       cin, dil);
    return hipGetLastError();
}
hipError_t launch_pool_head(const uint16_t* a, const float* head, float* pooled, float* out, int nb, int h, int w,
                            int mbh, int mbw, hipStream_t s) {
    const int n = nb * mbh * mbw;
    hipLaunchKernelGGL(k_pool_head, dim3((n + 127) / 128), dim3(128), 0, s, a, head, pooled, out, nb, h, w, mbh, mbw);
    return hipGetLastError();
}
hipError_t launch_pool_head_bwd(const float* dout, const float* pooled, const float* head, const uint16_t* a,
                                float* dhead, uint16_t* da, int nb, int h, int w, int mbh, int mbw, hipStream_t s) {
    const int n = nb * mbh * mbw;
    hipLaunchKernelGGL(k_pool_head_bwd, dim3((n + 127) / 128), dim3(128), 0, s, dout, pooled, head, a, dhead, da, nb, h,
                       w, mbh, mbw);
    return hipGetLastError();
}
hipError_t launch_loss(const float* out, const float* tgt, const uint8_t* valid, float* dout, float* loss, int n,
                       float inv_count, hipStream_t s) {
    hipLaunchKernelGGL(k_loss, dim3((n + 127) / 128), dim3(128), 0, s, out, tgt, valid, dout, loss, n, inv_count);
    return hipGetLastError();
}
hipError_t launch_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float c1,
                       float c2, float eps, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, c1, c2, eps);
    return hipGetLastError();
}
hipError_t launch_pack_weights(const float* wf, uint16_t* wt, uint16_t* wtt, int cin, hipStream_t s) {
    const int n = 32 * 9 * cin;
    hipLaunchKernelGGL(k_pack_weights, dim3((n + 255) / 256), dim3(256), 0, s, wf, wt, wtt, cin);
    return hipGetLastError();
}

}  // namespace pnet
