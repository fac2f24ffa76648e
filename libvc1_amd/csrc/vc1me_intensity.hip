// vc1me_intensity.hip — intensity (fade) compensation of the P analysis on
// gfx950: the sampled statistics and SAD grid of
// Vc1Encoder::estimate_intensity_comp (src/encoder_motion_compensation.cpp:296-348)
// and the plane remap of Vc1Encoder::intensity_compensated_reference
// (src/encoder_motion_compensation.cpp:195-294).
//
// Both are byte work bound by HBM (or by launch latency at these sizes): the
// sampler reads one luma byte of each frame per 16 pixels, the remap streams
// a frame through a 512-byte table held in LDS.
#include "vc1me_internal.h"

#include <algorithm>

namespace vc1me {
namespace {

// Sums over a 256-thread block of `count` 64-bit values, written by thread 0
// to out[0..count) (the block's partials; the host adds the blocks' rows).
template <int count>
__device__ void block_sum64(unsigned long long* out, const unsigned long long (&v)[count]) {
    __shared__ unsigned long long part[4][count];
    const int lane = static_cast<int>(threadIdx.x) & 63, wave = static_cast<int>(threadIdx.x) >> 6;
#pragma unroll
    for (int k = 0; k < count; ++k) {
        unsigned long long x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) part[wave][k] = x;
    }
    __syncthreads();
    for (int k = static_cast<int>(threadIdx.x); k < count; k += static_cast<int>(blockDim.x))
        out[k] = part[0][k] + part[1][k] + part[2][k] + part[3][k];
}

__device__ __forceinline__ int round_qpel_to_pixel(int q) { return q >= 0 ? (q + 2) / 4 : -((-q + 2) / 4); }

// One sample per (px, py) with px, py multiples of 4 inside the raster (the
// union over macroblocks of the reference's per-macroblock 4x4 lattice): the
// reference luma at the macroblock's vector rounded to whole pixels (clamped
// fetch) and the current luma.  Stores the pair (ref | cur << 8) and the
// block's sum(ref), sum(cur), sum(ref^2), sum(ref*cur) at sums[4 * block].
__global__ void __launch_bounds__(256) k_ic_samples(const uint8_t* __restrict__ cy, const uint8_t* __restrict__ ry,
                                                    int W, int H, const int2* __restrict__ mvs, int mbw,
                                                    uint16_t* __restrict__ samples, unsigned long long* sums) {
    const int cw = (W + 3) / 4, n = cw * ((H + 3) / 4);
    unsigned long long acc[4] = {0, 0, 0, 0};
    for (int idx = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); idx < n;
         idx += static_cast<int>(gridDim.x * blockDim.x)) {
        const int px = (idx % cw) * 4, py = (idx / cw) * 4;
        const int2 mv = mvs[(py >> 4) * mbw + (px >> 4)];
        const int rx = min(max(px + round_qpel_to_pixel(mv.x), 0), W - 1);
        const int ry_ = min(max(py + round_qpel_to_pixel(mv.y), 0), H - 1);
        const uint32_t r = ry[static_cast<size_t>(ry_) * W + rx];
        const uint32_t c = cy[static_cast<size_t>(py) * W + px];
        samples[idx] = static_cast<uint16_t>(r | (c << 8));
        acc[0] += r;
        acc[1] += c;
        acc[2] += r * r;
        acc[3] += r * c;
    }
    block_sum64<4>(sums + 4 * blockIdx.x, acc);
}

// Sampled SAD of every (LUMSCALE, LUMSHIFT) = (l0 + a, z0 + b), a < nl,
// b < nz (at most 7 x 7): |cur - map(ref)| with map(v) = clamp(((l + 32) * v
// + 32 >> 6) + z, 0, 255), which is intensity_map_y for l >= 1 (the shift
// term 64 z passes through the arithmetic shift unchanged).  The block's
// sums at sads[49 * block + a * 7 + b].
__global__ void __launch_bounds__(256) k_ic_sads(const uint16_t* __restrict__ samples, int n, int l0, int z0,
                                                 unsigned long long* sads) {
    uint32_t acc[7][7];
#pragma unroll
    for (int a = 0; a < 7; ++a)
#pragma unroll
        for (int b = 0; b < 7; ++b) acc[a][b] = 0;
    for (int idx = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); idx < n;
         idx += static_cast<int>(gridDim.x * blockDim.x)) {
        const uint32_t s = samples[idx];
        const int r = static_cast<int>(s & 255), c = static_cast<int>(s >> 8);
#pragma unroll
        for (int a = 0; a < 7; ++a) {
            const int base = ((l0 + a + 32) * r + 32) >> 6;
#pragma unroll
            for (int b = 0; b < 7; ++b) {
                const int m = min(max(base + z0 + b, 0), 255);
                acc[a][b] += static_cast<uint32_t>(abs(c - m));
            }
        }
    }
    // wave sums in 32 bits (a wave's partial is at most 64 x samples per
    // lane x 255), then the block's four waves in 64 bits
    __shared__ unsigned long long part[4][49];
    const int lane = static_cast<int>(threadIdx.x) & 63, wave = static_cast<int>(threadIdx.x) >> 6;
#pragma unroll
    for (int a = 0; a < 7; ++a)
#pragma unroll
        for (int b = 0; b < 7; ++b) {
            uint32_t x = acc[a][b];
            for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
            if (lane == 0) part[wave][a * 7 + b] = x;
        }
    __syncthreads();
    const int k = static_cast<int>(threadIdx.x);
    if (k < 49) sads[49 * blockIdx.x + k] = part[0][k] + part[1][k] + part[2][k] + part[3][k];
}

// dst = map(src) over `words16` 16-byte words of a frame buffer (luma then
// the two chroma planes, contiguous): bytes below luma_bytes through the
// luma table, the rest through the chroma table.  The tables are staged in
// LDS.
__global__ void __launch_bounds__(256) k_ic_remap(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                  size_t words16, size_t luma_bytes, IntensityLuts luts) {
    __shared__ uint32_t lut[128];
    if (threadIdx.x < 128) lut[threadIdx.x] = luts.w[threadIdx.x];
    __syncthreads();
    const uint8_t* tab = reinterpret_cast<const uint8_t*>(lut);
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < words16;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const uint4 v = src[i];
        uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int t = (i * 16 + k * 4 < luma_bytes) ? 0 : 256;
            const uint32_t x = w[k];
            w[k] = static_cast<uint32_t>(tab[t + (x & 255)]) | (static_cast<uint32_t>(tab[t + ((x >> 8) & 255)]) << 8) |
                   (static_cast<uint32_t>(tab[t + ((x >> 16) & 255)]) << 16) |
                   (static_cast<uint32_t>(tab[t + (x >> 24)]) << 24);
        }
        dst[i] = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

}  // namespace

int ic_sample_blocks(int n) { return std::max(1, std::min((n + 255) / 256, 1024)); }
int ic_sad_blocks(int n) { return std::max(1, std::min((n + 1023) / 1024, 256)); }

hipError_t launch_ic_samples(const uint8_t* cy, const uint8_t* ry, int W, int H, const int2* mvs, int mbw,
                             uint16_t* samples, unsigned long long* sums, hipStream_t s) {
    const int n = ((W + 3) / 4) * ((H + 3) / 4);
    hipLaunchKernelGGL(k_ic_samples, dim3(ic_sample_blocks(n)), dim3(256), 0, s, cy, ry, W, H, mvs, mbw, samples,
                       sums);
    return hipGetLastError();
}

hipError_t launch_ic_sads(const uint16_t* samples, int n, int l0, int z0, unsigned long long* sads, hipStream_t s) {
    hipLaunchKernelGGL(k_ic_sads, dim3(ic_sad_blocks(n)), dim3(256), 0, s, samples, n, l0, z0, sads);
    return hipGetLastError();
}

hipError_t launch_ic_remap(const uint8_t* src, uint8_t* dst, size_t bytes, size_t luma_bytes,
                           const IntensityLuts& luts, hipStream_t s) {
    const size_t words16 = (bytes + 15) / 16;
    const size_t blocks = std::min<size_t>((words16 + 255) / 256, 2048);
    hipLaunchKernelGGL(k_ic_remap, dim3(static_cast<unsigned>(std::max<size_t>(blocks, 1))), dim3(256), 0, s,
                       reinterpret_cast<const uint4*>(src), reinterpret_cast<uint4*>(dst), words16, luma_bytes,
                       luts);
    return hipGetLastError();
}

}  // namespace vc1me
