// Bloom and display transform for gfx950 (cuda/kernel.cu:17-84, 144-251).
//
// Two passes over the float4 HDR frame, both staged through LDS:
//  1. bloom_h: optional bright-pass max(c - threshold, 0), then the
//     horizontal truncated Gaussian with weights renormalised over the
//     in-bounds taps (kernel.cu:54-84, 144-173).  A block owns a 256-pixel
//     row segment and stages segment + 2R halo pixels (rgb) in LDS once.
//  2. post_v: vertical pass of the same Gaussian over bloom_h's output,
//     hdr + intensity * blur, ACES fit, sRGB OETF, round-to-nearest-even
//     bytes packed as 0x00RRGGBB (kernel.cu:175-251).  A block owns a
//     64-column x TH-row output tile and stages (TH + 2R) rows of it in LDS,
//     so every halo row is read from HBM once per tile instead of once per
//     tap.
// Tap weights exp(-i^2 / 2 sigma^2) are computed once on the host and
// passed in a small device table.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kt {

static constexpr int kRowSeg = 256;   // bloom_h: output pixels per block
static constexpr int kColTile = 64;   // post_v: output columns per block

__global__ __launch_bounds__(256) void bloom_h(const float4* __restrict__ hdr, float4* __restrict__ bloom, int W,
                                               int H, int R, const float* __restrict__ wts, float threshold,
                                               int enabled) {
    extern __shared__ float lds[];  // (kRowSeg + 2R) * 3 floats
    const int y = blockIdx.y;
    const int x0 = blockIdx.x * kRowSeg;
    const int span = kRowSeg + 2 * R;
    if (!enabled) {
        int x = x0 + (int)threadIdx.x;
        if (x < W) bloom[(size_t)y * W + x] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const float4* row = hdr + (size_t)y * W;
    for (int i = threadIdx.x; i < span; i += blockDim.x) {
        int sx = x0 - R + i;
        float3 v = make_float3(0.f, 0.f, 0.f);
        if (sx >= 0 && sx < W) {
            float4 c = row[sx];
            v = make_float3(c.x, c.y, c.z);
            if (threshold > 0.f) {
                v.x = fmaxf(v.x - threshold, 0.f);
                v.y = fmaxf(v.y - threshold, 0.f);
                v.z = fmaxf(v.z - threshold, 0.f);
            }
        }
        lds[i * 3 + 0] = v.x;
        lds[i * 3 + 1] = v.y;
        lds[i * 3 + 2] = v.z;
    }
    __syncthreads();
    const int x = x0 + (int)threadIdx.x;
    if (x >= W) return;
    float ax = 0.f, ay = 0.f, az = 0.f, ws = 0.f;
    const int lo = max(-R, -x), hi = min(R, W - 1 - x);
    for (int i = lo; i <= hi; ++i) {
        float w = wts[i + R];
        int li = (int)threadIdx.x + R + i;
        ax += lds[li * 3 + 0] * w;
        ay += lds[li * 3 + 1] * w;
        az += lds[li * 3 + 2] * w;
        ws += w;
    }
    if (ws > 0.f) {
        ax /= ws;
        ay /= ws;
        az /= ws;
    }
    bloom[(size_t)y * W + x] = make_float4(ax, ay, az, 0.f);
}

// cuda/kernel.cu:17-29 (Narkowicz ACES fit)
__device__ __forceinline__ float aces(float x) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    return (x * (a * x + b)) / (x * (c * x + d) + e);
}
// cuda/kernel.cu:30-35 (IEC 61966-2-1 OETF)
__device__ __forceinline__ float srgb_oetf(float x) {
    if (x <= 0.0031308f) return 12.92f * x;
    return 1.055f * powf(x, 0.4166666666666667f) - 0.055f;
}
// cuda/kernel.cu:44-47: saturate, round half to even
__device__ __forceinline__ uint32_t to_byte(float v) {
    float s = fminf(fmaxf(v, 0.f), 1.f);
    if (v != v) s = 0.f;
    return (uint32_t)rintf(s * 255.f);
}

__global__ __launch_bounds__(256) void post_v(const float4* __restrict__ hdr, const float4* __restrict__ bloom,
                                              uint32_t* __restrict__ out, int W, int H, int R, int TH,
                                              const float* __restrict__ wts, float intensity, int bloom_on) {
    extern __shared__ float lds[];  // (TH + 2R) rows x kColTile x 3 floats
    const int tx = threadIdx.x & (kColTile - 1);
    const int ty = threadIdx.x >> 6;  // 0..3
    const int x = blockIdx.x * kColTile + tx;
    const int y0 = blockIdx.y * TH;
    if (bloom_on) {
        const int rows = TH + 2 * R;
        for (int r = ty; r < rows; r += 4) {
            int sy = y0 - R + r;
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (x < W && sy >= 0 && sy < H) v = bloom[(size_t)sy * W + x];
            float* d = lds + ((size_t)r * kColTile + tx) * 3;
            d[0] = v.x;
            d[1] = v.y;
            d[2] = v.z;
        }
        __syncthreads();
    }
    if (x >= W) return;
    for (int oy = ty; oy < TH; oy += 4) {
        const int y = y0 + oy;
        if (y >= H) break;
        float4 base = hdr[(size_t)y * W + x];
        float fx = base.x, fy = base.y, fz = base.z;
        if (bloom_on) {
            float ax = 0.f, ay = 0.f, az = 0.f, ws = 0.f;
            const int lo = max(-R, -y), hi = min(R, H - 1 - y);
            for (int i = lo; i <= hi; ++i) {
                float w = wts[i + R];
                const float* s = lds + ((size_t)(oy + R + i) * kColTile + tx) * 3;
                ax += s[0] * w;
                ay += s[1] * w;
                az += s[2] * w;
                ws += w;
            }
            if (ws > 0.f) {
                ax /= ws;
                ay /= ws;
                az /= ws;
            }
            fx += ax * intensity;
            fy += ay * intensity;
            fz += az * intensity;
        }
        uint32_t r8 = to_byte(srgb_oetf(aces(fx)));
        uint32_t g8 = to_byte(srgb_oetf(aces(fy)));
        uint32_t b8 = to_byte(srgb_oetf(aces(fz)));
        out[(size_t)y * W + x] = (r8 << 16) | (g8 << 8) | b8;
    }
}

// LDS bytes of the two passes for radius R and tile height TH.
size_t bloom_h_lds_bytes(int R) { return (size_t)(kRowSeg + 2 * R) * 3 * sizeof(float); }
size_t post_v_lds_bytes(int R, int TH) { return (size_t)(TH + 2 * R) * kColTile * 3 * sizeof(float); }

hipError_t launch_bloom_h(const float* hdr, float* bloom, int W, int H, int R, const float* d_wts, float threshold,
                          int enabled, hipStream_t stream) {
    dim3 grid((W + kRowSeg - 1) / kRowSeg, H);
    size_t lds = enabled ? bloom_h_lds_bytes(R) : 0;
    hipLaunchKernelGGL(bloom_h, grid, dim3(256), lds, stream, reinterpret_cast<const float4*>(hdr),
                       reinterpret_cast<float4*>(bloom), W, H, R, d_wts, threshold, enabled);
    return hipGetLastError();
}

hipError_t launch_post_v(const float* hdr, const float* bloom, uint32_t* out, int W, int H, int R, int TH,
                         const float* d_wts, float intensity, int bloom_on, hipStream_t stream) {
    dim3 grid((W + kColTile - 1) / kColTile, (H + TH - 1) / TH);
    size_t lds = bloom_on ? post_v_lds_bytes(R, TH) : 0;
    hipLaunchKernelGGL(post_v, grid, dim3(256), lds, stream, reinterpret_cast<const float4*>(hdr),
                       reinterpret_cast<const float4*>(bloom), out, W, H, R, TH, d_wts, intensity, bloom_on);
    return hipGetLastError();
}

}  // namespace kt
