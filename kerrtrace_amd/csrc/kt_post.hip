// Bloom and display transform for gfx950 (cuda/kernel.cu:17-84, 144-251).
//
// Two passes over the float4 HDR frame, both staged through LDS:
//  1. bloom_h: optional bright-pass max(c - threshold, 0), then the
//     horizontal truncated Gaussian with weights renormalised over the
//     in-bounds taps (kernel.cu:54-84, 144-173).  A block owns a row
//     segment and stages segment + 2R halo pixels (rgb) in LDS once; each
//     thread computes kHK adjacent output pixels.  The intermediate is
//     packed RGB (3 floats, 12 bytes per pixel: the reference's float4
//     alpha is never read), in the caller's float4-sized bloom buffer.
//  2. post_v: vertical pass of the same Gaussian over bloom_h's output,
//     hdr + intensity * blur, ACES fit, sRGB OETF, round-to-nearest-even
//     bytes packed as 0x00RRGGBB (kernel.cu:175-251).  A block owns a
//     64-column x TH-row output tile and stages (TH + 2R) rows of it in LDS,
//     so every halo row is read from HBM once per tile instead of once per
//     tap; each lane computes K rows of its column.
//
// Register tiles: a thread's K outputs share their taps' inputs.  Output k
// of the tile needs inputs k .. k + 2R, so one pass over the K + 2R inputs,
// kept in a K-entry rotating window, feeds all K accumulators: per tap one
// LDS read serves K outputs.  Each output still sums its taps in tap order.
// Pixels outside the image are staged as zeros, so they add exact zeros;
// the renormalising sum of in-bounds weights comes from a per-row and a
// per-column table built on the host (the same sequential float sums).
// Tap weights exp(-i^2 / 2 sigma^2) are computed once on the host and
// passed in a small device table.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

namespace kt {

static constexpr int kHThreads = 128;  // bloom_h: threads per block
static constexpr int kHK = 5;          // bloom_h: adjacent outputs per thread (odd: conflict-free LDS stride)
static constexpr int kColTile = 64;    // post_v: output columns per block (one per lane)

// acc[k] += in(k + t) * w[t] for t = 0 .. T-1 in increasing t, k = 0 .. K-1,
// where in(m) is input m of the tile (3 channels, read by `ld`).  Inputs
// pass through a K-entry window indexed modulo K; the tap loop is unrolled
// by K so every window index is a compile-time constant.
template <int K, typename Load>
__device__ __forceinline__ void fir_tile(const Load& ld, const float* __restrict__ w, int T, float (&acc)[K][3]) {
    float win[K][3];
#pragma unroll
    for (int m = 0; m < K - 1; ++m) ld(m, win[m]);
    int tb = 0;
    for (; tb + K <= T; tb += K) {
#pragma unroll
        for (int u = 0; u < K; ++u) {
            ld(tb + u + K - 1, win[(u + K - 1) % K]);
            const float wt = w[tb + u];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const float* v = win[(u + k) % K];
                acc[k][0] = fmaf(v[0], wt, acc[k][0]);
                acc[k][1] = fmaf(v[1], wt, acc[k][1]);
                acc[k][2] = fmaf(v[2], wt, acc[k][2]);
            }
        }
    }
#pragma unroll
    for (int u = 0; u < K; ++u) {
        if (tb + u < T) {
            ld(tb + u + K - 1, win[(u + K - 1) % K]);
            const float wt = w[tb + u];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const float* v = win[(u + k) % K];
                acc[k][0] = fmaf(v[0], wt, acc[k][0]);
                acc[k][1] = fmaf(v[1], wt, acc[k][1]);
                acc[k][2] = fmaf(v[2], wt, acc[k][2]);
            }
        }
    }
}

// ws_col: [W] sums of the in-bounds horizontal weights of each column.
__global__ __launch_bounds__(kHThreads) void bloom_h(const float4* __restrict__ hdr, float* __restrict__ bloom, int W,
                                                     int R, const float* __restrict__ wts,
                                                     const float* __restrict__ ws_col, float threshold) {
    extern __shared__ float lds[];  // (kHThreads * kHK + 2R) * 3 floats
    const int y = blockIdx.y;
    const int x0 = blockIdx.x * (kHThreads * kHK);
    const int span = kHThreads * kHK + 2 * R;
    const float4* row = hdr + (size_t)y * W;
    for (int i = threadIdx.x; i < span; i += kHThreads) {
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
    const int xb = x0 + (int)threadIdx.x * kHK;
    if (xb >= W) return;
    const float* base = lds + (size_t)threadIdx.x * kHK * 3;
    float acc[kHK][3];
#pragma unroll
    for (int k = 0; k < kHK; ++k) acc[k][0] = acc[k][1] = acc[k][2] = 0.f;
    fir_tile<kHK>(
        [&](int m, float* v) {
            v[0] = base[m * 3 + 0];
            v[1] = base[m * 3 + 1];
            v[2] = base[m * 3 + 2];
        },
        wts, 2 * R + 1, acc);
    float* out = bloom + (size_t)y * W * 3;
#pragma unroll
    for (int k = 0; k < kHK; ++k) {
        const int x = xb + k;
        if (x < W) {
            const float ws = ws_col[x];
            out[x * 3 + 0] = acc[k][0] / ws;
            out[x * 3 + 1] = acc[k][1] / ws;
            out[x * 3 + 2] = acc[k][2] / ws;
        }
    }
}

// cuda/kernel.cu:17-29 (Narkowicz ACES fit)
__device__ __forceinline__ float aces(float x) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    return (x * (a * x + b)) / (x * (c * x + d) + e);
}
// cuda/kernel.cu:30-35 (IEC 61966-2-1 OETF).  x^(1/2.4) as
// 2^(log2(x) / 2.4) from the hardware base-2 logarithm and exponential
// (1 ulp each): on (0.0031308, 1.04], the range of the ACES fit above the
// linear segment, the power is within 1e-6 relative, far below the 1/255
// step of the byte it is rounded to.
__device__ __forceinline__ float srgb_oetf(float x) {
    if (x <= 0.0031308f) return 12.92f * x;
    if (!(x <= 2.0f)) return 1.055f * powf(x, 0.4166666666666667f) - 0.055f;  // outside the ACES range
    return 1.055f * __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * 0.4166666666666667f) - 0.055f;
}
// cuda/kernel.cu:44-47: saturate, round half to even
__device__ __forceinline__ uint32_t to_byte(float v) {
    float s = fminf(fmaxf(v, 0.f), 1.f);
    if (v != v) s = 0.f;
    return (uint32_t)rintf(s * 255.f);
}
__device__ __forceinline__ uint32_t display(float fx, float fy, float fz) {
    uint32_t r8 = to_byte(srgb_oetf(aces(fx)));
    uint32_t g8 = to_byte(srgb_oetf(aces(fy)));
    uint32_t b8 = to_byte(srgb_oetf(aces(fz)));
    return (r8 << 16) | (g8 << 8) | b8;
}

// Block: nw = blockDim.x / 64 waves; wave w computes rows w K .. w K + K - 1
// of the block's TH = nw K rows.  ws_row: [H] sums of the in-bounds
// vertical weights of each row.
template <int K>
__global__ __launch_bounds__(512) void post_v(const float4* __restrict__ hdr, const float* __restrict__ bloom,
                                              uint32_t* __restrict__ out, int W, int H, int R,
                                              const float* __restrict__ wts, const float* __restrict__ ws_row,
                                              float intensity) {
    extern __shared__ float lds[];  // (TH + 2R) rows x kColTile x 3 floats
    const int tx = threadIdx.x & (kColTile - 1);
    const int wv = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    const int TH = nw * K;
    const int x = blockIdx.x * kColTile + tx;
    const int y0 = blockIdx.y * TH;
    const int rows = TH + 2 * R;
    for (int r = wv; r < rows; r += nw) {
        int sy = y0 - R + r;
        float v0 = 0.f, v1 = 0.f, v2 = 0.f;
        if (x < W && sy >= 0 && sy < H) {
            const float* b = bloom + ((size_t)sy * W + x) * 3;
            v0 = b[0];
            v1 = b[1];
            v2 = b[2];
        }
        float* d = lds + ((size_t)r * kColTile + tx) * 3;
        d[0] = v0;
        d[1] = v1;
        d[2] = v2;
    }
    __syncthreads();
    const int yb = y0 + wv * K;
    if (x >= W || yb >= H) return;
    const float* base = lds + ((size_t)wv * K * kColTile + tx) * 3;
    float acc[K][3];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k][0] = acc[k][1] = acc[k][2] = 0.f;
    fir_tile<K>(
        [&](int m, float* v) {
            const float* s = base + (size_t)m * kColTile * 3;
            v[0] = s[0];
            v[1] = s[1];
            v[2] = s[2];
        },
        wts, 2 * R + 1, acc);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int y = yb + k;
        if (y < H) {
            const float ws = ws_row[y];
            const float4 c = hdr[(size_t)y * W + x];
            out[(size_t)y * W + x] = display(c.x + (acc[k][0] / ws) * intensity, c.y + (acc[k][1] / ws) * intensity,
                                             c.z + (acc[k][2] / ws) * intensity);
        }
    }
}

// Display transform alone (bloom off): one pixel per thread.
__global__ __launch_bounds__(256) void post_plain(const float4* __restrict__ hdr, uint32_t* __restrict__ out,
                                                  size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 c = hdr[i];
    out[i] = display(c.x, c.y, c.z);
}

// LDS bytes of the two passes for radius R and tile height TH.
size_t bloom_h_lds_bytes(int R) { return (size_t)(kHThreads * kHK + 2 * R) * 3 * sizeof(float); }
size_t post_v_lds_bytes(int R, int TH) { return (size_t)(TH + 2 * R) * kColTile * 3 * sizeof(float); }

// Tile plan of post_v for radius R within max_lds bytes of LDS: rows per
// lane (8, else 1) and waves per block.  Returns false if no plan fits.
bool post_v_plan(int R, size_t max_lds, int* k, int* waves) {
    static const int plans[][2] = {{8, 8}, {8, 4}, {8, 2}, {8, 1}, {1, 8}, {1, 4}};
    for (const auto& p : plans) {
        if (post_v_lds_bytes(R, p[0] * p[1]) <= max_lds) {
            *k = p[0];
            *waves = p[1];
            return true;
        }
    }
    return false;
}

// ---- library-owned bloom intermediate (kt_post_process_noalloc)
// The horizontal pass's output lives here instead of in a caller buffer,
// for callers that cannot allocate.  bloom_h writes every pixel of the
// frame before post_v reads it, so nothing is cleared between calls.  In
// host code `scratch` names a host-side stand-in: kernels get the device
// address that post_scratch_address resolves per device.
static constexpr size_t kNoAllocMaxPixels = size_t(1) << 21;  // 2,097,152 px of packed RGB: 24 MiB
static __device__ __align__(256) float scratch[3 * kNoAllocMaxPixels];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static float* g_scratch_ptr[kMaxDevices] = {};

size_t noalloc_max_pixels() { return kNoAllocMaxPixels; }

// Device address of `scratch` on `device`, which must be the current
// device: resolved once per device with hipGetSymbolAddress (each device
// holds its own copy of the array), then reused.
hipError_t post_scratch_address(int device, float** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

hipError_t launch_bloom_h(const float* hdr, float* bloom, int W, int H, int R, const float* d_wts,
                          const float* d_ws_col, float threshold, hipStream_t stream) {
    dim3 grid((W + kHThreads * kHK - 1) / (kHThreads * kHK), H);
    hipLaunchKernelGGL(bloom_h, grid, dim3(kHThreads), bloom_h_lds_bytes(R), stream,
                       reinterpret_cast<const float4*>(hdr), bloom, W, R, d_wts, d_ws_col,
                       threshold);
    return hipGetLastError();
}

hipError_t launch_post_v(const float* hdr, const float* bloom, uint32_t* out, int W, int H, int R, int K, int waves,
                         const float* d_wts, const float* d_ws_row, float intensity, int bloom_on,
                         hipStream_t stream) {
    if (!bloom_on) {
        const size_t n = (size_t)W * H;
        hipLaunchKernelGGL(post_plain, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                           reinterpret_cast<const float4*>(hdr), out, n);
        return hipGetLastError();
    }
    const int TH = K * waves;
    dim3 grid((W + kColTile - 1) / kColTile, (H + TH - 1) / TH);
    const size_t lds = post_v_lds_bytes(R, TH);
    const float4* h4 = reinterpret_cast<const float4*>(hdr);
    if (K == 8)
        hipLaunchKernelGGL(post_v<8>, grid, dim3(64 * waves), lds, stream, h4, bloom, out, W, H, R, d_wts, d_ws_row,
                           intensity);
    else
        hipLaunchKernelGGL(post_v<1>, grid, dim3(64 * waves), lds, stream, h4, bloom, out, W, H, R, d_wts, d_ws_row,
                           intensity);
    return hipGetLastError();
}

}  // namespace kt
