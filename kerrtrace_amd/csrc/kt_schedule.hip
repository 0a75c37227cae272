// Hardest-first ordering of the tracer's work queue.
//
// A persistent tracer's frame time is bounded below by its slowest ray,
// whose DP5(4) steps run one after another.  Rays that graze the photon
// ring orbit many times and cost 10-40x the mean, so they are started
// first: each tile of the queue gets a cost key predicted from the physics
// of its rays, and the tiles are visited in descending key order.
//
// Predictor: with E = 1, a ray's conserved azimuthal angular momentum
// lambda = L_y and Carter constant eta are taken from its initial
// direction n at the camera x (flat-space values, error O(M/r_cam)):
//   J = x cross n,  lambda = J_y,  eta = |J|^2 - J_y^2 - a^2 cos^2(theta_cam).
// Its radial potential (Boyer-Lindquist, M = 1)
//   R(r) = (r^2 + a^2 - a lambda)^2 - (r^2 - 2r + a^2)(eta + (lambda - a)^2)
// has a double root exactly on the critical curve (spherical photon
// orbits).  Near it R has a local minimum close to zero between r_+ and
// the photon region: a ray just outside the curve turns back at a shallow
// forbidden region, one just inside spirals over a shallow allowed one,
// and either way it orbits ~ -log|R(r*)| times.  R is sampled on a
// quadratic grid in [r_+, 12] (dense near the horizon, where the
// prograde orbits of a rapidly spinning hole sit), the lowest sample is
// refined with Newton steps on R' = 0, and key = -log(|R(r*)| / r*^4).
// (The smallest |R| over samples would not do: |R| vanishes at every
// simple turning point too.)  The key only orders work, it never changes
// a ray's arithmetic, so frames are identical with or without it.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kt_internal.hpp"
#include "kt_math.hpp"

namespace kt {

static constexpr int kBuckets = 256;

// One wave per global tile (tile edge T <= 8: T*T <= 64 lanes); key = max
// over the tile's pixels, quantised to 8 bits.  Renders are sorted by
// tile_begin; a binary search finds a tile's render.
__global__ __launch_bounds__(256) void cost_keys(const RenderParams* __restrict__ Ps, int n_renders,
                                                 uint8_t* __restrict__ keys, int n_tiles) {
    const int lane = threadIdx.x & 63;
    const int gtile = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (gtile >= n_tiles) return;
    int lo = 0, hi = n_renders - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if ((int)Ps[mid].tile_begin <= gtile) lo = mid; else hi = mid - 1;
    }
    const RenderParams& P = Ps[lo];
    const int tile = gtile - (int)P.tile_begin;
    const int T = P.tile;
    const int tx = tile % P.tiles_x, ty = tile / P.tiles_x;
    const int px = tx * T + lane % T, py = ty * T + lane / T;
    float key = -3.0e38f;
    if (lane < T * T && px < P.W && py < P.H) {
        const double u = (((double)px + 0.5) * P.inv_w_2 - 1.0) * P.aspect * P.fov_scale;
        const double v = (1.0 - ((double)py + 0.5) * P.inv_h_2) * P.fov_scale;
        double n[3];
        for (int c = 0; c < 3; ++c) n[c] = u * P.rgt[c] + v * P.up[c] + P.fwd[c];
        const double inv = 1.0 / sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        n[0] *= inv;
        n[1] *= inv;
        n[2] *= inv;
        key = (float)critical_key(P.cam, n, P.a * P.inv_M, P.inv_M);
    }
    for (int off = 32; off > 0; off >>= 1) key = fmaxf(key, __shfl_xor(key, off));
    if (lane == 0) {
        // -log spans about [-6, 20]; 8 levels per unit of log
        int q = (int)((key + 10.0f) * 8.0f);
        keys[gtile] = (uint8_t)min(max(q, 0), kBuckets - 1);
    }
}

// Counting sort of the tile keys in descending order as a block-local
// multi-split: each block owns a contiguous chunk of tiles, counts its keys
// in LDS (key_histogram), one block turns the (bucket, block) counts into
// start offsets, bucket-major (key_scan), and each block places its tiles
// at its offsets plus an LDS rank (key_scatter).  No global atomics.
static constexpr int kSortChunk = 4096;  // tiles per block

__global__ __launch_bounds__(256) void key_histogram(const uint8_t* __restrict__ keys, int n,
                                                     uint32_t* __restrict__ block_hist) {
    __shared__ uint32_t h[kBuckets];
    for (int i = threadIdx.x; i < kBuckets; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const int begin = blockIdx.x * kSortChunk, end = min(n, begin + kSortChunk);
    for (int i = begin + (int)threadIdx.x; i < end; i += blockDim.x) atomicAdd(&h[kBuckets - 1 - keys[i]], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < kBuckets; i += blockDim.x) block_hist[(size_t)blockIdx.x * kBuckets + i] = h[i];
}

// One block of 256 threads: thread b walks bucket b over all blocks, then
// an exclusive scan over the bucket totals places the buckets.
__global__ __launch_bounds__(256) void key_scan(uint32_t* __restrict__ block_hist, int n_blocks) {
    __shared__ uint32_t tot[kBuckets];
    const int b = threadIdx.x;
    uint32_t sum = 0;
    for (int k = 0; k < n_blocks; ++k) {
        uint32_t c = block_hist[(size_t)k * kBuckets + b];
        block_hist[(size_t)k * kBuckets + b] = sum;  // offset within the bucket
        sum += c;
    }
    tot[b] = sum;
    __syncthreads();
    for (int d = 1; d < kBuckets; d <<= 1) {
        uint32_t v = b >= d ? tot[b - d] : 0u;
        __syncthreads();
        tot[b] += v;
        __syncthreads();
    }
    const uint32_t bucket_start = tot[b] - sum;
    for (int k = 0; k < n_blocks; ++k) block_hist[(size_t)k * kBuckets + b] += bucket_start;
}

__global__ __launch_bounds__(256) void key_scatter(const uint8_t* __restrict__ keys, int n,
                                                   const uint32_t* __restrict__ block_off,
                                                   uint32_t* __restrict__ perm) {
    __shared__ uint32_t off[kBuckets];
    for (int i = threadIdx.x; i < kBuckets; i += blockDim.x) off[i] = block_off[(size_t)blockIdx.x * kBuckets + i];
    __syncthreads();
    const int begin = blockIdx.x * kSortChunk, end = min(n, begin + kSortChunk);
    for (int i = begin + (int)threadIdx.x; i < end; i += blockDim.x) {
        const uint32_t pos = atomicAdd(&off[kBuckets - 1 - keys[i]], 1u);
        perm[pos] = (uint32_t)i;
    }
}

// Builds perm (n_tiles entries): tile indices in descending predicted cost.
// d_hist holds kBuckets counters per sort block (ceil(n_tiles / 4096)).
hipError_t launch_cost_order(const RenderParams* d_params, int n_renders, int n_tiles, uint8_t* d_keys,
                             uint32_t* d_hist, uint32_t* d_perm, hipStream_t stream) {
    const int n_blocks = (n_tiles + kSortChunk - 1) / kSortChunk;
    hipLaunchKernelGGL(cost_keys, dim3((n_tiles + 3) / 4), dim3(256), 0, stream, d_params, n_renders, d_keys,
                       n_tiles);
    hipLaunchKernelGGL(key_histogram, dim3(n_blocks), dim3(256), 0, stream, d_keys, n_tiles, d_hist);
    hipLaunchKernelGGL(key_scan, dim3(1), dim3(kBuckets), 0, stream, d_hist, n_blocks);
    hipLaunchKernelGGL(key_scatter, dim3(n_blocks), dim3(256), 0, stream, d_keys, n_tiles, d_hist, d_perm);
    return hipGetLastError();
}

// Order of n items by descending 8-bit key (the same multi-split).
hipError_t launch_sort_desc(const uint8_t* d_keys, int n, uint32_t* d_hist, uint32_t* d_perm, hipStream_t stream) {
    const int n_blocks = (n + kSortChunk - 1) / kSortChunk;
    hipLaunchKernelGGL(key_histogram, dim3(n_blocks), dim3(256), 0, stream, d_keys, n, d_hist);
    hipLaunchKernelGGL(key_scan, dim3(1), dim3(kBuckets), 0, stream, d_hist, n_blocks);
    hipLaunchKernelGGL(key_scatter, dim3(n_blocks), dim3(256), 0, stream, d_keys, n, d_hist, d_perm);
    return hipGetLastError();
}

// Counter words the tile order needs for n_tiles tiles.
size_t cost_order_hist_words(int n_tiles) {
    return (size_t)((n_tiles + kSortChunk - 1) / kSortChunk) * kBuckets;
}

}  // namespace kt
