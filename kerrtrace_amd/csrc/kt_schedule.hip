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
        const double x = P.cam[0], y = P.cam[1], z = P.cam[2];
        const double jx = y * n[2] - z * n[1], jy = z * n[0] - x * n[2], jz = x * n[1] - y * n[0];
        const double rc2 = x * x + y * y + z * z;
        const double a = P.a * P.inv_M;  // in units of M
        const double cos2 = y * y / rc2;
        const double lam = jy * P.inv_M;
        const double eta = (jx * jx + jz * jz) * (P.inv_M * P.inv_M) - a * a * cos2;
        const double rp = 1.0 + sqrt(fmax(1.0 - a * a, 0.0));
        const double c2 = eta + (lam - a) * (lam - a);
        const double r_lo = rp * 1.0001, span = 12.0 - rp;
        double best = 3.0e38, r_best = r_lo;
        const int NS = 32;
        for (int i = 0; i < NS; ++i) {
            const double t = (double)i / (double)(NS - 1);
            const double r = r_lo + span * t * t;
            const double A = r * r + a * a - a * lam;
            const double R = A * A - (r * r - 2.0 * r + a * a) * c2;
            if (R < best) {
                best = R;
                r_best = r;
            }
        }
        double r = r_best;
        for (int it = 0; it < 3; ++it) {
            const double A = r * r + a * a - a * lam;
            const double d1 = 4.0 * r * A - (2.0 * r - 2.0) * c2;
            const double d2 = 4.0 * A + 8.0 * r * r - 2.0 * c2;
            if (d2 > 0.0) r = fmin(fmax(r - d1 / d2, r_lo), 12.0);
        }
        const double A = r * r + a * a - a * lam;
        const double R = fmin(A * A - (r * r - 2.0 * r + a * a) * c2, best);
        key = (float)(-log(fabs(R) / (r * r * r * r) + 1e-12));
    }
    for (int off = 32; off > 0; off >>= 1) key = fmaxf(key, __shfl_xor(key, off));
    if (lane == 0) {
        // -log spans about [-6, 20]; 8 levels per unit of log
        int q = (int)((key + 10.0f) * 8.0f);
        keys[gtile] = (uint8_t)min(max(q, 0), kBuckets - 1);
    }
}

__global__ void key_histogram(const uint8_t* __restrict__ keys, int n, uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[kBuckets];
    for (int i = threadIdx.x; i < kBuckets; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        atomicAdd(&h[kBuckets - 1 - keys[i]], 1u);  // descending key order
    __syncthreads();
    for (int i = threadIdx.x; i < kBuckets; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

// Exclusive scan of the 256 bucket counts (one block of 256 threads).
__global__ void key_scan(uint32_t* __restrict__ hist, uint32_t* __restrict__ offsets) {
    __shared__ uint32_t s[kBuckets];
    const int t = threadIdx.x;
    s[t] = hist[t];
    __syncthreads();
    for (int d = 1; d < kBuckets; d <<= 1) {
        uint32_t v = t >= d ? s[t - d] : 0u;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    offsets[t] = s[t] - hist[t];
}

__global__ void key_scatter(const uint8_t* __restrict__ keys, int n, uint32_t* __restrict__ offsets,
                            uint32_t* __restrict__ perm) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t pos = atomicAdd(&offsets[kBuckets - 1 - keys[i]], 1u);
        perm[pos] = (uint32_t)i;
    }
}

__global__ void zero_u32(uint32_t* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Builds perm (n_tiles entries): tile indices in descending predicted cost.
// scratch: keys (n_tiles bytes, 4-byte aligned region) + 2 * 256 uint32.
hipError_t launch_cost_order(const RenderParams* d_params, int n_renders, int n_tiles, uint8_t* d_keys,
                             uint32_t* d_hist, uint32_t* d_perm, hipStream_t stream) {
    uint32_t* d_off = d_hist + kBuckets;
    hipLaunchKernelGGL(zero_u32, dim3(2), dim3(256), 0, stream, d_hist, 2 * kBuckets);
    hipLaunchKernelGGL(cost_keys, dim3((n_tiles + 3) / 4), dim3(256), 0, stream, d_params, n_renders, d_keys,
                       n_tiles);
    int blocks = min(1024, (n_tiles + 255) / 256);
    hipLaunchKernelGGL(key_histogram, dim3(blocks), dim3(256), 0, stream, d_keys, n_tiles, d_hist);
    hipLaunchKernelGGL(key_scan, dim3(1), dim3(kBuckets), 0, stream, d_hist, d_off);
    hipLaunchKernelGGL(key_scatter, dim3(blocks), dim3(256), 0, stream, d_keys, n_tiles, d_off, d_perm);
    return hipGetLastError();
}

}  // namespace kt
