// Adjoint (reverse-mode) pass of the differentiable renderer on gfx950.
//
// The forward pass is the FP64 persistent tracer with its step log on
// (trace_fp64_persistent<false, true>): per ray, a checkpoint of the state
// every K committed steps, every committed step size and the transmittance
// before every step.  Here one lane takes one pixel and walks its chunks
// from the last to the first: it re-integrates the chunk's steps from the
// checkpoint with the logged step sizes (the same arithmetic as the
// forward pass, so the states are bit-identical), keeps them in its slice
// of a scratch buffer, then runs the chunk's steps backwards:
//   * escape: the sky term adds sky_int * Tr (piecewise constant in the
//     direction), so it only seeds the transmittance adjoint;
//   * crossing: reverse compositing (col += c Tr, Tr *= 1 - alpha), the
//     emission's dual-number pass over its 9 inputs, then the dense output
//     and its root;
//   * the DP5(4) update: six right-hand-side VJPs (kt_adjoint.hpp).
// At the ray's start the adjoint of (s0, pt, kt) is contracted with their
// dual-number tangents in spin, inclination and distance.  Each pixel's
// four gradient contributions go to grad_px; a deterministic reduction
// sums them.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kt_adjoint.hpp"
#include "kt_internal.hpp"
#include "kt_math.hpp"
#include "kt_tracer_common.hpp"

namespace kt {

// Per-lane scratch: element j of lane `tid` at sc[j * stride].  Step u of
// the segment being reversed holds its start state at j = 6 u; the state
// after the segment's last step sits at j = 6 K.  The stage inputs of a
// step are formed again from its start state when the step is reversed:
// five more right-hand sides per step instead of 30 more doubles of
// scratch traffic each way.
struct LaneScratch {
    double* p;
    uint32_t stride;
    __device__ __forceinline__ void put6(uint32_t j, const V6& v) const {
        p[(size_t)(j + 0) * stride] = v.x;
        p[(size_t)(j + 1) * stride] = v.y;
        p[(size_t)(j + 2) * stride] = v.z;
        p[(size_t)(j + 3) * stride] = v.px;
        p[(size_t)(j + 4) * stride] = v.py;
        p[(size_t)(j + 5) * stride] = v.pz;
    }
    __device__ __forceinline__ V6 get6(uint32_t j) const {
        return V6{p[(size_t)(j + 0) * stride], p[(size_t)(j + 1) * stride], p[(size_t)(j + 2) * stride],
                  p[(size_t)(j + 3) * stride], p[(size_t)(j + 4) * stride], p[(size_t)(j + 5) * stride]};
    }
};

// Adjoint of a step's equatorial crossing (kept out of line: crossings are
// about one step in a hundred).  Given the adjoint of the transmittance
// after the step, it returns the one before it and the direct adjoints of
// the step's start and end states, of k1, k3..k7 and of pt, kt, spin, mdot.
struct CrossAdj {
    V6 lk[6];
    V6 lam_s, lam_ns;
    double lamTr, lam_pt, lam_kt, g_a, g_m;
    bool lit;
};
template <bool MIXED>
static __device__ __forceinline__ CrossAdj crossing_adjoint(const RenderParams& P, const GradConsts& G, V6 sn, V6 ns,
                                                         V6 y3, V6 y4, V6 y5, V6 y6, double hs, double pt,
                                                         double kt, double Tr_n, double W0, double W1, double W2,
                                                         double lamTr) {
    CrossAdj o;
    o.lit = false;
    const Geo g = make_geo(P.a, P.aa, P.M, true);
    // the step's dense-output ingredients, as the forward pass formed them
    const V6 k1 = rhs_fwd<MIXED>(sn, pt, g), k3 = rhs_fwd<MIXED>(y3, pt, g), k4 = rhs_fwd<MIXED>(y4, pt, g);
    const V6 k5 = rhs_fwd<MIXED>(y5, pt, g), k6 = rhs_fwd<MIXED>(y6, pt, g), k7 = rhs_fwd<MIXED>(ns, pt, g);
#define KT_DS(c) dense_sum(k1.c, k3.c, k4.c, k5.c, k6.c, k7.c)
    const V6 dsum{KT_DS(x), KT_DS(y), KT_DS(z), KT_DS(px), KT_DS(py), KT_DS(pz)};
#undef KT_DS
    const Crossing hit = crossing_hit(sn, ns, k1, k7, dsum, hs, P.disk_root_mode);
    const ShadeD sv = shade_value(P, G, hit, pt, kt);
    if (!sv.lit) return o;
    // col += c Tr_n;  Tr_{n+1} = Tr_n (1 - alpha)
    const double c0 = sv.c[0].v, c1 = sv.c[1].v, c2 = sv.c[2].v, al = sv.alpha.v;
    const double lam_c[3] = {W0 * Tr_n, W1 * Tr_n, W2 * Tr_n};
    const double lam_alpha = -lamTr * Tr_n;
    o.lamTr = W0 * c0 + W1 * c1 + W2 * c2 + lamTr * (1.0 - al);
    o.lam_pt = o.lam_kt = o.g_a = o.g_m = 0.0;
    double lam_hit[5];
    shade_reverse(P, G, hit, pt, kt, lam_c, lam_alpha, lam_hit, o.lam_pt, o.lam_kt, o.g_a, o.g_m);
    for (int i = 0; i < 6; ++i) o.lk[i] = v6_zero();
    o.lam_s = v6_zero();
    o.lam_ns = v6_zero();
    crossing_reverse(sn, ns, k1, k7, dsum, hs, P.disk_root_mode, hit.t, lam_hit, o.lam_s, o.lam_ns, o.lk);
    o.lit = true;
    return o;
}

// One pixel's gradient contribution: out = d(sum_c w_c hdr_c)/d theta.
// MIXED: the forward pass was the mixed tracer (rhs32), whose arithmetic
// the re-integration repeats.  F32: the step reversals (right-hand-side
// VJPs, stage and state adjoints) in FP32; the re-integration, the crossing
// adjoint (dense-output root, emission) and the ray start stay FP64.
template <bool MIXED, bool F32>
__device__ __forceinline__ void ray_adjoint(const RenderParams& P, const GradConsts& G, int px, int py,
                                            const double w[3], const LaneScratch& sc, double out[4]) {
    out[0] = out[1] = out[2] = out[3] = 0.0;
    const GradLog& LG = P.log;
    const size_t pix = (size_t)py * (size_t)P.W + (size_t)px;
    const uint32_t N = LG.nsteps[pix];
    if (N == 0) return;  // no step logged (or the log overflowed)
    const uint32_t K = LG.K;
    const Geo g = make_geo(P.a, P.aa, P.M, true);
    const double pt = LG.pk[2 * pix], kt = LG.pk[2 * pix + 1];
    const uint8_t cls = P.cls ? P.cls[pix] : 0;
    // colour adjoint W = d loss / d colour
    const double W0 = w[0] * G.w_scale, W1 = w[1] * G.w_scale, W2 = w[2] * G.w_scale;
    // escape: col += sky_int * Tr_end on each channel where the grid is hit
    double lamTr = ((cls & 3) == 2 && (cls & 8)) ? P.sky_int * (W0 + W1 + W2) : 0.0;
    V6 lam = v6_zero();
    double lam_pt = 0.0, lam_kt = 0.0, g_a = 0.0, g_m = 0.0;
    V6f lamf = v6f_zero();  // F32: the state adjoint
    float lam_ptf = 0.f, g_af = 0.f;
    const float af = (float)P.a, a2f = (float)P.aa, Mf = (float)P.M, ptf = (float)pt;
    uint32_t chunk = LG.head[pix];
    if (chunk >= LG.capacity) return;
    uint32_t seg_start = ((N - 1) / K) * K;
    const uint32_t j_end = 6 * K;
    for (;;) {
        const double* ch = LG.arena + (size_t)chunk * LG.stride;
        const uint32_t len = min(K, N - seg_start);
        // re-integrate the segment from its checkpoint with the logged steps
        V6 s{ch[0], ch[1], ch[2], ch[3], ch[4], ch[5]};
        for (uint32_t u = 0; u < len; ++u) {
            sc.put6(6 * u, s);
            s = dp_step<MIXED>(s, ch[8 + u], pt, g, nullptr).ns;
        }
        sc.put6(j_end, s);
        for (int u = (int)len - 1; u >= 0; --u) {
            const V6 sn = sc.get6(6 * (uint32_t)u);
            const V6 ns = sc.get6((uint32_t)u + 1 < len ? 6 * ((uint32_t)u + 1) : j_end);
            const double hs = ch[8 + u];
            const double Tr_n = ch[8 + K + u];
            const double r_now = ks_r(ns.x, ns.y, ns.z, P.aa);
            const bool horizon = r_now < P.rh_eps;
            const bool crossing = !horizon && ((sn.y < 0.0 && ns.y >= 0.0) || (sn.y > 0.0 && ns.y <= 0.0));
            // the step's stage inputs, formed as the forward pass formed them
            // (FP32 reversal of a step without a crossing: formed again in
            // FP32, the points its FP32 VJPs linearise at)
            V6 Y[5];
            if (!F32 || crossing) (void)dp_step<MIXED>(sn, hs, pt, g, Y);
            V6 lk[6];
            V6 lam_s_dir = v6_zero();
            bool have_lk = false;
            if (crossing) {
                CrossAdj ca = crossing_adjoint<MIXED>(P, G, sn, ns, Y[1], Y[2], Y[3], Y[4], hs, pt, kt, Tr_n, W0, W1,
                                                      W2, lamTr);
                if (ca.lit) {
                    lamTr = ca.lamTr;
                    lam_pt += ca.lam_pt;
                    lam_kt += ca.lam_kt;
                    g_a += ca.g_a;
                    g_m += ca.g_m;
                    for (int i = 0; i < 6; ++i) lk[i] = ca.lk[i];
                    lam_s_dir = ca.lam_s;
                    if (F32) {
                        v6f_add(lamf, to_f(ca.lam_ns));
                    } else {
                        v6_add(lam, ca.lam_ns);
                    }
                    have_lk = true;
                }
            }
            if (F32) {
                V6f Yf[5], lkf[6];
                if (crossing) {
                    for (int i = 0; i < 5; ++i) Yf[i] = to_f(Y[i]);
                } else {
                    dp_stages32(to_f(sn), (float)hs, ptf, af, a2f, Mf, Yf);
                }
                if (have_lk)
                    for (int i = 0; i < 6; ++i) lkf[i] = to_f(lk[i]);
                lamf = dp_step_reverse32(to_f(sn), to_f(ns), Yf, (float)hs, ptf, af, a2f, Mf, lamf,
                                         have_lk ? lkf : nullptr, to_f(lam_s_dir), lam_ptf, g_af);
            } else {
                lam = dp_step_reverse(sn, ns, Y, hs, pt, g, lam, have_lk ? lk : nullptr, lam_s_dir, lam_pt, g_a);
            }
        }
        if (seg_start == 0) break;
        chunk = (uint32_t)__double_as_longlong(ch[6]);
        if (chunk >= LG.capacity) {  // a broken chain: no contribution
            out[0] = out[1] = out[2] = out[3] = 0.0;
            return;
        }
        seg_start -= K;
    }
    if (F32) {
        lam = to_d(lamf);
        lam_pt += (double)lam_ptf;
        g_a += (double)g_af;
    }
    // ray start: s0, pt, kt as functions of spin, inclination, distance
    for (int dir = 0; dir < 3; ++dir) {
        const StartD st = start_d(P, G, px, py, dir);
        double d = lam.x * st.s[0].d + lam.y * st.s[1].d + lam.z * st.s[2].d + lam.px * st.s[3].d +
                   lam.py * st.s[4].d + lam.pz * st.s[5].d + lam_pt * st.pt.d + lam_kt * st.kt.d;
        out[dir] += d;
    }
    out[0] += g_a;
    out[3] += g_m;
}

// Persistent adjoint grid: each wave claims one 8x8 pixel tile (64 work
// items) with one atomic; each lane reverses one ray.  Neighbouring rays
// cross the disk on the same steps, so tiles keep a wave's lanes together
// through the crossing path; the tiles are visited in descending order of
// their longest logged ray (grad_keys + a counting sort), longest first.
// One wave per 8x8 tile: key = min(longest logged ray of the tile, 255).
__global__ __launch_bounds__(256) void grad_keys(const uint32_t* __restrict__ nsteps, int W, int H, int tiles_x,
                                                 int n_tiles, uint8_t* __restrict__ keys) {
    const int lane = threadIdx.x & 63;
    const int tile = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (tile >= n_tiles) return;
    const int px = (tile % tiles_x) * 8 + (lane & 7), py = (tile / tiles_x) * 8 + (lane >> 3);
    uint32_t n = (px < W && py < H) ? nsteps[(size_t)py * W + px] : 0u;
    for (int off = 32; off > 0; off >>= 1) n = max(n, (uint32_t)__shfl_xor((int)n, off));
    if (lane == 0) keys[tile] = (uint8_t)min(n, 255u);
}

template <bool MIXED, bool F32>
__global__ __launch_bounds__(256) void grad_backward(const RenderParams* __restrict__ Ps, GradConsts G,
                                                     const float* __restrict__ dl_dhdr, double* __restrict__ grad_px,
                                                     double* __restrict__ scratch, uint32_t scratch_stride,
                                                     unsigned long long* __restrict__ queue, uint64_t total_items,
                                                     const uint32_t* __restrict__ order) {
    const RenderParams& P = Ps[0];
    const int lane = (int)(threadIdx.x & (kWave - 1));
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const LaneScratch sc{scratch + tid, scratch_stride};
    for (;;) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(queue, (unsigned long long)kWave);
        base = __shfl(base, 0);
        if (base >= total_items) break;
        const uint64_t item = base + (uint64_t)lane;
        const uint32_t tile = item < total_items ? order[item >> 6] : 0u;
        const int px = (int)(tile % (uint32_t)P.tiles_x) * 8 + (lane & 7);
        const int py = (int)(tile / (uint32_t)P.tiles_x) * 8 + (lane >> 3);
        if (item < total_items && px < P.W && py < P.H) {
            const size_t pix = (size_t)py * (size_t)P.W + (size_t)px;
            const float4 wv = *reinterpret_cast<const float4*>(dl_dhdr + 4 * pix);
            const double w[3] = {(double)wv.x, (double)wv.y, (double)wv.z};
            double out[4];
            ray_adjoint<MIXED, F32>(P, G, px, py, w, sc, out);
            double* o = grad_px + 4 * pix;
            o[0] = out[0];
            o[1] = out[1];
            o[2] = out[2];
            o[3] = out[3];
        }
    }
}

// Deterministic sum of the per-pixel contributions in two passes: block b
// of kReduceBlocks sums a fixed strided subset of pixels with a fixed tree
// into part[b]; one block then sums the parts the same way.
static constexpr int kReduceBlocks = 256;
__global__ __launch_bounds__(256) void grad_reduce_part(const double* __restrict__ grad_px, uint32_t n_pix,
                                                        double* __restrict__ part) {
    __shared__ double sh[4][256];
    const int t = threadIdx.x;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t i = blockIdx.x * 256u + (uint32_t)t; i < n_pix; i += kReduceBlocks * 256u)
        for (int k = 0; k < 4; ++k) acc[k] += grad_px[4 * (size_t)i + k];
    for (int k = 0; k < 4; ++k) sh[k][t] = acc[k];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off)
            for (int k = 0; k < 4; ++k) sh[k][t] += sh[k][t + off];
        __syncthreads();
    }
    if (t < 4) part[4 * blockIdx.x + t] = sh[t][0];
}
__global__ __launch_bounds__(256) void grad_reduce_final(const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double sh[4][256];
    const int t = threadIdx.x;
    for (int k = 0; k < 4; ++k) sh[k][t] = part[4 * t + k];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off)
            for (int k = 0; k < 4; ++k) sh[k][t] += sh[k][t + off];
        __syncthreads();
    }
    if (t < 4) out[t] = sh[t][0];
}

__global__ void zero_u64(unsigned long long* p) {
    if (threadIdx.x == 0) *p = 0ull;
}

__global__ void reset_log(uint32_t* pool_head, uint32_t* overflow, uint32_t first_pool_chunk) {
    if (threadIdx.x == 0) {
        *pool_head = first_pool_chunk;
        *overflow = 0u;
    }
}

hipError_t launch_reset_log(uint32_t* pool_head, uint32_t* overflow, uint32_t first_pool_chunk, hipStream_t stream) {
    hipLaunchKernelGGL(reset_log, dim3(1), dim3(64), 0, stream, pool_head, overflow, first_pool_chunk);
    return hipGetLastError();
}

hipError_t launch_sort_desc(const uint8_t* d_keys, int n, uint32_t* d_hist, uint32_t* d_perm, hipStream_t stream);

// keys/hist/order: one byte per 8x8 tile, the sort's counters, one word
// per tile; part: 4 * 256 doubles.
hipError_t launch_grad_backward(const RenderParams* d_params, const GradConsts& G, const float* dl_dhdr,
                                double* grad_px, double* scratch, uint32_t threads, unsigned long long* queue,
                                const uint32_t* nsteps, uint8_t* keys, uint32_t* hist, uint32_t* order, int W, int H,
                                uint32_t n_pix, double* part, double* grad_out, bool mixed_fwd, bool fp32,
                                hipStream_t stream) {
    const int tiles_x = (W + 7) / 8, n_tiles = tiles_x * ((H + 7) / 8);
    hipLaunchKernelGGL(grad_keys, dim3((n_tiles + 3) / 4), dim3(256), 0, stream, nsteps, W, H, tiles_x, n_tiles, keys);
    hipError_t e = launch_sort_desc(keys, n_tiles, hist, order, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, stream, queue);
    const dim3 grid(threads / 256), block(256);
    const uint64_t items = (uint64_t)n_tiles * 64;
    if (mixed_fwd && fp32)
        hipLaunchKernelGGL((grad_backward<true, true>), grid, block, 0, stream, d_params, G, dl_dhdr, grad_px,
                           scratch, threads, queue, items, order);
    else if (mixed_fwd)
        hipLaunchKernelGGL((grad_backward<true, false>), grid, block, 0, stream, d_params, G, dl_dhdr, grad_px,
                           scratch, threads, queue, items, order);
    else if (fp32)
        hipLaunchKernelGGL((grad_backward<false, true>), grid, block, 0, stream, d_params, G, dl_dhdr, grad_px,
                           scratch, threads, queue, items, order);
    else
        hipLaunchKernelGGL((grad_backward<false, false>), grid, block, 0, stream, d_params, G, dl_dhdr, grad_px,
                           scratch, threads, queue, items, order);
    hipLaunchKernelGGL(grad_reduce_part, dim3(kReduceBlocks), dim3(256), 0, stream, grad_px, n_pix, part);
    hipLaunchKernelGGL(grad_reduce_final, dim3(1), dim3(256), 0, stream, part, grad_out);
    return hipGetLastError();
}

}  // namespace kt
