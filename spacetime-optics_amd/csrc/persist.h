/* persist.h - persistent multi-frame tracer (device code, included by bhrt.hip).

   Workgroups stay resident and pull rays from per-frame work queues (one atomic counter per frame;
   queue entries are the rays of 8 x 8 pixel tiles, so the rays a wave takes at once are image
   neighbours). The tiles of a frame are queued in row order or in a given order (the launcher sorts
   them by their centre ray's impact parameter, or by its predicted length, longest first, so the
   rays that orbit the hole longest start first). Every `interval` integration steps a wave retires finished lanes (one ballot),
   writes their pixels, and refills the free lanes with new rays when at least `refill_min` lanes are
   free: lane 0 claims a contiguous range with one atomic add and each free lane takes the entry at
   its rank among the free lanes. A wave refills only from its current frame and moves to the next
   frame once its lanes are drained, so the frame parameters stay wave-uniform (scalar loads).
   Every wave exits when all frames are exhausted and its lanes are idle; each ray runs at most
   max_steps steps, so the grid always drains. With prio_age > 0 a wave that holds a ray older than
   prio_age steps raises its issue priority, so the longest rays run at their own latency instead
   of sharing the SIMD evenly with younger waves. What an entry means (which pixel and sample, where
   the finished ray is written) is a Work policy: frame batches, the tile interface's ray records,
   and the two passes of adaptive sampling. */
#pragma once
#include "../../include/bhrt.h"
#include "trace_fp64.h"
#include "trace_mixed.h"

namespace bhp {

constexpr int MAX_BLOCK = 256;   // largest workgroup of the persistent kernel
constexpr int NSTAT = 5;         // per-ray statistics summed per lane (steps, accepted, flow, escaped, FP64)

using bhd::RayResult;
using bhd::TraceParams;

struct BatchFrame {
    TraceParams P;         // scene + frame; P.w/P.h = frame size, P.spp = 1
    bhm32::MixedParams M;  // the frame's mixed-precision policy and emission constants
    unsigned int nentries; // queue entries of the frame
    int tiles_x;
    float4* out;           // W*H pixels (frame batches)
    unsigned int* order;   // queue position -> tile (null: row order)
};

__device__ __forceinline__ unsigned int queued_tile(const BatchFrame& F, unsigned int pos) {
    return F.order ? F.order[pos] : pos;
}

__device__ __forceinline__ void tile_pixel(int tiles_x, unsigned int tile, unsigned int within, int* px, int* py) {
    *px = (int)(tile % (unsigned)tiles_x) * 8 + (int)(within & 7u);
    *py = (int)(tile / (unsigned)tiles_x) * 8 + (int)(within >> 3);
}

/* Work mappings: entry e of frame f -> (pixel, sample index) and where a finished ray goes.
   Entries of 64 consecutive indices are 8 x 8 pixel tiles, so a wave's rays are image neighbours. */

// Frame batches: one pixel-centre (or jittered sample 0) ray per pixel, written as a float4 pixel.
struct BatchWork {
    const BatchFrame* frames;
    int n;
    __device__ __forceinline__ bool begin(int f, unsigned int e, int* px, int* py, int* s) const {
        tile_pixel(frames[f].tiles_x, queued_tile(frames[f], e >> 6), e & 63u, px, py);
        *s = 0;
        return *px < frames[f].P.w && *py < frames[f].P.h;
    }
    __device__ __forceinline__ void finish(int f, unsigned int e, const RayResult& r) const {
        int px, py;
        tile_pixel(frames[f].tiles_x, queued_tile(frames[f], e >> 6), e & 63u, &px, &py);
        frames[f].out[(size_t)py * frames[f].P.w + px] =
            make_float4((float)r.I[0], (float)r.I[1], (float)r.I[2], (float)r.I[3]);
    }
};

// Every per-ray output of a finished ray as the tile interface's record.
__device__ __forceinline__ void write_record(bh_ray_out* __restrict__ q, const RayResult& r) {
    for (int i = 0; i < 6; ++i) q->y[i] = r.y[i];
    q->th_inf = r.th_inf;
    q->ph_inf = r.ph_inf;
    q->I[0] = r.I[0]; q->I[1] = r.I[1]; q->I[2] = r.I[2]; q->I[3] = r.I[3];
    q->trans = r.trans;
    q->psi = r.psi;
    q->fate = r.fate;
    q->ncross = r.ncross;
    q->order = r.order;
    q->steps = r.steps;
    q->accepted = r.accepted;
    q->flow_steps = r.flow_steps;
    q->sky_u = r.sky_u;
    q->sky_v = r.sky_v;
}

// Ray records of a w x h region at (x0, y0) of a frame (the tile interface): P.w / P.h are the
// region's size and P.x0 / P.y0 its origin; one ray per pixel (centre, or jittered sample 0), its
// record at out[(py - y0) * w + (px - x0)].
struct RecordWork {
    const BatchFrame* frames;   // one frame
    int n;
    bh_ray_out* out;
    __device__ __forceinline__ bool begin(int f, unsigned int e, int* px, int* py, int* s) const {
        const TraceParams& P = frames[f].P;
        int lx, ly;
        tile_pixel(frames[f].tiles_x, queued_tile(frames[f], e >> 6), e & 63u, &lx, &ly);
        *px = P.x0 + lx;
        *py = P.y0 + ly;
        *s = 0;
        return lx < P.w && ly < P.h;
    }
    __device__ __forceinline__ void finish(int f, unsigned int e, const RayResult& r) const {
        int lx, ly;
        tile_pixel(frames[f].tiles_x, queued_tile(frames[f], e >> 6), e & 63u, &lx, &ly);
        write_record(out + (size_t)ly * frames[f].P.w + lx, r);
    }
};

// First pass of adaptive sampling: samples 0..spp-1 of every pixel (entry = tile, sample, pixel in
// tile), each kept in FP64 at out[pixel * spp + sample].
struct SampleWork {
    const BatchFrame* frames;   // one frame
    int n;
    int spp;
    double4* out;
    __device__ __forceinline__ bool begin(int f, unsigned int e, int* px, int* py, int* s) const {
        const unsigned int per_tile = 64u * (unsigned)spp, rem = e % per_tile;
        tile_pixel(frames[f].tiles_x, queued_tile(frames[f], e / per_tile), rem & 63u, px, py);
        *s = (int)(rem >> 6);
        return *px < frames[f].P.w && *py < frames[f].P.h;
    }
    __device__ __forceinline__ void finish(int f, unsigned int e, const RayResult& r) const {
        const unsigned int per_tile = 64u * (unsigned)spp, rem = e % per_tile;
        int px, py;
        tile_pixel(frames[f].tiles_x, queued_tile(frames[f], e / per_tile), rem & 63u, &px, &py);
        out[((size_t)py * frames[f].P.w + px) * spp + (rem >> 6)] = make_double4(r.I[0], r.I[1], r.I[2], r.I[3]);
    }
};

// Second pass: chunk g of the exclusive prefix sum `offsets` of per-pixel chunk counts belongs to
// the last pixel p with offsets[p] <= g; its 8 samples continue p's sequence at
// spp + 8 * (g - offsets[p]). Each sample is kept in FP64 at out[entry].
struct ChunkWork {
    const BatchFrame* frames;   // one frame
    int n;
    int spp;
    int npix;
    const unsigned int* offsets;
    double4* out;
    __device__ __forceinline__ bool begin(int f, unsigned int e, int* px, int* py, int* s) const {
        const unsigned int g = e >> 3;
        int lo = 0, hi = npix;   // upper bound of g in offsets
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (offsets[mid] <= g) lo = mid + 1;
            else hi = mid;
        }
        const int p = lo - 1;
        const int W = frames[f].P.w;
        *px = p % W;
        *py = p / W;
        *s = spp + 8 * (int)(g - offsets[p]) + (int)(e & 7u);
        return true;
    }
    __device__ __forceinline__ void finish(int, unsigned int e, const RayResult& r) const {
        out[e] = make_double4(r.I[0], r.I[1], r.I[2], r.I[3]);
    }
};

template <int PREC, int RADIO>
struct Tracer;

template <int RADIO>
struct Tracer<BH_PREC_FP64, RADIO> {
    using State = bhd::RayState;
    static __device__ __forceinline__ void init(const TraceParams& P, const bhm32::MixedParams&, double x, double y,
                                                State& R) {
        bhd::ray_init<RADIO>(P, x, y, R);
    }
    static __device__ __forceinline__ void step(const TraceParams& P, const bhm32::MixedParams&, State& R) {
        bhd::ray_step<RADIO>(P, R);
    }
    static __device__ __forceinline__ int result(const TraceParams& P, const State& R, RayResult& out) {
        bhd::ray_result(P, R, out);
        return out.steps;
    }
};

template <int RADIO>
struct Tracer<BH_PREC_MIXED, RADIO> {
    using State = bhm32::RayStateM;
    static __device__ __forceinline__ void init(const TraceParams& P, const bhm32::MixedParams& M, double x, double y,
                                                State& R) {
        bhm32::ray_init_m<RADIO>(P, M, x, y, R);
    }
    static __device__ __forceinline__ void step(const TraceParams& P, const bhm32::MixedParams& M, State& R) {
        bhm32::ray_step_m<RADIO>(P, M, R);
    }
    static __device__ __forceinline__ int result(const TraceParams& P, const State& R, RayResult& out) {
        return bhm32::ray_result_m(P, R, out);
    }
};

template <int PREC, int RADIO, class Work>
__device__ __forceinline__ void persist_body(const Work& Wk, unsigned int* __restrict__ queue, int interval,
                                             int refill_min, int prio_age, double* flux_out,
                                             unsigned long long* cnt_out) {
    using T = Tracer<PREC, RADIO>;
    const BatchFrame* __restrict__ frames = Wk.frames;
    const int nframes = Wk.n;
    const int lane = threadIdx.x & 63;
    int cur = 0;   // the wave's frame (wave-uniform)
    bool active = false;
    unsigned int ent = 0;   // the lane's queue entry (the work mapping recovers pixel and sample)
    typename T::State R;
    // per-lane sums over the lane's finished rays, in the lane's own LDS slots (updated once per ray,
    // so they need no registers during the integration)
    __shared__ double s_flux[MAX_BLOCK];
    __shared__ unsigned long long s_stat[NSTAT][MAX_BLOCK];
    const int tid = threadIdx.x;
    s_flux[tid] = 0.0;
    for (int i = 0; i < NSTAT; ++i) s_stat[i][tid] = 0ull;
    for (;;) {
        // retire-and-refill
        const unsigned long long freem = __ballot(!active);
        const int nfree = __popcll(freem);
        if (nfree >= refill_min || nfree == 64) {
            while (cur < nframes) {
                const unsigned int n = frames[cur].nentries;
                unsigned int base = 0;
                if (lane == 0) {
                    // read before adding, so an exhausted counter grows by at most one claim per wave
                    base = __hip_atomic_load(&queue[cur], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (base < n) base = atomicAdd(&queue[cur], (unsigned int)nfree);
                }
                base = __builtin_amdgcn_readfirstlane(__shfl(base, 0, 64));
                if (base < n) {
                    if (!active) {
                        const unsigned int rank = (unsigned int)__popcll(freem & ((1ull << lane) - 1ull));
                        const unsigned int e = base + rank;
                        int px, py, s = 0;
                        if (e < n && Wk.begin(cur, e, &px, &py, &s)) {
                            const TraceParams& P = frames[cur].P;
                            double ox, oy;
                            bhd::sample_offsets(P, px, py, s, &ox, &oy);
                            T::init(P, frames[cur].M, ox, oy, R);
                            ent = e;
                            active = true;
                        }
                    }
                    break;
                }
                if (nfree == 64) {
                    ++cur;   // this frame's queue is empty and the wave is idle: next frame
                    continue;
                }
                break;   // tail of the frame: keep stepping the live lanes
            }
        }
        if (__ballot(active) == 0ull && cur >= nframes) break;
        if (prio_age > 0) {   // wave-uniform: any lane with an old ray lifts the wave
            if (__ballot(active && R.it >= prio_age)) __builtin_amdgcn_s_setprio(2);
            else __builtin_amdgcn_s_setprio(0);
        }
        const int cf = __builtin_amdgcn_readfirstlane(cur < nframes ? cur : nframes - 1);
        // frame parameters through the constant address space: wave-uniform scalar loads
        using CFrame = const __attribute__((address_space(4))) BatchFrame;
        CFrame* cfr = (CFrame*)frames;
        const TraceParams& P = *(const TraceParams*)&cfr[cf].P;
        const bhm32::MixedParams& M = *(const bhm32::MixedParams*)&cfr[cf].M;
        for (int k = 0; k < interval; ++k) {
            if (active) {
                if (!R.done) T::step(P, M, R);
                if (R.done) {
                    RayResult r;
                    const int f64 = T::result(P, R, r);
                    Wk.finish(cf, ent, r);
                    s_flux[tid] += r.I[0];
                    s_stat[0][tid] += (unsigned long long)r.steps;
                    s_stat[1][tid] += (unsigned long long)r.accepted;
                    s_stat[2][tid] += (unsigned long long)r.flow_steps;
                    s_stat[3][tid] += r.fate == 0 ? 1ull : 0ull;
                    s_stat[4][tid] += (unsigned long long)f64;
                    active = false;
                }
            }
        }
    }
    if (prio_age > 0) __builtin_amdgcn_s_setprio(0);
    *flux_out = s_flux[tid];
    for (int i = 0; i < NSTAT; ++i) cnt_out[i] = s_stat[i][tid];
}

}  // namespace bhp
