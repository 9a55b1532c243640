/* ctx.h - the device context behind the C ABI (host-side definition shared by bhrt.hip and mlp.hip). */
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <string>
#include <utility>
#include <vector>

#include "../../include/bhrt.h"
#include "mlp.h"
#include "trace_fp64.h"

constexpr int NCNT = 5;   // counters per block: attempts, accepted, flow, limited, fp64 attempts

struct BlockOut {
    double flux;
    unsigned long long cnt[NCNT];
};


struct bh_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bh_scene S{};
    bh_settings rs{};
    float* d_k2 = nullptr;
    std::vector<float> k2_host;   // host copy of the K2 table (emission-radius estimate)
    std::vector<std::pair<std::array<double, 10>, double>> emit_cache;   // flow model -> emission radius
    int k2_n = 0;
    float k2_l0 = 0, k2_dl = 0;
    float* d_turb = nullptr;
    int turb_n = 0, turb_nt = 0;
    double turb_half = 0, turb_t0 = 0, turb_dt = 0;
    // sheared turbulence: slices baked on the device for the current scene and camera
    int turb_model = BH_TURB_GRID;
    int ts_dims[4] = {512, 192, 48, 24};
    float* d_ts = nullptr;
    size_t ts_cap = 0;
    bool ts_valid = false;
    std::array<double, 16> ts_key{};
    bhd::ShearedView* d_tsv = nullptr;   // the view the kernels read (TraceParams.ts)
    size_t tsv_cap = 0;
    bhd::ShearedView tsv_host{};
    float4* d_px = nullptr;
    size_t px_cap = 0;
    bh_ray_out* d_rays = nullptr;
    size_t rays_cap = 0;
    BlockOut* d_blk = nullptr;
    size_t blk_cap = 0;
    BlockOut* d_tot = nullptr;
    double* d_ph = nullptr;
    int* d_fates = nullptr;
    size_t ph_cap = 0;
    // surrogate: FP32 master parameters, Adam moments, normalisation and the baked table
    float* d_mlp = nullptr;
    float* d_adam_m = nullptr;
    float* d_adam_v = nullptr;
    bhmlp::Norm mlp_norm{};
    bool mlp_trained = false;
    float2* d_stab = nullptr;
    unsigned char* d_sbad = nullptr;
    bhmlp::SurrogateTable stab{};
    bool stab_ready = false;
    // persistent batches: frame descriptors, per-frame queues, output, cached graph
    unsigned char* d_bframes = nullptr;   // bhp::BatchFrame[]
    size_t bframes_cap = 0;               // bytes
    unsigned int* d_queue = nullptr;
    size_t queue_cap = 0;
    unsigned int* d_order = nullptr;      // queue position -> tile, all frames of a launch
    size_t order_cap = 0;
    unsigned int* d_ohist = nullptr;      // per-frame bin counters of the tile sort
    size_t ohist_cap = 0;
    // radio post: pyramid levels, scratch level / tone-map input, Gaussian taps, 8-bit output (bytes)
    void* d_post_lev = nullptr;
    size_t post_lev_cap = 0;
    void* d_post_tmp = nullptr;
    size_t post_tmp_cap = 0;
    void* d_post_wt = nullptr;
    size_t post_wt_cap = 0;
    void* d_post_rgb = nullptr;
    size_t post_rgb_cap = 0;
    void* d_hud = nullptr;                // HUD rectangles, glyphs and font atlas of bh_tonemap_hud
    size_t hud_cap = 0;
    float4* d_bout = nullptr;
    size_t bout_cap = 0;
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_def = nullptr;
    unsigned long long graph_key = 0;
    // adaptive sampling: per-sample pass-1 values, per-pixel sums and statistics, decisions, pass 2
    double4* d_samp1 = nullptr;
    size_t samp1_cap = 0;
    double4* d_acc = nullptr;
    size_t acc_cap = 0;
    float2* d_ystat = nullptr;
    size_t ystat_cap = 0;
    unsigned int *d_extra = nullptr, *d_chunks = nullptr, *d_offsets = nullptr, *d_bsum = nullptr, *d_nsamp = nullptr;
    size_t extra_cap = 0, chunks_cap = 0, offsets_cap = 0, bsum_cap = 0, nsamp_cap = 0;
    double* d_fpart = nullptr;
    size_t fpart_cap = 0;
    double4* d_samp2 = nullptr;
    size_t samp2_cap = 0;
    int n_cu = 0;
    bh_batch_cfg tuned[2]{};   // last bh_autotune result per precision (used when no cfg is given)
    bool tuned_ok[2]{};
    std::string err;
};
