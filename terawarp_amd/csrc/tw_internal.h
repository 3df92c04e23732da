// tw_internal.h -- shared definitions of libterawarp (HIP, gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <string>
#include <vector>

#include "../../include/terawarp.h"

typedef int64_t i64;

// Workspace slots cached on a context (grown on demand, freed at destroy).
enum TwSlot {
    WS_SRC = 0,     // staged source volume
    WS_DST,         // staged destination volume
    WS_GRID,        // staged / internal displacement grid
    WS_QTAB,        // FP16 grid block tables
    WS_SOLVE_A,     // TPS system matrix
    WS_SOLVE_B,     // TPS right-hand side / solution
    WS_SOLVE_AUX,   // pivots, landmark packs
    WS_SOLVE_AUX2,  // refinement vectors
    WS_SOLVE_L,     // the factor's off-diagonal panels (fused Cholesky steps)
    WS_PTS,         // landmark / coefficient packs for the grid kernel
    WS_DF,          // dense DF staging
    WS_WEIGHTS,     // B-spline weight tables
    WS_AUX_BOUNDS,  // grid layer bounds
    WS_COEF,        // device-side grid coefficients of the chained warp
    WS_BATCH_FIT,   // batch: landmark packs, fit jobs, fit states, coefficients
    WS_BATCH_SOLVE, // batch: per-system solve workspaces
    WS_BATCH_GRID,  // batch: grids (FP32 fields, FP16 planes, block tables)
    WS_BATCH_GTAB,  // batch: grid job tables
    WS_BATCH_RTAB,  // batch: resample job tables, queues, timestamps
    WS_SOLVE_FLAGS, // allocation-free mode: the pipelined solve's flags
    WS_COUNT
};

enum TwStage { ST_FIT = 0, ST_GRID, ST_QUANT, ST_RESAMPLE, ST_H2D, ST_D2H, ST_TOTAL, ST_COUNT };

// TW_MIXED fits factor projected systems of at least this order in FP32
// (MFMA trailing updates, FP64 refinement); smaller ones in FP64.  Measured
// on MI355X: the FP64 factor is faster up to 14000 landmarks (53.1 vs 55.8
// ms), even at 15000 (63.3 vs 63.5 ms), the FP32 one from 16000 (71.0 vs
// 74.5 ms); below, the block steps and the refinement's extra triangular
// solves are latency-bound.
constexpr int TW_FP32_FACTOR_MIN_DEFAULT = 15000;

// Bound of each wait between the pipelined solve's work-groups (µs).  A wait
// that runs out means other work held the CUs the solve's work-groups needed:
// the context then turns the pipelined solve off and the fit is redone with
// one launch per block.
constexpr int TW_SOLVE_WAIT_US_DEFAULT = 20000;

struct tw_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    bool timing = false;
    int tile = TW_TILE_AUTO;  // fused TPS resample tile (tw_set_tile)
    int tune_tz = 0;          // gather tile depth, single-volume launches (0: from the grid spacing)
    int tune_tz_batch = 0;    // the same, persistent launches
    int fp32_factor_min = TW_FP32_FACTOR_MIN_DEFAULT;  // TW_MIXED: FP32 factor from this order
    std::string err;
    void* ws[WS_COUNT] = {};
    size_t ws_cap[WS_COUNT] = {};
    hipEvent_t ev_begin[ST_COUNT] = {};
    hipEvent_t ev_end[ST_COUNT] = {};
    bool ev_used[ST_COUNT] = {};
    int refine_iters = 0;
    double fit_residual = 0.0;
    // Device properties used for launch shaping.
    int num_cus = 256;
    int pipe_cap64 = -1, pipe_cap32 = -1;  // resident work-groups of the pipelined solve (-1: not asked yet)
    int solve_wait_us = TW_SOLVE_WAIT_US_DEFAULT;  // bound of each wait of the pipelined solve
    int pipe_off = 0;  // a pipelined solve's wait ran out: solves take the per-block launches
    // B-spline weight tables on the device (WS_WEIGHTS) are for spacing
    // weights_f (0: none yet); their host image must outlive the copy.
    int weights_f = 0;
    int weights_basis = 0;
    // Per-block flags of the pipelined solve (2 * 128 ints, zeroed once) and
    // the epoch its calls count up.
    int* solve_flags = nullptr;
    unsigned solve_epoch = 0;
    std::vector<uint8_t> weights_host;
    // Pinned host staging for small uploads (landmark packs): pin_ev marks
    // the last upload from it, waited on before the buffer is refilled.
    // pin_state: a pinned mirror for reading a fit's state back.
    void* pin = nullptr;
    size_t pin_cap = 0;
    hipEvent_t pin_ev = nullptr;
    bool pin_busy = false;
    void* pin_state = nullptr;
    // Host images of a batch's tables, kept until the next batch (the batch
    // synchronises before it returns).
    std::vector<std::vector<uint8_t>> staging;
    // Allocation-free calls (tw_warp_tps_noalloc): while `arena` is set, every
    // workspace is carved from it (arena_cap bytes of static device memory,
    // bump-allocated from arena_off, one region per slot), host staging comes
    // from arena_host (static host memory) and nothing is allocated.  The
    // entry point sets and clears these around one call.
    uint8_t* arena = nullptr;
    size_t arena_cap = 0, arena_off = 0;
    void* arena_ws[WS_COUNT] = {};
    size_t arena_ws_cap[WS_COUNT] = {};
    uint8_t* arena_host = nullptr;
    size_t arena_host_cap = 0, arena_host_off = 0;
};

// Allocation-free mode: `bytes` of the context's device arena (256-byte
// aligned) or of its host staging arena; TW_E_NOMEM when the arena is full.
int tw_arena_carve(tw_ctx* ctx, size_t bytes, void** out);
int tw_arena_host(tw_ctx* ctx, size_t bytes, void** out);
// Zero `n` ints of device memory with a kernel on the context's stream.
int tw_fill_zero_i32(tw_ctx* ctx, int* p, int n);

// A zeroed host buffer of `bytes` that lives until tw_host_stage_reset.
uint8_t* tw_host_stage(tw_ctx* ctx, size_t bytes);
// Drop the staged host buffers (synchronises the stream first if any exist).
void tw_host_stage_reset(tw_ctx* ctx);

// A pinned host buffer of at least `bytes` to fill and upload from; waits
// for the previous upload from it first.  Call tw_pinned_issued after the
// upload is enqueued.
int tw_pinned(tw_ctx* ctx, size_t bytes, void** out);
int tw_pinned_issued(tw_ctx* ctx);
// Copy `bytes` (at most 4096) of device memory to `out` through the pinned
// state mirror and synchronise the stream.
int tw_read_small(tw_ctx* ctx, void* out, const void* dev, size_t bytes);

// Set the context error string and return code.
// Does a fused TPS resample run the LDS-streamed column tile?  TW_TILE_AUTO
// takes the per-tap gather tile, the faster of the two as measured on MI355X
// for the mixed form (the column tile ties it on the plain form).
inline bool tw_lds_tile(const tw_ctx* ctx, int interp) {
    return interp == TW_INTERP_LINEAR && ctx->tile == TW_TILE_LDS;
}

int tw_fail(tw_ctx* ctx, int code, const char* fmt, ...);

#define TW_HIP(ctx, call)                                                              \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return tw_fail((ctx), TW_E_DEVICE, "%s failed: %s (%s:%d)", #call,         \
                           hipGetErrorString(e_), __FILE__, __LINE__);                 \
    } while (0)

#define TW_LAUNCH_CHECK(ctx) TW_HIP(ctx, hipGetLastError())

// Grow-only device workspace of a context.
int tw_ws(tw_ctx* ctx, int slot, size_t bytes, void** out);

// Stage timing (no-ops unless timing is enabled on the context).
void tw_stage_begin(tw_ctx* ctx, int stage);
void tw_stage_end(tw_ctx* ctx, int stage);
void tw_stages_reset(tw_ctx* ctx);

// Make ctx's device current.
int tw_bind(tw_ctx* ctx);

inline int dtype_size(int dt) { return dt == TW_U8 ? 1 : (dt == TW_U16 ? 2 : 0); }

inline i64 vol_elems(const tw_volume* v) { return v->nx * v->ny * v->nz * v->nc; }

// Validate a volume descriptor; returns TW_OK or TW_E_ARG with message.
int tw_check_volume(tw_ctx* ctx, const tw_volume* v, const char* name);

// Bring a (possibly host) volume to the device: *dev points at device data.
int tw_stage_in(tw_ctx* ctx, const tw_volume* v, int slot, const void** dev);
// Device buffer for an output volume: v->data itself or a workspace slot.
int tw_stage_out_buffer(tw_ctx* ctx, const tw_volume* v, int slot, void** dev);
// Copy a staged output back to host memory (synchronises the stream).
int tw_stage_out_finish(tw_ctx* ctx, const tw_volume* v, const void* dev);

// Host-side FP64 cubic B-spline weights TB (n x 4), the reference's
// q_nonrigid_ini_bsplinebasis_3D T*B product (q_imgwarp_tps_quicksmallmemory.cpp:663-691).
void tw_bspline_tb(int n, double* TB);
// The n x 4 upsampling weights of a basis (TW_BASIS_BSPLINE: tw_bspline_tb;
// TW_BASIS_LINEAR: {0, 1 - t, t, 0}, t = i/n).
void tw_df_tb(int n, int basis, double* TB);

// Internal launchers shared between translation units.
int tw_grid_quantize_f16(tw_ctx* ctx, const float* df32, tw_grid* grid);

// Device-resident pieces of the chained warp (no host round trip, no
// stage reset).  A fit left on the device: the centres (float4), the
// grid-kernel coefficients (K x 3 FP64, columns 1..3 of c), the affine
// block d (4 x 4 FP64) and the fit state, checked later by
// tw_tps_fit_check (which synchronises the stream).
struct TwFitDev {
    const float4* pts;
    const double* cf;
    const double* d;
    void* state;
    int K;
    double bmax;  // largest |target| (scale of the fit residual)
};
int tw_tps_fit_dev(tw_ctx* ctx, const float* centres, const float* targets, int K, double lambda, int precision,
                   TwFitDev* out);
int tw_tps_fit_check(tw_ctx* ctx, const TwFitDev* fit);
int tw_grid_eval_dev(tw_ctx* ctx, const float4* pts, const double* cf, const double* d, int K, int precision,
                     tw_grid* grid);
// porigin (optional, xyz): the TW_EXACT sample position is the float sum
// (p - porigin + d) + porigin, the tera block warp's order with porigin the
// block's offset; null: p + d.
int tw_resample_tps_run(tw_ctx* ctx, const tw_volume* src, tw_volume* dst, const tw_grid* grid, int interp,
                        int precision, const i64* porigin = nullptr);
// Argument checks of the chained warp (tw_warp_tps and its allocation-free
// form); interp may carry TW_DF_LINEAR.
int tw_warp_tps_check(tw_ctx* ctx, const tw_volume* src, const tw_volume* dst, int f, int interp, int precision);
// The chained warp of tw_warp_tps on validated arguments (the allocation-free
// entry point runs it with the arena set).
int tw_warp_tps_chain(tw_ctx* ctx, const tw_volume* src, tw_volume* dst, const float* centres, const float* targets,
                      int K, int f, double lambda, int interp, int precision);
// Does a source window of nx*ny*nz voxels of `bpv` bytes fit the fused
// resampler's 32-bit tap offsets (else it runs in z-slabs)?
bool fits_fast(i64 nx, i64 ny, i64 nz, int bpv);

// ---- batches ---------------------------------------------------------------
// Fits of a batch: systems up to the small-system limit in one launch (one
// work-group each), larger ones by the single-system kernel sequence.  The
// batch fills each item's device centres (pts), grid-kernel coefficients
// (cf, K x 3), affine block (d, 4 x 4) and fit state, all in WS_BATCH_FIT.
struct TwFitItem {
    const float* centres;  // host, K x 3
    const float* targets;  // host, K x 3
    int K;
    double* cf;            // device out: grid-kernel coefficients (K x 3)
    double* d;             // device out: affine block (4 x 4)
    const float4* pts;     // device out: centres
    void* state;           // device out: fit state
    double bmax;           // out: residual scale
};
int tw_fit_batch(tw_ctx* ctx, TwFitItem* items, int n, double lambda, int precision);
// Check the states of a batch's fits (one device-to-host copy; the caller has
// synchronised).  status[i] = TW_OK or the item's error code; returns the
// first error (message names the item) or TW_OK.
int tw_fit_batch_check(tw_ctx* ctx, const TwFitItem* items, int n, int* status);

// Grids of a batch: grid.gx..gz, f, ox..oz, storage (and, for FP16, data =
// FP16 planes and qtab) describe the output; f32 holds the FP32 field.
struct TwGridItem {
    const float4* pts;
    const double* cf;
    const double* d;
    int K;
    tw_grid grid;
    float* f32;
};
int tw_grid_eval_batch(tw_ctx* ctx, const TwGridItem* items, int n);
