// tw_context.hip -- context, workspaces, staging and stage timing.
#include "tw_internal.h"

#include <cmath>
#include <cstring>

int tw_fail(tw_ctx* ctx, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    else fprintf(stderr, "terawarp: %s\n", buf);
    return code;
}

int tw_bind(tw_ctx* ctx) {
    TW_HIP(ctx, hipSetDevice(ctx->device));
    return TW_OK;
}

uint8_t* tw_host_stage(tw_ctx* ctx, size_t bytes) {
    ctx->staging.emplace_back(bytes ? bytes : 1, (uint8_t)0);
    return ctx->staging.back().data();
}

void tw_host_stage_reset(tw_ctx* ctx) {
    if (ctx->staging.empty()) return;
    (void)hipStreamSynchronize(ctx->stream);
    ctx->staging.clear();
}

namespace {
__global__ void k_fill_zero_i32(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}
}  // namespace

int tw_fill_zero_i32(tw_ctx* ctx, int* p, int n) {
    if (n <= 0) return TW_OK;
    hipLaunchKernelGGL(k_fill_zero_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, p, n);
    TW_LAUNCH_CHECK(ctx);
    return TW_OK;
}

int tw_arena_carve(tw_ctx* ctx, size_t bytes, void** out) {
    const size_t need = (bytes + 255) & ~size_t(255);
    if (need > ctx->arena_cap - ctx->arena_off)
        return tw_fail(ctx, TW_E_NOMEM,
                       "the allocation-free entry point's %zu MiB device scratch is full (%zu bytes more needed)",
                       ctx->arena_cap >> 20, need - (ctx->arena_cap - ctx->arena_off));
    *out = ctx->arena + ctx->arena_off;
    ctx->arena_off += need;
    return TW_OK;
}

int tw_arena_host(tw_ctx* ctx, size_t bytes, void** out) {
    const size_t need = (bytes + 255) & ~size_t(255);
    if (need > ctx->arena_host_cap - ctx->arena_host_off)
        return tw_fail(ctx, TW_E_NOMEM, "the allocation-free entry point's %zu KiB host staging is full",
                       ctx->arena_host_cap >> 10);
    *out = ctx->arena_host + ctx->arena_host_off;
    ctx->arena_host_off += need;
    return TW_OK;
}

int tw_ws(tw_ctx* ctx, int slot, size_t bytes, void** out) {
    if (bytes == 0) bytes = 256;
    if (ctx->arena) {
        if (!ctx->arena_ws[slot] || ctx->arena_ws_cap[slot] < bytes) {
            void* p = nullptr;
            if (int rc = tw_arena_carve(ctx, bytes, &p)) return rc;
            ctx->arena_ws[slot] = p;
            ctx->arena_ws_cap[slot] = bytes;
        }
        *out = ctx->arena_ws[slot];
        return TW_OK;
    }
    if (ctx->ws_cap[slot] < bytes) {
        if (ctx->ws[slot]) {
            TW_HIP(ctx, hipStreamSynchronize(ctx->stream));
            TW_HIP(ctx, hipFree(ctx->ws[slot]));
            ctx->ws[slot] = nullptr;
            ctx->ws_cap[slot] = 0;
        }
        size_t cap = (bytes + 0xFFFF) & ~size_t(0xFFFF);
        hipError_t e = hipMalloc(&ctx->ws[slot], cap);
        if (e != hipSuccess) {
            ctx->ws[slot] = nullptr;
            (void)hipGetLastError();
            return tw_fail(ctx, TW_E_NOMEM, "device allocation of %zu bytes failed: %s", cap,
                           hipGetErrorString(e));
        }
        ctx->ws_cap[slot] = cap;
    }
    *out = ctx->ws[slot];
    return TW_OK;
}

int tw_pinned(tw_ctx* ctx, size_t bytes, void** out) {
    // Allocation-free mode: a fresh region of the static host staging per
    // request (regions are not reused within a call, and the call has
    // synchronised before the next one refills them).
    if (ctx->arena) return tw_arena_host(ctx, bytes, out);
    if (ctx->pin_busy) {
        TW_HIP(ctx, hipEventSynchronize(ctx->pin_ev));
        ctx->pin_busy = false;
    }
    if (ctx->pin_cap < bytes) {
        if (ctx->pin) (void)hipHostFree(ctx->pin);
        ctx->pin = nullptr;
        ctx->pin_cap = 0;
        const size_t cap = (bytes + 0xFFFF) & ~size_t(0xFFFF);
        hipError_t e = hipHostMalloc(&ctx->pin, cap, hipHostMallocDefault);
        if (e != hipSuccess) {
            ctx->pin = nullptr;
            (void)hipGetLastError();
            return tw_fail(ctx, TW_E_NOMEM, "pinned host allocation of %zu bytes failed: %s", cap,
                           hipGetErrorString(e));
        }
        ctx->pin_cap = cap;
    }
    *out = ctx->pin;
    return TW_OK;
}

int tw_pinned_issued(tw_ctx* ctx) {
    if (ctx->arena) return TW_OK;
    if (!ctx->pin_ev) TW_HIP(ctx, hipEventCreateWithFlags(&ctx->pin_ev, hipEventDisableTiming));
    TW_HIP(ctx, hipEventRecord(ctx->pin_ev, ctx->stream));
    ctx->pin_busy = true;
    return TW_OK;
}

int tw_read_small(tw_ctx* ctx, void* out, const void* dev, size_t bytes) {
    if (bytes > 4096) return tw_fail(ctx, TW_E_ARG, "tw_read_small: %zu bytes", bytes);
    if (ctx->arena) {
        void* h = nullptr;
        if (int rc = tw_arena_host(ctx, bytes, &h)) return rc;
        TW_HIP(ctx, hipMemcpyAsync(h, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
        TW_HIP(ctx, hipStreamSynchronize(ctx->stream));
        memcpy(out, h, bytes);
        return TW_OK;
    }
    if (!ctx->pin_state) {
        hipError_t e = hipHostMalloc(&ctx->pin_state, 4096, hipHostMallocDefault);
        if (e != hipSuccess) {
            ctx->pin_state = nullptr;
            (void)hipGetLastError();
            return tw_fail(ctx, TW_E_NOMEM, "pinned host allocation failed: %s", hipGetErrorString(e));
        }
    }
    TW_HIP(ctx, hipMemcpyAsync(ctx->pin_state, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
    TW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    memcpy(out, ctx->pin_state, bytes);
    return TW_OK;
}

void tw_stages_reset(tw_ctx* ctx) {
    for (int i = 0; i < ST_COUNT; i++) ctx->ev_used[i] = false;
    ctx->refine_iters = 0;
    ctx->fit_residual = 0.0;
}

void tw_stage_begin(tw_ctx* ctx, int stage) {
    if (!ctx->timing) return;
    (void)hipEventRecord(ctx->ev_begin[stage], ctx->stream);
}

void tw_stage_end(tw_ctx* ctx, int stage) {
    if (!ctx->timing) return;
    (void)hipEventRecord(ctx->ev_end[stage], ctx->stream);
    ctx->ev_used[stage] = true;
}

int tw_check_volume(tw_ctx* ctx, const tw_volume* v, const char* name) {
    if (!v || !v->data) return tw_fail(ctx, TW_E_ARG, "%s: null volume or data", name);
    if (v->nx <= 0 || v->ny <= 0 || v->nz <= 0 || v->nc <= 0)
        return tw_fail(ctx, TW_E_ARG, "%s: non-positive extent", name);
    if (dtype_size(v->dtype) == 0) return tw_fail(ctx, TW_E_ARG, "%s: dtype must be TW_U8 or TW_U16", name);
    if (v->mem != TW_MEM_HOST && v->mem != TW_MEM_DEVICE)
        return tw_fail(ctx, TW_E_ARG, "%s: mem must be TW_MEM_HOST or TW_MEM_DEVICE", name);
    if (v->full_nx < 0 || v->full_ny < 0 || v->full_nz < 0)
        return tw_fail(ctx, TW_E_ARG, "%s: negative full extent", name);
    return TW_OK;
}

int tw_stage_in(tw_ctx* ctx, const tw_volume* v, int slot, const void** dev) {
    if (v->mem == TW_MEM_DEVICE) {
        *dev = v->data;
        return TW_OK;
    }
    const size_t bytes = (size_t)vol_elems(v) * dtype_size(v->dtype);
    void* p = nullptr;
    int rc = tw_ws(ctx, slot, bytes, &p);
    if (rc) return rc;
    tw_stage_begin(ctx, ST_H2D);
    TW_HIP(ctx, hipMemcpyAsync(p, v->data, bytes, hipMemcpyHostToDevice, ctx->stream));
    tw_stage_end(ctx, ST_H2D);
    *dev = p;
    return TW_OK;
}

int tw_stage_out_buffer(tw_ctx* ctx, const tw_volume* v, int slot, void** dev) {
    if (v->mem == TW_MEM_DEVICE) {
        *dev = v->data;
        return TW_OK;
    }
    const size_t bytes = (size_t)vol_elems(v) * dtype_size(v->dtype);
    return tw_ws(ctx, slot, bytes, dev);
}

int tw_stage_out_finish(tw_ctx* ctx, const tw_volume* v, const void* dev) {
    if (v->mem == TW_MEM_DEVICE) return TW_OK;
    const size_t bytes = (size_t)vol_elems(v) * dtype_size(v->dtype);
    tw_stage_begin(ctx, ST_D2H);
    TW_HIP(ctx, hipMemcpyAsync(v->data, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
    tw_stage_end(ctx, ST_D2H);
    TW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return TW_OK;
}

void tw_bspline_tb(int n, double* TB) {
    // B = M/6 with M the uniform cubic B-spline matrix; T(i,:) = [t^3 t^2 t 1],
    // t = i/n; TB = T*B summed over the inner index in ascending order.
    static const double Mi[4][4] = {{-1, 3, -3, 1}, {3, -6, 3, 0}, {-3, 0, 3, 0}, {1, 4, 1, 0}};
    double Bm[4][4];
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) Bm[a][b] = Mi[a][b] / 6.0;
    const double step = 1.0 / n;
    for (int i = 0; i < n; i++) {
        const double t = step * i;
        const double T[4] = {std::pow(t, 3.0), std::pow(t, 2.0), std::pow(t, 1.0), std::pow(t, 0.0)};
        for (int j = 0; j < 4; j++) {
            double s = 0.0;
            for (int m = 0; m < 4; m++) s += T[m] * Bm[m][j];
            TB[i * 4 + j] = s;
        }
    }
}

void tw_df_tb(int n, int basis, double* TB) {
    if (basis != TW_BASIS_LINEAR) {
        tw_bspline_tb(n, TB);
        return;
    }
    const double step = 1.0 / n;
    for (int i = 0; i < n; i++) {
        const double t = step * i;
        TB[i * 4 + 0] = 0.0;
        TB[i * 4 + 1] = 1.0 - t;
        TB[i * 4 + 2] = t;
        TB[i * 4 + 3] = 0.0;
    }
}

extern "C" {

const char* tw_version(void) { return "terawarp-mi355x 0.1 (gfx950)"; }

int tw_create(int device, tw_ctx** out) {
    if (!out) return TW_E_ARG;
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return tw_fail(nullptr, TW_E_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    }
    if (device < 0 || device >= n) return tw_fail(nullptr, TW_E_ARG, "device %d out of range [0,%d)", device, n);
    tw_ctx* ctx = new tw_ctx();
    ctx->device = device;
    ctx->err.reserve(512);  // tw_fail's messages (< 512 bytes) never allocate
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return tw_fail(nullptr, TW_E_DEVICE, "cannot initialise device %d", device);
    }
    ctx->stream = ctx->own_stream;
    for (int i = 0; i < ST_COUNT; i++) {
        if (hipEventCreate(&ctx->ev_begin[i]) != hipSuccess || hipEventCreate(&ctx->ev_end[i]) != hipSuccess) {
            delete ctx;
            return tw_fail(nullptr, TW_E_DEVICE, "cannot create HIP events");
        }
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        ctx->num_cus = prop.multiProcessorCount;
    *out = ctx;
    return TW_OK;
}

int tw_destroy(tw_ctx* ctx) {
    if (!ctx) return TW_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < WS_COUNT; i++)
        if (ctx->ws[i]) (void)hipFree(ctx->ws[i]);
    if (ctx->solve_flags) (void)hipFree(ctx->solve_flags);
    if (ctx->pin) (void)hipHostFree(ctx->pin);
    if (ctx->pin_state) (void)hipHostFree(ctx->pin_state);
    if (ctx->pin_ev) (void)hipEventDestroy(ctx->pin_ev);
    for (int i = 0; i < ST_COUNT; i++) {
        if (ctx->ev_begin[i]) (void)hipEventDestroy(ctx->ev_begin[i]);
        if (ctx->ev_end[i]) (void)hipEventDestroy(ctx->ev_end[i]);
    }
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return TW_OK;
}

int tw_set_stream(tw_ctx* ctx, void* hip_stream) {
    if (!ctx) return TW_E_ARG;
    ctx->stream = hip_stream ? (hipStream_t)hip_stream : ctx->own_stream;
    return TW_OK;
}

void* tw_get_stream(tw_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }
int tw_sync(tw_ctx* ctx) {
    if (!ctx) return TW_E_ARG;
    TW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return TW_OK;
}

int tw_set_timing(tw_ctx* ctx, int enable) {
    if (!ctx) return TW_E_ARG;
    ctx->timing = enable != 0;
    return TW_OK;
}

int tw_set_tile(tw_ctx* ctx, int tile) {
    if (!ctx) return TW_E_ARG;
    if (tile != TW_TILE_AUTO && tile != TW_TILE_GATHER && tile != TW_TILE_LDS)
        return tw_fail(ctx, TW_E_ARG, "tile must be TW_TILE_AUTO, TW_TILE_GATHER or TW_TILE_LDS");
    ctx->tile = tile;
    return TW_OK;
}

int tw_set_tune(tw_ctx* ctx, int key, int64_t value) {
    if (!ctx) return TW_E_ARG;
    switch (key) {
        case TW_TUNE_TILE:
            return tw_set_tile(ctx, (int)value);
        case TW_TUNE_TZ:
        case TW_TUNE_TZ_BATCH:
            if (value < 0 || value > 256)
                return tw_fail(ctx, TW_E_ARG, "tile depth takes 0 .. 256 (got %lld)", (long long)value);
            (key == TW_TUNE_TZ ? ctx->tune_tz : ctx->tune_tz_batch) = (int)value;
            return TW_OK;
        case TW_TUNE_FP32_FACTOR_MIN:
            if (value < 1 || value > (1ll << 30))
                return tw_fail(ctx, TW_E_ARG, "TW_TUNE_FP32_FACTOR_MIN takes 1 .. 2^30 (got %lld)", (long long)value);
            ctx->fp32_factor_min = (int)value;
            return TW_OK;
        case TW_TUNE_SOLVE_WAIT_US:
            if (value < 0 || value > 10000000)
                return tw_fail(ctx, TW_E_ARG, "TW_TUNE_SOLVE_WAIT_US takes 0 .. 10^7 (got %lld)", (long long)value);
            ctx->solve_wait_us = (int)value;
            ctx->pipe_off = 0;
            return TW_OK;
        case TW_TUNE_SOLVE_PIPELINED:
            return tw_fail(ctx, TW_E_ARG, "TW_TUNE_SOLVE_PIPELINED is read only (set TW_TUNE_SOLVE_WAIT_US)");
        default:
            return tw_fail(ctx, TW_E_ARG, "unknown tunable %d", key);
    }
}

int tw_get_tune(tw_ctx* ctx, int key, int64_t* value) {
    if (!ctx || !value) return TW_E_ARG;
    switch (key) {
        case TW_TUNE_TILE: *value = ctx->tile; return TW_OK;
        case TW_TUNE_TZ: *value = ctx->tune_tz; return TW_OK;
        case TW_TUNE_TZ_BATCH: *value = ctx->tune_tz_batch; return TW_OK;
        case TW_TUNE_FP32_FACTOR_MIN: *value = ctx->fp32_factor_min; return TW_OK;
        case TW_TUNE_SOLVE_WAIT_US: *value = ctx->solve_wait_us; return TW_OK;
        case TW_TUNE_SOLVE_PIPELINED: *value = ctx->pipe_off ? 0 : 1; return TW_OK;
        default: return tw_fail(ctx, TW_E_ARG, "unknown tunable %d", key);
    }
}

int tw_get_timings(tw_ctx* ctx, tw_timings* out) {
    if (!ctx || !out) return TW_E_ARG;
    memset(out, 0, sizeof *out);
    double* dst[ST_COUNT] = {&out->fit_ms, &out->grid_ms, &out->quant_ms, &out->resample_ms,
                             &out->h2d_ms, &out->d2h_ms, &out->total_ms};
    for (int i = 0; i < ST_COUNT; i++) {
        if (!ctx->ev_used[i]) continue;
        TW_HIP(ctx, hipEventSynchronize(ctx->ev_end[i]));
        float ms = 0.f;
        TW_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev_begin[i], ctx->ev_end[i]));
        *dst[i] = ms;
    }
    out->refine_iters = ctx->refine_iters;
    out->fit_residual = ctx->fit_residual;
    return TW_OK;
}

const char* tw_last_error(tw_ctx* ctx) { return ctx ? ctx->err.c_str() : ""; }

int64_t tw_grid_dim(int64_t n, int f) {
    if (n <= 0 || f <= 0) return 0;
    return (int64_t)std::ceil((double)n / f) + 3;
}

}  // extern "C"
