// tw_noalloc.hip -- the chained TPS warp without allocations.
//
// tw_warp_tps_noalloc runs tw_warp_tps's kernels (same arithmetic, same
// outputs) with every workspace carved from `scratch`, a static device array
// of kScratchBytes that each GPU holds from the first call on it until the
// process ends, and host staging from a static host array per device.  It
// makes no hipMalloc, hipHostMalloc, hipFree or other allocating call, for
// callers that cannot allocate: code under HIP graph capture, tight serving
// loops.  The caller passes device buffers for the volumes.  Every call on a
// GPU shares that GPU's scratch, which keeps the last call's contents (what
// must start at zero is cleared by a kernel in each call); calls that overlap
// on one GPU race.
#include "tw_internal.h"

#include <mutex>

namespace {
constexpr size_t kScratchBytes = 64ull << 20;  // device scratch per GPU
constexpr size_t kHostBytes = 256ull << 10;    // host staging per GPU
constexpr int kMaxDevices = 16;
constexpr int kMaxLandmarks = 1024;
// Upper bound of the fit's workspaces at K landmarks: the projected system and
// the fused steps' panel store (n x n FP64 each) plus 1 MiB for the vectors,
// the inverted diagonal blocks, the landmark pack and the coefficients.
inline size_t fit_bytes(int K) { return 2 * (size_t)K * K * sizeof(double) + (1ull << 20); }
}  // namespace

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

std::mutex g_setup;
uint8_t* g_scratch[kMaxDevices] = {};
alignas(256) uint8_t g_host[kMaxDevices][kHostBytes];

// This GPU's copy of `scratch`: its device address is looked up once per
// device, with that device current (each GPU holds the array at its own
// address); host code uses only that address, never the symbol itself.
int scratch_ptr(tw_ctx* ctx, uint8_t** out) {
    std::lock_guard<std::mutex> g(g_setup);
    if (!g_scratch[ctx->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return tw_fail(ctx, TW_E_DEVICE, "cannot resolve the scratch array on device %d: %s", ctx->device,
                           hipGetErrorString(e));
        }
        g_scratch[ctx->device] = (uint8_t*)p;
    }
    *out = g_scratch[ctx->device];
    return TW_OK;
}

// Sets the context's arena for one call and clears it on every exit path.
struct ArenaScope {
    tw_ctx* c;
    ArenaScope(tw_ctx* ctx, uint8_t* dev) : c(ctx) {
        c->arena = dev;
        c->arena_cap = kScratchBytes;
        c->arena_off = 0;
        c->arena_host = g_host[ctx->device];
        c->arena_host_cap = kHostBytes;
        c->arena_host_off = 0;
        for (int i = 0; i < WS_COUNT; i++) {
            c->arena_ws[i] = nullptr;
            c->arena_ws_cap[i] = 0;
        }
    }
    ~ArenaScope() {
        c->arena = nullptr;
        c->arena_host = nullptr;
        c->arena_cap = c->arena_off = c->arena_host_cap = c->arena_host_off = 0;
    }
};

}  // namespace

extern "C" int tw_warp_tps_noalloc(tw_ctx* ctx, const tw_volume* src, tw_volume* dst, const float* centres,
                                   const float* targets, int K, int f, double lambda, int interp, int precision) {
    if (!ctx) return TW_E_ARG;
    int rc;
    if ((rc = tw_warp_tps_check(ctx, src, dst, f, interp, precision))) return rc;
    if (src->mem != TW_MEM_DEVICE || dst->mem != TW_MEM_DEVICE)
        return tw_fail(ctx, TW_E_ARG, "tw_warp_tps_noalloc takes device buffers only (host ones would need staging)");
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
        return tw_fail(ctx, TW_E_UNSUPPORTED, "tw_warp_tps_noalloc serves devices 0..%d", kMaxDevices - 1);
    if (K > kMaxLandmarks)
        return tw_fail(ctx, TW_E_UNSUPPORTED, "tw_warp_tps_noalloc takes at most %d landmarks (got %d)",
                       kMaxLandmarks, K);
    if (precision != TW_EXACT && !fits_fast(src->nx, src->ny, src->nz, dtype_size(src->dtype)))
        return tw_fail(ctx, TW_E_UNSUPPORTED,
                       "tw_warp_tps_noalloc takes sources below 2 GiB (larger ones are resampled in slabs)");
    // Grid: FP32 field (12 B per node); the mixed form adds its FP16 planes
    // (6 B) and the block table (24 B per 8^3 nodes).
    const i64 gx = tw_grid_dim(dst->nx, f), gy = tw_grid_dim(dst->ny, f), gz = tw_grid_dim(dst->nz, f);
    const size_t plane = (size_t)(gx * gy * gz);
    const size_t grid_b = plane * 12 + (precision == TW_MIXED ? plane * 6 + (plane / 512 + 1) * 64 : 0);
    const size_t need = grid_b + fit_bytes(K > 5 ? K : 5);
    if (need > kScratchBytes)
        return tw_fail(ctx, TW_E_UNSUPPORTED,
                       "tw_warp_tps_noalloc: %lld grid nodes and %d landmarks need about %zu MiB of scratch; it "
                       "has %zu MiB",
                       (long long)plane, K, (need >> 20) + 1, kScratchBytes >> 20);
    if ((rc = tw_bind(ctx))) return rc;
    uint8_t* base = nullptr;
    if ((rc = scratch_ptr(ctx, &base))) return rc;
    ArenaScope scope(ctx, base);
    return tw_warp_tps_chain(ctx, src, dst, centres, targets, K, f, lambda, interp, precision);
}
