// tw_chain.hip -- the whole in-memory TPS chain on one context.
//
// Reference: GPU branch of imgwarp_smallmemory_image
// (q_imgwarp_tps_quicksmallmemory.cpp:1997-2162): STPS fit and grid DF
// (compute_df_stps_subsampled_volume_4bspline, 2066), then the fused
// upsample + resample (gpu_interpolation_stps, 2143).  Here every stage
// stays on the device and on the context's stream: the fit leaves its
// coefficients where the grid kernel reads them, the grid never leaves the
// device, nothing resets the device, and the stream is synchronised once,
// at the end, where the fit's status is checked.
#include "tw_internal.h"

#include <cmath>
#include <vector>

int tw_warp_tps_check(tw_ctx* ctx, const tw_volume* src, const tw_volume* dst, int f, int interp, int precision) {
    int rc;
    if ((rc = tw_check_volume(ctx, src, "src")) || (rc = tw_check_volume(ctx, dst, "dst"))) return rc;
    if (src->dtype != dst->dtype || src->nc != dst->nc)
        return tw_fail(ctx, TW_E_ARG, "src and dst must share dtype and channel count");
    interp &= ~TW_DF_LINEAR;
    if (interp != TW_INTERP_LINEAR && interp != TW_INTERP_NN)
        return tw_fail(ctx, TW_E_ARG, "interp must be TW_INTERP_LINEAR or TW_INTERP_NN (optionally | TW_DF_LINEAR)");
    if (precision != TW_EXACT && precision != TW_PLAIN && precision != TW_MIXED)
        return tw_fail(ctx, TW_E_ARG, "unknown precision %d", precision);
    if (f < 1) return tw_fail(ctx, TW_E_ARG, "grid spacing must be >= 1");
    // Block-size checks of imgwarp_smallmemory_image (2011-2020).
    if (f >= src->nx || f >= src->ny || f >= src->nz)
        return tw_fail(ctx, TW_E_ARG, "block size should be smaller than the image size");
    return TW_OK;
}

extern "C" int tw_warp_tps(tw_ctx* ctx, const tw_volume* src, tw_volume* dst, const float* centres,
                           const float* targets, int K, int f, double lambda, int interp, int precision) {
    if (!ctx) return TW_E_ARG;
    int rc;
    if ((rc = tw_warp_tps_check(ctx, src, dst, f, interp, precision))) return rc;
    if ((rc = tw_bind(ctx))) return rc;
    return tw_warp_tps_chain(ctx, src, dst, centres, targets, K, f, lambda, interp, precision);
}

int tw_warp_tps_chain(tw_ctx* ctx, const tw_volume* src, tw_volume* dst, const float* centres, const float* targets,
                      int K, int f, double lambda, int interp, int precision) {
    int rc;
    const int interp_arg = interp, pipe_was_off = ctx->pipe_off;
    const int basis = (interp & TW_DF_LINEAR) ? TW_BASIS_LINEAR : TW_BASIS_BSPLINE;
    interp &= ~TW_DF_LINEAR;
    tw_stages_reset(ctx);
    tw_stage_begin(ctx, ST_TOTAL);

    TwFitDev fit;
    const int fit_prec = precision == TW_MIXED ? TW_MIXED : TW_PLAIN;
    if ((rc = tw_tps_fit_dev(ctx, centres, targets, K, lambda, fit_prec, &fit))) return rc;

    // Grid over the output window (global coordinates of dst).
    tw_grid g = {};
    g.f = f;
    g.basis = basis;
    g.ox = dst->x0;
    g.oy = dst->y0;
    g.oz = dst->z0;
    g.gx = tw_grid_dim(dst->nx, f);
    g.gy = tw_grid_dim(dst->ny, f);
    g.gz = tw_grid_dim(dst->nz, f);
    g.mem = TW_MEM_DEVICE;
    g.storage = precision == TW_MIXED ? TW_DF_F16 : TW_DF_F32;
    const size_t plane = (size_t)g.gx * g.gy * g.gz;
    void* gbuf = nullptr;
    // The grid buffer lives in the WS_DF slot (the grid kernel stages its
    // FP32 field in WS_GRID when quantising); the FP16 block table in WS_QTAB.
    if ((rc = tw_ws(ctx, WS_DF, plane * 3 * (g.storage == TW_DF_F16 ? 2 : 4), &gbuf))) return rc;
    g.data = gbuf;
    if (g.storage == TW_DF_F16) {
        const size_t nb = (size_t)((g.gx + 7) / 8) * ((g.gy + 7) / 8) * ((g.gz + 7) / 8);
        void* q = nullptr;
        if ((rc = tw_ws(ctx, WS_QTAB, nb * 6 * sizeof(float), &q))) return rc;
        g.qtab = (float*)q;
    }
    if ((rc = tw_grid_eval_dev(ctx, fit.pts, fit.cf, fit.d, K, precision, &g))) return rc;
    if ((rc = tw_resample_tps_run(ctx, src, dst, &g, interp, precision))) return rc;
    tw_stage_end(ctx, ST_TOTAL);
    // One synchronisation: the fit's status (a failed fit leaves dst unspecified).
    rc = tw_tps_fit_check(ctx, &fit);
    // A pipelined solve whose wait ran out: once more, per-block launches.
    if (rc && ctx->pipe_off && !pipe_was_off)
        return tw_warp_tps_chain(ctx, src, dst, centres, targets, K, f, lambda, interp_arg, precision);
    return rc;
}
