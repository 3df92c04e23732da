// tw_chain.hip -- the whole in-memory TPS chain on one context.
//
// Reference: GPU branch of imgwarp_smallmemory_image
// (q_imgwarp_tps_quicksmallmemory.cpp:1997-2162): STPS fit and grid DF
// (compute_df_stps_subsampled_volume_4bspline, 2066), then the fused
// upsample + resample (gpu_interpolation_stps, 2143).  Here the volume is
// staged once, the grid never leaves the device, and no stage resets the
// device.
#include "tw_internal.h"

#include <cmath>
#include <vector>

extern "C" int tw_warp_tps(tw_ctx* ctx, const tw_volume* src, tw_volume* dst, const float* centres,
                           const float* targets, int K, int f, double lambda, int interp, int precision) {
    if (!ctx) return TW_E_ARG;
    int rc;
    if ((rc = tw_check_volume(ctx, src, "src")) || (rc = tw_check_volume(ctx, dst, "dst"))) return rc;
    if (f < 1) return tw_fail(ctx, TW_E_ARG, "grid spacing must be >= 1");
    // Block-size checks of imgwarp_smallmemory_image (2011-2020).
    if (f >= src->nx || f >= src->ny || f >= src->nz)
        return tw_fail(ctx, TW_E_ARG, "block size should be smaller than the image size");
    if ((rc = tw_bind(ctx))) return rc;

    std::vector<double> c((size_t)K * 4), d(16);
    const int fit_prec = precision == TW_MIXED ? TW_MIXED : TW_PLAIN;
    if ((rc = tw_tps_fit(ctx, centres, targets, K, lambda, fit_prec, c.data(), d.data()))) return rc;
    tw_timings tf;
    tw_get_timings(ctx, &tf);

    // Grid over the output window (global coordinates of dst).
    tw_grid g = {};
    g.f = f;
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
    // The grid buffer lives in the WS_DF slot: the grid kernel stages its own
    // FP32 copy in WS_GRID when quantising.
    if ((rc = tw_ws(ctx, WS_DF, plane * 3 * (g.storage == TW_DF_F16 ? 2 : 4), &gbuf))) return rc;
    g.data = gbuf;
    if ((rc = tw_tps_grid_eval(ctx, centres, K, c.data(), d.data(), precision, &g))) return rc;
    tw_timings tg;
    tw_get_timings(ctx, &tg);
    rc = tw_resample_tps(ctx, src, dst, &g, interp, precision);
    tw_timings tr;
    tw_get_timings(ctx, &tr);
    tw_grid_release(ctx, &g);
    // Report the stages of the whole chain.
    ctx->refine_iters = tf.refine_iters;
    ctx->fit_residual = tf.fit_residual;
    (void)tg;
    (void)tr;
    return rc;
}
