// k_affine.hip -- affine / rigid resample (FP64, the reference's operation order).
//
// Reference: CPU loop of q_imagewarp_affine_image (q_warp_affine_tps.cpp:1373-1437),
// GPU get_Displacement_affine (aff.cu:379-447) and the tera-block kernels
// get_Displacement_affine_new_{Z,Y,X} (aff.cu:136-375).
//
// Per output voxel: p = M [x y z 1] summed left to right, strict bounds test
// against the full subject extent, floor/ceil neighbours, weights
// l = 1-(p-floor), r = 1-l, and the blend tree
//   t*(l*v000 + r*v100) + b*(l*v010 + r*v110)   (slice z_s)
//   u*hi + d*lo, truncated to the voxel type (no +0.5).
// The library is compiled with -ffp-contract=off, so every product and sum
// rounds separately, exactly as in the FP64 CPU path; outputs are
// bit-identical to it for the same matrix.
//
// Roofline: 2*bpv algorithmic bytes per voxel (one subject read, one
// output write).  The fast path (tw_affine_dev.h) issues four pair loads per
// voxel, like the fused TPS kernel, and ~90 VALU instructions (FP64 weights
// and blend in the reference's order), so it is bound by the texture-address
// unit and the VALU, not by HBM bandwidth.
#include "tw_internal.h"

#include <cmath>
#include <cstring>

#include "tw_affine_dev.h"

namespace {

template <typename T, int INTERP>
__global__ __launch_bounds__(256) void k_affine(const AffineArgs a) {
    const i64 x = (i64)blockIdx.x * AFF_BX + threadIdx.x;
    const i64 y = (i64)blockIdx.y * AFF_BY + threadIdx.y;
    const i64 z = a.zbase + blockIdx.z;
    if (x >= a.dnx || y >= a.dny || z >= a.dnz) return;
    affine_voxel<T, INTERP>(a, x, y, z, 0, a.nc);
}

template <typename T, int INTERP>
__global__ __launch_bounds__(FT) void k_affine_fast(const AffineArgs a, const AffGeom g) {
    // XCD-major tile order, as the fused TPS kernel.
    const int wg = blockIdx.x;
    const int tile = (wg & 7) * g.per_xcd + (wg >> 3);
    if (tile >= g.ntiles) return;
    affine_tile_fast<T, INTERP>(a, g, tile, 0, (int)a.nc);
}

template <typename T, int INTERP>
__global__ __launch_bounds__(FT) void k_affine_lds(const AffineArgs a, const AffGeom g) {
    extern __shared__ __align__(16) uint8_t lds_box[];
    const int wg = blockIdx.x;
    const int tile = (wg & 7) * g.per_xcd + (wg >> 3);
    if (tile >= g.ntiles) return;
    affine_tile_lds<T, INTERP>(a, g, tile, 0, (int)a.nc, lds_box);
}

template <typename T>
int launch_affine(tw_ctx* ctx, const AffineArgs& a, int interp) {
    AffGeom g;
    const bool lds = affine_prefers_lds(a.m);
    if (a.nc <= 1024 && !affine_geom<T>(a, g, lds)) {
        const dim3 grid((unsigned)(8 * g.per_xcd)), block(TX, FTY, 1);
        if (lds) {
            const size_t sh = (size_t)g.lds_bytes;
            if (interp == TW_INTERP_NN)
                hipLaunchKernelGGL((k_affine_lds<T, TW_INTERP_NN>), grid, block, sh, ctx->stream, a, g);
            else
                hipLaunchKernelGGL((k_affine_lds<T, TW_INTERP_LINEAR>), grid, block, sh, ctx->stream, a, g);
        } else if (interp == TW_INTERP_NN) {
            hipLaunchKernelGGL((k_affine_fast<T, TW_INTERP_NN>), grid, block, 0, ctx->stream, a, g);
        } else {
            hipLaunchKernelGGL((k_affine_fast<T, TW_INTERP_LINEAR>), grid, block, 0, ctx->stream, a, g);
        }
        TW_LAUNCH_CHECK(ctx);
        return TW_OK;
    }
    // Windows beyond 32-bit offsets (or one voxel wide): 64-bit addressing.
    const dim3 block(AFF_BX, AFF_BY, 1);
    const i64 zmax = 65535;
    for (i64 z0 = 0; z0 < a.dnz; z0 += zmax) {
        AffineArgs b = a;
        b.zbase = z0;
        const i64 nz = (a.dnz - z0) < zmax ? (a.dnz - z0) : zmax;
        const dim3 grid((unsigned)((a.dnx + AFF_BX - 1) / AFF_BX), (unsigned)((a.dny + AFF_BY - 1) / AFF_BY),
                        (unsigned)nz);
        if (interp == TW_INTERP_NN)
            hipLaunchKernelGGL((k_affine<T, TW_INTERP_NN>), grid, block, 0, ctx->stream, b);
        else
            hipLaunchKernelGGL((k_affine<T, TW_INTERP_LINEAR>), grid, block, 0, ctx->stream, b);
        TW_LAUNCH_CHECK(ctx);
    }
    return TW_OK;
}

}  // namespace

extern "C" int tw_resample_affine(tw_ctx* ctx, const tw_volume* src, tw_volume* dst, const double* M,
                                  int interp) {
    if (!ctx) return TW_E_ARG;
    int rc;
    if ((rc = tw_check_volume(ctx, src, "src")) || (rc = tw_check_volume(ctx, dst, "dst"))) return rc;
    if (!M) return tw_fail(ctx, TW_E_ARG, "null matrix");
    if (src->dtype != dst->dtype || src->nc != dst->nc)
        return tw_fail(ctx, TW_E_ARG, "src and dst must share dtype and channel count");
    if (interp != TW_INTERP_LINEAR && interp != TW_INTERP_NN)
        return tw_fail(ctx, TW_E_ARG, "interp must be TW_INTERP_LINEAR or TW_INTERP_NN");
    for (int i = 0; i < 12; i++)
        if (!std::isfinite(M[i])) return tw_fail(ctx, TW_E_ARG, "matrix entry %d is not finite", i);
    if ((rc = tw_bind(ctx))) return rc;
    tw_stages_reset(ctx);
    tw_stage_begin(ctx, ST_TOTAL);

    AffineArgs a;
    memset(&a, 0, sizeof a);
    const void* sdev = nullptr;
    void* ddev = nullptr;
    if ((rc = tw_stage_in(ctx, src, WS_SRC, &sdev))) return rc;
    if ((rc = tw_stage_out_buffer(ctx, dst, WS_DST, &ddev))) return rc;
    a.src = sdev;
    a.dst = ddev;
    a.snx = src->nx; a.sny = src->ny; a.snz = src->nz;
    a.sx0 = src->x0; a.sy0 = src->y0; a.sz0 = src->z0;
    a.fnx = src->full_nx ? src->full_nx : src->x0 + src->nx;
    a.fny = src->full_ny ? src->full_ny : src->y0 + src->ny;
    a.fnz = src->full_nz ? src->full_nz : src->z0 + src->nz;
    a.dnx = dst->nx; a.dny = dst->ny; a.dnz = dst->nz;
    a.dx0 = dst->x0; a.dy0 = dst->y0; a.dz0 = dst->z0;
    a.nc = src->nc;
    for (int i = 0; i < 12; i++) a.m[i] = M[i];

    tw_stage_begin(ctx, ST_RESAMPLE);
    rc = src->dtype == TW_U8 ? launch_affine<uint8_t>(ctx, a, interp) : launch_affine<uint16_t>(ctx, a, interp);
    if (rc) return rc;
    tw_stage_end(ctx, ST_RESAMPLE);
    if ((rc = tw_stage_out_finish(ctx, dst, ddev))) return rc;
    tw_stage_end(ctx, ST_TOTAL);
    if (src->mem == TW_MEM_HOST) TW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return TW_OK;
}
