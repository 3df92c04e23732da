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
// output write); the FP64 weight algebra (~45 VALU ops per voxel) makes the
// uint8 case VALU-bound on MI355X, the uint16 case near the HBM/VALU
// balance point.
#include "tw_internal.h"

#include <cmath>
#include <cstring>

namespace {

struct AffineArgs {
    const void* src;
    void* dst;
    i64 snx, sny, snz, sx0, sy0, sz0;  // source window
    i64 fnx, fny, fnz;                 // full subject extent (bounds test)
    i64 dnx, dny, dnz, dx0, dy0, dz0;  // destination window
    i64 nc;
    i64 zbase;                         // first output slice of this launch
    double m[12];
};

constexpr int AFF_BX = 64, AFF_BY = 4;

template <typename T, int INTERP>
__global__ __launch_bounds__(256) void k_affine(const AffineArgs a) {
    const i64 x = (i64)blockIdx.x * AFF_BX + threadIdx.x;
    const i64 y = (i64)blockIdx.y * AFF_BY + threadIdx.y;
    const i64 z = a.zbase + blockIdx.z;
    if (x >= a.dnx || y >= a.dny || z >= a.dnz) return;
    const T* __restrict__ src = (const T*)a.src;
    T* __restrict__ dst = (T*)a.dst;
    const i64 dvol = a.dnx * a.dny * a.dnz;
    const i64 o = (z * a.dny + y) * a.dnx + x;

    const double gx = (double)(a.dx0 + x), gy = (double)(a.dy0 + y), gz = (double)(a.dz0 + z);
    const double p0 = a.m[0] * gx + a.m[1] * gy + a.m[2] * gz + a.m[3] * 1.0;
    const double p1 = a.m[4] * gx + a.m[5] * gy + a.m[6] * gz + a.m[7] * 1.0;
    const double p2 = a.m[8] * gx + a.m[9] * gy + a.m[10] * gz + a.m[11] * 1.0;

    if (p0 < 0 || p0 > (double)(a.fnx - 1) || p1 < 0 || p1 > (double)(a.fny - 1) || p2 < 0 ||
        p2 > (double)(a.fnz - 1)) {
        for (i64 c = 0; c < a.nc; c++) dst[c * dvol + o] = 0;
        return;
    }
    const i64 sxy = a.snx * a.sny, svol = sxy * a.snz;
    if (INTERP == TW_INTERP_NN) {
        i64 qx = (i64)(p0 + 0.5), qy = (i64)(p1 + 0.5), qz = (i64)(p2 + 0.5);
        qx = qx < a.fnx ? qx : a.fnx - 1;
        qy = qy < a.fny ? qy : a.fny - 1;
        qz = qz < a.fnz ? qz : a.fnz - 1;
        qx -= a.sx0; qy -= a.sy0; qz -= a.sz0;
        const bool in = qx >= 0 && qx < a.snx && qy >= 0 && qy < a.sny && qz >= 0 && qz < a.snz;
        for (i64 c = 0; c < a.nc; c++)
            dst[c * dvol + o] = in ? src[c * svol + qz * sxy + qy * a.snx + qx] : (T)0;
        return;
    }
    const double fx = floor(p0), fy = floor(p1), fz = floor(p2);
    const i64 xs = (i64)fx, xb = (i64)ceil(p0);
    const i64 ys = (i64)fy, yb = (i64)ceil(p1);
    const i64 zs = (i64)fz, zb = (i64)ceil(p2);
    const double lw = 1.0 - (p0 - (double)xs), rw = 1.0 - lw;
    const double tw = 1.0 - (p1 - (double)ys), bw = 1.0 - tw;
    const double uw = 1.0 - (p2 - (double)zs), dw = 1.0 - uw;

    const i64 lxs = xs - a.sx0, lxb = xb - a.sx0, lys = ys - a.sy0, lyb = yb - a.sy0;
    const i64 lzs = zs - a.sz0, lzb = zb - a.sz0;
    if (lxs < 0 || lxb >= a.snx || lys < 0 || lyb >= a.sny || lzs < 0 || lzb >= a.snz) {
        // Needed source voxels are outside the staged window.
        for (i64 c = 0; c < a.nc; c++) dst[c * dvol + o] = 0;
        return;
    }
    const i64 r00 = lzs * sxy + lys * a.snx, r01 = lzs * sxy + lyb * a.snx;
    const i64 r10 = lzb * sxy + lys * a.snx, r11 = lzb * sxy + lyb * a.snx;
    for (i64 c = 0; c < a.nc; c++) {
        const T* s = src + c * svol;
        const double v000 = (double)s[r00 + lxs], v100 = (double)s[r00 + lxb];
        const double v010 = (double)s[r01 + lxs], v110 = (double)s[r01 + lxb];
        const double v001 = (double)s[r10 + lxs], v101 = (double)s[r10 + lxb];
        const double v011 = (double)s[r11 + lxs], v111 = (double)s[r11 + lxb];
        const double hi = tw * (lw * v000 + rw * v100) + bw * (lw * v010 + rw * v110);
        const double lo = tw * (lw * v001 + rw * v101) + bw * (lw * v011 + rw * v111);
        dst[c * dvol + o] = (T)(uw * hi + dw * lo);
    }
}

template <typename T>
int launch_affine(tw_ctx* ctx, const AffineArgs& a, int interp) {
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
