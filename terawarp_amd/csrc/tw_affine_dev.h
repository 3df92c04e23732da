// tw_affine_dev.h -- per-voxel affine / rigid resample in the reference's
// FP64 operation order, shared by k_affine.hip and the batch kernel
// (k_batch.hip).  Reference and semantics: see k_affine.hip's header.
#pragma once
#include "tw_internal.h"

#include <cmath>

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

// Output voxel (x, y, z) of the destination window, channels [c0, c1).
template <typename T, int INTERP>
__device__ __forceinline__ void affine_voxel(const AffineArgs& a, i64 x, i64 y, i64 z, i64 c0, i64 c1) {
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
        for (i64 c = c0; c < c1; c++) dst[c * dvol + o] = 0;
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
        for (i64 c = c0; c < c1; c++)
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
        for (i64 c = c0; c < c1; c++) dst[c * dvol + o] = 0;
        return;
    }
    const i64 r00 = lzs * sxy + lys * a.snx, r01 = lzs * sxy + lyb * a.snx;
    const i64 r10 = lzb * sxy + lys * a.snx, r11 = lzb * sxy + lyb * a.snx;
    for (i64 c = c0; c < c1; c++) {
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

}  // namespace
