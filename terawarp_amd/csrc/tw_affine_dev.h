// tw_affine_dev.h -- affine / rigid resample in the reference's FP64
// operation order, shared by k_affine.hip and the batch kernel
// (k_batch.hip).  Reference and semantics: see k_affine.hip's header.
#pragma once
#include "tw_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "tw_tps_dev.h"  // buffer-load helpers and the tile shape

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

// ---- fast path ---------------------------------------------------------------
// The same per-voxel arithmetic with the fused TPS kernel's addressing:
// 32-bit byte offsets through four buffer resources (+0, +row, +plane,
// +row+plane), so a voxel's eight taps are four (x, x+1) pair loads, and four
// z-slices per step.  The reference reads (x_s, x_b) with x_b = ceil(p); where
// x_b = x_s its weight r = 1 - l is exactly 0, so the pair's x_s + 1 (or the
// next row / plane, likewise under a zero weight) leaves the result
// unchanged; a sample on the window's last column takes the pair (x_s - 1,
// x_s) instead, so that no pair leaves the channel (a range check would zero
// the whole access).  Bit-identical to affine_voxel.

struct AffGeom {
    int snx, sny, snz, sx0, sy0, sz0;   // source window
    uint32_t row_b, plane_b, chan_b;    // source strides / channel bytes
    uint32_t nrec01, nrec10, nrec11;    // channel bytes past +row, +plane, +row+plane
    uint32_t dchan_b;                   // destination channel bytes
    int tiles_x, tiles_y, ntiles, per_xcd;
    double fmx, fmy, fmz;               // full extent - 1 (bounds test)
    int fnx, fny, fnz;                  // full extent
    int lds;                            // 1: LDS-staged 16 x 16 x 8 tiles (affine_tile_lds), 0: 64 x 4 x 16
    int lds_bytes;                      // LDS per work-group for the staged box
    int whole;                          // the source window is the whole volume
    int ord0, ord1, ord2;               // LDS tiles: output axes from fastest to slowest in the tile order
    int tn[3];                          // tiles along output x, y, z
};

// Sample checks of the fast path: a tile's samples take the cheapest level
// its box allows.
constexpr int AFF_INTERIOR = 0;  // every sample a voxel inside the volume and the window: no tests
constexpr int AFF_BOUNDS = 1;    // whole-volume window: the volume bounds test only
constexpr int AFF_FULL = 2;      // bounds and window tests

constexpr int AFF_TZ = 16;  // tile depth (voxels along z)
// z-slices per step of the gather walk: their 16 pair loads go out together.
// Two slices per step need 65 instead of 101 VGPRs (7 instead of 5 waves per
// SIMD) but ran slower: the batch's affine items 6.07 against 5.79 ms.
constexpr int AFF_TG = 4;
// LDS-staged tiles: 16 x 16 x 8 output voxels; a work-group of 256 threads,
// each walking one (x, y) column of 8 slices.
constexpr int ALX = 16, ALY = 16, ALZ = 8;
constexpr int AFF_LDS_BYTES = 32 * 1024;  // box budget of the standalone kernel

// Whether the map's gathers scatter: along a row of output x (a gather's
// lanes) the sample moves by the matrix's first column; when its y and z
// parts move it to another source row at least every second lane, the
// gathers touch a cache line per lane or two and the LDS-staged tile pays
// for its staging (measured: near-identity maps and rotations up to 15
// degrees run faster on the 64 x 4 gather tile, large rotations on LDS).
inline bool affine_prefers_lds(const double* m) {
    return std::fabs(m[4]) + std::fabs(m[8]) > 0.5;
}

// Geometry of the fast path, or null if the window exceeds its limits (the
// caller then takes affine_voxel).
// aligned_order (LDS tiles of persistent batch launches): the tile order
// follows the map (below); single-volume launches keep x fastest.
template <typename T>
const char* affine_geom(const AffineArgs& a, AffGeom& g, bool lds = false, int lds_bytes = AFF_LDS_BYTES,
                        bool aligned_order = false) {
    const uint64_t schan = (uint64_t)a.snx * a.sny * a.snz * sizeof(T);
    const uint64_t taps = schan + (uint64_t)a.snx * (a.sny + 1) * sizeof(T) + 16;
    const uint64_t dchan = (uint64_t)a.dnx * a.dny * a.dnz * sizeof(T);
    const i64 cmax = std::max({a.dx0 + a.dnx, a.dy0 + a.dny, a.dz0 + a.dnz, a.fnx, a.fny, a.fnz});
    const i64 cmin = std::min({a.dx0, a.dy0, a.dz0, a.sx0, a.sy0, a.sz0});
    if (a.snx < 2 || taps > 0x7FFF0000ull || a.sny * a.snz >= (1ll << 24) || dchan > 0xF0000000ull ||
        cmax > (1ll << 24) || cmin < -(1ll << 24))
        return "window exceeds the fast affine path's 32-bit addressing";
    memset(&g, 0, sizeof g);
    g.snx = (int)a.snx; g.sny = (int)a.sny; g.snz = (int)a.snz;
    g.sx0 = (int)a.sx0; g.sy0 = (int)a.sy0; g.sz0 = (int)a.sz0;
    g.row_b = (uint32_t)(a.snx * sizeof(T));
    g.plane_b = (uint32_t)(a.snx * a.sny * sizeof(T));
    g.chan_b = (uint32_t)schan;
    auto past = [&](uint64_t skip) { return (uint32_t)(schan > skip ? schan - skip : 0); };
    g.nrec01 = past(g.row_b);
    g.nrec10 = past(g.plane_b);
    g.nrec11 = past((uint64_t)g.row_b + g.plane_b);
    g.dchan_b = (uint32_t)dchan;
    g.lds = lds ? 1 : 0;
    g.lds_bytes = lds_bytes;
    const int tx = lds ? ALX : TX, ty = lds ? ALY : FTY, tz = lds ? ALZ : AFF_TZ;
    g.tiles_x = (int)((a.dnx + tx - 1) / tx);
    g.tiles_y = (int)((a.dny + ty - 1) / ty);
    const i64 nt = (i64)g.tiles_x * g.tiles_y * ((a.dnz + tz - 1) / tz);
    g.tn[0] = g.tiles_x;
    g.tn[1] = g.tiles_y;
    g.tn[2] = (int)((a.dnz + tz - 1) / tz);
    // Aligned order: the output axis whose step moves the sample furthest
    // along source x fastest, then the one furthest along source y, so that
    // tiles that follow each other stage boxes along the same source rows,
    // whose cache lines they share while the lines are in the XCD's L2.
    // Each XCD takes a contiguous part of an item's tiles, so for one
    // volume this order leaves the XCDs whose slabs lie mostly outside the
    // source idle (a 1024 x 1024 x 256 rotation: 0.54 -> 0.94 ms); a batch's
    // XCDs walk their parts of every item in turn, which evens that out
    // (rigid items of the 256-item batch: 5.95 -> 5.2-5.6 ms).
    g.ord0 = 0;
    g.ord1 = 1;
    g.ord2 = 2;
    if (lds && aligned_order) {
        int f = 0;
        for (int j = 1; j < 3; j++)
            if (std::fabs(a.m[j]) > std::fabs(a.m[f])) f = j;
        const int r0 = f == 0 ? 1 : 0, r1 = f == 2 ? 1 : 2;
        const bool m0 = std::fabs(a.m[4 + r0]) >= std::fabs(a.m[4 + r1]);
        g.ord0 = f;
        g.ord1 = m0 ? r0 : r1;
        g.ord2 = m0 ? r1 : r0;
    }
    if (nt > (1ll << 30)) return "too many tiles for one launch";
    g.ntiles = (int)nt;
    g.per_xcd = (int)((nt + 7) / 8);
    g.fmx = (double)(a.fnx - 1); g.fmy = (double)(a.fny - 1); g.fmz = (double)(a.fnz - 1);
    g.fnx = (int)a.fnx; g.fny = (int)a.fny; g.fnz = (int)a.fnz;
    g.whole = a.sx0 == 0 && a.sy0 == 0 && a.sz0 == 0 && a.snx == a.fnx && a.sny == a.fny && a.snz == a.fnz;
    return nullptr;
}

struct ATap {
    uint32_t off;   // byte offset of (x_s, y_s, z_s) (or of x_s - 1: shift), OOB_OFF when outside
    bool ok;        // inside the volume and the window
    bool shift;     // last column: the pair is (x_s - 1, x_s)
    int ex, ey, ez; // AFF_INTERIOR / AFF_BOUNDS (when ok): window indices of (x_s, y_s, z_s)
    double lw, tw, uw;  // rw = 1 - lw, bw = 1 - tw, dw = 1 - uw: formed in the blend
};

// CHK = AFF_INTERIOR: the caller has shown that every sample of the tile
// lies a voxel inside the volume and the window (affine_box_interior), so the
// bounds tests, the window test and the last-column step are left out.
// CHK = AFF_BOUNDS: the window is the whole volume, so a sample inside the
// volume has its taps inside it (x_b = ceil(p) <= n - 1) and only the bounds
// test and the last-column step remain; outside, the offset is OOB_OFF and
// the result 0 (the floor and conversions of such a sample are not used).
// The arithmetic of the samples is the same at every level.
template <typename T, int INTERP, int CHK = AFF_FULL>
__device__ __forceinline__ ATap affine_tap(const AffineArgs& a, const AffGeom& g, double gx, double gy, double gz) {
    ATap t;
    const double p0 = a.m[0] * gx + a.m[1] * gy + a.m[2] * gz + a.m[3] * 1.0;
    const double p1 = a.m[4] * gx + a.m[5] * gy + a.m[6] * gz + a.m[7] * 1.0;
    const double p2 = a.m[8] * gx + a.m[9] * gy + a.m[10] * gz + a.m[11] * 1.0;
    if (CHK == AFF_BOUNDS) {
        const bool in = !(p0 < 0 || p0 > g.fmx || p1 < 0 || p1 > g.fmy || p2 < 0 || p2 > g.fmz);
        t.ok = in;
        if (INTERP == TW_INTERP_NN) {
            // Inside, p + 0.5 < n: the reference's clamp to n - 1 is a no-op.
            const uint32_t qx = (uint32_t)(int)(p0 + 0.5), qy = (uint32_t)(int)(p1 + 0.5),
                           qz = (uint32_t)(int)(p2 + 0.5);
            t.ex = (int)qx;
            t.ey = (int)qy;
            t.ez = (int)qz;
            t.shift = false;
            t.off = in ? (__umul24(qz * (uint32_t)g.sny + qy, (uint32_t)g.snx) + qx) * (uint32_t)sizeof(T) : OOB_OFF;
            t.lw = t.tw = t.uw = 0.0;
            return t;
        }
        const double fx = floor(p0), fy = floor(p1), fz = floor(p2);
        t.lw = 1.0 - (p0 - fx);
        t.tw = 1.0 - (p1 - fy);
        t.uw = 1.0 - (p2 - fz);
        const uint32_t xs = (uint32_t)(int)fx, ys = (uint32_t)(int)fy, zs = (uint32_t)(int)fz;
        t.ex = (int)xs;
        t.ey = (int)ys;
        t.ez = (int)zs;
        t.shift = xs == (uint32_t)g.snx - 1u;
        const uint32_t e = __umul24(zs * (uint32_t)g.sny + ys, (uint32_t)g.snx) + xs - (t.shift ? 1u : 0u);
        t.off = in ? e * (uint32_t)sizeof(T) : OOB_OFF;
        return t;
    }
    if (CHK == AFF_INTERIOR) {
        t.ok = true;
        t.shift = false;
        if (INTERP == TW_INTERP_NN) {
            const int qx = (int)(p0 + 0.5) - g.sx0, qy = (int)(p1 + 0.5) - g.sy0, qz = (int)(p2 + 0.5) - g.sz0;
            t.ex = qx;
            t.ey = qy;
            t.ez = qz;
            t.off = (__umul24((uint32_t)(qz * g.sny + qy), (uint32_t)g.snx) + (uint32_t)qx) * (uint32_t)sizeof(T);
            t.lw = t.tw = t.uw = 0.0;
            return t;
        }
        // floor(p) as a double equals the reference's (double)(long)floor(p).
        const double fx = floor(p0), fy = floor(p1), fz = floor(p2);
        t.lw = 1.0 - (p0 - fx);
        t.tw = 1.0 - (p1 - fy);
        t.uw = 1.0 - (p2 - fz);
        const int lxs = (int)fx - g.sx0, lys = (int)fy - g.sy0, lzs = (int)fz - g.sz0;
        t.ex = lxs;
        t.ey = lys;
        t.ez = lzs;
        t.off = (__umul24((uint32_t)(lzs * g.sny + lys), (uint32_t)g.snx) + (uint32_t)lxs) * (uint32_t)sizeof(T);
        return t;
    }
    const bool in = !(p0 < 0 || p0 > g.fmx || p1 < 0 || p1 > g.fmy || p2 < 0 || p2 > g.fmz);
    // Outside: evaluate at 0 (keeps the conversions in range); the result is 0.
    const double q0 = in ? p0 : 0.0, q1 = in ? p1 : 0.0, q2 = in ? p2 : 0.0;
    t.shift = false;
    if (INTERP == TW_INTERP_NN) {
        int qx = (int)(q0 + 0.5), qy = (int)(q1 + 0.5), qz = (int)(q2 + 0.5);
        qx = (qx < g.fnx ? qx : g.fnx - 1) - g.sx0;
        qy = (qy < g.fny ? qy : g.fny - 1) - g.sy0;
        qz = (qz < g.fnz ? qz : g.fnz - 1) - g.sz0;
        const bool w = qx >= 0 && qx < g.snx && qy >= 0 && qy < g.sny && qz >= 0 && qz < g.snz;
        t.ok = in && w;
        t.off = t.ok ? (__umul24((uint32_t)(qz * g.sny + qy), (uint32_t)g.snx) + (uint32_t)qx) * (uint32_t)sizeof(T)
                     : OOB_OFF;
        t.lw = t.tw = t.uw = 0.0;
        return t;
    }
    const int xs = (int)floor(q0), xb = (int)ceil(q0);
    const int ys = (int)floor(q1), yb = (int)ceil(q1);
    const int zs = (int)floor(q2), zb = (int)ceil(q2);
    t.lw = 1.0 - (q0 - (double)xs);
    t.tw = 1.0 - (q1 - (double)ys);
    t.uw = 1.0 - (q2 - (double)zs);
    const int lxs = xs - g.sx0, lxb = xb - g.sx0, lys = ys - g.sy0, lyb = yb - g.sy0;
    const int lzs = zs - g.sz0, lzb = zb - g.sz0;
    const bool w = !(lxs < 0 || lxb >= g.snx || lys < 0 || lyb >= g.sny || lzs < 0 || lzb >= g.snz);
    t.ok = in && w;
    t.shift = lxs == g.snx - 1;
    const uint32_t e = __umul24((uint32_t)(lzs * g.sny + lys), (uint32_t)g.snx) + (uint32_t)(lxs - (t.shift ? 1 : 0));
    t.off = t.ok ? e * (uint32_t)sizeof(T) : OOB_OFF;
    return t;
}

// The reference's blend of a tap (q_warp_affine_tps.cpp:1418-1433), truncated.
template <typename T, int INTERP, int CHK = AFF_FULL>
__device__ __forceinline__ uint32_t affine_blend(const SrcRs& rs, const ATap& t) {
    const uint32_t off = t.off;
    __builtin_assume(off <= OOB_OFF);
    if (INTERP == TW_INTERP_NN) return (CHK == AFF_INTERIOR || t.ok) ? bld<T>(rs.r00, off) : 0u;
    const uint32_t p00 = bld2<T>(rs.r00, off), p10 = bld2<T>(rs.r01, off);
    const uint32_t p01 = bld2<T>(rs.r10, off), p11 = bld2<T>(rs.r11, off);
    const bool sh = CHK != AFF_INTERIOR && t.shift;
    const double v000 = (double)(sh ? pair_hi<T>(p00) : pair_lo<T>(p00)), v100 = (double)pair_hi<T>(p00);
    const double v010 = (double)(sh ? pair_hi<T>(p10) : pair_lo<T>(p10)), v110 = (double)pair_hi<T>(p10);
    const double v001 = (double)(sh ? pair_hi<T>(p01) : pair_lo<T>(p01)), v101 = (double)pair_hi<T>(p01);
    const double v011 = (double)(sh ? pair_hi<T>(p11) : pair_lo<T>(p11)), v111 = (double)pair_hi<T>(p11);
    const double rw = 1.0 - t.lw, bw = 1.0 - t.tw, dw = 1.0 - t.uw;
    const double hi = t.tw * (t.lw * v000 + rw * v100) + bw * (t.lw * v010 + rw * v110);
    const double lo = t.tw * (t.lw * v001 + rw * v101) + bw * (t.lw * v011 + rw * v111);
    return (CHK == AFF_INTERIOR || t.ok) ? (uint32_t)(T)(t.uw * hi + dw * lo) : 0u;
}

// Whether no sample of the output box [lo, hi] lies inside the volume: along
// some axis the box's image (its corners' extremes) is more than half a voxel
// outside [0, n - 1], far beyond the rounding of the per-voxel evaluation.
// Every output voxel of such a box is 0.
__device__ __forceinline__ bool affine_box_outside(const AffineArgs& a, const AffGeom& g, const int* lo,
                                                   const int* hi) {
    const double c0[3] = {(double)(a.dx0 + lo[0]), (double)(a.dy0 + lo[1]), (double)(a.dz0 + lo[2])};
    const double c1[3] = {(double)(a.dx0 + hi[0]), (double)(a.dy0 + hi[1]), (double)(a.dz0 + hi[2])};
    const double fm[3] = {g.fmx, g.fmy, g.fmz};
    bool out = false;
    #pragma unroll
    for (int c = 0; c < 3; c++) {
        double pl = a.m[4 * c + 3], ph = a.m[4 * c + 3];
        #pragma unroll
        for (int k = 0; k < 3; k++) {
            const double u = a.m[4 * c + k] * c0[k], v = a.m[4 * c + k] * c1[k];
            pl += fmin(u, v);
            ph += fmax(u, v);
        }
        out = out || ph < -0.5 || pl > fm[c] + 0.5;
    }
    return out;
}

// Whether every sample of the output box [lo, hi] (destination window
// coordinates) lies a voxel inside the volume and the source window, with
// the last column excluded: the map is affine, so each component's extremes
// over the box sit at its corners; the one-voxel margin covers the rounding
// of the per-voxel evaluation.
__device__ __forceinline__ bool affine_box_interior(const AffineArgs& a, const AffGeom& g, const int* lo,
                                                    const int* hi) {
    const double c0[3] = {(double)(a.dx0 + lo[0]), (double)(a.dy0 + lo[1]), (double)(a.dz0 + lo[2])};
    const double c1[3] = {(double)(a.dx0 + hi[0]), (double)(a.dy0 + hi[1]), (double)(a.dz0 + hi[2])};
    const int s0[3] = {g.sx0, g.sy0, g.sz0}, sn[3] = {g.snx, g.sny, g.snz};
    const double fm[3] = {g.fmx, g.fmy, g.fmz};
    bool ok = true;
    #pragma unroll
    for (int c = 0; c < 3; c++) {
        double pl = a.m[4 * c + 3], ph = a.m[4 * c + 3];
        #pragma unroll
        for (int k = 0; k < 3; k++) {
            const double u = a.m[4 * c + k] * c0[k], v = a.m[4 * c + k] * c1[k];
            pl += fmin(u, v);
            ph += fmax(u, v);
        }
        ok = ok && pl - 1.0 >= fmax(0.0, (double)s0[c]) && ph + 1.0 <= fmin(fm[c], (double)(s0[c] + sn[c] - 2));
    }
    return ok;
}

__device__ __forceinline__ SrcRs affine_src_rs(const void* src, const AffGeom& g) {
    const char* p = (const char*)src;
    SrcRs r;
    r.r00 = __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)g.chan_b, BUF_FLAGS);
    r.r01 = __builtin_amdgcn_make_buffer_rsrc((void*)(p + g.row_b), (short)0, (int)g.nrec01, BUF_FLAGS);
    r.r10 = __builtin_amdgcn_make_buffer_rsrc((void*)(p + g.plane_b), (short)0, (int)g.nrec10, BUF_FLAGS);
    r.r11 = __builtin_amdgcn_make_buffer_rsrc((void*)(p + g.row_b + g.plane_b), (short)0, (int)g.nrec11, BUF_FLAGS);
    return r;
}

// One 64 x 4 x AFF_TZ tile of the output (x-fastest tile index), channels
// [c0, c1): each thread walks its (x, y) column four slices per step.
template <typename T, int INTERP>
__device__ __forceinline__ void affine_tile_fast(const AffineArgs& a, const AffGeom& g, int tile, int c0, int c1) {
    const int tix = tile % g.tiles_x, tiy = (tile / g.tiles_x) % g.tiles_y, tiz = tile / (g.tiles_x * g.tiles_y);
    const int x = tix * TX + (int)threadIdx.x, y = tiy * FTY + (int)threadIdx.y;
    const int dnx = (int)a.dnx, dny = (int)a.dny, dnz = (int)a.dnz;
    if (x >= dnx || y >= dny) return;
    const int z0 = tiz * AFF_TZ, z1 = min(z0 + AFF_TZ, dnz);
    // The tile's box (clipped to the destination): wave-uniform.
    const int blo[3] = {tix * TX, tiy * FTY, z0};
    const int bhi[3] = {min(blo[0] + TX, dnx) - 1, min(blo[1] + FTY, dny) - 1, z1 - 1};
    const uint32_t ostep = (uint32_t)dnx * (uint32_t)dny;
    if (affine_box_outside(a, g, blo, bhi)) {
        for (int c = c0; c < c1; c++) {
            const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(
                (uint8_t*)a.dst + (size_t)c * g.dchan_b, (short)0, (int)g.dchan_b, BUF_FLAGS);
            for (int z = z0; z < z1; z++)
                bst<T>(rd, ((uint32_t)z * ostep + (uint32_t)y * (uint32_t)dnx + (uint32_t)x) * (uint32_t)sizeof(T), 0u);
        }
        return;
    }
    // Coordinates are below 2^24 (affine_geom): 32-bit conversions are exact.
    const double gx = (double)(int)(a.dx0 + x), gy = (double)(int)(a.dy0 + y);
    const int gz0 = (int)a.dz0;
    auto walk = [&](auto check) {
        constexpr int CHK = decltype(check)::value;
        for (int z = z0; z < z1; z += AFF_TG) {
            const int nz = min(AFF_TG, z1 - z);
            ATap t[AFF_TG];
            #pragma unroll
            for (int j = 0; j < AFF_TG; j++)
                t[j] = affine_tap<T, INTERP, CHK>(a, g, gx, gy, (double)(gz0 + z + min(j, nz - 1)));
            const uint32_t o = (uint32_t)z * ostep + (uint32_t)y * (uint32_t)dnx + (uint32_t)x;
            for (int c = c0; c < c1; c++) {
                const SrcRs rs = affine_src_rs((const uint8_t*)a.src + (size_t)c * g.chan_b, g);
                const __amdgpu_buffer_rsrc_t rd =
                    __builtin_amdgcn_make_buffer_rsrc((uint8_t*)a.dst + (size_t)c * g.dchan_b, (short)0,
                                                      (int)g.dchan_b, BUF_FLAGS);
                uint32_t v[AFF_TG];
                #pragma unroll
                for (int j = 0; j < AFF_TG; j++) v[j] = affine_blend<T, INTERP, CHK>(rs, t[j]);
                #pragma unroll
                for (int j = 0; j < AFF_TG; j++)
                    if (j < nz) bst<T>(rd, (o + (uint32_t)j * ostep) * (uint32_t)sizeof(T), v[j]);
            }
        }
    };
    if (affine_box_interior(a, g, blo, bhi)) walk(std::integral_constant<int, AFF_INTERIOR>{});
    else if (g.whole) walk(std::integral_constant<int, AFF_BOUNDS>{});
    else walk(std::integral_constant<int, AFF_FULL>{});
}

// The source box [blo, bhi] (window coordinates, inclusive) that every
// sample of the output box [lo, hi] reaches with its +1 taps, and whether
// every sample lies a voxel inside the volume and the window (as
// affine_box_interior; then the box lies inside the window).
__device__ __forceinline__ bool affine_box_reach(const AffineArgs& a, const AffGeom& g, const int* lo, const int* hi,
                                                 int* blo, int* bhi) {
    const double c0[3] = {(double)(a.dx0 + lo[0]), (double)(a.dy0 + lo[1]), (double)(a.dz0 + lo[2])};
    const double c1[3] = {(double)(a.dx0 + hi[0]), (double)(a.dy0 + hi[1]), (double)(a.dz0 + hi[2])};
    const int s0[3] = {g.sx0, g.sy0, g.sz0}, sn[3] = {g.snx, g.sny, g.snz};
    const double fm[3] = {g.fmx, g.fmy, g.fmz};
    bool ok = true;
    #pragma unroll
    for (int c = 0; c < 3; c++) {
        double pl = a.m[4 * c + 3], ph = a.m[4 * c + 3];
        #pragma unroll
        for (int k = 0; k < 3; k++) {
            const double u = a.m[4 * c + k] * c0[k], v = a.m[4 * c + k] * c1[k];
            pl += fmin(u, v);
            ph += fmax(u, v);
        }
        ok = ok && pl - 1.0 >= fmax(0.0, (double)s0[c]) && ph + 1.0 <= fmin(fm[c], (double)(s0[c] + sn[c] - 2));
        // (Only used when ok: then both bounds are inside the window.)
        blo[c] = ok ? (int)floor(pl - 1.0) - s0[c] : 0;
        bhi[c] = ok ? (int)floor(ph + 1.0) + 1 - s0[c] : 0;
    }
    return ok;
}

// The source box [blo, bhi] of the output box [lo, hi] clipped to the volume
// (whole-volume window; the box is not affine_box_outside): every sample
// inside the volume has x_s in [blo, bhi] and, where x_s < bhi, x_s + 1
// (likewise y and z).  A sample with x_s = n - 1 lies on the last column
// exactly (p = n - 1), where the reference's x_b = ceil(p) = x_s.
__device__ __forceinline__ void affine_box_clip(const AffineArgs& a, const AffGeom& g, const int* lo, const int* hi,
                                                int* blo, int* bhi) {
    const double c0[3] = {(double)(a.dx0 + lo[0]), (double)(a.dy0 + lo[1]), (double)(a.dz0 + lo[2])};
    const double c1[3] = {(double)(a.dx0 + hi[0]), (double)(a.dy0 + hi[1]), (double)(a.dz0 + hi[2])};
    const int sn[3] = {g.snx, g.sny, g.snz};
    #pragma unroll
    for (int c = 0; c < 3; c++) {
        double pl = a.m[4 * c + 3], ph = a.m[4 * c + 3];
        #pragma unroll
        for (int k = 0; k < 3; k++) {
            const double u = a.m[4 * c + k] * c0[k], v = a.m[4 * c + k] * c1[k];
            pl += fmin(u, v);
            ph += fmax(u, v);
        }
        blo[c] = (int)fmax(0.0, floor(pl - 1.0));
        bhi[c] = (int)fmin((double)(sn[c] - 1), floor(ph + 1.0) + 1.0);
    }
}

// Division by d for 0 <= i < 2^20 and d >= 1, as div_f (exact).
__device__ __forceinline__ int aff_div(int i, float inv) { return (int)__builtin_fmaf((float)i, inv, 0.5f * inv); }

// One 16 x 16 x 8 tile of the output (x-fastest tile index), channels
// [c0, c1), through LDS: the source box the tile's samples reach is staged by
// 16-byte row pieces (the box's x start aligned down to a piece of the
// window's row), then every thread walks its (x, y) column reading the eight
// taps from LDS, with affine_tap / affine_blend's arithmetic.  A tile whose
// samples all lie outside the volume stores zeros; one whose samples partly
// leave a whole-volume source stages the box clipped to the volume and tests
// each sample's bounds.  A tile whose samples leave a partial window, or
// whose box exceeds g.lds_bytes, takes global gathers instead.  lds:
// g.lds_bytes, 16-byte aligned.  Every thread of the work-group calls it (one
// barrier inside for the staged path).
template <typename T, int INTERP>
__device__ __forceinline__ void affine_tile_lds(const AffineArgs& a, const AffGeom& g, int tile, int c0, int c1,
                                                uint8_t* lds) {
    constexpr int EPP = 16 / (int)sizeof(T);  // elements per 16-byte piece
    // Tile coordinates in the geometry's order (ord0 fastest).
    const int n0 = g.ord0 == 0 ? g.tn[0] : (g.ord0 == 1 ? g.tn[1] : g.tn[2]);
    const int n1 = g.ord1 == 0 ? g.tn[0] : (g.ord1 == 1 ? g.tn[1] : g.tn[2]);
    const int i0 = tile % n0, i1 = (tile / n0) % n1, i2 = tile / (n0 * n1);
    const int tix = g.ord0 == 0 ? i0 : (g.ord1 == 0 ? i1 : i2);
    const int tiy = g.ord0 == 1 ? i0 : (g.ord1 == 1 ? i1 : i2);
    const int tiz = g.ord0 == 2 ? i0 : (g.ord1 == 2 ? i1 : i2);
    const int dnx = (int)a.dnx, dny = (int)a.dny, dnz = (int)a.dnz;
    const int tid = (int)(threadIdx.y * TX + threadIdx.x);
    const int lo[3] = {tix * ALX, tiy * ALY, tiz * ALZ};
    const int hi[3] = {min(lo[0] + ALX, dnx) - 1, min(lo[1] + ALY, dny) - 1, min(lo[2] + ALZ, dnz) - 1};
    const int x = lo[0] + (tid % ALX), y = lo[1] + tid / ALX;
    const bool live = x <= hi[0] && y <= hi[1];
    const uint32_t ostep = (uint32_t)dnx * (uint32_t)dny;
    const uint32_t o0 = (uint32_t)lo[2] * ostep + (uint32_t)y * (uint32_t)dnx + (uint32_t)x;
    if (affine_box_outside(a, g, lo, hi)) {  // (work-group uniform: no barrier skipped)
        if (!live) return;
        for (int c = c0; c < c1; c++) {
            const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(
                (uint8_t*)a.dst + (size_t)c * g.dchan_b, (short)0, (int)g.dchan_b, BUF_FLAGS);
            for (int z = lo[2]; z <= hi[2]; z++)
                bst<T>(rd, (o0 + (uint32_t)(z - lo[2]) * ostep) * (uint32_t)sizeof(T), 0u);
        }
        return;
    }
    // Coordinates are below 2^24 (affine_geom): 32-bit conversions are exact.
    const double gx = (double)(int)(a.dx0 + x), gy = (double)(int)(a.dy0 + y);
    const int gz0 = (int)a.dz0;
    int blo[3], bhi[3];
    const bool inside = affine_box_reach(a, g, lo, hi, blo, bhi);
    const bool clip = !inside && g.whole;
    if (clip) affine_box_clip(a, g, lo, hi, blo, bhi);
    // Box rows: the x range aligned down to a piece of the window row.
    const int bx0 = blo[0] & ~(EPP - 1);
    const int ppr = (bhi[0] - bx0) / EPP + 1, bny = bhi[1] - blo[1] + 1, bnz = bhi[2] - blo[2] + 1;
    const int RB = ppr * 16, PB = bny * RB;  // LDS row / plane bytes
    const bool staged = (inside || clip) && (i64)bnz * PB <= g.lds_bytes;
    if (!staged) {
        // Global gathers (the 64 x 4 tile's arithmetic, this tile's voxels).
        if (!live) return;
        auto walk = [&](auto check) {
            constexpr int CHK = decltype(check)::value;
            for (int z = lo[2]; z <= hi[2]; z++) {
                const ATap t = affine_tap<T, INTERP, CHK>(a, g, gx, gy, (double)(gz0 + z));
                const uint32_t o = o0 + (uint32_t)(z - lo[2]) * ostep;
                for (int c = c0; c < c1; c++) {
                    const SrcRs rs = affine_src_rs((const uint8_t*)a.src + (size_t)c * g.chan_b, g);
                    const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(
                        (uint8_t*)a.dst + (size_t)c * g.dchan_b, (short)0, (int)g.dchan_b, BUF_FLAGS);
                    bst<T>(rd, o * (uint32_t)sizeof(T), affine_blend<T, INTERP, CHK>(rs, t));
                }
            }
        };
        if (inside) walk(std::integral_constant<int, AFF_INTERIOR>{});
        else if (g.whole) walk(std::integral_constant<int, AFF_BOUNDS>{});
        else walk(std::integral_constant<int, AFF_FULL>{});
        return;
    }
    const int npiece = bnz * bny * ppr;
    const float inv_ppr = 1.0f / (float)ppr, inv_bny = 1.0f / (float)bny;
    for (int c = c0; c < c1; c++) {
        const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
            (void*)((const uint8_t*)a.src + (size_t)c * g.chan_b), (short)0, (int)g.chan_b, BUF_FLAGS);
        const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc((uint8_t*)a.dst + (size_t)c * g.dchan_b,
                                                                            (short)0, (int)g.dchan_b, BUF_FLAGS);
        if (c > c0) __syncthreads();  // the previous channel's taps are read
        // Stage: piece i = (row, k), row = (z, y) of the box.  A piece past
        // the row's end reads the bytes after it into LDS bytes no tap
        // reads; one that crosses the channel's end is read byte by byte.
        for (int i = tid; i < npiece; i += 256) {
            const int row = aff_div(i, inv_ppr), k = i - row * ppr;
            const int zz = aff_div(row, inv_bny), yy = row - zz * bny;
            const uint32_t e = __umul24((uint32_t)((blo[2] + zz) * g.sny + blo[1] + yy), (uint32_t)g.snx) +
                               (uint32_t)(bx0 + k * EPP);
            const uint32_t at = e * (uint32_t)sizeof(T);
            const u32x4 v = at + 16u <= g.chan_b ? __builtin_amdgcn_raw_buffer_load_b128(rsrc, (int)at, 0, 0)
                                                 : load16_tail(rsrc, at);
            *(u32x4*)(lds + i * 16) = v;
        }
        __syncthreads();
        if (!live) continue;
        // Taps from LDS.  AFF_BOUNDS: a sample outside the volume is 0 and
        // reads LDS byte 0; the +1 neighbour of a tap on the box's last
        // column / row / plane is the tap itself (that sample lies on the
        // volume's last one, where the reference's ceil(p) = floor(p)).
        auto walk = [&](auto lvl) {
            constexpr int CHK = decltype(lvl)::value;
            constexpr int X1 = (int)sizeof(T);
            for (int z = lo[2]; z <= hi[2]; z++) {
                const ATap t = affine_tap<T, INTERP, CHK>(a, g, gx, gy, (double)(gz0 + z));
                int idx = (t.ez - blo[2]) * PB + (t.ey - blo[1]) * RB + (t.ex - bx0) * X1;
                int dx = X1, dy = RB, dz = PB;
                if (CHK != AFF_INTERIOR) {
                    idx = t.ok ? idx : 0;
                    dx = t.ok && t.ex < bhi[0] ? X1 : 0;
                    dy = t.ok && t.ey < bhi[1] ? RB : 0;
                    dz = t.ok && t.ez < bhi[2] ? PB : 0;
                }
                const uint8_t* q = lds + idx;
                auto at = [&](int off) -> uint32_t { return sizeof(T) == 1 ? (uint32_t)q[off] : (uint32_t)*(const uint16_t*)(q + off); };
                uint32_t v;
                if (INTERP == TW_INTERP_NN) {
                    v = at(0);
                } else {
                    const double v000 = (double)at(0), v100 = (double)at(dx);
                    const double v010 = (double)at(dy), v110 = (double)at(dy + dx);
                    const double v001 = (double)at(dz), v101 = (double)at(dz + dx);
                    const double v011 = (double)at(dz + dy), v111 = (double)at(dz + dy + dx);
                    const double rw = 1.0 - t.lw, bw = 1.0 - t.tw, dw = 1.0 - t.uw;
                    const double hi_ = t.tw * (t.lw * v000 + rw * v100) + bw * (t.lw * v010 + rw * v110);
                    const double lo_ = t.tw * (t.lw * v001 + rw * v101) + bw * (t.lw * v011 + rw * v111);
                    v = (uint32_t)(T)(t.uw * hi_ + dw * lo_);
                }
                if (CHK != AFF_INTERIOR) v = t.ok ? v : 0u;
                bst<T>(rd, (o0 + (uint32_t)(z - lo[2]) * ostep) * (uint32_t)sizeof(T), v);
            }
        };
        if (inside) walk(std::integral_constant<int, AFF_INTERIOR>{});
        else walk(std::integral_constant<int, AFF_BOUNDS>{});
    }
}

}  // namespace
