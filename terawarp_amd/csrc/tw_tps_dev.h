// tw_tps_dev.h -- device code of the fused B-spline upsample + resample,
// shared by the single-volume kernels (k_tps.hip) and the batch kernel
// (k_batch.hip).  Reference and semantics: see k_tps.hip's header.
#pragma once
#include "tw_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

namespace {


struct TpsArgs {
    const void* src;
    void* dst;
    i64 snx, sny, snz, sx0, sy0, sz0;  // source window
    i64 fnx, fny, fnz;                 // full subject extent
    i64 dnx, dny, dnz, dx0, dy0, dz0;  // destination window (global origin)
    i64 nc;
    i64 src_cstride, dst_cstride;      // bytes between channels (whole buffers)
    const void* grid;                  // 3 planes, float or half
    i64 gx, gy, gz, ox, oy, oz;
    int f;
    int basis;                         // TW_BASIS_BSPLINE / TW_BASIS_LINEAR (weight tables)
    int tz;                            // voxels per work-group along z
    const float* qtab;                 // FP16: per 8^3 node block (off xyz, scale xyz)
    i64 qbx, qby;
    const float* w32;                  // f x 4 FP32 weights
    const double* w64;                 // f x 4 FP64 weights TB
    float* dfout;                      // WRITE_DF: 3 planes of dnx*dny*dnz
    i64 zbase;                         // first z block of this launch (exact kernel)
    i64 px0, py0, pz0;                 // exact kernel: origin of the sample-position sum
};

constexpr int TX = 64, TY = 4;  // reference-order kernel block
// Fused kernel tile: TX x FTY threads, each walking one (x, y) column over
// TZ_CELLS grid cells of z.
#ifndef TW_FUSED_TY
#define TW_FUSED_TY 4
#endif
#ifndef TW_TZ_CELLS
#define TW_TZ_CELLS 4
#endif
#ifndef TW_FUSED_RY
#define TW_FUSED_RY 2
#endif
// FRY columns per thread: a tile is TX x FTH voxels in (x, y).
constexpr int FTY = TW_FUSED_TY, FT = TX * FTY, FRY = TW_FUSED_RY, FTH = FTY * FRY;
// Slices per group of the fused tile's z walk: the group's gathers are all
// issued before its blends.
#ifndef TW_FUSED_TG
#define TW_FUSED_TG 4
#endif
constexpr int TG = TW_FUSED_TG;

// ---------------------------------------------------------------------------
// Shared per-voxel resample tail.

template <typename T>
__device__ __forceinline__ double ld(const T* s, i64 i) {
    return (double)s[i];
}

// Reference blend in FP64 (q_imgwarp_tps_quicksmallmemory.cpp:133-160).
// The sample position is a float sum.  In-memory (origin 0) it is
// pos_warp + sx (106-108); the tera block warp adds the block's offset
// afterwards, (pos_warp + sx) + x_offset with pos_warp block-local, each
// step rounded to float (q_imgblockwarp_CYF, 2229-2235).  Adding the 0
// origin changes nothing, so one expression serves both.
template <typename T, int INTERP>
__device__ __forceinline__ void resample_fp64(const TpsArgs& a, i64 X, i64 Y, i64 Z, float d0, float d1, float d2,
                                              i64 o, i64 dvol) {
    const T* __restrict__ src = (const T*)a.src;
    T* __restrict__ dst = (T*)a.dst;
    const float q0 = (float)(X - a.px0) + d0, q1 = (float)(Y - a.py0) + d1, q2 = (float)(Z - a.pz0) + d2;
    const double p0 = q0 + (float)a.px0, p1 = q1 + (float)a.py0, p2 = q2 + (float)a.pz0;
    if (p0 < 0 || p0 > (double)(a.fnx - 1) || p1 < 0 || p1 > (double)(a.fny - 1) || p2 < 0 ||
        p2 > (double)(a.fnz - 1)) {
        for (i64 c = 0; c < a.nc; c++) dst[c * dvol + o] = 0;
        return;
    }
    const i64 sxy = a.snx * a.sny, svol = sxy * a.snz;
    if (INTERP == TW_INTERP_NN) {
        i64 qx = (i64)(p0 + 0.5), qy = (i64)(p1 + 0.5), qz = (i64)(p2 + 0.5);
        qx = (qx < a.fnx ? qx : a.fnx - 1) - a.sx0;
        qy = (qy < a.fny ? qy : a.fny - 1) - a.sy0;
        qz = (qz < a.fnz ? qz : a.fnz - 1) - a.sz0;
        const bool in = qx >= 0 && qx < a.snx && qy >= 0 && qy < a.sny && qz >= 0 && qz < a.snz;
        for (i64 c = 0; c < a.nc; c++) dst[c * dvol + o] = in ? src[c * svol + qz * sxy + qy * a.snx + qx] : (T)0;
        return;
    }
    const i64 xs = (i64)floor(p0), xb = (i64)ceil(p0);
    const i64 ys = (i64)floor(p1), yb = (i64)ceil(p1);
    const i64 zs = (i64)floor(p2), zb = (i64)ceil(p2);
    const double lw = 1.0 - (p0 - (double)xs), rw = 1.0 - lw;
    const double tw = 1.0 - (p1 - (double)ys), bw = 1.0 - tw;
    const double uw = 1.0 - (p2 - (double)zs), dw = 1.0 - uw;
    const i64 lxs = xs - a.sx0, lxb = xb - a.sx0, lys = ys - a.sy0, lyb = yb - a.sy0, lzs = zs - a.sz0,
              lzb = zb - a.sz0;
    if (lxs < 0 || lxb >= a.snx || lys < 0 || lyb >= a.sny || lzs < 0 || lzb >= a.snz) {
        for (i64 c = 0; c < a.nc; c++) dst[c * dvol + o] = 0;
        return;
    }
    const i64 r00 = lzs * sxy + lys * a.snx, r01 = lzs * sxy + lyb * a.snx;
    const i64 r10 = lzb * sxy + lys * a.snx, r11 = lzb * sxy + lyb * a.snx;
    for (i64 c = 0; c < a.nc; c++) {
        const T* s = src + c * svol;
        const double hi = tw * (lw * ld(s, r00 + lxs) + rw * ld(s, r00 + lxb)) +
                          bw * (lw * ld(s, r01 + lxs) + rw * ld(s, r01 + lxb));
        const double lo = tw * (lw * ld(s, r10 + lxs) + rw * ld(s, r10 + lxb)) +
                          bw * (lw * ld(s, r11 + lxs) + rw * ld(s, r11 + lxb));
        const double iv = uw * hi + dw * lo + 0.5;
        dst[c * dvol + o] = (T)iv;
    }
}

// ---------------------------------------------------------------------------
// Fused separable kernel (TW_PLAIN / TW_MIXED).
//
// Addressing: 32-bit byte offsets through buffer resources (one per
// channel, num_records = the channel's bytes), so one VGPR offset serves
// the eight taps: the x+1 taps use the instruction's immediate offset and
// the y+1 / z+1 taps a row / plane step that is zero on the last row /
// plane (where the tap's weight is zero anyway).  An x+1 tap past the end
// of the channel is caught by the buffer range check and reads 0, again
// under a zero weight.  The host only takes this path when every window
// fits in 32-bit byte offsets.

constexpr int BUF_FLAGS = 0x00020000;  // raw buffer, gfx9-family data format

// Upper bound of the node count one tile spans along an axis.
__host__ __device__ inline int node_span(int tile, int f) { return (tile - 1) / f + 5; }

template <typename T>
__device__ __forceinline__ uint32_t bld(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    if (sizeof(T) == 1) return (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(r, (int)off, 0, 0);
    return (uint32_t)(uint16_t)__builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0);
}

// The pair (x, x+1) in one access: bits 0-15 / 16-31 (uint16) or 0-7 /
// 8-15 (uint8).  The access may be unaligned; callers keep x+1 in the row.
template <typename T>
__device__ __forceinline__ uint32_t bld2(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    if (sizeof(T) == 1) return (uint32_t)(uint16_t)__builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0);
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0);
}
template <typename T>
__device__ __forceinline__ uint32_t pair_lo(uint32_t p) { return sizeof(T) == 1 ? (p & 0xffu) : (p & 0xffffu); }
template <typename T>
__device__ __forceinline__ uint32_t pair_hi(uint32_t p) { return sizeof(T) == 1 ? ((p >> 8) & 0xffu) : (p >> 16); }

template <typename T>
__device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t v) {
    if (sizeof(T) == 1) __builtin_amdgcn_raw_buffer_store_b8((uint8_t)v, r, (int)off, 0, 0);
    else __builtin_amdgcn_raw_buffer_store_b16((uint16_t)v, r, (int)off, 0, 0);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// 16 bytes at byte offset `at` of a buffer whose range ends inside them,
// byte by byte: a wider load that crosses the range's end returns 0 for the
// whole dword that crosses it, which would drop the channel's last bytes.
__device__ __noinline__ u32x4 load16_tail(__amdgpu_buffer_rsrc_t r, uint32_t at) {
    uint32_t w[4];
    #pragma unroll
    for (int q = 0; q < 4; q++) {
        w[q] = 0;
        #pragma unroll
        for (int b = 0; b < 4; b++)
            w[q] |= (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(r, (int)(at + 4 * q + b), 0, 0) << (8 * b);
    }
    return (u32x4){w[0], w[1], w[2], w[3]};
}

struct FastGeom {
    int snx, sny, snz, sx0, sy0, sz0;   // source window
    uint32_t row_b, plane_b, chan_b;    // source strides in bytes
    uint32_t nrec01, nrec10, nrec11;    // channel bytes past +row, +plane, +row+plane
    float fmx, fmy, fmz;                // full extent - 1 (bounds test)
    int imx, imy, imz;                  // the same as integers
    float invf;                         // 1 / f
    uint32_t dchan_b;                   // destination channel stride in bytes
    int tiles_x, tiles_y, ntiles, per_xcd;  // tile grid and XCD-major order
    int nxm, nym, nzm;                  // node-tile extents (maxima over the launch's tiles)
    float wt[4 * 64];                   // FP32 B-spline weights TB[t][k], f <= 64
};

// Tap geometry of one output voxel: byte offset of (x_s, y_s, z_s) in the
// channel and the exact fractional parts.  A sample that fails the bounds
// test gets OOB_OFF, past every buffer's range, so its taps read 0.
struct Tap {
    uint32_t off;
    float wx, wy, wz;
};
constexpr uint32_t OOB_OFF = 0x7FFF0000u;

// The source channel through four buffer resources based at +0, +row,
// +plane and +row+plane bytes, so the eight taps of a voxel share one
// offset register; each resource's range ends at the channel's end.
struct SrcRs {
    __amdgpu_buffer_rsrc_t r00, r01, r10, r11;
};

__device__ __forceinline__ SrcRs make_src_rs(const void* src, const FastGeom& g) {
    const char* p = (const char*)src;
    SrcRs r;
    r.r00 = __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)g.chan_b, BUF_FLAGS);
    r.r01 = __builtin_amdgcn_make_buffer_rsrc((void*)(p + g.row_b), (short)0, (int)g.nrec01, BUF_FLAGS);
    r.r10 = __builtin_amdgcn_make_buffer_rsrc((void*)(p + g.plane_b), (short)0, (int)g.nrec10, BUF_FLAGS);
    r.r11 = __builtin_amdgcn_make_buffer_rsrc((void*)(p + g.row_b + g.plane_b), (short)0, (int)g.nrec11, BUF_FLAGS);
    return r;
}

// floor(x) as int in one instruction (v_cvt_flr_i32_f32).
__device__ __forceinline__ int cvt_flr(float x) {
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// floor(x + 1/2) as int in one instruction (v_cvt_rpi_i32_f32; x >= 0 here).
__device__ __forceinline__ uint32_t cvt_rpi(float x) {
    int r;
    asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return (uint32_t)r;
}

// a * b + c on 24-bit operands (v_mad_u32_u24; b wave-uniform).
__device__ __forceinline__ uint32_t mad24(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
    return r;
}

// CHECK = false: the caller has shown that every sample of the cell lies
// inside the volume and the source window with a voxel to spare.
template <typename T, int INTERP, bool WIN, bool CHECK>
__device__ __forceinline__ Tap tap_setup(const FastGeom& g, float p0, float p1, float p2) {
    Tap t;
    bool ok = true;
    if (CHECK) ok = (p0 >= 0.f) & (p0 <= g.fmx) & (p1 >= 0.f) & (p1 <= g.fmy) & (p2 >= 0.f) & (p2 <= g.fmz);
    int ix, iy, iz;
    if (INTERP == TW_INTERP_NN) {
        // (long)(pos + 0.5), clamped to the extent; pos + 0.5 is exact in FP32
        // for positions below 2^23.
        ix = (int)fminf(floorf(p0 + 0.5f), g.fmx);
        iy = (int)fminf(floorf(p1 + 0.5f), g.fmy);
        iz = (int)fminf(floorf(p2 + 0.5f), g.fmz);
        t.wx = t.wy = t.wz = 0.f;
    } else {
        // pos - floor(pos) is exact for the non-negative in-bounds positions.
        ix = cvt_flr(p0);
        iy = cvt_flr(p1);
        iz = cvt_flr(p2);
        t.wx = __builtin_amdgcn_fractf(p0);
        t.wy = __builtin_amdgcn_fractf(p1);
        t.wz = __builtin_amdgcn_fractf(p2);
    }
    if (WIN) {
        ix -= g.sx0;
        iy -= g.sy0;
        iz -= g.sz0;
    }
    if (CHECK && INTERP != TW_INTERP_NN) {
        // A sample on the window's last column has weight 0 on x+1: take the
        // pair (x-1, x) with weight 1 instead, so no pair leaves its row.
        // The blend is unchanged: lerp(v(x-1), v(x), 1) = v(x) exactly.
        const bool last = ix == g.snx - 1;
        ix -= last ? 1 : 0;
        t.wx = last ? 1.0f : t.wx;
    }
    if (WIN) {
        if (CHECK) {
            if (INTERP == TW_INTERP_NN)
                ok &= (ix >= 0) & (iy >= 0) & (iz >= 0) & (ix < g.snx) & (iy < g.sny) & (iz < g.snz);
            else
                ok &= (ix >= 0) & (iy >= 0) & (iz >= 0) & (ix + (t.wx > 0.f) < g.snx) & (iy + (t.wy > 0.f) < g.sny) &
                      (iz + (t.wz > 0.f) < g.snz);
        }
    }
    const uint32_t e = __umul24(mad24((uint32_t)iz, (uint32_t)g.sny, (uint32_t)iy), (uint32_t)g.snx) + (uint32_t)ix;
    t.off = e * (uint32_t)sizeof(T);
    if (CHECK) t.off = ok ? t.off : OOB_OFF;
    return t;
}

// TW_MIXED blend.  The x step of each (x, x+1) pair is exact integer
// arithmetic on the matrix-free dot unit: the pair as two 16-bit lanes
// (v(x), v(x+1)) against the weights (2^15 - q, q), q = round(wx * 2^15),
// in one v_dot2_u32_u16 (below 65535 * 2^15, so no overflow).  The y and z
// steps are FP32 lerps of those sums, and the result is scaled back by 2^-15
// and rounded half up.  q carries the x fraction to 2^-16 voxel, far inside
// the FP16 field's own error.
typedef unsigned short tw_u16x2 __attribute__((ext_vector_type(2)));

// (2^15 - q) | q << 16: fma(wx, 2^15, 2^23) holds q in its low 24 bits for
// wx in [0, 1] (round to nearest even), which the 24-bit multiply reads.
__device__ __forceinline__ uint32_t wx_pair(float wx) {
    const uint32_t q = __float_as_uint(__builtin_fmaf(wx, 32768.0f, 8388608.0f)) & 0xFFFFFFu;
    return q * 0xFFFFu + 0x8000u;
}

// A pair as two 16-bit lanes: uint16 pairs already are; a uint8 pair (bits
// 0-7, 8-15) is spread to bits 0-7 and 16-23.
template <typename T>
__device__ __forceinline__ uint32_t pair16(uint32_t p) {
    if (sizeof(T) == 1) return __builtin_amdgcn_perm(0u, p, 0x0c010c00u);
    return p;
}

__device__ __forceinline__ float xlerp_q15(uint32_t p16, uint32_t w) {
    return (float)__builtin_amdgcn_udot2(__builtin_bit_cast(tw_u16x2, p16), __builtin_bit_cast(tw_u16x2, w), 0u,
                                         false);
}

__device__ __forceinline__ uint32_t blend_mixed(uint32_t q00, uint32_t q10, uint32_t q01, uint32_t q11,
                                                const Tap& t) {
    const uint32_t w = wx_pair(t.wx);
    const float a00 = xlerp_q15(q00, w), a10 = xlerp_q15(q10, w);
    const float a01 = xlerp_q15(q01, w), a11 = xlerp_q15(q11, w);
    const float b0 = __builtin_fmaf(t.wy, a10 - a00, a00);
    const float b1 = __builtin_fmaf(t.wy, a11 - a01, a01);
    return cvt_rpi(__builtin_fmaf(t.wz, b1 - b0, b0) * (1.0f / 32768.0f));
}

// Trilinear blend of the eight taps u[xyz] (x, y, z = 0 / 1).
template <int MODE>
__device__ __forceinline__ uint32_t blend8(uint32_t u000, uint32_t u100, uint32_t u010, uint32_t u110,
                                           uint32_t u001, uint32_t u101, uint32_t u011, uint32_t u111,
                                           const Tap& t) {
    if (MODE == TW_PLAIN) {
        // The reference's FP64 blend (q_imgwarp_tps_quicksmallmemory.cpp:141-160);
        // 1 - (pos - floor(pos)) is exact, so the weights equal the reference's.
        const double lw = 1.0 - (double)t.wx, rw = 1.0 - lw;
        const double tw = 1.0 - (double)t.wy, bw = 1.0 - tw;
        const double uw = 1.0 - (double)t.wz, dw = 1.0 - uw;
        const double hi = tw * (lw * (double)u000 + rw * (double)u100) + bw * (lw * (double)u010 + rw * (double)u110);
        const double lo = tw * (lw * (double)u001 + rw * (double)u101) + bw * (lw * (double)u011 + rw * (double)u111);
        return (uint32_t)(uw * hi + dw * lo + 0.5);
    }
    return blend_mixed(u000 | u100 << 16, u010 | u110 << 16, u001 | u101 << 16, u011 | u111 << 16, t);
}

// Trilinear blend of one voxel from its four (x, x+1) pairs: p00 / p10 the
// lower plane's rows y / y+1, p01 / p11 the upper plane's.  Equal to blend8
// on the unpacked taps.
template <typename T, int MODE>
__device__ __forceinline__ uint32_t blend_pairs(uint32_t p00, uint32_t p10, uint32_t p01, uint32_t p11,
                                                const Tap& t) {
    if (MODE == TW_PLAIN)
        return blend8<MODE>(pair_lo<T>(p00), pair_hi<T>(p00), pair_lo<T>(p10), pair_hi<T>(p10), pair_lo<T>(p01),
                            pair_hi<T>(p01), pair_lo<T>(p11), pair_hi<T>(p11), t);
    return blend_mixed(pair16<T>(p00), pair16<T>(p10), pair16<T>(p01), pair16<T>(p11), t);
}

// The eight taps of a voxel as four (x, x+1) pairs: y+1 / z+1 through the
// shifted resources.  Pairs of zero weight may land in the next plane or
// past the channel, where the range check returns 0.
template <typename T, int MODE, int INTERP>
__device__ __forceinline__ uint32_t blend(const SrcRs& rs, const Tap& t) {
    const uint32_t off = t.off;
    __builtin_assume(off <= OOB_OFF);
    if (INTERP == TW_INTERP_NN) return bld<T>(rs.r00, off);
    return blend_pairs<T, MODE>(bld2<T>(rs.r00, off), bld2<T>(rs.r01, off), bld2<T>(rs.r10, off),
                                bld2<T>(rs.r11, off), t);
}

// TG consecutive slices of one column (trilinear, every tap in bounds):
// where slice j's offset is slice j-1's plus one plane, its lower plane is
// slice j-1's upper plane (the same bytes), so those two pairs are not
// fetched again.  (have, poff, pu0, pu1) carry the previous group's last
// slice the same way.  Results are those of blend.
template <typename T, int MODE>
__device__ __forceinline__ void blend_shared(const SrcRs& rs, const Tap* tp, uint32_t plane_b, uint32_t* v,
                                              bool& have, uint32_t& poff, uint32_t& pu0, uint32_t& pu1) {
    uint32_t up0[TG], up1[TG], lo0[TG], lo1[TG];
    #pragma unroll
    for (int j = 0; j < TG; j++) {
        __builtin_assume(tp[j].off <= OOB_OFF);
        up0[j] = bld2<T>(rs.r10, tp[j].off);
        up1[j] = bld2<T>(rs.r11, tp[j].off);
    }
    // The lower-plane pairs are loaded by every lane, at a past-the-end
    // offset (reads 0) where the lane takes them from the slice before, and
    // selected after the loads: so no load waits for an earlier one.
    bool need[TG];
    uint32_t la[TG], lb[TG];
    need[0] = !(have && tp[0].off == poff + plane_b);
    #pragma unroll
    for (int j = 1; j < TG; j++) need[j] = tp[j].off != tp[j - 1].off + plane_b;
    #pragma unroll
    for (int j = 0; j < TG; j++) {
        const uint32_t off = need[j] ? tp[j].off : OOB_OFF;
        la[j] = bld2<T>(rs.r00, off);
        lb[j] = bld2<T>(rs.r01, off);
    }
    lo0[0] = need[0] ? la[0] : pu0;
    lo1[0] = need[0] ? lb[0] : pu1;
    #pragma unroll
    for (int j = 1; j < TG; j++) {
        lo0[j] = need[j] ? la[j] : up0[j - 1];
        lo1[j] = need[j] ? lb[j] : up1[j - 1];
    }
    have = true;
    poff = tp[TG - 1].off;
    pu0 = up0[TG - 1];
    pu1 = up1[TG - 1];
    #pragma unroll
    for (int j = 0; j < TG; j++) v[j] = blend_pairs<T, MODE>(lo0[j], lo1[j], up0[j], up1[j], tp[j]);
}

// floor(i / d) for 0 <= i < 2^21 as trunc((i + 1/2) / d) in FP32, with
// inv = fl(1/d), hinv = inv / 2: (i + 1/2) / d sits at least 1/(2d) from an
// integer and the FP32 evaluation errs by under 3 * 2^-24 * i / d < 3/(8d).
__device__ __forceinline__ int div_f(int i, float inv, float hinv) {
    return (int)__builtin_fmaf((float)i, inv, hinv);
}

__device__ __forceinline__ float wave_max(float v) {
    #pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

// One 64 x 4 x tz tile (index `tile` of the launch's x-fastest tile grid) of
// the fused kernel; src / dst: the channel's volumes.  lds: the node and
// x-collapse arrays (fused_lds_bytes), rmax_w: 6 * FTY floats.  Every thread of
// the work-group calls it (two work-group barriers inside).
template <typename T, int MODE, bool F16, int INTERP, bool WRITE_DF, bool WIN>
__device__ __forceinline__ void tps_tile(const TpsArgs& a, const FastGeom& g, int tile, const void* src, void* dst,
                                         float* lds, float* rmax_w) {
    const int f = a.f;
    const int NXM = g.nxm, NYM = g.nym, NZM = g.nzm;
    float* nodes = lds;                      // [3][NZM][NYM][NXM]
    float* xc = nodes + 3 * NZM * NYM * NXM; // [3][NZM][NYM][TX]

    const int lxi = threadIdx.x, lyi = threadIdx.y, tid = lyi * TX + lxi;
    const int dnx = (int)a.dnx, dny = (int)a.dny, dnz = (int)a.dnz;
    const int tix = tile % g.tiles_x, tiy = (tile / g.tiles_x) % g.tiles_y, tiz = tile / (g.tiles_x * g.tiles_y);
    const int tx0 = tix * TX, ty0 = tiy * FTH, tz0 = tiz * a.tz;
    const int tx1 = min(tx0 + TX, dnx) - 1, ty1 = min(ty0 + FTH, dny) - 1, tz1 = min(tz0 + a.tz, dnz) - 1;
    // Window-relative coordinates of the grid (r = global - origin >= 0).
    const int rx0 = (int)(a.dx0 - a.ox), ry0 = (int)(a.dy0 - a.oy), rz0 = (int)(a.dz0 - a.oz);
    const int nx0 = (rx0 + tx0) / f, ny0 = (ry0 + ty0) / f, nz0 = (rz0 + tz0) / f;
    const int nxn = (rx0 + tx1) / f + 4 - nx0;
    const int nyn = (ry0 + ty1) / f + 4 - ny0;
    const int nzn = (rz0 + tz1) / f + 4 - nz0;
    const float ivx = 1.0f / (float)nxn, ivy = 1.0f / (float)nyn;

    // Stage the tile's nodes (24-bit index arithmetic: the host keeps the
    // grid's planes below 2^31 values).
    const uint32_t plane = (uint32_t)(a.gx * a.gy * a.gz);
    const int NP = NZM * NYM * NXM;
    const int ntile = nxn * nyn * nzn;
    // Per component c: ext[c] = the largest -d_c, ext[3 + c] = the largest d_c
    // among this thread's nodes.
    float ext[6] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int i = tid; i < ntile; i += FT) {
        const int q = div_f(i, ivx, 0.5f * ivx), iz = div_f(q, ivy, 0.5f * ivy);
        const int ix = i - q * nxn, iy = q - iz * nyn;
        const uint32_t gi = mad24(mad24((uint32_t)(nz0 + iz), (uint32_t)a.gy, (uint32_t)(ny0 + iy)), (uint32_t)a.gx,
                                  (uint32_t)(nx0 + ix));
        const int li = (iz * NYM + iy) * NXM + ix;
        float v0, v1, v2;
        if (F16) {
            const __half* gh = (const __half*)a.grid;
            const uint32_t qb = mad24(mad24((uint32_t)((nz0 + iz) >> 3), (uint32_t)a.qby, (uint32_t)((ny0 + iy) >> 3)),
                                      (uint32_t)a.qbx, (uint32_t)((nx0 + ix) >> 3));
            const float* qt = a.qtab + qb * 6;
            v0 = __builtin_fmaf(qt[3], __half2float(gh[gi]), qt[0]);
            v1 = __builtin_fmaf(qt[4], __half2float(gh[plane + gi]), qt[1]);
            v2 = __builtin_fmaf(qt[5], __half2float(gh[2 * plane + gi]), qt[2]);
        } else {
            const float* gf = (const float*)a.grid;
            v0 = gf[gi];
            v1 = gf[plane + gi];
            v2 = gf[2 * plane + gi];
        }
        nodes[li] = v0;
        nodes[NP + li] = v1;
        nodes[2 * NP + li] = v2;
        ext[0] = fmaxf(ext[0], -v0);
        ext[1] = fmaxf(ext[1], -v1);
        ext[2] = fmaxf(ext[2], -v2);
        ext[3] = fmaxf(ext[3], v0);
        ext[4] = fmaxf(ext[4], v1);
        ext[5] = fmaxf(ext[5], v2);
    }
    #pragma unroll
    for (int k = 0; k < 6; k++) {
        ext[k] = wave_max(ext[k]);
        if (lxi == 0) rmax_w[lyi * 6 + k] = ext[k];
    }
    // Node index and phase of a column: base / f for the tile (scalar) plus
    // (base % f + lane) / f < 2 + 64 / f for the lane.
    auto split = [&](int base, int lane, int n0, int& b, int& ph) {
        const int q0 = base / f, r0 = base - q0 * f;
        const int rr = r0 + lane;
        const int qq = div_f(rr, g.invf, 0.5f * g.invf);
        b = q0 + qq - n0;
        ph = rr - qq * f;
    };
    __syncthreads();
    // Collapse along x for each of the tile's 64 columns.
    {
        int bxl, txp;
        split(rx0 + tx0, lxi, nx0, bxl, txp);
        const bool live = tx0 + lxi <= tx1;
        const float w0 = g.wt[txp * 4 + 0], w1 = g.wt[txp * 4 + 1], w2 = g.wt[txp * 4 + 2], w3 = g.wt[txp * 4 + 3];
        const int nstep = NYM * NXM, xstep = NYM * TX, xplane = NZM * NYM * TX;
        for (int iy = lyi; iy < nyn; iy += FTY) {
            const float* nrow = nodes + iy * NXM + bxl;
            float* xrow = xc + iy * TX + lxi;
            for (int iz = 0; iz < nzn; iz++) {
                #pragma unroll
                for (int c = 0; c < 3; c++) {
                    const float* nr = nrow + c * NP;
                    xrow[c * xplane] =
                        live ? __builtin_fmaf(w3, nr[3], __builtin_fmaf(w2, nr[2], __builtin_fmaf(w1, nr[1], w0 * nr[0])))
                             : 0.f;
                }
                nrow += nstep;
                xrow += xstep;
            }
        }
    }
    __syncthreads();
    if (tx0 + lxi > tx1) return;

    // Along axis a the samples of this tile lie in [X + dmin_a, X + dmax_a]
    // (+ one voxel for rounding), the tile's nodes bounding the field (a
    // convex combination of them).  Cells whose samples all stay a voxel
    // inside the volume and the source window skip the per-voxel bounds
    // test.  (Integer margins Rlo = ceil(-dmin) + 1 below and Rhi =
    // ceil(dmax) + 1 above the voxel, clamped past any coordinate, so the
    // tests below are scalar.)
    int Rlo[3], Rhi[3];
    #pragma unroll
    for (int c = 0; c < 3; c++) {
        float lo = rmax_w[c], hi = rmax_w[3 + c];
        #pragma unroll
        for (int w = 1; w < FTY; w++) {
            lo = fmaxf(lo, rmax_w[w * 6 + c]);
            hi = fmaxf(hi, rmax_w[w * 6 + 3 + c]);
        }
        Rlo[c] = __builtin_amdgcn_readfirstlane((int)ceilf(fminf(fmaxf(lo, -16777216.f), 16777216.f))) + 1;
        Rhi[c] = __builtin_amdgcn_readfirstlane((int)ceilf(fminf(fmaxf(hi, -16777216.f), 16777216.f))) + 1;
    }
    auto lo_bound = [](int s0) { return max(0, s0); };
    auto hi_bound = [](int im, int s0, int sn) { return min(im, s0 + sn - 1) - 1; };
    const int loz = lo_bound(g.sz0), hiz = hi_bound(g.imz, g.sz0, g.snz);
    const bool fast_xy = (int)a.dx0 + tx0 - Rlo[0] >= lo_bound(g.sx0) &&
                         (int)a.dx0 + tx1 + Rhi[0] <= hi_bound(g.imx, g.sx0, g.snx) &&
                         (int)a.dy0 + ty0 - Rlo[1] >= lo_bound(g.sy0) &&
                         (int)a.dy0 + ty1 + Rhi[1] <= hi_bound(g.imy, g.sy0, g.sny);
    // One channel per launch (the host offsets src / dst per channel).
    const SrcRs rs = make_src_rs(src, g);
    const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(dst, (short)0, (int)g.dchan_b, BUF_FLAGS);
    const int lx = tx0 + lxi;
    const float Xf = (float)((int)a.dx0 + lx);
    const uint32_t dvol = (uint32_t)dnx * (uint32_t)dny * (uint32_t)dnz;
    const int nzv = tz1 - tz0 + 1;
    const uint32_t ostep = (uint32_t)dnx * (uint32_t)dny;
    const float* xlane = xc + lxi;
    const int lstep = NYM * TX, cstep = NZM * NYM * TX;

    // The thread's FRY columns, rows lyi, lyi + FTY, ... of the tile, one
    // after the other; every column walks the tile's depth.
    for (int ry = lyi; ry < FTH; ry += FTY) {
        const int ly = ty0 + ry;
        if (ly > ty1) break;
        const float Yf = (float)((int)a.dy0 + ly);
        int byl, typ;
        split(ry0 + ty0, ry, ny0, byl, typ);
        const float wy0 = g.wt[typ * 4 + 0], wy1 = g.wt[typ * 4 + 1], wy2 = g.wt[typ * 4 + 2], wy3 = g.wt[typ * 4 + 3];

        // Layer k of the column: xcol + k * lstep (+ c * cstep).
        const float* xcol = xlane + byl * TX;
        auto layer = [&](int iz, float* L) {
            const float* col = xcol + iz * lstep;
            #pragma unroll
            for (int c = 0; c < 3; c++) {
                const float* cc = col + c * cstep;
                L[c] = __builtin_fmaf(wy3, cc[3 * TX],
                                      __builtin_fmaf(wy2, cc[2 * TX], __builtin_fmaf(wy1, cc[TX], wy0 * cc[0])));
            }
        };
        // z walk: the phase advances by one per slice, the node window by one
        // per f slices (wave-uniform).
        const int rz = rz0 + tz0;
        int bzl = rz / f - nz0;
        int tzp = rz - (rz / f) * f;
        float L0[3], L1[3], L2[3], L3[3];
        layer(bzl, L0); layer(bzl + 1, L1); layer(bzl + 2, L2); layer(bzl + 3, L3);
        uint32_t o = (uint32_t)tz0 * ostep + (uint32_t)ly * (uint32_t)dnx + (uint32_t)lx;
        float Zf = (float)((int)a.dz0 + tz0);

        auto df_ph = [&](float* d, int ph) {
            const float wz0 = g.wt[ph * 4 + 0], wz1 = g.wt[ph * 4 + 1], wz2 = g.wt[ph * 4 + 2],
                        wz3 = g.wt[ph * 4 + 3];
            #pragma unroll
            for (int c = 0; c < 3; c++)
                d[c] = __builtin_fmaf(wz3, L3[c], __builtin_fmaf(wz2, L2[c], __builtin_fmaf(wz1, L1[c], wz0 * L0[c])));
        };
        auto next_cell = [&]() {
            tzp = 0;
            bzl++;
            #pragma unroll
            for (int c = 0; c < 3; c++) { L0[c] = L1[c]; L1[c] = L2[c]; L2[c] = L3[c]; }
            layer(bzl + 3, L3);
        };

        if (WRITE_DF) {
            for (int lz = 0; lz < nzv; lz++) {
                float d[3];
                df_ph(d, tzp);
                a.dfout[o] = d[0];
                a.dfout[dvol + o] = d[1];
                a.dfout[2 * dvol + o] = d[2];
                o += ostep;
                if (++tzp == f && lz + 1 < nzv) next_cell();
            }
            continue;
        }
        // Cell by cell: the node-layer window moves once per cell; inside a
        // cell TG slices per iteration keep 8 TG taps in flight (a short last
        // group repeats its last slice and stores only the real ones).  The
        // last slice's upper-plane pairs carry to the next group.
        bool have = false;
        uint32_t poff = 0, pu0 = 0, pu1 = 0;
        auto group = [&](auto check, auto full, int nh) {
            constexpr bool CHECK = decltype(check)::value, FULL = decltype(full)::value;
            Tap tp[TG];
            #pragma unroll
            for (int j = 0; j < TG; j++) {
                const int jj = FULL ? j : min(j, nh - 1);
                float d[3];
                df_ph(d, tzp + jj);
                const float zj = FULL ? Zf + (float)j : Zf + (float)jj;
                tp[j] = tap_setup<T, INTERP, WIN, CHECK>(g, Xf + d[0], Yf + d[1], zj + d[2]);
            }
            uint32_t v[TG];
            if constexpr (!CHECK && INTERP != TW_INTERP_NN) {
                blend_shared<T, MODE>(rs, tp, g.plane_b, v, have, poff, pu0, pu1);
            } else {
                #pragma unroll
                for (int j = 0; j < TG; j++) v[j] = blend<T, MODE, INTERP>(rs, tp[j]);
                have = false;
            }
            #pragma unroll
            for (int j = 0; j < TG; j++)
                if (FULL || j < nh) bst<T>(rd, (o + (uint32_t)j * ostep) * (uint32_t)sizeof(T), v[j]);
            tzp += nh;
            o += (uint32_t)nh * ostep;
            Zf += (float)nh;
        };
        auto cell = [&](auto check, int cnt) {
            int k = 0;
            for (; k + TG <= cnt; k += TG) group(check, std::true_type{}, TG);
            if (k < cnt) group(check, std::false_type{}, cnt - k);
        };
        int lz = 0;
        while (true) {
            const int cnt = min(f - tzp, nzv - lz);
            const int zc = (int)a.dz0 + tz0 + lz;  // first slice of the cell
            if (fast_xy && zc - Rlo[2] >= loz && zc + cnt - 1 + Rhi[2] <= hiz)
                cell(std::false_type{}, cnt);
            else
                cell(std::true_type{}, cnt);
            lz += cnt;
            if (lz >= nzv) break;
            next_cell();
        }
    }
}

// ---- launch shaping (host) ----------------------------------------------------

inline size_t fused_lds_bytes(int f, int tz) {
    const int nxm = node_span(TX, f), nym = node_span(FTH, f), nzm = node_span(tz, f);
    return sizeof(float) * (size_t)(3 * nzm * nym * nxm + 3 * nzm * nym * TX);
}

// Largest node count a tile of `len` voxels spans along one axis over the
// tiles [t*len, min((t+1)*len, n)) of a window starting at grid offset r0.
inline int max_node_span(i64 r0, int len, i64 n, int f) {
    int best = 0;
    const i64 ntile = (n + len - 1) / len;
    // The span depends on the tile start modulo f and on the clipped last tile.
    const i64 period = f / std::__gcd(f, len);
    for (i64 t = 0; t < ntile && t < period; t++) {
        const i64 a = r0 + t * len, b = r0 + std::min<i64>((t + 1) * len, n) - 1;
        best = std::max(best, (int)(b / f - a / f + 4));
    }
    {
        const i64 t = ntile - 1;
        const i64 a = r0 + t * len, b = r0 + n - 1;
        best = std::max(best, (int)(b / f - a / f + 4));
    }
    return best;
}

// Voxels per work-group along z: `tuned` if given (halved until the LDS tile
// fits in 160 KiB), else 4 cells clamped to [16, 64], halved until the LDS
// tile fits in 64 KiB.
inline int pick_tz(int f, int tuned = 0) {
    int tz = tuned > 0 ? tuned : TW_TZ_CELLS * f;
    if (tuned <= 0) {
        if (tz < 16) tz = 16;
        if (tz > 64) tz = 64;
    }
    while (tz > 1 && fused_lds_bytes(f, tz) > (tuned > 0 ? 160 : 64) * 1024) tz /= 2;
    return tz;
}

// FP32 upsampling weights of spacing f and a basis into the geometry block.
inline void fill_weights(FastGeom& g, int f, int basis) {
    std::vector<double> tb((size_t)f * 4);
    tw_df_tb(f, basis, tb.data());
    for (size_t i = 0; i < tb.size(); i++) g.wt[i] = (float)tb[i];
}

// Geometry block of the fused kernel for the arguments a (a.tz set): node-tile
// extents, LDS bytes per work-group, strides, tile grid.  Returns null, or
// the limit the arguments break.
// (gather_lds = false: the caller runs another tile and sizes its own LDS.)
template <typename T, bool WRITE_DF>
const char* fused_geom(const TpsArgs& a, FastGeom& g, size_t& lds, bool gather_lds = true) {
    // Exact node-tile extents of this launch (tiles aligned to the grid need
    // fewer nodes than the worst case), which sets the LDS per work-group.
    const int nxm = max_node_span(a.dx0 - a.ox, TX, a.dnx, a.f);
    const int nym = max_node_span(a.dy0 - a.oy, FTH, a.dny, a.f);
    const int nzm = max_node_span(a.dz0 - a.oz, a.tz, a.dnz, a.f);
    lds = sizeof(float) * (size_t)(3 * nzm * nym * nxm + 3 * nzm * nym * TX);
    if (gather_lds && lds > 160 * 1024) return "the grid spacing needs more than 160 KiB of LDS per work-group";
    const i64 nbz = (a.dnz + a.tz - 1) / a.tz;
    // 32-bit addressing limits of the fused kernel: byte offsets of every tap
    // below 2 GiB (so the x+1 taps fold into the immediate offset), 24-bit
    // row indices and coordinates.
    const uint64_t schan = (uint64_t)a.snx * a.sny * a.snz * sizeof(T);
    const uint64_t dchan = (uint64_t)a.dnx * a.dny * a.dnz * (WRITE_DF ? 3 * sizeof(float) : sizeof(T));
    const uint64_t taps = schan + (uint64_t)a.snx * (a.sny + 1) * sizeof(T) + 16;  // offsets of every tap
    const i64 cmax = std::max({a.dx0 + a.dnx, a.dy0 + a.dny, a.dz0 + a.dnz, a.fnx, a.fny, a.fnz});
    if ((!WRITE_DF && (taps > 0x7FFF0000ull || a.sny * a.snz >= (1ll << 24))) || dchan > 0xF0000000ull ||
        cmax > (1ll << 24) || a.gy * a.gz >= (1ll << 24) || 3 * a.gx * a.gy * a.gz >= (1ll << 31))
        return "window exceeds the fused kernel's 32-bit addressing (split the volume into bricks)";
    memset(&g, 0, sizeof g);
    g.nxm = nxm; g.nym = nym; g.nzm = nzm;
    g.snx = (int)a.snx; g.sny = (int)a.sny; g.snz = (int)a.snz;
    g.sx0 = (int)a.sx0; g.sy0 = (int)a.sy0; g.sz0 = (int)a.sz0;
    g.row_b = (uint32_t)(a.snx * sizeof(T));
    g.plane_b = (uint32_t)(a.snx * a.sny * sizeof(T));
    g.chan_b = (uint32_t)schan;
    auto past = [&](uint64_t skip) { return (uint32_t)(schan > skip ? schan - skip : 0); };
    g.nrec01 = past(g.row_b);
    g.nrec10 = past(g.plane_b);
    g.nrec11 = past((uint64_t)g.row_b + g.plane_b);
    g.fmx = (float)(a.fnx - 1); g.fmy = (float)(a.fny - 1); g.fmz = (float)(a.fnz - 1);
    g.imx = (int)(a.fnx - 1); g.imy = (int)(a.fny - 1); g.imz = (int)(a.fnz - 1);
    g.invf = 1.0f / (float)a.f;
    g.dchan_b = (uint32_t)(a.dnx * a.dny * a.dnz * sizeof(T));
    fill_weights(g, a.f, a.basis);
    g.tiles_x = (int)((a.dnx + TX - 1) / TX);
    g.tiles_y = (int)((a.dny + FTH - 1) / FTH);
    const i64 ntiles = (i64)g.tiles_x * g.tiles_y * nbz;
    if (ntiles > (1ll << 30)) return "too many tiles for one launch";
    g.ntiles = (int)ntiles;
    g.per_xcd = (int)((ntiles + 7) / 8);
    return nullptr;
}

}  // namespace
