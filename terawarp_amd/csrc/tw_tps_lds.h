// tw_tps_lds.h -- fused B-spline upsample + trilinear resample with the
// source streamed through LDS (TW_PLAIN / TW_MIXED, linear image
// interpolation).  Results equal tps_tile's (tw_tps_dev.h) bit for bit: the
// same per-voxel displacement, sample position, bounds tests and blend; only
// where the eight taps come from differs.
//
// Why: a gather instruction costs the texture-address unit ~17-25 CU-cycles
// whatever its width and active lanes, and the gather tile issues four per
// voxel, which bounds it near 100 cycles per 64 voxels.  Here a work-group
// owns a column of 16 x 16 output voxels by SEG slices.  The field is a
// convex combination of the grid nodes, so the column's samples lie within
// [dmin, dmax] (its nodes' range per axis, plus a voxel of rounding and the
// +1 taps) of the column: a box of fixed x / y extent that slides along z.
// The box's planes live in an LDS ring of RING planes; the z walk proceeds
// in chunks of LCH slices, and while a chunk is resampled the planes the next
// chunk adds are loaded (16-byte row chunks, a few loads per 64 output
// voxels) and written to the ring after it.  The eight taps of a voxel are
// then eight LDS reads at immediate offsets from one address.  The node
// staging and x-collapse are paid once per column, not per chunk.  A column
// whose box exceeds the ring gathers its taps from global memory instead
// (same arithmetic).
#pragma once
#include "tw_tps_dev.h"

namespace {

constexpr int LTX = 16, LTY = 16;  // column footprint; the work-group is 64 x 4 threads
constexpr int LCH = 8;             // slices per chunk
constexpr int RING = 16;           // planes in the LDS ring (a power of two)

// Box pitches per dtype: X elements per row (CHK 16-byte chunks), Y rows per
// plane; RPI rows per staging instruction (RPI * CHK <= 64 lanes).
template <typename T>
struct BoxDims;
template <>
struct BoxDims<uint16_t> {
    static constexpr int X = 40, Y = 24, CHK = 5, RPI = 12;
};
template <>
struct BoxDims<uint8_t> {
    static constexpr int X = 48, Y = 32, CHK = 3, RPI = 21;
};

// LDS bytes of the ring: RING planes plus a copy of slot 0 after the last,
// so that the z+1 taps of slot RING - 1 sit at a fixed offset too, plus one
// row, which the zero-weight y+1 taps of that copy's last row read.
template <typename T>
constexpr uint32_t ring_bytes() {
    return (uint32_t)BoxDims<T>::X * (BoxDims<T>::Y * (RING + 1) + 1) * sizeof(T);
}
// Largest tap offset in the ring: slot RING - 1, row Y - 1, x X - 2.
template <typename T>
constexpr uint32_t ring_max_off() {
    return (uint32_t)(((RING - 1) * BoxDims<T>::Y + BoxDims<T>::Y - 1) * BoxDims<T>::X + BoxDims<T>::X - 2) * sizeof(T);
}

// Columns with displacements past LDS_DMAX or coordinates past LDS_CMAX take
// the gather path: the box bound leaves one voxel for the FP32 rounding of
// the field (~2^-20 of LDS_DMAX) and of the position (half an ulp below 2^23).
constexpr float LDS_DMAX = 65536.f;
constexpr int LDS_CMAX = 1 << 22;

// The voxel at byte offset `off` of the ring.  Element-aligned reads: a
// misaligned wider read of a tap pair stalls the LDS many times over, so the
// read is volatile, which keeps the compiler from merging neighbouring taps
// into one (it takes unaligned LDS access as supported).
template <typename T>
__device__ __forceinline__ uint32_t lds_at(const uint8_t* box, uint32_t off) {
    typedef const volatile __attribute__((address_space(3))) T lds_t;
    return *(lds_t*)(box + off);
}


// Floats of the x-collapse array (a multiple of 4, so that what follows it
// is 16-byte aligned).
__host__ __device__ inline int lds_xc_floats(int nym, int nzm) { return (3 * nzm * nym * LTX + 3) & ~3; }

// LDS bytes per work-group: the x-collapse array, the z weights (4 x 64
// floats: read from LDS, their loads stay in order with the tap reads), the
// output staging, then the ring and its trash area, which the column's nodes
// occupy before the ring is first filled.
constexpr int LDS_WT_FLOATS = 4 * 64;
// Staging writes of lanes without a chunk land in LDS_TRASH bytes after the
// ring (one 16-byte chunk per lane).
constexpr uint32_t LDS_TRASH = 64 * 16;
// Output staging: a chunk's LCH slices of the column's 16 x 16 voxels, [z][y][x].
template <typename T>
constexpr uint32_t obuf_bytes() {
    return (uint32_t)(LCH * LTY * LTX * sizeof(T));
}
template <typename T>
inline size_t lds_column_bytes(int nxm, int nym, int nzm) {
    const size_t nodes = (size_t)3 * nzm * nym * nxm * sizeof(float);
    return (size_t)(lds_xc_floats(nym, nzm) + LDS_WT_FLOATS) * sizeof(float) + obuf_bytes<T>() +
           std::max<size_t>(nodes, ring_bytes<T>() + LDS_TRASH);
}

// tap_setup<T, TW_INTERP_LINEAR, true, CHECK> with the tap's byte offset in
// the ring instead of the window: the same floor, fractions, last-column
// step and bounds tests.  bxy = by * RB + bx * sizeof(T) for the box's x / y
// origin (bx, by); plane z sits in ring slot z mod RING.  A failed sample
// gets offset 0 and ok = false; the offset is clamped to the ring, so no read
// leaves it whatever the position (NaN included).  (24-bit products: a
// sample inside the box has coordinates below 2^22.)
template <typename T, bool CHECK>
__device__ __forceinline__ Tap tap_ring(const FastGeom& g, float p0, float p1, float p2, uint32_t bxy, bool& ok) {
    using B = BoxDims<T>;
    constexpr uint32_t RB = B::X * sizeof(T), PB = B::Y * RB;
    Tap t;
    ok = true;
    if (CHECK) ok = (p0 >= 0.f) & (p0 <= g.fmx) & (p1 >= 0.f) & (p1 <= g.fmy) & (p2 >= 0.f) & (p2 <= g.fmz);
    int ix = cvt_flr(p0), iy = cvt_flr(p1), iz = cvt_flr(p2);
    t.wx = __builtin_amdgcn_fractf(p0);
    t.wy = __builtin_amdgcn_fractf(p1);
    t.wz = __builtin_amdgcn_fractf(p2);
    if (CHECK) {
        const bool last = ix - g.sx0 == g.snx - 1;
        ix -= last ? 1 : 0;
        t.wx = last ? 1.0f : t.wx;
        ok &= (ix >= g.sx0) & (iy >= g.sy0) & (iz >= g.sz0) & (ix - g.sx0 + (t.wx > 0.f) < g.snx) &
              (iy - g.sy0 + (t.wy > 0.f) < g.sny) & (iz - g.sz0 + (t.wz > 0.f) < g.snz);
    }
    uint32_t o = __umul24((uint32_t)iz & (RING - 1), PB) + __umul24((uint32_t)iy, RB) +
                 (uint32_t)ix * (uint32_t)sizeof(T) - bxy;
    o = min(o, ring_max_off<T>());
    t.off = CHECK ? (ok ? o : 0u) : o;
    return t;
}

// Geometry of the column kernel for the arguments a: the fused kernel's
// strides and limits (fused_geom), a.tz = SEG slices per column (the largest
// multiple of LCH up to 64 whose LDS fits lds_budget bytes), node-tile extents,
// column grid, LDS bytes per work-group.  Returns null, or the limit the
// arguments break.
template <typename T>
const char* fused_geom_lds(TpsArgs& a, FastGeom& g, size_t& lds, size_t lds_budget = 56 * 1024) {
    int seg = 64;
    for (;;) {
        const int nym = max_node_span(a.dy0 - a.oy, LTY, a.dny, a.f);
        const int nzm = max_node_span(a.dz0 - a.oz, seg, a.dnz, a.f);
        const int nxm = max_node_span(a.dx0 - a.ox, LTX, a.dnx, a.f);
        lds = lds_column_bytes<T>(nxm, nym, nzm);
        if (lds <= lds_budget || seg == LCH) break;
        seg -= LCH;
    }
    a.tz = seg;
    size_t unused = 0;
    if (const char* why = fused_geom<T, false>(a, g, unused, false)) return why;
    g.nxm = max_node_span(a.dx0 - a.ox, LTX, a.dnx, a.f);
    g.nym = max_node_span(a.dy0 - a.oy, LTY, a.dny, a.f);
    g.nzm = max_node_span(a.dz0 - a.oz, seg, a.dnz, a.f);
    if (lds > 160 * 1024) return "the grid spacing needs more than 160 KiB of LDS per work-group";
    g.tiles_x = (int)((a.dnx + LTX - 1) / LTX);
    g.tiles_y = (int)((a.dny + LTY - 1) / LTY);
    const i64 ntiles = (i64)g.tiles_x * g.tiles_y * ((a.dnz + seg - 1) / seg);
    if (ntiles > (1ll << 30)) return "too many tiles for one launch";
    g.ntiles = (int)ntiles;
    g.per_xcd = (int)((ntiles + 7) / 8);
    return nullptr;
}

// One column (index `tile` of the launch's x-fastest column grid; geometry
// from fused_geom_lds).  src / dst: the channel's volumes; lds: the
// x-collapse array and the ring; ext_w: 24 floats.  The work-group is 64 x 4
// threads, all of which call it and reach every work-group barrier inside.
template <typename T, int MODE, bool F16, bool WIN>
__device__ __forceinline__ void tps_column_lds(const TpsArgs& a, const FastGeom& g, int tile, const void* src,
                                               void* dst, float* lds, float* ext_w) {
    using B = BoxDims<T>;
    constexpr int BPV = (int)sizeof(T), RB = B::X * BPV, PB = B::Y * RB;
    constexpr int NW = 4;                             // waves per work-group
    constexpr int H = (B::Y + B::RPI - 1) / B::RPI;   // staging instructions per plane
    constexpr int KPF = (LCH * H + NW - 1) / NW;      // per wave for LCH planes
    const int f = a.f, SEG = a.tz;
    const int NXM = g.nxm, NYM = g.nym, NZM = g.nzm;
    float* xc = lds;                                               // [3][NZM][NYM][LTX]
    float* wts = lds + lds_xc_floats(NYM, NZM);                    // [f][4]
    T* obuf = (T*)(wts + LDS_WT_FLOATS);                           // [LCH][LTY][LTX] outputs of a chunk
    uint8_t* box = (uint8_t*)obuf + obuf_bytes<T>();               // ring [RING + 1][Y][X] of T
    float* nodes = (float*)box;                                    // [3][NZM][NYM][NXM] until the ring fills

    const int tid = threadIdx.y * TX + threadIdx.x, lane = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int lxi = tid % LTX, lyi = tid / LTX;
    const int dnx = (int)a.dnx, dny = (int)a.dny, dnz = (int)a.dnz;
    const int tix = tile % g.tiles_x, tiy = (tile / g.tiles_x) % g.tiles_y, tiz = tile / (g.tiles_x * g.tiles_y);
    const int tx0 = tix * LTX, ty0 = tiy * LTY, tz0 = tiz * SEG;
    const int tx1 = min(tx0 + LTX, dnx) - 1, ty1 = min(ty0 + LTY, dny) - 1, tz1 = min(tz0 + SEG, dnz) - 1;
    const int rx0 = (int)(a.dx0 - a.ox), ry0 = (int)(a.dy0 - a.oy), rz0 = (int)(a.dz0 - a.oz);
    const int nx0 = (rx0 + tx0) / f, ny0 = (ry0 + ty0) / f, nz0 = (rz0 + tz0) / f;
    const int nxn = (rx0 + tx1) / f + 4 - nx0;
    const int nyn = (ry0 + ty1) / f + 4 - ny0;
    const int nzn = (rz0 + tz1) / f + 4 - nz0;
    const float ivx = 1.0f / (float)nxn, ivy = 1.0f / (float)nyn;

    if (tid < 4 * f) wts[tid] = g.wt[tid];

    // 1. The column's nodes, and per component the largest -d and +d.
    const uint32_t plane = (uint32_t)(a.gx * a.gy * a.gz);
    const int NP = NZM * NYM * NXM;
    const int ntile = nxn * nyn * nzn;
    float ext[6] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int i = tid; i < ntile; i += NW * 64) {
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
        if (lane == 0) ext_w[wave * 6 + k] = ext[k];
    }
    __syncthreads();

    // 2. The column's displacement range per axis: fast-path margins (as in
    //    tps_tile), the box's x / y extent and the z reach of a chunk.
    int Rlo[3], Rhi[3];
    float dlo[3], dhi[3];
    bool small = true;
    #pragma unroll
    for (int c = 0; c < 3; c++) {
        float lo = ext_w[c], hi = ext_w[3 + c];
        #pragma unroll
        for (int w = 1; w < NW; w++) {
            lo = fmaxf(lo, ext_w[w * 6 + c]);
            hi = fmaxf(hi, ext_w[w * 6 + 3 + c]);
        }
        lo = fminf(fmaxf(lo, -16777216.f), 16777216.f);
        hi = fminf(fmaxf(hi, -16777216.f), 16777216.f);
        dlo[c] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, lo)));
        dhi[c] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, hi)));
        Rlo[c] = __builtin_amdgcn_readfirstlane((int)ceilf(lo)) + 1;
        Rhi[c] = __builtin_amdgcn_readfirstlane((int)ceilf(hi)) + 1;
        small &= fmaxf(fabsf(lo), fabsf(hi)) <= LDS_DMAX;
    }
    const int tl[3] = {(int)a.dx0 + tx0, (int)a.dy0 + ty0, (int)a.dz0 + tz0};
    const int th[3] = {(int)a.dx0 + tx1, (int)a.dy0 + ty1, (int)a.dz0 + tz1};
    const int wlo[3] = {max(0, g.sx0), max(0, g.sy0), max(0, g.sz0)};
    const int whi[3] = {g.sx0 + g.snx - 1, g.sy0 + g.sny - 1, g.sz0 + g.snz - 1};
    // Samples of voxels [u0, u1] along axis c lie in [u0 - lo, u1 + hi] up to
    // rounding: floor, one voxel of rounding below (two along x, for the
    // last-column step) and the +1 tap plus one above, clamped to the window.
    auto reach_lo = [&](int c, int u0) {
        return max(__builtin_amdgcn_readfirstlane((int)floorf((float)u0 - dlo[c])) - (c == 0 ? 2 : 1), wlo[c]);
    };
    auto reach_hi = [&](int c, int u1) {
        return min(__builtin_amdgcn_readfirstlane((int)floorf((float)u1 + dhi[c])) + 2, whi[c]);
    };
    // The box's x origin sits a multiple of 16 bytes from the window's, so
    // that the staging loads are 16-byte aligned where the window's rows are
    // (a misaligned 16-byte load costs the address unit several times over).
    const int bx0 = g.sx0 + (int)((uint32_t)((reach_lo(0, tl[0]) - g.sx0) * BPV) & ~15u) / BPV;
    const int by0 = reach_lo(1, tl[1]);
    const int bnx = reach_hi(0, th[0]) - bx0 + 1, bny = reach_hi(1, th[1]) - by0 + 1;
    // A chunk spans LCH slices plus the z range and four planes of margin.
    const int zspan = __builtin_amdgcn_readfirstlane((int)ceilf(dlo[2] + dhi[2])) + LCH + 5;
    const bool use_lds = small && th[0] < LDS_CMAX && th[1] < LDS_CMAX && th[2] < LDS_CMAX && bnx <= B::X &&
                         bny <= B::Y && zspan <= RING;
    const uint32_t bxy = (uint32_t)by0 * (uint32_t)RB + (uint32_t)bx0 * BPV;

    // 3. Collapse the nodes along x for each of the column's 16 x positions.
    auto split = [&](int base, int lanex, int n0, int& b, int& ph) {
        const int q0 = base / f, r0 = base - q0 * f;
        const int rr = r0 + lanex;
        const int qq = div_f(rr, g.invf, 0.5f * g.invf);
        b = q0 + qq - n0;
        ph = rr - qq * f;
    };
    {
        const int ncol = nzn * nyn * LTX;
        for (int e = tid; e < ncol; e += NW * 64) {
            const int col = e % LTX, r = e / LTX;
            const int iz = r / nyn, iy = r - iz * nyn;
            int bxl, txp;
            split(rx0 + tx0, col, nx0, bxl, txp);
            const bool live = tx0 + col <= tx1;
            bxl = live ? bxl : 0;
            const float w0 = g.wt[txp * 4 + 0], w1 = g.wt[txp * 4 + 1], w2 = g.wt[txp * 4 + 2], w3 = g.wt[txp * 4 + 3];
            #pragma unroll
            for (int c = 0; c < 3; c++) {
                const float* nr = nodes + c * NP + (iz * NYM + iy) * NXM + bxl;
                xc[c * NZM * NYM * LTX + (iz * NYM + iy) * LTX + col] =
                    live ? __builtin_fmaf(w3, nr[3], __builtin_fmaf(w2, nr[2], __builtin_fmaf(w1, nr[1], w0 * nr[0])))
                         : 0.f;
            }
        }
    }
    __syncthreads();  // xc complete; the nodes' space is free for the ring

    // 4. The thread's (x, y): y-collapsed node layers, then the z walk.  Lanes
    //    past the volume's edge walk too (their taps stay in the ring or the
    //    range-checked buffers) so that the walk's state stays wave-uniform;
    //    they store nothing.
    const bool live = tx0 + lxi <= tx1 && ty0 + lyi <= ty1;
    const int lx = tx0 + lxi, ly = ty0 + lyi;
    const float Xf = (float)((int)a.dx0 + lx), Yf = (float)((int)a.dy0 + ly);
    int byl, typ;
    split(ry0 + ty0, live ? lyi : 0, ny0, byl, typ);
    const float wy0 = g.wt[typ * 4 + 0], wy1 = g.wt[typ * 4 + 1], wy2 = g.wt[typ * 4 + 2], wy3 = g.wt[typ * 4 + 3];
    const float* xcol = xc + byl * LTX + lxi;
    const int lstep = NYM * LTX, cstep = NZM * NYM * LTX;
    auto layer = [&](int iz, float* L) {
        const float* col = xcol + iz * lstep;
        #pragma unroll
        for (int c = 0; c < 3; c++) {
            const float* cc = col + c * cstep;
            L[c] = __builtin_fmaf(wy3, cc[3 * LTX],
                                  __builtin_fmaf(wy2, cc[2 * LTX], __builtin_fmaf(wy1, cc[LTX], wy0 * cc[0])));
        }
    };
    const int rz = rz0 + tz0;
    int bzl = rz / f - nz0;
    int tzp = rz - (rz / f) * f;
    float L0[3], L1[3], L2[3], L3[3];
    layer(bzl, L0); layer(bzl + 1, L1); layer(bzl + 2, L2); layer(bzl + 3, L3);
    const int nzv = tz1 - tz0 + 1;
    const uint32_t ostep = (uint32_t)dnx * (uint32_t)dny;
    uint32_t o = (uint32_t)tz0 * ostep + (uint32_t)ly * (uint32_t)dnx + (uint32_t)lx;
    float Zf = (float)((int)a.dz0 + tz0);
    auto df_ph = [&](float* d, int ph) {
        const float4 wz = *(const float4*)(wts + ph * 4);
        const float wz0 = wz.x, wz1 = wz.y, wz2 = wz.z, wz3 = wz.w;
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
    auto lo_bound = [](int s0) { return max(0, s0); };
    auto hi_bound = [](int im, int s0, int sn) { return min(im, s0 + sn - 1) - 1; };
    const int loz = lo_bound(g.sz0), hiz = hi_bound(g.imz, g.sz0, g.snz);
    const bool fast_xy = (int)a.dx0 + tx0 - Rlo[0] >= lo_bound(g.sx0) &&
                         (int)a.dx0 + tx1 + Rhi[0] <= hi_bound(g.imx, g.sx0, g.snx) &&
                         (int)a.dy0 + ty0 - Rlo[1] >= lo_bound(g.sy0) &&
                         (int)a.dy0 + ty1 + Rhi[1] <= hi_bound(g.imy, g.sy0, g.sny);
    const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(dst, (short)0, (int)g.dchan_b, BUF_FLAGS);

    int lz = 0;  // slices of the column walked so far
    // LDS taps, four slices per group: the taps of all four first, then their
    // 32 reads, so that one LDS round trip serves the group (a short last
    // group repeats its last slice).  Outputs go to the chunk's staging at
    // slot = slice within the chunk.
    auto group_lds = [&](auto check, auto full, int nh) {
        constexpr bool CHECK = decltype(check)::value, FULL = decltype(full)::value;
        Tap tp[4];
        bool ok[4];
        #pragma unroll
        for (int j = 0; j < 4; j++) {
            const int jj = FULL ? j : min(j, nh - 1);
            float d[3];
            df_ph(d, tzp + jj);
            const float zj = FULL ? Zf + (float)j : Zf + (float)jj;
            tp[j] = tap_ring<T, CHECK>(g, Xf + d[0], Yf + d[1], zj + d[2], bxy, ok[j]);
        }
        uint32_t u[4][8];
        #pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t q = tp[j].off;
            u[j][0] = lds_at<T>(box, q);
            u[j][1] = lds_at<T>(box, q + BPV);
            u[j][2] = lds_at<T>(box, q + RB);
            u[j][3] = lds_at<T>(box, q + RB + BPV);
            u[j][4] = lds_at<T>(box, q + PB);
            u[j][5] = lds_at<T>(box, q + PB + BPV);
            u[j][6] = lds_at<T>(box, q + PB + RB);
            u[j][7] = lds_at<T>(box, q + PB + RB + BPV);
        }
        // Keep the scheduler from sinking each voxel's reads to their use.
        __builtin_amdgcn_sched_barrier(0);
        const int slot = lz % LCH;
        #pragma unroll
        for (int j = 0; j < 4; j++) {
            uint32_t v = blend8<MODE>(u[j][0], u[j][1], u[j][2], u[j][3], u[j][4], u[j][5], u[j][6], u[j][7], tp[j]);
            if (CHECK) v = ok[j] ? v : 0u;
            if (FULL || j < nh) obuf[((slot + j) * LTY + lyi) * LTX + lxi] = (T)v;
        }
        tzp += nh;
        lz += nh;
        Zf += (float)nh;
    };
    // Slices [lz, lz_end) by LDS taps, cell by cell (a cell: the slices
    // sharing one node-layer window; the fast form where every sample of the
    // cell stays a voxel inside the volume and the window).
    auto walk_lds = [&](int lz_end) {
        while (lz < lz_end) {
            const int cnt = min(f - tzp, lz_end - lz);
            const int zc = (int)a.dz0 + tz0 + lz;
            const bool fast = fast_xy && zc - Rlo[2] >= loz && zc + cnt - 1 + Rhi[2] <= hiz;
            int k = 0;
            if (fast) {
                for (; k + 4 <= cnt; k += 4) group_lds(std::false_type{}, std::true_type{}, 4);
                if (k < cnt) group_lds(std::false_type{}, std::false_type{}, cnt - k);
            } else {
                for (; k + 4 <= cnt; k += 4) group_lds(std::true_type{}, std::true_type{}, 4);
                if (k < cnt) group_lds(std::true_type{}, std::false_type{}, cnt - k);
            }
            if (tzp == f && lz < nzv) next_cell();
        }
    };
    const SrcRs rs = make_src_rs(src, g);
    auto group_gather = [&](auto check, auto full, int nh) {
        constexpr bool CHECK = decltype(check)::value, FULL = decltype(full)::value;
        uint32_t v[4];
        #pragma unroll
        for (int j = 0; j < 4; j++) {
            const int jj = FULL ? j : min(j, nh - 1);
            float d[3];
            df_ph(d, tzp + jj);
            const float zj = FULL ? Zf + (float)j : Zf + (float)jj;
            v[j] = blend<T, MODE, TW_INTERP_LINEAR>(
                rs, tap_setup<T, TW_INTERP_LINEAR, WIN, CHECK>(g, Xf + d[0], Yf + d[1], zj + d[2]));
        }
        #pragma unroll
        for (int j = 0; j < 4; j++)
            if (live && (FULL || j < nh)) bst<T>(rd, (o + (uint32_t)j * ostep) * (uint32_t)BPV, v[j]);
        tzp += nh;
        o += (uint32_t)nh * ostep;
        Zf += (float)nh;
    };
    // Gather taps: the column cell by cell (a cell: the slices sharing one
    // node-layer window; the fast form where every sample of the cell stays a
    // voxel inside the volume and the window).
    auto walk = [&]() {
        while (lz < nzv) {
            const int cnt = min(f - tzp, nzv - lz);
            const int zc = (int)a.dz0 + tz0 + lz;
            const bool fast = fast_xy && zc - Rlo[2] >= loz && zc + cnt - 1 + Rhi[2] <= hiz;
            int k = 0;
            if (fast) {
                for (; k + 4 <= cnt; k += 4) group_gather(std::false_type{}, std::true_type{}, 4);
                if (k < cnt) group_gather(std::false_type{}, std::false_type{}, cnt - k);
            } else {
                for (; k + 4 <= cnt; k += 4) group_gather(std::true_type{}, std::true_type{}, 4);
                if (k < cnt) group_gather(std::true_type{}, std::false_type{}, cnt - k);
            }
            lz += cnt;
            if (tzp == f && lz < nzv) next_cell();
        }
    };
    if (!use_lds) {
        walk();
        return;
    }

    // 5. The ring.  Planes [pa, pa + np) are staged by instructions i = 0 ..
    //    np * H - 1: plane pa + i / H, rows (i % H) * RPI + rsub, 16-byte
    //    chunk `chunk` of each (lane = rsub * CHK + chunk); wave w issues i =
    //    w, w + NW, ...  Chunks past a row's end read the bytes after it in
    //    the window (0 past the channel) into ring bytes no sample reads.
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)src, (short)0, (int)g.chan_b, BUF_FLAGS);
    const int rsub = lane / B::CHK, chunk = lane - rsub * B::CHK;
    const uint32_t voff = (uint32_t)rsub * g.row_b + (uint32_t)chunk * 16u;
    const uint32_t wxy = (uint32_t)(by0 - g.sy0) * g.row_b + (uint32_t)(bx0 - g.sx0) * BPV;
    constexpr uint32_t OOB_V = 0x7FFFFFF0u;  // a vector offset past every range: loads read 0, stores drop
    // Every lane writes (lanes without a piece write to the trash area): a
    // write under a lane mask would leave its load counted as pending on the
    // path that skips it, and the compiler would then wait for every vector
    // memory operation there.
    uint8_t* trash = box + ring_bytes<T>();
    auto st_store = [&](int i, int pa, bool on, const u32x4& v) {
        const int pz = i / H, hh = i - pz * H;
        const int slot = (pa + pz) & (RING - 1);
        uint8_t* p = box + hh * B::RPI * RB + lane * 16;
        *(u32x4*)(on ? p + slot * PB : trash + lane * 16) = v;
        if (slot == 0) *(u32x4*)(on ? p + RING * PB : trash + lane * 16) = v;
    };
    // Staging load i of planes [pa, pa + np).  EXACT: rows of whole aligned
    // 16-byte pieces (no piece crosses the channel's end), one load per lane
    // whatever `on` (idle lanes read past the range); otherwise a piece that
    // crosses the channel's end is read byte by byte.
    auto st_load = [&](auto exact, int i, int pa, int np, bool& on) -> u32x4 {
        const int pz = i / H, hh = i - pz * H;
        on = pz < np && hh * B::RPI + rsub < bny && rsub < B::RPI;
        // (Wave-uniform; said so, or the compiler wraps the load in a loop
        // over the lanes' values.)
        const uint32_t soff = __builtin_amdgcn_readfirstlane(wxy + (uint32_t)(pa + pz - g.sz0) * g.plane_b +
                                                             (uint32_t)(hh * B::RPI) * g.row_b);
        if constexpr (decltype(exact)::value) {
            return __builtin_amdgcn_raw_buffer_load_b128(rsrc, (int)(on ? voff : OOB_V), (int)soff, 0);
        } else {
            u32x4 v = {0u, 0u, 0u, 0u};
            if (on) {
                if (voff + soff + 16u <= g.chan_b) v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, (int)voff, (int)soff, 0);
                else v = load16_tail(rsrc, voff + soff);
            }
            return v;
        }
    };
    // Planes chunk c (slices [c * LCH, c * LCH + LCH) of the column) reaches:
    // the first chunk's reach shifted by c * LCH (exact integers), so that a
    // chunk adds at most LCH planes to its predecessor's.  (The last chunk
    // may reach past its slices: at most the extra planes of a full chunk.)
    const int zl0 = __builtin_amdgcn_readfirstlane((int)floorf((float)tl[2] - dlo[2])) - 1;
    const int zh0 = __builtin_amdgcn_readfirstlane((int)floorf((float)(tl[2] + LCH - 1) + dhi[2])) + 2;
    auto chunk_planes = [&](int c, int& pl, int& ph) {
        pl = max(zl0 + c * LCH, wlo[2]);
        ph = min(zh0 + c * LCH, whi[2]);
    };
    constexpr int EPC = 16 / BPV, CPR = LTX / EPC;  // voxels per 16-byte piece, pieces per row
    const bool full_x = tx0 + LTX <= dnx;
    const bool rows16 = (g.row_b & 15u) == 0 && ((uintptr_t)src & 15u) == 0;
    const int nch = (nzv + LCH - 1) / LCH;
    // Planes in flight for a chunk: loaded into v (planes [na, na + min(nn,
    // LCH))), written to the ring at the end of the chunk before it.
    struct Inflight {
        int na, nn;
        u32x4 v[KPF];
        bool on[KPF];
    };
    // The chunk loop.  EXACT (aligned rows, a full 16-wide column): every
    // chunk issues the same vector memory operations in the same order, so
    // that the compiler counts exactly and the refill waits only for the
    // planes it writes; loads then run two chunks ahead of their use, since
    // one chunk's resampling is shorter than a loaded HBM round trip.
    auto ring_loop = [&](auto exact) {
        constexpr bool EXACT = decltype(exact)::value;
        // Chunk cc's planes after `after` (the highest plane staged or in
        // flight), the first LCH of them loaded now.
        auto issue = [&](int cc, int& after, Inflight& fl) {
            int npl, nph;
            chunk_planes(cc, npl, nph);
            fl.na = max(after + 1, npl);
            fl.nn = cc < nch ? nph - fl.na + 1 : 0;
            #pragma unroll
            for (int j = 0; j < KPF; j++) fl.v[j] = st_load(exact, wave + NW * j, fl.na, min(fl.nn, LCH), fl.on[j]);
            if (fl.nn > 0) after = nph;
        };
        // Chunk c's staged outputs in 16-byte row pieces (EXACT: one store per
        // thread, idle threads' past the range); otherwise voxel by voxel
        // where the volume's x edge cuts the column.
        auto store_chunk = [&](int c, int cnt) {
            const int k = tid / (LTY * CPR), y = (tid / CPR) % LTY, h = tid % CPR;
            const bool ok = k < cnt && ty0 + y <= ty1;
            const uint32_t e0 = ((uint32_t)(tz0 + c * LCH + k) * (uint32_t)dny + (uint32_t)(ty0 + y)) * (uint32_t)dnx +
                                (uint32_t)(tx0 + h * EPC);
            const T* src_row = obuf + (min(k, LCH - 1) * LTY + y) * LTX + h * EPC;
            if constexpr (EXACT) {
                __builtin_amdgcn_raw_buffer_store_b128(*(const u32x4*)src_row, rd, (int)(ok ? e0 * BPV : OOB_V), 0, 0);
            } else if (ok) {
                if (full_x) {
                    __builtin_amdgcn_raw_buffer_store_b128(*(const u32x4*)src_row, rd, (int)(e0 * BPV), 0, 0);
                } else {
                    #pragma unroll
                    for (int e = 0; e < EPC; e++)
                        if (tx0 + h * EPC + e < dnx) bst<T>(rd, (e0 + (uint32_t)e) * BPV, src_row[e]);
                }
            }
        };
        int pl, ph;
        chunk_planes(0, pl, ph);
        // The first chunk's planes, synchronously.
        for (int i0 = 0; i0 < (ph - pl + 1) * H; i0 += NW * KPF) {
            u32x4 v[KPF];
            bool on[KPF];
            #pragma unroll
            for (int j = 0; j < KPF; j++) v[j] = st_load(exact, i0 + wave + NW * j, pl, ph - pl + 1, on[j]);
            #pragma unroll
            for (int j = 0; j < KPF; j++) st_store(i0 + wave + NW * j, pl, on[j], v[j]);
        }
        int top = max(ph, pl - 1);  // the highest plane staged or in flight
        Inflight fa, fb;
        issue(1, top, fa);
        __syncthreads();
        // Chunk c with `cur` holding chunk c + 1's planes; loads chunk c + 2's
        // into `nxt`.  Two calls per iteration swap the roles, so that no
        // in-flight register is copied (a copy would wait for its load).
        auto step = [&](int c, Inflight& cur, Inflight& nxt) {
            issue(c + 2, top, nxt);
            const int cnt = min(LCH, nzv - c * LCH);
            walk_lds(c * LCH + cnt);
            __syncthreads();  // chunk c done: its planes may be overwritten, its outputs are staged
            if (EXACT || cur.nn > 0) {
                #pragma unroll
                for (int j = 0; j < KPF; j++) st_store(wave + NW * j, cur.na, cur.on[j], cur.v[j]);
            }
            store_chunk(c, cnt);
            __syncthreads();
        };
        for (int c = 0; c < nch; c += 2) {
            step(c, fa, fb);
            if (c + 1 < nch) step(c + 1, fb, fa);
        }
    };
    if (rows16 && full_x) ring_loop(std::true_type{});
    else ring_loop(std::false_type{});
}

}  // namespace
