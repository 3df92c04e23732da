// k_tps.hip -- fused cubic-B-spline upsample of the grid DF + resample.
//
// Reference: in-memory CPU loop q_dfblcokinterp_bspline + q_imgblockwarp
// (q_imgwarp_tps_quicksmallmemory.cpp:85-203, driver 2119-2133) and its
// CUDA port get_Displacement_brightness_stps (chazhi.cu:226-338).
//
// Semantics kept from the reference:
//  * voxel X uses grid nodes b..b+3 with b = (X-o)/f and phase t = (X-o)%f;
//    the upsample is the tensor product of the cubic B-spline weights
//    TB[t][k] (uniform, approximating);
//  * the displacement is stored as float, and the sample position is the
//    FLOAT sum of the integer coordinate and the float displacement
//    (pos_warp + sx is long + float in the reference, 106-108);
//  * strict bounds test against the full subject extent, floor/ceil
//    neighbours, weights 1-frac, the t/b, l/r, u/d blend tree, +0.5 and
//    truncation (135-160); nearest neighbour rounds pos+0.5 and clamps.
//
// Variants:
//  TW_EXACT  one thread per voxel, the 64-term FP64 basis product in the
//            reference's order (nodes z->x->y, basis (TB_z*TB_x)*TB_y),
//            rounded to float: bit-identical to the FP64 CPU path.
//  TW_PLAIN  FP32 separable upsample (x, then y, then z), FP64 blend tree
//            in the reference's order.
//  TW_MIXED  as PLAIN with an FP32 blend (7 lerps) and optional
//            offset-scaled FP16 nodes.
//
// Fused kernel layout (PLAIN/MIXED): a work-group of 64x4 threads owns a
// tile of 64 (x) by 4 (y) by TZ (z) output voxels.  It stages the tile's
// grid nodes in LDS, collapses them along x for each of its 64 columns
// (XC, in LDS), and then every thread walks its (x, y) column along z,
// collapsing y per node layer and keeping a sliding window of four layers
// in registers, so a voxel costs 12 FMAs for the z combination plus a few
// amortised ones.  Output rows are written coalesced along x.
//
// Roofline: B_alg = 2*bpv + 3*s_df*nodes/voxels bytes per voxel (subject
// read + output write + grid read); HBM-bound when the per-voxel VALU work
// stays under ~33 ops at 60% of the 8 TB/s peak.
#include "tw_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "tw_tps_dev.h"
#include "tw_tps_lds.h"

namespace {

// ---------------------------------------------------------------------------
// TW_EXACT: one voxel per thread, reference-order FP64 upsample.

template <typename T, int INTERP, bool WRITE_DF>
__global__ __launch_bounds__(256) void k_tps_exact(const TpsArgs a) {
    const i64 lx = (i64)blockIdx.x * TX + threadIdx.x;
    const i64 ly = (i64)blockIdx.y * TY + threadIdx.y;
    const i64 lz = a.zbase + blockIdx.z;
    if (lx >= a.dnx || ly >= a.dny || lz >= a.dnz) return;
    const i64 X = a.dx0 + lx, Y = a.dy0 + ly, Z = a.dz0 + lz;
    const int f = a.f;
    const i64 rx = X - a.ox, ry = Y - a.oy, rz = Z - a.oz;
    const i64 bx = rx / f, by = ry / f, bz = rz / f;
    const int tx = (int)(rx - bx * f), ty = (int)(ry - by * f), tz = (int)(rz - bz * f);
    const float* g = (const float*)a.grid;
    const i64 plane = a.gx * a.gy * a.gz;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int dz = 0; dz < 4; dz++)
        for (int dx = 0; dx < 4; dx++)
            for (int dy = 0; dy < 4; dy++) {
                const double w = (a.w64[tz * 4 + dz] * a.w64[tx * 4 + dx]) * a.w64[ty * 4 + dy];
                const i64 idx = ((bz + dz) * a.gy + (by + dy)) * a.gx + (bx + dx);
                s0 += w * (double)g[idx];
                s1 += w * (double)g[plane + idx];
                s2 += w * (double)g[2 * plane + idx];
            }
    const float d0 = (float)s0, d1 = (float)s1, d2 = (float)s2;
    const i64 dvol = a.dnx * a.dny * a.dnz;
    const i64 o = (lz * a.dny + ly) * a.dnx + lx;
    if (WRITE_DF) {
        a.dfout[o] = d0;
        a.dfout[dvol + o] = d1;
        a.dfout[2 * dvol + o] = d2;
    } else {
        resample_fp64<T, INTERP>(a, X, Y, Z, d0, d1, d2, o, dvol);
    }
}


template <typename T, int MODE, bool F16, int INTERP, bool WRITE_DF, bool WIN>
__global__ __launch_bounds__(FT) void k_tps_fused(const TpsArgs a, const FastGeom g) {
    extern __shared__ float lds[];
    __shared__ float rmax_w[FTY * 6];
    // Work-groups are dispatched round-robin over the 8 XCDs; give each XCD
    // a contiguous run of tiles (x fastest, then y, then z) so that tiles
    // sharing source rows share that XCD's L2.
    const int wg = blockIdx.x;
    const int tile = (wg & 7) * g.per_xcd + (wg >> 3);
    if (tile >= g.ntiles) return;
    tps_tile<T, MODE, F16, INTERP, WRITE_DF, WIN>(a, g, tile, a.src, a.dst, lds, rmax_w);
}

// The fused kernel with the source streamed through LDS (linear PLAIN /
// MIXED): one 16 x 16 x SEG column per work-group, XCD-major as above.
template <typename T, int MODE, bool F16, bool WIN>
__global__ __launch_bounds__(256) void k_tps_lds(const TpsArgs a, const FastGeom g) {
    extern __shared__ __align__(16) float lds16[];
    __shared__ float ext_w[4 * 6];
    const int wg = blockIdx.x;
    const int tile = (wg & 7) * g.per_xcd + (wg >> 3);
    if (tile >= g.ntiles) return;
    tps_column_lds<T, MODE, F16, WIN>(a, g, tile, a.src, a.dst, lds16, ext_w);
}

template <typename T>
int launch_tps_lds(tw_ctx* ctx, TpsArgs a, int precision, bool f16) {
    FastGeom g;
    size_t lds = 0;
    if (const char* why = fused_geom_lds<T>(a, g, lds)) return tw_fail(ctx, TW_E_UNSUPPORTED, "%s", why);
    const bool win = !(a.sx0 == 0 && a.sy0 == 0 && a.sz0 == 0 && a.snx == a.fnx && a.sny == a.fny && a.snz == a.fnz);
    const dim3 grid((unsigned)(8 * g.per_xcd), 1, 1), block(TX, 4, 1);
    const void* src0 = a.src;
    void* dst0 = a.dst;
    for (int ch = 0; ch < (int)a.nc; ch++) {
        a.src = (const void*)((const uint8_t*)src0 + (size_t)ch * a.src_cstride);
        a.dst = (void*)((uint8_t*)dst0 + (size_t)ch * a.dst_cstride);
#define TW_LDS(M, H, W) hipLaunchKernelGGL((k_tps_lds<T, M, H, W>), grid, block, lds, ctx->stream, a, g)
        if (precision == TW_PLAIN) {
            if (f16) {
                if (win) TW_LDS(TW_PLAIN, true, true); else TW_LDS(TW_PLAIN, true, false);
            } else {
                if (win) TW_LDS(TW_PLAIN, false, true); else TW_LDS(TW_PLAIN, false, false);
            }
        } else {
            if (f16) {
                if (win) TW_LDS(TW_MIXED, true, true); else TW_LDS(TW_MIXED, true, false);
            } else {
                if (win) TW_LDS(TW_MIXED, false, true); else TW_LDS(TW_MIXED, false, false);
            }
        }
#undef TW_LDS
        TW_LAUNCH_CHECK(ctx);
    }
    return TW_OK;
}

template <typename T, bool WRITE_DF>
int launch_tps(tw_ctx* ctx, TpsArgs a, int interp, int precision, bool f16) {
    const dim3 block(TX, TY, 1);
    // A source window one voxel wide has no (x, x+1) pairs: the reference-
    // order kernel serves it (its results are the FP64 reference's).
    if (precision == TW_EXACT || (!WRITE_DF && a.snx < 2)) {
        const i64 zmax = 65535;
        for (i64 z0 = 0; z0 < a.dnz; z0 += zmax) {
            a.zbase = z0;
            const i64 nz = std::min(zmax, a.dnz - z0);
            const dim3 grid((unsigned)((a.dnx + TX - 1) / TX), (unsigned)((a.dny + TY - 1) / TY), (unsigned)nz);
            if (interp == TW_INTERP_NN)
                hipLaunchKernelGGL((k_tps_exact<T, TW_INTERP_NN, WRITE_DF>), grid, block, 0, ctx->stream, a);
            else
                hipLaunchKernelGGL((k_tps_exact<T, TW_INTERP_LINEAR, WRITE_DF>), grid, block, 0, ctx->stream, a);
            TW_LAUNCH_CHECK(ctx);
        }
        return TW_OK;
    }
    if (a.f > 64) return tw_fail(ctx, TW_E_UNSUPPORTED, "grid spacing %d above 64 is not supported", a.f);
    if (!WRITE_DF && tw_lds_tile(ctx, interp))
        return launch_tps_lds<T>(ctx, a, precision, f16);
    a.tz = pick_tz(a.f, ctx->tune_tz);
    FastGeom g;
    size_t lds = 0;
    if (const char* why = fused_geom<T, WRITE_DF>(a, g, lds)) return tw_fail(ctx, TW_E_UNSUPPORTED, "%s", why);
    // The source window is the whole volume: no window test needed.
    const bool win = !(a.sx0 == 0 && a.sy0 == 0 && a.sz0 == 0 && a.snx == a.fnx && a.sny == a.fny && a.snz == a.fnz);
    const dim3 grid((unsigned)(8 * g.per_xcd), 1, 1), fblock(TX, FTY, 1);
    const void* src0 = a.src;
    void* dst0 = a.dst;
    for (int ch = 0; ch < (WRITE_DF ? 1 : (int)a.nc); ch++) {
    a.src = src0 ? (const void*)((const uint8_t*)src0 + (size_t)ch * a.src_cstride) : nullptr;
    a.dst = dst0 ? (void*)((uint8_t*)dst0 + (size_t)ch * a.dst_cstride) : nullptr;
#define TW_FUSED(M, H, I, W) \
    hipLaunchKernelGGL((k_tps_fused<T, M, H, I, WRITE_DF, W>), grid, fblock, lds, ctx->stream, a, g)
    if (WRITE_DF) {
        if (f16) TW_FUSED(TW_MIXED, true, TW_INTERP_LINEAR, false);
        else TW_FUSED(TW_MIXED, false, TW_INTERP_LINEAR, false);
    } else if (interp == TW_INTERP_NN) {
        if (f16) TW_FUSED(TW_PLAIN, true, TW_INTERP_NN, true);
        else TW_FUSED(TW_PLAIN, false, TW_INTERP_NN, true);
    } else if (precision == TW_PLAIN) {
        if (f16) {
            if (win) TW_FUSED(TW_PLAIN, true, TW_INTERP_LINEAR, true); else TW_FUSED(TW_PLAIN, true, TW_INTERP_LINEAR, false);
        } else {
            if (win) TW_FUSED(TW_PLAIN, false, TW_INTERP_LINEAR, true); else TW_FUSED(TW_PLAIN, false, TW_INTERP_LINEAR, false);
        }
    } else {
        if (f16) {
            if (win) TW_FUSED(TW_MIXED, true, TW_INTERP_LINEAR, true); else TW_FUSED(TW_MIXED, true, TW_INTERP_LINEAR, false);
        } else {
            if (win) TW_FUSED(TW_MIXED, false, TW_INTERP_LINEAR, true); else TW_FUSED(TW_MIXED, false, TW_INTERP_LINEAR, false);
        }
    }
#undef TW_FUSED
    TW_LAUNCH_CHECK(ctx);
    }
    return TW_OK;
}

// FP64 and FP32 B-spline weight tables of spacing f on the device, cached
// per context: uploaded only when f changes (ordered on the stream after
// every earlier use of the old tables).
int upload_weights(tw_ctx* ctx, int f, int basis, const float** w32, const double** w64) {
    const size_t n = (size_t)f * 4, b64 = n * sizeof(double), b32 = n * sizeof(float);
    void* p = nullptr;
    int rc = tw_ws(ctx, WS_WEIGHTS, b64 + b32, &p);
    if (rc) return rc;
    if (ctx->arena) {
        // Allocation-free mode: the tables go to the arena on every call,
        // from the static host staging (not the context's cached copy).
        void* h = nullptr;
        if ((rc = tw_arena_host(ctx, b64 + b32, &h))) return rc;
        double tb[4 * 64];
        if (f > 64) return tw_fail(ctx, TW_E_UNSUPPORTED, "grid spacing %d above 64 is not supported", f);
        tw_df_tb(f, basis, tb);
        memcpy(h, tb, b64);
        float* h32 = (float*)((uint8_t*)h + b64);
        for (size_t i = 0; i < n; i++) h32[i] = (float)tb[i];
        TW_HIP(ctx, hipMemcpyAsync(p, h, b64 + b32, hipMemcpyHostToDevice, ctx->stream));
        *w64 = (const double*)p;
        *w32 = (const float*)((uint8_t*)p + b64);
        return TW_OK;
    }
    if (ctx->weights_f != f || ctx->weights_basis != basis || ctx->weights_host.size() != b64 + b32) {
        std::vector<double> tb(n);
        tw_df_tb(f, basis, tb.data());
        ctx->weights_host.assign(b64 + b32, 0);
        memcpy(ctx->weights_host.data(), tb.data(), b64);
        float* h32 = (float*)(ctx->weights_host.data() + b64);
        for (size_t i = 0; i < n; i++) h32[i] = (float)tb[i];
        TW_HIP(ctx, hipMemcpyAsync(p, ctx->weights_host.data(), b64 + b32, hipMemcpyHostToDevice, ctx->stream));
        ctx->weights_f = f;
        ctx->weights_basis = basis;
    }
    *w64 = (const double*)p;
    *w32 = (const float*)((uint8_t*)p + b64);
    return TW_OK;
}

int check_grid_cover(tw_ctx* ctx, const tw_grid* g, i64 x0, i64 y0, i64 z0, i64 nx, i64 ny, i64 nz) {
    if (!g || !g->data) return tw_fail(ctx, TW_E_ARG, "null grid");
    if (g->f < 1) return tw_fail(ctx, TW_E_ARG, "grid spacing must be >= 1");
    if (g->basis != TW_BASIS_BSPLINE && g->basis != TW_BASIS_LINEAR)
        return tw_fail(ctx, TW_E_ARG, "grid basis must be TW_BASIS_BSPLINE or TW_BASIS_LINEAR");
    if (g->mem != TW_MEM_DEVICE && g->mem != TW_MEM_HOST) return tw_fail(ctx, TW_E_ARG, "bad grid mem");
    if (g->storage == TW_DF_F16 && (g->mem != TW_MEM_DEVICE || !g->qtab))
        return tw_fail(ctx, TW_E_ARG, "an FP16 grid must be device-resident with its block table");
    if (x0 < g->ox || y0 < g->oy || z0 < g->oz)
        return tw_fail(ctx, TW_E_ARG, "output window starts before the grid origin");
    const i64 f = g->f;
    if ((x0 + nx - 1 - g->ox) / f + 3 >= g->gx || (y0 + ny - 1 - g->oy) / f + 3 >= g->gy ||
        (z0 + nz - 1 - g->oz) / f + 3 >= g->gz)
        return tw_fail(ctx, TW_E_ARG, "grid (%lld x %lld x %lld, f=%d) does not cover the output window",
                       (long long)g->gx, (long long)g->gy, (long long)g->gz, g->f);
    return TW_OK;
}

int stage_grid(tw_ctx* ctx, const tw_grid* g, const void** dev) {
    if (g->mem == TW_MEM_DEVICE) {
        *dev = g->data;
        return TW_OK;
    }
    const size_t bytes = (size_t)g->gx * g->gy * g->gz * 3 * sizeof(float);
    void* p = nullptr;
    int rc = tw_ws(ctx, WS_GRID, bytes, &p);
    if (rc) return rc;
    TW_HIP(ctx, hipMemcpyAsync(p, g->data, bytes, hipMemcpyHostToDevice, ctx->stream));
    *dev = p;
    return TW_OK;
}

void fill_grid_args(TpsArgs& a, const tw_grid* g, const void* gdev) {
    a.grid = gdev;
    a.gx = g->gx; a.gy = g->gy; a.gz = g->gz;
    a.ox = g->ox; a.oy = g->oy; a.oz = g->oz;
    a.f = g->f;
    a.basis = g->basis;
    a.qtab = g->qtab;
    a.qbx = (g->gx + 7) / 8;
    a.qby = (g->gy + 7) / 8;
}

}  // namespace

// Per node layer k: min and max of each displacement component over the
// layer (out[k*6 + 2c], out[k*6 + 2c + 1]).  The upsampled field is a convex
// combination of 4x4x4 nodes, so these bound the field of every voxel whose
// nodes lie in the layers.
__global__ __launch_bounds__(256) void k_layer_bounds(const void* __restrict__ grid, int f16,
                                                      const float* __restrict__ qtab, i64 qbx, i64 qby, i64 gx,
                                                      i64 gy, i64 gz, float* __restrict__ out) {
    __shared__ float red[6][256];
    const i64 k = blockIdx.x, plane = gx * gy * gz;
    float v[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
    for (i64 e = threadIdx.x; e < gx * gy; e += 256) {
        const i64 i = e % gx, j = e / gx, gi = (k * gy + j) * gx + i;
        for (int c = 0; c < 3; c++) {
            float x;
            if (f16) {
                const float* q = qtab + (((k >> 3) * qby + (j >> 3)) * qbx + (i >> 3)) * 6;
                x = q[c] + q[3 + c] * __half2float(((const __half*)grid)[c * plane + gi]);
            } else {
                x = ((const float*)grid)[c * plane + gi];
            }
            v[2 * c] = fminf(v[2 * c], x);
            v[2 * c + 1] = fmaxf(v[2 * c + 1], x);
        }
    }
    for (int c = 0; c < 6; c++) red[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int c = 0; c < 6; c++)
                red[c][threadIdx.x] = (c & 1) ? fmaxf(red[c][threadIdx.x], red[c][threadIdx.x + st])
                                              : fminf(red[c][threadIdx.x], red[c][threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x < 6) out[k * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// Does a source window of nx*ny*nz voxels of `bpv` bytes fit the fused
// kernel's 32-bit tap offsets?
bool fits_fast(i64 nx, i64 ny, i64 nz, int bpv) {
    const uint64_t taps = (uint64_t)nx * ny * nz * bpv + (uint64_t)nx * (ny + 1) * bpv + 16;
    return taps <= 0x7FFF0000ull && ny * nz < (1ll << 24);
}

// Resample dst in z-slabs whose source windows (from the grid's layer bounds)
// fit the fused kernel; used when the whole source does not.
template <typename T>
int resample_slabs(tw_ctx* ctx, TpsArgs a, const tw_grid* grid, int interp, int precision, bool f16) {
    if (ctx->arena)
        return tw_fail(ctx, TW_E_UNSUPPORTED,
                       "a source window past 2 GiB is resampled in slabs, which the allocation-free entry point "
                       "does not take");
    const i64 gz = grid->gz;
    void* bp = nullptr;
    int rc = tw_ws(ctx, WS_AUX_BOUNDS, (size_t)gz * 6 * sizeof(float), &bp);
    if (rc) return rc;
    hipLaunchKernelGGL(k_layer_bounds, dim3((unsigned)gz), dim3(256), 0, ctx->stream, a.grid, f16 ? 1 : 0, a.qtab,
                       a.qbx, a.qby, a.gx, a.gy, gz, (float*)bp);
    TW_LAUNCH_CHECK(ctx);
    std::vector<float> b((size_t)gz * 6);
    TW_HIP(ctx, hipMemcpyAsync(b.data(), bp, b.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    TW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const i64 plane_s = a.snx * a.sny * (i64)sizeof(T), plane_d = a.dnx * a.dny * (i64)sizeof(T);
    const TpsArgs full = a;
    i64 slab = 256;
    for (i64 z0 = 0; z0 < full.dnz;) {
        i64 n = std::min(slab, full.dnz - z0);
        i64 lo = 0, hi = 0;
        for (;;) {
            const i64 Zs = full.dz0 + z0, Ze = full.dz0 + z0 + n - 1;
            const i64 kmin = (Zs - full.oz) / full.f, kmax = std::min<i64>((Ze - full.oz) / full.f + 3, gz - 1);
            float dmin = INFINITY, dmax = -INFINITY;
            for (i64 k = kmin; k <= kmax; k++) {
                dmin = std::min(dmin, b[k * 6 + 4]);
                dmax = std::max(dmax, b[k * 6 + 5]);
            }
            if (!std::isfinite(dmin) || !std::isfinite(dmax)) { dmin = -1e30f; dmax = 1e30f; }
            // Two voxels of margin cover the float rounding of the position.
            lo = std::max<i64>(full.sz0, (i64)std::floor((double)Zs + dmin) - 2);
            hi = std::min<i64>(full.sz0 + full.snz, (i64)std::ceil((double)Ze + dmax) + 3);
            if (hi <= lo) { lo = full.sz0; hi = full.sz0 + 1; }
            if (fits_fast(full.snx, full.sny, hi - lo, sizeof(T)) || n == 1) break;
            n = std::max<i64>(1, n / 2);
            slab = n;
        }
        if (!fits_fast(full.snx, full.sny, hi - lo, sizeof(T)))
            return tw_fail(ctx, TW_E_UNSUPPORTED, "displacement range too large for a 2 GiB source window per slab");
        TpsArgs s = full;
        s.src = (const uint8_t*)full.src + (lo - full.sz0) * plane_s;
        s.sz0 = lo;
        s.snz = hi - lo;
        s.dst = (uint8_t*)full.dst + z0 * plane_d;
        s.dz0 = full.dz0 + z0;
        s.dnz = n;
        if ((rc = launch_tps<T, false>(ctx, s, interp, precision, f16))) return rc;
        z0 += n;
    }
    return TW_OK;
}

extern "C" int tw_resample_tps(tw_ctx* ctx, const tw_volume* src, tw_volume* dst, const tw_grid* grid, int interp,
                               int precision) {
    if (!ctx) return TW_E_ARG;
    int rc;
    if ((rc = tw_check_volume(ctx, src, "src")) || (rc = tw_check_volume(ctx, dst, "dst"))) return rc;
    if (src->dtype != dst->dtype || src->nc != dst->nc)
        return tw_fail(ctx, TW_E_ARG, "src and dst must share dtype and channel count");
    if (interp != TW_INTERP_LINEAR && interp != TW_INTERP_NN)
        return tw_fail(ctx, TW_E_ARG, "interp must be TW_INTERP_LINEAR or TW_INTERP_NN");
    if (precision != TW_EXACT && precision != TW_PLAIN && precision != TW_MIXED)
        return tw_fail(ctx, TW_E_ARG, "unknown precision %d", precision);
    if ((rc = check_grid_cover(ctx, grid, dst->x0, dst->y0, dst->z0, dst->nx, dst->ny, dst->nz))) return rc;
    if (precision == TW_EXACT && grid->storage != TW_DF_F32)
        return tw_fail(ctx, TW_E_ARG, "TW_EXACT needs an FP32 grid");
    if ((rc = tw_bind(ctx))) return rc;
    tw_stages_reset(ctx);
    tw_stage_begin(ctx, ST_TOTAL);
    if ((rc = tw_resample_tps_run(ctx, src, dst, grid, interp, precision))) return rc;
    tw_stage_end(ctx, ST_TOTAL);
    if (src->mem == TW_MEM_HOST || grid->mem == TW_MEM_HOST) TW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return TW_OK;
}

// The resample of tw_resample_tps on validated arguments, without a stage
// reset (the chained warp records its stages in one set).
int tw_resample_tps_run(tw_ctx* ctx, const tw_volume* src, tw_volume* dst, const tw_grid* grid, int interp,
                        int precision, const i64* porigin) {
    int rc;
    TpsArgs a;
    memset(&a, 0, sizeof a);
    const void* sdev = nullptr;
    void* ddev = nullptr;
    const void* gdev = nullptr;
    if ((rc = tw_stage_in(ctx, src, WS_SRC, &sdev))) return rc;
    if ((rc = tw_stage_out_buffer(ctx, dst, WS_DST, &ddev))) return rc;
    if ((rc = stage_grid(ctx, grid, &gdev))) return rc;
    if ((rc = upload_weights(ctx, grid->f, grid->basis, &a.w32, &a.w64))) return rc;
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
    if (porigin) {
        a.px0 = porigin[0];
        a.py0 = porigin[1];
        a.pz0 = porigin[2];
    }
    a.src_cstride = src->nx * src->ny * src->nz * (i64)dtype_size(src->dtype);
    a.dst_cstride = dst->nx * dst->ny * dst->nz * (i64)dtype_size(dst->dtype);
    fill_grid_args(a, grid, gdev);

    const bool f16 = grid->storage == TW_DF_F16;
    const bool slabs = precision != TW_EXACT && !fits_fast(a.snx, a.sny, a.snz, dtype_size(src->dtype));
    tw_stage_begin(ctx, ST_RESAMPLE);
    if (slabs)
        rc = src->dtype == TW_U8 ? resample_slabs<uint8_t>(ctx, a, grid, interp, precision, f16)
                                 : resample_slabs<uint16_t>(ctx, a, grid, interp, precision, f16);
    else
        rc = src->dtype == TW_U8 ? launch_tps<uint8_t, false>(ctx, a, interp, precision, f16)
                                 : launch_tps<uint16_t, false>(ctx, a, interp, precision, f16);
    if (rc) return rc;
    tw_stage_end(ctx, ST_RESAMPLE);
    return tw_stage_out_finish(ctx, dst, ddev);
}

extern "C" int tw_upsample_df(tw_ctx* ctx, const tw_grid* grid, int64_t nx, int64_t ny, int64_t nz, int64_t x0,
                              int64_t y0, int64_t z0, int precision, float* out, int mem) {
    if (!ctx) return TW_E_ARG;
    if (!out || nx <= 0 || ny <= 0 || nz <= 0) return tw_fail(ctx, TW_E_ARG, "bad output window");
    int rc;
    if ((rc = check_grid_cover(ctx, grid, x0, y0, z0, nx, ny, nz))) return rc;
    if (precision == TW_EXACT && grid->storage != TW_DF_F32)
        return tw_fail(ctx, TW_E_ARG, "TW_EXACT needs an FP32 grid");
    if ((rc = tw_bind(ctx))) return rc;
    tw_stages_reset(ctx);
    TpsArgs a;
    memset(&a, 0, sizeof a);
    const void* gdev = nullptr;
    if ((rc = stage_grid(ctx, grid, &gdev))) return rc;
    if ((rc = upload_weights(ctx, grid->f, grid->basis, &a.w32, &a.w64))) return rc;
    const size_t bytes = (size_t)nx * ny * nz * 3 * sizeof(float);
    float* dout = out;
    if (mem == TW_MEM_HOST) {
        void* p = nullptr;
        if ((rc = tw_ws(ctx, WS_DF, bytes, &p))) return rc;
        dout = (float*)p;
    }
    a.dnx = nx; a.dny = ny; a.dnz = nz;
    a.dx0 = x0; a.dy0 = y0; a.dz0 = z0;
    a.dfout = dout;
    a.nc = 1;
    a.fnx = a.fny = a.fnz = 1;
    fill_grid_args(a, grid, gdev);
    rc = launch_tps<uint8_t, true>(ctx, a, TW_INTERP_LINEAR, precision, grid->storage == TW_DF_F16);
    if (rc) return rc;
    if (mem == TW_MEM_HOST) {
        TW_HIP(ctx, hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, ctx->stream));
    }
    TW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return TW_OK;
}
