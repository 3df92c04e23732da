// k_batch.hip -- many independent warps in one launch per stage.
//
// The reference warps one volume (or one tera block) per call: its block
// driver (q_imgwarp_tps_quicksmallmemory.cpp:1161-1994) fits, evaluates the
// grid and resamples block after block, each through its own set of GPU
// wrapper calls (gpu_computedistance weiyijisuan.cu:118-219,
// gpu_interpolation chazhi.cu:557-663, gpu_interpolation_new aff.cu:449-536).
// Here a batch of volumes of different sizes, dtypes and modes runs as
//   1. every TPS fit in one launch (one work-group per system),
//   2. every grid in one launch (and every FP16 quantisation in one),
//   3. every output voxel in one persistent launch,
// with the same per-voxel, per-node and per-system arithmetic as the
// single-volume entry points, so each item's output equals theirs.
//
// Persistent resample: the work is the items' output tiles (TPS: the fused
// kernel's 64x4xtz tiles; affine: the fast affine path's 64x4x16 tiles), per
// channel.  Each
// item's tiles are split in eight contiguous parts, part x served by the
// work-groups of XCD x (work-groups are dispatched round-robin over the 8
// XCDs), which walk their parts item after item with a static stride; all
// XCDs so work on the same item at once and items complete in order.  The
// job table and the part tables are read through the constant address
// space (scalar loads; the kernel never writes them).
#include "tw_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "tw_affine_dev.h"
#include "tw_tps_dev.h"
#include "tw_tps_lds.h"

namespace {

enum { BK_TPS = 0, BK_AFFINE = 1 };

struct BatchJob {
    int kind, dtype;
    int ntile;              // tiles per channel
    int pad;
    i64 src_cstride, dst_cstride;  // bytes between channels
    TpsArgs t;
    FastGeom g;
    AffineArgs af;
    AffGeom ag;
};

#define TW_CONST __attribute__((address_space(4)))

// parts: for XCD x at parts + x * (2 * njobs + 1): P[0..njobs] (prefix sums of
// the part sizes) then S[0..njobs-1] (first tile of each part).  t_end[j]:
// device clock when the last tile of job j ended.
// LDS: TPS tiles are tps_column_lds's 16 x 16 x SEG columns (linear
// interpolation; the jobs' geometry from fused_geom_lds), else tps_tile's.
template <int MODE, bool F16, int INTERP, bool WIN = false, bool LDS = false>
__global__ __launch_bounds__(FT) void k_resample_batch(const BatchJob* __restrict__ jobs_g,
                                                       const int* __restrict__ parts_g, int njobs, int nwg_x,
                                                       unsigned long long* __restrict__ t_end) {
    static_assert(FT == 256 && (!LDS || INTERP == TW_INTERP_LINEAR), "LDS tiles: 64 x 4 threads, linear");
    extern __shared__ __align__(16) float lds[];
    __shared__ float rmax_w[FTY * 6];
    const TW_CONST BatchJob* jobs = (const TW_CONST BatchJob*)jobs_g;
    const TW_CONST int* P = (const TW_CONST int*)parts_g + (blockIdx.x & 7) * (2 * njobs + 1);
    const TW_CONST int* S = P + njobs + 1;
    const bool stamp = threadIdx.x == 0 && threadIdx.y == 0;
    const int Q = P[njobs];
    int j = 0, last = -1;
    for (int q = (int)(blockIdx.x >> 3); q < Q; q += nwg_x) {
        while (P[j + 1] <= q) j++;
        const int tile = S[j] + (q - P[j]);
        __syncthreads();  // the previous tile is done (LDS reuse, its time stamp)
        if (stamp && last >= 0) atomicMax(&t_end[last], (unsigned long long)wall_clock64());
        last = j;
        const TW_CONST BatchJob& J = jobs[j];
        const int ch = tile / J.ntile, t = tile - ch * J.ntile;
        if (J.kind == BK_TPS) {
            const TpsArgs& a = *(const TpsArgs*)&J.t;
            const FastGeom& g = *(const FastGeom*)&J.g;
            const void* src = (const uint8_t*)a.src + ch * J.src_cstride;
            void* dst = (uint8_t*)a.dst + ch * J.dst_cstride;
            if constexpr (LDS) {
                if (J.dtype == TW_U8)
                    tps_column_lds<uint8_t, MODE, F16, WIN>(a, g, t, src, dst, lds, rmax_w);
                else
                    tps_column_lds<uint16_t, MODE, F16, WIN>(a, g, t, src, dst, lds, rmax_w);
            } else {
                if (J.dtype == TW_U8)
                    tps_tile<uint8_t, MODE, F16, INTERP, false, WIN>(a, g, t, src, dst, lds, rmax_w);
                else
                    tps_tile<uint16_t, MODE, F16, INTERP, false, WIN>(a, g, t, src, dst, lds, rmax_w);
            }
        } else if constexpr (!WIN) {  // (windowed launches are brick layers: TPS items only)
            const AffineArgs& af = *(const AffineArgs*)&J.af;
            const AffGeom& ag = *(const AffGeom*)&J.ag;
            if (ag.lds) {
                if (J.dtype == TW_U8) affine_tile_lds<uint8_t, INTERP>(af, ag, t, ch, ch + 1, (uint8_t*)lds);
                else affine_tile_lds<uint16_t, INTERP>(af, ag, t, ch, ch + 1, (uint8_t*)lds);
            } else if (J.dtype == TW_U8) {
                affine_tile_fast<uint8_t, INTERP>(af, ag, t, ch, ch + 1);
            } else {
                affine_tile_fast<uint16_t, INTERP>(af, ag, t, ch, ch + 1);
            }
        }
    }
    __syncthreads();
    if (stamp && last >= 0) atomicMax(&t_end[last], (unsigned long long)wall_clock64());
}

// t0 = device clock now; t_end[0..n) = 0.
__global__ void k_batch_init(unsigned long long* __restrict__ t0, unsigned long long* __restrict__ t_end, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) t_end[i] = 0;
    if (i == 0) *t0 = (unsigned long long)wall_clock64();
}

inline size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

// Geometry of a TPS job for the tile the launch runs (the column kernel sets
// a.tz to its column depth).
const char* tps_job_geom(TpsArgs& a, int dtype, bool lds_tile, FastGeom& g, size_t& lds) {
    if (lds_tile) return dtype == TW_U8 ? fused_geom_lds<uint8_t>(a, g, lds) : fused_geom_lds<uint16_t>(a, g, lds);
    return dtype == TW_U8 ? fused_geom<uint8_t, false>(a, g, lds) : fused_geom<uint16_t, false>(a, g, lds);
}

int check_item(tw_ctx* ctx, const tw_warp_item& it, int i, int f) {
    int rc;
    if ((rc = tw_check_volume(ctx, &it.src, "src")) || (rc = tw_check_volume(ctx, &it.dst, "dst")))
        return tw_fail(ctx, rc, "item %d: %s", i, tw_last_error(ctx));
    const tw_volume &s = it.src, &d = it.dst;
    if (s.mem != TW_MEM_DEVICE || d.mem != TW_MEM_DEVICE)
        return tw_fail(ctx, TW_E_ARG, "item %d: batch volumes must be in device memory", i);
    if (s.dtype != d.dtype || s.nc != d.nc)
        return tw_fail(ctx, TW_E_ARG, "item %d: src and dst must share dtype and channel count", i);
    if (s.x0 || s.y0 || s.z0 || (s.full_nx && s.full_nx != s.nx) || (s.full_ny && s.full_ny != s.ny) ||
        (s.full_nz && s.full_nz != s.nz))
        return tw_fail(ctx, TW_E_ARG, "item %d: a batch source must be a whole volume", i);
    if (it.kind == TW_WARP_TPS) {
        if (!it.centres || !it.targets) return tw_fail(ctx, TW_E_ARG, "item %d: null landmarks", i);
        // Block-size checks of imgwarp_smallmemory_image (2011-2020).
        if (f >= s.nx || f >= s.ny || f >= s.nz)
            return tw_fail(ctx, TW_E_ARG, "item %d: block size should be smaller than the image size", i);
    } else if (it.kind == TW_WARP_AFFINE) {
        for (int k = 0; k < 12; k++)
            if (!std::isfinite(it.M[k])) return tw_fail(ctx, TW_E_ARG, "item %d: matrix entry %d is not finite", i, k);
    } else {
        return tw_fail(ctx, TW_E_ARG, "item %d: unknown kind %d", i, it.kind);
    }
    return TW_OK;
}

}  // namespace

extern "C" int tw_warp_batch(tw_ctx* ctx, const tw_warp_item* items, int n, int f, double lambda, int interp,
                             int precision, int32_t* status, double* done_ms) {
    if (!ctx) return TW_E_ARG;
    if (!items || n < 1) return tw_fail(ctx, TW_E_ARG, "empty batch");
    const int interp_arg = interp, pipe_was_off = ctx->pipe_off;
    const int basis = (interp & TW_DF_LINEAR) ? TW_BASIS_LINEAR : TW_BASIS_BSPLINE;
    interp &= ~TW_DF_LINEAR;
    if (interp != TW_INTERP_LINEAR && interp != TW_INTERP_NN)
        return tw_fail(ctx, TW_E_ARG, "interp must be TW_INTERP_LINEAR or TW_INTERP_NN (optionally | TW_DF_LINEAR)");
    if (precision != TW_PLAIN && precision != TW_MIXED)
        return tw_fail(ctx, TW_E_ARG, "a batch runs TW_PLAIN or TW_MIXED (got %d)", precision);
    if (f < 1 || f > 64) return tw_fail(ctx, TW_E_ARG, "grid spacing must be in [1, 64]");
    int rc;
    for (int i = 0; i < n; i++)
        if ((rc = check_item(ctx, items[i], i, f))) return rc;
    if ((rc = tw_bind(ctx))) return rc;
    tw_host_stage_reset(ctx);
    tw_stages_reset(ctx);
    tw_stage_begin(ctx, ST_TOTAL);
    if (status)
        for (int i = 0; i < n; i++) status[i] = TW_OK;

    // Layout of the grids (WS_BATCH_GRID) and of the resample tables (WS_BATCH_RTAB).
    const bool f16 = precision == TW_MIXED;
    std::vector<int> tps_of(n, -1);
    std::vector<TwFitItem> fits;
    std::vector<TwGridItem> grids;
    std::vector<size_t> goff;
    size_t gtotal = 0;
    for (int i = 0; i < n; i++) {
        const tw_warp_item& it = items[i];
        if (it.kind != TW_WARP_TPS) continue;
        tps_of[i] = (int)fits.size();
        TwFitItem fi = {};
        fi.centres = it.centres;
        fi.targets = it.targets;
        fi.K = it.K;
        fits.push_back(fi);
        TwGridItem gi = {};
        tw_grid& g = gi.grid;
        g.f = f;
        g.ox = it.dst.x0; g.oy = it.dst.y0; g.oz = it.dst.z0;
        g.gx = tw_grid_dim(it.dst.nx, f);
        g.gy = tw_grid_dim(it.dst.ny, f);
        g.gz = tw_grid_dim(it.dst.nz, f);
        g.mem = TW_MEM_DEVICE;
        g.storage = f16 ? TW_DF_F16 : TW_DF_F32;
        grids.push_back(gi);
        const size_t plane = (size_t)g.gx * g.gy * g.gz;
        const size_t nb = (size_t)((g.gx + 7) / 8) * ((g.gy + 7) / 8) * ((g.gz + 7) / 8);
        goff.push_back(gtotal);
        gtotal += up256(plane * 3 * sizeof(float));
        if (f16) gtotal += up256(plane * 3 * sizeof(__half)) + up256(nb * 6 * sizeof(float));
    }
    const int ntps = (int)fits.size();
    const size_t jobs_b = up256((size_t)n * sizeof(BatchJob));
    const size_t parts_b = up256((size_t)8 * (2 * n + 1) * sizeof(int));
    const size_t ts_b = up256((size_t)(n + 1) * sizeof(unsigned long long));
    void *gbase = nullptr, *rbase = nullptr;
    if (ntps && (rc = tw_ws(ctx, WS_BATCH_GRID, gtotal, &gbase))) return rc;
    if ((rc = tw_ws(ctx, WS_BATCH_RTAB, jobs_b + parts_b + ts_b, &rbase))) return rc;
    uint8_t* rdev = (uint8_t*)rbase;
    unsigned long long* t0 = (unsigned long long*)(rdev + jobs_b + parts_b);
    unsigned long long* t_end = t0 + 1;
    hipLaunchKernelGGL(k_batch_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, t0, t_end, n);
    TW_LAUNCH_CHECK(ctx);

    // 1. Fits, 2. grids.
    if (ntps) {
        tw_stage_begin(ctx, ST_FIT);
        const int fit_prec = precision == TW_MIXED ? TW_MIXED : TW_PLAIN;
        if ((rc = tw_fit_batch(ctx, fits.data(), ntps, lambda, fit_prec))) return rc;
        tw_stage_end(ctx, ST_FIT);
        for (int k = 0; k < ntps; k++) {
            TwGridItem& gi = grids[k];
            tw_grid& g = gi.grid;
            uint8_t* p = (uint8_t*)gbase + goff[k];
            const size_t plane = (size_t)g.gx * g.gy * g.gz;
            gi.f32 = (float*)p;
            if (f16) {
                g.data = p + up256(plane * 3 * sizeof(float));
                g.qtab = (float*)(p + up256(plane * 3 * sizeof(float)) + up256(plane * 3 * sizeof(__half)));
            } else {
                g.data = p;
            }
            gi.pts = fits[k].pts;
            gi.cf = fits[k].cf;
            gi.d = fits[k].d;
            gi.K = fits[k].K;
        }
        if ((rc = tw_grid_eval_batch(ctx, grids.data(), ntps))) return rc;
    }

    // 3. Job table and XCD parts.
    uint8_t* host = tw_host_stage(ctx, jobs_b + parts_b);
    BatchJob* hj = (BatchJob*)host;
    int* parts = (int*)(host + jobs_b);
    std::vector<i64> tiles((size_t)n);
    size_t lds = 0;
    const int tz = pick_tz(f, ctx->tune_tz_batch);
    const bool lds_tile = tw_lds_tile(ctx, interp);
    bool any_lds_affine = false;
    for (int i = 0; i < n; i++) {
        const tw_warp_item& it = items[i];
        BatchJob& J = hj[i];
        memset(&J, 0, sizeof J);
        J.dtype = it.src.dtype;
        const int bpv = dtype_size(it.src.dtype);
        J.src_cstride = it.src.nx * it.src.ny * it.src.nz * (i64)bpv;
        J.dst_cstride = it.dst.nx * it.dst.ny * it.dst.nz * (i64)bpv;
        i64 per_ch = 0;
        if (it.kind == TW_WARP_TPS) {
            const tw_grid& g = grids[tps_of[i]].grid;
            J.kind = BK_TPS;
            TpsArgs& a = J.t;
            a.src = it.src.data;
            a.dst = it.dst.data;
            a.snx = it.src.nx; a.sny = it.src.ny; a.snz = it.src.nz;
            a.fnx = it.src.nx; a.fny = it.src.ny; a.fnz = it.src.nz;
            a.dnx = it.dst.nx; a.dny = it.dst.ny; a.dnz = it.dst.nz;
            a.dx0 = it.dst.x0; a.dy0 = it.dst.y0; a.dz0 = it.dst.z0;
            a.nc = it.src.nc;
            a.src_cstride = J.src_cstride;
            a.dst_cstride = J.dst_cstride;
            a.grid = g.data;
            a.gx = g.gx; a.gy = g.gy; a.gz = g.gz;
            a.ox = g.ox; a.oy = g.oy; a.oz = g.oz;
            a.f = f;
            a.basis = basis;
            a.tz = tz;
            a.qtab = g.qtab;
            a.qbx = (g.gx + 7) / 8;
            a.qby = (g.gy + 7) / 8;
            size_t l = 0;
            const char* why = tps_job_geom(a, it.src.dtype, lds_tile, J.g, l);
            if (why) return tw_fail(ctx, TW_E_UNSUPPORTED, "item %d: %s", i, why);
            lds = std::max(lds, l);
            per_ch = J.g.ntiles;
        } else {
            J.kind = BK_AFFINE;
            AffineArgs& a = J.af;
            a.src = it.src.data;
            a.dst = it.dst.data;
            a.snx = it.src.nx; a.sny = it.src.ny; a.snz = it.src.nz;
            a.fnx = it.src.nx; a.fny = it.src.ny; a.fnz = it.src.nz;
            a.dnx = it.dst.nx; a.dny = it.dst.ny; a.dnz = it.dst.nz;
            a.dx0 = it.dst.x0; a.dy0 = it.dst.y0; a.dz0 = it.dst.z0;
            a.nc = it.src.nc;
            for (int k = 0; k < 12; k++) a.m[k] = it.M[k];
            // The box budget is the launch's LDS, set after every job (below).
            const bool al = affine_prefers_lds(a.m);
            const char* why = it.src.dtype == TW_U8 ? affine_geom<uint8_t>(a, J.ag, al, AFF_LDS_BYTES, true)
                                                    : affine_geom<uint16_t>(a, J.ag, al, AFF_LDS_BYTES, true);
            if (why) return tw_fail(ctx, TW_E_UNSUPPORTED, "item %d: %s", i, why);
            any_lds_affine = any_lds_affine || al;
            per_ch = J.ag.ntiles;
        }
        if (per_ch * it.src.nc >= (1ll << 30)) return tw_fail(ctx, TW_E_UNSUPPORTED, "item %d: too many tiles", i);
        J.ntile = (int)per_ch;
        tiles[i] = per_ch * it.src.nc;
    }
    // LDS-staged affine tiles get the launch's LDS (at least their own
    // budget) for their source boxes.
    if (any_lds_affine) {
        lds = std::max(lds, (size_t)AFF_LDS_BYTES);
        for (int i = 0; i < n; i++)
            if (hj[i].kind == BK_AFFINE && hj[i].ag.lds) hj[i].ag.lds_bytes = (int)lds;
    }
    for (int x = 0; x < 8; x++) {
        int* P = parts + x * (2 * n + 1);
        int* S = P + n + 1;
        i64 acc = 0;
        P[0] = 0;
        for (int i = 0; i < n; i++) {
            const i64 a = tiles[i] * x / 8, b = tiles[i] * (x + 1) / 8;
            S[i] = (int)a;
            acc += b - a;
            if (acc >= (1ll << 31)) return tw_fail(ctx, TW_E_UNSUPPORTED, "batch has too many tiles for one launch");
            P[i + 1] = (int)acc;
        }
    }
    TW_HIP(ctx, hipMemcpyAsync(rdev, host, jobs_b + parts_b, hipMemcpyHostToDevice, ctx->stream));

    // Persistent grid: as many work-groups as stay resident, a multiple of 8.
    const void* kern;
    if (interp == TW_INTERP_NN)
        kern = f16 ? (const void*)k_resample_batch<TW_PLAIN, true, TW_INTERP_NN>
                   : (const void*)k_resample_batch<TW_PLAIN, false, TW_INTERP_NN>;
    else if (lds_tile)
        kern = f16 ? (const void*)k_resample_batch<TW_MIXED, true, TW_INTERP_LINEAR, false, true>
                   : (const void*)k_resample_batch<TW_PLAIN, false, TW_INTERP_LINEAR, false, true>;
    else
        kern = f16 ? (const void*)k_resample_batch<TW_MIXED, true, TW_INTERP_LINEAR>
                   : (const void*)k_resample_batch<TW_PLAIN, false, TW_INTERP_LINEAR>;
    int per_cu = 0;
    TW_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, FT, lds));
    int nwg = std::max(1, per_cu) * ctx->num_cus;
    nwg = std::max(8, nwg - nwg % 8);
    const int nwg_x = nwg / 8;
    const BatchJob* djobs = (const BatchJob*)rdev;
    const int* dparts = (const int*)(rdev + jobs_b);
    tw_stage_begin(ctx, ST_RESAMPLE);
    const dim3 grid((unsigned)nwg), block(TX, FTY, 1);
#define TW_BATCH(M, H, I, L) \
    hipLaunchKernelGGL((k_resample_batch<M, H, I, false, L>), grid, block, lds, ctx->stream, djobs, dparts, n, nwg_x, \
                       t_end)
    if (interp == TW_INTERP_NN) {
        if (f16) TW_BATCH(TW_PLAIN, true, TW_INTERP_NN, false); else TW_BATCH(TW_PLAIN, false, TW_INTERP_NN, false);
    } else if (lds_tile) {
        if (f16) TW_BATCH(TW_MIXED, true, TW_INTERP_LINEAR, true); else TW_BATCH(TW_PLAIN, false, TW_INTERP_LINEAR, true);
    } else {
        if (f16) TW_BATCH(TW_MIXED, true, TW_INTERP_LINEAR, false); else TW_BATCH(TW_PLAIN, false, TW_INTERP_LINEAR, false);
    }
#undef TW_BATCH
    TW_LAUNCH_CHECK(ctx);
    tw_stage_end(ctx, ST_RESAMPLE);
    tw_stage_end(ctx, ST_TOTAL);

    // One synchronisation: fit states, then the device clock stamps.
    std::vector<unsigned long long> ts((size_t)n + 1);
    TW_HIP(ctx, hipMemcpyAsync(ts.data(), t0, ts.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                               ctx->stream));
    TW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int first = TW_OK;
    if (ntps) {
        std::vector<int> st((size_t)ntps, TW_OK);
        first = tw_fit_batch_check(ctx, fits.data(), ntps, st.data());
        if (status)
            for (int i = 0; i < n; i++)
                if (tps_of[i] >= 0) status[i] = st[tps_of[i]];
        if (first != TW_OK) {
            // Name the item, not the fit's position among the TPS items.
            for (int i = 0; i < n; i++)
                if (tps_of[i] >= 0 && st[tps_of[i]] != TW_OK) {
                    const std::string m = tw_last_error(ctx);
                    const size_t c = m.find(':');
                    tw_fail(ctx, first, "item %d%s", i, c == std::string::npos ? "" : m.substr(c).c_str());
                    break;
                }
        }
    }
    if (done_ms) {
        int khz = 0;
        TW_HIP(ctx, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
        for (int i = 0; i < n; i++)
            done_ms[i] = khz > 0 && ts[1 + i] >= ts[0] ? (double)(ts[1 + i] - ts[0]) / (double)khz : 0.0;
    }
    // A pipelined solve whose wait ran out: once more, per-block launches.
    if (first != TW_OK && ctx->pipe_off && !pipe_was_off)
        return tw_warp_batch(ctx, items, n, f, lambda, interp_arg, precision, status, done_ms);
    return first;
}

// ---- bricks of one volume against one grid ----------------------------------
// The tera-block drivers warp one block per call (q_imgwarp_tps_quicksmallmemory.cpp:
// 1309-1848: per block gpu_interpolation, chazhi.cu:557-663).  Here every brick's
// output tiles go through one persistent launch of the batch kernel, each job a
// source window (global origin, bounds tested against the whole volume) and an
// output brick, all sharing the caller's grid.
extern "C" int tw_resample_tps_bricks(tw_ctx* ctx, const tw_volume* srcs, const tw_volume* dsts, int n,
                                      const tw_grid* grid, int interp, int precision) {
    if (!ctx) return TW_E_ARG;
    if (!srcs || !dsts || n < 1) return tw_fail(ctx, TW_E_ARG, "no bricks");
    if (interp != TW_INTERP_LINEAR && interp != TW_INTERP_NN)
        return tw_fail(ctx, TW_E_ARG, "interp must be TW_INTERP_LINEAR or TW_INTERP_NN");
    if (precision != TW_PLAIN && precision != TW_MIXED)
        return tw_fail(ctx, TW_E_ARG, "bricks run TW_PLAIN or TW_MIXED (TW_EXACT: tw_resample_tps per brick)");
    if (!grid || !grid->data || grid->mem != TW_MEM_DEVICE || grid->f < 1 || grid->f > 64)
        return tw_fail(ctx, TW_E_ARG, "the grid must be device-resident with spacing 1..64");
    const bool f16 = grid->storage == TW_DF_F16;
    if (f16 && (!grid->qtab || precision != TW_MIXED))
        return tw_fail(ctx, TW_E_ARG, "an FP16 grid needs its block table and TW_MIXED");
    int rc;
    for (int i = 0; i < n; i++) {
        const tw_volume &s = srcs[i], &d = dsts[i];
        if ((rc = tw_check_volume(ctx, &s, "src")) || (rc = tw_check_volume(ctx, &d, "dst")))
            return tw_fail(ctx, rc, "brick %d: %s", i, tw_last_error(ctx));
        if (s.mem != TW_MEM_DEVICE || d.mem != TW_MEM_DEVICE)
            return tw_fail(ctx, TW_E_ARG, "brick %d: volumes must be in device memory", i);
        if (s.dtype != d.dtype || s.nc != d.nc)
            return tw_fail(ctx, TW_E_ARG, "brick %d: src and dst must share dtype and channel count", i);
        const i64 f = grid->f;
        if (d.x0 < grid->ox || d.y0 < grid->oy || d.z0 < grid->oz ||
            (d.x0 + d.nx - 1 - grid->ox) / f + 3 >= grid->gx || (d.y0 + d.ny - 1 - grid->oy) / f + 3 >= grid->gy ||
            (d.z0 + d.nz - 1 - grid->oz) / f + 3 >= grid->gz)
            return tw_fail(ctx, TW_E_ARG, "brick %d: the grid does not cover its output box", i);
    }
    if ((rc = tw_bind(ctx))) return rc;
    tw_host_stage_reset(ctx);
    tw_stages_reset(ctx);
    tw_stage_begin(ctx, ST_TOTAL);
    const size_t jobs_b = up256((size_t)n * sizeof(BatchJob));
    const size_t parts_b = up256((size_t)8 * (2 * n + 1) * sizeof(int));
    const size_t ts_b = up256((size_t)(n + 1) * sizeof(unsigned long long));
    void* rbase = nullptr;
    if ((rc = tw_ws(ctx, WS_BATCH_RTAB, jobs_b + parts_b + ts_b, &rbase))) return rc;
    uint8_t* rdev = (uint8_t*)rbase;
    uint8_t* host = tw_host_stage(ctx, jobs_b + parts_b);
    BatchJob* hj = (BatchJob*)host;
    int* parts = (int*)(host + jobs_b);
    std::vector<i64> tiles((size_t)n);
    size_t lds = 0;
    const int tz = pick_tz(grid->f, ctx->tune_tz_batch);
    const bool lds_tile = tw_lds_tile(ctx, interp);
    for (int i = 0; i < n; i++) {
        const tw_volume &s = srcs[i], &d = dsts[i];
        BatchJob& J = hj[i];
        memset(&J, 0, sizeof J);
        J.kind = BK_TPS;
        J.dtype = s.dtype;
        const int bpv = dtype_size(s.dtype);
        J.src_cstride = s.nx * s.ny * s.nz * (i64)bpv;
        J.dst_cstride = d.nx * d.ny * d.nz * (i64)bpv;
        TpsArgs& a = J.t;
        a.src = s.data;
        a.dst = d.data;
        a.snx = s.nx; a.sny = s.ny; a.snz = s.nz;
        a.sx0 = s.x0; a.sy0 = s.y0; a.sz0 = s.z0;
        a.fnx = s.full_nx ? s.full_nx : s.x0 + s.nx;
        a.fny = s.full_ny ? s.full_ny : s.y0 + s.ny;
        a.fnz = s.full_nz ? s.full_nz : s.z0 + s.nz;
        a.dnx = d.nx; a.dny = d.ny; a.dnz = d.nz;
        a.dx0 = d.x0; a.dy0 = d.y0; a.dz0 = d.z0;
        a.nc = s.nc;
        a.src_cstride = J.src_cstride;
        a.dst_cstride = J.dst_cstride;
        a.grid = grid->data;
        a.gx = grid->gx; a.gy = grid->gy; a.gz = grid->gz;
        a.ox = grid->ox; a.oy = grid->oy; a.oz = grid->oz;
        a.f = grid->f;
        a.basis = grid->basis;
        a.tz = tz;
        a.qtab = grid->qtab;
        a.qbx = (grid->gx + 7) / 8;
        a.qby = (grid->gy + 7) / 8;
        if (s.nx < 2) return tw_fail(ctx, TW_E_UNSUPPORTED, "brick %d: a source window needs two columns", i);
        size_t l = 0;
        const char* why = tps_job_geom(a, s.dtype, lds_tile, J.g, l);
        if (why) return tw_fail(ctx, TW_E_UNSUPPORTED, "brick %d: %s", i, why);
        lds = std::max(lds, l);
        if ((i64)J.g.ntiles * s.nc >= (1ll << 30)) return tw_fail(ctx, TW_E_UNSUPPORTED, "brick %d: too many tiles", i);
        J.ntile = J.g.ntiles;
        tiles[i] = (i64)J.g.ntiles * s.nc;
    }
    for (int x = 0; x < 8; x++) {
        int* P = parts + x * (2 * n + 1);
        int* S = P + n + 1;
        i64 acc = 0;
        P[0] = 0;
        for (int i = 0; i < n; i++) {
            const i64 a = tiles[i] * x / 8, b = tiles[i] * (x + 1) / 8;
            S[i] = (int)a;
            acc += b - a;
            if (acc >= (1ll << 31)) return tw_fail(ctx, TW_E_UNSUPPORTED, "too many tiles for one launch");
            P[i + 1] = (int)acc;
        }
    }
    TW_HIP(ctx, hipMemcpyAsync(rdev, host, jobs_b + parts_b, hipMemcpyHostToDevice, ctx->stream));
    unsigned long long* t0 = (unsigned long long*)(rdev + jobs_b + parts_b);
    unsigned long long* t_end = t0 + 1;
    hipLaunchKernelGGL(k_batch_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, t0, t_end, n);
    TW_LAUNCH_CHECK(ctx);
    const void* kern;
    if (interp == TW_INTERP_NN)
        kern = f16 ? (const void*)k_resample_batch<TW_PLAIN, true, TW_INTERP_NN, true>
                   : (const void*)k_resample_batch<TW_PLAIN, false, TW_INTERP_NN, true>;
    else if (precision == TW_MIXED)
        kern = lds_tile ? (f16 ? (const void*)k_resample_batch<TW_MIXED, true, TW_INTERP_LINEAR, true, true>
                               : (const void*)k_resample_batch<TW_MIXED, false, TW_INTERP_LINEAR, true, true>)
                        : (f16 ? (const void*)k_resample_batch<TW_MIXED, true, TW_INTERP_LINEAR, true>
                               : (const void*)k_resample_batch<TW_MIXED, false, TW_INTERP_LINEAR, true>);
    else
        kern = lds_tile ? (const void*)k_resample_batch<TW_PLAIN, false, TW_INTERP_LINEAR, true, true>
                        : (const void*)k_resample_batch<TW_PLAIN, false, TW_INTERP_LINEAR, true>;
    int per_cu = 0;
    TW_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, FT, lds));
    int nwg = std::max(1, per_cu) * ctx->num_cus;
    nwg = std::max(8, nwg - nwg % 8);
    const int nwg_x = nwg / 8;
    const BatchJob* djobs = (const BatchJob*)rdev;
    const int* dparts = (const int*)(rdev + jobs_b);
    tw_stage_begin(ctx, ST_RESAMPLE);
    const dim3 lg((unsigned)nwg), lb(TX, FTY, 1);
#define TW_BRICKS(M, H, I, L) \
    hipLaunchKernelGGL((k_resample_batch<M, H, I, true, L>), lg, lb, lds, ctx->stream, djobs, dparts, n, nwg_x, t_end)
    if (interp == TW_INTERP_NN) {
        if (f16) TW_BRICKS(TW_PLAIN, true, TW_INTERP_NN, false); else TW_BRICKS(TW_PLAIN, false, TW_INTERP_NN, false);
    } else if (precision == TW_MIXED) {
        if (lds_tile) {
            if (f16) TW_BRICKS(TW_MIXED, true, TW_INTERP_LINEAR, true); else TW_BRICKS(TW_MIXED, false, TW_INTERP_LINEAR, true);
        } else {
            if (f16) TW_BRICKS(TW_MIXED, true, TW_INTERP_LINEAR, false); else TW_BRICKS(TW_MIXED, false, TW_INTERP_LINEAR, false);
        }
    } else {
        if (lds_tile) TW_BRICKS(TW_PLAIN, false, TW_INTERP_LINEAR, true); else TW_BRICKS(TW_PLAIN, false, TW_INTERP_LINEAR, false);
    }
#undef TW_BRICKS
    TW_LAUNCH_CHECK(ctx);
    tw_stage_end(ctx, ST_RESAMPLE);
    tw_stage_end(ctx, ST_TOTAL);
    return TW_OK;
}
