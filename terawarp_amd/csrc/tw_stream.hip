// tw_stream.hip -- bricked streaming of volumes larger than HBM.
//
// Reference: the tera-block drivers of the TPS path
// (q_imgwarp_tps_quicksmallmemory.cpp:1161-1994: per block, estimate the
// source window from 14 probes of the field, read it, compute the block's
// grid with offsets (1084-1152), warp it with q_imgblockwarp_CYF
// (2208-2304) and write the block) and of the affine path
// (q_warp_affine_tps.cpp:851-1329).  Here:
//   1. one fit and one grid over the whole volume on the device (a brick's
//      grid with its origin on the node lattice is a sub-grid of it);
//   2. per brick, the bounds of the three displacement components over the
//      nodes that reach it (one work-group per brick), so each brick's source
//      window is exact: brick + [floor(dmin) - 2, ceil(dmax) + 3];
//   3. a pipeline over NSLOT slots of pinned host and device buffers: a
//      reader thread fills brick b's window (its pool of threads calling the
//      read callback on z-ranges) while the issuing thread copies earlier
//      windows to the device (copy stream), resamples them (the context's
//      stream) and copies the bricks back (a second copy stream), and a
//      writer thread hands each finished brick to the write callback.
// Every wait between the stages is on a HIP event recorded before the wait
// starts or on a host counter guarded by a mutex; an error in any stage
// stops all of them.
#include "tw_internal.h"

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr int NSLOT = 3;
// Resample events per brick: pair b % NEV, read last by the writer of brick
// b.  The issuer records brick b's pair after its iteration for brick
// b - NSLOT, which ended by waiting for the writer to take brick b - 2 NSLOT,
// the previous user of the pair; so no pair is re-recorded before the
// writer has read it.
constexpr int NEV = 2 * NSLOT;

// Brick b (x fastest): min / max of each displacement component over the
// nodes X0/f .. (X1-1)/f + 3 (per axis) that reach its voxels.
__global__ __launch_bounds__(256) void k_brick_bounds(const void* __restrict__ grid, int f16,
                                                      const float* __restrict__ qtab, i64 qbx, i64 qby, i64 gx,
                                                      i64 gy, i64 gz, int f, int brick, i64 nbx, i64 nby, i64 nx,
                                                      i64 ny, i64 nz, float* __restrict__ out) {
    __shared__ float red[6][256];
    const i64 b = blockIdx.x;
    const i64 bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    const i64 X0 = bx * brick, Y0 = by * brick, Z0 = bz * brick;
    const i64 X1 = X0 + brick < nx ? X0 + brick : nx, Y1 = Y0 + brick < ny ? Y0 + brick : ny;
    const i64 Z1 = Z0 + brick < nz ? Z0 + brick : nz;
    const i64 i0 = X0 / f, j0 = Y0 / f, k0 = Z0 / f;
    const i64 i1 = (X1 - 1) / f + 3 < gx - 1 ? (X1 - 1) / f + 3 : gx - 1;
    const i64 j1 = (Y1 - 1) / f + 3 < gy - 1 ? (Y1 - 1) / f + 3 : gy - 1;
    const i64 k1 = (Z1 - 1) / f + 3 < gz - 1 ? (Z1 - 1) / f + 3 : gz - 1;
    const i64 ni = i1 - i0 + 1, nj = j1 - j0 + 1, cnt = ni * nj * (k1 - k0 + 1), plane = gx * gy * gz;
    float v[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
    for (i64 e = threadIdx.x; e < cnt; e += 256) {
        const i64 i = i0 + e % ni, j = j0 + (e / ni) % nj, k = k0 + e / (ni * nj);
        const i64 gi = (k * gy + j) * gx + i;
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
    if (threadIdx.x < 6) out[b * 6 + threadIdx.x] = red[threadIdx.x][0];
}

struct Brick {
    i64 x0, y0, z0, nx, ny, nz;        // output box
    i64 sx0, sy0, sz0, snx, sny, snz;  // source window
};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Buffers, streams and events of one call; released on every exit path.
struct Pipe {
    int device = 0;
    void* hsrc[NSLOT] = {};
    void* hdst[NSLOT] = {};
    void* dsrc[NSLOT] = {};
    void* ddst[NSLOT] = {};
    hipStream_t h2d = nullptr, d2h = nullptr;
    hipEvent_t ev_h2d[NSLOT] = {}, ev_d2h[NSLOT] = {};
    hipEvent_t ev_beg[NEV] = {}, ev_end[NEV] = {};
    ~Pipe() {
        (void)hipSetDevice(device);
        if (h2d) (void)hipStreamSynchronize(h2d);
        if (d2h) (void)hipStreamSynchronize(d2h);
        (void)hipDeviceSynchronize();
        for (int s = 0; s < NSLOT; s++) {
            if (hsrc[s]) (void)hipHostFree(hsrc[s]);
            if (hdst[s]) (void)hipHostFree(hdst[s]);
            if (dsrc[s]) (void)hipFree(dsrc[s]);
            if (ddst[s]) (void)hipFree(ddst[s]);
            for (hipEvent_t e : {ev_h2d[s], ev_d2h[s]})
                if (e) (void)hipEventDestroy(e);
        }
        for (int e = 0; e < NEV; e++) {
            if (ev_beg[e]) (void)hipEventDestroy(ev_beg[e]);
            if (ev_end[e]) (void)hipEventDestroy(ev_end[e]);
        }
        if (h2d) (void)hipStreamDestroy(h2d);
        if (d2h) (void)hipStreamDestroy(d2h);
    }
};

// Progress counters shared by the three stages (-1: nothing yet).
struct Shared {
    std::mutex m;
    std::condition_variable cv;
    i64 ready = -1;                 // last brick whose window is filled
    i64 h2d_issued[NSLOT] = {-1, -1, -1};  // last brick whose upload was issued, per slot
    i64 d2h_issued = -1;            // last brick whose download was issued
    i64 written = -1;               // last brick handed to the write callback
    int err = TW_OK;
    std::string msg;
    double read_s = 0, write_s = 0, resample_ms = 0;

    void fail(int code, const std::string& m2) {
        std::lock_guard<std::mutex> g(m);
        if (err == TW_OK) {
            err = code;
            msg = m2;
        }
        cv.notify_all();
    }
    // Wait until pred() or an error; returns false on error.
    template <class P>
    bool wait(P pred) {
        std::unique_lock<std::mutex> g(m);
        cv.wait(g, [&] { return err != TW_OK || pred(); });
        return err == TW_OK;
    }
    template <class F>
    void set(F f) {
        {
            std::lock_guard<std::mutex> g(m);
            f();
        }
        cv.notify_all();
    }
};

}  // namespace

extern "C" int tw_warp_stream(tw_ctx* ctx, const tw_stream_desc* d, tw_stream_stats* stats) {
    if (!ctx) return TW_E_ARG;
    if (!d) return tw_fail(ctx, TW_E_ARG, "null stream descriptor");
    if (d->nx < 2 || d->ny < 1 || d->nz < 1) return tw_fail(ctx, TW_E_ARG, "bad volume extent");
    if (d->dtype != TW_U8 && d->dtype != TW_U16) return tw_fail(ctx, TW_E_ARG, "bad dtype");
    if (d->f < 1 || d->brick < 1 || d->brick % d->f)
        return tw_fail(ctx, TW_E_ARG, "the brick edge must be a positive multiple of the grid spacing");
    if (d->f >= d->nx || d->f >= d->ny || d->f >= d->nz)
        return tw_fail(ctx, TW_E_ARG, "block size should be smaller than the image size");
    const int interp = d->interp & ~TW_DF_LINEAR;
    if (interp != TW_INTERP_LINEAR && interp != TW_INTERP_NN)
        return tw_fail(ctx, TW_E_ARG, "interp must be TW_INTERP_LINEAR or TW_INTERP_NN (optionally | TW_DF_LINEAR)");
    if (d->precision != TW_EXACT && d->precision != TW_PLAIN && d->precision != TW_MIXED)
        return tw_fail(ctx, TW_E_ARG, "unknown precision %d", d->precision);
    if (!d->read || !d->write) return tw_fail(ctx, TW_E_ARG, "null read or write callback");
    if (d->flags & ~TW_STREAM_BLOCK_ORIGIN) return tw_fail(ctx, TW_E_ARG, "unknown stream flags 0x%x", d->flags);
    const int threads = d->threads < 1 ? 1 : (d->threads > 256 ? 256 : d->threads);
    int rc;
    if ((rc = tw_bind(ctx))) return rc;
    const int pipe_was_off = ctx->pipe_off;
    const double t_start = now_s();
    tw_stages_reset(ctx);
    tw_host_stage_reset(ctx);

    // 1. Fit and the grid over the whole volume.
    TwFitDev fit;
    const int fit_prec = d->precision == TW_MIXED ? TW_MIXED : TW_PLAIN;
    if ((rc = tw_tps_fit_dev(ctx, d->centres, d->targets, d->K, d->lambda, fit_prec, &fit))) return rc;
    tw_grid g = {};
    g.f = d->f;
    g.basis = (d->interp & TW_DF_LINEAR) ? TW_BASIS_LINEAR : TW_BASIS_BSPLINE;
    g.gx = tw_grid_dim(d->nx, d->f);
    g.gy = tw_grid_dim(d->ny, d->f);
    g.gz = tw_grid_dim(d->nz, d->f);
    g.mem = TW_MEM_DEVICE;
    g.storage = d->precision == TW_MIXED ? TW_DF_F16 : TW_DF_F32;
    const size_t plane = (size_t)g.gx * g.gy * g.gz;
    const i64 qbx = (g.gx + 7) / 8, qby = (g.gy + 7) / 8, qbz = (g.gz + 7) / 8;
    void* gbuf = nullptr;
    if ((rc = tw_ws(ctx, WS_DF, plane * 3 * (g.storage == TW_DF_F16 ? 2 : 4), &gbuf))) return rc;
    g.data = gbuf;
    if (g.storage == TW_DF_F16) {
        void* q = nullptr;
        if ((rc = tw_ws(ctx, WS_QTAB, (size_t)(qbx * qby * qbz) * 6 * sizeof(float), &q))) return rc;
        g.qtab = (float*)q;
    }
    if ((rc = tw_grid_eval_dev(ctx, fit.pts, fit.cf, fit.d, d->K, d->precision, &g))) return rc;

    // 2. Bricks and their exact source windows.
    const i64 B = d->brick;
    const i64 nbx = (d->nx + B - 1) / B, nby = (d->ny + B - 1) / B, nbz = (d->nz + B - 1) / B;
    const i64 nb = nbx * nby * nbz;
    if (nb > (1ll << 31) - 1) return tw_fail(ctx, TW_E_UNSUPPORTED, "too many bricks");
    void* bp = nullptr;
    if ((rc = tw_ws(ctx, WS_AUX_BOUNDS, (size_t)nb * 6 * sizeof(float), &bp))) return rc;
    hipLaunchKernelGGL(k_brick_bounds, dim3((unsigned)nb), dim3(256), 0, ctx->stream, g.data,
                       g.storage == TW_DF_F16 ? 1 : 0, g.qtab, qbx, qby, g.gx, g.gy, g.gz, d->f, d->brick, nbx, nby,
                       d->nx, d->ny, d->nz, (float*)bp);
    TW_LAUNCH_CHECK(ctx);
    std::vector<float> bounds((size_t)nb * 6);
    TW_HIP(ctx, hipMemcpyAsync(bounds.data(), bp, bounds.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    if ((rc = tw_tps_fit_check(ctx, &fit))) {  // synchronises
        // A pipelined solve whose wait ran out: once more, per-block launches
        // (no brick has been read or written yet).
        if (ctx->pipe_off && !pipe_was_off) return tw_warp_stream(ctx, d, stats);
        return rc;
    }
    const int bpv = dtype_size(d->dtype);
    std::vector<Brick> bricks((size_t)nb);
    size_t max_src = 0, max_dst = 0;
    double max_margin = 0;
    for (i64 b = 0; b < nb; b++) {
        Brick& k = bricks[(size_t)b];
        const i64 bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
        const i64 o[3] = {bx * B, by * B, bz * B}, n[3] = {d->nx, d->ny, d->nz};
        i64 lo[3], hi[3], len[3];
        for (int a = 0; a < 3; a++) {
            len[a] = std::min(B, n[a] - o[a]);
            const float dmin = bounds[(size_t)b * 6 + 2 * a], dmax = bounds[(size_t)b * 6 + 2 * a + 1];
            if (std::isfinite(dmin) && std::isfinite(dmax)) {
                // Two voxels of margin cover the float rounding of the position.
                lo[a] = std::max<i64>(0, (i64)std::floor((double)o[a] + dmin) - 2);
                hi[a] = std::min<i64>(n[a], (i64)std::ceil((double)(o[a] + len[a] - 1) + dmax) + 3);
            } else {
                lo[a] = 0;
                hi[a] = n[a];
            }
            if (hi[a] <= lo[a]) {  // every sample outside: any one-voxel window
                lo[a] = std::min(o[a], n[a] - 1);
                hi[a] = lo[a] + 1;
            }
            max_margin = std::max({max_margin, (double)(o[a] - lo[a]), (double)(hi[a] - o[a] - len[a])});
        }
        k = {o[0], o[1], o[2], len[0], len[1], len[2], lo[0], lo[1], lo[2], hi[0] - lo[0], hi[1] - lo[1],
             hi[2] - lo[2]};
        if (k.snx < 2) {  // the fused kernel pairs x with x + 1 (nx >= 2)
            k.sx0 = std::min(k.sx0, d->nx - 2);
            k.snx = 2;
        }
        // Windows past the fused kernel's 32-bit offsets are resampled in
        // z-slabs by tw_resample_tps_run (TW_EXACT addresses in 64 bits).
        const uint64_t sb = (uint64_t)k.snx * k.sny * k.snz * bpv;
        max_src = std::max<size_t>(max_src, sb);
        max_dst = std::max<size_t>(max_dst, (size_t)(k.nx * k.ny * k.nz * bpv));
    }
    const double t_setup = now_s();

    // 3. The pipeline.
    Pipe P;
    P.device = ctx->device;
    for (int s = 0; s < NSLOT; s++) {
        TW_HIP(ctx, hipHostMalloc(&P.hsrc[s], max_src, hipHostMallocDefault));
        TW_HIP(ctx, hipHostMalloc(&P.hdst[s], max_dst, hipHostMallocDefault));
        TW_HIP(ctx, hipMalloc(&P.dsrc[s], max_src));
        TW_HIP(ctx, hipMalloc(&P.ddst[s], max_dst));
        TW_HIP(ctx, hipEventCreateWithFlags(&P.ev_h2d[s], hipEventDisableTiming));
        TW_HIP(ctx, hipEventCreateWithFlags(&P.ev_d2h[s], hipEventDisableTiming));
    }
    for (int e = 0; e < NEV; e++) {
        TW_HIP(ctx, hipEventCreate(&P.ev_beg[e]));
        TW_HIP(ctx, hipEventCreate(&P.ev_end[e]));
    }
    TW_HIP(ctx, hipStreamCreateWithFlags(&P.h2d, hipStreamNonBlocking));
    TW_HIP(ctx, hipStreamCreateWithFlags(&P.d2h, hipStreamNonBlocking));
    Shared S;

    // Reader: fills slot b % NSLOT once the slot's previous upload completed.
    std::thread reader([&] {
        (void)hipSetDevice(ctx->device);
        for (i64 b = 0; b < nb; b++) {
            const int s = (int)(b % NSLOT);
            if (b >= NSLOT) {
                if (!S.wait([&] { return S.h2d_issued[s] >= b - NSLOT; })) return;
                if (hipEventSynchronize(P.ev_h2d[s]) != hipSuccess) return S.fail(TW_E_DEVICE, "upload failed");
            }
            const Brick& k = bricks[(size_t)b];
            const double t0 = now_s();
            const int nt = (int)std::min<i64>(threads, k.snz);
            std::vector<std::thread> pool;
            std::vector<int> rcs((size_t)nt, TW_OK);
            const size_t zbytes = (size_t)k.snx * k.sny * bpv;
            for (int t = 0; t < nt; t++) {
                const i64 za = k.snz * t / nt, zb = k.snz * (t + 1) / nt;
                pool.emplace_back([&, t, za, zb] {
                    rcs[(size_t)t] = d->read(d->read_user, k.sx0, k.sy0, k.sz0 + za, k.snx, k.sny, zb - za,
                                             (uint8_t*)P.hsrc[s] + (size_t)za * zbytes);
                });
            }
            for (auto& th : pool) th.join();
            for (int r : rcs)
                if (r != TW_OK) return S.fail(r, "read callback failed for brick " + std::to_string(b));
            const double dt = now_s() - t0;
            S.set([&] {
                S.read_s += dt;
                S.ready = b;
            });
        }
    });
    // Writer: hands each brick to the callback once its download completed.
    std::thread writer([&] {
        (void)hipSetDevice(ctx->device);
        for (i64 b = 0; b < nb; b++) {
            const int s = (int)(b % NSLOT);
            if (!S.wait([&] { return S.d2h_issued >= b; })) return;
            if (hipEventSynchronize(P.ev_d2h[s]) != hipSuccess) return S.fail(TW_E_DEVICE, "download failed");
            float ms = 0.f;
            const hipError_t te = hipEventElapsedTime(&ms, P.ev_beg[b % NEV], P.ev_end[b % NEV]);
            if (te != hipSuccess)
                return S.fail(TW_E_DEVICE, std::string("resample timing of brick ") + std::to_string(b) + ": " +
                                               hipGetErrorString(te));
            const Brick& k = bricks[(size_t)b];
            const double t0 = now_s();
            const int r = d->write(d->write_user, b, k.x0, k.y0, k.z0, k.nx, k.ny, k.nz, P.hdst[s]);
            if (r != TW_OK) return S.fail(r, "write callback failed for brick " + std::to_string(b));
            const double dt = now_s() - t0;
            S.set([&] {
                S.write_s += dt;
                S.resample_ms += ms;
                S.written = b;
            });
        }
    });
    // Issuing (this thread): upload, resample, download.
    double h2d_bytes = 0, d2h_bytes = 0, wait_s = 0;
    // The issuing thread's waits on the reader and writer (excluded from issue_s).
    auto wait_for = [&](auto pred) {
        const double t0 = now_s();
        const bool ok = S.wait(pred);
        wait_s += now_s() - t0;
        return ok;
    };
    const double t_issue = now_s();
    auto issue = [&]() -> int {
        for (i64 b = 0; b < nb; b++) {
            const int s = (int)(b % NSLOT);
            const Brick& k = bricks[(size_t)b];
            if (!wait_for([&] { return S.ready >= b; })) return S.err;
            const size_t sb = (size_t)(k.snx * k.sny * k.snz * bpv), db = (size_t)(k.nx * k.ny * k.nz * bpv);
            // The device window of slot s is free once brick b - NSLOT's resample ended.
            if (b >= NSLOT) TW_HIP(ctx, hipStreamWaitEvent(P.h2d, P.ev_end[(b - NSLOT) % NEV], 0));
            TW_HIP(ctx, hipMemcpyAsync(P.dsrc[s], P.hsrc[s], sb, hipMemcpyHostToDevice, P.h2d));
            TW_HIP(ctx, hipEventRecord(P.ev_h2d[s], P.h2d));
            S.set([&] { S.h2d_issued[s] = b; });
            h2d_bytes += (double)sb;
            TW_HIP(ctx, hipStreamWaitEvent(ctx->stream, P.ev_h2d[s], 0));
            if (b >= NSLOT) TW_HIP(ctx, hipStreamWaitEvent(ctx->stream, P.ev_d2h[s], 0));
            tw_volume vs = {P.dsrc[s], k.snx, k.sny, k.snz, 1, k.sx0, k.sy0, k.sz0, d->nx, d->ny, d->nz,
                            d->dtype, TW_MEM_DEVICE};
            tw_volume vd = {P.ddst[s], k.nx, k.ny, k.nz, 1, k.x0, k.y0, k.z0, 0, 0, 0, d->dtype, TW_MEM_DEVICE};
            const i64 po[3] = {k.x0, k.y0, k.z0};
            const bool block_origin = (d->flags & TW_STREAM_BLOCK_ORIGIN) != 0;
            TW_HIP(ctx, hipEventRecord(P.ev_beg[b % NEV], ctx->stream));
            int r = tw_resample_tps_run(ctx, &vs, &vd, &g, interp, d->precision, block_origin ? po : nullptr);
            if (r != TW_OK) return r;
            TW_HIP(ctx, hipEventRecord(P.ev_end[b % NEV], ctx->stream));
            // The pinned brick of slot s is free once the writer took brick b - NSLOT.
            if (b >= NSLOT && !wait_for([&] { return S.written >= b - NSLOT; })) return S.err;
            TW_HIP(ctx, hipStreamWaitEvent(P.d2h, P.ev_end[b % NEV], 0));
            TW_HIP(ctx, hipMemcpyAsync(P.hdst[s], P.ddst[s], db, hipMemcpyDeviceToHost, P.d2h));
            TW_HIP(ctx, hipEventRecord(P.ev_d2h[s], P.d2h));
            S.set([&] { S.d2h_issued = b; });
            d2h_bytes += (double)db;
        }
        return TW_OK;
    };
    rc = issue();
    const double issue_s = now_s() - t_issue - wait_s;
    if (rc != TW_OK && S.err == TW_OK) S.fail(rc, ctx->err);
    reader.join();
    writer.join();
    if (S.err != TW_OK) return tw_fail(ctx, S.err, "%s", S.msg.c_str());
    if (stats) {
        stats->bricks = nb;
        stats->wall_s = now_s() - t_start;
        stats->setup_ms = (t_setup - t_start) * 1e3;
        stats->resample_ms = S.resample_ms;
        stats->read_s = S.read_s;
        stats->write_s = S.write_s;
        stats->h2d_bytes = h2d_bytes;
        stats->d2h_bytes = d2h_bytes;
        stats->max_margin = max_margin;
        stats->issue_s = issue_s;
    }
    return TW_OK;
}
