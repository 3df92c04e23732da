// k_grid.hip -- TPS displacement at the coarse grid nodes.
//
// Reference: CPU loop of compute_df_stps_subsampled_volume_4bspline
// (q_imgwarp_tps_quicksmallmemory.cpp:588-649; block variant with offsets
// 1084-1152) and its CUDA port get_cd / gpu_computedistance
// (weiyijisuan.cu:91-219).
//
// Node (i, j, k) sits at P = ((i-1)f + ox, (j-1)f + oy, (k-1)f + oz).
// x_stps = [1 P] d + sum_n -|P - centre_n| c_n and DF = x_stps - P.
// The kernel distance takes float differences of the integer node
// coordinate and the float centre, as the reference does.
//
// TW_EXACT: FP64, the reference's summation order (affine part, then the
//   landmark sum in ascending n, then their sum), one thread per node.
// Otherwise the distances are FP32 (float differences, FP32 squares and
//   v_sqrt_f32) and the landmark sum is compensated by accumulating in FP64:
//   on MI355X an FP64 FMA issues at the FP32 rate, so the FP64 accumulator
//   (one cvt + three FMAs per landmark) is cheaper than TwoSum compensation
//   (six FP32 ops per component) and rounds the products only once.  The
//   affine part is evaluated in FP64 as d0 + P (d_lin - I) so that the large
//   identity term never enters the field.
// Landmarks are read with wave-uniform addresses (scalar loads).
//
// Roofline: compute-bound; ~11 VALU issue slots per (node, landmark) pair
// (the quarter-rate sqrt counts four), so nodes*K*11 / 3.9e13 per second.
#include "tw_internal.h"

#include <cmath>
#include <cstring>

namespace {

struct GridArgs {
    const float4* pts;   // K centres (x, y, z, 0)
    const double* c64;   // K x 3 coefficients (cols 1..3), TW_EXACT
    const double* cf;    // K x 3 coefficients (cols 1..3), FP32-distance path
    int K;
    int f;
    i64 gx, gy, gz, ox, oy, oz;
    const double* d;     // affine block, row-major 4x4 (device)
    float* out;          // 3 planes gx*gy*gz
};

template <bool EXACT>
__global__ __launch_bounds__(256) void k_grid_eval(const GridArgs a) {
    const i64 plane = a.gx * a.gy * a.gz;
    const i64 idx = (i64)blockIdx.x * 256 + threadIdx.x;
    if (idx >= plane) return;
    const i64 i = idx % a.gx;
    const i64 j = (idx / a.gx) % a.gy;
    const i64 k = idx / (a.gx * a.gy);
    const i64 X = (i - 1) * a.f + a.ox, Y = (j - 1) * a.f + a.oy, Z = (k - 1) * a.f + a.oz;
    const float Xf = (float)X, Yf = (float)Y, Zf = (float)Z;
    if (EXACT) {
        const double xo[4] = {1.0, (double)X, (double)Y, (double)Z};
        double b1 = 0.0, b2 = 0.0, b3 = 0.0;
        for (int n = 0; n < a.K; n++) {
            const float4 p = a.pts[n];
            const double dx = (float)(Xf - p.x), dy = (float)(Yf - p.y), dz = (float)(Zf - p.z);
            const double kr = -sqrt(dx * dx + dy * dy + dz * dz);
            b1 += kr * a.c64[3 * n + 0];
            b2 += kr * a.c64[3 * n + 1];
            b3 += kr * a.c64[3 * n + 2];
        }
        double a1 = 0.0, a2 = 0.0, a3 = 0.0;
        for (int m = 0; m < 4; m++) {
            a1 += xo[m] * a.d[m * 4 + 1];
            a2 += xo[m] * a.d[m * 4 + 2];
            a3 += xo[m] * a.d[m * 4 + 3];
        }
        a.out[idx] = (float)((a1 + b1) - (double)X);
        a.out[plane + idx] = (float)((a2 + b2) - (double)Y);
        a.out[2 * plane + idx] = (float)((a3 + b3) - (double)Z);
    } else {
        double s1 = 0.0, s2 = 0.0, s3 = 0.0;
        for (int n = 0; n < a.K; n++) {
            const float4 p = a.pts[n];
            const float dx = Xf - p.x, dy = Yf - p.y, dz = Zf - p.z;
            const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
            const double r = (double)__builtin_amdgcn_sqrtf(r2);
            s1 = fma(-r, a.cf[3 * n + 0], s1);
            s2 = fma(-r, a.cf[3 * n + 1], s2);
            s3 = fma(-r, a.cf[3 * n + 2], s3);
        }
        // Affine part minus identity, in FP64.
        const double Xd = (double)X, Yd = (double)Y, Zd = (double)Z;
        const double l1 = a.d[1] + Xd * (a.d[5] - 1.0) + Yd * a.d[9] + Zd * a.d[13];
        const double l2 = a.d[2] + Xd * a.d[6] + Yd * (a.d[10] - 1.0) + Zd * a.d[14];
        const double l3 = a.d[3] + Xd * a.d[7] + Yd * a.d[11] + Zd * (a.d[15] - 1.0);
        a.out[idx] = (float)(l1 + s1);
        a.out[plane + idx] = (float)(l2 + s2);
        a.out[2 * plane + idx] = (float)(l3 + s3);
    }
}

// FP32-distance grid kernel, tiled: a work-group owns 16 GNP x 16 nodes of
// one z-plane (wave w: rows 4w..4w+3, 16 GNP nodes each, a thread GNP nodes
// 16 apart in one row, so each landmark's scalar loads and LDS entry serve
// GNP nodes).  For every chunk of 256 landmarks it stages (Y - py)^2 + (Z - pz)^2
// per row in LDS, so a node-landmark pair costs dx, one FMA, the square root
// and the three FP64 accumulations.
constexpr int GT = 16, GCH = 256, GRS = GCH + 4;  // tile rows, landmark chunk, LDS row stride
// A thread owns nodes i + GT q, q < GNP, of one row: GNP = 2 for a single
// large grid (more nodes per thread cost occupancy there: K = 4096 at
// 259x259x67 nodes took 4.93 ms at 4 against 4.60 ms at 2), 4 for the
// batch's many small grids (0.79 against 1.19 ms for 158 grids).
constexpr int GNP_ONE = 2, GNP_BATCH = 4;

template <int GNP>
__device__ __forceinline__ void grid_tiled_body(const GridArgs& a, int bx, int by, int bz, float (*syz)[GRS]) {
    constexpr int GTX = GNP * GT;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const i64 i = (i64)bx * GTX + (lane & 15);
    const i64 j = (i64)by * GT + 4 * w + (lane >> 4);
    const i64 k = bz;
    const bool live = i < a.gx && j < a.gy;
    const i64 X = (i - 1) * a.f + a.ox, Z = (k - 1) * a.f + a.oz;
    float Xf[GNP];
    #pragma unroll
    for (int q = 0; q < GNP; q++) Xf[q] = (float)(X + q * GT * a.f);
    const float Zf = (float)Z;
    const int row = 4 * w + (lane >> 4);
    const float Y0 = (float)(((i64)by * GT - 1) * a.f + a.oy);  // row 0 of the tile
    double s[GNP][3];
    #pragma unroll
    for (int q = 0; q < GNP; q++) s[q][0] = s[q][1] = s[q][2] = 0.0;
    for (int n0 = 0; n0 < a.K; n0 += GCH) {
        const int cnt = min(GCH, a.K - n0);
        __syncthreads();
        if (t < cnt) {
            const float4 p = a.pts[n0 + t];
            const float dz = Zf - p.z;
            const float dz2 = dz * dz;
            #pragma unroll
            for (int r = 0; r < GT; r++) {
                const float dy = (Y0 + (float)(r * a.f)) - p.y;
                syz[r][t] = __builtin_fmaf(dy, dy, dz2);
            }
        }
        __syncthreads();
        if (live) {
            // Every node's terms are formed whether or not it exists (same
            // row, so the same landmark data and LDS entry); only live nodes
            // are stored.
            const float* sr = syz[row];
            constexpr int UNROLL = 32 / GNP;
            #pragma unroll(UNROLL)
            for (int n = 0; n < cnt; n++) {
                const int m = n0 + n;
                const float px = a.pts[m].x, syzn = sr[n];
                const double c1 = a.cf[3 * m + 0], c2 = a.cf[3 * m + 1], c3 = a.cf[3 * m + 2];
                #pragma unroll
                for (int q = 0; q < GNP; q++) {
                    const float dx = Xf[q] - px;
                    const double r = (double)__builtin_amdgcn_sqrtf(__builtin_fmaf(dx, dx, syzn));
                    s[q][0] = fma(-r, c1, s[q][0]);
                    s[q][1] = fma(-r, c2, s[q][1]);
                    s[q][2] = fma(-r, c3, s[q][2]);
                }
            }
        }
    }
    if (!live) return;
    const i64 plane = a.gx * a.gy * a.gz, idx = (k * a.gy + j) * a.gx + i;
    const double Yd = (double)((j - 1) * a.f + a.oy), Zd = (double)Z;
    #pragma unroll
    for (int q = 0; q < GNP; q++) {
        if (i + q * GT >= a.gx) break;
        const double Xd = (double)(X + q * GT * a.f);
        const double l1 = a.d[1] + Xd * (a.d[5] - 1.0) + Yd * a.d[9] + Zd * a.d[13];
        const double l2 = a.d[2] + Xd * a.d[6] + Yd * (a.d[10] - 1.0) + Zd * a.d[14];
        const double l3 = a.d[3] + Xd * a.d[7] + Yd * a.d[11] + Zd * (a.d[15] - 1.0);
        const i64 o = idx + q * GT;
        a.out[o] = (float)(l1 + s[q][0]);
        a.out[plane + o] = (float)(l2 + s[q][1]);
        a.out[2 * plane + o] = (float)(l3 + s[q][2]);
    }
}

__global__ __launch_bounds__(256) void k_grid_tiled(const GridArgs a) {
    __shared__ __align__(16) float syz[GT][GRS];
    grid_tiled_body<GNP_ONE>(a, blockIdx.x, blockIdx.y, blockIdx.z, syz);
}

// FP16 offset-scaled storage: per block of 8x8x8 nodes and component,
// value = offset + scale * h with h in [-1, 1].
constexpr int QB = 8;

struct QuantLds {
    float red[2][3][256];
    float prm[6];
};

__device__ __forceinline__ void quantize_body(const float* __restrict__ df, __half* __restrict__ out,
                                              float* __restrict__ qtab, i64 gx, i64 gy, i64 gz, i64 qbx, i64 qby,
                                              i64 b, QuantLds& S) {
    auto& red = S.red;
    auto& prm = S.prm;
    const i64 plane = gx * gy * gz;
    const i64 bx = b % qbx, by = (b / qbx) % qby, bz = b / (qbx * qby);
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    // A thread's two nodes stay in registers for the quantising pass.
    constexpr int NPT = QB * QB * QB / 256;
    float v[NPT][3];
    i64 idx[NPT];
    bool in[NPT];
    #pragma unroll
    for (int u = 0; u < NPT; u++) {
        const int t = threadIdx.x + 256 * u;
        const i64 i = bx * QB + (t % QB), j = by * QB + (t / QB) % QB, k = bz * QB + t / (QB * QB);
        in[u] = i < gx && j < gy && k < gz;
        idx[u] = (k * gy + j) * gx + i;
        #pragma unroll
        for (int c = 0; c < 3; c++) {
            v[u][c] = in[u] ? df[c * plane + idx[u]] : 0.f;
            if (in[u]) {
                mn[c] = fminf(mn[c], v[u][c]);
                mx[c] = fmaxf(mx[c], v[u][c]);
            }
        }
    }
    // Min / max over the work-group: within each wave by shuffles, then the
    // four waves' results (order-independent, so the same bits as any tree).
    for (int c = 0; c < 3; c++) {
        for (int o = 32; o > 0; o >>= 1) {
            mn[c] = fminf(mn[c], __shfl_xor(mn[c], o, 64));
            mx[c] = fmaxf(mx[c], __shfl_xor(mx[c], o, 64));
        }
        if ((threadIdx.x & 63) == 0) {
            red[0][c][threadIdx.x >> 6] = mn[c];
            red[1][c][threadIdx.x >> 6] = mx[c];
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int c = threadIdx.x;
        const float lo = fminf(fminf(red[0][c][0], red[0][c][1]), fminf(red[0][c][2], red[0][c][3]));
        const float hi = fmaxf(fmaxf(red[1][c][0], red[1][c][1]), fmaxf(red[1][c][2], red[1][c][3]));
        const float off = 0.5f * (lo + hi);
        float sc = fmaxf(hi - off, off - lo);
        if (!(sc > 0.f)) sc = 1.f;
        prm[c] = off;
        prm[3 + c] = sc;
        qtab[b * 6 + c] = off;
        qtab[b * 6 + 3 + c] = sc;
    }
    __syncthreads();
    #pragma unroll
    for (int u = 0; u < NPT; u++)
        if (in[u])
            #pragma unroll
            for (int c = 0; c < 3; c++) {
                const float h = (v[u][c] - prm[c]) / prm[3 + c];
                out[c * plane + idx[u]] = __float2half_rn(fminf(1.f, fmaxf(-1.f, h)));
            }
}

// One grid's quantisation: a run of QR = 8 blocks along x per work-group,
// so that a wave reads and writes whole 64-node rows (quantize_body, which
// the batch runs one block per work-group, covers eight 8-node rows of one
// block per wave).  Lane l owns node x = 64 rx + l of
// block 8 rx + l / 8 and the wave's 16 (y, z) rows of the block row; the
// blocks' extremes and parameters are those of quantize_body (min / max are
// exact in any order), so the output and the table are bit-identical.
constexpr int QR = 8;

struct QuantRowLds {
    float red[2][3][4][QR];  // per wave: min / max per component and block
    float prm[QR][6];
};

__global__ __launch_bounds__(256) void k_quantize_rows(const float* __restrict__ df, __half* __restrict__ out,
                                                       float* __restrict__ qtab, i64 gx, i64 gy, i64 gz, i64 qbx,
                                                       i64 qby, i64 nrx) {
    __shared__ QuantRowLds S;
    const i64 wg = blockIdx.x;
    const i64 rx = wg % nrx, by = (wg / nrx) % qby, bz = wg / (nrx * qby);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, blk = lane / QB;
    const i64 plane = gx * gy * gz;
    const i64 i = rx * QB * QR + lane;
    constexpr int RPW = QB * QB / 4;  // (y, z) rows per wave
    float v[RPW][3];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    #pragma unroll
    for (int r = 0; r < RPW; r++) {
        const int row = w * RPW + r;
        const i64 j = by * QB + row % QB, k = bz * QB + row / QB;
        const bool in = i < gx && j < gy && k < gz;
        const i64 idx = (k * gy + j) * gx + i;
        #pragma unroll
        for (int c = 0; c < 3; c++) {
            v[r][c] = in ? df[c * plane + idx] : 0.f;
            if (in) {
                mn[c] = fminf(mn[c], v[r][c]);
                mx[c] = fmaxf(mx[c], v[r][c]);
            }
        }
    }
    #pragma unroll
    for (int c = 0; c < 3; c++) {
        #pragma unroll
        for (int o = 1; o < QB; o <<= 1) {  // the 8 lanes of a block
            mn[c] = fminf(mn[c], __shfl_xor(mn[c], o, 64));
            mx[c] = fmaxf(mx[c], __shfl_xor(mx[c], o, 64));
        }
        if ((lane & (QB - 1)) == 0) {
            S.red[0][c][w][blk] = mn[c];
            S.red[1][c][w][blk] = mx[c];
        }
    }
    __syncthreads();
    if (threadIdx.x < 3 * QR) {
        const int c = threadIdx.x % 3, bk = threadIdx.x / 3;
        const i64 bx = rx * QR + bk;
        const float lo = fminf(fminf(S.red[0][c][0][bk], S.red[0][c][1][bk]), fminf(S.red[0][c][2][bk], S.red[0][c][3][bk]));
        const float hi = fmaxf(fmaxf(S.red[1][c][0][bk], S.red[1][c][1][bk]), fmaxf(S.red[1][c][2][bk], S.red[1][c][3][bk]));
        const float off = 0.5f * (lo + hi);
        float sc = fmaxf(hi - off, off - lo);
        if (!(sc > 0.f)) sc = 1.f;
        S.prm[bk][c] = off;
        S.prm[bk][3 + c] = sc;
        if (bx < qbx) {
            const i64 b = (bz * qby + by) * qbx + bx;
            qtab[b * 6 + c] = off;
            qtab[b * 6 + 3 + c] = sc;
        }
    }
    __syncthreads();
    #pragma unroll
    for (int r = 0; r < RPW; r++) {
        const int row = w * RPW + r;
        const i64 j = by * QB + row % QB, k = bz * QB + row / QB;
        if (i < gx && j < gy && k < gz) {
            const i64 idx = (k * gy + j) * gx + i;
            #pragma unroll
            for (int c = 0; c < 3; c++) {
                const float h = (v[r][c] - S.prm[blk][c]) / S.prm[blk][3 + c];
                out[c * plane + idx] = __float2half_rn(fminf(1.f, fmaxf(-1.f, h)));
            }
        }
    }
}

// ---- batches: every system's grid in one launch ---------------------------
// Work-group w serves tile w - pref[j] of job j, pref[j] <= w < pref[j + 1]
// (pref: exclusive prefix sums of the jobs' tile counts).
struct GridJob {
    GridArgs a;         // a.out: the FP32 planes
    int tiles_x, tiles_y;
    i64 qbx, qby;       // FP16 block grid
    __half* h;          // FP16 planes, or null (FP32 grid)
    float* qtab;
};

__device__ __forceinline__ int job_of(const int* __restrict__ pref, int njobs, int w) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pref[mid] <= w) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_grid_batch(const GridJob* __restrict__ jobs, const int* __restrict__ pref,
                                                    int njobs) {
    __shared__ __align__(16) float syz[GT][GRS];
    const int w = blockIdx.x;
    const int j = job_of(pref, njobs, w);
    const GridJob& J = jobs[j];
    const int t = w - pref[j];
    const int bx = t % J.tiles_x, by = (t / J.tiles_x) % J.tiles_y, bz = t / (J.tiles_x * J.tiles_y);
    grid_tiled_body<GNP_BATCH>(J.a, bx, by, bz, syz);
}

__global__ __launch_bounds__(256) void k_quantize_batch(const GridJob* __restrict__ jobs, const int* __restrict__ pref,
                                                        int njobs) {
    __shared__ QuantLds S;
    const int w = blockIdx.x;
    const int j = job_of(pref, njobs, w);
    const GridJob& J = jobs[j];
    quantize_body(J.a.out, J.h, J.qtab, J.a.gx, J.a.gy, J.a.gz, J.qbx, J.qby, w - pref[j], S);
}

}  // namespace

int tw_grid_quantize_f16(tw_ctx* ctx, const float* df32, tw_grid* grid) {
    const i64 qbx = (grid->gx + QB - 1) / QB, qby = (grid->gy + QB - 1) / QB, qbz = (grid->gz + QB - 1) / QB;
    const i64 nb = qbx * qby * qbz;
    if (!grid->qtab) {
        float* t = nullptr;
        TW_HIP(ctx, hipMalloc(&t, (size_t)nb * 6 * sizeof(float)));
        grid->qtab = t;
    }
    tw_stage_begin(ctx, ST_QUANT);
    const i64 nrx = (qbx + QR - 1) / QR;
    hipLaunchKernelGGL(k_quantize_rows, dim3((unsigned)(nrx * qby * qbz)), dim3(256), 0, ctx->stream, df32,
                       (__half*)grid->data, grid->qtab, grid->gx, grid->gy, grid->gz, qbx, qby, nrx);
    TW_LAUNCH_CHECK(ctx);
    tw_stage_end(ctx, ST_QUANT);
    return TW_OK;
}

extern "C" int tw_grid_release(tw_ctx* ctx, tw_grid* grid) {
    if (!grid) return TW_E_ARG;
    if (grid->qtab) {
        if (ctx) {
            (void)hipSetDevice(ctx->device);
            (void)hipStreamSynchronize(ctx->stream);
        }
        (void)hipFree(grid->qtab);
        grid->qtab = nullptr;
    }
    return TW_OK;
}

int tw_grid_eval_dev(tw_ctx* ctx, const float4* pts, const double* cf, const double* d, int K, int precision,
                     tw_grid* grid) {
    int rc;
    const bool exact = precision == TW_EXACT;
    const i64 plane = grid->gx * grid->gy * grid->gz;
    GridArgs a;
    memset(&a, 0, sizeof a);
    a.pts = pts;
    a.c64 = cf;
    a.cf = cf;
    a.K = K;
    a.f = grid->f;
    a.gx = grid->gx; a.gy = grid->gy; a.gz = grid->gz;
    a.ox = grid->ox; a.oy = grid->oy; a.oz = grid->oz;
    a.d = d;

    // FP32 output: the grid itself (device F32), else a workspace.
    float* out32 = nullptr;
    const bool direct = grid->storage == TW_DF_F32 && grid->mem == TW_MEM_DEVICE;
    if (direct) {
        out32 = (float*)grid->data;
    } else {
        void* p = nullptr;
        if ((rc = tw_ws(ctx, WS_GRID, (size_t)plane * 3 * sizeof(float), &p))) return rc;
        out32 = (float*)p;
    }
    a.out = out32;
    tw_stage_begin(ctx, ST_GRID);
    const unsigned nblk = (unsigned)((plane + 255) / 256);
    if (exact) {
        hipLaunchKernelGGL(k_grid_eval<true>, dim3(nblk), dim3(256), 0, ctx->stream, a);
    } else {
        if (grid->gz > 65535 || (grid->gy + GT - 1) / GT > 65535)
            return tw_fail(ctx, TW_E_UNSUPPORTED, "grid too large for one launch");
        const dim3 g((unsigned)((grid->gx + GNP_ONE * GT - 1) / (GNP_ONE * GT)), (unsigned)((grid->gy + GT - 1) / GT), (unsigned)grid->gz);
        hipLaunchKernelGGL(k_grid_tiled, g, dim3(256), 0, ctx->stream, a);
    }
    TW_LAUNCH_CHECK(ctx);
    tw_stage_end(ctx, ST_GRID);

    if (grid->storage == TW_DF_F16) {
        if (grid->mem != TW_MEM_DEVICE) return tw_fail(ctx, TW_E_ARG, "an FP16 grid must live in device memory");
        if ((rc = tw_grid_quantize_f16(ctx, out32, grid))) return rc;
    } else if (!direct) {
        TW_HIP(ctx, hipMemcpyAsync(grid->data, out32, (size_t)plane * 3 * sizeof(float), hipMemcpyDeviceToHost,
                                   ctx->stream));
    }
    return TW_OK;
}

extern "C" int tw_tps_grid_eval(tw_ctx* ctx, const float* centres, int K, const double* c, const double* d,
                                int precision, tw_grid* grid) {
    if (!ctx) return TW_E_ARG;
    if (!centres || !c || !d || !grid || !grid->data) return tw_fail(ctx, TW_E_ARG, "null argument");
    if (K < 1) return tw_fail(ctx, TW_E_ARG, "K must be >= 1");
    if (grid->gx < 1 || grid->gy < 1 || grid->gz < 1 || grid->f < 1)
        return tw_fail(ctx, TW_E_ARG, "grid extent and spacing must be positive");
    if (grid->storage != TW_DF_F32 && grid->storage != TW_DF_F16)
        return tw_fail(ctx, TW_E_ARG, "grid storage must be TW_DF_F32 or TW_DF_F16");
    if (grid->storage == TW_DF_F16 && precision == TW_EXACT)
        return tw_fail(ctx, TW_E_ARG, "TW_EXACT needs an FP32 grid");
    int rc;
    if ((rc = tw_bind(ctx))) return rc;
    tw_stages_reset(ctx);
    tw_stage_begin(ctx, ST_TOTAL);

    // Landmark pack: centres (float4), coefficients (FP64 x3), affine block.
    const size_t pts_bytes = (size_t)K * sizeof(float4);
    const size_t coef_bytes = (size_t)K * 3 * sizeof(double);
    void* pack = nullptr;
    if ((rc = tw_ws(ctx, WS_PTS, pts_bytes + coef_bytes + 16 * sizeof(double), &pack))) return rc;
    std::vector<uint8_t> host(pts_bytes + coef_bytes + 16 * sizeof(double));
    float4* hp = (float4*)host.data();
    for (int n = 0; n < K; n++) hp[n] = make_float4(centres[3 * n], centres[3 * n + 1], centres[3 * n + 2], 0.f);
    double* hc = (double*)(host.data() + pts_bytes);
    for (int n = 0; n < K; n++)
        for (int q = 0; q < 3; q++) hc[3 * n + q] = c[4 * n + 1 + q];
    memcpy(host.data() + pts_bytes + coef_bytes, d, 16 * sizeof(double));
    TW_HIP(ctx, hipMemcpyAsync(pack, host.data(), host.size(), hipMemcpyHostToDevice, ctx->stream));
    const uint8_t* pk = (const uint8_t*)pack;
    rc = tw_grid_eval_dev(ctx, (const float4*)pk, (const double*)(pk + pts_bytes),
                          (const double*)(pk + pts_bytes + coef_bytes), K, precision, grid);
    if (rc) return rc;
    tw_stage_end(ctx, ST_TOTAL);
    // The host pack must outlive the copy; host grids need the result now.
    TW_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return TW_OK;
}

// Grids of a batch of systems: every FP32 field in one launch of the tiled
// kernel, then (FP16 items) every quantisation in one launch.  Per node the
// arithmetic is that of tw_grid_eval_dev.  The job tables are staged in
// host memory that the context keeps until the next batch.
int tw_grid_eval_batch(tw_ctx* ctx, const TwGridItem* items, int n) {
    int rc;
    std::vector<GridJob> jobs((size_t)n);
    std::vector<int> tpref((size_t)n + 1, 0), qpref((size_t)n + 1, 0);
    bool any16 = false;
    for (int i = 0; i < n; i++) {
        const TwGridItem& it = items[i];
        const tw_grid& g = it.grid;
        GridJob& J = jobs[i];
        memset(&J, 0, sizeof J);
        J.a.pts = it.pts;
        J.a.c64 = it.cf;
        J.a.cf = it.cf;
        J.a.K = it.K;
        J.a.f = g.f;
        J.a.gx = g.gx; J.a.gy = g.gy; J.a.gz = g.gz;
        J.a.ox = g.ox; J.a.oy = g.oy; J.a.oz = g.oz;
        J.a.d = it.d;
        J.a.out = it.f32;
        J.tiles_x = (int)((g.gx + GNP_BATCH * GT - 1) / (GNP_BATCH * GT));
        J.tiles_y = (int)((g.gy + GT - 1) / GT);
        J.qbx = (g.gx + QB - 1) / QB;
        J.qby = (g.gy + QB - 1) / QB;
        const i64 qbz = (g.gz + QB - 1) / QB;
        if (g.storage == TW_DF_F16) {
            J.h = (__half*)g.data;
            J.qtab = g.qtab;
            any16 = true;
        }
        const i64 nt = (i64)J.tiles_x * J.tiles_y * g.gz;
        const i64 nq = g.storage == TW_DF_F16 ? J.qbx * J.qby * qbz : 0;
        if (tpref[i] + nt >= (1ll << 31) || qpref[i] + nq >= (1ll << 31))
            return tw_fail(ctx, TW_E_UNSUPPORTED, "batch grids too large for one launch");
        tpref[i + 1] = (int)(tpref[i] + nt);
        qpref[i + 1] = (int)(qpref[i] + nq);
    }
    const size_t jb = ((size_t)n * sizeof(GridJob) + 255) & ~(size_t)255;
    const size_t pb = ((size_t)(n + 1) * sizeof(int) + 255) & ~(size_t)255;
    uint8_t* host = tw_host_stage(ctx, jb + 2 * pb);
    memcpy(host, jobs.data(), (size_t)n * sizeof(GridJob));
    memcpy(host + jb, tpref.data(), (size_t)(n + 1) * sizeof(int));
    memcpy(host + jb + pb, qpref.data(), (size_t)(n + 1) * sizeof(int));
    void* dev = nullptr;
    if ((rc = tw_ws(ctx, WS_BATCH_GTAB, jb + 2 * pb, &dev))) return rc;
    TW_HIP(ctx, hipMemcpyAsync(dev, host, jb + 2 * pb, hipMemcpyHostToDevice, ctx->stream));
    const GridJob* djobs = (const GridJob*)dev;
    const int* dt = (const int*)((uint8_t*)dev + jb);
    const int* dq = (const int*)((uint8_t*)dev + jb + pb);
    tw_stage_begin(ctx, ST_GRID);
    if (tpref[n] > 0) hipLaunchKernelGGL(k_grid_batch, dim3((unsigned)tpref[n]), dim3(256), 0, ctx->stream, djobs, dt, n);
    TW_LAUNCH_CHECK(ctx);
    tw_stage_end(ctx, ST_GRID);
    if (any16 && qpref[n] > 0) {
        tw_stage_begin(ctx, ST_QUANT);
        hipLaunchKernelGGL(k_quantize_batch, dim3((unsigned)qpref[n]), dim3(256), 0, ctx->stream, djobs, dq, n);
        TW_LAUNCH_CHECK(ctx);
        tw_stage_end(ctx, ST_QUANT);
    }
    return TW_OK;
}
