/*
 * terawarp.h -- C ABI of the MI355X warp path (libterawarp.so).
 *
 * Plain C types only: POD descriptors, caller-owned buffers, int status.
 * Every entry point replaces one group of the reference's `extern "C"` GPU
 * wrappers (declared at q_imgwarp_tps_quicksmallmemory.cpp:23-40 and
 * q_warp_affine_tps.cpp:21-26), which took newmat Matrix& and
 * DisplaceFieldF3D*** arguments, allocated and copied on every call, ended
 * with cudaDeviceReset() and always returned success.  Here:
 *   - a context owns the device, a HIP stream and cached device workspaces;
 *     nothing resets the device;
 *   - a buffer is host memory (the library stages it through the device,
 *     like the reference wrappers did) or device memory (no copies);
 *   - errors come back as negative status codes, with a message available
 *     from tw_last_error();
 *   - calls on one context are serialised by the caller (one context per
 *     host thread); the calls that take a context are asynchronous on the
 *     context's stream only when every buffer is in device memory, and
 *     synchronous otherwise.
 *
 * Coordinates: voxel (x, y, z) of a volume with origin (x0, y0, z0) is at
 * global position (x0 + x, y0 + y, z0 + z).  Volumes are [c][z][y][x] with x
 * fastest.  All transforms are backward maps: output voxel p reads the
 * subject volume at T(p), T mapping target space to subject space.
 */
#ifndef TERAWARP_H
#define TERAWARP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define TW_OK              0
#define TW_E_ARG          -1  /* invalid argument / shape mismatch */
#define TW_E_SINGULAR     -2  /* singular TPS system or degenerate affine fit */
#define TW_E_NOMEM        -3  /* device or host allocation failed */
#define TW_E_DEVICE       -4  /* HIP runtime error / no device */
#define TW_E_IO           -5  /* file I/O or format error */
#define TW_E_UNSUPPORTED  -6  /* valid request this build does not implement */
#define TW_E_NOCONVERGE   -7  /* iterative refinement did not reach FP64 accuracy */

/* ---- enums ------------------------------------------------------------ */
#define TW_U8   1
#define TW_U16  2

#define TW_MEM_HOST    0
#define TW_MEM_DEVICE  1

/* Precision forms.
 * TW_EXACT : FP64 throughout, in the reference's operation order (parity).
 * TW_PLAIN : FP64 TPS solve; FP32 grid field; FP32 B-spline upsample;
 *            FP64 trilinear in the reference's order.
 * TW_MIXED : the FP64 solve up to 4096 projected unknowns (MI355X runs
 *            FP64 vector FMAs at the FP32 rate), FP32 Cholesky + FP64
 *            iterative refinement above; FP32 grid field with FP64
 *            accumulation stored as offset-scaled FP16; FP32 upsample and
 *            trilinear. */
#define TW_EXACT  2
#define TW_PLAIN  0
#define TW_MIXED  1

#define TW_INTERP_LINEAR  0
#define TW_INTERP_NN      1

#define TW_DF_F32  0
#define TW_DF_F16  1

/* How a grid's node displacements are upsampled to voxels (tw_grid.basis).
 * The reference declares both (i_interpmethod_df: 0 linear, 1 B-spline;
 * q_imgwarp_tps_quicksmallmemory.h:14-17, .cpp:1158, 2021-2025) but only
 * implements the B-spline (q_dfblcokinterp_bspline, .cpp:169-203; the
 * linear q_dfblcokinterp_linear is declared and never defined).
 * TW_BASIS_BSPLINE: voxel b*f + t of a block uses nodes b .. b+3 with the
 *   uniform cubic B-spline weights TB[t] (q_nonrigid_ini_bsplinebasis_3D).
 * TW_BASIS_LINEAR: the same voxel interpolates linearly between nodes b+1
 *   and b+2 (the nodes at b*f and (b+1)*f) with weights 1 - t/f and t/f
 *   per axis, i.e. TB[t] = {0, 1 - t/f, t/f, 0}. */
#define TW_BASIS_BSPLINE  0
#define TW_BASIS_LINEAR   1

/* OR into the interp argument of the one-call warps (tw_warp_tps,
 * tw_warp_batch, tw_warp_stream) to upsample their grids with
 * TW_BASIS_LINEAR. */
#define TW_DF_LINEAR      0x100

typedef struct tw_ctx tw_ctx;

/* A volume or a window of one.  full_* is the extent of the whole volume
 * the window belongs to (0 = the window itself); it sets the out-of-bounds
 * test of the resamplers (reference: sz_img_ori in q_imgblockwarp_CYF,
 * q_imgwarp_tps_quicksmallmemory.cpp:2208-2304). */
typedef struct {
    void*   data;
    int64_t nx, ny, nz, nc;
    int64_t x0, y0, z0;
    int64_t full_nx, full_ny, full_nz;
    int32_t dtype;  /* TW_U8 / TW_U16 */
    int32_t mem;    /* TW_MEM_HOST / TW_MEM_DEVICE */
} tw_volume;

/* Coarse displacement grid (reference: Vol3DSimple<DisplaceFieldF3D>,
 * q_littleQuickWarp_common.h:42-70, 16-byte AoS records).  Here three
 * planes sx, sy, sz of gx*gy*gz values, x fastest.  Node (i, j, k) sits at
 * ((i-1)*f + ox, (j-1)*f + oy, (k-1)*f + oz)
 * (q_imgwarp_tps_quicksmallmemory.cpp:628-630, 1126-1143).
 * TW_DF_F16: value = offset[b][c] + scale[b][c] * half, where b indexes
 * blocks of 8x8x8 nodes (nbx = ceil(gx/8), ...); the library allocates and
 * owns the per-block tables of a grid it fills (tw_grid_release). */
typedef struct {
    void*   data;
    int64_t gx, gy, gz;
    int32_t f;                 /* node spacing = B-spline block size */
    int32_t storage;           /* TW_DF_F32 / TW_DF_F16 */
    int64_t ox, oy, oz;
    int32_t mem;               /* where data lives */
    int32_t basis;             /* TW_BASIS_BSPLINE (0) or TW_BASIS_LINEAR */
    float*  qtab;              /* F16: device table, per node block 6 floats
                                  (offset xyz, scale xyz); owned by library */
} tw_grid;

/* Per-call timings of the last call on a context, from HIP events on the
 * context's stream (milliseconds; 0 when a stage did not run). */
typedef struct {
    double fit_ms, grid_ms, quant_ms, resample_ms, h2d_ms, d2h_ms, total_ms;
    int32_t refine_iters;      /* TW_MIXED solve: refinement sweeps used */
    int32_t reserved;
    double  fit_residual;      /* relative FP64 residual of the solved system */
} tw_timings;

/* ---- context ------------------------------------------------------------ */
int         tw_create(int device, tw_ctx** out);
int         tw_destroy(tw_ctx* ctx);
/* Use an existing hipStream_t (as void*); NULL restores the context's own. */
int         tw_set_stream(tw_ctx* ctx, void* hip_stream);
/* The hipStream_t (as void*) the context launches on, so that a caller can
 * order its own streams against it (NULL for a null context). */
void*       tw_get_stream(tw_ctx* ctx);
int         tw_sync(tw_ctx* ctx);
/* 1: record per-stage HIP events on every call (tw_get_timings). */
int         tw_set_timing(tw_ctx* ctx, int enable);
/* Output tile of the fused TPS resample (TW_PLAIN / TW_MIXED, linear
 * interpolation): TW_TILE_GATHER reads every tap from global memory (64x4
 * tiles); TW_TILE_LDS streams each 16x16 column's source box through an LDS
 * ring of planes; TW_TILE_AUTO (default) is TW_TILE_GATHER, the faster on
 * MI355X as measured.  Outputs are identical; only speed differs.  No
 * reference counterpart (the reference has one CPU loop). */
#define TW_TILE_AUTO    0
#define TW_TILE_GATHER  1
#define TW_TILE_LDS     2
int         tw_set_tile(tw_ctx* ctx, int tile);
/* Tunables of the hot paths, per context.  None changes a TW_PLAIN or
 * TW_EXACT result; TW_TUNE_FP32_FACTOR_MIN moves TW_MIXED fits between two
 * factorisations that both meet its 1e-8 bar.  The reference hard-codes its
 * counterparts (threads per block ele.cuh:4, block size and read halo
 * q_imgwarp_tps_quicksmallmemory.cpp:1264,1279); terawarp_amd.tune measures
 * and persists these per device.
 *   TW_TUNE_TILE             TW_TILE_* (as tw_set_tile)
 *   TW_TUNE_TZ               gather tile depth in voxels along z of
 *                            single-volume launches, 1..256; 0: 4 grid cells
 *                            clamped to [16, 64] (default)
 *   TW_TUNE_TZ_BATCH         the same for the persistent launches
 *                            (tw_warp_batch, tw_resample_tps_bricks)
 *   TW_TUNE_FP32_FACTOR_MIN  TW_MIXED: smallest projected system order
 *                            (landmarks - 4) factored in FP32 with FP64
 *                            refinement; smaller ones factor in FP64
 *   TW_TUNE_SOLVE_WAIT_US    bound of each wait between the work-groups of
 *                            the pipelined triangular solve, 0..10^7 µs
 *                            (20000 by default).  A wait that runs out
 *                            (other work held the CUs) turns the pipelined
 *                            solve off for the context and the fit is redone
 *                            with one launch per block, same result to
 *                            rounding; setting this tunable turns it back
 *                            on.
 *   TW_TUNE_SOLVE_PIPELINED  read only: 1 while the context uses the
 *                            pipelined solve, 0 after one of its waits ran
 *                            out. */
#define TW_TUNE_TILE             1
#define TW_TUNE_TZ               2
#define TW_TUNE_FP32_FACTOR_MIN  3
#define TW_TUNE_TZ_BATCH         4
#define TW_TUNE_SOLVE_WAIT_US    5
#define TW_TUNE_SOLVE_PIPELINED  6
int         tw_set_tune(tw_ctx* ctx, int key, int64_t value);
int         tw_get_tune(tw_ctx* ctx, int key, int64_t* value);
int         tw_get_timings(tw_ctx* ctx, tw_timings* out);
const char* tw_last_error(tw_ctx* ctx);
const char* tw_version(void);

/* ---- affine / rigid ---------------------------------------------------- */
/* DLT fit with Hartley normalisation, pts_b ~= M pts_a; M is 4x4 row-major.
 * Replaces q_affine_compute_affinmatrix_3D (q_warp_affine_tps.cpp:649-731);
 * at the plugin call site pts_a = target markers, pts_b = subject markers,
 * so M maps target space to subject space (q_warp_affine_tps.cpp:1340-1351).
 * Host-side math (3N x 13 SVD).  n >= 4. */
int tw_affine_fit(const float* pts_a, const float* pts_b, int n, double* M);

/* Matrix of the README-only modes (README.md:85-100), target -> subject,
 * 4x4 row-major, for a subject volume of size (nx, ny, nz):
 *   mode 0 rotation   : angles para (deg) about x, then y, then z, about the
 *                       image centre ((n-1)/2 per axis)
 *   mode 1 translation: output(p) = subject(p - para)
 *   mode 2 scaling    : output(p) = subject(p / para), about the origin
 * para[3]: per-axis parameters. */
int tw_mode_matrix(int mode, const double* para, int64_t nx, int64_t ny, int64_t nz, double* M);

/* Trilinear (or nearest) affine resample, FP64 in the reference's order,
 * truncating like the reference's affine path (q_warp_affine_tps.cpp:1433).
 * Replaces gpu_interpolation_affine (aff.cu:539-619) and the tera-block
 * gpu_interpolation_new (aff.cu:449-536).  src and dst share dtype and nc. */
int tw_resample_affine(tw_ctx* ctx, const tw_volume* src, tw_volume* dst, const double* M, int interp);

/* ---- TPS ---------------------------------------------------------------- */
/* Regularised 3-D STPS fit, kernel -|r|: solves the bordered system
 *   [[K + lambda I, X], [X', 0]] [c; d] = [Y; 0],  X = [1 centres], Y = [1 targets]
 * whose solution equals the reference's null-space computation
 * (q_imgwarp_tps_quicksmallmemory.cpp:449-586: gpu_QR, gpu_extendornormal,
 * matrixMultiply, gpu_A_i_new; lambda = 0.2 at 581).  centres are the
 * kernel centres (target space; the file-target markers at the plugin call
 * site, plugin_Bigimagewarp.cpp:197), targets the subject-space points.
 * c: K x 4 row-major, d: 4 x 4 row-major, host memory (the reference's
 * xnx4_c, x4x4_d).  Computed by the reference's null-space method
 * (Householder QR of X, Cholesky of q2'(K + lambda I)q2); precision:
 * TW_PLAIN/TW_EXACT (FP64 factor) or TW_MIXED (the FP64 factor below
 * TW_TUNE_FP32_FACTOR_MIN projected unknowns, 15000 by default, the measured
 * crossover; from there the FP32 factor with matrix-core updates + FP64
 * iterative refinement, an FP64 refactor if the refinement stalls).
 * K <= 16384. */
int tw_tps_fit(tw_ctx* ctx, const float* centres, const float* targets, int K, double lambda,
               int precision, double* c, double* d);

/* Evaluate the TPS displacement at every grid node.  Replaces
 * gpu_computedistance (weiyijisuan.cu:118-219).  grid->gx..gz, f, ox..oz,
 * storage and mem are inputs; grid->data must hold 3*gx*gy*gz values of
 * the storage type.  precision TW_EXACT: FP64 in the reference's order;
 * otherwise FP32 node-landmark distances and square roots with the landmark
 * sum accumulated in FP64. */
int tw_tps_grid_eval(tw_ctx* ctx, const float* centres, int K, const double* c, const double* d,
                     int precision, tw_grid* grid);

/* Release library-owned tables of a grid (qtab). */
int tw_grid_release(tw_ctx* ctx, tw_grid* grid);

/* Fused cubic-B-spline upsample of the grid + trilinear/NN resample of src
 * into dst (TPS rounding: +0.5 then truncate).  Replaces
 * gpu_interpolation_stps (chazhi.cu:844-1040) and the tera-block
 * gpu_interpolation (chazhi.cu:557-663).  precision: TW_EXACT (FP64
 * 64-term upsample in the reference's order), TW_PLAIN, TW_MIXED. */
int tw_resample_tps(tw_ctx* ctx, const tw_volume* src, tw_volume* dst, const tw_grid* grid,
                    int interp, int precision);

/* Dense per-voxel displacement of the upsample over dst's window: three
 * float planes of nx*ny*nz (device or host per mem).  Stage check only. */
int tw_upsample_df(tw_ctx* ctx, const tw_grid* grid, int64_t nx, int64_t ny, int64_t nz,
                   int64_t x0, int64_t y0, int64_t z0, int precision, float* out, int mem);

/* Whole in-memory chain of imgwarp_smallmemory_image's GPU branch
 * (q_imgwarp_tps_quicksmallmemory.cpp:1997-2162): fit, grid, resample.
 * Grid spacing f (the reference uses 4), lambda 0.2.  Every stage stays on
 * the device (the coefficients and the grid never cross to the host) and
 * the stream is synchronised once, at the end; a failed fit is reported
 * then and leaves dst unspecified.  tw_get_timings gives each stage. */
int tw_warp_tps(tw_ctx* ctx, const tw_volume* src, tw_volume* dst, const float* centres,
                const float* targets, int K, int f, double lambda, int interp, int precision);

/* The chained warp of tw_warp_tps without any allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and outputs as tw_warp_tps; src and dst must be device
 * buffers.  Every workspace comes from a static 64 MiB device array that each
 * GPU holds from its first call until the process ends, host staging from a
 * static host array; no hipMalloc, hipHostMalloc or hipFree is called.
 * Limits: K <= 1024 landmarks, a source below 2 GiB, and the grid plus the
 * fit within the 64 MiB (about 2.5 M grid nodes for TW_MIXED at K = 1024,
 * e.g. 1024 x 1024 x 256 at f = 8 uses 0.6 M); larger requests are rejected
 * with TW_E_UNSUPPORTED before any work.  Calls on one GPU share its array:
 * calls that overlap on one GPU race. */
int tw_warp_tps_noalloc(tw_ctx* ctx, const tw_volume* src, tw_volume* dst, const float* centres,
                        const float* targets, int K, int f, double lambda, int interp, int precision);

/* ---- batches ------------------------------------------------------------- */
#define TW_WARP_TPS     0
#define TW_WARP_AFFINE  1

/* One warp of a batch.  src is a whole volume (origin 0) and dst a volume or
 * window (origin x0, y0, z0) of the same dtype and channel count, both in
 * device memory.  TW_WARP_TPS: K landmark pairs (host memory, as for
 * tw_tps_fit), fitted with the batch's lambda, grid of the batch's spacing f
 * over dst.  TW_WARP_AFFINE: M, the 3 x 4 row-major target -> subject matrix
 * (rigid transforms and the README modes through tw_mode_matrix). */
typedef struct {
    tw_volume    src;
    tw_volume    dst;
    int32_t      kind;      /* TW_WARP_TPS / TW_WARP_AFFINE */
    int32_t      K;
    const float* centres;   /* TPS: K x 3 */
    const float* targets;   /* TPS: K x 3 */
    double       M[12];     /* AFFINE */
} tw_warp_item;

/* Many independent warps of different sizes, dtypes and modes, one launch per
 * stage: every TPS fit in one launch (a work-group per system up to 256
 * landmarks), every grid in one, and every output voxel of every item in one
 * persistent resample launch that finishes the items in order.  The
 * reference warps one volume or tera block per call (its block driver,
 * q_imgwarp_tps_quicksmallmemory.cpp:1161-1994, calls gpu_interpolation,
 * chazhi.cu:557-663, or gpu_interpolation_new, aff.cu:449-536, per block);
 * each item's output equals tw_warp_tps (TPS, same f, lambda, interp and
 * precision) or tw_resample_affine (AFFINE) on the same arguments.
 * precision: TW_PLAIN or TW_MIXED.  Synchronises once, at the end.
 * status (optional, n entries): per-item TW_OK or error code (a failed fit
 * leaves its item's dst unspecified; the call returns the first error).
 * done_ms (optional, n entries): device-clock time from the batch's first
 * kernel to the end of the item's last output tile. */
int tw_warp_batch(tw_ctx* ctx, const tw_warp_item* items, int n, int f, double lambda, int interp,
                  int precision, int32_t* status, double* done_ms);

/* Many output bricks of one volume against one grid in one persistent launch.
 * Replaces the tera-block drivers' per-block gpu_interpolation calls
 * (chazhi.cu:557-663, called once per block from
 * q_imgwarp_tps_quicksmallmemory.cpp:1309-1848).  srcs[i] is brick i's source
 * window (origin x0/y0/z0 and full_* the whole volume, which sets the bounds
 * test), dsts[i] its output box (global origin); every volume in device
 * memory.  The grid (device, F32 or F16 with TW_MIXED) covers every output
 * box.  Each brick's output equals tw_resample_tps on the same window, box and
 * grid.  precision: TW_PLAIN or TW_MIXED.  Asynchronous on the context's
 * stream. */
int tw_resample_tps_bricks(tw_ctx* ctx, const tw_volume* srcs, const tw_volume* dsts, int n, const tw_grid* grid,
                           int interp, int precision);

/* ---- streaming --------------------------------------------------------- */
/* A volume larger than HBM (or than host memory), warped brick by brick.
 * The reference's tera-block drivers (q_imgwarp_tps_quicksmallmemory.cpp:
 * 1161-1994, q_warp_affine_tps.cpp:851-1329) read each block's window
 * (probing 14 points of the field to guess its extent), warp it and write
 * block files, one block at a time.  Here one fit, one grid over the whole
 * volume (the bricks' grids are its sub-grids: brick origins lie on the node
 * lattice), exact per-brick source windows from the field's bounds over each
 * brick's nodes, and a pipeline of pinned host buffers: host threads fill
 * brick windows through `read` while earlier bricks are copied to the device,
 * resampled and copied back, and a writer thread hands finished bricks to
 * `write`.  The output equals tw_warp_tps on the whole volume (same grid,
 * same per-voxel arithmetic). */

/* Fill the box [x0, x0+nx) x [y0, y0+ny) x [z0, z0+nz) of the subject volume
 * into out ([z][y][x], x fastest).  Called concurrently from several threads
 * for disjoint z-ranges of one window; returns TW_OK or an error code. */
typedef int (*tw_read_fn)(void* user, int64_t x0, int64_t y0, int64_t z0, int64_t nx, int64_t ny, int64_t nz,
                          void* out);
/* Receive output brick `brick` (box origin and size, data [z][y][x]); called
 * from one writer thread, in brick order; data is valid during the call. */
typedef int (*tw_write_fn)(void* user, int64_t brick, int64_t x0, int64_t y0, int64_t z0, int64_t nx, int64_t ny,
                           int64_t nz, const void* data);

/* tw_stream_desc.flags: with TW_EXACT, each sample position is the float sum
 * (p - o + d) + o with o the brick's origin, the order of the reference's
 * tera block warp, whose blocks are the bricks (pos_warp block-local plus
 * the displacement, then plus x_offset: q_imgblockwarp_CYF,
 * q_imgwarp_tps_quicksmallmemory.cpp:2229-2235); without it, p + d as in the
 * in-memory warp.  TW_PLAIN / TW_MIXED form positions in FP32 from the
 * global coordinate either way. */
#define TW_STREAM_BLOCK_ORIGIN 1

typedef struct {
    int64_t      nx, ny, nz;      /* whole volume (one channel) */
    int32_t      dtype;           /* TW_U8 / TW_U16 */
    int32_t      brick;           /* output brick edge; a multiple of f (edge bricks are smaller) */
    int32_t      f;               /* grid spacing */
    int32_t      interp;
    int32_t      precision;       /* TW_EXACT, TW_PLAIN or TW_MIXED */
    int32_t      threads;         /* host threads filling windows */
    int32_t      K;
    int32_t      flags;           /* TW_STREAM_BLOCK_ORIGIN or 0 */
    double       lambda;
    const float* centres;         /* K x 3, host */
    const float* targets;         /* K x 3, host */
    tw_read_fn   read;
    void*        read_user;
    tw_write_fn  write;
    void*        write_user;
} tw_stream_desc;

typedef struct {
    int64_t bricks;
    double  wall_s;               /* whole call */
    double  setup_ms;             /* fit + grid + bounds */
    double  resample_ms;          /* sum of the bricks' resample kernels (device events) */
    double  read_s, write_s;      /* time spent in the callbacks (summed over bricks) */
    double  h2d_bytes, d2h_bytes; /* bytes copied */
    double  max_margin;           /* largest source margin used (voxels past the brick) */
    double  issue_s;              /* host time the issuing thread spent enqueuing the bricks'
                                     copies, kernels and events (its waits on the reader and
                                     writer excluded) */
} tw_stream_stats;

int tw_warp_stream(tw_ctx* ctx, const tw_stream_desc* desc, tw_stream_stats* stats);

/* Procedural subject volume for streaming runs: uniform noise hashed from
 * (seed, x, y, z) plus Gaussian blobs drawn per cell of `cell` voxels; any
 * window holds the same values as the whole volume would. */
typedef struct {
    uint64_t seed;
    int64_t  nx, ny, nz;
    int32_t  dtype;
    int32_t  cell;
    int32_t  blobs_per_cell;
    int32_t  pad;
    float    sigma_lo, sigma_hi, amp_lo, amp_hi, noise;
} tw_synth;

int tw_synth_fill(const tw_synth* s, int64_t x0, int64_t y0, int64_t z0, int64_t nx, int64_t ny, int64_t nz,
                  void* out);
/* tw_read_fn over a tw_synth (user = const tw_synth*). */
int tw_synth_read(void* user, int64_t x0, int64_t y0, int64_t z0, int64_t nx, int64_t ny, int64_t nz, void* out);

/* Position-weighted 64-bit checksum of a buffer: the sum of its 64-bit
 * little-endian words w_i (last one zero-padded) times (2i + 1), modulo 2^64,
 * xor the byte count. */
uint64_t tw_checksum64(const void* data, int64_t bytes);

/* tw_write_fn that records each brick's tw_checksum64 in sums[brick] and
 * copies the bricks listed in keep_ids to keep_ptrs (user = tw_checksum_sink*). */
typedef struct {
    uint64_t*       sums;
    int64_t         nbricks;
    int32_t         bytes_per_voxel;
    int32_t         nkeep;
    const int64_t*  keep_ids;
    void* const*    keep_ptrs;
} tw_checksum_sink;
int tw_checksum_write(void* user, int64_t brick, int64_t x0, int64_t y0, int64_t z0, int64_t nx, int64_t ny,
                      int64_t nz, const void* data);

/* Grid node count per axis for a volume extent n and spacing f:
 * ceil(n / f) + 3 (q_imgwarp_tps_quicksmallmemory.cpp:592). */
int64_t tw_grid_dim(int64_t n, int f);

#ifdef __cplusplus
}
#endif

#endif /* TERAWARP_H */
