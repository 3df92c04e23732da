// k_solve.hip -- STPS coefficient solve on the GPU.
//
// Reference: the GPU branch of compute_df_stps_subsampled_volume_4bspline
// (q_imgwarp_tps_quicksmallmemory.cpp:449-553: gpu_QR qr.cu:8-131,
// gpu_extendornormal extendornormal.cu:55-150, matrixMultiply aff.cu:37-133,
// gpu_A_i_new GET_A_i.cu:80-161) and the CPU branch (554-586) it mirrors:
//   A = q2' K q2 + 0.2 I,  c = q2 A^-1 q2' Y,  d = r^-1 q1' (Y - K c).
// Because q1' q2 = 0, that (c, d) is the unique solution of the bordered
// system  [[K + lambda I, X], [X', 0]] [c; d] = [Y; 0]  (X = [1 p],
// Y = [1 q]).  This file solves the bordered system directly, which needs
// no orthonormal basis at all:
//
//  1. assemble M' in scaled coordinates p' = (p - mu)/s (K' = K/s,
//     lambda' = lambda/s, X' = [1 p']); the solution maps back through
//     c = c'/s, d = T^-1 d' with X = X' T;
//  2. blocked right-looking LU with partial pivoting (panel of 64 columns in
//     one work-group, row swaps, TRSM of U12, GEMM update of A22);
//  3. blocked forward / backward substitution for the 4 right-hand sides;
//  4. TW_MIXED: the factorisation runs in FP32 and the FP64 residual
//     b - M' x (M' regenerated in FP64 from the landmarks) drives iterative
//     refinement until the correction stalls; if it does not converge the
//     system is refactored in FP64.
//
// K_ij = -|p_i - p_j| is formed from float coordinate differences widened to
// FP64, as in the reference (434-444).
#include "tw_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace {

constexpr int NB = 64;  // panel width / block size

struct SysDesc {
    const float4* pts;   // N centres (x, y, z, 0), original coordinates
    const double* y;     // N x 4 right-hand side rows [1 qx qy qz]
    int N, m;            // m = N + 4
    double inv_s;        // 1/s
    double mu[3];        // centroid
    double lam;          // lambda / s
};

__device__ __forceinline__ double sys_entry(const SysDesc& s, int r, int c) {
    const int N = s.N;
    if (r < N && c < N) {
        const float4 a = s.pts[r], b = s.pts[c];
        const double dx = (float)(a.x - b.x), dy = (float)(a.y - b.y), dz = (float)(a.z - b.z);
        double v = -sqrt(dx * dx + dy * dy + dz * dz) * s.inv_s;
        if (r == c) v += s.lam;
        return v;
    }
    if (r >= N && c >= N) return 0.0;
    const int i = r < N ? r : c;      // landmark index
    const int a = r < N ? c - N : r - N;  // column of X
    if (a == 0) return 1.0;
    const float4 p = s.pts[i];
    const double v = a == 1 ? (double)p.x : (a == 2 ? (double)p.y : (double)p.z);
    return (v - s.mu[a - 1]) * s.inv_s;
}

template <typename T>
__global__ __launch_bounds__(256) void k_assemble(T* __restrict__ A, i64 lda, const SysDesc s) {
    const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
    if (e >= (i64)s.m * s.m) return;
    const int r = (int)(e % s.m), c = (int)(e / s.m);
    A[(i64)c * lda + r] = (T)sys_entry(s, r, c);
}

// b = [Y; 0] in precision T (m x 4, column-major, ld m).
template <typename T>
__global__ __launch_bounds__(256) void k_rhs(T* __restrict__ B, const SysDesc s) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= s.m) return;
    for (int j = 0; j < 4; j++) B[(i64)j * s.m + r] = r < s.N ? (T)s.y[r * 4 + j] : (T)0;
}

// Unblocked GEPP of the panel A[j0:m, j0:j0+jb] in one work-group.
template <typename T>
__global__ __launch_bounds__(1024) void k_panel(T* __restrict__ A, i64 lda, int m, int j0, int jb,
                                                int* __restrict__ ipiv, int* __restrict__ info) {
    __shared__ double rv[1024];
    __shared__ int ri[1024];
    __shared__ int s_prow;
    const int tid = threadIdx.x;
    for (int j = 0; j < jb; j++) {
        const int col = j0 + j;
        T* Acol = A + (i64)col * lda;
        double best = -1.0;
        int bi = m;
        for (int r = col + tid; r < m; r += 1024) {
            const double v = fabs((double)Acol[r]);
            if (v > best) { best = v; bi = r; }
        }
        rv[tid] = best;
        ri[tid] = bi;
        __syncthreads();
        for (int s = 512; s > 0; s >>= 1) {
            if (tid < s) {
                const double v2 = rv[tid + s];
                const int i2 = ri[tid + s];
                if (v2 > rv[tid] || (v2 == rv[tid] && i2 < ri[tid])) { rv[tid] = v2; ri[tid] = i2; }
            }
            __syncthreads();
        }
        if (tid == 0) {
            s_prow = ri[0];
            ipiv[col] = ri[0];
            if (!(rv[0] > 0.0) && *info == 0) *info = col + 1;
        }
        __syncthreads();
        const int prow = s_prow;
        if (prow != col && tid < jb) {
            T* c = A + (i64)(j0 + tid) * lda;
            const T t = c[col];
            c[col] = c[prow];
            c[prow] = t;
        }
        __syncthreads();
        const T piv = Acol[col];
        if (piv != (T)0)
            for (int r = col + 1 + tid; r < m; r += 1024) Acol[r] = Acol[r] / piv;
        __syncthreads();
        const int nrows = m - col - 1, ncols = j0 + jb - col - 1;
        if (nrows > 0 && ncols > 0) {
            const i64 total = (i64)nrows * ncols;
            for (i64 e = tid; e < total; e += 1024) {
                const int cc = col + 1 + (int)(e / nrows);
                const int r = col + 1 + (int)(e % nrows);
                T* Ac = A + (i64)cc * lda;
                Ac[r] -= Acol[r] * Ac[col];
            }
        }
        __syncthreads();
    }
}

// Apply the panel's row swaps to every column outside it.
template <typename T>
__global__ __launch_bounds__(256) void k_swap(T* __restrict__ A, i64 lda, int ncols, int j0, int jb,
                                              const int* __restrict__ ipiv) {
    int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= ncols - jb) return;
    if (c >= j0) c += jb;
    T* Ac = A + (i64)c * lda;
    for (int j = 0; j < jb; j++) {
        const int p = ipiv[j0 + j];
        if (p != j0 + j) {
            const T t = Ac[j0 + j];
            Ac[j0 + j] = Ac[p];
            Ac[p] = t;
        }
    }
}

// U12 = L11^-1 A12 for 64 columns per work-group.
template <typename T>
__global__ __launch_bounds__(256) void k_trsm(T* __restrict__ A, i64 lda, int ncols, int j0, int jb) {
    __shared__ T Ls[NB * (NB + 1)];
    __shared__ T Xs[NB * (NB + 1)];
    const int tid = threadIdx.x;
    const int c0 = j0 + jb + blockIdx.x * NB;
    for (int e = tid; e < jb * jb; e += 256) {
        const int r = e % jb, k = e / jb;
        Ls[k * (NB + 1) + r] = A[(i64)(j0 + k) * lda + j0 + r];
    }
    for (int e = tid; e < jb * NB; e += 256) {
        const int r = e % jb, cc = e / jb;
        const int col = c0 + cc;
        Xs[cc * (NB + 1) + r] = col < ncols ? A[(i64)col * lda + j0 + r] : (T)0;
    }
    __syncthreads();
    for (int k = 0; k < jb - 1; k++) {
        const int rem = jb - k - 1;
        for (int e = tid; e < rem * NB; e += 256) {
            const int r = k + 1 + e % rem, cc = e / rem;
            Xs[cc * (NB + 1) + r] -= Ls[k * (NB + 1) + r] * Xs[cc * (NB + 1) + k];
        }
        __syncthreads();
    }
    for (int e = tid; e < jb * NB; e += 256) {
        const int r = e % jb, cc = e / jb;
        const int col = c0 + cc;
        if (col < ncols) A[(i64)col * lda + j0 + r] = Xs[cc * (NB + 1) + r];
    }
}

// A22 -= L21 U12 over rows [r0, m) and columns [c0, ncols), inner dim jb.
template <typename T>
__global__ __launch_bounds__(256) void k_gemm_update(T* __restrict__ A, i64 lda, int m, int ncols, int j0, int jb) {
    __shared__ T As[NB][NB + 1];
    __shared__ T Bs[NB][NB + 1];
    const int tid = threadIdx.x;
    const int r0 = j0 + jb, c0 = j0 + jb;
    const int R = r0 + blockIdx.x * NB, C = c0 + blockIdx.y * NB;
    for (int e = tid; e < NB * jb; e += 256) {
        const int r = e % NB, k = e / NB;
        As[k][r] = (R + r < m) ? A[(i64)(j0 + k) * lda + R + r] : (T)0;
    }
    for (int e = tid; e < NB * jb; e += 256) {
        const int k = e % jb, c = e / jb;
        Bs[k][c] = (C + c < ncols) ? A[(i64)(C + c) * lda + j0 + k] : (T)0;
    }
    __syncthreads();
    const int tx = tid % 16, ty = tid / 16;
    T acc[4][4];
    #pragma unroll
    for (int i = 0; i < 4; i++)
        #pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0;
    for (int k = 0; k < jb; k++) {
        T av[4], bv[4];
        #pragma unroll
        for (int i = 0; i < 4; i++) av[i] = As[k][tx + 16 * i];
        #pragma unroll
        for (int j = 0; j < 4; j++) bv[j] = Bs[k][ty + 16 * j];
        #pragma unroll
        for (int i = 0; i < 4; i++)
            #pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = fma(av[i], bv[j], acc[i][j]);
    }
    #pragma unroll
    for (int j = 0; j < 4; j++) {
        const int c = C + ty + 16 * j;
        if (c >= ncols) continue;
        #pragma unroll
        for (int i = 0; i < 4; i++) {
            const int r = R + tx + 16 * i;
            if (r < m) A[(i64)c * lda + r] -= acc[i][j];
        }
    }
}

// ---- solves with the factors, 4 right-hand sides (m x 4, column-major) ----

template <typename T>
__global__ void k_apply_pivots(T* __restrict__ B, int m, const int* __restrict__ ipiv) {
    const int j = threadIdx.x;
    if (j >= 4) return;
    T* b = B + (i64)j * m;
    for (int i = 0; i < m; i++) {
        const int p = ipiv[i];
        if (p != i) {
            const T t = b[i];
            b[i] = b[p];
            b[p] = t;
        }
    }
}

// Diagonal block solve: lower unit (UPPER=false) or upper (UPPER=true).
template <typename T, bool UPPER>
__global__ __launch_bounds__(256) void k_diag_solve(const T* __restrict__ A, i64 lda, T* __restrict__ B, int m,
                                                    int i0, int ib) {
    __shared__ T Ls[NB * (NB + 1)];
    __shared__ T Xs[4 * (NB + 1)];
    const int tid = threadIdx.x;
    for (int e = tid; e < ib * ib; e += 256) {
        const int r = e % ib, k = e / ib;
        Ls[k * (NB + 1) + r] = A[(i64)(i0 + k) * lda + i0 + r];
    }
    for (int e = tid; e < ib * 4; e += 256) {
        const int r = e % ib, j = e / ib;
        Xs[j * (NB + 1) + r] = B[(i64)j * m + i0 + r];
    }
    __syncthreads();
    if (!UPPER) {
        for (int k = 0; k < ib - 1; k++) {
            const int rem = ib - k - 1;
            for (int e = tid; e < rem * 4; e += 256) {
                const int r = k + 1 + e % rem, j = e / rem;
                Xs[j * (NB + 1) + r] -= Ls[k * (NB + 1) + r] * Xs[j * (NB + 1) + k];
            }
            __syncthreads();
        }
    } else {
        for (int k = ib - 1; k >= 0; k--) {
            if (tid < 4) Xs[tid * (NB + 1) + k] = Xs[tid * (NB + 1) + k] / Ls[k * (NB + 1) + k];
            __syncthreads();
            for (int e = tid; e < k * 4; e += 256) {
                const int r = e % k, j = e / k;
                Xs[j * (NB + 1) + r] -= Ls[k * (NB + 1) + r] * Xs[j * (NB + 1) + k];
            }
            __syncthreads();
        }
    }
    for (int e = tid; e < ib * 4; e += 256) {
        const int r = e % ib, j = e / ib;
        B[(i64)j * m + i0 + r] = Xs[j * (NB + 1) + r];
    }
}

// B[rows] -= A[rows, i0:i0+ib] * B[i0:i0+ib] for rows in [ra, rb).
template <typename T>
__global__ __launch_bounds__(256) void k_block_update(const T* __restrict__ A, i64 lda, T* __restrict__ B, int m,
                                                      int i0, int ib, int ra, int rb) {
    const int r = ra + blockIdx.x * 256 + threadIdx.x;
    if (r >= rb) return;
    T acc[4] = {0, 0, 0, 0};
    for (int k = 0; k < ib; k++) {
        const T a = A[(i64)(i0 + k) * lda + r];
        #pragma unroll
        for (int j = 0; j < 4; j++) acc[j] = fma(a, B[(i64)j * m + i0 + k], acc[j]);
    }
    #pragma unroll
    for (int j = 0; j < 4; j++) B[(i64)j * m + r] -= acc[j];
}

// FP64 residual  r = b - M' x  with M' regenerated from the landmarks.
// One work-group per row; also returns max |r| and max |b| per row.
__global__ __launch_bounds__(256) void k_residual(const SysDesc s, const double* __restrict__ x,
                                                  double* __restrict__ res) {
    __shared__ double red[4][256];
    const int r = blockIdx.x, tid = threadIdx.x;
    double acc[4] = {0, 0, 0, 0};
    for (int k = tid; k < s.m; k += 256) {
        const double a = sys_entry(s, r, k);
        if (a != 0.0)
            for (int j = 0; j < 4; j++) acc[j] += a * x[(i64)j * s.m + k];
    }
    for (int j = 0; j < 4; j++) red[j][tid] = acc[j];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (tid < st)
            for (int j = 0; j < 4; j++) red[j][tid] += red[j][tid + st];
        __syncthreads();
    }
    if (tid < 4) {
        const double b = r < s.N ? s.y[r * 4 + tid] : 0.0;
        res[(i64)tid * s.m + r] = b - red[tid][0];
    }
}

template <typename Src, typename Dst>
__global__ void k_convert(const Src* __restrict__ a, Dst* __restrict__ b, i64 n) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = (Dst)a[i];
}

__global__ void k_axpy_f32(double* __restrict__ x, const float* __restrict__ d, i64 n) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i < n) x[i] += (double)d[i];
}

inline unsigned nblk(i64 n, int t = 256) { return (unsigned)((n + t - 1) / t); }

template <typename T>
int lu_factor(tw_ctx* ctx, T* A, int m, int* ipiv, int* info) {
    const i64 lda = m;
    for (int j0 = 0; j0 < m; j0 += NB) {
        const int jb = std::min(NB, m - j0);
        hipLaunchKernelGGL(k_panel<T>, dim3(1), dim3(1024), 0, ctx->stream, A, lda, m, j0, jb, ipiv, info);
        if (m - jb > 0)
            hipLaunchKernelGGL(k_swap<T>, dim3(nblk(m - jb)), dim3(256), 0, ctx->stream, A, lda, m, j0, jb, ipiv);
        const int rest = m - j0 - jb;
        if (rest > 0) {
            hipLaunchKernelGGL(k_trsm<T>, dim3(nblk(rest, NB)), dim3(256), 0, ctx->stream, A, lda, m, j0, jb);
            hipLaunchKernelGGL(k_gemm_update<T>, dim3(nblk(rest, NB), nblk(rest, NB)), dim3(256), 0, ctx->stream, A,
                               lda, m, m, j0, jb);
        }
        TW_LAUNCH_CHECK(ctx);
    }
    return TW_OK;
}

template <typename T>
int lu_solve(tw_ctx* ctx, const T* A, int m, const int* ipiv, T* B) {
    const i64 lda = m;
    hipLaunchKernelGGL(k_apply_pivots<T>, dim3(1), dim3(64), 0, ctx->stream, B, m, ipiv);
    for (int i0 = 0; i0 < m; i0 += NB) {
        const int ib = std::min(NB, m - i0);
        hipLaunchKernelGGL((k_diag_solve<T, false>), dim3(1), dim3(256), 0, ctx->stream, A, lda, B, m, i0, ib);
        if (i0 + ib < m)
            hipLaunchKernelGGL(k_block_update<T>, dim3(nblk(m - i0 - ib)), dim3(256), 0, ctx->stream, A, lda, B, m,
                               i0, ib, i0 + ib, m);
    }
    const int last = ((m - 1) / NB) * NB;
    for (int i0 = last; i0 >= 0; i0 -= NB) {
        const int ib = std::min(NB, m - i0);
        hipLaunchKernelGGL((k_diag_solve<T, true>), dim3(1), dim3(256), 0, ctx->stream, A, lda, B, m, i0, ib);
        if (i0 > 0)
            hipLaunchKernelGGL(k_block_update<T>, dim3(nblk(i0)), dim3(256), 0, ctx->stream, A, lda, B, m, i0, ib, 0,
                               i0);
    }
    TW_LAUNCH_CHECK(ctx);
    return TW_OK;
}

// max |v| over an m x 4 host array
double amax(const std::vector<double>& v) {
    double r = 0;
    for (double x : v) r = std::max(r, std::fabs(x));
    return r;
}

}  // namespace

extern "C" int tw_tps_fit(tw_ctx* ctx, const float* centres, const float* targets, int K, double lambda,
                          int precision, double* c, double* d) {
    if (!ctx) return TW_E_ARG;
    if (!centres || !targets || !c || !d) return tw_fail(ctx, TW_E_ARG, "null argument");
    if (K < 5) return tw_fail(ctx, TW_E_ARG, "the STPS fit needs at least 5 landmarks (got %d)", K);
    if (!(lambda >= 0.0) || !std::isfinite(lambda)) return tw_fail(ctx, TW_E_ARG, "lambda must be finite and >= 0");
    if (precision != TW_EXACT && precision != TW_PLAIN && precision != TW_MIXED)
        return tw_fail(ctx, TW_E_ARG, "unknown precision %d", precision);
    for (int i = 0; i < 3 * K; i++)
        if (!std::isfinite(centres[i]) || !std::isfinite(targets[i]))
            return tw_fail(ctx, TW_E_ARG, "landmark coordinate %d is not finite", i);
    int rc;
    if ((rc = tw_bind(ctx))) return rc;
    tw_stages_reset(ctx);
    tw_stage_begin(ctx, ST_TOTAL);
    tw_stage_begin(ctx, ST_FIT);

    const int N = K, m = N + 4;
    // Scaling: centroid and max deviation of the centres.
    double mu[3] = {0, 0, 0};
    for (int i = 0; i < N; i++)
        for (int a = 0; a < 3; a++) mu[a] += centres[3 * i + a];
    for (int a = 0; a < 3; a++) mu[a] /= N;
    double s = 0;
    for (int i = 0; i < N; i++)
        for (int a = 0; a < 3; a++) s = std::max(s, std::fabs(centres[3 * i + a] - mu[a]));
    if (!(s > 0)) return tw_fail(ctx, TW_E_SINGULAR, "all centres coincide");

    // Device packs: centres (float4), Y rows (double x4).
    const size_t pb = (size_t)N * sizeof(float4), yb = (size_t)N * 4 * sizeof(double);
    void* pack = nullptr;
    if ((rc = tw_ws(ctx, WS_PTS, pb + yb, &pack))) return rc;
    std::vector<uint8_t> host(pb + yb);
    float4* hp = (float4*)host.data();
    double* hy = (double*)(host.data() + pb);
    for (int i = 0; i < N; i++) {
        hp[i] = make_float4(centres[3 * i], centres[3 * i + 1], centres[3 * i + 2], 0.f);
        hy[4 * i] = 1.0;
        for (int a = 0; a < 3; a++) hy[4 * i + 1 + a] = targets[3 * i + a];
    }
    TW_HIP(ctx, hipMemcpyAsync(pack, host.data(), host.size(), hipMemcpyHostToDevice, ctx->stream));
    SysDesc sd;
    sd.pts = (const float4*)pack;
    sd.y = (const double*)((uint8_t*)pack + pb);
    sd.N = N;
    sd.m = m;
    sd.inv_s = 1.0 / s;
    for (int a = 0; a < 3; a++) sd.mu[a] = mu[a];
    sd.lam = lambda / s;

    // Workspaces.
    const bool mixed = precision == TW_MIXED;
    const size_t esz = mixed ? sizeof(float) : sizeof(double);
    void *Aw = nullptr, *Bw = nullptr, *aux = nullptr, *aux2 = nullptr;
    if ((rc = tw_ws(ctx, WS_SOLVE_A, (size_t)m * m * esz, &Aw))) return rc;
    if ((rc = tw_ws(ctx, WS_SOLVE_B, (size_t)m * 4 * sizeof(double) * 2, &Bw))) return rc;
    if ((rc = tw_ws(ctx, WS_SOLVE_AUX, (size_t)(m + 16) * sizeof(int), &aux))) return rc;
    if ((rc = tw_ws(ctx, WS_SOLVE_AUX2, (size_t)m * 4 * sizeof(double), &aux2))) return rc;
    int* ipiv = (int*)aux;
    int* info = ipiv + m;
    double* x64 = (double*)Bw;              // solution, FP64
    double* r64 = x64 + (size_t)m * 4;      // residual, FP64
    TW_HIP(ctx, hipMemsetAsync(info, 0, sizeof(int), ctx->stream));

    std::vector<double> hx((size_t)m * 4), hr((size_t)m * 4);
    auto fetch = [&](const double* dev, std::vector<double>& h) -> int {
        TW_HIP(ctx, hipMemcpyAsync(h.data(), dev, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        TW_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return TW_OK;
    };
    int hinfo = 0;
    double bmax = 0;
    for (int i = 0; i < N; i++)
        for (int a = 0; a < 4; a++) bmax = std::max(bmax, std::fabs(hy[4 * i + a]));

    auto solve_fp64 = [&]() -> int {
        double* A = (double*)Aw;
        if (!mixed) {
            hipLaunchKernelGGL(k_assemble<double>, dim3(nblk((i64)m * m)), dim3(256), 0, ctx->stream, A, (i64)m, sd);
        } else {
            // Refactor in FP64: the FP32 workspace is too small, grow it.
            int r2 = tw_ws(ctx, WS_SOLVE_A, (size_t)m * m * sizeof(double), &Aw);
            if (r2) return r2;
            A = (double*)Aw;
            hipLaunchKernelGGL(k_assemble<double>, dim3(nblk((i64)m * m)), dim3(256), 0, ctx->stream, A, (i64)m, sd);
        }
        TW_HIP(ctx, hipMemsetAsync(info, 0, sizeof(int), ctx->stream));
        int r2 = lu_factor<double>(ctx, A, m, ipiv, info);
        if (r2) return r2;
        TW_HIP(ctx, hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        TW_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (hinfo) return tw_fail(ctx, TW_E_SINGULAR, "singular TPS system (zero pivot at column %d)", hinfo - 1);
        hipLaunchKernelGGL(k_rhs<double>, dim3(nblk(m)), dim3(256), 0, ctx->stream, x64, sd);
        return lu_solve<double>(ctx, A, m, ipiv, x64);
    };

    if (!mixed) {
        if ((rc = solve_fp64())) return rc;
        hipLaunchKernelGGL(k_residual, dim3(m), dim3(256), 0, ctx->stream, sd, x64, r64);
        TW_LAUNCH_CHECK(ctx);
        if ((rc = fetch(r64, hr))) return rc;
        ctx->fit_residual = amax(hr) / bmax;
        ctx->refine_iters = 0;
    } else {
        float* A = (float*)Aw;
        float* b32 = (float*)aux2;
        hipLaunchKernelGGL(k_assemble<float>, dim3(nblk((i64)m * m)), dim3(256), 0, ctx->stream, A, (i64)m, sd);
        if ((rc = lu_factor<float>(ctx, A, m, ipiv, info))) return rc;
        TW_HIP(ctx, hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        TW_HIP(ctx, hipStreamSynchronize(ctx->stream));
        bool ok = hinfo == 0;
        int iters = 0;
        if (ok) {
            hipLaunchKernelGGL(k_rhs<float>, dim3(nblk(m)), dim3(256), 0, ctx->stream, b32, sd);
            if ((rc = lu_solve<float>(ctx, A, m, ipiv, b32))) return rc;
            hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk((i64)m * 4)), dim3(256), 0, ctx->stream, b32,
                               x64, (i64)m * 4);
            double prev = INFINITY;
            ok = false;
            for (iters = 1; iters <= 30; iters++) {
                hipLaunchKernelGGL(k_residual, dim3(m), dim3(256), 0, ctx->stream, sd, x64, r64);
                TW_LAUNCH_CHECK(ctx);
                if ((rc = fetch(r64, hr))) return rc;
                const double rel = amax(hr) / bmax;
                if (rel < 1e-14 || !(rel < 0.5 * prev)) {
                    ok = rel < 1e-10;
                    ctx->fit_residual = rel;
                    break;
                }
                prev = rel;
                hipLaunchKernelGGL((k_convert<double, float>), dim3(nblk((i64)m * 4)), dim3(256), 0, ctx->stream, r64,
                                   b32, (i64)m * 4);
                if ((rc = lu_solve<float>(ctx, A, m, ipiv, b32))) return rc;
                hipLaunchKernelGGL(k_axpy_f32, dim3(nblk((i64)m * 4)), dim3(256), 0, ctx->stream, x64, b32, (i64)m * 4);
            }
        }
        ctx->refine_iters = iters;
        if (!ok) {
            // FP32 factors too inaccurate for this system: FP64 refactor.
            if ((rc = solve_fp64())) return rc;
            hipLaunchKernelGGL(k_residual, dim3(m), dim3(256), 0, ctx->stream, sd, x64, r64);
            if ((rc = fetch(r64, hr))) return rc;
            ctx->fit_residual = amax(hr) / bmax;
            ctx->refine_iters = -iters;  // negative: fell back to FP64 factorisation
        }
    }
    if ((rc = fetch(x64, hx))) return rc;
    // Undo the scaling: c = c'/s, d = T^-1 d'.
    for (int i = 0; i < N; i++)
        for (int j = 0; j < 4; j++) c[4 * i + j] = hx[(size_t)j * m + i] / s;
    for (int j = 0; j < 4; j++) {
        const double d0 = hx[(size_t)j * m + N], d1 = hx[(size_t)j * m + N + 1], d2 = hx[(size_t)j * m + N + 2],
                     d3 = hx[(size_t)j * m + N + 3];
        d[0 * 4 + j] = d0 - (mu[0] * d1 + mu[1] * d2 + mu[2] * d3) / s;
        d[1 * 4 + j] = d1 / s;
        d[2 * 4 + j] = d2 / s;
        d[3 * 4 + j] = d3 / s;
    }
    tw_stage_end(ctx, ST_FIT);
    tw_stage_end(ctx, ST_TOTAL);
    for (int i = 0; i < 4 * N; i++)
        if (!std::isfinite(c[i])) return tw_fail(ctx, TW_E_SINGULAR, "non-finite TPS coefficients");
    return TW_OK;
}
