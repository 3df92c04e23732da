// k_solve.hip -- STPS coefficient solve on the GPU (null-space method).
//
// Reference: compute_df_stps_subsampled_volume_4bspline
// (q_imgwarp_tps_quicksmallmemory.cpp:449-586).  Its CPU branch computes
//   QRZ(X) -> q1, r;  extend_orthonormal -> q2;
//   A = q2' K q2 + 0.2 I;  c = q2 (A^-1 q2' Y);  d = r^-1 q1' (Y - K c),
// and its GPU branch ports those steps piecewise (gpu_QR qr.cu:8-131,
// gpu_extendornormal extendornormal.cu:55-150, matrixMultiply aff.cu:37-133,
// gpu_A_i_new GET_A_i.cu:80-161).  Same algorithm here, laid out for the GPU:
//
//  * X' = [1 p'] in scaled coordinates p' = (p - mu)/s (K' = K/s,
//    lambda' = lambda/s; the solution maps back through c = c'/s and
//    d = T^-1 d' with X = X' T);
//  * Householder QR of X' (one work-group): Q = H1..H4 = I - V T V' in
//    compact WY form, so q2 is never formed: the projected matrix is
//      B = [Q' (K' + lambda' I) Q]_(4:,4:) = [A' - W V' - V W']_(4:,4:),
//      W = A' V T - 1/2 V (T' V' A' V T),
//    with A' = K' + lambda' I generated on the fly from the landmarks (one
//    matrix-free pass for A' V, one for B);
//  * B is symmetric positive definite (-|r| is conditionally positive
//    definite of order 1 and q2 is orthogonal to the constants), so it is
//    factored by blocked Cholesky with no pivoting: a 64x64 diagonal block
//    factored and inverted in LDS (FP64 on the matrix cores), the panel below
//    it as a product with that inverse, an update of the lower trailing
//    tiles; up to 16 blocks each block column is one fused launch that also
//    runs the forward substitution;
//  * z = B^-1 [Q' Y]_(4:) by block substitution with the inverted diagonal
//    blocks (resident work-groups handing blocks on through flags),
//    c' = Q [0; z], d' = R^-1 [Q' (Y - A' c')]_(0:4);
//  * TW_MIXED factors B in FP32 and refines z with FP64 residuals
//    r = [Q'Y]_(4:) - B z, B applied matrix-free in FP64; a stall of the
//    refinement refactors in FP64.
//
// K_ij = -|p_i - p_j| is formed from float coordinate differences widened to
// FP64, as in the reference (434-444).  The (c, d) so computed is the unique
// solution of the bordered system [[K + lambda I, X], [X', 0]] [c; d] = [Y; 0]
// (independent of the basis chosen for the null space); its FP64 residual
// is reported.
#include "tw_internal.h"

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstring>

namespace {

constexpr int NB = 64;  // Cholesky block size

struct SysDesc {
    const float4* pts;   // N centres (x, y, z, 0), original coordinates
    const double* y;     // N x 4 right-hand side rows [1 qx qy qz] (row-major)
    int N, m;            // m = N + 4
    double inv_s;        // 1/s
    double mu[3];        // centroid
    double lam;          // lambda / s
};

// Entry (r, c) of the scaled bordered matrix [[K' + lambda' I, X'], [X'^T, 0]].
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
    const int i = r < N ? r : c;
    const int a = r < N ? c - N : r - N;
    if (a == 0) return 1.0;
    const float4 p = s.pts[i];
    const double v = a == 1 ? (double)p.x : (a == 2 ? (double)p.y : (double)p.z);
    const double mu = a == 1 ? s.mu[0] : (a == 2 ? s.mu[1] : s.mu[2]);  // no indexed access to s
    return (v - mu) * s.inv_s;
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over a work-group of up to 1024 threads; every thread gets the total.
__device__ __forceinline__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    double t = 0;
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; i++) t += red[i];
    return t;
}

// Value of lane `lane` (wave-uniform) in every lane, through an SGPR.
__device__ __forceinline__ float bcast(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double bcast(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sums of NV values over a work-group of up to 1024 threads; red must hold 17 * NV
// doubles.  Every thread gets the totals in v.
template <int NV>
__device__ __forceinline__ void block_sum_n(double* v, double* red) {
    #pragma unroll
    for (int i = 0; i < NV; i++) v[i] = wave_sum(v[i]);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if (l == 0)
        #pragma unroll
        for (int i = 0; i < NV; i++) red[w * NV + i] = v[i];
    __syncthreads();
    if (threadIdx.x < NV) {
        double t = 0;
        for (int k = 0; k < nw; k++) t += red[k * NV + threadIdx.x];
        red[16 * NV + threadIdx.x] = t;
    }
    __syncthreads();
    #pragma unroll
    for (int i = 0; i < NV; i++) v[i] = red[16 * NV + i];
}

// Small per-fit state shared between kernels (device memory).
struct FitState {
    double tau[4];
    double R[16];   // row-major upper triangle of X' = Q [R; 0]
    double T[16];   // row-major compact-WY factor, Q = I - V T V'
    int info;       // 1: rank-deficient X; 2 + k: non-positive pivot at column k
    int ir_stop;    // refinement: 0 running, 1 converged, 2 stalled
    double rnorm, bnorm;
    double ir_prev; // last relative residual
    int ir_iters;   // corrections applied
    int sync_timeout;  // a bounded wait between work-groups ran out
    double fit_rmax;  // max |residual| of the bordered system
};

// Householder QR of X' (N x 4) and the compact-WY factor T.  One work-group.
template <int NT>
__device__ __forceinline__ void hh_x_body(const SysDesc& s, double* __restrict__ V, FitState* st, double* red, double* sh) {
    const int N = s.N, tid = threadIdx.x;
    // X' into V's storage (column-major N x 4).
    for (int i = tid; i < N; i += NT) {
        const float4 p = s.pts[i];
        V[i] = 1.0;
        V[N + i] = ((double)p.x - s.mu[0]) * s.inv_s;
        V[2 * N + i] = ((double)p.y - s.mu[1]) * s.inv_s;
        V[3 * N + i] = ((double)p.z - s.mu[2]) * s.inv_s;
    }
    double R[16] = {0};
    double tau[4];
    double scale0 = 0.0;
    for (int k = 0; k < 4; k++) {
        double* xk = V + (size_t)k * N;
        __syncthreads();
        double part = 0;
        for (int i = k + 1 + tid; i < N; i += NT) part += xk[i] * xk[i];
        const double xnorm2 = block_sum(part, red);
        const double alpha = xk[k];
        double beta, t, scale;
        if (xnorm2 == 0.0) {
            beta = alpha; t = 0.0; scale = 0.0;
        } else {
            beta = -copysign(sqrt(alpha * alpha + xnorm2), alpha);
            t = (beta - alpha) / beta;
            scale = 1.0 / (alpha - beta);
        }
        tau[k] = t;
        R[k * 4 + k] = beta;
        if (k == 0) scale0 = fabs(beta);
        __syncthreads();
        for (int i = tid; i < N; i += NT) {
            if (i < k) xk[i] = 0.0;
            else if (i == k) xk[i] = 1.0;
            else xk[i] *= scale;
        }
        __syncthreads();
        // Apply H_k to the remaining columns j > k.
        for (int j = k + 1; j < 4; j++) {
            double* xj = V + (size_t)j * N;
            double pw = 0;
            for (int i = k + tid; i < N; i += NT) pw += xk[i] * xj[i];
            const double w = block_sum(pw, red);
            __syncthreads();
            for (int i = k + tid; i < N; i += NT) xj[i] -= t * xk[i] * w;
            __syncthreads();
            if (tid == 0) sh[j] = xj[k];
            __syncthreads();
            R[k * 4 + j] = sh[j];
        }
    }
    // V'V for T.
    double G[16];
    for (int a = 0; a < 4; a++)
        for (int b = a; b < 4; b++) {
            double p = 0;
            for (int i = tid; i < N; i += NT) p += V[(size_t)a * N + i] * V[(size_t)b * N + i];
            G[a * 4 + b] = block_sum(p, red);
            G[b * 4 + a] = G[a * 4 + b];
        }
    if (tid == 0) {
        double T[16] = {0};
        for (int k = 0; k < 4; k++) {
            T[k * 4 + k] = tau[k];
            for (int i = 0; i < k; i++) {
                double acc = 0;
                for (int j = i; j < k; j++) acc += T[i * 4 + j] * G[j * 4 + k];
                T[i * 4 + k] = -tau[k] * acc;
            }
        }
        for (int i = 0; i < 16; i++) { st->R[i] = R[i]; st->T[i] = T[i]; }
        for (int k = 0; k < 4; k++) st->tau[k] = tau[k];
        int bad = 0;
        for (int k = 1; k < 4; k++)
            if (!(fabs(R[k * 4 + k]) > 1e-9 * scale0)) bad = 1;
        st->info = bad;
    }
}

__global__ __launch_bounds__(1024) void k_hh_x(const SysDesc s, double* __restrict__ V, FitState* st) {
    __shared__ double red[16];
    __shared__ double sh[8];
    hh_x_body<1024>(s, V, st, red, sh);
}


// P = A' V (N x 4), A' = K' + lambda' I generated on the fly; one row per work-group.
__global__ __launch_bounds__(256) void k_av(const SysDesc s, const double* __restrict__ V, double* __restrict__ P) {
    __shared__ double red[4][4];
    const int i = blockIdx.x, tid = threadIdx.x, N = s.N;
    double acc[4] = {0, 0, 0, 0};
    for (int j = tid; j < N; j += 256) {
        const double a = sys_entry(s, i, j);
        #pragma unroll
        for (int k = 0; k < 4; k++) acc[k] += a * V[(size_t)k * N + j];
    }
    #pragma unroll
    for (int k = 0; k < 4; k++) acc[k] = wave_sum(acc[k]);
    if ((tid & 63) == 0)
        for (int k = 0; k < 4; k++) red[tid >> 6][k] = acc[k];
    __syncthreads();
    if (tid < 4) P[(size_t)tid * N + i] = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
}

// S = V'P, M = T' S T, W = P T - 1/2 V M.  One work-group.
template <int NT>
__device__ __forceinline__ void build_w_body(const double* __restrict__ V, const double* __restrict__ P, double* __restrict__ W,
                             const FitState* st, int N, double* red, double* M, double* Ts) {
    const int tid = threadIdx.x;
    double S[16];
    for (int i = 0; i < 16; i++) S[i] = 0;
    for (int i = tid; i < N; i += NT) {
        double v[4], p[4];
        for (int k = 0; k < 4; k++) { v[k] = V[(size_t)k * N + i]; p[k] = P[(size_t)k * N + i]; }
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) S[a * 4 + b] += v[a] * p[b];
    }
    block_sum_n<16>(S, red);
    if (tid == 0) {
        double TS[16];
        for (int i = 0; i < 16; i++) Ts[i] = st->T[i];
        for (int a = 0; a < 4; a++)   // (T' S)
            for (int b = 0; b < 4; b++) {
                double acc = 0;
                for (int k = 0; k < 4; k++) acc += Ts[k * 4 + a] * S[k * 4 + b];
                TS[a * 4 + b] = acc;
            }
        for (int a = 0; a < 4; a++)   // (T' S) T
            for (int b = 0; b < 4; b++) {
                double acc = 0;
                for (int k = 0; k < 4; k++) acc += TS[a * 4 + k] * Ts[k * 4 + b];
                M[a * 4 + b] = acc;
            }
    }
    __syncthreads();
    for (int i = tid; i < N; i += NT) {
        double p[4], v[4];
        for (int k = 0; k < 4; k++) { p[k] = P[(size_t)k * N + i]; v[k] = V[(size_t)k * N + i]; }
        for (int b = 0; b < 4; b++) {
            double acc = 0;
            for (int k = 0; k < 4; k++) acc += p[k] * Ts[k * 4 + b] - 0.5 * v[k] * M[k * 4 + b];
            W[(size_t)b * N + i] = acc;
        }
    }
}

__global__ __launch_bounds__(1024) void k_build_w(const double* __restrict__ V, const double* __restrict__ P,
                                                  double* __restrict__ W, const FitState* st, int N) {
    __shared__ double red[17 * 16];
    __shared__ double M[16], Ts[16];
    build_w_body<1024>(V, P, W, st, N, red, M, Ts);
}


// B = [A' - W V' - V W']_(4:,4:), n x n column-major in precision T.
template <typename T>
__global__ __launch_bounds__(256) void k_build_b(const SysDesc s, const double* __restrict__ V,
                                                 const double* __restrict__ W, T* __restrict__ B) {
    const int n = s.N - 4, N = s.N;
    const i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
    if (e >= (i64)n * n) return;
    const int i = (int)(e % n) + 4, j = (int)(e / n) + 4;
    double v = sys_entry(s, i, j);
    #pragma unroll
    for (int k = 0; k < 4; k++)
        v -= W[(size_t)k * N + i] * V[(size_t)k * N + j] + V[(size_t)k * N + i] * W[(size_t)k * N + j];
    B[e] = (T)v;
}

// ---- blocked Cholesky (lower), column-major n x n -----------------------
// Each 64 x 64 diagonal block is factored and inverted in registers; Dinv
// holds, per block b, X = L_bb^-1 as a row-major 64 x 64 tile (rows and
// columns past the matrix are the identity).  With X the panel solve is a
// product and the triangular solves are block products.


// Keeps a step's arithmetic on the 16 values inside the step: the
// scheduler may not sink it past the step's barrier (which would keep the
// step's loaded operands alive across many steps).
template <typename T>
__device__ __forceinline__ void pin16(T* v) {
    asm volatile("" : "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3]), "+v"(v[4]), "+v"(v[5]), "+v"(v[6]), "+v"(v[7]));
    asm volatile("" : "+v"(v[8]), "+v"(v[9]), "+v"(v[10]), "+v"(v[11]), "+v"(v[12]), "+v"(v[13]), "+v"(v[14]),
                 "+v"(v[15]));
}

// Lane s's value in all four lanes of each quad (DPP quad permute; s is a
// constant once the caller's loop is unrolled).
__device__ __forceinline__ int quad_bcast32(int v, int s) {
    switch (s) {
        case 0: return __builtin_amdgcn_mov_dpp(v, 0x00, 0xF, 0xF, true);
        case 1: return __builtin_amdgcn_mov_dpp(v, 0x55, 0xF, 0xF, true);
        case 2: return __builtin_amdgcn_mov_dpp(v, 0xAA, 0xF, 0xF, true);
        default: return __builtin_amdgcn_mov_dpp(v, 0xFF, 0xF, 0xF, true);
    }
}
__device__ __forceinline__ float quad_bcast(float v, int s) { return __int_as_float(quad_bcast32(__float_as_int(v), s)); }
__device__ __forceinline__ double quad_bcast(double v, int s) {
    const long long b = __double_as_longlong(v);
    const int lo = quad_bcast32((int)(b & 0xffffffff), s), hi = quad_bcast32((int)(b >> 32), s);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// 1/d in a few instructions: the hardware reciprocal and one Newton step
// (d > 0 and normal here; the result is within an ulp of 1/d).
__device__ __forceinline__ double recip(double d) {
    double r = __builtin_amdgcn_rcp(d);
    const double e = fma(-d, r, 1.0);
    return fma(r, e, r);
}
__device__ __forceinline__ float recip(float d) {
    float r = __builtin_amdgcn_rcpf(d);
    const float e = fmaf(-d, r, 1.0f);
    return fmaf(r, e, r);
}

typedef double v4d __attribute__((ext_vector_type(4)));

// D += A B for a 16 x 4 A and a 4 x 16 B on one wave's FP64 matrix core:
// lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16], and item m of
// the result holds D[4m + l / 16][l % 16], so item s is lane l's B operand
// for rows 4s..4s+3 of a following product.
__device__ __forceinline__ v4d mfma16x4(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Factor and invert the diagonal block at j0.  Four waves; lane r owns row
// r; the 16 blocks of 4 columns are dealt round-robin, wave w holding
// blocks b = 4t + w (columns 16t + 4w + j).  The factor is built as
// L~ D^-1 L~' (L~[r][k] the Schur-complement column, d_k = L~[k][k]); per
// block step the owner factors its 4 columns locally (pivots through
// readlane), publishes them through LDS, and every wave applies the 4
// rank-1 updates to its columns right of the block: 16 barriers in all.
// L = L~ D^-1/2 at the end (lane r forms d_r^-1/2 once; column c takes
// lane c's value).  Then X = L^-1 by columns, right-looking: column
// c = 16w + r/4 belongs to the quad r/4 of wave w, whose lane p keeps the
// running sums acc_i = sum_(j<q) L[4i+p][j] x[j] of rows 4i + p.  At row q
// the lane holding row q forms x[q] = (delta_qc - acc) / L[q][q], the quad
// shares it (DPP), and every lane adds L[4i+p][q] x[q] to its rows' sums,
// so a row's critical path is a subtract, a multiply, the exchange and one
// FMA.
constexpr int LS = NB + 2;  // row stride of the diagonal block's L (16-byte aligned rows)

template <typename T>
struct PotrfLds {
    T colb[2][4][NB];                  // L~[r][k] (0 for r < k), by step parity
    __align__(16) T colp[2][4][NB];    // L~[c][k] / d_k for columns c right of the block, else 0
    T dsh[NB];                         // pivots d_k
    T rdsh[NB];                        // 1 / d_k (FP64 factor)
    union {
        // Lc[q][16p + i] = L[4i + p][q]: lane p of a quad reads its rows' terms contiguously.
        __align__(16) T Lc[NB][LS];    // L by columns
        __align__(16) T Am[NB][LS];    // FP64 factor: the working block, row-major
    };
    __align__(16) T Xs[NB][LS];        // X = L^-1, row-major (FP64 blocked inverse)
    T rsqsh[NB];                       // 1 / L[c][c] = d_c^-1/2
    int badw[4];                       // first non-positive pivot each wave met (NB: none)
};

// X = L^-1 in 16 x 16 blocks (FP64), L in Lc (by columns, see PotrfLds)
// and 1 / L[q][q] in rsq.  The diagonal blocks X_II = L_II^-1 by
// substitution, wave I: column 16I + r/4 in the quad r/4, lane p keeping
// rows 4i + p, the quad sharing each solved row through DPP.  Then
// X_IJ = -X_II sum_(K=J..I-1) L_IK X_KJ on the matrix cores, by distance
// I - J: (1,0) (2,1) (3,2), then (2,0) (3,1), then (3,0), whose first two
// terms wave 3 forms during the earlier passes.  X goes to Xs (for the
// later passes) and to D (row-major, the upper blocks zero).
__device__ __forceinline__ void inverse_blocked(const double (*Lc)[LS], double (*Xs)[LS], const double* rsq,
                                                double* __restrict__ D, int r, int w) {
    {
        const int cc = r >> 2, p = r & 3;
        double acc[4] = {0, 0, 0, 0}, xr[4] = {0, 0, 0, 0};
        #pragma unroll
        for (int q = 0; q < 16; q++) {
            const int iq = q >> 2, pq = q & 3;
            const double xq = quad_bcast(((q == cc ? 1.0 : 0.0) - acc[iq]) * rsq[16 * w + q], pq);
            xr[iq] = p == pq ? xq : xr[iq];
            const double* lq = &Lc[16 * w + q][16 * p + 4 * w];  // L[16w + 4i + p][16w + q] at i
            #pragma unroll
            for (int i = iq & ~1; i < 4; i++) acc[i] = fma(lq[i], xq, acc[i]);
        }
        #pragma unroll
        for (int i = 0; i < 4; i++) {
            const int row = 16 * w + 4 * i + p;
            Xs[row][16 * w + cc] = xr[i];
            D[row * NB + 16 * w + cc] = xr[i];
            for (int J = w + 1; J < 4; J++) D[row * NB + 16 * J + cc] = 0;
        }
    }
    __syncthreads();
    const int lr = r & 15, lk = r >> 4;
    auto lx = [&](int I, int K, int J, v4d y) {  // y + L_IK X_KJ
        #pragma unroll
        for (int s = 0; s < 4; s++) {
            const int kk = 16 * K + 4 * s + lk;
            y = mfma16x4(Lc[kk][16 * (lr & 3) + 4 * I + (lr >> 2)], Xs[kk][16 * J + lr], y);
        }
        return y;
    };
    auto finish = [&](int I, int J, v4d y) {  // X_IJ = -X_II y
        v4d o = {0, 0, 0, 0};
        #pragma unroll
        for (int s = 0; s < 4; s++) o = mfma16x4(Xs[16 * I + lr][16 * I + 4 * s + lk], y[s], o);
        #pragma unroll
        for (int m = 0; m < 4; m++) {
            const int row = 16 * I + 4 * m + lk;
            Xs[row][16 * J + lr] = -o[m];
            D[row * NB + 16 * J + lr] = -o[m];
        }
    };
    const v4d z = {0, 0, 0, 0};
    v4d y30 = z;
    if (w < 3) finish(w + 1, w, lx(w + 1, w, w, z));
    else y30 = lx(3, 0, 0, z);
    __syncthreads();
    if (w < 2) finish(w + 2, w, lx(w + 2, w + 1, w, lx(w + 2, w, w, z)));
    else if (w == 3) y30 = lx(3, 1, 0, y30);
    __syncthreads();
    if (w == 3) finish(3, 0, lx(3, 2, 0, y30));
}

// FP64 diagonal block with contiguous ownership: wave s owns columns
// 16s .. 16s+15 of the block (staged row-major in LDS) and factors them in
// four 4-column steps inside the wave, with no work-group barrier: each
// step eliminates its 4 x 4 diagonal part on wave-uniform values (each
// pivot's reciprocal waits only on the previous one), forms the columns
// lane-parallel, writes L~ back into the block, and drops the step from the
// wave's later columns (the multipliers go through a wave-private LDS row,
// read in order by the same wave).  Then the trailing 16 x 16 tiles of the
// later sub-blocks take the rank-16 update on the matrix cores (A = the
// sub-block's L~, B = L~ / d), dealt over the four waves: four barriers per
// block instead of sixteen, and no broadcast reads of pivot rows.  L, the
// pivots and X = L^-1 follow as in potrf_diag_body.
// staged: the caller has written the block into S.Am (identity past jb).
__device__ __forceinline__ void potrf_diag_fp64(double* __restrict__ B, double* __restrict__ Dinv, int n, int j0,
                                                int jb, FitState* st, PotrfLds<double>& S, bool staged = false) {
    auto& Am = S.Am;
    auto& dsh = S.dsh;
    auto& rdsh = S.rdsh;
    auto& mw = S.colp[0];
    const int tid = threadIdx.x, r = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lr = r & 15, lk = r >> 4;
    const bool rin = r < jb;
    if (!staged) {
        double v[NB * NB / 256];  // clamped loads, all in flight at once
        #pragma unroll
        for (int u = 0; u < NB * NB / 256; u++) {  // column-major B: consecutive threads take consecutive rows
            const int e = tid + 256 * u, c = e >> 6, rr = e & 63;
            v[u] = B[(i64)(j0 + min(c, jb - 1)) * n + j0 + min(rr, jb - 1)];
        }
        #pragma unroll
        for (int u = 0; u < NB * NB / 256; u++) {
            const int e = tid + 256 * u, c = e >> 6, rr = e & 63;
            Am[rr][c] = (rr < jb && c < jb) ? v[u] : (rr == c ? 1.0 : 0.0);
        }
    }
    int mybad = NB;  // first non-positive pivot this wave met (wave-uniform)
    #pragma unroll
    for (int s = 0; s < 4; s++) {
        __syncthreads();
        if (w == s) {
            double a[16];  // column 16s + c at row r
            #pragma unroll
            for (int c = 0; c < 16; c++) a[c] = Am[r][16 * s + c];
            #pragma unroll
            for (int t = 0; t < 4; t++) {
                const int k0 = 16 * s + 4 * t;
                double* ab = &a[4 * t];
                const double s00 = bcast(ab[0], k0), s10 = bcast(ab[0], k0 + 1), s20 = bcast(ab[0], k0 + 2);
                const double s30 = bcast(ab[0], k0 + 3), s11 = bcast(ab[1], k0 + 1), s21 = bcast(ab[1], k0 + 2);
                const double s31 = bcast(ab[1], k0 + 3), s22 = bcast(ab[2], k0 + 2), s32 = bcast(ab[2], k0 + 3);
                const double s33 = bcast(ab[3], k0 + 3);
                const double r0 = recip(s00);
                const double l10 = s10 * r0, l20 = s20 * r0, l30 = s30 * r0;
                const double d1 = fma(-l10, s10, s11), t21 = fma(-l20, s10, s21), t31 = fma(-l30, s10, s31);
                const double t22 = fma(-l20, s20, s22), t32 = fma(-l30, s20, s32), t33 = fma(-l30, s30, s33);
                const double r1 = recip(d1);
                const double l21 = t21 * r1, l31 = t31 * r1;
                const double d2 = fma(-l21, t21, t22), u32 = fma(-l31, t21, t32), u33 = fma(-l31, t31, t33);
                const double r2 = recip(d2);
                const double l32 = u32 * r2;
                const double d3 = fma(-l32, u32, u33);
                const double r3 = recip(d3);
                mybad = min(mybad, s00 > 0.0 ? NB : k0);
                mybad = min(mybad, d1 > 0.0 ? NB : k0 + 1);
                mybad = min(mybad, d2 > 0.0 ? NB : k0 + 2);
                mybad = min(mybad, d3 > 0.0 ? NB : k0 + 3);
                const double lr0 = r >= k0 ? ab[0] : 0.0, m0 = lr0 * r0;
                ab[1] = fma(-m0, s10, ab[1]);
                ab[2] = fma(-m0, s20, ab[2]);
                ab[3] = fma(-m0, s30, ab[3]);
                const double lr1 = r >= k0 + 1 ? ab[1] : 0.0, m1 = lr1 * r1;
                ab[2] = fma(-m1, t21, ab[2]);
                ab[3] = fma(-m1, t31, ab[3]);
                const double lr2 = r >= k0 + 2 ? ab[2] : 0.0, m2 = lr2 * r2;
                ab[3] = fma(-m2, u32, ab[3]);
                const double lr3 = r >= k0 + 3 ? ab[3] : 0.0, m3 = lr3 * r3;
                Am[r][k0 + 0] = lr0;
                Am[r][k0 + 1] = lr1;
                Am[r][k0 + 2] = lr2;
                Am[r][k0 + 3] = lr3;
                if (r < 4) {
                    dsh[k0 + r] = r == 0 ? s00 : r == 1 ? d1 : r == 2 ? d2 : d3;
                    rdsh[k0 + r] = r == 0 ? r0 : r == 1 ? r1 : r == 2 ? r2 : r3;
                }
                if (t < 3) {
                    mw[0][r] = m0;
                    mw[1][r] = m1;
                    mw[2][r] = m2;
                    mw[3][r] = m3;
                    #pragma unroll
                    for (int c = 4 * t + 4; c < 16; c++) {
                        const int cr = 16 * s + c;
                        double v = fma(-lr0, mw[0][cr], a[c]);
                        v = fma(-lr1, mw[1][cr], v);
                        v = fma(-lr2, mw[2][cr], v);
                        a[c] = fma(-lr3, mw[3][cr], v);
                    }
                }
            }
        }
        __syncthreads();
        if (s < 3) {
            // Tiles (I, J), s < J <= I <= 3, dealt to the waves in order.
            int p = 0;
            #pragma unroll
            for (int J = s + 1; J < 4; J++)
                #pragma unroll
                for (int I = J; I < 4; I++, p++) {
                    if ((p & 3) != w) continue;
                    v4d c, d = {0, 0, 0, 0};
                    #pragma unroll
                    for (int m = 0; m < 4; m++) c[m] = Am[16 * I + 4 * m + lk][16 * J + lr];
                    #pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int k = 16 * s + 4 * q + lk;
                        d = mfma16x4(Am[16 * I + lr][k], Am[16 * J + lr][k] * rdsh[k], d);
                    }
                    #pragma unroll
                    for (int m = 0; m < 4; m++) Am[16 * I + 4 * m + lk][16 * J + lr] = c[m] - d[m];
                }
        }
    }
    // L = L~ D^-1/2 (lower part; the rest is zero), row r, columns 16w + c.
    if (r == 0) S.badw[w] = mybad;
    const double rsr = 1.0 / sqrt(dsh[r]);
    if (w == 0) S.rsqsh[r] = rsr;
    double Lv[16];
    #pragma unroll
    for (int c = 0; c < 16; c++) {
        const int cc = 16 * w + c;
        Lv[c] = r >= cc ? Am[r][cc] * bcast(rsr, cc) : 0.0;
        if (rin && cc < jb && cc <= r) B[(i64)(j0 + cc) * n + j0 + r] = Lv[c];
    }
    __syncthreads();  // the block is read: Lc takes its storage
    #pragma unroll
    for (int c = 0; c < 16; c++) S.Lc[16 * w + c][16 * (r & 3) + (r >> 2)] = Lv[c];
    __syncthreads();
    if (threadIdx.x == 0) {
        const int bad = min(min(S.badw[0], S.badw[1]), min(S.badw[2], S.badw[3]));
        if (bad < NB && st->info == 0) st->info = 2 + j0 + bad;
    }
    inverse_blocked(S.Lc, S.Xs, S.rsqsh, Dinv + (i64)(j0 / NB) * NB * NB, r, w);
}

template <typename T>
__device__ __forceinline__ void potrf_diag_body(T* __restrict__ B, T* __restrict__ Dinv, int n, int j0, int jb,
                                                FitState* st, PotrfLds<T>& S) {
    if constexpr (std::is_same<T, double>::value) {
        potrf_diag_fp64(B, Dinv, n, j0, jb, st, S);
        return;
    }
    auto& colb = S.colb;
    auto& colp = S.colp;
    auto& dsh = S.dsh;
    auto& Lc = S.Lc;
    auto& rsqsh = S.rsqsh;
    const int r = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const bool rin = r < jb;
    T a[NB / 4];  // a[4t + j]: column 16t + 4w + j
    {
        const T* col = B + (i64)j0 * n + j0 + r;
        #pragma unroll
        for (int t = 0; t < 4; t++)
            #pragma unroll
            for (int j = 0; j < 4; j++) {
                const int c = 16 * t + 4 * w + j;
                a[4 * t + j] = (rin && c < jb) ? col[(i64)c * n] : (T)(r == c ? 1 : 0);
            }
    }
    int mybad = NB;  // first non-positive pivot this wave met (wave-uniform)
    #pragma unroll
    for (int b = 0; b < NB / 4; b++) {
        const int buf = b & 1, tb = b >> 2, k0 = 4 * b;
        if (w == (b & 3)) {
            // The block's 4 x 4 diagonal part, read once and eliminated on
            // uniform values with the same operations the columns take below,
            // so each pivot's reciprocal waits only on the one before.
            T* ab = &a[4 * tb];
            const T s00 = bcast(ab[0], k0), s10 = bcast(ab[0], k0 + 1), s20 = bcast(ab[0], k0 + 2);
            const T s30 = bcast(ab[0], k0 + 3), s11 = bcast(ab[1], k0 + 1), s21 = bcast(ab[1], k0 + 2);
            const T s31 = bcast(ab[1], k0 + 3), s22 = bcast(ab[2], k0 + 2), s32 = bcast(ab[2], k0 + 3);
            const T s33 = bcast(ab[3], k0 + 3);
            const T r0 = recip(s00);
            const T l10 = s10 * r0, l20 = s20 * r0, l30 = s30 * r0;
            const T d1 = fma(-l10, s10, s11), t21 = fma(-l20, s10, s21), t31 = fma(-l30, s10, s31);
            const T t22 = fma(-l20, s20, s22), t32 = fma(-l30, s20, s32), t33 = fma(-l30, s30, s33);
            const T r1 = recip(d1);
            const T l21 = t21 * r1, l31 = t31 * r1;
            const T d2 = fma(-l21, t21, t22), u32 = fma(-l31, t21, t32), u33 = fma(-l31, t31, t33);
            const T r2 = recip(d2);
            const T l32 = u32 * r2;
            const T d3 = fma(-l32, u32, u33);
            const T r3 = recip(d3);
            // A non-positive pivot marks the factor as failed; what follows
            // it is not used.
            mybad = min(mybad, s00 > (T)0 ? NB : k0);
            mybad = min(mybad, d1 > (T)0 ? NB : k0 + 1);
            mybad = min(mybad, d2 > (T)0 ? NB : k0 + 2);
            mybad = min(mybad, d3 > (T)0 ? NB : k0 + 3);
            // The columns: L~[r][k] for r >= k (0 above), m_rk = L~[r][k] / d_k.
            const T lr0 = r >= k0 ? ab[0] : (T)0, m0 = lr0 * r0;
            ab[1] = fma(-m0, s10, ab[1]);
            ab[2] = fma(-m0, s20, ab[2]);
            ab[3] = fma(-m0, s30, ab[3]);
            const T lr1 = r >= k0 + 1 ? ab[1] : (T)0, m1 = lr1 * r1;
            ab[2] = fma(-m1, t21, ab[2]);
            ab[3] = fma(-m1, t31, ab[3]);
            const T lr2 = r >= k0 + 2 ? ab[2] : (T)0, m2 = lr2 * r2;
            ab[3] = fma(-m2, u32, ab[3]);
            const T lr3 = r >= k0 + 3 ? ab[3] : (T)0, m3 = lr3 * r3;
            const bool below = r > k0 + 3;
            colb[buf][0][r] = lr0;
            colb[buf][1][r] = lr1;
            colb[buf][2][r] = lr2;
            colb[buf][3][r] = lr3;
            colp[buf][0][r] = below ? m0 : (T)0;
            colp[buf][1][r] = below ? m1 : (T)0;
            colp[buf][2][r] = below ? m2 : (T)0;
            colp[buf][3][r] = below ? m3 : (T)0;
            if (r < 4) dsh[k0 + r] = r == 0 ? s00 : r == 1 ? d1 : r == 2 ? d2 : d3;
        }
        __syncthreads();
        // Only columns right of the block change: the wave's groups t > tb,
        // and group tb when the wave comes after the owner.
        #pragma unroll
        for (int j = 0; j < 4; j++) {
            const T l = colb[buf][j][r];
            #pragma unroll
            for (int t = tb; t < 4; t++) {
                if (t == tb && w <= (b & 3)) continue;
                #pragma unroll
                for (int jj = 0; jj < 4; jj++)
                    a[4 * t + jj] = fma(-l, colp[buf][j][16 * t + 4 * w + jj], a[4 * t + jj]);
            }
        }
        pin16(a);
    }
    // L = L~ D^-1/2 (lower part; the rest is zero).
    const T rsr = (T)1 / sqrt(dsh[r]);
    if (w == 0) rsqsh[r] = rsr;
    #pragma unroll
    for (int t = 0; t < 4; t++)
        #pragma unroll
        for (int j = 0; j < 4; j++) {
            const int c = 16 * t + 4 * w + j;
            a[4 * t + j] = r >= c ? a[4 * t + j] * bcast(rsr, c) : (T)0;
            Lc[c][16 * (r & 3) + (r >> 2)] = a[4 * t + j];
        }
    {
        T* col = B + (i64)j0 * n + j0 + r;
        #pragma unroll
        for (int t = 0; t < 4; t++)
            #pragma unroll
            for (int j = 0; j < 4; j++) {
                const int c = 16 * t + 4 * w + j;
                if (rin && c < jb && c <= r) col[(i64)c * n] = a[4 * t + j];
            }
    }
    if (r == 0) S.badw[w] = mybad;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int bad = min(min(S.badw[0], S.badw[1]), min(S.badw[2], S.badw[3]));
        if (bad < NB && st->info == 0) st->info = 2 + j0 + bad;
    }
    if constexpr (std::is_same<T, double>::value) {
        inverse_blocked(S.Lc, S.Xs, rsqsh, Dinv + (i64)(j0 / NB) * NB * NB, r, w);
    } else {
        const int p = r & 3, c = 16 * w + (r >> 2);
        T acc[NB / 4], xr[NB / 4];  // xr[i] = X[4i + p][c]
        #pragma unroll
        for (int i = 0; i < NB / 4; i++) acc[i] = xr[i] = 0;
        // Row q's terms L[4i+p][q] (the rows below q: i >= q/4, from an even i
        // for 16-byte reads) and 1 / L[q][q], read one row ahead.
        T cur[NB / 4], nxt[NB / 4], rc, rn;
        #pragma unroll
        for (int i = 0; i < NB / 4; i++) nxt[i] = 0;
        auto fetch = [&](int q, T* v, T& rq) {
            const T* lq = &Lc[q][16 * p];
            #pragma unroll
            for (int i = (q >> 2) & ~1; i < NB / 4; i++) v[i] = lq[i];
            rq = rsqsh[q];
        };
        fetch(0, cur, rc);
        #pragma unroll
        for (int q = 0; q < NB; q++) {
            const int iq = q >> 2, pq = q & 3;
            if (q + 1 < NB) fetch(q + 1, nxt, rn);
            asm volatile("" ::: "memory");  // the next row's reads are issued here, before this row's work
            const T xq = quad_bcast(((q == c ? (T)1 : (T)0) - acc[iq]) * rc, pq);
            xr[iq] = p == pq ? xq : xr[iq];
            // Rows at or above q take harmless updates: their sums are no longer read.
            #pragma unroll
            for (int i = iq & ~1; i < NB / 4; i++) acc[i] = fma(cur[i], xq, acc[i]);
            #pragma unroll
            for (int i = 0; i < NB / 4; i++) cur[i] = nxt[i];
            rc = rn;
        }
        // Dinv row-major: D[row][c].
        T* D = Dinv + (i64)(j0 / NB) * NB * NB + c;
        #pragma unroll
        for (int i = 0; i < NB / 4; i++) D[(4 * i + p) * NB] = xr[i];
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_potrf_diag(T* __restrict__ B, T* __restrict__ Dinv, int n, int j0, int jb,
                                                    FitState* st) {
    __shared__ PotrfLds<T> S;
    potrf_diag_body<T>(B, Dinv, n, j0, jb, st, S);
}

// Panel below the diagonal block: L21 = B21 X', 64 rows per work-group; wave
// q computes columns c = 4t + q of every row (row i in lane i).
template <typename T>
struct TrsmLds {
    T Xs[NB][NB];
};

// acc: the lane's row of L21, columns 4t + q (for the forward substitution).
template <typename T, typename Lds>
__device__ __forceinline__ void trsm_body(T* __restrict__ B, const T* __restrict__ Dinv, int n, int j0, int jb,
                                          int bx, Lds& S, T (&acc)[NB / 4]) {
    auto& Xs = S.Xs;
    const int lane = threadIdx.x & 63;
    const int q = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = j0 + jb + bx * NB + lane;
    const bool ok = i < n;
    const T* D = Dinv + (i64)(j0 / NB) * NB * NB;
    T xv[NB * NB / 256];
    #pragma unroll
    for (int u = 0; u < NB * NB / 256; u++) xv[u] = D[threadIdx.x + 256 * u];
    T a[NB];
    #pragma unroll
    for (int j = 0; j < NB; j++) a[j] = (ok && j < jb) ? B[(i64)(j0 + j) * n + i] : (T)0;
    #pragma unroll
    for (int u = 0; u < NB * NB / 256; u++) (&Xs[0][0])[threadIdx.x + 256 * u] = xv[u];
    __syncthreads();  // also: all waves have read the panel before any writes it
    #pragma unroll
    for (int t = 0; t < NB / 4; t++) {
        const T* xr = &Xs[4 * t + q][0];
        T s = 0;
        #pragma unroll
        for (int j = 0; j < 4 * t + 4; j++) {
            s = fma(a[j], xr[j], s);
            if ((j & 7) == 7) asm volatile("" : "+v"(s)::"memory");  // bound the LDS reads in flight
        }
        acc[t] = s;
    }
    if (ok)
        #pragma unroll
        for (int t = 0; t < NB / 4; t++)
            if (4 * t + q < jb) B[(i64)(j0 + 4 * t + q) * n + i] = acc[t];
}

template <typename T>
__global__ __launch_bounds__(256) void k_trsm_chol(T* __restrict__ B, const T* __restrict__ Dinv, int n, int j0,
                                                   int jb) {
    __shared__ TrsmLds<T> S;
    T acc[NB / 4];
    trsm_body<T>(B, Dinv, n, j0, jb, blockIdx.x, S, acc);
}

// The FP64 panel with the forward substitution of step k riding on it
// (rhs: n x 4, column-major): every work-group forms y_k = X_k r_k from the
// block's right-hand side (final after the earlier steps; no work-group of
// this launch writes it), work-group 0 stores it to yout, and each drops
// L_Rk y_k from its own rows R.  The pipelined solve then runs only the
// backward pass (and the last block's forward step).  Block k is full: it is
// not the last.
struct TrsmFwdLds {
    double Xs[NB][NB];
    double rk[4][NB];       // r_k, transposed
    double yk[NB][4];
    double part[4][NB][4];  // per wave: the row's sums over its 16 columns
};

__global__ __launch_bounds__(256) void k_trsm_fwd(double* __restrict__ B, const double* __restrict__ Dinv, int n,
                                                  int j0, int jb, double* __restrict__ rhs,
                                                  double* __restrict__ yout) {
    __shared__ TrsmFwdLds S;
    const int lane = threadIdx.x & 63;
    const int q = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = j0 + jb + blockIdx.x * NB + lane;
    S.rk[q][lane] = rhs[(i64)q * n + j0 + lane];  // visible after trsm_body's barrier
    const double ri = i < n ? rhs[(i64)q * n + i] : 0.0;
    double acc[NB / 4];
    trsm_body<double>(B, Dinv, n, j0, jb, blockIdx.x, S, acc);
    {
        // Lane c sums X[c][j] r[j] from j = c on (cyclically): the lanes'
        // LDS reads fall in different banks.
        double s = 0;
        #pragma unroll 8
        for (int j = 0; j < NB; j++) {
            const int jj = (j + lane) & (NB - 1);
            s = fma(S.Xs[lane][jj], S.rk[q][jj], s);
        }
        S.yk[lane][q] = s;
        if (blockIdx.x == 0) yout[(i64)q * n + j0 + lane] = s;
    }
    __syncthreads();
    double p[4] = {0, 0, 0, 0};
    #pragma unroll
    for (int t = 0; t < NB / 4; t++)
        #pragma unroll
        for (int c = 0; c < 4; c++) p[c] = fma(acc[t], S.yk[4 * t + q][c], p[c]);
    #pragma unroll
    for (int c = 0; c < 4; c++) S.part[q][lane][c] = p[c];
    __syncthreads();
    if (i < n)
        rhs[(i64)q * n + i] = ri - ((S.part[0][lane][q] + S.part[1][lane][q]) + (S.part[2][lane][q] + S.part[3][lane][q]));
}

// Lower trailing tiles: B[R, C] -= L21[R] L21[C]' for tile C <= tile R.
template <typename T>
struct SyrkLds {
    T As[NB][NB + 1];  // As[k][r]
    T Bs[NB][NB + 1];  // Bs[k][c]
};

template <typename T>
__device__ __forceinline__ void syrk_body(T* __restrict__ B, int n, int j0, int jb, int bx, int by, SyrkLds<T>& S) {
    auto& As = S.As;
    auto& Bs = S.Bs;
    const int tid = threadIdx.x;
    const int r0 = j0 + jb;
    const int R = r0 + bx * NB, C = r0 + by * NB;
    const int tx = tid % 16, ty = tid / 16;
    T va[NB * NB / 256], vb[NB * NB / 256], old[4][4];
    #pragma unroll
    for (int u = 0; u < NB * NB / 256; u++) {
        const int e = tid + 256 * u, r = e & (NB - 1), k = e / NB;
        va[u] = (k < jb && R + r < n) ? B[(i64)(j0 + k) * n + R + r] : (T)0;
        vb[u] = (k < jb && C + r < n) ? B[(i64)(j0 + k) * n + C + r] : (T)0;
    }
    #pragma unroll
    for (int j = 0; j < 4; j++)
        #pragma unroll
        for (int i = 0; i < 4; i++) {
            const int r = R + tx + 16 * i, c = C + ty + 16 * j;
            old[i][j] = (r < n && c < n && c <= r) ? B[(i64)c * n + r] : (T)0;
        }
    #pragma unroll
    for (int u = 0; u < NB * NB / 256; u++) {
        const int e = tid + 256 * u;
        As[e / NB][e & (NB - 1)] = va[u];
        Bs[e / NB][e & (NB - 1)] = vb[u];
    }
    __syncthreads();
    T acc[4][4];
    #pragma unroll
    for (int i = 0; i < 4; i++)
        #pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0;
    #pragma unroll 8
    for (int k = 0; k < NB; k++) {
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
    for (int j = 0; j < 4; j++)
        #pragma unroll
        for (int i = 0; i < 4; i++) {
            const int r = R + tx + 16 * i, c = C + ty + 16 * j;
            if (r < n && c < n && c <= r) B[(i64)c * n + r] = old[i][j] - acc[i][j];
        }
}

// D += A B for a 16 x 4 A and a 4 x 16 B on one wave's matrix core, FP64 or
// FP32: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]; item m of
// the result holds D[mma_row<T>(m, l / 16)][l % 16] (FP64: 4m + l/16; FP32:
// 4 (l/16) + m, as measured on gfx950).
typedef float v4f __attribute__((ext_vector_type(4)));
template <typename T>
struct Mma;
template <>
struct Mma<double> {
    typedef v4d V;
    static __device__ __forceinline__ V op(double a, double b, V c) { return mfma16x4(a, b, c); }
    static __device__ __forceinline__ int row(int m, int lk) { return 4 * m + lk; }
};
template <>
struct Mma<float> {
    typedef v4f V;
    static __device__ __forceinline__ V op(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int m, int lk) { return 4 * lk + m; }
};

// The trailing update of syrk_body on the matrix cores (FP64 or FP32 with
// FP32 accumulation): tile (bx, by) of B_RC -= L_R L_C', formed transposed
// (D = L_C L_R', wave w owning columns 16w .. 16w + 15 of the tile), so that
// a lane's results run down B's columns: the tile's loads and stores take
// 16 consecutive elements per column.
template <typename T>
__device__ __forceinline__ void syrk_mfma_body(T* __restrict__ B, int n, int j0, int jb, int bx, int by,
                                               SyrkLds<T>& S) {
    typedef typename Mma<T>::V V;
    auto& As = S.As;
    auto& Bs = S.Bs;
    const int tid = threadIdx.x, lr = tid & 15, lk = (tid >> 4) & 3;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int r0 = j0 + jb;
    const int R = r0 + bx * NB, C = r0 + by * NB;
    T va[NB * NB / 256], vb[NB * NB / 256], old[4][4];
    #pragma unroll
    for (int u = 0; u < NB * NB / 256; u++) {
        const int e = tid + 256 * u, r = e & (NB - 1), k = e / NB;
        va[u] = (k < jb && R + r < n) ? B[(i64)(j0 + k) * n + R + r] : (T)0;
        vb[u] = (k < jb && C + r < n) ? B[(i64)(j0 + k) * n + C + r] : (T)0;
    }
    #pragma unroll
    for (int J = 0; J < 4; J++)
        #pragma unroll
        for (int m = 0; m < 4; m++) {
            const int c = C + 16 * w + Mma<T>::row(m, lk), r = R + 16 * J + lr;
            old[J][m] = (r < n && c < n && c <= r) ? B[(i64)c * n + r] : (T)0;
        }
    #pragma unroll
    for (int u = 0; u < NB * NB / 256; u++) {
        const int e = tid + 256 * u;
        As[e / NB][e & (NB - 1)] = va[u];
        Bs[e / NB][e & (NB - 1)] = vb[u];
    }
    __syncthreads();
    V acc[4];
    #pragma unroll
    for (int J = 0; J < 4; J++) acc[J] = V{0, 0, 0, 0};
    #pragma unroll
    for (int s = 0; s < NB / 4; s++) {
        const T a = Bs[4 * s + lk][16 * w + lr];
        #pragma unroll
        for (int J = 0; J < 4; J++) acc[J] = Mma<T>::op(a, As[4 * s + lk][16 * J + lr], acc[J]);
    }
    #pragma unroll
    for (int J = 0; J < 4; J++)
        #pragma unroll
        for (int m = 0; m < 4; m++) {
            const int c = C + 16 * w + Mma<T>::row(m, lk), r = R + 16 * J + lr;
            if (r < n && c < n && c <= r) B[(i64)c * n + r] = old[J][m] - acc[J][m];
        }
}

template <typename T>
__global__ __launch_bounds__(256) void k_syrk(T* __restrict__ B, int n, int j0, int jb) {
    if (blockIdx.y > blockIdx.x) return;
    __shared__ SyrkLds<T> S;
    syrk_mfma_body<T>(B, n, j0, jb, blockIdx.x, blockIdx.y, S);
}

// The FP64 trailing update of step k with the factor of the next diagonal
// block: work-group (0, 0), dispatched first, updates that block's tile and
// then factors and inverts it (k_potrf_diag's arithmetic on the same stored
// values) while the other work-groups update theirs, so the diagonal factor
// leaves the chain of dependent launches.
union SyrkPotrfLds {
    SyrkLds<double> syrk;
    PotrfLds<double> potrf;
};

__global__ __launch_bounds__(256) void k_syrk_potrf(double* __restrict__ B, double* __restrict__ Dinv, int n, int j0,
                                                    int jb, FitState* st) {
    if (blockIdx.y > blockIdx.x) return;
    __shared__ SyrkPotrfLds S;
    syrk_mfma_body<double>(B, n, j0, jb, blockIdx.x, blockIdx.y, S.syrk);
    if (blockIdx.x == 0 && blockIdx.y == 0) {
        const int r0 = j0 + jb;
        __syncthreads();  // the tile's stores are visible to the work-group; its LDS is free
        potrf_diag_fp64(B, Dinv, n, r0, min(NB, n - r0), st, S.potrf);
    }
}

// One right-looking step of the blocked Cholesky in one launch, for small
// systems (latency-bound): after the diagonal block k (rows j0 .. j0+jb-1) is
// factored and inverted, work-group (bx, by) (by <= bx) of the trailing tiles
// forms its two panel blocks L_R = B_Rk X_k' and L_C = B_Ck X_k' itself (no
// separate panel launch; the panel is computed once per tile instead of once
// per row block), writes L_R to the panel store Lp when by == 0, updates its
// tile B_RC -= L_R L_C', and tile (0, 0) then factors and inverts the next
// diagonal block.  The panel input B_Rk is only read in a step (the factor's
// off-diagonal blocks go to Lp, not over B), so tiles of one row block never
// race.  Lp: n x n column-major like B; the diagonal blocks stay in B and
// their inverses in Dinv.  Both products run on the FP64 matrix cores, wave w
// owning rows 16w .. 16w + 15 of the tile (FP64 only: jb = NB).
template <typename T>
struct StepLds {
    union {
        struct {
            T Xs[NB][NB + 1];  // X_k = L_kk^-1, row-major
            T As[NB][NB + 1];  // As[c][r] = L_R[r][c]
            T Bs[NB][NB + 1];  // Bs[c][r] = L_C[r][c]
            T rk[NB][4];       // forward substitution: r_k, then y_k
            T yk[NB][4];
        } u;
        PotrfLds<T> potrf;
    };
};

// Rows 16w .. 16w + 15 of L_R = B_Rk X_k': a[s] is the lane's A operand
// B_Rk[16w + l%16][4s + l/16]; output tile (w, c) sums the chunks s <= 4c + 3
// (X_k' is upper triangular), the four tiles' chains interleaved.
// out[c][r] = L_R[r][c]; Lp, when given, gets L_R.
__device__ __forceinline__ void step_panel(const double (&a)[NB / 4], const double (*Xs)[NB + 1],
                                           double (*out)[NB + 1], double* __restrict__ Lp, int n, int j0, int R,
                                           int w, int lr, int lk) {
    v4d acc[4];
    #pragma unroll
    for (int c = 0; c < 4; c++) acc[c] = v4d{0, 0, 0, 0};
    #pragma unroll
    for (int s = 0; s < NB / 4; s++)
        #pragma unroll
        for (int c = s / 4; c < 4; c++) acc[c] = mfma16x4(a[s], Xs[16 * c + lr][4 * s + lk], acc[c]);
    #pragma unroll
    for (int c = 0; c < 4; c++)
        #pragma unroll
        for (int m = 0; m < 4; m++) {
            const int row = 16 * w + 4 * m + lk, col = 16 * c + lr;
            out[col][row] = acc[c][m];
            if (Lp && R + row < n) Lp[(i64)(j0 + col) * n + R + row] = acc[c][m];
        }
}

template <typename T>
__global__ __launch_bounds__(256) void k_chol_step(T* __restrict__ B, T* __restrict__ Lp, T* __restrict__ Dinv,
                                                   int n, int j0, int jb, FitState* st, T* __restrict__ rhs,
                                                   T* __restrict__ yout) {
    static_assert(std::is_same<T, double>::value, "the fused step runs on the FP64 matrix cores");
    const int bx = blockIdx.x, by = blockIdx.y;
    if (by > bx) return;
    __shared__ StepLds<T> S;
    auto& Xs = S.u.Xs;
    auto& As = S.u.As;
    auto& Bs = S.u.Bs;
    const int tid = threadIdx.x, lr = tid & 15, lk = (tid >> 4) & 3;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int r0 = j0 + jb, R = r0 + bx * NB, C = r0 + by * NB;
    // The wave's A operands of both panels and its part of the tile B_RC, in
    // flight while X_k is staged (rows and columns past n are clamped for the
    // load and zeroed or not stored).
    T ar[NB / 4], ac[NB / 4], old[4][4];
    {
        const int rr = R + 16 * w + lr, rc = C + 16 * w + lr;
        const int rrc = min(rr, n - 1), rcc = min(rc, n - 1);
        #pragma unroll
        for (int s = 0; s < NB / 4; s++) {
            const i64 col = (i64)(j0 + 4 * s + lk) * n;
            const T va = B[col + rrc], vc = B[col + rcc];
            ar[s] = rr < n ? va : (T)0;
            ac[s] = rc < n ? vc : (T)0;
        }
        #pragma unroll
        for (int J = 0; J < 4; J++)
            #pragma unroll
            for (int m = 0; m < 4; m++)
                old[J][m] = B[(i64)min(C + 16 * J + lr, n - 1) * n + min(R + 16 * w + 4 * m + lk, n - 1)];
    }
    const T* D = Dinv + (i64)(j0 / NB) * NB * NB;
    #pragma unroll
    for (int u = 0; u < NB * NB / 256; u++) {
        const int e = tid + 256 * u;
        Xs[e / NB][e % NB] = D[e];
    }
    __syncthreads();
    step_panel(ar, Xs, As, by == 0 ? Lp : nullptr, n, j0, R, w, lr, lk);
    if (by != bx) step_panel(ac, Xs, Bs, nullptr, n, j0, C, w, lr, lk);
    __syncthreads();
    T (*Cs)[NB + 1] = by == bx ? As : Bs;
    // B_RC -= L_R L_C': tiles (w, J).  A diagonal tile needs only J <= w; the
    // waves form all four anyway (wave 3 needs them, so the others cost
    // nothing) and the store keeps the lower part.
    v4d acc[4];
    #pragma unroll
    for (int J = 0; J < 4; J++) acc[J] = v4d{0, 0, 0, 0};
    #pragma unroll
    for (int s = 0; s < NB / 4; s++) {
        const T a = As[4 * s + lk][16 * w + lr];
        #pragma unroll
        for (int J = 0; J < 4; J++) acc[J] = mfma16x4(a, Cs[4 * s + lk][16 * J + lr], acc[J]);
    }
    #pragma unroll
    for (int J = 0; J < 4; J++)
        #pragma unroll
        for (int m = 0; m < 4; m++) {
            const int r = R + 16 * w + 4 * m + lk, c = C + 16 * J + lr;
            if (r < n && c < n && c <= r) B[(i64)c * n + r] = old[J][m] - acc[J][m];
        }
    // The forward substitution rides on the steps (rhs: n x 4, column-major):
    // y_k = X_k r_k (r_k took every earlier step's update), then each trailing
    // row block drops L_Rk y_k.  Work-groups (bx, 0), bx >= 1, take rows R,
    // (1, 0) also rows r0 (its L_C) and writes y_k; with one trailing block
    // (0, 0) does both before it factors.  These work-groups would otherwise
    // idle while (0, 0) factors the next diagonal block.
    const int nt = gridDim.x;
    if (rhs && by == 0 && (bx >= 1 || nt == 1)) {
        auto& rk = S.u.rk;
        auto& yk = S.u.yk;
        const int i = tid & (NB - 1), q = tid >> 6;
        rk[i][q] = rhs[(i64)q * n + j0 + i];  // block k is full (jb = NB)
        __syncthreads();
        T y = 0;
        #pragma unroll 8
        for (int m = 0; m < NB; m++) y = fma(Xs[i][m], rk[m][q], y);
        yk[i][q] = y;
        if (bx == 1 || nt == 1) yout[(i64)q * n + j0 + i] = y;
        __syncthreads();
        auto drop = [&](const T (*Ls)[NB + 1], int R0) {
            if (R0 + i >= n) return;
            T v = rhs[(i64)q * n + R0 + i];
            #pragma unroll 8
            for (int c = 0; c < NB; c++) v = fma(-Ls[c][i], yk[c][q], v);
            rhs[(i64)q * n + R0 + i] = v;
        };
        drop(As, R);
        if (bx == 1) drop(Bs, C);
    }
    if (bx == 0 && by == 0) {
        // The updated tile is the next diagonal block: it goes from the
        // registers into the factor's LDS block (identity past its extent)
        // once the step's LDS is free.
        const int jb1 = min(NB, n - r0);
        __syncthreads();
        #pragma unroll
        for (int J = 0; J < 4; J++)
            #pragma unroll
            for (int m = 0; m < 4; m++) {
                const int rr = 16 * w + 4 * m + lk, cc = 16 * J + lr;
                S.potrf.Am[rr][cc] = (rr < jb1 && cc < jb1) ? old[J][m] - acc[J][m] : (rr == cc ? 1.0 : 0.0);
            }
        potrf_diag_fp64(B, Dinv, n, r0, jb1, st, S.potrf, true);
    }
}

// Systems of at most this many blocks factor with the fused steps.
constexpr int STEP_MAX_BLOCKS = 16;

constexpr int SOLVE_THREADS = 512;
constexpr int BUF_FLAGS = 0x00020000;  // raw buffer, gfx9-family data format

// One float / double through a buffer resource: byte offsets voff (per
// lane) + soff (wave-uniform).
template <typename T>
__device__ __forceinline__ T ld_buf(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
    } else {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
    }
}

// Solve L L' x = b for 4 right-hand sides (n x 4 column-major, in place).
// One work-group.  Per 64-row block: wave 0 multiplies the block's
// right-hand side by X (forward) or X' (backward), then every thread
// updates one remaining row with the block's 64 entries of L.  Those
// entries do not depend on the solution, so they are loaded (together with
// the block's X) before the block's product and their latency overlaps it.
// skip (optional): a nonzero *skip makes the call a no-op.
template <typename T>
struct SolveLds {
    T Xs[NB][NB + 1];
    T vk[NB][4];
    T yk[NB][4];
};

template <typename T, int NT>
__device__ __forceinline__ void chol_solve_body(const T* __restrict__ L, const T* __restrict__ Dinv, int n,
                                                T* __restrict__ b, SolveLds<T>& S) {
    static_assert(NT >= 4 * NB && NT % NB == 0, "one thread per right-hand-side entry of a block");
    auto& Xs = S.Xs;
    auto& vk = S.vk;
    auto& yk = S.yk;
    const int tid = threadIdx.x, lane = tid & 63;
    const int nb = (n + NB - 1) / NB;
    // L through a buffer resource (the host keeps n * n * sizeof(T) < 2^31).
    const uint32_t colb = (uint32_t)n * (uint32_t)sizeof(T);
    const __amdgpu_buffer_rsrc_t rl =
        __builtin_amdgcn_make_buffer_rsrc((void*)L, (short)0, (int)(colb * (uint32_t)n), BUF_FLAGS);
    for (int pass = 0; pass < 2; pass++) {
        for (int s = 0; s < nb; s++) {
            const int kb = pass == 0 ? s : nb - 1 - s;
            const int k0 = kb * NB, ib = min(NB, n - k0);
            // Rows to update: below the block (forward; ib == NB whenever
            // rows remain) or above it (backward).
            const int r0 = pass == 0 ? k0 + NB : 0, r1 = pass == 0 ? n : k0;
            int i = r0 + tid;
            T l[NB];
            auto load_row = [&](int row) {
                if (pass == 0) {
                    // L[row][k0 + c]: one lane offset, the column in the scalar offset.
                    const uint32_t v = (uint32_t)row * (uint32_t)sizeof(T);
                    #pragma unroll
                    for (int c = 0; c < NB; c++) l[c] = ld_buf<T>(rl, v, (uint32_t)(k0 + c) * colb);
                } else {
                    // Column `row` of L, rows k0..k0+ib-1 (contiguous).
                    const uint32_t v = ((uint32_t)row * (uint32_t)n + (uint32_t)k0) * (uint32_t)sizeof(T);
                    #pragma unroll
                    for (int c = 0; c < NB; c++) l[c] = c < ib ? ld_buf<T>(rl, v + (uint32_t)(c * sizeof(T)), 0u) : (T)0;
                }
            };
            // The block's X and right-hand side first: loads retire in order,
            // so the row loads issued after them stay in flight past the
            // barrier.
            const T* D = Dinv + (i64)kb * NB * NB;
            T xv[NB * NB / NT];
            #pragma unroll
            for (int u = 0; u < NB * NB / NT; u++) xv[u] = D[tid + NT * u];
            T bv = 0;
            if (tid < 4 * NB) {
                const int r = tid & (NB - 1), j = tid / NB;
                bv = r < ib ? b[(i64)j * n + k0 + r] : (T)0;
            }
            if (i < r1) load_row(i);
            if (tid < 4 * NB) vk[tid & (NB - 1)][tid / NB] = bv;
            #pragma unroll
            for (int u = 0; u < NB * NB / NT; u++) {
                const int e = tid + NT * u;
                Xs[e / NB][e & (NB - 1)] = xv[u];
            }
            __syncthreads();
            if (tid < 64) {
                T y[4] = {0, 0, 0, 0};
                #pragma unroll 4
                for (int c = 0; c < NB; c++) {
                    const T xe = pass == 0 ? Xs[lane][c] : Xs[c][lane];
                    #pragma unroll
                    for (int j = 0; j < 4; j++) y[j] = fma(xe, vk[c][j], y[j]);
                    if ((c & 3) == 3)  // the row loads of l stay live here: few LDS reads in flight
                        asm volatile("" : "+v"(y[0]), "+v"(y[1]), "+v"(y[2]), "+v"(y[3])::"memory");
                }
                #pragma unroll
                for (int j = 0; j < 4; j++) yk[lane][j] = y[j];
                if (lane < ib)
                    #pragma unroll
                    for (int j = 0; j < 4; j++) b[(i64)j * n + k0 + lane] = y[j];
            }
            __syncthreads();
            for (bool first = true; i < r1; i += NT, first = false) {
                if (!first) load_row(i);
                T acc[4] = {0, 0, 0, 0};
                #pragma unroll
                for (int c = 0; c < NB; c++) {
                    #pragma unroll
                    for (int j = 0; j < 4; j++) acc[j] = fma(l[c], yk[c][j], acc[j]);
                    if ((c & 7) == 7)  // bound the LDS reads in flight
                        asm volatile("" : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3])::"memory");
                }
                #pragma unroll
                for (int j = 0; j < 4; j++) b[(i64)j * n + i] -= acc[j];
            }
            __syncthreads();
        }
    }
}

template <typename T>
__global__ __launch_bounds__(SOLVE_THREADS) void k_chol_solve(const T* __restrict__ L, const T* __restrict__ Dinv,
                                                              int n, T* __restrict__ b, const int* skip) {
    if (skip && *skip) return;
    __shared__ SolveLds<T> S;
    chol_solve_body<T, SOLVE_THREADS>(L, Dinv, n, b, S);
}


// The same solve for large n, one launch per 64-row block so that many
// work-groups share the rows to update (the single work-group above streams
// the whole factor through one CU).  Forward step kb: every work-group forms
// y_kb = X_kb b_kb from LDS (work-group 0 stores it in y) and updates its 256
// rows below the block in b.  Backward step kb: x_kb = X_kb' y_kb (stored in b
// by work-group 0), rows above the block updated in y.  skip as above.
constexpr int STEP_THREADS = 256;

template <typename T, bool FWD>
__global__ __launch_bounds__(STEP_THREADS) void k_solve_step(const T* __restrict__ L, const T* __restrict__ Dinv,
                                                             int n, int kb, T* __restrict__ b, T* __restrict__ y,
                                                             const int* skip) {
    if (skip && *skip) return;
    __shared__ T Xs[NB][NB + 1];
    __shared__ T vk[NB][4];
    __shared__ T yk[NB][4];
    const int tid = threadIdx.x;
    const int k0 = kb * NB, ib = min(NB, n - k0);
    const uint32_t colb = (uint32_t)n * (uint32_t)sizeof(T);
    const __amdgpu_buffer_rsrc_t rl =
        __builtin_amdgcn_make_buffer_rsrc((void*)L, (short)0, (int)(colb * (uint32_t)n), BUF_FLAGS);
    // The block's X and right-hand side first (loads retire in order), then
    // this thread's row of L.
    const T* D = Dinv + (i64)kb * NB * NB;
    T xv[NB * NB / STEP_THREADS];
    #pragma unroll
    for (int u = 0; u < NB * NB / STEP_THREADS; u++) xv[u] = D[tid + STEP_THREADS * u];
    const T* rhs = FWD ? b : y;  // the block's current right-hand side
    const int r = tid & (NB - 1), j = tid / NB;
    const T bv = r < ib ? rhs[(i64)j * n + k0 + r] : (T)0;
    const int i = FWD ? k0 + NB + blockIdx.x * STEP_THREADS + tid : blockIdx.x * STEP_THREADS + tid;
    const bool row = FWD ? i < n : i < k0;
    T l[NB];
    if (row) {
        if (FWD) {
            const uint32_t v = (uint32_t)i * (uint32_t)sizeof(T);
            #pragma unroll
            for (int c = 0; c < NB; c++) l[c] = ld_buf<T>(rl, v, (uint32_t)(k0 + c) * colb);
        } else {
            const uint32_t v = ((uint32_t)i * (uint32_t)n + (uint32_t)k0) * (uint32_t)sizeof(T);
            #pragma unroll
            for (int c = 0; c < NB; c++) l[c] = c < ib ? ld_buf<T>(rl, v + (uint32_t)(c * sizeof(T)), 0u) : (T)0;
        }
    }
    #pragma unroll
    for (int u = 0; u < NB * NB / STEP_THREADS; u++) {
        const int e = tid + STEP_THREADS * u;
        Xs[e / NB][e & (NB - 1)] = xv[u];
    }
    vk[r][j] = bv;
    __syncthreads();
    {
        T acc = 0;
        #pragma unroll 8
        for (int c = 0; c < NB; c++) acc = fma(FWD ? Xs[r][c] : Xs[c][r], vk[c][j], acc);
        yk[r][j] = acc;
        if (blockIdx.x == 0 && r < ib) (FWD ? y : b)[(i64)j * n + k0 + r] = acc;
    }
    __syncthreads();
    if (row) {
        T acc[4] = {0, 0, 0, 0};
        #pragma unroll
        for (int c = 0; c < NB; c++) {
            #pragma unroll
            for (int q = 0; q < 4; q++) acc[q] = fma(l[c], yk[c][q], acc[q]);
            if ((c & 7) == 7)  // bound the LDS reads in flight
                asm volatile("" : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3])::"memory");
        }
        T* out = FWD ? b : y;
        #pragma unroll
        for (int q = 0; q < 4; q++) out[(i64)q * n + i] -= acc[q];
    }
}

// Pipelined block substitution: one work-group per 64-row block, all
// resident at once.  Forward: work-group j subtracts L_jk y_k for k < j as
// each y_k is published (per block a flag, stored with release order after
// y_k), forms y_j = X_j r_j (X_j the inverted diagonal block) and publishes
// it.  Backward: x_j = X_j' (y_j - sum_{k>j} L_kj' x_k), published the same
// way.  The blocks of L do not depend on the solution: each is loaded before
// the wait for the vector it multiplies.  The critical path is one block
// product and one flag hand-over per block instead of the single work-group's
// stream of the whole factor through one CU.  Flags hold the call's epoch
// (no reset between calls).  Every wait is bounded (wait_ticks of the
// 100 MHz device clock): one that runs out marks the fit state and stops
// waiting, and the host redoes the fit with one launch per block.
constexpr int PIPE_THREADS = 256, PIPE_MAX_BLOCKS = 128;

template <typename T>
__global__ __launch_bounds__(PIPE_THREADS) void k_solve_pipe(const T* __restrict__ L, const T* __restrict__ Dinv, int n,
                                                             T* __restrict__ b, T* __restrict__ y, int* flags,
                                                             int epoch, FitState* st, const int* skip, int fwd_done,
                                                             unsigned long long wait_ticks) {
    if (skip && *skip) return;
    __shared__ T vs[NB][4];       // the published block being applied (y_k or x_k)
    __shared__ T part[4][NB][4];  // partial sums of the four quarters
    __shared__ T accs[NB][4];     // the block's reduced right-hand side
    __shared__ T ys[NB][4];       // y_j, the backward pass's start
    __shared__ T Xs[NB][NB + 1];  // X_j, staged before the first wait
    __shared__ int timed_out;
    const int tid = threadIdx.x, r = tid & (NB - 1), q = tid / NB;
    const int nb = (n + NB - 1) / NB, j = blockIdx.x, j0 = j * NB, ib = min(NB, n - j0);
    int* fwd = flags;
    int* bwd = flags + nb;
    if (tid == 0) timed_out = 0;
    __syncthreads();
    auto wait_for = [&](int* f) {
        if (tid == 0 && !timed_out) {
            const unsigned long long t0 = wall_clock64();
            while (__hip_atomic_load(f, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
                if (wall_clock64() - t0 >= wait_ticks) {
                    timed_out = 1;
                    st->sync_timeout = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    };
    auto publish = [&](int* f) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (tid == 0) __hip_atomic_store(f, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    };
    // acc (thread r, quarter q) -> accs[r][c] = sum over the quarters.
    auto reduce = [&](const T* acc) {
        #pragma unroll
        for (int c = 0; c < 4; c++) part[q][r][c] = acc[c];
        __syncthreads();
        accs[r][q] = (part[0][r][q] + part[1][r][q]) + (part[2][r][q] + part[3][r][q]);
        __syncthreads();
    };
    {
        const T* X = Dinv + (i64)j * NB * NB;  // row-major; rows past ib are the identity's
        #pragma unroll
        for (int u = 0; u < NB * NB / PIPE_THREADS; u++) {
            const int e = tid + PIPE_THREADS * u;
            Xs[e / NB][e % NB] = X[e];
        }
    }
    T acc[4], l[NB / 4];
    // ---- forward: y_j = X_j (b_j - sum_{k<j} L_jk y_k).  With fwd_done the
    // factor steps have done it for every block but the last (y holds y_j,
    // b the last block's reduced right-hand side).
    if (!fwd_done || j == nb - 1) {
        #pragma unroll
        for (int c = 0; c < 4; c++) acc[c] = (q == 0 && r < ib) ? b[(i64)c * n + j0 + r] : (T)0;
        for (int k = 0; k < (fwd_done ? 0 : j); k++) {
            const int k0 = k * NB;
            #pragma unroll
            for (int m = 0; m < NB / 4; m++) l[m] = r < ib ? L[(i64)(k0 + 16 * q + m) * n + j0 + r] : (T)0;
            wait_for(&fwd[k]);
            vs[r][q] = y[(i64)q * n + k0 + r];  // block k is full (k < j)
            __syncthreads();
            #pragma unroll
            for (int m = 0; m < NB / 4; m++)
                #pragma unroll
                for (int c = 0; c < 4; c++) acc[c] = fma(-l[m], vs[16 * q + m][c], acc[c]);
            __syncthreads();
        }
        reduce(acc);
        {
            T s = 0;
            #pragma unroll 8
            for (int m = 0; m < NB; m++) s = fma(Xs[r][m], accs[m][q], s);
            ys[r][q] = s;
            if (r < ib) y[(i64)q * n + j0 + r] = s;
        }
        if (!fwd_done) publish(&fwd[j]);
    } else {
        ys[r][q] = r < ib ? y[(i64)q * n + j0 + r] : (T)0;
    }
    __syncthreads();
    // ---- backward: x_j = X_j' (y_j - sum_{k>j} L_kj' x_k)
    #pragma unroll
    for (int c = 0; c < 4; c++) acc[c] = q == 0 ? ys[r][c] : (T)0;
    for (int k = nb - 1; k > j; k--) {
        const int k0 = k * NB, ibk = min(NB, n - k0);
        #pragma unroll
        for (int m = 0; m < NB / 4; m++)
            l[m] = (r < ib && 16 * q + m < ibk) ? L[(i64)(j0 + r) * n + k0 + 16 * q + m] : (T)0;
        wait_for(&bwd[k]);
        vs[r][q] = r < ibk ? b[(i64)q * n + k0 + r] : (T)0;
        __syncthreads();
        #pragma unroll
        for (int m = 0; m < NB / 4; m++)
            #pragma unroll
            for (int c = 0; c < 4; c++) acc[c] = fma(-l[m], vs[16 * q + m][c], acc[c]);
        __syncthreads();
    }
    reduce(acc);
    {
        T s = 0;
        #pragma unroll 8
        for (int m = 0; m < NB; m++) s = fma(Xs[m][r], accs[m][q], s);
        if (r < ib) b[(i64)q * n + j0 + r] = s;
    }
    publish(&bwd[j]);
}

// Systems above this order use the per-block launches.
constexpr int SOLVE_MULTI_MIN = 1025;

// L L' x = b (n x 4 column-major, in place); y: n x 4 scratch.  Up to
// PIPE_MAX_BLOCKS blocks the pipelined kernel runs (one launch); above, one
// launch per block.
// fwd_done: the factor steps ran the forward substitution (pipelined path only).
// The pipelined solve needs all its nb work-groups resident at once (work-
// group j waits on blocks dispatched after it); pipe_fits says whether the
// device holds that many (once per context and type), and not after one of
// its waits ran out on this context.
template <typename T>
bool pipe_fits(tw_ctx* ctx, int nb) {
    if (ctx->pipe_off) return false;
    int& cap = sizeof(T) == 8 ? ctx->pipe_cap64 : ctx->pipe_cap32;
    if (cap < 0) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_solve_pipe<T>, PIPE_THREADS, 0) != hipSuccess)
            per_cu = 0;
        cap = per_cu * ctx->num_cus;
    }
    return nb <= cap;
}

template <typename T>
int chol_solve(tw_ctx* ctx, const T* L, const T* Dinv, int n, T* b, T* y, const int* skip, FitState* st,
               int fwd_done = 0) {
    const int nb = (n + NB - 1) / NB;
    // fwd_done (at most 16 blocks, from the fused factor steps) leaves only
    // the pipelined kernel's backward pass: the fallbacks run both passes.
    const bool pipe = nb <= PIPE_MAX_BLOCKS && pipe_fits<T>(ctx, nb);
    if (fwd_done && !pipe)
        return tw_fail(ctx, TW_E_DEVICE, "the device cannot hold the %d work-groups of the pipelined solve", nb);
    const unsigned long long wait_ticks = 100ull * (unsigned long long)ctx->solve_wait_us;  // 100 MHz clock
    if (pipe && ctx->arena) {
        // Allocation-free mode: the flags come from the arena and are cleared
        // by a kernel before each solve (the arena keeps earlier calls' data).
        void* p = nullptr;
        int rc;
        if ((rc = tw_ws(ctx, WS_SOLVE_FLAGS, 2 * PIPE_MAX_BLOCKS * sizeof(int), &p))) return rc;
        if ((rc = tw_fill_zero_i32(ctx, (int*)p, 2 * PIPE_MAX_BLOCKS))) return rc;
        hipLaunchKernelGGL(k_solve_pipe<T>, dim3(nb), dim3(PIPE_THREADS), 0, ctx->stream, L, Dinv, n, b, y, (int*)p,
                           1, st, skip, fwd_done, wait_ticks);
    } else if (pipe) {
        if (!ctx->solve_flags) {
            void* p = nullptr;
            TW_HIP(ctx, hipMalloc(&p, 2 * PIPE_MAX_BLOCKS * sizeof(int)));
            TW_HIP(ctx, hipMemset(p, 0, 2 * PIPE_MAX_BLOCKS * sizeof(int)));
            ctx->solve_flags = (int*)p;
        }
        if (++ctx->solve_epoch == 0) ctx->solve_epoch = 1;  // a new value per call, never the initial 0
        const int epoch = (int)ctx->solve_epoch;
        hipLaunchKernelGGL(k_solve_pipe<T>, dim3(nb), dim3(PIPE_THREADS), 0, ctx->stream, L, Dinv, n, b, y,
                           ctx->solve_flags, epoch, st, skip, fwd_done, wait_ticks);
    } else if (n < SOLVE_MULTI_MIN) {
        hipLaunchKernelGGL(k_chol_solve<T>, dim3(1), dim3(SOLVE_THREADS), 0, ctx->stream, L, Dinv, n, b, skip);
    } else {
        const int nb = (n + NB - 1) / NB;
        for (int kb = 0; kb < nb; kb++) {
            const int below = n - (kb + 1) * NB;
            const unsigned g = (unsigned)std::max(1, (below + STEP_THREADS - 1) / STEP_THREADS);
            hipLaunchKernelGGL((k_solve_step<T, true>), dim3(g), dim3(STEP_THREADS), 0, ctx->stream, L, Dinv, n, kb,
                               b, y, skip);
        }
        for (int kb = nb - 1; kb >= 0; kb--) {
            const int above = kb * NB;
            const unsigned g = (unsigned)std::max(1, (above + STEP_THREADS - 1) / STEP_THREADS);
            hipLaunchKernelGGL((k_solve_step<T, false>), dim3(g), dim3(STEP_THREADS), 0, ctx->stream, L, Dinv, n, kb,
                               b, y, skip);
        }
    }
    TW_LAUNCH_CHECK(ctx);
    return TW_OK;
}

// U (N x 4, column-major) <- Q' U (trans = 1) or Q U (trans = 0), Q = I - V T V'.
template <int NT>
__device__ __forceinline__ void q_apply_body(const double* __restrict__ V, const FitState* st, double* __restrict__ U, int N,
                             int trans, double* red, double* H) {
    const int tid = threadIdx.x;
    double G[16];
    for (int i = 0; i < 16; i++) G[i] = 0;
    for (int i = tid; i < N; i += NT) {
        double v[4], u[4];
        for (int k = 0; k < 4; k++) { v[k] = V[(size_t)k * N + i]; u[k] = U[(size_t)k * N + i]; }
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) G[a * 4 + b] += v[a] * u[b];
    }
    block_sum_n<16>(G, red);
    if (tid == 0)
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) {
                double acc = 0;
                for (int k = 0; k < 4; k++) acc += (trans ? st->T[k * 4 + a] : st->T[a * 4 + k]) * G[k * 4 + b];
                H[a * 4 + b] = acc;
            }
    __syncthreads();
    for (int i = tid; i < N; i += NT)
        for (int b = 0; b < 4; b++) {
            double acc = 0;
            for (int k = 0; k < 4; k++) acc += V[(size_t)k * N + i] * H[k * 4 + b];
            U[(size_t)b * N + i] -= acc;
        }
}

__global__ __launch_bounds__(1024) void k_q_apply(const double* __restrict__ V, const FitState* st,
                                                  double* __restrict__ U, int N, int trans) {
    __shared__ double red[17 * 16];
    __shared__ double H[16];
    q_apply_body<1024>(V, st, U, N, trans, red, H);
}


// out = (with_y ? Y : 0) - A' x, x and out N x 4 column-major.
__global__ __launch_bounds__(256) void k_matvec(const SysDesc s, const double* __restrict__ x,
                                                double* __restrict__ out, int with_y, const int* skip) {
    if (skip && *skip) return;
    __shared__ double red[4][4];
    const int i = blockIdx.x, tid = threadIdx.x, N = s.N;
    double acc[4] = {0, 0, 0, 0};
    for (int j = tid; j < N; j += 256) {
        const double a = sys_entry(s, i, j);
        #pragma unroll
        for (int k = 0; k < 4; k++) acc[k] += a * x[(size_t)k * N + j];
    }
    #pragma unroll
    for (int k = 0; k < 4; k++) acc[k] = wave_sum(acc[k]);
    if ((tid & 63) == 0)
        for (int k = 0; k < 4; k++) red[tid >> 6][k] = acc[k];
    __syncthreads();
    if (tid < 4) {
        const double ax = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
        out[(size_t)tid * N + i] = (with_y ? s.y[i * 4 + tid] : 0.0) - ax;
    }
}

// U = Y (N x 4 column-major from the row-major Y rows).
__global__ void k_load_y(const SysDesc s, double* __restrict__ U) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < s.N)
        for (int k = 0; k < 4; k++) U[(size_t)k * s.N + i] = s.y[i * 4 + k];
}

// U = [0; z] (z is n x 4 column-major, FP64).
__global__ void k_embed(const double* __restrict__ z, double* __restrict__ U, int N, const int* skip) {
    if (skip && *skip) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    const int n = N - 4;
    for (int k = 0; k < 4; k++) U[(size_t)k * N + i] = i < 4 ? 0.0 : z[(size_t)k * n + i - 4];
}

// b2 (n x 4) = rows 4.. of U; also converts to the factor precision.
template <typename T>
__global__ void k_tail(const double* __restrict__ U, double* __restrict__ b64, T* __restrict__ bt, int N) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int n = N - 4;
    if (i >= n) return;
    for (int k = 0; k < 4; k++) {
        const double v = U[(size_t)k * N + 4 + i];
        if (b64) b64[(size_t)k * n + i] = v;
        bt[(size_t)k * n + i] = (T)v;
    }
}

// Refinement residual r = b2 - B z with B = [A' - W V' - V W']_(4:,4:):
// given Au = -A' [0; z] (from k_matvec), r_i = b2_i + Au_(i+4) + W (V'u) + V (W'u).
// Writes r in FP32 for the correction solve and the max norms to st.
__global__ __launch_bounds__(1024) void k_ir_residual(const double* __restrict__ V, const double* __restrict__ W,
                                                      const double* __restrict__ U, const double* __restrict__ negAu,
                                                      const double* __restrict__ b2, float* __restrict__ r32,
                                                      FitState* st, int N) {
    if (st->ir_stop) return;
    __shared__ double red[17 * 32];
    __shared__ double G1[16], G2[16];
    const int tid = threadIdx.x, n = N - 4;
    double G[32];
    #pragma unroll
    for (int i = 0; i < 32; i++) G[i] = 0;
    for (int i = tid; i < N; i += 1024) {
        double v[4], w[4], u[4];
        #pragma unroll
        for (int k = 0; k < 4; k++) {
            v[k] = V[(size_t)k * N + i];
            w[k] = W[(size_t)k * N + i];
            u[k] = U[(size_t)k * N + i];
        }
        #pragma unroll
        for (int a = 0; a < 4; a++)
            #pragma unroll
            for (int b = 0; b < 4; b++) {
                G[a * 4 + b] += v[a] * u[b];
                G[16 + a * 4 + b] += w[a] * u[b];
            }
    }
    block_sum_n<32>(G, red);
    if (tid < 16) {  // the totals, from LDS (G indexed by tid would live in scratch)
        G1[tid] = red[16 * 32 + tid];
        G2[tid] = red[16 * 32 + 16 + tid];
    }
    __syncthreads();
    double rmax = 0, bmax = 0;
    for (int i = tid; i < n; i += 1024)
        #pragma unroll
        for (int b = 0; b < 4; b++) {
            double corr = 0;
            #pragma unroll
            for (int k = 0; k < 4; k++)
                corr += W[(size_t)k * N + 4 + i] * G1[k * 4 + b] + V[(size_t)k * N + 4 + i] * G2[k * 4 + b];
            const double bb = b2[(size_t)b * n + i];
            const double r = bb + negAu[(size_t)b * N + 4 + i] + corr;
            r32[(size_t)b * n + i] = (float)r;
            rmax = fmax(rmax, fabs(r));
            bmax = fmax(bmax, fabs(bb));
        }
    for (int o = 32; o > 0; o >>= 1) {
        rmax = fmax(rmax, __shfl_xor(rmax, o, 64));
        bmax = fmax(bmax, __shfl_xor(bmax, o, 64));
    }
    __syncthreads();
    if ((tid & 63) == 0) { red[tid >> 6] = rmax; }
    __syncthreads();
    if (tid == 0) {
        double m = 0;
        for (int w = 0; w < 16; w++) m = fmax(m, red[w]);
        st->rnorm = m;
    }
    __syncthreads();
    if ((tid & 63) == 0) { red[tid >> 6] = bmax; }
    __syncthreads();
    if (tid == 0) {
        double m = 0;
        for (int w = 0; w < 16; w++) m = fmax(m, red[w]);
        st->bnorm = m;
        // Stop on convergence to FP64 level or when a sweep no longer
        // halves the residual; 1 marks a residual usable as FP64.
        const double rel = st->rnorm / (m > 0 ? m : 1.0);
        if (rel < 1e-15 || !(rel < 0.5 * st->ir_prev)) {
            st->ir_stop = rel < 1e-11 ? 1 : 2;
        } else {
            st->ir_prev = rel;
            st->ir_iters += 1;
        }
    }
}

__device__ __forceinline__ void fit_state_init(FitState* st) {
    st->ir_stop = 0;
    st->ir_prev = INFINITY;
    st->ir_iters = 0;
    st->info = 0;
    st->sync_timeout = 0;
    st->fit_rmax = 0.0;
}

__global__ void k_ir_init(FitState* st) { fit_state_init(st); }

__global__ void k_axpy_f32(double* __restrict__ x, const float* __restrict__ d, i64 n, const int* skip) {
    if (skip && *skip) return;
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i < n) x[i] += (double)d[i];
}

template <typename Src, typename Dst>
__global__ void k_convert(const Src* __restrict__ a, Dst* __restrict__ b, i64 n) {
    const i64 i = (i64)blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = (Dst)a[i];
}

// FP64 residual of the bordered system  r = [Y; 0] - M' x  (x is m x 4),
// one row per work-group (k_unscale takes the largest |r|).
__global__ __launch_bounds__(256) void k_residual(const SysDesc s, const double* __restrict__ x,
                                                  double* __restrict__ res) {
    __shared__ double red[4][4];
    const int r = blockIdx.x, tid = threadIdx.x;
    double acc[4] = {0, 0, 0, 0};
    for (int k = tid; k < s.m; k += 256) {
        const double a = sys_entry(s, r, k);
        #pragma unroll
        for (int j = 0; j < 4; j++) acc[j] += a * x[(i64)j * s.m + k];
    }
    #pragma unroll
    for (int j = 0; j < 4; j++) acc[j] = wave_sum(acc[j]);
    if ((tid & 63) == 0)
        for (int j = 0; j < 4; j++) red[tid >> 6][j] = acc[j];
    __syncthreads();
    if (tid < 4) {
        const double b = r < s.N ? s.y[r * 4 + tid] : 0.0;
        res[(i64)tid * s.m + r] = b - (red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid]);
    }
}

// x (m x 4) = [c'; d'] with d' = R^-1 [Q'(Y - A'c')]_(0:4); U = c', U2 = Q'(Y - A'c').
__global__ void k_finish(const double* __restrict__ U, const double* __restrict__ U2, const FitState* st,
                         double* __restrict__ x, int N) {
    const int m = N + 4;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < N)
        #pragma unroll
        for (int j = 0; j < 4; j++) x[(i64)j * m + i] = U[(i64)j * N + i];
    if (blockIdx.x == 0 && threadIdx.x < 4) {
        const int j = threadIdx.x;
        double dp[4];
        for (int a = 3; a >= 0; a--) {
            double v = U2[(i64)j * N + a];
            for (int q = a + 1; q < 4; q++) v -= st->R[a * 4 + q] * dp[q];
            dp[a] = v / st->R[a * 4 + a];
        }
        for (int a = 0; a < 4; a++) x[(i64)j * m + N + a] = dp[a];
    }
}

// ---- the fit's one-work-group kernels for N <= QR_FAST_MAX -----------------
// NT threads (the QR: one wave up to N = 1024; the products, whose time is
// global-load latency, four waves); rows tid, tid + NT,
// ... of an N x 4 operand belong to thread tid, and each phase's sums go
// through one group reduction (DPP and readlane within one wave, no
// barrier).  The small 4 x 4 products are formed by every thread (no
// broadcast barrier).
constexpr int QR_FAST_MAX = 4096, QR_WAVE_MAX = 1024, QRPT = 16;

template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffff), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over one wave, the same value in every lane: quad exchanges, then the
// half-row and row mirrors leave each row of 16 with its sum (a + b == b + a,
// so every lane of a row holds the same bits), then the four rows in order.
__device__ __forceinline__ double wave_sum_dpp(double v) {
    v += dpp_d<0xB1>(v);   // quad_perm [1, 0, 3, 2]
    v += dpp_d<0x4E>(v);   // quad_perm [2, 3, 0, 1]
    v += dpp_d<0x141>(v);  // row_half_mirror
    v += dpp_d<0x140>(v);  // row_mirror
    return (bcast(v, 0) + bcast(v, 16)) + (bcast(v, 32) + bcast(v, 48));
}

template <int NT, int NV>
__device__ __forceinline__ void group_sum(double* v, double* red) {
    if constexpr (NT == 64) {
        #pragma unroll
        for (int i = 0; i < NV; i++) v[i] = wave_sum_dpp(v[i]);
    } else {
        block_sum_n<NV>(v, red);
    }
}

// State init, X' = [1 p'] (scaled) and its Householder QR (V, R, tau) and
// compact-WY factor T; X' stays in registers.  Same steps as hh_x_body,
// with the four column reductions of a step fused into two.
template <int QT>
__global__ __launch_bounds__(QT) void k_qr_fast(const SysDesc s, double* __restrict__ V, FitState* st) {
    __shared__ double red[17 * 10];
    const int N = s.N, tid = threadIdx.x;
    if (tid == 0) fit_state_init(st);
    double x[QRPT][4];
    #pragma unroll
    for (int u = 0; u < QRPT; u++) {
        const int i = tid + QT * u;
        x[u][0] = x[u][1] = x[u][2] = x[u][3] = 0.0;
        if (QT * u < N && i < N) {
            const float4 p = s.pts[i];
            x[u][0] = 1.0;
            x[u][1] = ((double)p.x - s.mu[0]) * s.inv_s;
            x[u][2] = ((double)p.y - s.mu[1]) * s.inv_s;
            x[u][3] = ((double)p.z - s.mu[2]) * s.inv_s;
        }
    }
    double R[16] = {0};
    double tau[4];
    double scale0 = 0.0;
    #pragma unroll
    for (int k = 0; k < 4; k++) {
        double nv[2] = {0.0, 0.0};  // |x_k below row k|^2, x_k[k]
        #pragma unroll
        for (int u = 0; u < QRPT; u++) {
            const int i = tid + QT * u;
            if (QT * u < N) {
                if (i > k) nv[0] += x[u][k] * x[u][k];
                else if (i == k) nv[1] = x[u][k];
            }
        }
        group_sum<QT, 2>(nv, red);
        const double xnorm2 = nv[0], alpha = nv[1];
        double beta, t, scale;
        if (xnorm2 == 0.0) {
            beta = alpha; t = 0.0; scale = 0.0;
        } else {
            beta = -copysign(sqrt(alpha * alpha + xnorm2), alpha);
            t = (beta - alpha) / beta;
            scale = 1.0 / (alpha - beta);
        }
        tau[k] = t;
        R[k * 4 + k] = beta;
        if (k == 0) scale0 = fabs(beta);
        double d[6] = {0, 0, 0, 0, 0, 0};  // v_k . x_j and x_j[k] for j = k+1..3
        #pragma unroll
        for (int u = 0; u < QRPT; u++) {
            const int i = tid + QT * u;
            x[u][k] = i < k ? 0.0 : (i == k ? 1.0 : x[u][k] * scale);
            if (QT * u < N)
                #pragma unroll
                for (int j = k + 1; j < 4; j++) {
                    if (i >= k) d[j - k - 1] += x[u][k] * x[u][j];
                    if (i == k) d[3 + j - k - 1] = x[u][j];
                }
        }
        group_sum<QT, 6>(d, red);
        #pragma unroll
        for (int j = k + 1; j < 4; j++) {
            const double w = d[j - k - 1];
            R[k * 4 + j] = d[3 + j - k - 1] - t * w;
            #pragma unroll
            for (int u = 0; u < QRPT; u++) x[u][j] -= t * x[u][k] * w;
        }
    }
    double G[10];  // V'V, upper triangle by rows
    {
        int g = 0;
        #pragma unroll
        for (int a = 0; a < 4; a++)
            #pragma unroll
            for (int b = a; b < 4; b++) {
                double p = 0;
                #pragma unroll
                for (int u = 0; u < QRPT; u++)
                    if (QT * u < N) p += x[u][a] * x[u][b];
                G[g++] = p;
            }
    }
    group_sum<QT, 10>(G, red);
    #pragma unroll
    for (int u = 0; u < QRPT; u++) {
        const int i = tid + QT * u;
        if (QT * u < N && i < N)
            #pragma unroll
            for (int k = 0; k < 4; k++) V[(size_t)k * N + i] = x[u][k];
    }
    if (tid == 0) {
        auto gv = [&](int a, int b) { return a <= b ? G[a * 4 - a * (a - 1) / 2 + (b - a)] : G[b * 4 - b * (b - 1) / 2 + (a - b)]; };
        double T[16] = {0};
        for (int k = 0; k < 4; k++) {
            T[k * 4 + k] = tau[k];
            for (int i = 0; i < k; i++) {
                double acc = 0;
                for (int j = i; j < k; j++) acc += T[i * 4 + j] * gv(j, k);
                T[i * 4 + k] = -tau[k] * acc;
            }
        }
        for (int i = 0; i < 16; i++) { st->R[i] = R[i]; st->T[i] = T[i]; }
        for (int k = 0; k < 4; k++) st->tau[k] = tau[k];
        int bad = 0;
        for (int k = 1; k < 4; k++)
            if (!(fabs(R[k * 4 + k]) > 1e-9 * scale0)) bad = 1;
        st->info = bad;
    }
}

// H = T' G (trans) or T G, every thread.
__device__ __forceinline__ void small_tg(const double* __restrict__ T, const double* G, int trans, double* H) {
    #pragma unroll
    for (int a = 0; a < 4; a++)
        #pragma unroll
        for (int b = 0; b < 4; b++) {
            double acc = 0;
            #pragma unroll
            for (int k = 0; k < 4; k++) acc += (trans ? T[k * 4 + a] : T[a * 4 + k]) * G[k * 4 + b];
            H[a * 4 + b] = acc;
        }
}

// W = P T - 1/2 V M (M = T' (V'P) T) and U = Q'Y = Y - V T' (V'Y), one pass
// of sums for both; z (n x 4) gets rows 4.. of U (the FP64 right-hand side).
template <int QT>
__global__ __launch_bounds__(QT) void k_w_qy_fast(const SysDesc s, const double* __restrict__ V,
                                                  const double* __restrict__ P, double* __restrict__ W,
                                                  double* __restrict__ U, const FitState* __restrict__ st,
                                                  double* __restrict__ z) {
    __shared__ double red[17 * 32];
    const int N = s.N, tid = threadIdx.x;
    double S[32];
    #pragma unroll
    for (int i = 0; i < 32; i++) S[i] = 0;
    for (int i = tid; i < N; i += QT) {
        double v[4], p[4], y[4];
        #pragma unroll
        for (int k = 0; k < 4; k++) { v[k] = V[(size_t)k * N + i]; p[k] = P[(size_t)k * N + i]; y[k] = s.y[i * 4 + k]; }
        #pragma unroll
        for (int a = 0; a < 4; a++)
            #pragma unroll
            for (int b = 0; b < 4; b++) {
                S[a * 4 + b] += v[a] * p[b];
                S[16 + a * 4 + b] += v[a] * y[b];
            }
    }
    group_sum<QT, 32>(S, red);
    double Ts[16], TS[16], M[16], H[16];
    #pragma unroll
    for (int i = 0; i < 16; i++) Ts[i] = st->T[i];
    small_tg(Ts, S, 1, TS);  // T' S
    #pragma unroll
    for (int a = 0; a < 4; a++)  // (T' S) T
        #pragma unroll
        for (int b = 0; b < 4; b++) {
            double acc = 0;
            #pragma unroll
            for (int k = 0; k < 4; k++) acc += TS[a * 4 + k] * Ts[k * 4 + b];
            M[a * 4 + b] = acc;
        }
    small_tg(Ts, S + 16, 1, H);
    for (int i = tid; i < N; i += QT) {
        double v[4], p[4], y[4];
        #pragma unroll
        for (int k = 0; k < 4; k++) { v[k] = V[(size_t)k * N + i]; p[k] = P[(size_t)k * N + i]; y[k] = s.y[i * 4 + k]; }
        #pragma unroll
        for (int b = 0; b < 4; b++) {
            double acc = 0, q = 0;
            #pragma unroll
            for (int k = 0; k < 4; k++) {
                acc += p[k] * Ts[k * 4 + b] - 0.5 * v[k] * M[k * 4 + b];
                q += v[k] * H[k * 4 + b];
            }
            W[(size_t)b * N + i] = acc;
            U[(size_t)b * N + i] = y[b] - q;
            if (i >= 4) z[(size_t)b * (N - 4) + i - 4] = y[b] - q;
        }
    }
}

// U <- Q'U (trans) or QU for an N x 4 U.  With z, U starts as [0; z] (z:
// n x 4, n = N - 4).  With x, the result's first four rows give
// d' = R^-1 [Q'U]_(0:4) and x (m x 4, m = N + 4) = [Uc; d'].
template <int QT>
__global__ __launch_bounds__(QT) void k_q_fast(const double* __restrict__ V, const FitState* __restrict__ st,
                                               double* __restrict__ U, int N, int trans, const double* __restrict__ z,
                                               const double* __restrict__ Uc, double* __restrict__ x) {
    __shared__ double red[17 * 16];
    __shared__ double top[4][4];
    const int tid = threadIdx.x, n = N - 4, m = N + 4;
    auto row = [&](int i, double* u) {
        #pragma unroll
        for (int k = 0; k < 4; k++) u[k] = z ? (i < 4 ? 0.0 : z[(size_t)k * n + i - 4]) : U[(size_t)k * N + i];
    };
    double G[16];
    #pragma unroll
    for (int i = 0; i < 16; i++) G[i] = 0;
    for (int i = tid; i < N; i += QT) {
        double v[4], u[4];
        row(i, u);
        #pragma unroll
        for (int k = 0; k < 4; k++) v[k] = V[(size_t)k * N + i];
        #pragma unroll
        for (int a = 0; a < 4; a++)
            #pragma unroll
            for (int b = 0; b < 4; b++) G[a * 4 + b] += v[a] * u[b];
    }
    group_sum<QT, 16>(G, red);
    double Ts[16], H[16];
    #pragma unroll
    for (int i = 0; i < 16; i++) Ts[i] = st->T[i];
    small_tg(Ts, G, trans, H);
    for (int i = tid; i < N; i += QT) {
        double v[4], u[4];
        row(i, u);
        #pragma unroll
        for (int k = 0; k < 4; k++) v[k] = V[(size_t)k * N + i];
        #pragma unroll
        for (int b = 0; b < 4; b++) {
            double acc = 0;
            #pragma unroll
            for (int k = 0; k < 4; k++) acc += v[k] * H[k * 4 + b];
            const double r = u[b] - acc;
            U[(size_t)b * N + i] = r;
            if (x) {
                if (i < 4) top[i][b] = r;
                x[(i64)b * m + i] = Uc[(size_t)b * N + i];
            }
        }
    }
    if (x) {
        __syncthreads();
        if (tid < 4) {
            const int j = tid;
            double dp[4];
            for (int a = 3; a >= 0; a--) {
                double v = top[a][j];
                for (int q = a + 1; q < 4; q++) v -= st->R[a * 4 + q] * dp[q];
                dp[a] = v / st->R[a * 4 + a];
            }
            for (int a = 0; a < 4; a++) x[(i64)j * m + N + a] = dp[a];
        }
    }
}

inline unsigned nblk(i64 n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// Blocked Cholesky with the fused steps (small systems): the off-diagonal
// blocks of the factor go to Lp.
// With rhs (n x 4), the steps also run the forward substitution: yout gets
// y = L^-1 rhs for every block but the last, and rhs the last block's
// reduced right-hand side (the solve finishes with fwd_done).
template <typename T>
void cholesky_steps(tw_ctx* ctx, T* B, T* Lp, T* Dinv, int n, FitState* st, T* rhs, T* yout) {
    hipLaunchKernelGGL(k_potrf_diag<T>, dim3(1), dim3(256), 0, ctx->stream, B, Dinv, n, 0, std::min(NB, n), st);
    for (int j0 = 0; j0 + NB < n; j0 += NB) {
        const unsigned nt = nblk(n - j0 - NB, NB);
        hipLaunchKernelGGL(k_chol_step<T>, dim3(nt, nt), dim3(256), 0, ctx->stream, B, Lp, Dinv, n, j0, NB, st, rhs,
                           yout);
    }
}

// FP64: each trailing update also factors the next diagonal block
// (k_syrk_potrf), and with rhs (n x 4) each panel runs the forward
// substitution (k_trsm_fwd: yout gets y_k for every block but the last, rhs
// the last block's reduced right-hand side); FP32: one launch per phase.
template <typename T>
void cholesky(tw_ctx* ctx, T* B, T* Dinv, int n, FitState* st, T* rhs = nullptr, T* yout = nullptr) {
    constexpr bool FUSE = std::is_same<T, double>::value;
    for (int j0 = 0; j0 < n; j0 += NB) {
        const int jb = std::min(NB, n - j0);
        if (!FUSE || j0 == 0)
            hipLaunchKernelGGL(k_potrf_diag<T>, dim3(1), dim3(256), 0, ctx->stream, B, Dinv, n, j0, jb, st);
        const int rest = n - j0 - jb;
        if (rest > 0) {
            if (FUSE && rhs)
                hipLaunchKernelGGL(k_trsm_fwd, dim3(nblk(rest, NB)), dim3(256), 0, ctx->stream, (double*)B,
                                   (const double*)Dinv, n, j0, jb, (double*)rhs, (double*)yout);
            else
                hipLaunchKernelGGL(k_trsm_chol<T>, dim3(nblk(rest, NB)), dim3(256), 0, ctx->stream, B, Dinv, n, j0,
                                   jb);
            const dim3 tiles(nblk(rest, NB), nblk(rest, NB));
            if constexpr (FUSE)
                hipLaunchKernelGGL(k_syrk_potrf, tiles, dim3(256), 0, ctx->stream, (double*)B, (double*)Dinv, n, j0,
                                   jb, st);
            else
                hipLaunchKernelGGL(k_syrk<T>, tiles, dim3(256), 0, ctx->stream, B, n, j0, jb);
        }
    }
}


// Grid-kernel coefficients from the scaled solution x (m x 4, column-major):
// cf[3i + q] = c'[i][q+1] / s and d = T^-1 d' (d[0][j] = d'0j - (mu . d'[1..3][j]) / s,
// d[a][j] = d'[a][j] / s), as the host unscaling of tw_tps_fit.
__device__ __forceinline__ void unscale_coef(const double* __restrict__ x, int N, double s, double* __restrict__ cf,
                                             int i) {
    const int m = N + 4;
    #pragma unroll
    for (int q = 0; q < 3; q++) cf[3 * i + q] = x[(i64)(q + 1) * m + i] / s;
}
__device__ __forceinline__ void unscale_affine(const double* __restrict__ x, int N, double s, double mu0, double mu1,
                                               double mu2, double* __restrict__ d, int j) {
    const int m = N + 4;
    const double d0 = x[(i64)j * m + N], d1 = x[(i64)j * m + N + 1], d2 = x[(i64)j * m + N + 2],
                 d3 = x[(i64)j * m + N + 3];
    d[0 * 4 + j] = d0 - (mu0 * d1 + mu1 * d2 + mu2 * d3) / s;
    d[1 * 4 + j] = d1 / s;
    d[2 * 4 + j] = d2 / s;
    d[3 * 4 + j] = d3 / s;
}

// st->fit_rmax = max |res| over cnt entries (256 threads).
__device__ __forceinline__ void res_max_body(const double* __restrict__ res, int cnt, FitState* st) {
    __shared__ double red[4];
    double m = 0;
    for (int k = threadIdx.x; k < cnt; k += 256) m = fmax(m, fabs(res[k]));
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) st->fit_rmax = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

__global__ __launch_bounds__(256) void k_res_max(const double* __restrict__ res, int cnt, FitState* st) {
    res_max_body(res, cnt, st);
}

// With res, work-group 0 also sets st->fit_rmax = max |res| over cnt entries.
__global__ __launch_bounds__(256) void k_unscale(const double* __restrict__ x, int N, double s, double mu0,
                                                 double mu1, double mu2, double* __restrict__ cf,
                                                 double* __restrict__ d, const double* __restrict__ res, int cnt,
                                                 FitState* st) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < N) unscale_coef(x, N, s, cf, i);
    if (blockIdx.x == 0 && threadIdx.x < 4) unscale_affine(x, N, s, mu0, mu1, mu2, d, threadIdx.x);
    if (blockIdx.x == 0 && res) res_max_body(res, cnt, st);
}

// ---- small systems: the whole fit in one work-group ------------------------
// The FP64 steps of fit_run (QR of X', W, B, blocked Cholesky with inverted
// diagonal blocks, substitution, c' and d', bordered residual) run in one
// work-group of 256 threads with work-group barriers between them, so a
// single small system costs one launch instead of a kernel sequence and a
// batch of systems takes one launch with a work-group per system.  The
// landmarks and every N x 4 vector stay in LDS; B and the inverted diagonal
// blocks live in the system's workspace.  The tile kernels' bodies are
// shared with the large-system path; the matrix-free products go one row per
// thread with the column index uniform across the wave (LDS broadcasts).
constexpr int FIT_SMALL_THREADS = 256;
constexpr int FIT_SMALL_MAX = 256;  // largest K fitted this way (one row per thread)

struct SmallFitJob {
    SysDesc s;
    double* ws;       // small_ws_doubles(N) doubles
    FitState* st;
    double* cf;       // grid-kernel coefficients (K x 3) and affine block (4 x 4), or null
    double* d;
    double scale;     // s (the solution is in scaled coordinates)
};

// Workspace of one small system: x (m x 4), B (n x n, lower triangle), the
// inverted diagonal blocks.
inline size_t small_ws_doubles(int N) {
    const size_t n = (size_t)N - 4, m = (size_t)N + 4, nb = (n + NB - 1) / NB;
    return 4 * m + n * n + nb * NB * NB;
}
// Offset of x in that workspace.
inline size_t small_ws_x(int) { return 0; }

union SmallFitLds {
    PotrfLds<double> potrf;
    TrsmLds<double> trsm;
    SyrkLds<double> syrk;
    SolveLds<double> solve;
    struct {
        double red[17 * 16];
        double M[16], Ts[16], H[16], sh[8];
    } v;
};

// The system's vectors (column-major N x 4 with stride N; x: m x 4, stride m).
struct SmallVecs {
    float4 pts[FIT_SMALL_MAX];
    double y[4 * FIT_SMALL_MAX];  // rows [1 target]
    double V[4 * FIT_SMALL_MAX], P[4 * FIT_SMALL_MAX], W[4 * FIT_SMALL_MAX], U[4 * FIT_SMALL_MAX],
        U2[4 * FIT_SMALL_MAX];
    double x[4 * (FIT_SMALL_MAX + 4)];
    double z[4 * FIT_SMALL_MAX];
};

// acc = row i < N of A' x (x: N x 4 column-major, stride N); the entries of
// A' as sys_entry forms them, four columns per step so that their LDS reads
// are in flight together.
__device__ __forceinline__ void a_row(const SysDesc& s, const double* __restrict__ x, int N, int i, double* acc) {
    const float4 pi = s.pts[i];
    acc[0] = acc[1] = acc[2] = acc[3] = 0.0;
    #pragma unroll 4
    for (int k = 0; k < N; k++) {
        const float4 pk = s.pts[k];
        const double dx = (float)(pi.x - pk.x), dy = (float)(pi.y - pk.y), dz = (float)(pi.z - pk.z);
        double a = -sqrt(dx * dx + dy * dy + dz * dz) * s.inv_s;
        if (k == i) a += s.lam;
        #pragma unroll
        for (int j = 0; j < 4; j++) acc[j] += a * x[j * N + k];
    }
}

__device__ __forceinline__ void fit_small_body(const SmallFitJob& J, SmallFitLds& L, SmallVecs& S) {
    constexpr int NT = FIT_SMALL_THREADS;
    static_assert(FIT_SMALL_MAX <= NT, "one row per thread");
    const int N = J.s.N, n = N - 4, m = N + 4, tid = threadIdx.x;
    double* xg = J.ws;
    double* B = xg + (size_t)m * 4;
    double* Dinv = B + (size_t)n * n;
    FitState* st = J.st;
    if (tid < N) {
        S.pts[tid] = J.s.pts[tid];
        #pragma unroll
        for (int k = 0; k < 4; k++) S.y[4 * tid + k] = J.s.y[4 * tid + k];
    }
    SysDesc s = J.s;
    s.pts = S.pts;
    s.y = S.y;
    double *V = S.V, *P = S.P, *W = S.W, *U = S.U, *U2 = S.U2, *x = S.x, *z = S.z;
    if (tid == 0) {
        st->ir_stop = 0;
        st->ir_prev = INFINITY;
        st->ir_iters = 0;
        st->info = 0;
        st->sync_timeout = 0;
    }
    __syncthreads();
    // 1. QR of X', P = A' V, W.
    hh_x_body<NT>(s, V, st, L.v.red, L.v.sh);
    __syncthreads();
    if (tid < N) {
        double acc[4];
        a_row(s, V, N, tid, acc);
        #pragma unroll
        for (int k = 0; k < 4; k++) P[k * N + tid] = acc[k];
    }
    __syncthreads();
    build_w_body<NT>(V, P, W, st, N, L.v.red, L.v.M, L.v.Ts);
    // 2. U = Q'Y; B (lower triangle, row i - 4 in thread i - 4); z = U[4:].
    if (tid < N)
        #pragma unroll
        for (int k = 0; k < 4; k++) U[k * N + tid] = s.y[tid * 4 + k];
    __syncthreads();
    q_apply_body<NT>(V, st, U, N, 1, L.v.red, L.v.H);
    __syncthreads();
    if (tid < n) {
        const int i = tid + 4;
        double wi[4], vi[4];
        #pragma unroll
        for (int k = 0; k < 4; k++) {
            wi[k] = W[k * N + i];
            vi[k] = V[k * N + i];
        }
        const float4 pi = s.pts[i];
        #pragma unroll 4
        for (int j = 4; j <= i; j++) {
            const float4 pj = s.pts[j];
            const double dx = (float)(pi.x - pj.x), dy = (float)(pi.y - pj.y), dz = (float)(pi.z - pj.z);
            double v = -sqrt(dx * dx + dy * dy + dz * dz) * s.inv_s;
            if (i == j) v += s.lam;
            #pragma unroll
            for (int k = 0; k < 4; k++) v -= wi[k] * V[k * N + j] + vi[k] * W[k * N + j];
            B[(size_t)(j - 4) * n + tid] = v;
        }
        #pragma unroll
        for (int k = 0; k < 4; k++) z[k * n + tid] = U[k * N + i];
    }
    __syncthreads();
    // 3. Blocked Cholesky of B, tile by tile.
    for (int j0 = 0; j0 < n; j0 += NB) {
        const int jb = min(NB, n - j0);
        potrf_diag_body<double>(B, Dinv, n, j0, jb, st, L.potrf);
        __syncthreads();
        const int nt = (n - j0 - jb + NB - 1) / NB;
        for (int bx = 0; bx < nt; bx++) {
            double acc[NB / 4];
            trsm_body<double>(B, Dinv, n, j0, jb, bx, L.trsm, acc);
            __syncthreads();
        }
        for (int bx = 0; bx < nt; bx++)
            for (int by = 0; by <= bx; by++) {
                syrk_body<double>(B, n, j0, jb, bx, by, L.syrk);
                __syncthreads();
            }
    }
    chol_solve_body<double, NT>(B, Dinv, n, z, L.solve);
    __syncthreads();
    // 4. c' = Q [0; z]; U2 = Q'(Y - A'c'); d' = R^-1 U2[0:4].
    if (tid < N)
        #pragma unroll
        for (int k = 0; k < 4; k++) U[k * N + tid] = tid < 4 ? 0.0 : z[k * n + tid - 4];
    __syncthreads();
    q_apply_body<NT>(V, st, U, N, 0, L.v.red, L.v.H);
    __syncthreads();
    if (tid < N) {
        double acc[4];
        a_row(s, U, N, tid, acc);
        #pragma unroll
        for (int k = 0; k < 4; k++) U2[k * N + tid] = s.y[tid * 4 + k] - acc[k];
    }
    __syncthreads();
    q_apply_body<NT>(V, st, U2, N, 1, L.v.red, L.v.H);
    __syncthreads();
    if (tid < N)
        #pragma unroll
        for (int j = 0; j < 4; j++) x[j * m + tid] = U[j * N + tid];
    if (tid < 4) {
        const int j = tid;
        double dp[4];
        #pragma unroll
        for (int a = 3; a >= 0; a--) {
            double v = U2[j * N + a];
            #pragma unroll
            for (int q = a + 1; q < 4; q++) v -= st->R[a * 4 + q] * dp[q];
            dp[a] = v / st->R[a * 4 + a];
        }
        #pragma unroll
        for (int a = 0; a < 4; a++) x[j * m + N + a] = dp[a];
    }
    __syncthreads();
    // 5. Bordered residual max: rows < N one per thread, the four border rows
    //    (X'^T c') by a work-group sum; x and the unscaled coefficients out.
    double rm = 0.0, bs[16];
    #pragma unroll
    for (int e = 0; e < 16; e++) bs[e] = 0.0;
    if (tid < N) {
        // Row tid of the bordered product: A' c' (stride m) then X' d'.
        const float4 pi = s.pts[tid];
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        #pragma unroll 4
        for (int k = 0; k < N; k++) {
            const float4 pk = s.pts[k];
            const double dx = (float)(pi.x - pk.x), dy = (float)(pi.y - pk.y), dz = (float)(pi.z - pk.z);
            double a = -sqrt(dx * dx + dy * dy + dz * dz) * s.inv_s;
            if (k == tid) a += s.lam;
            #pragma unroll
            for (int j = 0; j < 4; j++) acc[j] += a * x[j * m + k];
        }
        #pragma unroll
        for (int a = 0; a < 4; a++) {
            const double xa = sys_entry(s, tid, N + a);
            #pragma unroll
            for (int j = 0; j < 4; j++) acc[j] += xa * x[j * m + N + a];
        }
        #pragma unroll
        for (int j = 0; j < 4; j++) rm = fmax(rm, fabs(s.y[tid * 4 + j] - acc[j]));
        #pragma unroll
        for (int a = 0; a < 4; a++) {
            const double xa = sys_entry(s, N + a, tid);
            #pragma unroll
            for (int j = 0; j < 4; j++) bs[a * 4 + j] = xa * x[j * m + tid];
        }
    }
    block_sum_n<16>(bs, L.v.red);
    #pragma unroll
    for (int e = 0; e < 16; e++) rm = fmax(rm, fabs(bs[e]));
    for (int o = 32; o > 0; o >>= 1) rm = fmax(rm, __shfl_xor(rm, o, 64));
    __syncthreads();
    if ((tid & 63) == 0) L.v.red[tid >> 6] = rm;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int w = 0; w < NT / 64; w++) t = fmax(t, L.v.red[w]);
        st->fit_rmax = t;
    }
    for (int i = tid; i < 4 * m; i += NT) xg[i] = x[i];
    if (J.cf) {
        if (tid < N) unscale_coef(x, N, J.scale, J.cf, tid);
        if (tid < 4) unscale_affine(x, N, J.scale, s.mu[0], s.mu[1], s.mu[2], J.d, tid);
    }
}

__global__ __launch_bounds__(FIT_SMALL_THREADS) void k_fit_small(const SmallFitJob job) {
    __shared__ SmallFitLds L;
    __shared__ SmallVecs S;
    fit_small_body(job, L, S);
}

__global__ __launch_bounds__(FIT_SMALL_THREADS) void k_fit_small_batch(const SmallFitJob* __restrict__ jobs) {
    __shared__ SmallFitLds L;
    __shared__ SmallVecs S;
    const SmallFitJob job = jobs[blockIdx.x];
    fit_small_body(job, L, S);
}

// Host-side scaling of a system: centroid mu and half-extent s of the
// centres; bmax = the largest |entry| of [1 targets] (the residual's scale).
struct HostSys {
    double mu[3], s, bmax;
};
bool host_sys(const float* centres, const float* targets, int N, HostSys& h) {
    for (int a = 0; a < 3; a++) h.mu[a] = 0;
    for (int i = 0; i < N; i++)
        for (int a = 0; a < 3; a++) h.mu[a] += centres[3 * i + a];
    for (int a = 0; a < 3; a++) h.mu[a] /= N;
    h.s = 0;
    for (int i = 0; i < N; i++)
        for (int a = 0; a < 3; a++) h.s = std::max(h.s, std::fabs(centres[3 * i + a] - h.mu[a]));
    h.bmax = 1.0;
    for (int i = 0; i < 3 * N; i++) h.bmax = std::max(h.bmax, std::fabs((double)targets[i]));
    return h.s > 0;
}

// Landmark pack of a system: N centres as float4, then N rows [1 target] (FP64).
inline size_t pack_pts_bytes(int N) { return ((size_t)N * sizeof(float4) + 255) & ~(size_t)255; }
inline size_t pack_y_bytes(int N) { return ((size_t)N * 4 * sizeof(double) + 255) & ~(size_t)255; }
void pack_host(const float* centres, const float* targets, int N, uint8_t* out) {
    float4* hp = (float4*)out;
    double* hy = (double*)(out + pack_pts_bytes(N));
    for (int i = 0; i < N; i++) {
        hp[i] = make_float4(centres[3 * i], centres[3 * i + 1], centres[3 * i + 2], 0.f);
        hy[4 * i] = 1.0;
        for (int a = 0; a < 3; a++) hy[4 * i + 1 + a] = targets[3 * i + a];
    }
}
SysDesc sys_desc(const HostSys& h, const void* pack, int N, double lambda) {
    SysDesc sd;
    sd.pts = (const float4*)pack;
    sd.y = (const double*)((const uint8_t*)pack + pack_pts_bytes(N));
    sd.N = N;
    sd.m = N + 4;
    sd.inv_s = 1.0 / h.s;
    for (int a = 0; a < 3; a++) sd.mu[a] = h.mu[a];
    sd.lam = lambda / h.s;
    return sd;
}

// Device results of one fit run (scaled coordinates: c = c'/s, d = T^-1 d').
struct FitRun {
    double* x;        // m x 4 column-major: [c'; d'], device
    FitState* st;     // device
    const double* res = nullptr;  // bordered residual (m x 4) whose max k_unscale takes; none: set already
    const float4* pts;  // centres, device
    double s, mu[3];
    double bmax;      // largest |entry| of [1 targets]
    int N, m, iters;
};

int check_fit_args(tw_ctx* ctx, const float* centres, const float* targets, int K, double lambda, int precision) {
    if (!centres || !targets) return tw_fail(ctx, TW_E_ARG, "null argument");
    if (K < 5) return tw_fail(ctx, TW_E_ARG, "the STPS fit needs at least 5 landmarks (got %d)", K);
    if (K > 16384)  // the factor's byte offsets stay below 2^31
        return tw_fail(ctx, TW_E_UNSUPPORTED, "the STPS fit takes at most 16384 landmarks (got %d)", K);
    if (!(lambda >= 0.0) || !std::isfinite(lambda)) return tw_fail(ctx, TW_E_ARG, "lambda must be finite and >= 0");
    if (precision != TW_EXACT && precision != TW_PLAIN && precision != TW_MIXED)
        return tw_fail(ctx, TW_E_ARG, "unknown precision %d", precision);
    for (int i = 0; i < 3 * K; i++)
        if (!std::isfinite(centres[i]) || !std::isfinite(targets[i]))
            return tw_fail(ctx, TW_E_ARG, "landmark coordinate %d is not finite", i);
    return TW_OK;
}

// Every kernel of the fit on ctx's stream, through the bordered-system
// residual; synchronises only for the refinement control of a large
// TW_MIXED system.
int fit_run(tw_ctx* ctx, const float* centres, const float* targets, int K, double lambda, int precision,
            FitRun& out) {
    int rc;
    const int N = K, n = N - 4, m = N + 4;
    HostSys hs;
    if (!host_sys(centres, targets, K, hs)) return tw_fail(ctx, TW_E_SINGULAR, "all centres coincide");
    const double s = hs.s;
    const double* mu = hs.mu;

    // Landmark packs.
    const size_t pb = pack_pts_bytes(N), yb = pack_y_bytes(N);
    void* pack = nullptr;
    if ((rc = tw_ws(ctx, WS_PTS, pb + yb, &pack))) return rc;
    void* host = nullptr;
    if ((rc = tw_pinned(ctx, pb + yb, &host))) return rc;
    pack_host(centres, targets, N, (uint8_t*)host);
    TW_HIP(ctx, hipMemcpyAsync(pack, host, pb + yb, hipMemcpyHostToDevice, ctx->stream));
    if ((rc = tw_pinned_issued(ctx))) return rc;
    const SysDesc sd = sys_desc(hs, pack, N, lambda);
    out.pts = sd.pts;
    out.s = s;
    for (int a = 0; a < 3; a++) out.mu[a] = mu[a];
    out.bmax = hs.bmax;
    out.N = N;
    out.m = m;
    out.iters = 0;

    if (K <= FIT_SMALL_MAX) {
        // One work-group, FP64 (every precision form).
        void *arena = nullptr, *aux = nullptr;
        if ((rc = tw_ws(ctx, WS_SOLVE_B, small_ws_doubles(N) * sizeof(double), &arena))) return rc;
        if ((rc = tw_ws(ctx, WS_SOLVE_AUX, sizeof(FitState), &aux))) return rc;
        SmallFitJob job;
        job.s = sd;
        job.ws = (double*)arena;
        job.st = (FitState*)aux;
        job.cf = job.d = nullptr;
        job.scale = s;
        hipLaunchKernelGGL(k_fit_small, dim3(1), dim3(FIT_SMALL_THREADS), 0, ctx->stream, job);
        TW_LAUNCH_CHECK(ctx);
        out.x = (double*)arena + small_ws_x(N);
        out.st = job.st;
        return TW_OK;
    }

    // Workspaces: B (n x n, FP64 or FP32); vectors in one FP64 arena;
    // state, FP32 vectors and the inverted diagonal blocks in aux.
    // The block steps and triangular solves are latency-bound at these
    // orders, so below ctx->fp32_factor_min the FP64 factor (one solve) beats
    // the FP32 factor plus refinement sweeps (several solves and residuals);
    // TW_MIXED then takes it too.
    const bool mixed = precision == TW_MIXED && n >= ctx->fp32_factor_min;
    void *Bw = nullptr, *arena = nullptr, *aux = nullptr;
    if ((rc = tw_ws(ctx, WS_SOLVE_A, (size_t)n * n * (mixed ? sizeof(float) : sizeof(double)), &Bw))) return rc;
    const size_t NV = (size_t)N * 4;
    // V, P, W, U, U2, x(m x 4), res(m x 4), b2(n x 4), z(n x 4)
    const size_t arena_d = 5 * NV + 2 * (size_t)m * 4 + 2 * (size_t)n * 4;
    if ((rc = tw_ws(ctx, WS_SOLVE_B, arena_d * sizeof(double), &arena))) return rc;
    const int nblocks = (n + NB - 1) / NB;
    const size_t st_b = (sizeof(FitState) + 255) & ~(size_t)255;
    const size_t f32_b = ((size_t)n * 4 * sizeof(float) + 255) & ~(size_t)255;
    const size_t dinv_b = (size_t)nblocks * NB * NB * sizeof(double);
    const size_t ys_b = ((size_t)n * 4 * sizeof(double) + 255) & ~(size_t)255;  // solve scratch (large n)
    if ((rc = tw_ws(ctx, WS_SOLVE_AUX, st_b + 2 * f32_b + ys_b + dinv_b, &aux))) return rc;
    double* V = (double*)arena;
    double* P = V + NV;
    double* W = P + NV;
    double* U = W + NV;
    double* U2 = U + NV;
    double* x = U2 + NV;
    double* res = x + (size_t)m * 4;
    double* b2 = res + (size_t)m * 4;
    double* z = b2 + (size_t)n * 4;
    FitState* st = (FitState*)aux;
    float* f32a = (float*)((uint8_t*)aux + st_b);
    float* f32b = (float*)((uint8_t*)aux + st_b + f32_b);
    void* ys = (uint8_t*)aux + st_b + 2 * f32_b;
    void* dinv = (uint8_t*)aux + st_b + 2 * f32_b + ys_b;
    const int* stop = &st->ir_stop;

    FitState hst;
    auto fetch_state = [&]() -> int {
        if ((rc = tw_read_small(ctx, &hst, st, sizeof hst))) return rc;
        return TW_OK;
    };

    // 1. QR of X', W, B.
    // 2. Q'Y -> U (b2 = rows 4.. of U)
    const bool qfast = N <= QR_FAST_MAX, qwave = N <= QR_WAVE_MAX;
    if (qfast) {
        if (qwave) hipLaunchKernelGGL(k_qr_fast<64>, dim3(1), dim3(64), 0, ctx->stream, sd, V, st);
        else hipLaunchKernelGGL(k_qr_fast<256>, dim3(1), dim3(256), 0, ctx->stream, sd, V, st);
        hipLaunchKernelGGL(k_av, dim3(N), dim3(256), 0, ctx->stream, sd, V, P);
        hipLaunchKernelGGL(k_w_qy_fast<256>, dim3(1), dim3(256), 0, ctx->stream, sd, V, P, W, U,
                           (const FitState*)st, z);
    } else {
        hipLaunchKernelGGL(k_ir_init, dim3(1), dim3(1), 0, ctx->stream, st);
        hipLaunchKernelGGL(k_hh_x, dim3(1), dim3(1024), 0, ctx->stream, sd, V, st);
        hipLaunchKernelGGL(k_av, dim3(N), dim3(256), 0, ctx->stream, sd, V, P);
        hipLaunchKernelGGL(k_build_w, dim3(1), dim3(1024), 0, ctx->stream, V, P, W, st, N);
        hipLaunchKernelGGL(k_load_y, dim3(nblk(N)), dim3(256), 0, ctx->stream, sd, U);
        hipLaunchKernelGGL(k_q_apply, dim3(1), dim3(1024), 0, ctx->stream, V, st, U, N, 1);
    }
    TW_LAUNCH_CHECK(ctx);

    auto factor_solve_fp64 = [&]() -> int {
        double* B = (double*)Bw;
        if (mixed) {
            int r2 = tw_ws(ctx, WS_SOLVE_A, (size_t)n * n * sizeof(double), &Bw);
            if (r2) return r2;
            B = (double*)Bw;
        }
        hipLaunchKernelGGL(k_build_b<double>, dim3(nblk((i64)n * n)), dim3(256), 0, ctx->stream, sd, V, W, B);
        const double* Lf = B;  // where the solve reads the factor's off-diagonal blocks
        // The right-hand side first (k_w_qy_fast wrote it unless the FP32
        // refinement has used z since): the fused steps run the forward
        // substitution.
        if (!qfast || mixed)
            hipLaunchKernelGGL(k_tail<double>, dim3(nblk(n)), dim3(256), 0, ctx->stream, U, (double*)nullptr, z, N);
        const int nbk = (n + NB - 1) / NB;
        const bool fused = nbk <= STEP_MAX_BLOCKS;
        // The factor's forward substitution (fused steps, or the panels of
        // the blocked factor) leaves the backward pass to the pipelined
        // solve; without the pipelined solve the solve runs both passes.
        const bool fwd = nbk <= PIPE_MAX_BLOCKS && pipe_fits<double>(ctx, nbk);
        if (fused) {
            void* lp = nullptr;
            int r2 = tw_ws(ctx, WS_SOLVE_L, (size_t)n * n * sizeof(double), &lp);
            if (r2) return r2;
            cholesky_steps<double>(ctx, B, (double*)lp, (double*)dinv, n, st, fwd ? z : nullptr,
                                   fwd ? (double*)ys : nullptr);
            Lf = (const double*)lp;
        } else {
            cholesky<double>(ctx, B, (double*)dinv, n, st, fwd ? z : nullptr, fwd ? (double*)ys : nullptr);
        }
        return chol_solve<double>(ctx, Lf, (const double*)dinv, n, z, (double*)ys, nullptr, st, fwd ? 1 : 0);
    };

    int iters = 0;
    if (!mixed) {
        if ((rc = factor_solve_fp64())) return rc;
    } else {
        float* B = (float*)Bw;
        float* Df = (float*)dinv;
        hipLaunchKernelGGL(k_build_b<float>, dim3(nblk((i64)n * n)), dim3(256), 0, ctx->stream, sd, V, W, B);
        cholesky<float>(ctx, B, Df, n, st);
        hipLaunchKernelGGL(k_tail<float>, dim3(nblk(n)), dim3(256), 0, ctx->stream, U, b2, f32a, N);
        if ((rc = chol_solve<float>(ctx, B, (const float*)Df, n, f32a, (float*)ys, nullptr, st))) return rc;
        hipLaunchKernelGGL((k_convert<float, double>), dim3(nblk((i64)n * 4)), dim3(256), 0, ctx->stream, f32a, z,
                           (i64)n * 4);
        TW_LAUNCH_CHECK(ctx);
        // Refinement sweeps run in batches; the device stops a sweep (and
        // turns the rest of its batch into no-ops) once the residual reaches
        // FP64 level or stalls, so the host checks once per batch.
        constexpr int kBatch = 3, kMaxSweeps = 21;
        bool ok = false;
        for (int done = 0; done < kMaxSweeps; done += kBatch) {
            for (int it = 0; it < kBatch; it++) {
                hipLaunchKernelGGL(k_embed, dim3(nblk(N)), dim3(256), 0, ctx->stream, z, U2, N, stop);
                hipLaunchKernelGGL(k_matvec, dim3(N), dim3(256), 0, ctx->stream, sd, U2, P, 0, stop);  // -A'[0;z]
                hipLaunchKernelGGL(k_ir_residual, dim3(1), dim3(1024), 0, ctx->stream, V, W, U2, P, b2, f32b, st,
                                   N);
                if ((rc = chol_solve<float>(ctx, B, (const float*)Df, n, f32b, (float*)ys, stop, st))) return rc;
                hipLaunchKernelGGL(k_axpy_f32, dim3(nblk((i64)n * 4)), dim3(256), 0, ctx->stream, z, f32b,
                                   (i64)n * 4, stop);
            }
            TW_LAUNCH_CHECK(ctx);
            if ((rc = fetch_state())) return rc;
            if (hst.info == 1) return tw_fail(ctx, TW_E_SINGULAR, "landmark centres are coplanar (rank-deficient X)");
            if (hst.info != 0) break;  // FP32 factor not positive definite
            if (hst.ir_stop) {
                ok = hst.ir_stop == 1;
                break;
            }
        }
        iters = hst.ir_iters;
        if (!ok) {
            // FP32 factors too weak for this system: factor in FP64 (U still
            // holds Q'Y).
            hipLaunchKernelGGL(k_ir_init, dim3(1), dim3(1), 0, ctx->stream, st);
            if ((rc = factor_solve_fp64())) return rc;
            iters = -std::max(iters, 1);
        }
    }
    // 3. c' = Q [0; z];  t = Y - A' c';  Q' t (first 4 rows give q1' t);
    //    d' = R^-1 (q1' t); FP64 residual of the bordered system.
    if (qfast) {
        hipLaunchKernelGGL(k_q_fast<256>, dim3(1), dim3(256), 0, ctx->stream, V, (const FitState*)st, U, N, 0,
                           (const double*)z, (const double*)nullptr, (double*)nullptr);
        hipLaunchKernelGGL(k_matvec, dim3(N), dim3(256), 0, ctx->stream, sd, U, U2, 1, (const int*)nullptr);
        hipLaunchKernelGGL(k_q_fast<256>, dim3(1), dim3(256), 0, ctx->stream, V, (const FitState*)st, U2, N, 1,
                           (const double*)nullptr, (const double*)U, x);
    } else {
        hipLaunchKernelGGL(k_embed, dim3(nblk(N)), dim3(256), 0, ctx->stream, z, U, N, (const int*)nullptr);
        hipLaunchKernelGGL(k_q_apply, dim3(1), dim3(1024), 0, ctx->stream, V, st, U, N, 0);
        hipLaunchKernelGGL(k_matvec, dim3(N), dim3(256), 0, ctx->stream, sd, U, U2, 1, (const int*)nullptr);
        hipLaunchKernelGGL(k_q_apply, dim3(1), dim3(1024), 0, ctx->stream, V, st, U2, N, 1);
        hipLaunchKernelGGL(k_finish, dim3(nblk(N)), dim3(256), 0, ctx->stream, U, U2, st, x, N);
    }
    hipLaunchKernelGGL(k_residual, dim3(m), dim3(256), 0, ctx->stream, sd, x, res);
    TW_LAUNCH_CHECK(ctx);
    out.x = x;
    out.res = res;
    out.st = st;
    out.iters = iters;
    return TW_OK;
}

// Fit state after a run: errors, residual, refinement count (synchronises).
int fit_finish(tw_ctx* ctx, const FitState* st, double bmax, int iters) {
    FitState hst;
    if (int rc = tw_read_small(ctx, &hst, st, sizeof hst)) return rc;
    if (hst.info == 1) return tw_fail(ctx, TW_E_SINGULAR, "landmark centres are coplanar (rank-deficient X)");
    if (hst.info >= 2)
        return tw_fail(ctx, TW_E_SINGULAR, "projected TPS system not positive definite (column %d)", hst.info - 2);
    if (hst.sync_timeout) {
        ctx->pipe_off = 1;  // the callers redo the fit with one launch per block
        return tw_fail(ctx, TW_E_DEVICE, "the solve's work-groups did not meet (wait bound reached)");
    }
    ctx->fit_residual = hst.fit_rmax / bmax;
    ctx->refine_iters = iters;
    if (!std::isfinite(ctx->fit_residual)) return tw_fail(ctx, TW_E_SINGULAR, "non-finite TPS coefficients");
    return TW_OK;
}

}  // namespace

extern "C" int tw_tps_fit(tw_ctx* ctx, const float* centres, const float* targets, int K, double lambda,
                          int precision, double* c, double* d) {
    if (!ctx) return TW_E_ARG;
    if (!c || !d) return tw_fail(ctx, TW_E_ARG, "null argument");
    int rc;
    if ((rc = check_fit_args(ctx, centres, targets, K, lambda, precision))) return rc;
    if ((rc = tw_bind(ctx))) return rc;
    const int pipe_was_off = ctx->pipe_off;
    tw_stages_reset(ctx);
    tw_stage_begin(ctx, ST_TOTAL);
    tw_stage_begin(ctx, ST_FIT);
    FitRun r;
    if ((rc = fit_run(ctx, centres, targets, K, lambda, precision, r))) return rc;
    const int N = r.N, m = r.m;
    if (r.res) hipLaunchKernelGGL(k_res_max, dim3(1), dim3(256), 0, ctx->stream, r.res, m * 4, r.st);
    std::vector<double> hx((size_t)m * 4);
    TW_HIP(ctx, hipMemcpyAsync(hx.data(), r.x, hx.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    tw_stage_end(ctx, ST_FIT);
    tw_stage_end(ctx, ST_TOTAL);
    if ((rc = fit_finish(ctx, r.st, r.bmax, r.iters))) {
        // A pipelined solve whose wait ran out: once more, per-block launches.
        if (ctx->pipe_off && !pipe_was_off) return tw_tps_fit(ctx, centres, targets, K, lambda, precision, c, d);
        return rc;
    }
    // Undo the scaling: c = c'/s, d = T^-1 d'.
    const double s = r.s;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < 4; j++) c[4 * i + j] = hx[(size_t)j * m + i] / s;
    for (int j = 0; j < 4; j++) {
        const double d0 = hx[(size_t)j * m + N], d1 = hx[(size_t)j * m + N + 1], d2 = hx[(size_t)j * m + N + 2],
                     d3 = hx[(size_t)j * m + N + 3];
        d[0 * 4 + j] = d0 - (r.mu[0] * d1 + r.mu[1] * d2 + r.mu[2] * d3) / s;
        d[1 * 4 + j] = d1 / s;
        d[2 * 4 + j] = d2 / s;
        d[3 * 4 + j] = d3 / s;
    }
    for (int i = 0; i < 4 * N; i++)
        if (!std::isfinite(c[i])) return tw_fail(ctx, TW_E_SINGULAR, "non-finite TPS coefficients");
    return TW_OK;
}

int tw_tps_fit_dev(tw_ctx* ctx, const float* centres, const float* targets, int K, double lambda, int precision,
                   TwFitDev* out) {
    int rc;
    if ((rc = check_fit_args(ctx, centres, targets, K, lambda, precision))) return rc;
    tw_stage_begin(ctx, ST_FIT);
    FitRun r;
    if ((rc = fit_run(ctx, centres, targets, K, lambda, precision, r))) return rc;
    void* p = nullptr;
    if ((rc = tw_ws(ctx, WS_COEF, ((size_t)K * 3 + 16) * sizeof(double), &p))) return rc;
    double* cf = (double*)p;
    double* d = cf + (size_t)K * 3;
    hipLaunchKernelGGL(k_unscale, dim3(nblk(K)), dim3(256), 0, ctx->stream, r.x, K, r.s, r.mu[0], r.mu[1], r.mu[2],
                       cf, d, r.res, (K + 4) * 4, r.st);
    TW_LAUNCH_CHECK(ctx);
    tw_stage_end(ctx, ST_FIT);
    out->pts = r.pts;
    out->cf = cf;
    out->d = d;
    out->state = r.st;
    out->K = K;
    out->bmax = r.bmax;
    ctx->refine_iters = r.iters;
    return TW_OK;
}

int tw_tps_fit_check(tw_ctx* ctx, const TwFitDev* fit) {
    return fit_finish(ctx, (const FitState*)fit->state, fit->bmax, ctx->refine_iters);
}

// ---- batches -----------------------------------------------------------------
// Every system up to FIT_SMALL_MAX landmarks in one launch of the one-work-
// group kernel (arithmetic identical to tw_tps_fit_dev's for the same
// system); larger systems through fit_run one after another.  Layout of
// WS_BATCH_FIT: landmark packs, small-system jobs, fit states, coefficients.
namespace {
constexpr size_t STATE_BYTES = (sizeof(FitState) + 255) & ~(size_t)255;
inline size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

int tw_fit_batch(tw_ctx* ctx, TwFitItem* items, int n, double lambda, int precision) {
    int rc;
    std::vector<HostSys> hs((size_t)n);
    std::vector<size_t> pack_off((size_t)n), ws_off((size_t)n), coef_off((size_t)n);
    size_t pack_total = 0, ws_total = 0, coef_total = 0;
    int nsmall = 0;
    for (int i = 0; i < n; i++) {
        const TwFitItem& it = items[i];
        if ((rc = check_fit_args(ctx, it.centres, it.targets, it.K, lambda, precision)))
            return tw_fail(ctx, rc, "item %d: %s", i, tw_last_error(ctx));
        if (!host_sys(it.centres, it.targets, it.K, hs[i]))
            return tw_fail(ctx, TW_E_SINGULAR, "item %d: all centres coincide", i);
        pack_off[i] = pack_total;
        pack_total += pack_pts_bytes(it.K) + pack_y_bytes(it.K);
        coef_off[i] = coef_total;
        coef_total += up256(((size_t)it.K * 3 + 16) * sizeof(double));
        if (it.K <= FIT_SMALL_MAX) {
            ws_off[i] = ws_total;
            ws_total += up256(small_ws_doubles(it.K) * sizeof(double));
            nsmall++;
        }
    }
    const size_t jobs_off = pack_total, jobs_b = up256((size_t)nsmall * sizeof(SmallFitJob));
    const size_t st_off = jobs_off + jobs_b, coefs_off = st_off + (size_t)n * STATE_BYTES;
    void *base = nullptr, *wsb = nullptr;
    if ((rc = tw_ws(ctx, WS_BATCH_FIT, coefs_off + coef_total, &base))) return rc;
    if ((rc = tw_ws(ctx, WS_BATCH_SOLVE, ws_total, &wsb))) return rc;
    uint8_t* dev = (uint8_t*)base;
    uint8_t* host = tw_host_stage(ctx, st_off);
    SmallFitJob* hjobs = (SmallFitJob*)(host + jobs_off);
    int js = 0;
    for (int i = 0; i < n; i++) {
        TwFitItem& it = items[i];
        pack_host(it.centres, it.targets, it.K, host + pack_off[i]);
        it.pts = (const float4*)(dev + pack_off[i]);
        it.cf = (double*)(dev + coefs_off + coef_off[i]);
        it.d = it.cf + (size_t)it.K * 3;
        it.state = dev + st_off + (size_t)i * STATE_BYTES;
        it.bmax = hs[i].bmax;
        if (it.K <= FIT_SMALL_MAX) {
            SmallFitJob& J = hjobs[js++];
            J.s = sys_desc(hs[i], dev + pack_off[i], it.K, lambda);
            J.ws = (double*)((uint8_t*)wsb + ws_off[i]);
            J.st = (FitState*)it.state;
            J.cf = it.cf;
            J.d = it.d;
            J.scale = hs[i].s;
        }
    }
    TW_HIP(ctx, hipMemcpyAsync(dev, host, st_off, hipMemcpyHostToDevice, ctx->stream));
    if (nsmall > 0) {
        hipLaunchKernelGGL(k_fit_small_batch, dim3((unsigned)nsmall), dim3(FIT_SMALL_THREADS), 0, ctx->stream,
                           (const SmallFitJob*)(dev + jobs_off));
        TW_LAUNCH_CHECK(ctx);
    }
    int iters = 0;
    for (int i = 0; i < n; i++) {
        TwFitItem& it = items[i];
        if (it.K <= FIT_SMALL_MAX) continue;
        FitRun r;
        if ((rc = fit_run(ctx, it.centres, it.targets, it.K, lambda, precision, r)))
            return tw_fail(ctx, rc, "item %d: %s", i, tw_last_error(ctx));
        hipLaunchKernelGGL(k_unscale, dim3(nblk(it.K)), dim3(256), 0, ctx->stream, r.x, it.K, r.s, r.mu[0], r.mu[1],
                           r.mu[2], it.cf, it.d, r.res, (it.K + 4) * 4, r.st);
        TW_LAUNCH_CHECK(ctx);
        TW_HIP(ctx, hipMemcpyAsync(it.state, r.st, sizeof(FitState), hipMemcpyDeviceToDevice, ctx->stream));
        iters = std::max(iters, r.iters);
    }
    ctx->refine_iters = iters;
    return TW_OK;
}

int tw_fit_batch_check(tw_ctx* ctx, const TwFitItem* items, int n, int* status) {
    if (n <= 0) return TW_OK;
    const uint8_t* st0 = (const uint8_t*)items[0].state;
    std::vector<uint8_t> h((size_t)n * STATE_BYTES);
    TW_HIP(ctx, hipMemcpy(h.data(), st0, h.size(), hipMemcpyDeviceToHost));
    int first = TW_OK;
    double worst = 0.0;
    char msg[256] = {0};
    for (int i = 0; i < n; i++) {
        FitState hst;
        memcpy(&hst, h.data() + (size_t)i * STATE_BYTES, sizeof hst);
        int code = TW_OK;
        if (hst.info == 1) {
            code = TW_E_SINGULAR;
            snprintf(msg, sizeof msg, "item %d: landmark centres are coplanar (rank-deficient X)", i);
        } else if (hst.sync_timeout) {
            code = TW_E_DEVICE;
            ctx->pipe_off = 1;  // the caller redoes the fits with one launch per block
            snprintf(msg, sizeof msg, "item %d: the solve's work-groups did not meet (wait bound reached)", i);
        } else if (hst.info >= 2) {
            code = TW_E_SINGULAR;
            snprintf(msg, sizeof msg, "item %d: projected TPS system not positive definite (column %d)", i,
                     hst.info - 2);
        } else {
            const double res = hst.fit_rmax / items[i].bmax;
            if (!std::isfinite(res)) {
                code = TW_E_SINGULAR;
                snprintf(msg, sizeof msg, "item %d: non-finite TPS coefficients", i);
            } else {
                worst = std::max(worst, res);
            }
        }
        if (status) status[i] = code;
        if (code != TW_OK && first == TW_OK) first = tw_fail(ctx, code, "%s", msg);
    }
    ctx->fit_residual = worst;
    return first;
}
