// tw_oracle.cpp -- FP64 CPU restatement of TeraWarp's warp path.
//
// TEST INFRASTRUCTURE ONLY.  This file is what the GPU results are compared
// with: it is loaded by the tests, by __graft_entry__.smoke() and by the
// CPU-baseline leg of bench.py, never by the terawarp_amd product library,
// which has no CPU fallback.
//
// Every function restates one function of the reference's CPU path
// (q_imgwarp_tps_quicksmallmemory.cpp, q_warp_affine_tps.cpp) with plain
// loops in IEEE double, compiled with -ffp-contract=off and no fast-math,
// so that each arithmetic expression rounds exactly as written.  newmat's
// Matrix algebra is replaced by explicit loops that sum in ascending index
// order (the order newmat's row-by-column products use).
//
// Parity pins (tests/test_oracle_*.py):
//   or_stps_fit        bordered-system residual, numpy.linalg.solve of the
//                      bordered system, identical / exactly-affine landmark
//                      sets, agreement with or_stps_fit_bordered.
//   or_grid_eval       interpolation identity f(p_i) = y_i - lambda*c_i at
//                      the landmarks, affine fit gives an affine field.
//   or_bspline_basis   closed-form weights, partition of unity, Kronecker
//                      ordering against a hand-built 4-node example.
//   or_tps_resample    linear precision of the upsample, integer
//                      positions, out-of-bounds strictness, brute force.
//   or_tps_resample_window  equal to or_tps_resample on the whole volume and
//                      on sub-windows whose source covers every tap; zero
//                      where the taps leave the window.
//   or_tps_warp_tera   one block over the volume = the in-memory chain of the
//                      4x markers; a block's grid = the volume grid's
//                      sub-grid; with small blocks, voxels differ from the
//                      in-memory warp only where the block-local float order
//                      of the position rounds differently.
//   or_affine_fit      exact recovery of a known matrix, order invariance.
//   or_affine_resample brute force on tiny volumes, identity/shift cases.
// newmat's own rounding sequences (QRZ, .i(), SVD) cannot be reproduced
// without newmat; they move results by ~1e-12 relative only.
//
// Layouts: volumes are [c][z][y][x], x fastest; the displacement grid is
// three float planes (sx, sy, sz), each [gz][gy][gx], x fastest; points are
// float xyz triples.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t i64;
// OR-ed into a resample's interp: upsample the grid with the linear DF basis
// (or_df_basis kind 1) instead of the cubic B-spline.
constexpr int DF_LINEAR_FLAG = 0x100;

namespace {

void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

// C = A (m x k) * B (k x n), all row-major.  Each C(i,j) is summed over k
// in ascending order, as newmat's Matrix product does.
void matmul(const double* A, const double* B, double* C, i64 m, i64 k, i64 n) {
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < m; i++) {
        double* c = C + i * n;
        for (i64 j = 0; j < n; j++) c[j] = 0.0;
        for (i64 p = 0; p < k; p++) {
            const double a = A[i * k + p];
            const double* b = B + p * n;
            for (i64 j = 0; j < n; j++) c[j] += a * b[j];
        }
    }
}

// LU with partial pivoting, in place on the n x n row-major matrix A, then
// solve A X = B for the n x m right-hand side B (overwritten by X).
// Returns 0, or -1 for an exactly singular pivot.  Stands for newmat's
// .i(), which solves through an LU decomposition.
int lu_solve(double* A, double* B, i64 n, i64 m) {
    std::vector<i64> piv(n);
    for (i64 col = 0; col < n; col++) {
        i64 p = col;
        double best = std::fabs(A[col * n + col]);
        for (i64 r = col + 1; r < n; r++) {
            double v = std::fabs(A[r * n + col]);
            if (v > best) { best = v; p = r; }
        }
        if (best == 0.0) return -1;
        piv[col] = p;
        if (p != col) {
            for (i64 j = 0; j < n; j++) std::swap(A[col * n + j], A[p * n + j]);
            for (i64 j = 0; j < m; j++) std::swap(B[col * m + j], B[p * m + j]);
        }
        const double d = A[col * n + col];
#pragma omp parallel for schedule(static) if (n - col > 256)
        for (i64 r = col + 1; r < n; r++) {
            double f = A[r * n + col] / d;
            A[r * n + col] = f;
            if (f != 0.0) {
                for (i64 j = col + 1; j < n; j++) A[r * n + j] -= f * A[col * n + j];
                for (i64 j = 0; j < m; j++) B[r * m + j] -= f * B[col * m + j];
            }
        }
    }
    for (i64 r = n - 1; r >= 0; r--) {
        for (i64 j = 0; j < m; j++) {
            double s = B[r * m + j];
            for (i64 q = r + 1; q < n; q++) s -= A[r * n + q] * B[q * m + j];
            B[r * m + j] = s / A[r * n + r];
        }
    }
    return 0;
}

// K_ij = -|p_i - p_j|.  The coordinate differences are taken in float, as
// the reference subtracts two MYFLOAT_JBA values before widening to double
// (q_imgwarp_tps_quicksmallmemory.cpp:434-444).
void build_K(const float* P, i64 N, double* K) {
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < N; i++)
        for (i64 j = 0; j < N; j++) {
            double dx = (float)(P[3 * i + 0] - P[3 * j + 0]);
            double dy = (float)(P[3 * i + 1] - P[3 * j + 1]);
            double dz = (float)(P[3 * i + 2] - P[3 * j + 2]);
            K[i * N + j] = -std::sqrt(dx * dx + dy * dy + dz * dz);
        }
}

template <class T>
inline double px(const T* v, i64 sx, i64 sxy, i64 x, i64 y, i64 z) {
    return (double)v[z * sxy + y * sx + x];
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// STPS fit.  Restates compute_df_stps_subsampled_volume_4bspline, CPU branch
// (q_imgwarp_tps_quicksmallmemory.cpp:377-586):
//   K = -|p_i - p_j| over the TPS centres p (434-444)
//   X = [1 p], Y = [1 q] (560-572)
//   QRZ(X) -> q1 (N x 4), r (4 x 4)          newmat QRZ, modified Gram-Schmidt
//   extend_orthonormal(Q, 4) -> q2 (N x N-4)  (575; algorithm as in the
//                                             commented body, extendornormal.cu:109-122)
//   A = q2' K q2 + lambda I                   (581)
//   c = q2 (A^-1 q2' Y)                       (582)
//   d = r^-1 q1' (Y - K c)                    (584)
// centres: K x 3 float (the TPS kernel centres; at the plugin call site
// these are the file-target markers, plugin_Bigimagewarp.cpp:197).
// targets: K x 3 float (points the centres map to: file-subject markers).
// c_out: K x 4 row-major; d_out: 4 x 4 row-major.  Returns 0, -1 bad args,
// -2 singular (coplanar centres or singular A).
int or_stps_fit(const float* centres, const float* targets, int K, double lambda,
                double* c_out, double* d_out, int nthreads) {
    if (K < 5 || !centres || !targets || !c_out || !d_out) return -1;
    set_threads(nthreads);
    const i64 N = K, n = N - 4;
    std::vector<double> Km((size_t)N * N);
    build_K(centres, N, Km.data());

    std::vector<double> Q((size_t)N * N, 0.0), Y((size_t)N * 4);
    for (i64 i = 0; i < N; i++) {
        Q[i * N + 0] = 1.0;
        Q[i * N + 1] = centres[3 * i + 0];
        Q[i * N + 2] = centres[3 * i + 1];
        Q[i * N + 3] = centres[3 * i + 2];
        Y[i * 4 + 0] = 1.0;
        Y[i * 4 + 1] = targets[3 * i + 0];
        Y[i * 4 + 2] = targets[3 * i + 1];
        Y[i * 4 + 3] = targets[3 * i + 2];
    }
    // QRZ on the four live columns: for column i, U(i, i..3) = X_i' X_j,
    // U(i,i) = sqrt, U(i, j>i) /= U(i,i), X_i /= U(i,i), X_j -= U(i,j) X_i.
    double R[16] = {0};
    for (int i = 0; i < 4; i++) {
        double u[4] = {0, 0, 0, 0};
        for (i64 r = 0; r < N; r++) {
            const double xi = Q[r * N + i];
            for (int j = i; j < 4; j++) u[j] += xi * Q[r * N + j];
        }
        const double s = std::sqrt(u[i]);
        if (s == 0.0) return -2;
        R[i * 4 + i] = s;
        for (int j = i + 1; j < 4; j++) R[i * 4 + j] = u[j] / s;
        for (i64 r = 0; r < N; r++) {
            const double xi = Q[r * N + i] / s;
            Q[r * N + i] = xi;
            for (int j = i + 1; j < 4; j++) Q[r * N + j] -= R[i * 4 + j] * xi;
        }
    }
    for (int i = 1; i < 4; i++)
        if (R[i * 4 + i] <= 1e-12 * R[0]) return -2;  // rank-deficient X
    // extend_orthonormal(Q, 4)
    std::vector<double> ssr(N, 0.0), X(N);
    for (i64 r = 0; r < N; r++)
        for (int j = 0; j < 4; j++) ssr[r] += Q[r * N + j] * Q[r * N + j];
    for (i64 i = 4; i < N; i++) {
        i64 k = 0;
        double mn = ssr[0];
        for (i64 r = 1; r < N; r++)
            if (ssr[r] < mn) { mn = ssr[r]; k = r; }
#pragma omp parallel for schedule(static) if (N > 512)
        for (i64 r = 0; r < N; r++) {
            double s = 0.0;
            for (i64 j = 0; j < i; j++) s += -Q[r * N + j] * Q[k * N + j];
            X[r] = s;
        }
        X[k] += 1.0;
        double ss = 0.0;
        for (i64 r = 0; r < N; r++) ss += X[r] * X[r];
        const double nrm = std::sqrt(ss);
        for (i64 r = 0; r < N; r++) {
            X[r] /= nrm;
            ssr[r] += X[r] * X[r];
            Q[r * N + i] = X[r];
        }
    }
    // q2 (N x n), q2t (n x N)
    std::vector<double> q2((size_t)N * n), q2t((size_t)n * N);
    for (i64 r = 0; r < N; r++)
        for (i64 j = 0; j < n; j++) {
            q2[r * n + j] = Q[r * N + 4 + j];
            q2t[j * N + r] = Q[r * N + 4 + j];
        }
    std::vector<double> KQ((size_t)N * n), A((size_t)n * n);
    matmul(Km.data(), q2.data(), KQ.data(), N, N, n);
    matmul(q2t.data(), KQ.data(), A.data(), n, N, n);
    for (i64 i = 0; i < n; i++) A[i * n + i] += lambda;
    std::vector<double> Z((size_t)n * 4);
    matmul(q2t.data(), Y.data(), Z.data(), n, N, 4);
    if (lu_solve(A.data(), Z.data(), n, 4) != 0) return -2;
    matmul(q2.data(), Z.data(), c_out, N, n, 4);
    // d = r^-1 q1' (Y - K c)
    std::vector<double> Kc((size_t)N * 4), W((size_t)N * 4);
    matmul(Km.data(), c_out, Kc.data(), N, N, 4);
    for (i64 i = 0; i < N * 4; i++) W[i] = Y[i] - Kc[i];
    double V[16] = {0};
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) {
            double s = 0.0;
            for (i64 r = 0; r < N; r++) s += Q[r * N + a] * W[r * 4 + b];
            V[a * 4 + b] = s;
        }
    for (int b = 0; b < 4; b++)
        for (int a = 3; a >= 0; a--) {
            double s = V[a * 4 + b];
            for (int q = a + 1; q < 4; q++) s -= R[a * 4 + q] * d_out[q * 4 + b];
            d_out[a * 4 + b] = s / R[a * 4 + a];
        }
    return 0;
}

// Direct FP64 solve of the bordered system that the null-space method above
// solves implicitly:  [[K + lambda I, X], [X', 0]] [c; d] = [Y; 0].
// Independent cross-check of or_stps_fit; same outputs and layouts.
int or_stps_fit_bordered(const float* centres, const float* targets, int K, double lambda,
                         double* c_out, double* d_out, int nthreads) {
    if (K < 5 || !centres || !targets || !c_out || !d_out) return -1;
    set_threads(nthreads);
    const i64 N = K, M = N + 4;
    std::vector<double> Km((size_t)N * N), B((size_t)M * M, 0.0), rhs((size_t)M * 4, 0.0);
    build_K(centres, N, Km.data());
    for (i64 i = 0; i < N; i++) {
        for (i64 j = 0; j < N; j++) B[i * M + j] = Km[i * N + j];
        B[i * M + i] += lambda;
        const double xr[4] = {1.0, centres[3 * i], centres[3 * i + 1], centres[3 * i + 2]};
        for (int a = 0; a < 4; a++) {
            B[i * M + N + a] = xr[a];
            B[(N + a) * M + i] = xr[a];
        }
        rhs[i * 4 + 0] = 1.0;
        for (int a = 0; a < 3; a++) rhs[i * 4 + 1 + a] = targets[3 * i + a];
    }
    if (lu_solve(B.data(), rhs.data(), M, 4) != 0) return -2;
    for (i64 i = 0; i < N * 4; i++) c_out[i] = rhs[i];
    for (int i = 0; i < 16; i++) d_out[i] = rhs[N * 4 + i];
    return 0;
}

// ---------------------------------------------------------------------------
// Grid displacement field.  Restates the CPU triple loop of
// compute_df_stps_subsampled_volume_4bspline (q_imgwarp_tps_quicksmallmemory.cpp:588-649):
// node (i,j,k) sits at ((i-1)g + ox, (j-1)g + oy, (k-1)g + oz) (628-630;
// offsets as in the block variant, 1126-1143); the kernel distance uses
// float differences of the integer node coordinate and the float centre
// (635-638); x_stps = x_ori*d + xmxn_K*c (642); DF = x_stps - node,
// stored as float (643-645).
// df: three planes of gz*gy*gx floats (sx, then sy, then sz).
int or_grid_eval(const float* centres, int K, const double* c, const double* d,
                 i64 gx, i64 gy, i64 gz, int fx, int fy, int fz,
                 i64 ox, i64 oy, i64 oz, float* df, int nthreads) {
    if (K < 1 || gx < 1 || gy < 1 || gz < 1 || !df) return -1;
    set_threads(nthreads);
    const i64 plane = gx * gy * gz;
#pragma omp parallel for schedule(dynamic)
    for (i64 k = 0; k < gz; k++) {
        std::vector<double> kr(K);
        for (i64 j = 0; j < gy; j++)
            for (i64 i = 0; i < gx; i++) {
                const i64 X = (i - 1) * fx + ox, Yc = (j - 1) * fy + oy, Z = (k - 1) * fz + oz;
                const double xo[4] = {1.0, (double)X, (double)Yc, (double)Z};
                for (i64 n = 0; n < K; n++) {
                    double dx = (float)((float)X - centres[3 * n + 0]);
                    double dy = (float)((float)Yc - centres[3 * n + 1]);
                    double dz = (float)((float)Z - centres[3 * n + 2]);
                    kr[n] = -std::sqrt(dx * dx + dy * dy + dz * dz);
                }
                double xs[4];
                for (int col = 1; col < 4; col++) {
                    double a = 0.0;
                    for (int m = 0; m < 4; m++) a += xo[m] * d[m * 4 + col];
                    double b = 0.0;
                    for (i64 n = 0; n < K; n++) b += kr[n] * c[n * 4 + col];
                    xs[col] = a + b;
                }
                const i64 idx = (k * gy + j) * gx + i;
                df[idx] = (float)(xs[1] - (double)X);
                df[plane + idx] = (float)(xs[2] - (double)Yc);
                df[2 * plane + idx] = (float)(xs[3] - (double)Z);
            }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Cubic B-spline basis.  Restates q_nonrigid_ini_bsplinebasis_3D
// (q_imgwarp_tps_quicksmallmemory.cpp:663-699): B = M/6 (673-678),
// T(i,:) = [t^3 t^2 t 1], t = i/n (681-688), TB = T*B (691),
// BxBxB = KP(KP(TB,TB),TB) (694-696).
// TB: n x 4; basis: n^3 x 64 (row (i1*n+i2)*n+i3, col (j1*4+j2)*4+j3).
//
// kind 1: the linear DF upsample.  The reference declares it
// (i_interpmethod_df = 0, q_imgwarp_tps_quicksmallmemory.h:14-17,
// .cpp:1158, 2021-2025) and never defines q_dfblcokinterp_linear; this
// restatement defines it on the B-spline's block layout: voxel b*n + i uses
// nodes b+1 and b+2 (at b*n and (b+1)*n) with weights 1 - t and t, t = i/n,
// so TB(i,:) = [0, 1 - t, t, 0] and BxBxB is built as for the B-spline.
int or_df_basis(int n, int kind, double* TB, double* basis) {
    if (n <= 0 || !TB || (kind != 0 && kind != 1)) return -1;
    const double step = 1.0 / n;
    if (kind == 1) {
        for (int i = 0; i < n; i++) {
            const double t = step * i;
            TB[i * 4 + 0] = 0.0;
            TB[i * 4 + 1] = 1.0 - t;
            TB[i * 4 + 2] = t;
            TB[i * 4 + 3] = 0.0;
        }
    } else {
        static const double Mi[4][4] = {{-1, 3, -3, 1}, {3, -6, 3, 0}, {-3, 0, 3, 0}, {1, 4, 1, 0}};
        double Bm[4][4];
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) Bm[a][b] = Mi[a][b] / 6.0;
        for (int i = 0; i < n; i++) {
            const double t = step * i;
            const double T[4] = {std::pow(t, 3.0), std::pow(t, 2.0), std::pow(t, 1.0), std::pow(t, 0.0)};
            for (int j = 0; j < 4; j++) {
                double s = 0.0;
                for (int m = 0; m < 4; m++) s += T[m] * Bm[m][j];
                TB[i * 4 + j] = s;
            }
        }
    }
    if (basis) {
        for (int i1 = 0; i1 < n; i1++)
            for (int i2 = 0; i2 < n; i2++)
                for (int i3 = 0; i3 < n; i3++)
                    for (int j1 = 0; j1 < 4; j1++)
                        for (int j2 = 0; j2 < 4; j2++)
                            for (int j3 = 0; j3 < 4; j3++) {
                                const i64 row = ((i64)i1 * n + i2) * n + i3;
                                const int col = (j1 * 4 + j2) * 4 + j3;
                                basis[row * 64 + col] = (TB[i1 * 4 + j1] * TB[i2 * 4 + j2]) * TB[i3 * 4 + j3];
                            }
    }
    return 0;
}

int or_bspline_basis(int n, double* TB, double* basis) { return or_df_basis(n, 0, TB, basis); }

}  // extern "C"

namespace {

// One block of q_dfblcokinterp_bspline (q_imgwarp_tps_quicksmallmemory.cpp:169-203):
// gather the 4x4x4 nodes in dep(z) -> col(x) -> row(y) order (176-184),
// multiply by the basis (188), scatter rows in zz -> xx -> yy order
// (192-200).  out: n^3 x 3 floats indexed ((zz*n + yy)*n + xx)*3 + comp.
void df_block(const float* df, i64 gx, i64 gy, i64 gz, const double* basis, int n,
              i64 bx, i64 by, i64 bz, float* out) {
    const i64 plane = gx * gy * gz;
    double P[64][3];
    int ind = 0;
    for (i64 dep = bz; dep < bz + 4; dep++)
        for (i64 col = bx; col < bx + 4; col++)
            for (i64 row = by; row < by + 4; row++) {
                const i64 idx = (dep * gy + row) * gx + col;
                P[ind][0] = df[idx];
                P[ind][1] = df[plane + idx];
                P[ind][2] = df[2 * plane + idx];
                ind++;
            }
    i64 r = 0;
    for (int zz = 0; zz < n; zz++)
        for (int xx = 0; xx < n; xx++)
            for (int yy = 0; yy < n; yy++, r++) {
                const double* b = basis + r * 64;
                for (int comp = 0; comp < 3; comp++) {
                    double s = 0.0;
                    for (int q = 0; q < 64; q++) s += b[q] * P[q][comp];
                    out[(((i64)zz * n + yy) * n + xx) * 3 + comp] = (float)s;
                }
            }
}

// q_imgblockwarp (q_imgwarp_tps_quicksmallmemory.cpp:85-167) for one block.
template <class T>
void block_warp(const T* src, const i64* sz, const float* dfb, int n, i64 bx, i64 by, i64 bz,
                int interp, T* dst) {
    const i64 sx = sz[0], sxy = sz[0] * sz[1], vol = sxy * sz[2];
    for (int z = 0; z < n; z++)
        for (int y = 0; y < n; y++)
            for (int x = 0; x < n; x++) {
                const i64 wx = bx * n + x, wy = by * n + y, wz = bz * n + z;
                if (wx >= sz[0] || wy >= sz[1] || wz >= sz[2]) continue;
                const float* dv = dfb + (((i64)z * n + y) * n + x) * 3;
                // pos_warp (long) + sx (float) is a float sum in the
                // reference (106-108), widened to double afterwards.
                double p[3];
                p[0] = (float)wx + dv[0];
                p[1] = (float)wy + dv[1];
                p[2] = (float)wz + dv[2];
                const i64 o = wz * sxy + wy * sx + wx;
                if (p[0] < 0 || p[0] > sz[0] - 1 || p[1] < 0 || p[1] > sz[1] - 1 ||
                    p[2] < 0 || p[2] > sz[2] - 1) {
                    for (i64 ch = 0; ch < sz[3]; ch++) dst[ch * vol + o] = 0;
                    continue;
                }
                if (interp == 1) {  // nearest neighbour (119-129)
                    i64 q[3];
                    for (int a = 0; a < 3; a++) {
                        q[a] = (i64)(p[a] + 0.5);
                        q[a] = q[a] < sz[a] ? q[a] : sz[a] - 1;
                    }
                    for (i64 ch = 0; ch < sz[3]; ch++)
                        dst[ch * vol + o] = src[ch * vol + q[2] * sxy + q[1] * sx + q[0]];
                    continue;
                }
                const i64 xs = (i64)std::floor(p[0]), xb = (i64)std::ceil(p[0]);
                const i64 ys = (i64)std::floor(p[1]), yb = (i64)std::ceil(p[1]);
                const i64 zs = (i64)std::floor(p[2]), zb = (i64)std::ceil(p[2]);
                const double lw = 1.0 - (p[0] - xs), rw = 1.0 - lw;
                const double tw = 1.0 - (p[1] - ys), bw = 1.0 - tw;
                const double uw = 1.0 - (p[2] - zs), dw = 1.0 - uw;
                for (i64 ch = 0; ch < sz[3]; ch++) {
                    const T* s = src + ch * vol;
                    const double hi = tw * (lw * px(s, sx, sxy, xs, ys, zs) + rw * px(s, sx, sxy, xb, ys, zs)) +
                                      bw * (lw * px(s, sx, sxy, xs, yb, zs) + rw * px(s, sx, sxy, xb, yb, zs));
                    const double lo = tw * (lw * px(s, sx, sxy, xs, ys, zb) + rw * px(s, sx, sxy, xb, ys, zb)) +
                                      bw * (lw * px(s, sx, sxy, xs, yb, zb) + rw * px(s, sx, sxy, xb, yb, zb));
                    const double iv = uw * hi + dw * lo + 0.5;
                    dst[ch * vol + o] = (T)iv;
                }
            }
}

template <class T>
int tps_resample(const T* src, const i64* sz, const float* df, i64 gx, i64 gy, i64 gz, int n,
                 int interp, T* dst, i64 zb0, i64 zb1) {
    std::vector<double> TB((size_t)n * 4), basis((size_t)n * n * n * 64);
    or_df_basis(n, (interp & DF_LINEAR_FLAG) ? 1 : 0, TB.data(), basis.data());
    interp &= ~DF_LINEAR_FLAG;
    const i64 nbx = (sz[0] + n - 1) / n, nby = (sz[1] + n - 1) / n, nbz = (sz[2] + n - 1) / n;
    if (nbx + 3 > gx || nby + 3 > gy || nbz + 3 > gz) return -1;
    if (zb1 < 0 || zb1 > nbz) zb1 = nbz;
#pragma omp parallel for schedule(dynamic)
    for (i64 bz = zb0; bz < zb1; bz++) {
        std::vector<float> blk((size_t)n * n * n * 3);
        for (i64 by = 0; by < nby; by++)
            for (i64 bx = 0; bx < nbx; bx++) {
                df_block(df, gx, gy, gz, basis.data(), n, bx, by, bz, blk.data());
                block_warp(src, sz, blk.data(), n, bx, by, bz, interp, dst);
            }
    }
    return 0;
}

// Tera-block form of the warp (q_imgblockwarp_CYF,
// q_imgwarp_tps_quicksmallmemory.cpp:2208-2304): the output window
// [d0, d0 + dsz) of a volume of full extent fsz, read from the source window
// [s0, s0 + ssz) of that volume, with a grid whose node (i, j, k) sits at
// ((i-1)n + g0x, (j-1)n + g0y, (k-1)n + g0z).  A voxel's sample position is
// the float sum of its global coordinate and the upsampled displacement (the
// in-memory path's arithmetic, 106-108); a sample outside the full extent
// gives 0 (2235-2242), and so does one whose taps leave the source window
// (2243-2251).  Blocks are those of the grid (origin g0), visited only where
// they hold output voxels.
template <class T>
int tps_resample_window(const T* src, const i64* ssz, const i64* s0, const i64* fsz, const float* df, i64 gx,
                        i64 gy, i64 gz, const i64* g0, int n, int interp, T* dst, const i64* dsz, const i64* d0) {
    std::vector<double> TB((size_t)n * 4), basis((size_t)n * n * n * 64);
    or_df_basis(n, (interp & DF_LINEAR_FLAG) ? 1 : 0, TB.data(), basis.data());
    interp &= ~DF_LINEAR_FLAG;
    for (int a = 0; a < 3; a++)
        if (d0[a] < g0[a] || dsz[a] < 1 || ssz[a] < 1) return -1;
    // Block range (grid-relative) holding output voxels.
    const i64 bx0 = (d0[0] - g0[0]) / n, bx1 = (d0[0] + dsz[0] - 1 - g0[0]) / n;
    const i64 by0 = (d0[1] - g0[1]) / n, by1 = (d0[1] + dsz[1] - 1 - g0[1]) / n;
    const i64 bz0 = (d0[2] - g0[2]) / n, bz1 = (d0[2] + dsz[2] - 1 - g0[2]) / n;
    if (bx1 + 3 >= gx || by1 + 3 >= gy || bz1 + 3 >= gz) return -1;
    const i64 ssx = ssz[0], ssxy = ssz[0] * ssz[1];
    const i64 dsx = dsz[0], dsxy = dsz[0] * dsz[1];
#pragma omp parallel for schedule(dynamic)
    for (i64 bz = bz0; bz <= bz1; bz++) {
        std::vector<float> blk((size_t)n * n * n * 3);
        for (i64 by = by0; by <= by1; by++)
            for (i64 bx = bx0; bx <= bx1; bx++) {
                df_block(df, gx, gy, gz, basis.data(), n, bx, by, bz, blk.data());
                for (int z = 0; z < n; z++)
                    for (int y = 0; y < n; y++)
                        for (int x = 0; x < n; x++) {
                            const i64 X = g0[0] + bx * n + x, Y = g0[1] + by * n + y, Z = g0[2] + bz * n + z;
                            const i64 lx = X - d0[0], ly = Y - d0[1], lz = Z - d0[2];
                            if (lx < 0 || ly < 0 || lz < 0 || lx >= dsz[0] || ly >= dsz[1] || lz >= dsz[2]) continue;
                            const float* dv = blk.data() + (((i64)z * n + y) * n + x) * 3;
                            double p[3];
                            p[0] = (float)X + dv[0];
                            p[1] = (float)Y + dv[1];
                            p[2] = (float)Z + dv[2];
                            T& out = dst[lz * dsxy + ly * dsx + lx];
                            if (p[0] < 0 || p[0] > fsz[0] - 1 || p[1] < 0 || p[1] > fsz[1] - 1 || p[2] < 0 ||
                                p[2] > fsz[2] - 1) {
                                out = 0;
                                continue;
                            }
                            if (interp == 1) {
                                i64 q[3];
                                bool in = true;
                                for (int a = 0; a < 3; a++) {
                                    q[a] = (i64)(p[a] + 0.5);
                                    q[a] = (q[a] < fsz[a] ? q[a] : fsz[a] - 1) - s0[a];
                                    in = in && q[a] >= 0 && q[a] < ssz[a];
                                }
                                out = in ? src[q[2] * ssxy + q[1] * ssx + q[0]] : (T)0;
                                continue;
                            }
                            const i64 xs = (i64)std::floor(p[0]), xb = (i64)std::ceil(p[0]);
                            const i64 ys = (i64)std::floor(p[1]), yb = (i64)std::ceil(p[1]);
                            const i64 zs = (i64)std::floor(p[2]), zb = (i64)std::ceil(p[2]);
                            const double lw = 1.0 - (p[0] - xs), rw = 1.0 - lw;
                            const double tw = 1.0 - (p[1] - ys), bw = 1.0 - tw;
                            const double uw = 1.0 - (p[2] - zs), dw = 1.0 - uw;
                            const i64 lxs = xs - s0[0], lxb = xb - s0[0], lys = ys - s0[1], lyb = yb - s0[1];
                            const i64 lzs = zs - s0[2], lzb = zb - s0[2];
                            if (lxs < 0 || lxb >= ssz[0] || lys < 0 || lyb >= ssz[1] || lzs < 0 || lzb >= ssz[2]) {
                                out = 0;
                                continue;
                            }
                            const double hi =
                                tw * (lw * px(src, ssx, ssxy, lxs, lys, lzs) + rw * px(src, ssx, ssxy, lxb, lys, lzs)) +
                                bw * (lw * px(src, ssx, ssxy, lxs, lyb, lzs) + rw * px(src, ssx, ssxy, lxb, lyb, lzs));
                            const double lo =
                                tw * (lw * px(src, ssx, ssxy, lxs, lys, lzb) + rw * px(src, ssx, ssxy, lxb, lys, lzb)) +
                                bw * (lw * px(src, ssx, ssxy, lxs, lyb, lzb) + rw * px(src, ssx, ssxy, lxb, lyb, lzb));
                            const double iv = uw * hi + dw * lo + 0.5;
                            out = (T)iv;
                        }
            }
    }
    return 0;
}

// Tera-block TPS driver: the CPU branch of imgwarp_smallmemory_CYF
// (q_imgwarp_tps_quicksmallmemory.cpp:1161-1994) over a volume held in
// memory, after the fit (c, d of the 4x-scaled centres cen):
//  * output blocks of `block` voxels stepped z, then y, then x from 0
//    (per_block_siza, 1264, 1309-1317), edge blocks clipped;
//  * each block gets its own grid at spacing n with the block's offset
//    (compute_df_stps_subsampled_volume_4bspline_block, 1084-1152:
//    ceil(size/n) + 3 nodes per axis, node i at (i-1)n + offset);
//  * per n^3 sub-block (substart < gsz - 3, STPS_interpolate_CYF, 66-79) the
//    B-spline upsample and q_imgblockwarp_CYF (2208-2304): the sample
//    position is (pos_warp + sx) + x_offset with pos_warp block-local, each
//    addition rounded to float (2229-2235), tested against the whole
//    volume's extent (2237-2242); then floor/ceil, 1 - frac weights, the
//    t/b, l/r, u/d tree, +0.5 and truncation (2266-2296).
// The reference reads a block's source from a window it estimates from 14
// probes of the field with a 60-voxel margin (1279, 1403-1711) and zeroes
// samples outside it (2243-2250).  A position minus the window's integer
// origin is an exact float difference for every sample inside the window,
// so there the arithmetic equals this whole-volume form; samples the
// estimate misses are not reproduced.  Linear image interpolation only:
// the reference's nearest-neighbour branch clamps window-relative indices
// against the block extent (2254-2263).
template <class T>
int tps_warp_tera(const T* src, const i64* sz, const float* cen, const double* c, const double* d, int K, int n,
                  i64 block, int kind, T* dst) {
    std::vector<double> TB((size_t)n * 4), basis((size_t)n * n * n * 64);
    or_df_basis(n, kind, TB.data(), basis.data());
    const i64 sx = sz[0], sxy = sz[0] * sz[1];
    for (i64 oz = 0; oz < sz[2]; oz += block)
        for (i64 oy = 0; oy < sz[1]; oy += block)
            for (i64 ox = 0; ox < sz[0]; ox += block) {
                const i64 bs[3] = {std::min(block, sz[0] - ox), std::min(block, sz[1] - oy),
                                   std::min(block, sz[2] - oz)};
                const i64 gx = (bs[0] + n - 1) / n + 3, gy = (bs[1] + n - 1) / n + 3, gz = (bs[2] + n - 1) / n + 3;
                std::vector<float> df((size_t)(3 * gx * gy * gz));
                if (or_grid_eval(cen, K, c, d, gx, gy, gz, n, n, n, ox, oy, oz, df.data(), 0)) return -1;
#pragma omp parallel for schedule(dynamic)
                for (i64 bz = 0; bz < gz - 3; bz++) {
                    std::vector<float> blk((size_t)n * n * n * 3);
                    for (i64 by = 0; by < gy - 3; by++)
                        for (i64 bx = 0; bx < gx - 3; bx++) {
                            df_block(df.data(), gx, gy, gz, basis.data(), n, bx, by, bz, blk.data());
                            for (int z = 0; z < n; z++)
                                for (int y = 0; y < n; y++)
                                    for (int x = 0; x < n; x++) {
                                        const i64 w[3] = {bx * n + x, by * n + y, bz * n + z};
                                        if (w[0] >= bs[0] || w[1] >= bs[1] || w[2] >= bs[2]) continue;
                                        const float* dv = blk.data() + (((i64)z * n + y) * n + x) * 3;
                                        const i64 off[3] = {ox, oy, oz};
                                        double p[3];
                                        for (int a = 0; a < 3; a++) {
                                            const float local = (float)w[a] + dv[a];
                                            p[a] = local + (float)off[a];
                                        }
                                        T& out = dst[(oz + w[2]) * sxy + (oy + w[1]) * sx + (ox + w[0])];
                                        if (p[0] < 0 || p[0] > sz[0] - 1 || p[1] < 0 || p[1] > sz[1] - 1 ||
                                            p[2] < 0 || p[2] > sz[2] - 1) {
                                            out = 0;
                                            continue;
                                        }
                                        const i64 xs = (i64)std::floor(p[0]), xb = (i64)std::ceil(p[0]);
                                        const i64 ys = (i64)std::floor(p[1]), yb = (i64)std::ceil(p[1]);
                                        const i64 zs = (i64)std::floor(p[2]), zb = (i64)std::ceil(p[2]);
                                        const double lw = 1.0 - (p[0] - xs), rw = 1.0 - lw;
                                        const double tw = 1.0 - (p[1] - ys), bw = 1.0 - tw;
                                        const double uw = 1.0 - (p[2] - zs), dw = 1.0 - uw;
                                        const double hi =
                                            tw * (lw * px(src, sx, sxy, xs, ys, zs) + rw * px(src, sx, sxy, xb, ys, zs)) +
                                            bw * (lw * px(src, sx, sxy, xs, yb, zs) + rw * px(src, sx, sxy, xb, yb, zs));
                                        const double lo =
                                            tw * (lw * px(src, sx, sxy, xs, ys, zb) + rw * px(src, sx, sxy, xb, ys, zb)) +
                                            bw * (lw * px(src, sx, sxy, xs, yb, zb) + rw * px(src, sx, sxy, xb, yb, zb));
                                        const double iv = uw * hi + dw * lo + 0.5;
                                        out = (T)iv;
                                    }
                        }
                }
            }
    return 0;
}

}  // namespace

extern "C" {

// The tera-block TPS driver (tps_warp_tera above) from the marker files'
// coordinates: both marker sets are scaled by 4 (x_para = y_para = z_para =
// 4, q_imgwarp_tps_quicksmallmemory.cpp:1194-1204; int times float), the
// target-space markers are the TPS centres (the plugin calls
// imgwarp_smallmemory_CYF(sub, tar, ...) into parameters (tar, sub),
// plugin_Bigimagewarp.cpp:294), then or_stps_fit with lambda and the block
// warp.  tar, sub: K x 3 float; sz = {x, y, z}; one channel; dtype 1 / 2;
// interp 0 (linear) or DF_LINEAR_FLAG (linear DF basis).  Returns 0, -1 bad
// arguments, -2 singular fit.
int or_tps_warp_tera(const void* src, const i64* sz, int dtype, const float* tar, const float* sub, int K,
                     int n, i64 block, double lambda, int interp, void* dst, int nthreads) {
    if (!src || !sz || !tar || !sub || !dst || K < 1 || n <= 0 || block < n || block % n) return -1;
    if ((interp & ~DF_LINEAR_FLAG) != 0) return -1;
    set_threads(nthreads);
    std::vector<float> cen((size_t)K * 3), tgt((size_t)K * 3);
    for (int i = 0; i < 3 * K; i++) {
        cen[(size_t)i] = 4 * tar[i];
        tgt[(size_t)i] = 4 * sub[i];
    }
    std::vector<double> c((size_t)K * 4), d(16);
    const int rc = or_stps_fit(cen.data(), tgt.data(), K, lambda, c.data(), d.data(), nthreads);
    if (rc) return rc;
    const int kind = (interp & DF_LINEAR_FLAG) ? 1 : 0;
    if (dtype == 1)
        return tps_warp_tera((const uint8_t*)src, sz, cen.data(), c.data(), d.data(), K, n, block, kind,
                             (uint8_t*)dst);
    if (dtype == 2)
        return tps_warp_tera((const uint16_t*)src, sz, cen.data(), c.data(), d.data(), K, n, block, kind,
                             (uint16_t*)dst);
    return -1;
}

// In-memory TPS warp, CPU branch of imgwarp_smallmemory_image
// (q_imgwarp_tps_quicksmallmemory.cpp:2119-2133): for every n^3 block,
// B-spline-upsample the grid DF and warp the block.  Only blocks that hold
// output voxels are visited (the reference also visits blocks past the
// volume, which write nothing).  dtype 1 = uint8, 2 = uint16.
// sz = {x, y, z, c}.  zblock range [zb0, zb1) restricts the work to a slab
// of block rows (zb1 < 0: to the end); the rest of dst is left untouched.
int or_tps_resample(const void* src, const i64* sz, int dtype, const float* df,
                    i64 gx, i64 gy, i64 gz, int n, int interp, void* dst,
                    i64 zb0, i64 zb1, int nthreads) {
    if (!src || !sz || !df || !dst || n <= 0) return -1;
    set_threads(nthreads);
    if (dtype == 1)
        return tps_resample((const uint8_t*)src, sz, df, gx, gy, gz, n, interp, (uint8_t*)dst, zb0, zb1);
    if (dtype == 2)
        return tps_resample((const uint16_t*)src, sz, df, gx, gy, gz, n, interp, (uint16_t*)dst, zb0, zb1);
    return -1;
}

// Windowed form (tps_resample_window above).  dtype 1 = uint8, 2 = uint16;
// ssz, s0, fsz, g0, dsz, d0: xyz triples.
int or_tps_resample_window(const void* src, const i64* ssz, const i64* s0, const i64* fsz, int dtype,
                           const float* df, i64 gx, i64 gy, i64 gz, const i64* g0, int n, int interp, void* dst,
                           const i64* dsz, const i64* d0, int nthreads) {
    if (!src || !ssz || !s0 || !fsz || !df || !g0 || !dst || !dsz || !d0 || n <= 0) return -1;
    set_threads(nthreads);
    if (dtype == 1)
        return tps_resample_window((const uint8_t*)src, ssz, s0, fsz, df, gx, gy, gz, g0, n, interp, (uint8_t*)dst,
                                   dsz, d0);
    if (dtype == 2)
        return tps_resample_window((const uint16_t*)src, ssz, s0, fsz, df, gx, gy, gz, g0, n, interp,
                                   (uint16_t*)dst, dsz, d0);
    return -1;
}

// Dense per-voxel DF of the B-spline upsample (the pppDFBlock values of
// q_dfblcokinterp_bspline for every voxel; kind 1: the linear upsample of
// or_df_basis), three float planes of sz0*sz1*sz2 each.  Used to state the
// displacement error of each stage.
int or_dense_df(const i64* sz, const float* df, i64 gx, i64 gy, i64 gz, int n, float* out,
                int nthreads, int kind) {
    if (!sz || !df || !out || n <= 0) return -1;
    set_threads(nthreads);
    std::vector<double> TB((size_t)n * 4), basis((size_t)n * n * n * 64);
    if (or_df_basis(n, kind, TB.data(), basis.data())) return -1;
    const i64 nbx = (sz[0] + n - 1) / n, nby = (sz[1] + n - 1) / n, nbz = (sz[2] + n - 1) / n;
    if (nbx + 3 > gx || nby + 3 > gy || nbz + 3 > gz) return -1;
    const i64 sx = sz[0], sxy = sz[0] * sz[1], vol = sxy * sz[2];
#pragma omp parallel for schedule(dynamic)
    for (i64 bz = 0; bz < nbz; bz++) {
        std::vector<float> blk((size_t)n * n * n * 3);
        for (i64 by = 0; by < nby; by++)
            for (i64 bx = 0; bx < nbx; bx++) {
                df_block(df, gx, gy, gz, basis.data(), n, bx, by, bz, blk.data());
                for (int z = 0; z < n; z++)
                    for (int y = 0; y < n; y++)
                        for (int x = 0; x < n; x++) {
                            const i64 wx = bx * n + x, wy = by * n + y, wz = bz * n + z;
                            if (wx >= sz[0] || wy >= sz[1] || wz >= sz[2]) continue;
                            const float* dv = blk.data() + (((i64)z * n + y) * n + x) * 3;
                            const i64 o = wz * sxy + wy * sx + wx;
                            out[o] = dv[0];
                            out[vol + o] = dv[1];
                            out[2 * vol + o] = dv[2];
                        }
            }
    }
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Affine fit: q_normalize_points_3D (q_warp_affine_tps.cpp:493-570) and
// q_affine_compute_affinmatrix_3D (649-731).  The 3N x 13 DLT matrix is
// decomposed by a one-sided Jacobi SVD (newmat's SVD is Golub-Kahan; the
// null vector is the same up to sign and rounding, and the sign cancels in
// h /= h13).
namespace {

void normalize_points(const float* p, int n, std::vector<double>& out, double T[16]) {
    double cx = 0, cy = 0, cz = 0;
    for (int i = 0; i < n; i++) { cx += (double)p[3 * i]; cy += (double)p[3 * i + 1]; cz += (double)p[3 * i + 2]; }
    cx /= n; cy /= n; cz /= n;
    out.resize((size_t)3 * n);
    for (int i = 0; i < n; i++) {
        out[3 * i] = (double)p[3 * i] - cx;
        out[3 * i + 1] = (double)p[3 * i + 1] - cy;
        out[3 * i + 2] = (double)p[3 * i + 2] - cz;
    }
    double avg = 0;
    for (int i = 0; i < n; i++)
        avg += std::sqrt(out[3 * i] * out[3 * i] + out[3 * i + 1] * out[3 * i + 1] + out[3 * i + 2] * out[3 * i + 2]);
    avg /= n;
    const double s = 1.0 / avg;
    for (int i = 0; i < 3 * n; i++) out[i] *= s;
    for (int i = 0; i < 16; i++) T[i] = 0;
    T[0] = s; T[3] = -s * cx;
    T[5] = s; T[7] = -s * cy;
    T[10] = s; T[11] = -s * cz;
    T[15] = 1;
}

// One-sided Jacobi SVD of the m x 13 matrix A (row-major, destroyed);
// returns V (13 x 13, row-major) and the column norms (singular values).
void jacobi_svd13(std::vector<double>& A, int m, double V[169], double sv[13]) {
    const int n = 13;
    for (int i = 0; i < n * n; i++) V[i] = 0;
    for (int i = 0; i < n; i++) V[i * n + i] = 1;
    for (int sweep = 0; sweep < 80; sweep++) {
        double off = 0;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double a = 0, b = 0, g = 0;
                for (int r = 0; r < m; r++) {
                    const double x = A[(size_t)r * n + p], y = A[(size_t)r * n + q];
                    a += x * x; b += y * y; g += x * y;
                }
                if (g == 0.0) continue;
                const double rel = std::fabs(g) / std::sqrt(a * b);
                off = std::max(off, rel);
                if (rel < 1e-15) continue;
                const double zeta = (b - a) / (2 * g);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1 + zeta * zeta));
                const double cs = 1 / std::sqrt(1 + t * t), sn = cs * t;
                for (int r = 0; r < m; r++) {
                    const double x = A[(size_t)r * n + p], y = A[(size_t)r * n + q];
                    A[(size_t)r * n + p] = cs * x - sn * y;
                    A[(size_t)r * n + q] = sn * x + cs * y;
                }
                for (int r = 0; r < n; r++) {
                    const double x = V[r * n + p], y = V[r * n + q];
                    V[r * n + p] = cs * x - sn * y;
                    V[r * n + q] = sn * x + cs * y;
                }
            }
        if (off < 1e-15) break;
    }
    for (int j = 0; j < n; j++) {
        double s = 0;
        for (int r = 0; r < m; r++) s += A[(size_t)r * n + j] * A[(size_t)r * n + j];
        sv[j] = std::sqrt(s);
    }
}

void inv4_scale_translate(const double T[16], double Ti[16]) {
    // T = [s 0 0 tx; 0 s 0 ty; 0 0 s tz; 0 0 0 1]
    for (int i = 0; i < 16; i++) Ti[i] = 0;
    for (int a = 0; a < 3; a++) {
        Ti[a * 4 + a] = 1.0 / T[a * 4 + a];
        Ti[a * 4 + 3] = -T[a * 4 + 3] / T[a * 4 + a];
    }
    Ti[15] = 1;
}

void mul4(const double* A, const double* B, double* C) {
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double s = 0;
            for (int k = 0; k < 4; k++) s += A[i * 4 + k] * B[k * 4 + j];
            C[i * 4 + j] = s;
        }
}

}  // namespace

extern "C" {

// vec_A -> vec_B affine, B = M A (q_warp_affine_tps.cpp:649).  At the plugin
// call site vec_A are the target markers and vec_B the subject markers
// (q_warp_affine_tps.cpp:1340-1351), so M maps target space to subject space.
// M: 4 x 4 row-major.  Returns 0, -1 bad args, -2 degenerate.
int or_affine_fit(const float* A, const float* B, int n, double* M) {
    if (!A || !B || !M || n < 4) return -1;
    std::vector<double> an, bn;
    double TA[16], TB[16];
    normalize_points(A, n, an, TA);
    normalize_points(B, n, bn, TB);
    const int m = 3 * n;
    std::vector<double> D((size_t)std::max(m, 13) * 13, 0.0);
    for (int i = 0; i < n; i++) {
        for (int r = 0; r < 3; r++) {
            double* row = &D[(size_t)(3 * i + r) * 13];
            row[4 * r + 0] = an[3 * i];
            row[4 * r + 1] = an[3 * i + 1];
            row[4 * r + 2] = an[3 * i + 2];
            row[4 * r + 3] = 1.0;
            row[12] = -bn[3 * i + r];
        }
    }
    double V[169], sv[13];
    jacobi_svd13(D, std::max(m, 13), V, sv);
    int jmin = 0;
    for (int j = 1; j < 13; j++)
        if (sv[j] < sv[jmin]) jmin = j;
    double h[13];
    for (int r = 0; r < 13; r++) h[r] = V[r * 13 + jmin];
    if (h[12] == 0.0) return -2;
    for (int r = 0; r < 13; r++) h[r] /= h[12];
    double Mn[16] = {h[0], h[1], h[2], h[3], h[4], h[5], h[6], h[7], h[8], h[9], h[10], h[11], 0, 0, 0, 1};
    double TBi[16], tmp[16];
    inv4_scale_translate(TB, TBi);
    mul4(TBi, Mn, tmp);
    mul4(tmp, TA, M);
    return 0;
}

}  // extern "C"

namespace {

// CPU loop of q_imagewarp_affine_image (q_warp_affine_tps.cpp:1373-1437):
// cur = M*[x y z 1] (1382-1388); strict out-of-bounds test (1399-1405);
// trilinear with floor/ceil weights (1410-1431); the result is truncated,
// with no +0.5 (1433).
template <class T>
void affine_resample(const T* src, const i64* ssz, const double* M, T* dst, const i64* dsz,
                     i64 z0, i64 z1) {
    const i64 sx = ssz[0], sxy = ssz[0] * ssz[1], svol = sxy * ssz[2];
    const i64 dx = dsz[0], dxy = dsz[0] * dsz[1], dvol = dxy * dsz[2];
#pragma omp parallel for schedule(dynamic)
    for (i64 z = z0; z < z1; z++)
        for (i64 y = 0; y < dsz[1]; y++)
            for (i64 x = 0; x < dsz[0]; x++) {
                double p[3];
                for (int r = 0; r < 3; r++)
                    p[r] = M[r * 4 + 0] * (double)x + M[r * 4 + 1] * (double)y + M[r * 4 + 2] * (double)z +
                           M[r * 4 + 3] * 1.0;
                const i64 o = z * dxy + y * dx + x;
                if (p[0] < 0 || p[0] > ssz[0] - 1 || p[1] < 0 || p[1] > ssz[1] - 1 || p[2] < 0 ||
                    p[2] > ssz[2] - 1) {
                    for (i64 ch = 0; ch < dsz[3]; ch++) dst[ch * dvol + o] = 0;
                    continue;
                }
                const i64 xs = (i64)std::floor(p[0]), xb = (i64)std::ceil(p[0]);
                const i64 ys = (i64)std::floor(p[1]), yb = (i64)std::ceil(p[1]);
                const i64 zs = (i64)std::floor(p[2]), zb = (i64)std::ceil(p[2]);
                const double lw = 1.0 - (p[0] - xs), rw = 1.0 - lw;
                const double tw = 1.0 - (p[1] - ys), bw = 1.0 - tw;
                const double uw = 1.0 - (p[2] - zs), dw = 1.0 - uw;
                for (i64 ch = 0; ch < dsz[3]; ch++) {
                    const T* s = src + ch * svol;
                    const double hi = tw * (lw * px(s, sx, sxy, xs, ys, zs) + rw * px(s, sx, sxy, xb, ys, zs)) +
                                      bw * (lw * px(s, sx, sxy, xs, yb, zs) + rw * px(s, sx, sxy, xb, yb, zs));
                    const double lo = tw * (lw * px(s, sx, sxy, xs, ys, zb) + rw * px(s, sx, sxy, xb, ys, zb)) +
                                      bw * (lw * px(s, sx, sxy, xs, yb, zb) + rw * px(s, sx, sxy, xb, yb, zb));
                    dst[ch * dvol + o] = (T)(uw * hi + dw * lo);
                }
            }
}

}  // namespace

extern "C" {

// M: 4 x 4 row-major target->subject matrix (only rows 0-2 are used).
// ssz/dsz = {x, y, z, c}; the channel count of dst must equal src's.
// z range [z0, z1) of output slices (z1 < 0: to the end).
int or_affine_resample(const void* src, const i64* ssz, int dtype, const double* M, void* dst,
                       const i64* dsz, i64 z0, i64 z1, int nthreads) {
    if (!src || !ssz || !M || !dst || !dsz || dsz[3] != ssz[3]) return -1;
    set_threads(nthreads);
    if (z1 < 0 || z1 > dsz[2]) z1 = dsz[2];
    if (dtype == 1)
        affine_resample((const uint8_t*)src, ssz, M, (uint8_t*)dst, dsz, z0, z1);
    else if (dtype == 2)
        affine_resample((const uint16_t*)src, ssz, M, (uint16_t*)dst, dsz, z0, z1);
    else
        return -1;
    return 0;
}

int or_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
