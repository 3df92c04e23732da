// host_fit.cpp -- host-side transform construction (affine DLT fit and the
// README's rigid / translation / scaling modes).  Small dense FP64 math; the
// per-voxel work runs on the GPU (k_affine.hip).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/terawarp.h"

namespace {

// Hartley normalisation (q_normalize_points_3D, q_warp_affine_tps.cpp:493-570):
// centre on the centroid, scale the mean distance to the origin to 1.
// T is the 4x4 normalising matrix (row-major): p_n = T p.
void hartley(const float* p, int n, std::vector<double>& out, double T[16]) {
    double c[3] = {0, 0, 0};
    for (int i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) c[a] += (double)p[3 * i + a];
    for (int a = 0; a < 3; a++) c[a] /= n;
    out.assign((size_t)3 * n, 0.0);
    double mean = 0;
    for (int i = 0; i < n; i++) {
        double r2 = 0;
        for (int a = 0; a < 3; a++) {
            out[3 * i + a] = (double)p[3 * i + a] - c[a];
            r2 += out[3 * i + a] * out[3 * i + a];
        }
        mean += std::sqrt(r2);
    }
    mean /= n;
    const double s = 1.0 / mean;
    for (double& v : out) v *= s;
    std::memset(T, 0, 16 * sizeof(double));
    for (int a = 0; a < 3; a++) {
        T[a * 4 + a] = s;
        T[a * 4 + 3] = -s * c[a];
    }
    T[15] = 1.0;
}

// Householder QR of the m x 13 row-major matrix A (m >= 13); returns R (13x13).
void householder_r(std::vector<double>& A, int m, double R[169]) {
    const int n = 13;
    for (int k = 0; k < n; k++) {
        double nrm = 0;
        for (int i = k; i < m; i++) nrm += A[(size_t)i * n + k] * A[(size_t)i * n + k];
        nrm = std::sqrt(nrm);
        if (nrm == 0.0) continue;
        const double alpha = A[(size_t)k * n + k] > 0 ? -nrm : nrm;
        std::vector<double> v(m - k);
        for (int i = k; i < m; i++) v[i - k] = A[(size_t)i * n + k];
        v[0] -= alpha;
        double vv = 0;
        for (double x : v) vv += x * x;
        if (vv == 0.0) continue;
        for (int j = k; j < n; j++) {
            double dot = 0;
            for (int i = k; i < m; i++) dot += v[i - k] * A[(size_t)i * n + j];
            const double f = 2.0 * dot / vv;
            for (int i = k; i < m; i++) A[(size_t)i * n + j] -= f * v[i - k];
        }
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) R[i * n + j] = j >= i ? A[(size_t)i * n + j] : 0.0;
}

// Right singular vector of the smallest singular value of the 13x13 R by
// cyclic one-sided Jacobi (orthogonalise the columns of R, accumulate V).
void smallest_right_sv(double R[169], double h[13]) {
    const int n = 13;
    double V[169];
    for (int i = 0; i < n * n; i++) V[i] = 0;
    for (int i = 0; i < n; i++) V[i * n + i] = 1;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double a = 0, b = 0, g = 0;
                for (int r = 0; r < n; r++) {
                    a += R[r * n + p] * R[r * n + p];
                    b += R[r * n + q] * R[r * n + q];
                    g += R[r * n + p] * R[r * n + q];
                }
                if (g == 0.0 || a == 0.0 || b == 0.0) continue;
                const double rel = std::fabs(g) / std::sqrt(a * b);
                off = std::max(off, rel);
                if (rel < 1e-16) continue;
                const double zeta = (b - a) / (2 * g);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1 + zeta * zeta));
                const double cs = 1 / std::sqrt(1 + t * t), sn = cs * t;
                for (int r = 0; r < n; r++) {
                    const double x = R[r * n + p], y = R[r * n + q];
                    R[r * n + p] = cs * x - sn * y;
                    R[r * n + q] = sn * x + cs * y;
                    const double vx = V[r * n + p], vy = V[r * n + q];
                    V[r * n + p] = cs * vx - sn * vy;
                    V[r * n + q] = sn * vx + cs * vy;
                }
            }
        if (off < 1e-16) break;
    }
    int jmin = 0;
    double smin = INFINITY;
    for (int j = 0; j < n; j++) {
        double s = 0;
        for (int r = 0; r < n; r++) s += R[r * n + j] * R[r * n + j];
        if (s < smin) { smin = s; jmin = j; }
    }
    for (int r = 0; r < n; r++) h[r] = V[r * n + jmin];
}

void mat4_mul(const double* A, const double* B, double* C) {
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double s = 0;
            for (int k = 0; k < 4; k++) s += A[i * 4 + k] * B[k * 4 + j];
            C[i * 4 + j] = s;
        }
}

// sin/cos of an angle in degrees, exact at multiples of 90 degrees.
void sincos_deg(double deg, double* s, double* c) {
    const double q = deg / 90.0;
    if (q == std::floor(q)) {
        const long k = ((long)q % 4 + 4) % 4;
        static const double S[4] = {0, 1, 0, -1}, C[4] = {1, 0, -1, 0};
        *s = S[k];
        *c = C[k];
        return;
    }
    const double r = deg * (M_PI / 180.0);
    *s = std::sin(r);
    *c = std::cos(r);
}

}  // namespace

extern "C" int tw_affine_fit(const float* pts_a, const float* pts_b, int n, double* M) {
    if (!pts_a || !pts_b || !M || n < 4) return TW_E_ARG;
    for (int i = 0; i < 3 * n; i++)
        if (!std::isfinite(pts_a[i]) || !std::isfinite(pts_b[i])) return TW_E_ARG;
    std::vector<double> an, bn;
    double TA[16], TB[16];
    hartley(pts_a, n, an, TA);
    hartley(pts_b, n, bn, TB);
    if (!std::isfinite(TA[0]) || !std::isfinite(TB[0])) return TW_E_SINGULAR;
    // Rows [x y z 1 0...0 -x'] etc. (q_warp_affine_tps.cpp:680-700), padded
    // with zero rows to at least 13.
    const int m = std::max(3 * n, 13);
    std::vector<double> D((size_t)m * 13, 0.0);
    for (int i = 0; i < n; i++)
        for (int r = 0; r < 3; r++) {
            double* row = &D[(size_t)(3 * i + r) * 13];
            for (int a = 0; a < 3; a++) row[4 * r + a] = an[3 * i + a];
            row[4 * r + 3] = 1.0;
            row[12] = -bn[3 * i + r];
        }
    double R[169], h[13];
    householder_r(D, m, R);
    smallest_right_sv(R, h);
    if (h[12] == 0.0 || !std::isfinite(h[12])) return TW_E_SINGULAR;
    for (int r = 0; r < 13; r++) h[r] /= h[12];
    const double Mn[16] = {h[0], h[1], h[2], h[3], h[4], h[5], h[6], h[7], h[8], h[9], h[10], h[11], 0, 0, 0, 1};
    // M = TB^-1 Mn TA (denormalise, 728).
    double TBi[16] = {0};
    for (int a = 0; a < 3; a++) {
        TBi[a * 4 + a] = 1.0 / TB[a * 4 + a];
        TBi[a * 4 + 3] = -TB[a * 4 + 3] / TB[a * 4 + a];
    }
    TBi[15] = 1.0;
    double tmp[16];
    mat4_mul(TBi, Mn, tmp);
    mat4_mul(tmp, TA, M);
    for (int i = 0; i < 16; i++)
        if (!std::isfinite(M[i])) return TW_E_SINGULAR;
    return TW_OK;
}

extern "C" int tw_mode_matrix(int mode, const double* para, int64_t nx, int64_t ny, int64_t nz, double* M) {
    if (!para || !M || nx <= 0 || ny <= 0 || nz <= 0) return TW_E_ARG;
    for (int a = 0; a < 3; a++)
        if (!std::isfinite(para[a])) return TW_E_ARG;
    std::memset(M, 0, 16 * sizeof(double));
    M[15] = 1.0;
    if (mode == 1) {  // translation: output(p) = subject(p - t)
        for (int a = 0; a < 3; a++) {
            M[a * 4 + a] = 1.0;
            M[a * 4 + 3] = -para[a];
        }
        return TW_OK;
    }
    if (mode == 2) {  // scaling about the origin: output(p) = subject(p / s)
        for (int a = 0; a < 3; a++) {
            if (!(para[a] > 0)) return TW_E_ARG;
            M[a * 4 + a] = 1.0 / para[a];
        }
        return TW_OK;
    }
    if (mode != 0) return TW_E_ARG;
    // Rotation R = Rz Ry Rx about the centre; backward map p = R^T (q - c) + c.
    double sx, cx, sy, cy, sz, cz;
    sincos_deg(para[0], &sx, &cx);
    sincos_deg(para[1], &sy, &cy);
    sincos_deg(para[2], &sz, &cz);
    const double Rx[16] = {1, 0, 0, 0, 0, cx, -sx, 0, 0, sx, cx, 0, 0, 0, 0, 1};
    const double Ry[16] = {cy, 0, sy, 0, 0, 1, 0, 0, -sy, 0, cy, 0, 0, 0, 0, 1};
    const double Rz[16] = {cz, -sz, 0, 0, sz, cz, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    double t[16], R[16];
    mat4_mul(Ry, Rx, t);
    mat4_mul(Rz, t, R);
    const double c[3] = {(nx - 1) / 2.0, (ny - 1) / 2.0, (nz - 1) / 2.0};
    for (int i = 0; i < 3; i++) {
        double off = c[i];
        for (int j = 0; j < 3; j++) {
            M[i * 4 + j] = R[j * 4 + i];  // R^T
            off -= R[j * 4 + i] * c[j];
        }
        M[i * 4 + 3] = off;
    }
    return TW_OK;
}
