// polar.h — rotation of a 3x3 matrix by the fp64 polar decomposition of the
// reference's robust shape-matching path (forward.cu:5991-6086): Jacobi on
// A^T A with a relative stopping test (1e-12), U = A V S^-1, the smallest
// axis rebuilt from the other two when S0 < 0.08 Smax, reflection fix.
#pragma once
#include "mat3.h"

namespace xpbd {

__device__ __forceinline__ bool polar_rotation_f64(const M3& A, M3& Rout)
{
    double a[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) a[r][c] = (double)A.m[c][r];
    double m[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) m[r][c] = a[0][r] * a[0][c] + a[1][r] * a[1][c] + a[2][r] * a[2][c];
    double v[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
    for (int iter = 0; iter < 50; ++iter) {
        bool done = true;
        int p = 0, q = 1;
        double maxOff = -1.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int P = (k == 2) ? 1 : 0, Q = (k == 0) ? 1 : 2;
            const double off = fabs(m[P][Q]);
            if (off > 1e-12 * sqrt(fabs(m[P][P] * m[Q][Q])) + 1e-300) done = false;
            if (off > maxOff) { maxOff = off; p = P; q = Q; }
        }
        if (done) break;
        const double app = m[p][p], aqq = m[q][q], apq = m[p][q];
        const double phi = 0.5 * (aqq - app) / (apq + 1e-300);
        const double t = (phi >= 0.0) ? 1.0 / (phi + sqrt(1.0 + phi * phi)) : 1.0 / (phi - sqrt(1.0 + phi * phi));
        const double c = 1.0 / sqrt(1.0 + t * t);
        const double s = t * c;
        m[p][p] = app - t * apq;
        m[q][q] = aqq + t * apq;
        m[p][q] = m[q][p] = 0.0;
        const int r3 = 3 - p - q;
        const double arp = m[r3][p], arq = m[r3][q];
        m[r3][p] = m[p][r3] = c * arp - s * arq;
        m[r3][q] = m[q][r3] = c * arq + s * arp;
        for (int r = 0; r < 3; ++r) {
            const double vrp = v[r][p], vrq = v[r][q];
            v[r][p] = c * vrp - s * vrq;
            v[r][q] = s * vrp + c * vrq;
        }
    }
    double ev[3] = {m[0][0], m[1][1], m[2][2]};
    for (int i = 0; i < 2; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (ev[i] > ev[j]) {
                const double te = ev[i]; ev[i] = ev[j]; ev[j] = te;
                for (int r = 0; r < 3; ++r) { const double tv = v[r][i]; v[r][i] = v[r][j]; v[r][j] = tv; }
            }
    double S[3], Sinv[3];
    for (int i = 0; i < 3; ++i) S[i] = sqrt(ev[i] > 0.0 ? ev[i] : 0.0);
    const double Smax = (S[2] > 1e-300) ? S[2] : 1e-300;
    for (int i = 0; i < 3; ++i) Sinv[i] = (S[i] > 1e-12 * Smax) ? 1.0 / S[i] : 0.0;
    double u[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) u[r][c] = (a[r][0] * v[0][c] + a[r][1] * v[1][c] + a[r][2] * v[2][c]) * Sinv[c];
    if (S[0] < 0.08 * Smax) {
        const double cx = u[1][1] * u[2][2] - u[2][1] * u[1][2];
        const double cy = u[2][1] * u[0][2] - u[0][1] * u[2][2];
        const double cz = u[0][1] * u[1][2] - u[1][1] * u[0][2];
        const double l = sqrt(cx * cx + cy * cy + cz * cz);
        if (!(l > 1e-7)) return false;
        u[0][0] = cx / l; u[1][0] = cy / l; u[2][0] = cz / l;
    }
    double R[3][3];
    for (int pass = 0; pass < 2; ++pass) {
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) R[r][c] = u[r][0] * v[c][0] + u[r][1] * v[c][1] + u[r][2] * v[c][2];
        const double det = R[0][0] * (R[1][1] * R[2][2] - R[1][2] * R[2][1]) -
                           R[0][1] * (R[1][0] * R[2][2] - R[1][2] * R[2][0]) +
                           R[0][2] * (R[1][0] * R[2][1] - R[1][1] * R[2][0]);
        if (det >= 0.0 || pass == 1) break;
        for (int r = 0; r < 3; ++r) u[r][0] = -u[r][0];
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            if (!isfinite(R[r][c])) return false;
            Rout.m[c][r] = (float)R[r][c];
        }
    return true;
}

}  // namespace xpbd
