// polar.h — rotation factor of a 3x3 matrix in fp64, the robust path of the
// reference's shape matching (forward.cu:5991-6086).
//
// With A = U diag(s) V^T the rotation is R = U V^T, made proper:
//   1. B = A^T A in double, B = V diag(s^2) V^T by the largest-pair Jacobi
//      schedule of jacobi3.h, stopped once every off-diagonal entry is at
//      most 1e-12 sqrt(|b_pp b_qq|); singular values ascending;
//   2. U = A V diag(1/s), 1/s taken as 0 for s <= 1e-12 s_max;
//   3. a nearly flat A (s_min < 0.08 s_max) gets U's first column rebuilt as
//      the cross product of the other two, which fails when they are
//      degenerate;
//   4. R = U V^T; when det R < 0, U's first column is negated once.
// Returns false, leaving Rout untouched, when step 3 fails or R is not finite.
#pragma once
#include "jacobi3.h"
#include "mat3.h"

namespace xpbd {

struct SettledRelative {
    __host__ __device__ bool operator()(const double (&b)[3][3], double) const
    {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int p = pair_p(k), q = pair_q(k);
            if (fabs(b[p][q]) > 1e-12 * sqrt(fabs(b[p][p] * b[q][q])) + 1e-300) return false;
        }
        return true;
    }
};

// R = U V^T (rows of U against rows of V)
__device__ __forceinline__ void frames_product(const double (&u)[3][3], const double (&v)[3][3], double (&R)[3][3])
{
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) R[r][c] = u[r][0] * v[c][0] + u[r][1] * v[c][1] + u[r][2] * v[c][2];
}

__device__ __forceinline__ double det_rows(const double (&R)[3][3])
{
    return R[0][0] * (R[1][1] * R[2][2] - R[1][2] * R[2][1]) - R[0][1] * (R[1][0] * R[2][2] - R[1][2] * R[2][0]) +
           R[0][2] * (R[1][0] * R[2][1] - R[1][1] * R[2][0]);
}

__device__ __forceinline__ bool polar_rotation_f64(const M3& A, M3& Rout)
{
    double a[3][3], b[3][3], v[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) a[r][c] = (double)A.m[c][r];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) b[r][c] = a[0][r] * a[0][c] + a[1][r] * a[1][c] + a[2][r] * a[2][c];
    set_identity(v);
    jacobi_largest_pair<double, true>(b, v, 1e-300, SettledRelative{});
    double sv[3] = {b[0][0], b[1][1], b[2][2]};
    sort_ascending<double, true>(sv, v);
#pragma unroll
    for (int i = 0; i < 3; ++i) sv[i] = sqrt(sv[i] > 0.0 ? sv[i] : 0.0);
    const double top = sv[2] > 1e-300 ? sv[2] : 1e-300;
    double u[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double inv = sv[c] > 1e-12 * top ? 1.0 / sv[c] : 0.0;
#pragma unroll
        for (int r = 0; r < 3; ++r) u[r][c] = (a[r][0] * v[0][c] + a[r][1] * v[1][c] + a[r][2] * v[2][c]) * inv;
    }
    if (sv[0] < 0.08 * top) {
        const double n0 = u[1][1] * u[2][2] - u[2][1] * u[1][2];
        const double n1 = u[2][1] * u[0][2] - u[0][1] * u[2][2];
        const double n2 = u[0][1] * u[1][2] - u[1][1] * u[0][2];
        const double len = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
        if (!(len > 1e-7)) return false;
        u[0][0] = n0 / len;
        u[1][0] = n1 / len;
        u[2][0] = n2 / len;
    }
    double R[3][3];
    frames_product(u, v, R);
    if (det_rows(R) < 0.0) {
#pragma unroll
        for (int r = 0; r < 3; ++r) u[r][0] = -u[r][0];
        frames_product(u, v, R);
    }
    bool finite = true;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) finite = finite && isfinite(R[r][c]);
    if (!finite) return false;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) Rout.m[c][r] = (float)R[r][c];
    return true;
}

}  // namespace xpbd
