// mat3.h — 3x3 float matrix helpers for the constraint kernels.
//
// Column-major storage m[col][row] with the operand order of the matrix
// library the reference solver uses (GLM 0.9.9: determinant in
// func_matrix.inl:253-262, mat3 * vec3 in type_mat3x3.inl:468-474,
// outerProduct accumulation column by column), so cluster statistics computed
// here equal the reference's bit for bit when summed in the same order.
#pragma once
#include <hip/hip_runtime.h>

#include "jacobi3.h"

namespace xpbd {

struct M3 {
    float m[3][3];
};

__host__ __device__ __forceinline__ M3 m3_zero()
{
    M3 a;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) a.m[c][r] = 0.0f;
    return a;
}

// S += v v^T, column by column
__host__ __device__ __forceinline__ void m3_add_outer(M3& S, float x, float y, float z)
{
    const float v[3] = {x, y, z};
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) S.m[c][r] += v[r] * v[c];
}

__host__ __device__ __forceinline__ void m3_scale(M3& S, float s)
{
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) S.m[c][r] *= s;
}

__host__ __device__ __forceinline__ float m3_det(const M3& M)
{
    return +M.m[0][0] * (M.m[1][1] * M.m[2][2] - M.m[2][1] * M.m[1][2])
           - M.m[1][0] * (M.m[0][1] * M.m[2][2] - M.m[2][1] * M.m[0][2])
           + M.m[2][0] * (M.m[0][1] * M.m[1][2] - M.m[1][1] * M.m[0][2]);
}

// Sigma = s * SigmaHat with s = trace/3 (forward.cu:7666-7676)
__host__ __device__ __forceinline__ bool m3_scale_shape_det(const M3& S, float& s, float& detHat)
{
    const float sc = (S.m[0][0] + S.m[1][1] + S.m[2][2]) * (1.0f / 3.0f);
    if (!isfinite(sc) || sc <= 1e-24f) return false;
    M3 h;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) h.m[c][r] = S.m[c][r] / sc;
    const float d = m3_det(h);
    if (!isfinite(d) || d <= 1e-6f) return false;
    s = sc;
    detHat = d;
    return true;
}

// adjugate inverse with an absolute determinant floor (forward.cu:1696-1720)
__host__ __device__ __forceinline__ bool m3_inverse_safe(const M3& M, M3& inv, float epsDet)
{
    const float a00 = M.m[0][0], a01 = M.m[0][1], a02 = M.m[0][2];
    const float a10 = M.m[1][0], a11 = M.m[1][1], a12 = M.m[1][2];
    const float a20 = M.m[2][0], a21 = M.m[2][1], a22 = M.m[2][2];
    const float det = a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20) + a02 * (a10 * a21 - a11 * a20);
    if (fabsf(det) < epsDet) return false;
    const float id = 1.0f / det;
    inv.m[0][0] = (a11 * a22 - a12 * a21) * id;
    inv.m[0][1] = -(a01 * a22 - a02 * a21) * id;
    inv.m[0][2] = (a01 * a12 - a02 * a11) * id;
    inv.m[1][0] = -(a10 * a22 - a12 * a20) * id;
    inv.m[1][1] = (a00 * a22 - a02 * a20) * id;
    inv.m[1][2] = -(a00 * a12 - a02 * a10) * id;
    inv.m[2][0] = (a10 * a21 - a11 * a20) * id;
    inv.m[2][1] = -(a00 * a21 - a01 * a20) * id;
    inv.m[2][2] = (a00 * a11 - a01 * a10) * id;
    return true;
}

__host__ __device__ __forceinline__ float3 m3_mul_vec(const M3& M, float x, float y, float z)
{
    return make_float3(M.m[0][0] * x + M.m[1][0] * y + M.m[2][0] * z,
                       M.m[0][1] * x + M.m[1][1] * y + M.m[2][1] * z,
                       M.m[0][2] * x + M.m[1][2] * y + M.m[2][2] * z);
}

__host__ __device__ __forceinline__ M3 m3_diag(float a, float b, float c)
{
    M3 m = m3_zero();
    m.m[0][0] = a; m.m[1][1] = b; m.m[2][2] = c;
    return m;
}

// A * B with the operand order of GLM's mat3 product (type_mat3x3.inl:486-518)
__host__ __device__ __forceinline__ M3 m3_mul(const M3& A, const M3& B)
{
    M3 R;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) R.m[c][r] = A.m[0][r] * B.m[c][0] + A.m[1][r] * B.m[c][1] + A.m[2][r] * B.m[c][2];
    return R;
}

__host__ __device__ __forceinline__ M3 m3_transpose(const M3& A)
{
    M3 R;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) R.m[c][r] = A.m[r][c];
    return R;
}

__host__ __device__ __forceinline__ bool m3_finite(const M3& A)
{
    bool ok = true;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) ok = ok && isfinite(A.m[c][r]);
    return ok;
}

// Eigen-decomposition of a symmetric matrix in float: largest-pair Jacobi
// (jacobi3.h) until every off-diagonal entry is below 1e-6, at most 50
// rotations, values ascending (the float solver of forward.cu:782-869).
// m3_eig_sym also returns the eigenvectors as the columns of V.
struct SettledBelow {
    float eps;
    __host__ __device__ bool operator()(const float (&)[3][3], float largest) const { return largest < eps; }
};

template <bool kVectors>
__host__ __device__ __forceinline__ void m3_eig_sym_f32(const M3& S, float ev[3], M3* V)
{
    float a[3][3], v[3][3];
    set_identity(v);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) a[r][c] = S.m[c][r];
    jacobi_largest_pair<float, kVectors>(a, v, 1e-20f, SettledBelow{1e-6f});
    float e[3] = {a[0][0], a[1][1], a[2][2]};
    sort_ascending<float, kVectors>(e, v);
    ev[0] = e[0]; ev[1] = e[1]; ev[2] = e[2];
    if (kVectors)
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int r = 0; r < 3; ++r) V->m[c][r] = v[r][c];
}

__host__ __device__ __forceinline__ void m3_eig_sym_values(const M3& S, float ev[3])
{
    m3_eig_sym_f32<false>(S, ev, nullptr);
}

__host__ __device__ __forceinline__ void m3_eig_sym(const M3& S, float ev[3], M3& V)
{
    m3_eig_sym_f32<true>(S, ev, &V);
}

// symmetric storage [00, 11, 22, 10, 20, 21] (forward.cu:7632-7652)
__host__ __device__ __forceinline__ M3 m3_load_sym(float s0, float s1, float s2, float s3, float s4, float s5)
{
    M3 S;
    S.m[0][0] = s0; S.m[1][1] = s1; S.m[2][2] = s2;
    S.m[0][1] = S.m[1][0] = s3;
    S.m[0][2] = S.m[2][0] = s4;
    S.m[1][2] = S.m[2][1] = s5;
    return S;
}

}  // namespace xpbd
