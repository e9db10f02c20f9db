// mat3.h — 3x3 float matrix helpers for the constraint kernels.
//
// Column-major storage m[col][row] with the operand order of the matrix
// library the reference solver uses (GLM 0.9.9: determinant in
// func_matrix.inl:253-262, mat3 * vec3 in type_mat3x3.inl:468-474,
// outerProduct accumulation column by column), so cluster statistics computed
// here equal the reference's bit for bit when summed in the same order.
#pragma once
#include <hip/hip_runtime.h>

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

// Symmetric 3x3 eigenvalues by cyclic-max Jacobi rotations, ascending
// (the scheme of forward.cu:782-897: largest off-diagonal first, stop below
// 1e-6, at most 50 rotations).
__host__ __device__ __forceinline__ void m3_eig_sym_values(const M3& S, float ev[3])
{
    float A[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) A[r][c] = S.m[c][r];
    for (int iter = 0; iter < 50; ++iter) {
        int p = 0, q = 1;
        float maxOff = fabsf(A[0][1]);
        if (fabsf(A[0][2]) > maxOff) { p = 0; q = 2; maxOff = fabsf(A[0][2]); }
        if (fabsf(A[1][2]) > maxOff) { p = 1; q = 2; maxOff = fabsf(A[1][2]); }
        if (maxOff < 1e-6f) break;
        const float app = A[p][p], aqq = A[q][q], apq = A[p][q];
        const float phi = 0.5f * (aqq - app) / (apq + 1e-20f);
        const float t = (phi >= 0.0f) ? (1.0f / (phi + sqrtf(1.0f + phi * phi))) : (1.0f / (phi - sqrtf(1.0f + phi * phi)));
        const float c = 1.0f / sqrtf(1.0f + t * t);
        const float s = t * c;
        A[p][p] = app - t * apq;
        A[q][q] = aqq + t * apq;
        A[p][q] = A[q][p] = 0.0f;
        const int r = 3 - p - q;
        const float arp = A[r][p], arq = A[r][q];
        A[r][p] = A[p][r] = c * arp - s * arq;
        A[r][q] = A[q][r] = c * arq + s * arp;
    }
    ev[0] = A[0][0];
    ev[1] = A[1][1];
    ev[2] = A[2][2];
    if (ev[0] > ev[1]) { const float t = ev[0]; ev[0] = ev[1]; ev[1] = t; }
    if (ev[0] > ev[2]) { const float t = ev[0]; ev[0] = ev[2]; ev[2] = t; }
    if (ev[1] > ev[2]) { const float t = ev[1]; ev[1] = ev[2]; ev[2] = t; }
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

// Same Jacobi scheme as m3_eig_sym_values, also returning the eigenvectors as
// columns of V, sorted with the values (ascending).
__host__ __device__ __forceinline__ void m3_eig_sym(const M3& S, float ev[3], M3& Vout)
{
    float A[3][3], V[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            A[r][c] = S.m[c][r];
            V[r][c] = (r == c) ? 1.0f : 0.0f;
        }
    for (int iter = 0; iter < 50; ++iter) {
        int p = 0, q = 1;
        float maxOff = fabsf(A[0][1]);
        if (fabsf(A[0][2]) > maxOff) { p = 0; q = 2; maxOff = fabsf(A[0][2]); }
        if (fabsf(A[1][2]) > maxOff) { p = 1; q = 2; maxOff = fabsf(A[1][2]); }
        if (maxOff < 1e-6f) break;
        const float app = A[p][p], aqq = A[q][q], apq = A[p][q];
        const float phi = 0.5f * (aqq - app) / (apq + 1e-20f);
        const float t = (phi >= 0.0f) ? (1.0f / (phi + sqrtf(1.0f + phi * phi))) : (1.0f / (phi - sqrtf(1.0f + phi * phi)));
        const float c = 1.0f / sqrtf(1.0f + t * t);
        const float s = t * c;
        A[p][p] = app - t * apq;
        A[q][q] = aqq + t * apq;
        A[p][q] = A[q][p] = 0.0f;
        const int r3 = 3 - p - q;
        const float arp = A[r3][p], arq = A[r3][q];
        A[r3][p] = A[p][r3] = c * arp - s * arq;
        A[r3][q] = A[q][r3] = c * arq + s * arp;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const float vrp = V[r][p], vrq = V[r][q];
            V[r][p] = c * vrp - s * vrq;
            V[r][q] = s * vrp + c * vrq;
        }
    }
    ev[0] = A[0][0];
    ev[1] = A[1][1];
    ev[2] = A[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (ev[i] > ev[j]) {
                const float te = ev[i]; ev[i] = ev[j]; ev[j] = te;
#pragma unroll
                for (int r = 0; r < 3; ++r) { const float tv = V[r][i]; V[r][i] = V[r][j]; V[r][j] = tv; }
            }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) Vout.m[c][r] = V[r][c];
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
