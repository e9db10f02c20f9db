// jacobi3.h — eigen-decomposition of a symmetric 3x3 matrix by Jacobi plane
// rotations: the one rotation primitive behind every eigen and polar routine
// of the solver, in float and in double.
//
// The rotation in the (p, q) plane that zeroes a[p][q] (Golub & Van Loan,
// "Matrix Computations", 4th ed., sec. 8.5.2): with
// zeta = (a[q][q] - a[p][p]) / (2 a[p][q]), t = tan(theta) is the root of
// t^2 + 2 zeta t - 1 = 0 of smaller magnitude, c = 1 / sqrt(1 + t^2), s = t c.
// Two schedules drive it:
//   * largest pair first (jacobi_largest_pair): the rotated pair is updated in
//     closed form (a[p][p] - t a[p][q], a[q][q] + t a[p][q], 0) and only the
//     third row and column are turned.  The reference's float eigen solver
//     (forward.cu:782-869) and its fp64 polar decomposition (5991-6086) follow
//     this schedule, each with its own stopping rule;
//   * cyclic sweeps (jacobi_cyclic): pairs (0,1), (0,2), (1,2) in turn, the
//     rotation applied to both sides of the full matrix, as the reference's
//     fp64 eigen solver for deformed covariances (13022-13061).
// Every step performs the floating-point operations of those routines in the
// same order (products rounded before they are added: built without fp
// contraction), so the results agree with them bit for bit.
#pragma once
#include <hip/hip_runtime.h>

namespace xpbd {

__host__ __device__ __forceinline__ float j3_sqrt(float x) { return sqrtf(x); }
__host__ __device__ __forceinline__ double j3_sqrt(double x) { return sqrt(x); }
__host__ __device__ __forceinline__ float j3_abs(float x) { return fabsf(x); }
__host__ __device__ __forceinline__ double j3_abs(double x) { return fabs(x); }

// off-diagonal pair k = 0, 1, 2 of a 3x3 matrix: (0,1), (0,2), (1,2)
__host__ __device__ __forceinline__ int pair_p(int k) { return k >> 1; }
__host__ __device__ __forceinline__ int pair_q(int k) { return k == 0 ? 1 : 2; }

template <class T>
struct PlaneRotation {
    T c, s, t;
};

// Rotation zeroing the (p, q) entry of a symmetric matrix with diagonal app,
// aqq; apq carries whatever guard against a zero divisor the caller adds.
template <class T>
__host__ __device__ __forceinline__ PlaneRotation<T> zeroing_rotation(T app, T aqq, T apq)
{
    const T zeta = T(0.5) * (aqq - app) / apq;
    const T hyp = j3_sqrt(T(1) + zeta * zeta);
    PlaneRotation<T> g;
    g.t = T(1) / (zeta >= T(0) ? zeta + hyp : zeta - hyp);
    g.c = T(1) / j3_sqrt(T(1) + g.t * g.t);
    g.s = g.t * g.c;
    return g;
}

// (x, y) <- (c x - s y, s x + c y)
template <class T>
__host__ __device__ __forceinline__ void turn(const PlaneRotation<T>& g, T& x, T& y)
{
    const T nx = g.c * x - g.s * y;
    const T ny = g.s * x + g.c * y;
    x = nx;
    y = ny;
}

// the off-diagonal pair of largest magnitude (the earlier pair on ties)
template <class T>
__host__ __device__ __forceinline__ int largest_pair(const T (&a)[3][3], T& mag)
{
    const T m1 = j3_abs(a[0][2]), m2 = j3_abs(a[1][2]);
    int k = 0;
    mag = j3_abs(a[0][1]);
    if (m1 > mag) { k = 1; mag = m1; }
    if (m2 > mag) { k = 2; mag = m2; }
    return k;
}

// Largest-pair schedule, at most 50 rotations, ending as soon as
// settled(a, largest |off-diagonal|) holds.  a stays symmetric; v (columns)
// accumulates the rotations when kVectors.  The pivot differs from lane to
// lane, so the step reads and writes the entries through selects on the pair
// index rather than through run-time array indices (which would put the
// matrices in scratch memory).
template <class T, bool kVectors, class Settled>
__host__ __device__ __forceinline__ void jacobi_largest_pair(T (&a)[3][3], T (&v)[3][3], T guard,
                                                            const Settled& settled)
{
    for (int it = 0; it < 50; ++it) {
        T mag;
        const int k = largest_pair(a, mag);
        if (settled(a, mag)) break;
        // pair (p, q) and the third index r: k = 0: (0,1) r 2; k = 1: (0,2) r 1; k = 2: (1,2) r 0
        const T app = k == 2 ? a[1][1] : a[0][0];
        const T aqq = k == 0 ? a[1][1] : a[2][2];
        const T apq = k == 0 ? a[0][1] : (k == 1 ? a[0][2] : a[1][2]);
        T arp = k == 0 ? a[0][2] : a[0][1];
        T arq = k == 2 ? a[0][2] : a[1][2];
        const PlaneRotation<T> g = zeroing_rotation(app, aqq, apq + guard);
        const T npp = app - g.t * apq, nqq = aqq + g.t * apq;
        turn(g, arp, arq);
        a[0][0] = k == 2 ? a[0][0] : npp;
        a[1][1] = k == 0 ? nqq : (k == 2 ? npp : a[1][1]);
        a[2][2] = k == 0 ? a[2][2] : nqq;
        a[0][1] = a[1][0] = k == 0 ? T(0) : arp;
        a[0][2] = a[2][0] = k == 1 ? T(0) : (k == 0 ? arp : arq);
        a[1][2] = a[2][1] = k == 2 ? T(0) : arq;
        if (kVectors)
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                T vp = k == 2 ? v[m][1] : v[m][0];
                T vq = k == 0 ? v[m][1] : v[m][2];
                turn(g, vp, vq);
                v[m][0] = k == 2 ? v[m][0] : vp;
                v[m][1] = k == 0 ? vq : (k == 2 ? vp : v[m][1]);
                v[m][2] = k == 0 ? v[m][2] : vq;
            }
    }
}

// Cyclic schedule: at most `sweeps` sweeps, ending when
// sum |off-diagonal| <= relTol * sum |diagonal| + absTol; a pair whose entry
// is at most absTol is skipped.  v (columns) accumulates the rotations.
template <class T>
__host__ __device__ __forceinline__ void jacobi_cyclic(T (&a)[3][3], T (&v)[3][3], int sweeps, T relTol, T absTol)
{
    for (int sw = 0; sw < sweeps; ++sw) {
        const T off = j3_abs(a[0][1]) + j3_abs(a[0][2]) + j3_abs(a[1][2]);
        const T diag = j3_abs(a[0][0]) + j3_abs(a[1][1]) + j3_abs(a[2][2]);
        if (off <= relTol * diag + absTol) break;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int p = pair_p(k), q = pair_q(k);
            if (j3_abs(a[p][q]) <= absTol) continue;
            const PlaneRotation<T> g = zeroing_rotation(a[p][p], a[q][q], a[p][q]);
#pragma unroll
            for (int m = 0; m < 3; ++m) turn(g, a[m][p], a[m][q]);   // columns p, q
#pragma unroll
            for (int m = 0; m < 3; ++m) turn(g, a[p][m], a[q][m]);   // rows p, q
#pragma unroll
            for (int m = 0; m < 3; ++m) turn(g, v[m][p], v[m][q]);
        }
    }
}

// eigenvalues ascending, eigenvector columns moved with them (pairwise
// exchanges in the order (0,1), (0,2), (1,2))
template <class T, bool kVectors>
__host__ __device__ __forceinline__ void sort_ascending(T (&ev)[3], T (&v)[3][3])
{
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int i = pair_p(k), j = pair_q(k);
        if (ev[i] > ev[j]) {
            const T e = ev[i]; ev[i] = ev[j]; ev[j] = e;
            if (kVectors)
#pragma unroll
                for (int m = 0; m < 3; ++m) { const T x = v[m][i]; v[m][i] = v[m][j]; v[m][j] = x; }
        }
    }
}

template <class T>
__host__ __device__ __forceinline__ void set_identity(T (&v)[3][3])
{
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) v[r][c] = r == c ? T(1) : T(0);
}

}  // namespace xpbd
