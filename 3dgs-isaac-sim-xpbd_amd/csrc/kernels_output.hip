// kernels_output.hip — frame outputs written into caller-owned device buffers
// (the positions and deformed shapes a renderer or scene writer consumes each
// frame), with the two scalars such a consumer needs from them: the centre of
// the positions and the largest coordinate change since the last frame it
// wrote.  The reference's runtime copies positions and shapes to the host and
// derives both there (runtime_loop.py:387-402, 938-946); here nothing but the
// four scalars leaves the device.
#include "kernels.h"

namespace xpbd {

constexpr int kOutThreads = 256;

// Sum over the 64 lanes of a wave (butterfly: every lane ends with the total,
// added in a fixed order).
__device__ __forceinline__ double wave_sum_d(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

__device__ __forceinline__ float wave_max_f(float v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmaxf(v, __shfl_xor(v, m, 64));
    return v;
}

// One thread per caller-order node c < n: copy the position to dPos, compare
// with dPrev, and leave per-block partials (sum x, y, z in double, largest
// |change|) in part[4 * block ..].
__global__ void __launch_bounds__(kOutThreads) frame_positions_kernel(int n, const float3* __restrict__ pos,
                                                                     float* __restrict__ dPos,
                                                                     float* __restrict__ dPrev,
                                                                     double* __restrict__ part)
{
    __shared__ double red[kOutThreads / 64][4];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    float mv = 0.0f;
    if (c < n) {
        const float3 p = pos[c];
        sx = p.x; sy = p.y; sz = p.z;
        if (dPos) {
            dPos[3 * c] = p.x; dPos[3 * c + 1] = p.y; dPos[3 * c + 2] = p.z;
        }
        if (dPrev)
            mv = fmaxf(fabsf(p.x - dPrev[3 * c]), fmaxf(fabsf(p.y - dPrev[3 * c + 1]), fabsf(p.z - dPrev[3 * c + 2])));
    }
    sx = wave_sum_d(sx); sy = wave_sum_d(sy); sz = wave_sum_d(sz);
    mv = wave_max_f(mv);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[w][0] = sx; red[w][1] = sy; red[w][2] = sz; red[w][3] = (double)mv;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double v = red[0][threadIdx.x];
        for (int k = 1; k < kOutThreads / 64; ++k)
            v = threadIdx.x < 3 ? v + red[k][threadIdx.x] : fmax(v, red[k][threadIdx.x]);
        part[4 * (size_t)blockIdx.x + threadIdx.x] = v;
    }
}

// One block: the block partials combined in a fixed order into out[0..3].
__global__ void __launch_bounds__(64) frame_total_kernel(int blocks, const double* __restrict__ part,
                                                         double* __restrict__ out)
{
    double sx = 0.0, sy = 0.0, sz = 0.0, mv = 0.0;
    for (int k = threadIdx.x; k < blocks; k += 64) {
        sx += part[4 * (size_t)k];
        sy += part[4 * (size_t)k + 1];
        sz += part[4 * (size_t)k + 2];
        mv = fmax(mv, part[4 * (size_t)k + 3]);
    }
    sx = wave_sum_d(sx); sy = wave_sum_d(sy); sz = wave_sum_d(sz);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) mv = fmax(mv, __shfl_xor(mv, m, 64));
    if (threadIdx.x == 0) {
        out[0] = sx; out[1] = sy; out[2] = sz; out[3] = mv;
    }
}

// dPrev takes the new positions when the largest change (out[3]) exceeds
// moveEps: the decision a scene writer makes before writing a frame, taken on
// the device.
__global__ void __launch_bounds__(256) frame_prev_kernel(int n, const float3* __restrict__ pos,
                                                        float* __restrict__ dPrev, const double* __restrict__ out,
                                                        float moveEps)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n || !(out[3] > (double)moveEps)) return;
    const float3 p = pos[c];
    dPrev[3 * c] = p.x; dPrev[3 * c + 1] = p.y; dPrev[3 * c + 2] = p.z;
}

// Deformed shapes of device nodes [nodeBegin, nodeBegin + n) to caller order;
// quaternions (w, x, y, z) as they are, or as (x, y, z, w).
__global__ void __launch_bounds__(256) shapes_to_caller_kernel(int nodeBegin, int n, int callerBase,
                                                              const int* __restrict__ callerPos,
                                                              const float* __restrict__ scale,
                                                              const float* __restrict__ quat, float* __restrict__ dScale,
                                                              float* __restrict__ dQuat, int xyzw)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n) return;
    const int c = callerPos[nodeBegin + d] - callerBase;
    if (dScale) {
        dScale[3 * c] = scale[3 * d]; dScale[3 * c + 1] = scale[3 * d + 1]; dScale[3 * c + 2] = scale[3 * d + 2];
    }
    if (dQuat) {
        const float w = quat[4 * d], x = quat[4 * d + 1], y = quat[4 * d + 2], z = quat[4 * d + 3];
        if (xyzw) {
            dQuat[4 * c] = x; dQuat[4 * c + 1] = y; dQuat[4 * c + 2] = z; dQuat[4 * c + 3] = w;
        } else {
            dQuat[4 * c] = w; dQuat[4 * c + 1] = x; dQuat[4 * c + 2] = y; dQuat[4 * c + 3] = z;
        }
    }
}

int frame_output_blocks(int n) { return (n + kOutThreads - 1) / kOutThreads; }

void launch_frame_positions(hipStream_t s, int n, const float3* pos, float* dPos, float* dPrev, float moveEps,
                            double* part, double* out)
{
    const int blocks = frame_output_blocks(n);
    if (blocks > 0) frame_positions_kernel<<<blocks, kOutThreads, 0, s>>>(n, pos, dPos, dPrev, part);
    frame_total_kernel<<<1, 64, 0, s>>>(blocks, part, out);
    if (dPrev && blocks > 0) frame_prev_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, pos, dPrev, out, moveEps);
}

void launch_shapes_to_caller(hipStream_t s, int nodeBegin, int n, int callerBase, const int* callerPos,
                             const float* scale, const float* quat, float* dScale, float* dQuat, bool xyzw)
{
    if (n > 0)
        shapes_to_caller_kernel<<<(n + 255) / 256, 256, 0, s>>>(nodeBegin, n, callerBase, callerPos, scale, quat,
                                                                dScale, dQuat, xyzw ? 1 : 0);
}

}  // namespace xpbd
