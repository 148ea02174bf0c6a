// Per-spin affine maps of reused RF pulses, in float64 (replaces the reference's pulse_maps,
// src/blochparty/engines/gpu/cuda/pulse.cu.in:3-32).
//
// A pulse of K RF substeps maps every spin's magnetization affinely, m' = m + D m + v (Strang steps are affine: the
// rotations and the relaxation are linear, the recovery to PD is the constant part). Per spin, the substeps are
// composed on four states at once (0, e_x, e_y, e_z) in float64: v = m'(0) and the columns of D are
// m'(e_c) - v - e_c. One lane per spin, one grid row per distinct pulse; all lanes read the same table rows.
// Output maps[(p * 12 + r) * n_pad + spin]: r = 0..8 the rows of D (D[i][k] at 3 i + k), r = 9..11 v, float32.
// With a spin list, only the listed spins are composed, one lane each (the surrogate's fallback and its judge).
// D is formed as a difference so that its float32 rounding is relative to the (small) change, not to m.
//
// Gradients through the maps (the reference's reused pulses are forward only): the adjoint sweep sums, per spin and
// pulse, the seeds lambda m^T and lambda of every occurrence (in the pulse's frame); the parameter gradients are then
// <seeds, d map / d theta>, with the derivatives of the float64 composition by central differences (steps set by
// the caller for a ~1e-4 rad change of the pulse's rotation: truncation ~1e-9, rounding ~1e-11 relative).
#include "bloch_device.h"

namespace hb {

// Pulse rows [off, off + K) of table composed for one spin: D (D[i][k] at 3 i + k) and v in float64.
__device__ __forceinline__ void compose_map(const double* __restrict__ table, int off, int K, double px, double py,
                                           double pz, double pd, double r1, double r2, double db0, double b1,
                                           double (&D)[9], double (&v)[3]) {
    double m[4][3] = {{0.0, 0.0, 0.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
    const double two_pi = 6.283185307179586;
    for (int k = 0; k < K; ++k) {
        const double* row = table + (size_t)(off + k) * 8;
        const double dt = row[0];
        const double e1 = exp(-0.5 * dt * r1), e2 = exp(-0.5 * dt * r2);
        const double w = -two_pi * dt;
        const double wx = w * b1 * row[1], wy = w * b1 * row[2];
        const double wz = w * (px * row[3] + py * row[4] + pz * row[5] + db0 + row[6]);
        const double p2 = wx * wx + wy * wy + wz * wz;
        double a, b;
        if (p2 < 1e-8) {
            a = 1.0 - p2 / 6.0;
            b = 0.5 - p2 / 24.0;
        } else {
            const double phi = sqrt(p2);
            const double sh = sin(0.5 * phi);
            a = sin(phi) / phi;
            b = 2.0 * sh * sh / p2;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            double x = m[i][0] * e2, y = m[i][1] * e2, z = pd + (m[i][2] - pd) * e1;
            const double cx = wy * z - wz * y, cy = wz * x - wx * z, cz = wx * y - wy * x;
            const double ex = wy * cz - wz * cy, ey = wz * cx - wx * cz, ez = wx * cy - wy * cx;
            x += a * cx + b * ex;
            y += a * cy + b * ey;
            z += a * cz + b * ez;
            m[i][0] = x * e2, m[i][1] = y * e2, m[i][2] = pd + (z - pd) * e1;
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int i = 0; i < 3; ++i) D[3 * i + c] = m[c + 1][i] - m[0][i] - (i == c ? 1.0 : 0.0);
#pragma unroll
    for (int i = 0; i < 3; ++i) v[i] = m[0][i];
}

// A spin's parameters in the order of the adjoint's per-spin gradients: px, py, pz, db0, pd, r1, r2, b1.
__device__ __forceinline__ void spin_params(const float* __restrict__ spins, const float* __restrict__ pos_lo,
                                            int n_pad, int j, double (&q)[8]) {
    q[0] = (double)spins[j] + (double)pos_lo[j];
    q[1] = (double)spins[n_pad + j] + (double)pos_lo[n_pad + j];
    q[2] = (double)spins[2 * n_pad + j] + (double)pos_lo[2 * n_pad + j];
    q[3] = spins[6 * n_pad + j];
    q[4] = spins[3 * n_pad + j];
    q[5] = spins[4 * n_pad + j];
    q[6] = spins[5 * n_pad + j];
    q[7] = spins[7 * n_pad + j];
}

__device__ __forceinline__ void compose_q(const double* __restrict__ table, int off, int K, const double (&q)[8],
                                         double (&D)[9], double (&v)[3]) {
    compose_map(table, off, K, q[0], q[1], q[2], q[4], q[5], q[6], q[3], q[7], D, v);
}

__global__ __launch_bounds__(256) void pulse_maps_kernel(const float* __restrict__ spins,
                                                         const float* __restrict__ pos_lo, int n_pad,
                                                         const double* __restrict__ table,
                                                         const int* __restrict__ index,
                                                         const int* __restrict__ list, int n_lanes,
                                                         float* __restrict__ maps) {
    const int lane_id = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane_id >= n_lanes) return;
    const int j = list ? list[lane_id] : lane_id;
    const int p = blockIdx.y;
    double q[8], D[9], v[3];
    spin_params(spins, pos_lo, n_pad, j, q);
    compose_q(table, index[2 * p], index[2 * p + 1], q, D, v);
    float* out = maps + (size_t)p * 12 * n_pad + j;
#pragma unroll
    for (int r = 0; r < 9; ++r) out[r * (size_t)n_pad] = (float)D[r];
#pragma unroll
    for (int i = 0; i < 3; ++i) out[(9 + i) * (size_t)n_pad] = (float)v[i];
}

// gs[a][j] += sum over pulses of <seeds_p, (map_p(q + h e_a) - map_p(q - h e_a)) / 2 h>, parameter a = blockIdx.y
// (a zero step: the map does not depend on the parameter). The pulses are summed in order in one lane.
__global__ __launch_bounds__(256) void pulse_map_grads_kernel(const float* __restrict__ spins,
                                                              const float* __restrict__ pos_lo, int n, int n_pad,
                                                              const double* __restrict__ table,
                                                              const int* __restrict__ index, int n_pulses,
                                                              const double* __restrict__ steps,
                                                              const double* __restrict__ seeds,
                                                              double* __restrict__ gs) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int a = blockIdx.y;
    double q[8];
    spin_params(spins, pos_lo, n_pad, j, q);
    double acc = 0.0;
    for (int p = 0; p < n_pulses; ++p) {
        const double h = steps[p * 8 + a];
        if (h == 0.0) continue;
        double qp[8], qm[8], Dp[9], vp[3], Dm[9], vm[3];
#pragma unroll
        for (int k = 0; k < 8; ++k) qp[k] = q[k] + (k == a ? h : 0.0), qm[k] = q[k] - (k == a ? h : 0.0);
        const int off = index[2 * p], K = index[2 * p + 1];
        compose_q(table, off, K, qp, Dp, vp);
        compose_q(table, off, K, qm, Dm, vm);
        const double* sd = seeds + (size_t)p * 12 * n_pad + j;
        double t = 0.0;
#pragma unroll
        for (int r = 0; r < 9; ++r) t += sd[r * (size_t)n_pad] * (Dp[r] - Dm[r]);
#pragma unroll
        for (int i = 0; i < 3; ++i) t += sd[(9 + i) * (size_t)n_pad] * (vp[i] - vm[i]);
        acc += t / (2.0 * h);
    }
    gs[(size_t)a * n_pad + j] += acc;
}

// maps_d = b1 d map / d b1 by central differences at b1 (1 +- rel).
__global__ __launch_bounds__(256) void pulse_map_dlnb1_kernel(const float* __restrict__ spins,
                                                              const float* __restrict__ pos_lo, int n_pad,
                                                              const double* __restrict__ table,
                                                              const int* __restrict__ index, double rel,
                                                              float* __restrict__ maps_d) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_pad) return;
    const int p = blockIdx.y;
    double q[8], Dp[9], vp[3], Dm[9], vm[3];
    spin_params(spins, pos_lo, n_pad, j, q);
    const double b1 = q[7];
    const int off = index[2 * p], K = index[2 * p + 1];
    q[7] = b1 * (1.0 + rel);
    compose_q(table, off, K, q, Dp, vp);
    q[7] = b1 * (1.0 - rel);
    compose_q(table, off, K, q, Dm, vm);
    float* out = maps_d + (size_t)p * 12 * n_pad + j;
#pragma unroll
    for (int r = 0; r < 9; ++r) out[r * (size_t)n_pad] = (float)((Dp[r] - Dm[r]) / (2.0 * rel));
#pragma unroll
    for (int i = 0; i < 3; ++i) out[(9 + i) * (size_t)n_pad] = (float)((vp[i] - vm[i]) / (2.0 * rel));
}

}  // namespace hb

extern "C" int hb_pulse_maps(const hb_case* cs, const double* table, const int32_t* index, int32_t n_pulses,
                             const int32_t* spin_list, int32_t n_list, float* maps, void* stream) {
    using namespace hb;
    if (!cs || cs->n_pad <= 0 || cs->n_pad % 256 || n_pulses < 0 || n_pulses > 65535) return hipErrorInvalidValue;
    if (spin_list && (n_list < 0 || n_list > cs->n_pad)) return hipErrorInvalidValue;
    const int n_lanes = spin_list ? n_list : cs->n_pad;
    if (n_pulses == 0 || n_lanes == 0) return hipSuccess;
    hipLaunchKernelGGL(pulse_maps_kernel, dim3((n_lanes + 255) / 256, n_pulses), dim3(256), 0, (hipStream_t)stream,
                       cs->spins, cs->pos_lo, cs->n_pad, table, index, spin_list, n_lanes, maps);
    return hipGetLastError();
}

extern "C" int hb_pulse_map_grads(const hb_case* cs, const double* table, const int32_t* index, int32_t n_pulses,
                                  const double* steps, const double* seeds, double* gs, void* stream) {
    using namespace hb;
    if (!cs || cs->n_pad <= 0 || cs->n_pad % 256 || cs->n < 0 || cs->n > cs->n_pad || n_pulses < 0 ||
        (n_pulses > 0 && (!table || !index || !steps || !seeds || !gs)))
        return hipErrorInvalidValue;
    if (n_pulses == 0 || cs->n == 0) return hipSuccess;
    hipLaunchKernelGGL(pulse_map_grads_kernel, dim3((cs->n + 255) / 256, 8), dim3(256), 0, (hipStream_t)stream,
                       cs->spins, cs->pos_lo, cs->n, cs->n_pad, table, index, n_pulses, steps, seeds, gs);
    return hipGetLastError();
}

extern "C" int hb_pulse_map_dlnb1(const hb_case* cs, const double* table, const int32_t* index, int32_t n_pulses,
                                  double rel, float* maps_d, void* stream) {
    using namespace hb;
    if (!cs || cs->n_pad <= 0 || cs->n_pad % 256 || n_pulses < 0 || n_pulses > 65535 || !(rel > 0.0 && rel < 0.1) ||
        (n_pulses > 0 && (!table || !index || !maps_d)))
        return hipErrorInvalidValue;
    if (n_pulses == 0) return hipSuccess;
    hipLaunchKernelGGL(pulse_map_dlnb1_kernel, dim3(cs->n_pad / 256, n_pulses), dim3(256), 0, (hipStream_t)stream,
                       cs->spins, cs->pos_lo, cs->n_pad, table, index, rel, maps_d);
    return hipGetLastError();
}
