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
#include "bloch_device.h"

namespace hb {

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
    const double px = (double)spins[j] + (double)pos_lo[j];
    const double py = (double)spins[n_pad + j] + (double)pos_lo[n_pad + j];
    const double pz = (double)spins[2 * n_pad + j] + (double)pos_lo[2 * n_pad + j];
    const double pd = spins[3 * n_pad + j], r1 = spins[4 * n_pad + j], r2 = spins[5 * n_pad + j];
    const double db0 = spins[6 * n_pad + j], b1 = spins[7 * n_pad + j];
    const int off = index[2 * p], K = index[2 * p + 1];
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
    float* out = maps + (size_t)p * 12 * n_pad + j;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int i = 0; i < 3; ++i)
            out[(3 * i + c) * (size_t)n_pad] = (float)(m[c + 1][i] - m[0][i] - (i == c ? 1.0 : 0.0));
#pragma unroll
    for (int i = 0; i < 3; ++i) out[(9 + i) * (size_t)n_pad] = (float)m[0][i];
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
