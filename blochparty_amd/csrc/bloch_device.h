// Device-side physics shared by the forward, replay and adjoint kernels: one isochromat per lane, state in
// registers, the segment record read through the scalar cache (every lane of a wave steps the same segment).
//
// Physics follows the reference's core (src/blochparty/engines/core.py:71-150, 268-319): RF segments by Strang
// splitting (relax dt/2, rotate about w = -2 pi dt (b1 rf_x, b1 rf_y, bz), relax dt/2) with the increments added by
// two-sum into a low part, as the reference's GPU step (shaders/common.metal:18-23, 83-103); free segments exactly
// (precession about z, relaxation). The rotation is written in half-angle (Cayley-Klein / spinor) form: with
// phi = |w|, h = phi / 2, the increment is a (w x m) + b w x (w x m), a = sin(phi)/phi = 2 cos(h) sin(h)/phi,
// b = (1 - cos phi)/phi^2 = 2 (sin(h)/phi)^2, from a single sincos of h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/blochparty_amd.h"

#define HB_TWO_PI 6.283185307179586f

namespace hb {

// hi + lo += d without losing the rounding of hi (Knuth's two-sum on d + lo).
__device__ __forceinline__ void two_sum(float& hi, float& lo, float d) {
    d += lo;
    const float t = hi + d;
    const float bp = t - hi;
    lo = (hi - (t - bp)) + (d - bp);
    hi = t;
}

struct Spin {
    float px, py, pz, pd, r1, r2, db0, b1;
    float lx, ly, lz;  // float64 position = p + l
};

__device__ __forceinline__ Spin load_spin(const float* __restrict__ spins, const float* __restrict__ pos_lo,
                                          int n_pad, int j) {
    Spin p;
    p.px = spins[j];
    p.py = spins[n_pad + j];
    p.pz = spins[2 * n_pad + j];
    p.pd = spins[3 * n_pad + j];
    p.r1 = spins[4 * n_pad + j];
    p.r2 = spins[5 * n_pad + j];
    p.db0 = spins[6 * n_pad + j];
    p.b1 = spins[7 * n_pad + j];
    p.lx = pos_lo[j];
    p.ly = pos_lo[n_pad + j];
    p.lz = pos_lo[2 * n_pad + j];
    return p;
}

// Off-resonance [Hz] of a spin in segment q, float32 (RF substeps: phi per step is small).
__device__ __forceinline__ float offset(const Spin& p, const hb_segment& q) {
    return p.px * q.g_x + p.py * q.g_y + p.pz * q.g_z + p.db0 + q.b0;
}

// Phase [cycles] of a free segment in float64 from the segment's float64 gradient moment, reduced to [-1/2, 1/2]:
// cycles = k . r + db0 dt + int b0_offset dt.
__device__ __forceinline__ float free_cycles(const Spin& p, const hb_segment& q) {
    const double rx = (double)p.px + (double)p.lx, ry = (double)p.py + (double)p.ly, rz = (double)p.pz + (double)p.lz;
    double cyc = rx * q.m_x + ry * q.m_y + rz * q.m_z + (double)p.db0 * q.dt64 + q.b0_area;
    cyc -= rint(cyc);
    return (float)cyc;
}

// Factors kept across runs of identical segments.
struct Cache {
    float c, s, e1, e2;   // free: cos, sin of the precession, expm1 of the relaxation over dt
    float h1, h2;         // RF: expm1 of the relaxation over dt / 2
};

// m += a (w x m) + b w x (w x m), each component by two-sum.
__device__ __forceinline__ void rotate_inc(float wx, float wy, float wz, float& mx, float& my, float& mz, float& lx,
                                           float& ly, float& lz) {
    const float phi2 = wx * wx + wy * wy + wz * wz;
    float a, b;
    if (phi2 < 1e-6f) {  // phi < 1e-3: series to phi^4 / 120
        a = 1.0f - phi2 * (1.0f / 6.0f);
        b = 0.5f - phi2 * (1.0f / 24.0f);
    } else {
        const float phi = sqrtf(phi2);
        float sh, ch;
        sincosf(0.5f * phi, &sh, &ch);
        const float sg = sh / phi;
        a = 2.0f * ch * sg;
        b = 2.0f * sg * sg;
    }
    const float cx = wy * mz - wz * my, cy = wz * mx - wx * mz, cz = wx * my - wy * mx;
    const float dx = wy * cz - wz * cy, dy = wz * cx - wx * cz, dz = wx * cy - wy * cx;
    two_sum(mx, lx, a * cx + b * dx);
    two_sum(my, ly, a * cy + b * dy);
    two_sum(mz, lz, a * cz + b * dz);
}

struct State {
    float mx, my, mz, lx, ly, lz;
};

// One segment (not including the spoil).
__device__ __forceinline__ void step(const hb_segment& q, int f, const Spin& p, State& m, Cache& k) {
    if (f & HB_RF) {
        if (!(f & HB_SAME_RF_DT)) {
            k.h1 = expm1f(-0.5f * q.dt * p.r1);
            k.h2 = expm1f(-0.5f * q.dt * p.r2);
        }
        two_sum(m.mx, m.lx, m.mx * k.h2);
        two_sum(m.my, m.ly, m.my * k.h2);
        two_sum(m.mz, m.lz, (m.mz - p.pd) * k.h1);
        const float w = -HB_TWO_PI * q.dt;
        const float wb = w * p.b1;
        rotate_inc(wb * q.rf_x, wb * q.rf_y, w * offset(p, q), m.mx, m.my, m.mz, m.lx, m.ly, m.lz);
        two_sum(m.mx, m.lx, m.mx * k.h2);
        two_sum(m.my, m.ly, m.my * k.h2);
        two_sum(m.mz, m.lz, (m.mz - p.pd) * k.h1);
    } else {
        // Mxy is rotated: fold its low parts. Mz keeps its low part: near equilibrium a free segment adds an
        // increment far below one ulp of Mz, and rounding Mz to float32 after every segment biases the steady state
        // (2e-5 of the signal of a 264-TR spoiled GRE), where two-sum keeps it at 5e-7.
        m.mx += m.lx;
        m.my += m.ly;
        m.lx = m.ly = 0.0f;
        if (!(f & HB_SAME_FREE)) {
            sincospif(-2.0f * free_cycles(p, q), &k.s, &k.c);
            k.e1 = expm1f(-q.dt * p.r1);
            k.e2 = expm1f(-q.dt * p.r2);
        }
        const float nx = k.c * m.mx - k.s * m.my, ny = k.s * m.mx + k.c * m.my;
        m.mx = nx + nx * k.e2;
        m.my = ny + ny * k.e2;
        two_sum(m.mz, m.lz, (m.mz - p.pd) * k.e1);
    }
}

}  // namespace hb
