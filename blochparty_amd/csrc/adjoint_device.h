// Reverse mode of one segment (the forward step of bloch_device.h), one isochromat per lane.
//
// A segment maps the magnetization affinely, m' = A m + b. With lambda = dL/dm' (spoil and ADC source already
// applied), the reverse gives dL/dm = A^T lambda and the parameter gradients of the segment from the state m before
// it. The calculus (the reference's reverse sweep, shaders/adjoint_common.metal:31-62, 211-315, states the same):
//
//   free:  m'xy = E2 R(phi) mxy, m'z = pd + (mz - pd) E1, phi = -2 pi dt bz, bz = r.g + db0 + b0_offset
//   RF:    m1 = relax(dt/2) m, m2 = R(w) m1, m' = relax(dt/2) m2, w = -2 pi dt (b1 rf_x, b1 rf_y, bz)
//   R(w) m = c m + s1 (w x m) + s2 w (w.m), c = cos|w|, s1 = sin|w|/|w|, s2 = (1 - cos|w|)/|w|^2;
//   for a = dL/d(R m): dL/dm = R^T a, dL/dw = k w + s1 (m x a) + s2 ((w.a) m + (w.m) a),
//   k = -s1 (a.m) + f1 (a.(w x m)) + f2 (a.w)(w.m), f1 = (c - s1)/|w|^2, f2 = (s1 - 2 s2)/|w|^2
//   (f1, f2 by their Taylor series below |w|^2 = 1, where the quotients cancel).
#pragma once

#include "bloch_device.h"

namespace hb {

// per-spin gradient accumulators: d/d(px, py, pz, db0, pd, r1, r2, b1)
struct Acc {
    float v[8];
};
enum { A_PX = 0, A_PY, A_PZ, A_DB0, A_PD, A_R1, A_R2, A_B1 };
// per-segment gradients: d/d(dt, rf_x, rf_y, g_x, g_y, g_z, b0_offset)
enum { G_DT = 0, G_RX, G_RY, G_GX, G_GY, G_GZ, G_B0, N_SEG_GRADS };

struct RevCache {
    float c, sn, d1, d2;  // free: cos, sin, expm1(-dt r1), expm1(-dt r2)
    float h1, h2;         // RF: expm1(-dt/2 r1), expm1(-dt/2 r2)
};

__device__ __forceinline__ void rot_factors(float p2, float& c, float& s1, float& s2, float& f1, float& f2) {
    rot_ab(p2, s1, s2);
    c = 1.0f - s2 * p2;
    if (p2 < 1.0f) {
        f1 = -1.0f / 3.0f + p2 * (1.0f / 30.0f + p2 * (-1.0f / 840.0f + p2 * (1.0f / 45360.0f + p2 * (-1.0f / 3991680.0f))));
        f2 = -1.0f / 12.0f + p2 * (1.0f / 180.0f + p2 * (-1.0f / 6720.0f + p2 * (1.0f / 453600.0f + p2 * (-1.0f / 47900160.0f))));
    } else {
        f1 = (c - s1) / p2;
        f2 = (s1 - 2.0f * s2) / p2;
    }
}

// Reverse through segment q from the state (x0, y0, z0) before it. (lx, ly, lz): dL/d(state after) in, dL/d(state
// before) out. Per-spin gradients into acc; per-segment gradients into sg (when SEG).
template <bool SEG>
__device__ __forceinline__ void reverse_step(const hb_segment& q, int f, const Spin& p, float x0, float y0, float z0,
                                             float& lx, float& ly, float& lz, Acc& acc, float (&sg)[N_SEG_GRADS],
                                             RevCache& rc) {
    const float dt = q.dt;
    const float w = -HB_TWO_PI * dt;
    const float bz = offset(p, q);
    float gb;  // dL/dbz
    if (f & HB_RF) {
        const float h = 0.5f * dt;
        if (!(f & HB_SAME_RF_DT)) {
            rc.h1 = expm1x(-h * p.r1);
            rc.h2 = expm1x(-h * p.r2);
        }
        const float d1 = rc.h1, d2 = rc.h2, e1 = 1.0f + d1, e2 = 1.0f + d2;
        const float wx0 = w * q.rf_x, wy0 = w * q.rf_y;
        const float wx = wx0 * p.b1, wy = wy0 * p.b1, wz = w * bz;
        float c, s1, s2, f1, f2;
        rot_factors(wx * wx + wy * wy + wz * wz, c, s1, s2, f1, f2);
        // forward recomputation: m1 (after the first half relaxation), n = R m1
        const float m1x = x0 + x0 * d2, m1y = y0 + y0 * d2, m1z = z0 + (z0 - p.pd) * d1;
        const float dot = wx * m1x + wy * m1y + wz * m1z;
        const float nx = c * m1x + s1 * (wy * m1z - wz * m1y) + s2 * wx * dot;
        const float ny = c * m1y + s1 * (wz * m1x - wx * m1z) + s2 * wy * dot;
        const float nz = c * m1z + s1 * (wx * m1y - wy * m1x) + s2 * wz * dot;
        // second half relaxation, reversed (lambda + lambda d: one rounding, where a factor 1 + d rounded to float32
        // would bias every substep the same way)
        float gE2 = lx * nx + ly * ny, gE1 = lz * (nz - p.pd), gpd = -lz * d1;
        const float ax = lx + lx * d2, ay = ly + ly * d2, az = lz + lz * d1;
        // rotation, reversed: R^T a = a - s1 (w x a) + s2 w x (w x a), an increment to a (c = 1 - s2 |w|^2 is not
        // formed: rounded near 1 it would bias the norm of lambda at every substep)
        const float wa = wx * ax + wy * ay + wz * az;
        const float cxa = wy * az - wz * ay, cya = wz * ax - wx * az, cza = wx * ay - wy * ax;
        const float exa = wy * cza - wz * cya, eya = wz * cxa - wx * cza, eza = wx * cya - wy * cxa;
        const float bx = ax + (s2 * exa - s1 * cxa);
        const float by = ay + (s2 * eya - s1 * cya);
        const float bzz = az + (s2 * eza - s1 * cza);
        const float am = ax * m1x + ay * m1y + az * m1z;
        const float awm = ax * (wy * m1z - wz * m1y) + ay * (wz * m1x - wx * m1z) + az * (wx * m1y - wy * m1x);
        const float kk = -s1 * am + f1 * awm + f2 * wa * dot;
        const float gwx = kk * wx + s1 * (m1y * az - m1z * ay) + s2 * (wa * m1x + dot * ax);
        const float gwy = kk * wy + s1 * (m1z * ax - m1x * az) + s2 * (wa * m1y + dot * ay);
        const float gwz = kk * wz + s1 * (m1x * ay - m1y * ax) + s2 * (wa * m1z + dot * az);
        // first half relaxation, reversed
        gE2 += bx * x0 + by * y0;
        gE1 += bzz * (z0 - p.pd);
        gpd -= bzz * d1;
        lx = bx + bx * d2;
        ly = by + by * d2;
        lz = bzz + bzz * d1;
        acc.v[A_R2] -= gE2 * h * e2;
        acc.v[A_R1] -= gE1 * h * e1;
        acc.v[A_PD] += gpd;
        acc.v[A_B1] += gwx * wx0 + gwy * wy0;
        gb = gwz * w;
        if (SEG) {
            sg[G_DT] += -HB_TWO_PI * (gwx * p.b1 * q.rf_x + gwy * p.b1 * q.rf_y + gwz * bz) -
                        0.5f * (gE2 * p.r2 * e2 + gE1 * p.r1 * e1);
            sg[G_RX] += gwx * w * p.b1;
            sg[G_RY] += gwy * w * p.b1;
        }
    } else {
        if (!(f & HB_SAME_FREE)) {
            sincospi(-2.0f * free_cycles(p, q), &rc.sn, &rc.c);
            rc.d1 = expm1x(-dt * p.r1);
            rc.d2 = expm1x(-dt * p.r2);
        }
        const float c = rc.c, sn = rc.sn, e2 = 1.0f + rc.d2, e1 = 1.0f + rc.d1;
        const float ux = c * x0 - sn * y0, uy = sn * x0 + c * y0;
        const float gE2 = lx * ux + ly * uy;
        const float gphi = e2 * (ly * ux - lx * uy);
        const float gE1 = lz * (z0 - p.pd);
        const float ax = c * lx + sn * ly, ay = c * ly - sn * lx;
        acc.v[A_PD] -= lz * rc.d1;
        lx = ax + ax * rc.d2;
        ly = ay + ay * rc.d2;
        lz = lz + lz * rc.d1;
        acc.v[A_R2] -= gE2 * dt * e2;
        acc.v[A_R1] -= gE1 * dt * e1;
        gb = gphi * w;
        if (SEG) sg[G_DT] += -HB_TWO_PI * gphi * bz - gE2 * p.r2 * e2 - gE1 * p.r1 * e1;
    }
    // bz = r.g + db0 + b0_offset
    acc.v[A_PX] += gb * q.g_x;
    acc.v[A_PY] += gb * q.g_y;
    acc.v[A_PZ] += gb * q.g_z;
    acc.v[A_DB0] += gb;
    if (SEG) {
        sg[G_GX] += gb * p.px;
        sg[G_GY] += gb * p.py;
        sg[G_GZ] += gb * p.pz;
        sg[G_B0] += gb;
    }
}

}  // namespace hb
