// Device-side physics shared by the forward, replay and adjoint kernels: one isochromat per lane, state in
// registers; every lane of a wave steps the same segment, so control flow never diverges.
//
// Physics follows the reference's core (src/blochparty/engines/core.py:71-150, 268-319): RF segments by Strang
// splitting (relax dt/2, rotate about w = -2 pi dt (b1 rf_x, b1 rf_y, bz), relax dt/2), Mz's increments added by
// two-sum into a low part (as the reference's GPU step keeps low parts, shaders/common.metal:18-23, 83-103); free
// segments exactly (precession about z, relaxation). A rotation by w is applied as the increment
// a (w x m) + b w x (w x m) with a = sin(phi)/phi and b = (1 - cos phi)/phi^2, phi = |w|: even functions of phi, so
// polynomials in phi^2 for the small angles of shaped-pulse substeps, and the half-angle (Cayley-Klein) form
// a = 2 cos(h) sin(h)/phi, b = 2 (sin(h)/phi)^2, h = phi / 2, from one sincos, for larger ones.
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

// sin(pi x) and cos(pi x) for |x| < 2^22: x = r + k / 2 with k = rint(2 x) (r = x - k / 2 is exact, |r| <= 1/4),
// the Taylor series sin(pi r) = r (pi + r^2 P(r^2)) to r^9 and cos(pi r) = 1 + r^2 (-pi^2 / 2 + r^2 Q(r^2)) to
// r^10 (truncation under 2e-9 at |r| = 1/4), then the quadrant k mod 4 by swapping and sign flips. The leading
// coefficients pi and -pi^2 / 2 enter as float32 hi + lo pairs: rounded to float32 alone they would make
// s^2 + c^2 - 1 average +1.6e-8, a bias that a free precession's magnitude picks up at every step. Within 1 ulp of
// the exact values, s^2 + c^2 - 1 unbiased to 5e-10 (checked in float32 against float64); about 26 instructions.
constexpr float kPiHi = 3.141592741e+00f, kPiLo = -8.742277657e-08f;   // pi = hi + lo
constexpr float kC2Hi = -4.934802055e+00f, kC2Lo = -1.451857941e-07f;  // -pi^2 / 2 = hi + lo
constexpr float kS3 = -5.167712688e+00f, kS5 = 2.550163984e+00f;        // -pi^3 / 3!, pi^5 / 5!
constexpr float kS7 = -5.992645025e-01f, kS9 = 8.214588463e-02f;        // -pi^7 / 7!, pi^9 / 9!
constexpr float kC4 = 4.058712006e+00f, kC6 = -1.335262775e+00f;        // pi^4 / 4!, -pi^6 / 6!
constexpr float kC8 = 2.353306264e-01f, kC10 = -2.580689080e-02f;       // pi^8 / 8!, -pi^10 / 10!

__device__ __forceinline__ void sincospi(float x, float* s, float* c) {
    const float k = rintf(2.0f * x);
    const float r = fmaf(-0.5f, k, x);
    const float r2 = r * r;
    float ps = fmaf(kS9, r2, kS7);
    ps = fmaf(ps, r2, kS5);
    ps = fmaf(ps, r2, kS3);
    ps = fmaf(ps, r2, kPiLo);
    const float sv = fmaf(r, kPiHi, r * ps);
    float pc = fmaf(kC10, r2, kC8);
    pc = fmaf(pc, r2, kC6);
    pc = fmaf(pc, r2, kC4);
    pc = fmaf(pc, r2, kC2Lo);
    const float cv = fmaf(r2, kC2Hi, r2 * pc) + 1.0f;
    const int q = (int)k;  // (sin, cos) of pi (r + q / 2): q = 0: (S, C), 1: (C, -S), 2: (-S, -C), 3: (-C, S)
    const bool odd = q & 1;
    const float s0 = odd ? cv : sv, c0 = odd ? sv : cv;
    *s = __int_as_float(__float_as_int(s0) ^ ((q & 2) << 30));
    *c = __int_as_float(__float_as_int(c0) ^ (((q + 1) & 2) << 30));
}

// e^x - 1: for |x| < 1/4 its Taylor series to x^7 (truncation under 4e-10 relative: full relative accuracy for the
// small arguments of short segments, where a rounded 1 + e would lose e's low bits), beyond it the hardware exp2
// minus one (|result| > 0.22: within 7e-8 absolute). Within 3 ulp relative over x <= 0 (checked in float32 against
// float64); about 12 instructions.
__device__ __forceinline__ float expm1x(float x) {
    float p = 1.0f / 5040.0f;
    p = fmaf(p, x, 1.0f / 720.0f);
    p = fmaf(p, x, 1.0f / 120.0f);
    p = fmaf(p, x, 1.0f / 24.0f);
    p = fmaf(p, x, 1.0f / 6.0f);
    p = fmaf(p, x, 0.5f);
    const float small = fmaf(p, x * x, x);
    const float big = __builtin_amdgcn_exp2f(x * 1.4426950408889634f) - 1.0f;
    return fabsf(x) < 0.25f ? small : big;
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

// A spin's float64 position (hi + lo), formed once by a kernel that takes many free-segment phases.
struct Pos64 {
    double x, y, z;
};

__device__ __forceinline__ Pos64 pos64(const Spin& p) {
    return {(double)p.px + (double)p.lx, (double)p.py + (double)p.ly, (double)p.pz + (double)p.lz};
}

// Phase [cycles] of a free segment in float64 from the segment's float64 gradient moment, unreduced:
// cycles = k . r + db0 dt + int b0_offset dt.
__device__ __forceinline__ double free_cycles64(const Pos64& r, const Spin& p, const hb_segment& q) {
    return r.x * q.m_x + r.y * q.m_y + r.z * q.m_z + (double)p.db0 * q.dt64 + q.b0_area;
}

// The same reduced to [-1/2, 1/2].
__device__ __forceinline__ float free_cycles(const Pos64& r, const Spin& p, const hb_segment& q) {
    double cyc = free_cycles64(r, p, q);
    cyc -= rint(cyc);
    return (float)cyc;
}

__device__ __forceinline__ float free_cycles(const Spin& p, const hb_segment& q) {
    return free_cycles(pos64(p), p, q);
}

// Factors kept across runs of identical segments.
struct Cache {
    float c, s, e1, e2;   // free: cos, sin of the precession, expm1 of the relaxation over dt
    float h1, h2;         // RF: expm1 of the relaxation over dt / 2
};

// a = sin(phi)/phi and b = (1 - cos(phi))/phi^2 of a rotation by phi, from p2 = phi^2. Below phi = 1/2 their
// Taylor series to p2^4 (the next terms are under 3e-11): no square root, sincos or division on the hot path.
__device__ __forceinline__ void rot_ab(float p2, float& a, float& b) {
    if (p2 < 0.25f) {
        a = 1.0f + p2 * (-1.0f / 6.0f + p2 * (1.0f / 120.0f + p2 * (-1.0f / 5040.0f + p2 * (1.0f / 362880.0f))));
        b = 0.5f + p2 * (-1.0f / 24.0f + p2 * (1.0f / 720.0f + p2 * (-1.0f / 40320.0f + p2 * (1.0f / 3628800.0f))));
    } else {
        const float phi = sqrtf(p2);
        float sh, ch;
        sincospi(phi * 0.15915494309189535f, &sh, &ch);  // the half angle phi/2 in units of pi: reduced by rint
        const float sg = sh / phi;
        a = 2.0f * ch * sg;
        b = 2.0f * sg * sg;
    }
}

struct State {
    float mx, my, mz, lz;  // lz: the two-sum low part of mz
};

// The state from a (6, n_pad) buffer (mx my mz and their low parts): the transverse low parts are folded in.
__device__ __forceinline__ State load_state(const float* __restrict__ st, int n_pad, int j) {
    State m;
    m.mx = st[j] + st[3 * n_pad + j];
    m.my = st[n_pad + j] + st[4 * n_pad + j];
    m.mz = st[2 * n_pad + j];
    m.lz = st[5 * n_pad + j];
    return m;
}

__device__ __forceinline__ void store_state(float* __restrict__ st, int n_pad, int j, const State& m) {
    st[j] = m.mx, st[n_pad + j] = m.my, st[2 * n_pad + j] = m.mz;
    st[3 * n_pad + j] = 0.0f, st[4 * n_pad + j] = 0.0f, st[5 * n_pad + j] = m.lz;
}

// The jump over a readout run of L = q.pad[3] samples (q its first segment): one exact free step of L dt, phase
// L cyc (cyc the segment's float64 phase in cycles), Mxy zeroed instead when the run ends spoiled or its Mxy is never
// read; Mz relaxes by e1 = expm1(-L dt R1).
__device__ __forceinline__ void run_jump(const hb_segment& q, const Spin& p, double cyc, float e1, State& m) {
    if (q.flags & (HB_XY_DEAD | HB_PULSE_SPOIL)) {
        m.mx = m.my = 0.0f;
    } else {
        const float T = (float)q.pad[3] * q.dt;
        double cl = cyc * (double)q.pad[3];
        cl -= rint(cl);
        float s2, c2;
        sincospi(-2.0f * (float)cl, &s2, &c2);
        const float e2 = expm1x(-T * p.r2);
        const float nx = c2 * m.mx - s2 * m.my, ny = s2 * m.mx + c2 * m.my;
        m.mx = nx + nx * e2;
        m.my = ny + ny * e2;
    }
    two_sum(m.mz, m.lz, (m.mz - p.pd) * e1);
}

// A reused pulse's occurrence with RF phase psi (c, sn its cosine and sine): m += Rz(psi) (D Rz(-psi) m + v), U the
// spin's map (rows n_pad apart: D[i][k] at 3 i + k, then v), read per occurrence so that no registers hold it.
__device__ __forceinline__ void apply_map(const float* __restrict__ U, int np, float c, float sn, State& m) {
    const float ux = c * m.mx + sn * m.my, uy = c * m.my - sn * m.mx, uz = m.mz;
    const float dux = U[0] * ux + U[np] * uy + U[2 * np] * uz + U[9 * np];
    const float duy = U[3 * np] * ux + U[4 * np] * uy + U[5 * np] * uz + U[10 * np];
    const float duz = U[6 * np] * ux + U[7 * np] * uy + U[8 * np] * uz + U[11 * np];
    m.mx += c * dux - sn * duy;
    m.my += sn * dux + c * duy;
    two_sum(m.mz, m.lz, duz);
}

// One segment (not including the spoil).
//
// RF: the three Strang increments (relax dt/2, rotate, relax dt/2) are formed from the running state in float32 and
// summed, then added to the state; Mz's sum goes through two-sum. The rotation vector is (b1 w_rf, w_g . r + w db0 +
// w_b0) from the record's host-formed factors, and phi^2 = b1^2 |w_rf|^2 + wz^2.
// Free: Mxy is rotated by the cached cos / sin and decayed as n + n e2 (e2 = expm1: a float32 factor 1 + e2 would
// lose e2's low bits, and the same factor repeats over a readout, so its rounding would add up).
// Mz keeps a low part through both: near equilibrium a segment adds an increment far below one ulp of Mz, and
// rounding Mz to float32 after every segment biases the steady state (2e-5 of the signal of a 264-TR spoiled GRE,
// against 5e-7 with two-sum); the transverse components are rotated and spoiled, so their rounding does not build up
// (a 5,000-substep CW saturation stays within 1e-6 of float64 without it).
// The fields of a record an RF step reads (dt, flags and the rotation factors): a train loop loads the next one while
// it steps the current one.
struct RfRec {
    float dt, w, w_rf_x, w_rf_y, w_g_x, w_g_y, w_g_z, w_b0, w_rf2;
    int flags;
};

__device__ __forceinline__ RfRec load_rf(const hb_segment* __restrict__ r) {
    const float4 a = *reinterpret_cast<const float4*>(&r->w), b = *reinterpret_cast<const float4*>(&r->w_g_y);
    return {r->dt, a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, r->flags};
}

template <class R>
__device__ __forceinline__ void rf_step(const R& q, int f, const Spin& p, float b1sq, State& m, Cache& k) {
    {
        if (!(f & HB_SAME_RF_DT)) {
            k.h1 = expm1x(-0.5f * q.dt * p.r1);
            k.h2 = expm1x(-0.5f * q.dt * p.r2);
        }
        const float dx1 = m.mx * k.h2, dy1 = m.my * k.h2, dz1 = (m.mz - p.pd) * k.h1;
        const float x1 = m.mx + dx1, y1 = m.my + dy1, z1 = m.mz + dz1;
        const float wx = p.b1 * q.w_rf_x, wy = p.b1 * q.w_rf_y;
        const float wz = p.px * q.w_g_x + p.py * q.w_g_y + p.pz * q.w_g_z + (p.db0 * q.w + q.w_b0);
        float a, b;
        rot_ab(b1sq * q.w_rf2 + wz * wz, a, b);
        const float cx = wy * z1 - wz * y1, cy = wz * x1 - wx * z1, cz = wx * y1 - wy * x1;
        const float ex = wy * cz - wz * cy, ey = wz * cx - wx * cz, ez = wx * cy - wy * cx;
        const float drx = a * cx + b * ex, dry = a * cy + b * ey, drz = a * cz + b * ez;
        const float x2 = x1 + drx, y2 = y1 + dry, z2 = z1 + drz;
        m.mx += dx1 + drx + x2 * k.h2;
        m.my += dy1 + dry + y2 * k.h2;
        two_sum(m.mz, m.lz, dz1 + drz + (z2 - p.pd) * k.h1);
    }
}

__device__ __forceinline__ void free_step(const hb_segment& q, int f, const Spin& p, const Pos64& r, State& m,
                                          Cache& k) {
    {
        if (!(f & HB_SAME_FREE)) {
            sincospi(-2.0f * free_cycles(r, p, q), &k.s, &k.c);
            if (!(f & HB_SAME_FREE_DT)) {
                k.e1 = expm1x(-q.dt * p.r1);
                k.e2 = expm1x(-q.dt * p.r2);
            }
        }
        const float nx = k.c * m.mx - k.s * m.my, ny = k.s * m.mx + k.c * m.my;
        m.mx = nx + nx * k.e2;
        m.my = ny + ny * k.e2;
        two_sum(m.mz, m.lz, (m.mz - p.pd) * k.e1);
    }
}

__device__ __forceinline__ void step(const hb_segment& q, int f, const Spin& p, const Pos64& r, float b1sq, State& m,
                                     Cache& k) {
    if (f & HB_RF)
        rf_step(q, f, p, b1sq, m, k);
    else
        free_step(q, f, p, r, m, k);
}

// The same for kernels that keep no float64 position (it is formed per free segment).
__device__ __forceinline__ void step(const hb_segment& q, int f, const Spin& p, float b1sq, State& m, Cache& k) {
    if (f & HB_RF)
        rf_step(q, f, p, b1sq, m, k);
    else
        free_step(q, f, p, pos64(p), m, k);
}

}  // namespace hb
