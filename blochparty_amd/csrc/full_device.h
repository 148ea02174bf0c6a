// Float64 per-spin physics of the full variant (a rate bound pool, a scanner's concomitant fields, bound-only RF)
// and its reverse mode, for the full variant's adjoint kernel (full_adjoint.hip). Plain C++ on doubles, callable from
// host code too, so that the same functions can be checked on the CPU.
//
// Forward (as forward_mt_kernel's float32 step, mt.hip; the reference's step_mt and field model,
// shaders/common.metal:113-208, 563-601, and engines/core.py:115-121):
//   lin = r . g + db0 + b0_offset, f_c = the concomitant field (leading: c . mono(g); exact: t^2 / (D + u),
//   t = (ax . g, ay . g), u = f0 + lin, D = sqrt(u^2 + |t|^2)), bz = lin + f_c
//   RF: Strang split, half a step of transverse decay E2 = exp(-dt/2 R2) and of longitudinal exchange X(dt/2, W),
//       W = (2 pi b1 |rf|)^2 G(f_rf - bz - dw) (without a bound pool: T1 recovery E1), the rotation R(w),
//       w = -2 pi dt (b1 rf_x, b1 rf_y, bz) ((0, 0, w_z) for RF only the bound pool feels), the other half.
//   free: precession by phi = -2 pi (r . m + db0 dt + b0_area + f_c dt) (m the segment's gradient moment), decay E2
//       = exp(-dt R2), the unsaturated exchange X(dt, 0) (or T1 recovery).
//   X(t, W): z' = z_ss + exp(A t) (z - z_ss), A = [[-(R1 + k f), k], [k f, -(R1b + k + W)]], c = (R1 PD, R1b f PD),
//       z_ss = -A^-1 c; exp(A t) = P I + R (A - m I), m = tr A / 2, Q = (a11 - a22)^2 / 4 + a12 a21 >= 0,
//       P = e^{m t} cosh(sqrt(Q) t), R = e^{m t} sinh(sqrt(Q) t) / sqrt(Q) (even in sqrt(Q): series for small Q t^2).
// Reverse: for lambda = dL/d(state after), dL/d(state before) and the gradients of every parameter the step reads,
// by hand-written reverse mode through the same intermediates (the rotation's as adjoint_device.h states it).
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define HB_HD __host__ __device__
#else
#define HB_HD
#endif

namespace hb {
namespace full {

constexpr double kTwoPi = 6.283185307179586;
constexpr int kRf = 1, kBoundOnly = 262144;  // HB_RF, HB_BOUND_ONLY
constexpr int kLorentzian = 1;               // HB_LINE_LORENTZIAN
constexpr int kLeading = 1, kExact = 2;      // HB_CONC_LEADING, HB_CONC_EXACT

// per-spin gradient slots (the rows of hb_full_adjoint.gs)
enum { G_PX = 0, G_PY, G_PZ, G_DB0, G_PD, G_R1, G_R2, G_B1, G_CC, G_F = G_CC + 7, G_K, G_R1B, G_T2B, G_DW, N_G };
// per-segment gradient slots (hb_seg_reduce's rows)
enum { S_DT = 0, S_RX, S_RY, S_GX, S_GY, S_GZ, S_B0, S_FRF, N_S };

struct Params {
    double px, py, pz, pd, r1, r2, db0, b1;
    double f, k, r1b, t2b, dw;  // the bound pool (unused without one)
    double cc[7];               // concomitant coefficients of the mode
};

struct Seg {
    double dt, rfx, rfy, gx, gy, gz, b0, frf;
    double mx, my, mz, b0a;  // float64 gradient moment and b0_offset area of the segment
    int flags;
};

struct Line {
    const float* table;  // ln F(x) at ln x = lx0 + i dlx (super-Lorentzian)
    double lx0, dlx, cutoff;
    int n, kind;
};

struct Model {
    bool pool;
    int conc;  // HB_CONC_*
    Line line;
};

struct St {
    double x, y, z, b;
};

// ---- lineshape: G(2 pi delta) [s] and its derivatives along delta and T2b ----
HB_HD inline double line_g(const Line& L, double delta, double t2b, double& gd, double& gt) {
    if (L.kind == kLorentzian) {
        const double x = kTwoPi * delta * t2b, den = 1.0 + x * x, G = t2b / den;
        gd = -2.0 * x * G / den * kTwoPi * t2b;
        gt = (1.0 - x * x) / (den * den);
        return G;
    }
    const double ad = fabs(delta);
    const bool held = ad < L.cutoff;  // G is held at the cut
    const double x = kTwoPi * (held ? L.cutoff : ad) * t2b;
    const double pos = (log(x) - L.lx0) / L.dlx;
    if (!(pos <= (double)(L.n - 1))) {
        gd = gt = 0.0;
        return 0.0;
    }
    const double pc = pos > 0.0 ? pos : 0.0;
    int i = (int)pc;
    if (i > L.n - 2) i = L.n - 2;
    const double a = pc - (double)i;
    const double t0 = L.table[i], t1 = L.table[i + 1];
    const double G = t2b * exp(t0 * (1.0 - a) + t1 * a);
    const double dlnx = pos > 0.0 ? G * (t1 - t0) / L.dlx : 0.0;  // dG / d ln x
    gt = (G + dlnx) / t2b;                                          // d ln x / d T2b = 1 / T2b
    gd = held ? 0.0 : dlnx / delta;                                 // d ln x / d delta = 1 / delta
    return G;
}

// ---- concomitant field f_c of gradient g at offset lin, with df_c/dlin, df_c/dg and df_c/dc ----
struct Conc {
    double fc, dlin, dg[3], dc[7];
};

HB_HD inline Conc conc_field(int mode, const double (&c)[7], double gx, double gy, double gz, double lin) {
    Conc o = {0.0, 0.0, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0}};
    if (mode == kLeading) {
        o.dc[0] = gx * gx, o.dc[1] = gy * gy, o.dc[2] = gz * gz;
        o.dc[3] = gx * gy, o.dc[4] = gx * gz, o.dc[5] = gy * gz;
        for (int i = 0; i < 6; ++i) o.fc += c[i] * o.dc[i];
        o.dg[0] = 2.0 * c[0] * gx + c[3] * gy + c[4] * gz;
        o.dg[1] = 2.0 * c[1] * gy + c[3] * gx + c[5] * gz;
        o.dg[2] = 2.0 * c[2] * gz + c[4] * gx + c[5] * gy;
    } else if (mode == kExact) {
        const double tx = c[0] * gx + c[1] * gy + c[2] * gz, ty = c[3] * gx + c[4] * gy + c[5] * gz;
        const double t2 = tx * tx + ty * ty, u = c[6] + lin;
        const double D = sqrt(u * u + t2), den = D + u;
        if (den > 0.0 && D > 0.0) {
            o.fc = t2 / den;
            const double ax = tx / D, ay = ty / D, du = -o.fc / D;
            o.dc[0] = ax * gx, o.dc[1] = ax * gy, o.dc[2] = ax * gz;
            o.dc[3] = ay * gx, o.dc[4] = ay * gy, o.dc[5] = ay * gz, o.dc[6] = du;
            o.dlin = du;
            o.dg[0] = ax * c[0] + ay * c[3];
            o.dg[1] = ax * c[1] + ay * c[4];
            o.dg[2] = ax * c[2] + ay * c[5];
        }
    }
    return o;
}

// ---- the exchange step X(t, W) of (Mz, Mzb) ----
struct Xch {
    double a11, a12, a21, a22, c1, c2, inv, z1, z2, m, hh, Q, P, R, dRdQ, t;
};

HB_HD inline Xch xch_setup(double t, double pd, double r1, double f, double k, double r1b, double w) {
    Xch X;
    X.t = t;
    X.a11 = -(r1 + k * f), X.a12 = k, X.a21 = k * f, X.a22 = -(r1b + k + w);
    X.c1 = r1 * pd, X.c2 = r1b * f * pd;
    X.inv = 1.0 / (X.a11 * X.a22 - X.a12 * X.a21);  // det = R1 R1b + ... > 0 for R1 > 0
    X.z1 = (X.a12 * X.c2 - X.a22 * X.c1) * X.inv;
    X.z2 = (X.a21 * X.c1 - X.a11 * X.c2) * X.inv;
    X.m = 0.5 * (X.a11 + X.a22), X.hh = 0.5 * (X.a11 - X.a22);
    X.Q = X.hh * X.hh + X.a12 * X.a21;
    const double x = X.Q * t * t;
    if (x < 0.01) {  // cosh, sinh / q and d(sinh / q)/dQ by their series in x = Q t^2 (truncation < 3e-14)
        const double E = exp(X.m * t);
        const double C = 1.0 + x * (1.0 / 2 + x * (1.0 / 24 + x * (1.0 / 720 + x * (1.0 / 40320 + x / 3628800))));
        const double S = t * (1.0 + x * (1.0 / 6 + x * (1.0 / 120 + x * (1.0 / 5040 + x * (1.0 / 362880)))));
        const double dS = t * t * t *
                          (1.0 / 6 + x * (1.0 / 60 + x * (1.0 / 1680 + x * (1.0 / 90720 + x * (1.0 / 7983360)))));
        X.P = E * C, X.R = E * S, X.dRdQ = E * dS;
    } else {  // both eigenvalues m +- q are <= 0: no overflow
        const double q = sqrt(X.Q);
        const double e1 = exp((X.m + q) * t), e2 = exp((X.m - q) * t);
        X.P = 0.5 * (e1 + e2), X.R = 0.5 * (e1 - e2) / q;
        X.dRdQ = (t * X.P - X.R) / (2.0 * X.Q);
    }
    return X;
}

HB_HD inline void xch_apply(const Xch& X, double& z, double& b) {
    const double d1 = z - X.z1, d2 = b - X.z2;
    z = X.z1 + X.P * d1 + X.R * (X.hh * d1 + X.a12 * d2);
    b = X.z2 + X.P * d2 + X.R * (X.a21 * d1 - X.hh * d2);
}

// Reverse of xch_apply (set up from p's pool) from the state (z, b) before it: (lz, lb) = dL/d(after) in,
// dL/d(before) out. The gradients along pd, R1, f, k, R1b, W and t are added to g[0..6].
HB_HD inline void xch_reverse(const Xch& X, const Params& p, double z, double b, double& lz, double& lb,
                              double (&g)[7]) {
    const double d1 = z - X.z1, d2 = b - X.z2;
    const double Lz = lz, Lb = lb;
    const double bP = Lz * d1 + Lb * d2;
    const double bR = Lz * (X.hh * d1 + X.a12 * d2) + Lb * (X.a21 * d1 - X.hh * d2);
    const double bd1 = Lz * (X.P + X.R * X.hh) + Lb * X.R * X.a21;
    const double bd2 = Lz * X.R * X.a12 + Lb * (X.P - X.R * X.hh);
    double bhh = X.R * (Lz * d1 - Lb * d2);
    double b12 = Lz * X.R * d2, b21 = Lb * X.R * d1;
    const double bz1 = Lz - bd1, bz2 = Lb - bd2;
    lz = bd1, lb = bd2;
    const double bm = X.t * (bP * X.P + bR * X.R);
    const double bQ = bP * 0.5 * X.t * X.R + bR * X.dRdQ;
    const double bt = bP * (X.m * X.P + X.Q * X.R) + bR * (X.m * X.R + X.P);
    bhh += 2.0 * X.hh * bQ;
    b12 += X.a21 * bQ;
    b21 += X.a12 * bQ;
    double b11 = 0.5 * (bm + bhh), b22 = 0.5 * (bm - bhh);
    const double bc1 = (-X.a22 * bz1 + X.a21 * bz2) * X.inv;
    const double bc2 = (X.a12 * bz1 - X.a11 * bz2) * X.inv;
    const double bdet = -(X.z1 * bz1 + X.z2 * bz2) * X.inv;
    b12 += X.c2 * bz1 * X.inv;
    b22 -= X.c1 * bz1 * X.inv;
    b21 += X.c1 * bz2 * X.inv;
    b11 -= X.c2 * bz2 * X.inv;
    b11 += X.a22 * bdet;
    b22 += X.a11 * bdet;
    b12 -= X.a21 * bdet;
    b21 -= X.a12 * bdet;
    // a11 = -(R1 + k f), a12 = k, a21 = k f, a22 = -(R1b + k + W), c1 = R1 PD, c2 = R1b f PD
    g[0] += p.r1 * bc1 + p.r1b * p.f * bc2;                     // pd
    g[1] += -b11 + p.pd * bc1;                                  // R1
    g[2] += p.k * (b21 - b11) + p.r1b * p.pd * bc2;             // f
    g[3] += -p.f * b11 + b12 + p.f * b21 - b22;                 // k
    g[4] += -b22 + p.f * p.pd * bc2;                            // R1b
    g[5] += -b22;                                               // W
    g[6] += bt;                                              // t
}

// ---- rotation by w: R(w) m = m + s1 (w x m) + s2 w x (w x m) ----
struct Rot {
    double s1, s2, f1, f2;
};

HB_HD inline Rot rot_factors(double p2) {
    Rot r;
    if (p2 < 0.1) {  // Taylor series (truncation < 2e-16)
        r.s1 = 1.0 + p2 * (-1.0 / 6 + p2 * (1.0 / 120 + p2 * (-1.0 / 5040 + p2 * (1.0 / 362880 - p2 / 39916800))));
        r.s2 = 0.5 + p2 * (-1.0 / 24 + p2 * (1.0 / 720 + p2 * (-1.0 / 40320 + p2 * (1.0 / 3628800 - p2 / 479001600))));
        r.f1 = -1.0 / 3 + p2 * (1.0 / 30 + p2 * (-1.0 / 840 + p2 * (1.0 / 45360 + p2 * (-1.0 / 3991680 +
                                                                                       p2 / 518918400))));
        r.f2 = -1.0 / 12 + p2 * (1.0 / 180 + p2 * (-1.0 / 6720 + p2 * (1.0 / 453600 + p2 * (-1.0 / 47900160 +
                                                                                         p2 / 7264857600))));
    } else {
        const double phi = sqrt(p2), sh = sin(0.5 * phi);
        const double c = cos(phi);
        r.s1 = sin(phi) / phi;
        r.s2 = 2.0 * sh * sh / p2;
        r.f1 = (c - r.s1) / p2;
        r.f2 = (r.s1 - 2.0 * r.s2) / p2;
    }
    return r;
}

HB_HD inline void rotate(const Rot& r, double wx, double wy, double wz, double& x, double& y, double& z) {
    const double cx = wy * z - wz * y, cy = wz * x - wx * z, cz = wx * y - wy * x;
    const double ex = wy * cz - wz * cy, ey = wz * cx - wx * cz, ez = wx * cy - wy * cx;
    x += r.s1 * cx + r.s2 * ex;
    y += r.s1 * cy + r.s2 * ey;
    z += r.s1 * cz + r.s2 * ez;
}

// Reverse of rotate from m (before): a = dL/d(R m) in, dL/dm out; dL/dw into gw.
HB_HD inline void rotate_reverse(const Rot& r, double wx, double wy, double wz, double mx, double my, double mz,
                                 double& ax, double& ay, double& az, double (&gw)[3]) {
    const double wa = wx * ax + wy * ay + wz * az, wm = wx * mx + wy * my + wz * mz;
    const double am = ax * mx + ay * my + az * mz;
    const double awm = ax * (wy * mz - wz * my) + ay * (wz * mx - wx * mz) + az * (wx * my - wy * mx);
    const double k = -r.s1 * am + r.f1 * awm + r.f2 * wa * wm;
    gw[0] = k * wx + r.s1 * (my * az - mz * ay) + r.s2 * (wa * mx + wm * ax);
    gw[1] = k * wy + r.s1 * (mz * ax - mx * az) + r.s2 * (wa * my + wm * ay);
    gw[2] = k * wz + r.s1 * (mx * ay - my * ax) + r.s2 * (wa * mz + wm * az);
    const double cx = wy * az - wz * ay, cy = wz * ax - wx * az, cz = wx * ay - wy * ax;
    const double ex = wy * cz - wz * cy, ey = wz * cx - wx * cz, ez = wx * cy - wy * cx;
    ax += r.s2 * ex - r.s1 * cx;
    ay += r.s2 * ey - r.s1 * cy;
    az += r.s2 * ez - r.s1 * cz;
}

// ---- one segment: the forward with the intermediates the reverse reads ----
struct Step {
    Conc cc;
    double lin, bz, e2, e1;
    // RF
    double w[3], W, Wk, G, gd, gt;
    Rot rot;
    St m1, m2;  // after the first half, after the rotation
    Xch X;
    // free
    double cs, sn, ux, uy;
    St after;   // after the step (before a spoil)
};

// sin and cos of pi x (device: the library's sincospi; host: through sin, cos)
HB_HD inline void sincospi_d(double x, double* s, double* c) {
#if defined(__HIP_DEVICE_COMPILE__)
    ::sincospi(x, s, c);
#else
    *s = sin(M_PI * x);
    *c = cos(M_PI * x);
#endif
}

HB_HD inline void longitudinal(const Model& md, const Params& p, const Xch& X, double e1, double& z, double& b) {
    if (md.pool)
        xch_apply(X, z, b);
    else
        z = p.pd + (z - p.pd) * e1;
}

// Factors that depend only on a segment's length and the spin, kept while consecutive segments of the kind have the
// same dt: the free segment's decay and unsaturated exchange, the RF segment's half-step decay.
struct Cache {
    double free_dt, rf_dt;  // the lengths the factors belong to (< 0: none yet)
    double fe2, fe1, re2, re1;
    Xch fx;
};

HB_HD inline Cache no_cache() {
    Cache c;
    c.free_dt = c.rf_dt = -1.0;
    c.fe2 = c.fe1 = c.re2 = c.re1 = 0.0;
    return c;
}

HB_HD inline Step step(const Model& md, const Params& p, const Seg& q, const St& m0, Cache& k) {
    Step s;
    s.lin = p.px * q.gx + p.py * q.gy + p.pz * q.gz + p.db0 + q.b0;
    s.cc = conc_field(md.conc, p.cc, q.gx, q.gy, q.gz, s.lin);
    s.bz = s.lin + s.cc.fc;
    St m = m0;
    if (q.flags & kRf) {
        const double h = 0.5 * q.dt;
        if (k.rf_dt != q.dt) {
            k.re2 = exp(-h * p.r2);
            k.re1 = md.pool ? 1.0 : exp(-h * p.r1);
            k.rf_dt = q.dt;
        }
        s.e2 = k.re2, s.e1 = k.re1;
        if (md.pool) {
            s.Wk = kTwoPi * kTwoPi * p.b1 * p.b1 * (q.rfx * q.rfx + q.rfy * q.rfy);
            s.G = line_g(md.line, q.frf - s.bz - p.dw, p.t2b, s.gd, s.gt);
            s.W = s.Wk * s.G;
            s.X = xch_setup(h, p.pd, p.r1, p.f, p.k, p.r1b, s.W);
        }
        m.x *= s.e2, m.y *= s.e2;
        longitudinal(md, p, s.X, s.e1, m.z, m.b);
        s.m1 = m;
        const bool water = !(q.flags & kBoundOnly);
        s.w[0] = water ? -kTwoPi * q.dt * p.b1 * q.rfx : 0.0;
        s.w[1] = water ? -kTwoPi * q.dt * p.b1 * q.rfy : 0.0;
        s.w[2] = -kTwoPi * q.dt * s.bz;
        s.rot = rot_factors(s.w[0] * s.w[0] + s.w[1] * s.w[1] + s.w[2] * s.w[2]);
        rotate(s.rot, s.w[0], s.w[1], s.w[2], m.x, m.y, m.z);
        s.m2 = m;
        m.x *= s.e2, m.y *= s.e2;
        longitudinal(md, p, s.X, s.e1, m.z, m.b);
    } else {
        double cyc = p.px * q.mx + p.py * q.my + p.pz * q.mz + p.db0 * q.dt + q.b0a + s.cc.fc * q.dt;
        cyc -= rint(cyc);
        sincospi_d(-2.0 * cyc, &s.sn, &s.cs);
        if (k.free_dt != q.dt) {
            k.fe2 = exp(-q.dt * p.r2);
            k.fe1 = md.pool ? 1.0 : exp(-q.dt * p.r1);
            if (md.pool) k.fx = xch_setup(q.dt, p.pd, p.r1, p.f, p.k, p.r1b, 0.0);
            k.free_dt = q.dt;
        }
        s.e2 = k.fe2, s.e1 = k.fe1;
        if (md.pool) s.X = k.fx;
        s.ux = s.cs * m.x - s.sn * m.y, s.uy = s.sn * m.x + s.cs * m.y;
        m.x = s.e2 * s.ux, m.y = s.e2 * s.uy;
        longitudinal(md, p, s.X, s.e1, m.z, m.b);
    }
    s.after = m;
    return s;
}

HB_HD inline Step step(const Model& md, const Params& p, const Seg& q, const St& m0) {
    Cache k = no_cache();
    return step(md, p, q, m0, k);
}

// T1 recovery reversed (no bound pool): z' = pd + (z - pd) e1 over t
HB_HD inline void t1_reverse(const Params& p, double e1, double t, double z, double& lz, double (&g)[N_G],
                             double& gdt_t) {
    const double gE1 = lz * (z - p.pd);
    g[G_PD] += lz * (1.0 - e1);
    g[G_R1] -= t * e1 * gE1;
    gdt_t -= p.r1 * e1 * gE1;
    lz *= e1;
}

// The offset's gradient g_bz onto the spin (positions, db0, concomitant coefficients) and the segment (g, b0).
HB_HD inline void offset_reverse(const Params& p, const Seg& q, const Step& s, double g_bz, double (&g)[N_G],
                                 double (&sg)[N_S]) {
    const double g_lin = g_bz * (1.0 + s.cc.dlin);
    for (int i = 0; i < 7; ++i) g[G_CC + i] += g_bz * s.cc.dc[i];
    g[G_PX] += g_lin * q.gx;
    g[G_PY] += g_lin * q.gy;
    g[G_PZ] += g_lin * q.gz;
    g[G_DB0] += g_lin;
    sg[S_GX] += g_bz * s.cc.dg[0] + g_lin * p.px;
    sg[S_GY] += g_bz * s.cc.dg[1] + g_lin * p.py;
    sg[S_GZ] += g_bz * s.cc.dg[2] + g_lin * p.pz;
    sg[S_B0] += g_lin;
}

// Reverse of step s (from the state m0 before it): l = dL/d(after) in (spoil and ADC source applied), dL/d(before)
// out; spin gradients added to g, segment gradients to sg.
HB_HD inline void step_reverse(const Model& md, const Params& p, const Seg& q, const St& m0, const Step& s, St& l,
                               double (&g)[N_G], double (&sg)[N_S]) {
    if (q.flags & kRf) {
        const double h = 0.5 * q.dt;
        double gx[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // exchange: pd, R1, f, k, R1b, W, t
        double gt1 = 0.0;                                     // T1 recovery's d/dt
        // second half
        double gE2 = l.x * s.m2.x + l.y * s.m2.y;
        l.x *= s.e2, l.y *= s.e2;
        if (md.pool)
            xch_reverse(s.X, p, s.m2.z, s.m2.b, l.z, l.b, gx);
        else
            t1_reverse(p, s.e1, h, s.m2.z, l.z, g, gt1);
        // rotation
        double gw[3];
        rotate_reverse(s.rot, s.w[0], s.w[1], s.w[2], s.m1.x, s.m1.y, s.m1.z, l.x, l.y, l.z, gw);
        // first half
        gE2 += l.x * m0.x + l.y * m0.y;
        l.x *= s.e2, l.y *= s.e2;
        if (md.pool)
            xch_reverse(s.X, p, m0.z, m0.b, l.z, l.b, gx);
        else
            t1_reverse(p, s.e1, h, m0.z, l.z, g, gt1);
        g[G_R2] -= h * s.e2 * gE2;
        sg[S_DT] += -0.5 * p.r2 * s.e2 * gE2 + 0.5 * gt1;
        const double w = -kTwoPi * q.dt;
        double g_bz = gw[2] * w;
        sg[S_DT] += -kTwoPi * gw[2] * s.bz;
        if (!(q.flags & kBoundOnly)) {  // the water's nutation (RF only the bound pool feels has none)
            sg[S_DT] += -kTwoPi * p.b1 * (gw[0] * q.rfx + gw[1] * q.rfy);
            g[G_B1] += w * (gw[0] * q.rfx + gw[1] * q.rfy);
            sg[S_RX] += w * p.b1 * gw[0];
            sg[S_RY] += w * p.b1 * gw[1];
        }
        if (md.pool) {
            g[G_PD] += gx[0], g[G_R1] += gx[1], g[G_F] += gx[2], g[G_K] += gx[3], g[G_R1B] += gx[4];
            sg[S_DT] += 0.5 * gx[6];
            // W = Wk G(f_rf - bz - dw; T2b), Wk = (2 pi b1)^2 |rf|^2
            const double gW = gx[5], gG = gW * s.Wk;
            const double k2 = kTwoPi * kTwoPi * 2.0 * gW * s.G;
            g[G_B1] += k2 * p.b1 * (q.rfx * q.rfx + q.rfy * q.rfy);
            sg[S_RX] += k2 * p.b1 * p.b1 * q.rfx;
            sg[S_RY] += k2 * p.b1 * p.b1 * q.rfy;
            const double gdel = gG * s.gd;
            sg[S_FRF] += gdel;
            g_bz -= gdel;
            g[G_DW] -= gdel;
            g[G_T2B] += gG * s.gt;
        }
        offset_reverse(p, q, s, g_bz, g, sg);
    } else {
        double gx[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        double gt1 = 0.0;
        const double gE2 = l.x * s.ux + l.y * s.uy;
        const double lux = s.e2 * l.x, luy = s.e2 * l.y;
        const double gphi = luy * s.ux - lux * s.uy;
        l.x = s.cs * lux + s.sn * luy;
        l.y = s.cs * luy - s.sn * lux;
        if (md.pool)
            xch_reverse(s.X, p, m0.z, m0.b, l.z, l.b, gx);
        else
            t1_reverse(p, s.e1, q.dt, m0.z, l.z, g, gt1);
        g[G_R2] -= q.dt * s.e2 * gE2;
        sg[S_DT] += -p.r2 * s.e2 * gE2 + gt1;
        if (md.pool) {
            g[G_PD] += gx[0], g[G_R1] += gx[1], g[G_F] += gx[2], g[G_K] += gx[3], g[G_R1B] += gx[4];
            sg[S_DT] += gx[6];
        }
        const double gcyc = -kTwoPi * gphi;  // phi = -2 pi cyc, cyc = dt bz
        sg[S_DT] += gcyc * s.bz;
        offset_reverse(p, q, s, gcyc * q.dt, g, sg);
    }
}

}  // namespace full
}  // namespace hb
