// The full-variant forward kernel: a rate bound pool (magnetization transfer) and / or a scanner's concomitant fields.
// Replaces the reference's V = 1 variant of bloch_chunk_mma (src/blochparty/engines/gpu/shaders/common.metal:113-208:
// lineshape, the exchange step, step_mt; :563-601 the field model; engines/core.py:115-121, 190-217).
//
// One lane per isochromat, workgroups of 256 (4 waves of 64), every lane walking the same plain segment records. Per
// spin the free pool's (Mx, My, Mz) and, with a bound pool, its Mzb; Mz and Mzb carry two-sum low parts. Without a
// bound pool the longitudinal step is plain T1 recovery.
// - Concomitant field f_c [Hz] of a segment with gradient g: leading order c . (gx^2, gy^2, gz^2, gx gy, gx gz,
//   gy gz), or exact |B| - f0 - lin = t^2 / (sqrt(u^2 + t^2) + u) (t^2 = (ax . g)^2 + (ay . g)^2, u = f0 + lin, lin the
//   rest of the offset); it adds to an RF segment's offset (and so to the bound pool's saturation offset) and, times
//   the segment's length, to a free segment's float64 phase.
// - RF segment: Strang split. Half a segment of transverse decay and of longitudinal exchange, with the bound pool
//   saturated at W = (2 pi b1 |rf|)^2 G(f_rf - bz - dw); the water's rotation (only the precession about z for RF
//   that only the bound pool feels); the other half.
// - Free segment: exact precession (float64 phase, as the plain kernels') and transverse decay; the unsaturated
//   exchange over the segment.
// - The exchange over t is exact: z_ss + exp(A t) (z - z_ss) for the 2 x 2 generator A of (Mz, Mzb), in increment
//   form (exp(A t) - I)(z - z_ss) = (e^{mt} cosh(qt) - 1)(z - z_ss) + e^{mt} sinh(qt)/q (A - m I)(z - z_ss), with
//   m +- q the eigenvalues and both factors from expm1 of m +- q, so that short steps keep their low bits.
// - Factors that depend only on the segment's length and the spin (the free segment's decay and unsaturated exchange,
//   the RF segment's half-step decay) are kept across segments the records mark as having the previous one's dt
//   (HB_SAME_FREE_DT, HB_SAME_RF_DT): the cache restarts with every dispatch, as the hints do.
// - ADC: every sample's coil sums over the workgroup (over each wave, then the 4 waves in order in LDS) go to the
//   partial rows that hb_signal_reduce sums in a fixed order: no atomics. Up to 16 coils the sensitivities sit in
//   registers and the wave's 2 C sums are one reduce-scatter (each step halves the values a lane carries: 17
//   shuffles for 8 coils, against 96 for a butterfly per value); more coils take a butterfly per value.
// - The next segment record is loaded while the current one is stepped.
#include "bloch_device.h"

#include "../../include/blochparty_amd_mt.h"

namespace hb {
namespace {

constexpr int kBlockM = 256;
constexpr int kWavesM = kBlockM / 64;
constexpr int kMaxCoilsM = 128;

// The exchange step over t: the steady state and the two increment factors.
struct XStep {
    float z1, z2;      // steady state -A^-1 c
    float chm1, sh;    // e^{mt} cosh(qt) - 1, e^{mt} sinh(qt) / q
    float h, a12, a21; // A - m I = [[h, a12], [a21, -h]]
};

__device__ __forceinline__ XStep xstep(float t, float pd, float r1, float f, float k, float r1b, float w) {
    const float kf = k * f;
    const float a11 = -(r1 + kf), a12 = k, a21 = kf, a22 = -(r1b + k + w);
    const float c1 = r1 * pd, c2 = r1b * f * pd;
    const float det = a11 * a22 - a12 * a21;
    XStep x;
    x.z1 = (a12 * c2 - a22 * c1) / det;
    x.z2 = (a21 * c1 - a11 * c2) / det;
    const float m = 0.5f * (a11 + a22), h = 0.5f * (a11 - a22);
    const float q = sqrtf(fmaf(h, h, a12 * a21));
    x.chm1 = 0.5f * (expm1x((m + q) * t) + expm1x((m - q) * t));
    const float lo = __expf((m - q) * t);
    if (2.0f * q * t < 1.0f)
        x.sh = q > 0.0f ? lo * 0.5f * expm1x(2.0f * q * t) / q : t * __expf(m * t);
    else
        x.sh = 0.5f * (__expf((m + q) * t) - lo) / q;
    x.h = h, x.a12 = a12, x.a21 = a21;
    return x;
}

__device__ __forceinline__ void exchange(const XStep& x, State& m, float& mb, float& lb) {
    const float d1 = m.mz - x.z1, d2 = mb - x.z2;
    two_sum(m.mz, m.lz, x.chm1 * d1 + x.sh * (x.h * d1 + x.a12 * d2));
    two_sum(mb, lb, x.chm1 * d2 + x.sh * (x.a21 * d1 - x.h * d2));
}

// G(2 pi delta) [s]: the Lorentzian, or T2b F(x) from the table of ln F (held at the cut, 0 past the table's end).
__device__ __forceinline__ float lineshape_g(const hb_mt& mt, float delta, float t2b) {
    if (mt.lineshape == HB_LINE_LORENTZIAN) {
        const float x = HB_TWO_PI * delta * t2b;
        return t2b / fmaf(x, x, 1.0f);
    }
    const float x = HB_TWO_PI * fmaxf(fabsf(delta), mt.cutoff) * t2b;
    const float pos = (logf(x) - mt.lx0) / mt.dlx;
    if (!(pos <= (float)(mt.n_table - 1))) return 0.0f;
    const float pc = fmaxf(pos, 0.0f);
    const int i = min((int)pc, mt.n_table - 2);
    const float a = pc - (float)i;
    return t2b * expf(mt.table[i] * (1.0f - a) + mt.table[i + 1] * a);
}

// Rotation of the state by the rotation vector w (right-handed, angle |w|), as rf_step's: m + a (w x m) + b w x (w x m).
__device__ __forceinline__ void rotate_state(State& m, float wx, float wy, float wz) {
    float a, b;
    rot_ab(wx * wx + wy * wy + wz * wz, a, b);
    const float cx = wy * m.mz - wz * m.my, cy = wz * m.mx - wx * m.mz, cz = wx * m.my - wy * m.mx;
    const float ex = wy * cz - wz * cy, ey = wz * cx - wx * cz, ez = wx * cy - wy * cx;
    m.mx += a * cx + b * ex;
    m.my += a * cy + b * ey;
    two_sum(m.mz, m.lz, a * cz + b * ez);
}

// The spin's concomitant field in segment q, lin the rest of its offset there [Hz]; cc the spin's coefficients.
__device__ __forceinline__ float conc_field(int mode, const float (&cc)[7], const hb_segment& q, float lin) {
    const float gx = q.g_x, gy = q.g_y, gz = q.g_z;
    if (mode == HB_CONC_LEADING)
        return cc[0] * gx * gx + cc[1] * gy * gy + cc[2] * gz * gz + cc[3] * gx * gy + cc[4] * gx * gz +
               cc[5] * gy * gz;
    const float tx = cc[0] * gx + cc[1] * gy + cc[2] * gz, ty = cc[3] * gx + cc[4] * gy + cc[5] * gz;
    const float t2 = tx * tx + ty * ty, u = cc[6] + lin;
    const float den = sqrtf(u * u + t2) + u;  // > 0 for any real field (u = f0 + lin > 0)
    return den > 0.0f ? t2 / den : 0.0f;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v >> 1); }

// The wave's sums of V values (V a power of two <= 64), in a fixed order: over lane bits 5, 4, ... each step keeps
// half the values (a lane with the bit set the upper half) and adds the partner's copy of them, then the remaining
// low bits are summed by a butterfly. Value i ends on the lanes with lane >> (6 - log2 V) == i, in v[0].
// (Each step is its own instantiation, so that every index into v is a compile-time constant and v stays in
// registers without select chains.)
template <int V, int ST>
__device__ __forceinline__ void scatter_step(float (&v)[V], int lane) {
    if constexpr ((V >> ST) > 1) {
        constexpr int bit = 1 << (5 - ST), half = V >> (ST + 1);
        const bool up = lane & bit;
#pragma unroll
        for (int k = 0; k < half; ++k) {
            const float send = up ? v[k] : v[k + half], keep = up ? v[k + half] : v[k];
            v[k] = keep + __shfl_xor(send, bit, 64);
        }
        scatter_step<V, ST + 1>(v, lane);
    }
}

template <int V>
__device__ __forceinline__ void wave_scatter_sums(float (&v)[V], int lane) {
    scatter_step<V, 0>(v, lane);
#pragma unroll
    for (int b = 5 - ilog2(V); b >= 0; --b) v[0] += __shfl_xor(v[0], 1 << b, 64);
}

// NC: the coils' register width (1, 2, 4, 8, 16; C <= NC, the rest zero), or 0 for any count read per sample.
template <int NC>
__global__ __launch_bounds__(kBlockM) void forward_mt_kernel(hb_case cs, hb_mt mt, int s0, int s1,
                                                             float* __restrict__ partial, int rows, int row0) {
    __shared__ float red[kWavesM][2 * kMaxCoilsM];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = blockIdx.x * kBlockM + tid;
    const int np = cs.n_pad;
    const Spin p = load_spin(cs.spins, cs.pos_lo, np, j);
    const Pos64 r64 = pos64(p);
    const bool pool = mt.pool != nullptr;  // uniform
    const float f = pool ? mt.pool[j] : 0.0f, k = pool ? mt.pool[np + j] : 0.0f;
    const float r1b = pool ? mt.pool[2 * np + j] : 1.0f, t2b = pool ? mt.pool[3 * np + j] : 1.0f;
    const float dw = pool ? mt.pool[4 * np + j] : 0.0f;
    const float b1sq = p.b1 * p.b1;
    State m = load_state(cs.state, np, j);
    float mb = pool ? mt.mb[j] : 0.0f, lb = pool ? mt.mb[np + j] : 0.0f;
    const int cmode = mt.conc_mode;
    float cc[7] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (cmode != HB_CONC_NONE) {
#pragma unroll
        for (int i = 0; i < 7; ++i) cc[i] = mt.conc[(size_t)i * np + j];
    }
    const int C = cs.n_coils, vals = 16 * cs.n_tiles;
    float sr[NC > 0 ? NC : 1], si[NC > 0 ? NC : 1];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        sr[c] = c < C ? mt.sens[(size_t)(2 * c) * np + j] : 0.0f;
        si[c] = c < C ? mt.sens[(size_t)(2 * c + 1) * np + j] : 0.0f;
    }
    XStep xf;                       // free segments: the unsaturated exchange over dt (with a bound pool)
    float e1f = 0.0f, e2f = 0.0f;   // free segments: expm1 of -dt R1 (without a bound pool), -dt R2
    float e1r = 0.0f, e2r = 0.0f;   // RF segments: expm1 of -dt/2 R1 (without a bound pool), -dt/2 R2
    bool free_ok = false, rf_ok = false;
    hb_segment nq = cs.segs[s0];
    for (int s = s0; s < s1; ++s) {
        const hb_segment q = nq;
        if (s + 1 < s1) nq = cs.segs[s + 1];
        const int fl = q.flags;
        if (fl & HB_RF) {
            float bz = offset(p, q);
            const float fc = cmode != HB_CONC_NONE ? conc_field(cmode, cc, q, bz) : 0.0f;
            bz += fc;
            if (!(rf_ok && (fl & HB_SAME_RF_DT))) {
                e2r = expm1x(-0.5f * q.dt * p.r2);
                if (!pool) e1r = expm1x(-0.5f * q.dt * p.r1);
                rf_ok = true;
            }
            const float e2 = e2r, e1 = e1r;
            XStep x;
            if (pool) {
                const float w1sq = HB_TWO_PI * HB_TWO_PI * b1sq * (q.rf_x * q.rf_x + q.rf_y * q.rf_y);
                const float sat = w1sq * lineshape_g(mt, mt.rf_freq[s] - bz - dw, t2b);
                x = xstep(0.5f * q.dt, p.pd, p.r1, f, k, r1b, sat);
            }
            m.mx += m.mx * e2, m.my += m.my * e2;
            if (pool)
                exchange(x, m, mb, lb);
            else
                two_sum(m.mz, m.lz, (m.mz - p.pd) * e1);
            const float wz =
                p.px * q.w_g_x + p.py * q.w_g_y + p.pz * q.w_g_z + (p.db0 * q.w + q.w_b0) + q.w * fc;
            if (fl & HB_BOUND_ONLY)
                rotate_state(m, 0.0f, 0.0f, wz);
            else
                rotate_state(m, p.b1 * q.w_rf_x, p.b1 * q.w_rf_y, wz);
            m.mx += m.mx * e2, m.my += m.my * e2;
            if (pool)
                exchange(x, m, mb, lb);
            else
                two_sum(m.mz, m.lz, (m.mz - p.pd) * e1);
        } else {
            double cyc = free_cycles64(r64, p, q);
            if (cmode != HB_CONC_NONE) cyc += (double)(conc_field(cmode, cc, q, offset(p, q)) * q.dt);
            cyc -= rint(cyc);
            float sn, cs_;
            sincospi(-2.0f * (float)cyc, &sn, &cs_);
            if (!(free_ok && (fl & HB_SAME_FREE_DT))) {
                e2f = expm1x(-q.dt * p.r2);
                if (pool)
                    xf = xstep(q.dt, p.pd, p.r1, f, k, r1b, 0.0f);
                else
                    e1f = expm1x(-q.dt * p.r1);
                free_ok = true;
            }
            const float nx = cs_ * m.mx - sn * m.my, ny = sn * m.mx + cs_ * m.my;
            m.mx = nx + nx * e2f;
            m.my = ny + ny * e2f;
            if (pool)
                exchange(xf, m, mb, lb);
            else
                two_sum(m.mz, m.lz, (m.mz - p.pd) * e1f);
        }
        if (fl & HB_ADC) {  // the workgroup's coil sums of this sample (uniform branch: every lane takes it)
            const int r = q.pad[0] - row0;
            if constexpr (NC > 0) {
                float v[2 * NC];
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    v[2 * c] = sr[c] * m.mx - si[c] * m.my;
                    v[2 * c + 1] = sr[c] * m.my + si[c] * m.mx;
                }
                wave_scatter_sums<2 * NC>(v, lane);
                constexpr int low = 6 - ilog2(2 * NC);
                if ((lane & ((1 << low) - 1)) == 0) red[wave][lane >> low] = v[0];
            } else {
                for (int c = 0; c < C; ++c) {
                    const float s_r = mt.sens[(size_t)(2 * c) * np + j], s_i = mt.sens[(size_t)(2 * c + 1) * np + j];
                    const float re = wave_sum(s_r * m.mx - s_i * m.my), im = wave_sum(s_r * m.my + s_i * m.mx);
                    if (lane == 0) red[wave][2 * c] = re, red[wave][2 * c + 1] = im;
                }
            }
            __syncthreads();
            if (r >= 0 && r < rows && tid < vals) {
                float v = 0.0f;
                if (tid < 2 * C) {
                    v = red[0][tid];
#pragma unroll
                    for (int w = 1; w < kWavesM; ++w) v += red[w][tid];
                }
                partial[((size_t)blockIdx.x * rows + r) * vals + tid] = v;
            }
            __syncthreads();
        }
        if (fl & HB_SPOIL) m.mx = m.my = 0.0f;
    }
    store_state(cs.state, np, j, m);
    if (pool) mt.mb[j] = mb, mt.mb[np + j] = lb;
}

}  // namespace
}  // namespace hb

extern "C" int hb_forward_mt(const hb_case* cs, const hb_mt* mt, int32_t s0, int32_t s1, float* partial,
                             int32_t rows, int32_t row0, void* stream) {
    using namespace hb;
    if (!cs || !mt || !mt->sens) return hipErrorInvalidValue;
    if (mt->pool && (!mt->mb || !mt->rf_freq)) return hipErrorInvalidValue;
    if (mt->conc_mode < HB_CONC_NONE || mt->conc_mode > HB_CONC_EXACT || (mt->conc_mode && !mt->conc))
        return hipErrorInvalidValue;
    if (cs->n_pad <= 0 || cs->n_pad % kBlockM || cs->n < 0 || cs->n > cs->n_pad) return hipErrorInvalidValue;
    if (cs->n_coils < 1 || cs->n_coils > kMaxCoilsM || cs->n_tiles < 1 || 8 * cs->n_tiles < cs->n_coils)
        return hipErrorInvalidValue;
    if (s0 < 0 || s1 < s0 || s1 > cs->n_segments || rows < 0 || row0 < 0) return hipErrorInvalidValue;
    if (mt->pool && mt->lineshape == HB_LINE_SUPERLORENTZIAN && (!mt->table || mt->n_table < 2 || !(mt->dlx > 0.0f)))
        return hipErrorInvalidValue;
    if (mt->pool && mt->lineshape != HB_LINE_SUPERLORENTZIAN && mt->lineshape != HB_LINE_LORENTZIAN)
        return hipErrorInvalidValue;
    if (s1 == s0) return hipSuccess;
    const dim3 grid(cs->n_pad / kBlockM), block(kBlockM);
    hipStream_t st = (hipStream_t)stream;
    const int C = cs->n_coils;
    if (C <= 1)
        hipLaunchKernelGGL(forward_mt_kernel<1>, grid, block, 0, st, *cs, *mt, s0, s1, partial, rows, row0);
    else if (C <= 2)
        hipLaunchKernelGGL(forward_mt_kernel<2>, grid, block, 0, st, *cs, *mt, s0, s1, partial, rows, row0);
    else if (C <= 4)
        hipLaunchKernelGGL(forward_mt_kernel<4>, grid, block, 0, st, *cs, *mt, s0, s1, partial, rows, row0);
    else if (C <= 8)
        hipLaunchKernelGGL(forward_mt_kernel<8>, grid, block, 0, st, *cs, *mt, s0, s1, partial, rows, row0);
    else if (C <= 16)
        hipLaunchKernelGGL(forward_mt_kernel<16>, grid, block, 0, st, *cs, *mt, s0, s1, partial, rows, row0);
    else
        hipLaunchKernelGGL(forward_mt_kernel<0>, grid, block, 0, st, *cs, *mt, s0, s1, partial, rows, row0);
    return hipGetLastError();
}
