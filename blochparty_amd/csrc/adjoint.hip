// Adjoint kernels: the reverse sweep of one checkpoint interval and the in-order sum of the per-segment gradients.
// Replace the reference's bloch_replay + bloch_adjoint_c{NC} (src/blochparty/engines/gpu/cuda/replay.cu.in:3-33,
// adjoint.cu.in:5-121) and seg_reduce (cuda/seg_reduce.cu:3-11).
//
// One launch per interval [s0, s1) of the forward's checkpoints, last interval first; one lane per isochromat:
//   1. replay the interval from its checkpoint, the state before every 8th segment into HBM scratch (r8);
//   2. for every block of 8 segments, last first: replay the block from its r8 state with the 9 states it passes
//      through held in registers (the loop is unrolled, so the indices are static), then sweep it in reverse:
//      unspoil, add the ADC sample's source term sum_c h_c conj(S_c) and the coil gradients h_c conj(Mxy), reverse
//      the step (adjoint_device.h).
// Per-spin gradients are summed in registers and added to float64 accumulators at every segment index that is a
// multiple of 64 (when something was added since the last time); the accumulators (8 per-spin gradients and the 2 NC
// coil gradients) live in LDS for the whole launch, each lane its own slots, read from HBM at the start and written
// back at the end (one read and one write per interval instead of one of each per flush: the same float64 additions
// in the same order); per-segment gradients are summed over each wave (DPP
// + one MFMA per segment), then over the workgroup's 4 waves in LDS in a fixed order, one partial per workgroup and
// segment, and over workgroups in a fixed order by seg_reduce. Nothing depends on the interval length (a multiple of
// 64) or on scheduling: the gradients are the same bit for bit for any checkpoint_every and every run.
// Reused pulses (maps) and NUFFT readout runs are one record each, their other segments identity bodies: the replay
// jumps over them, the reverse skips the 8-blocks inside them (pulse_reverse, nufft_reverse below).
#include "adjoint_device.h"

namespace hb {

constexpr int kBlockA = 256;
constexpr int kWavesA = kBlockA / 64;

struct Stash {
    float x, y, z;
};
struct StashXY {
    float x, y;
};

typedef float f4 __attribute__((ext_vector_type(4)));

template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Per-segment gradients summed over the wave, in a fixed order, without LDS traffic. Within each 16-lane row, a
// reduce-scatter over lane bits 2, 1, 0 (DPP: half-row mirror, then quad permutations; each step keeps half the
// values and adds the partner's copy of them) leaves lane l with value l & 7 summed over its 8-lane group, and a
// row rotation by 8 adds the other group. One MFMA (v_mfma_f32_16x16x4f32, B = those row sums: B[k][n] is lane
// 16 k + n) then adds the four rows, its A operand a one-hot row selector (A[m][k] = 1 for m = i) putting segment i
// of the current 8-block into row i of the accumulator: D[i][n] = the wave's sum of value n & 7 of segment i.
__device__ __forceinline__ void seg_add(f4& d, const float (&sg)[N_SEG_GRADS], int i, int lane) {
    float v[8];
#pragma unroll
    for (int g = 0; g < 8; ++g) v[g] = g < N_SEG_GRADS ? sg[g] : 0.0f;
    {
        const bool up = lane & 4;  // partner: 7 - (lane & 7) in the half row
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float send = up ? v[k] : v[k + 4], keep = up ? v[k + 4] : v[k];
            v[k] = keep + dpp<0x141>(send);
        }
    }
    {
        const bool up = lane & 2;  // partner: lane ^ 2
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const float send = up ? v[k] : v[k + 2], keep = up ? v[k + 2] : v[k];
            v[k] = keep + dpp<0x4E>(send);
        }
    }
    {
        const bool up = lane & 1;  // partner: lane ^ 1
        const float send = up ? v[0] : v[1], keep = up ? v[1] : v[0];
        v[0] = keep + dpp<0xB1>(send);
    }
    const float r = v[0] + dpp<0x128>(v[0]);  // + lane (l + 8) mod 16 of the row
    const float a = (lane & 15) == i ? 1.0f : 0.0f;
    d = __builtin_amdgcn_mfma_f32_16x16x4f32(a, r, d, 0, 0, 0);
}

// The accumulator's rows 0..7 (lanes 0..15 hold rows 0..3, lanes 16..31 rows 4..7, one per register; column
// n = lane & 15) into segbuf[8][8].
__device__ __forceinline__ void seg_store(const f4& d, float (*segbuf)[8], int lane) {
    if (lane < 32 && (lane & 15) < 8) {
#pragma unroll
        for (int r = 0; r < 4; ++r) segbuf[4 * (lane >> 4) + r][lane & 15] = d[r];
    }
}

// r8 after a record at s whose segments [s, e) it covers (a reused pulse's occurrence, a NUFFT run): the reverse sweep
// skips every 8-block wholly inside them, so only the block that starts inside and ends past e needs its state.
__device__ __forceinline__ void r8_after_span(float* __restrict__ r8, int np, int j, int s0, int s1, int s, int e,
                                              const State& m) {
    const int tl = s0 + ((e - 1 - s0) & ~7);  // the last block start before e
    if (tl > s && e < s1 && e - tl < 8) {
        float* r = r8 + (size_t)((tl - s0) >> 3) * 3 * np;
        r[j] = m.mx, r[np + j] = m.my, r[2 * np + j] = m.mz + m.lz;
    }
}

// Effective flags: cache hints honoured only while the cache holds the factors they refer to.
__device__ __forceinline__ int live_flags(int f, bool free_ok, bool rf_ok, int same_free, int same_rf) {
    if (!free_ok) f &= ~(same_free | HB_SAME_FREE_DT);
    if (!rf_ok) f &= ~same_rf;
    return f;
}

// Reverse through a reused pulse's occurrence (head record q): m_after = Rz(psi) ((I + D) Rz(-psi) m + v). With
// (lx, ly, lz) = dL/dm_after in, dL/dm out: lambda <- Rz(psi) (I + D)^T Rz(-psi) lambda. The seeds of the pulse's
// parameter gradients, in its frame (t = Rz(-psi) lambda_after, u = Rz(-psi) m_before): t u^T on D's entries and t
// on v's, added in float64. With SEG: sg[G_RX] = t^T (Dd u + vd) (the map's derivative along the pulse's amplitude,
// maps_d), sg[G_RY] = dL/dpsi = (lambda_after x m_after)_z - (lambda_before x m_before)_z (the frame's turn).
// in_lds: with a single reused pulse, its seeds are held in LDS for the launch (sl, the workgroup's array: slot a of
// lane tid at a * kBlockA + tid); else they are added in HBM. The two are separate branches, so that each pointer
// keeps its address space (LDS and global instructions, not flat ones that wait on both counters).
__device__ __forceinline__ void seeds_add(double* __restrict__ sd, size_t ss, const float (&t3)[3],
                                          const float (&u3)[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int k = 0; k < 3; ++k) sd[(3 * i + k) * ss] += (double)t3[i] * (double)u3[k];
        sd[(9 + i) * ss] += (double)t3[i];
    }
}

template <bool SEG>
__device__ __forceinline__ void pulse_reverse(const hb_segment& q, const hb_case& cs, const hb_adjoint& ad, int j,
                                              const Stash& mb, const StashXY& ma, float& lx, float& ly, float& lz,
                                              float (&sg)[N_SEG_GRADS], double* sl, bool in_lds) {
    const int np = cs.n_pad;
    const float c = q.w_rf_x, sn = q.w_rf_y;
    const size_t base = (size_t)q.pad[1] * 12 * np + j;
    const float* U = cs.maps + base;
    const float tx = c * lx + sn * ly, ty = c * ly - sn * lx, tz = lz;
    const float ux = c * mb.x + sn * mb.y, uy = c * mb.y - sn * mb.x, uz = mb.z;
    const float t3[3] = {tx, ty, tz}, u3[3] = {ux, uy, uz};
    if (in_lds)
        seeds_add(sl + threadIdx.x, kBlockA, t3, u3);
    else
        seeds_add(ad.seeds + base, (size_t)np, t3, u3);
    const float bx = tx + U[0] * tx + U[3 * np] * ty + U[6 * np] * tz;
    const float by = ty + U[np] * tx + U[4 * np] * ty + U[7 * np] * tz;
    const float bz = tz + U[2 * np] * tx + U[5 * np] * ty + U[8 * np] * tz;
    const float nlx = c * bx - sn * by, nly = sn * bx + c * by;
    if (SEG) {
#pragma unroll
        for (int g = 0; g < N_SEG_GRADS; ++g) sg[g] = 0.0f;
        const float* M = ad.maps_d + base;
        sg[G_RX] = tx * (M[0] * ux + M[np] * uy + M[2 * np] * uz + M[9 * np]) +
                   ty * (M[3 * np] * ux + M[4 * np] * uy + M[5 * np] * uz + M[10 * np]) +
                   tz * (M[6 * np] * ux + M[7 * np] * uy + M[8 * np] * uz + M[11 * np]);
        sg[G_RY] = (ly * ma.x - lx * ma.y) - (nly * mb.x - nlx * mb.y);
    }
    lx = nlx, ly = nly, lz = bz;
}

// Two waves per SIMD need at most 256 registers each, VGPRs and AGPRs together: the attribute holds the allocation
// there (one wave per SIMD halves the sweep's speed); 16 coils do not fit without spilling and keep one wave.
// Reverse through a NUFFT readout run (head record q, m0 the state at the run's start; (lx, ly, lz) = dL/dm after the
// run in, at its start out). The forward takes the samples s_ic = sum over spins of K_ic S_c m0 (the spreading, the
// grid's DFT, the decay terms, the demodulation), then jumps over the run. Reversed:
//   1. the jump (an exact free step of L dt, phase L cyc), with its R2, R1, PD and phase gradients;
//   2. A_c = e^{2 pi i ph} sum_m e^{-R2 tau_m} sum_u psi(z_u) G[m][u][c] over the workgroup's window (the transpose of
//      the spreading, G from hb_nufft_adjoint_grid): lambda += sum_c conj(S_c) A_c, the coil gradients += conj(m0)
//      A_c; for a parameter theta of K, dL/dtheta = Re(m0 conj(sum_c conj(S_c) dA_c/dtheta)): R2 through the decay
//      terms (-tau_m e^{-R2 tau_m}), the phase per sample through ph and the cells' z (psi', dz/dfr = -ng 2 / width).
// The phase per sample cyc = bz dt (bz = g.r + db0 + b0_offset) gives the position and db0 gradients.
template <int NC>
__device__ __forceinline__ void nufft_reverse(const hb_segment& q, const hb_adjoint& ad, const Spin& p,
                                              const Stash& m0, float& lx, float& ly, float& lz, Acc& acc,
                                              const float (&sr)[NC > 0 ? NC : 1], const float (&si)[NC > 0 ? NC : 1],
                                              float (&gr)[NC > 0 ? NC : 1], float (&gi)[NC > 0 ? NC : 1]) {
    const float* par = ad.nu_params;
    const int L = q.pad[3], kind = q.pad[2], run = q.pad[4];
    const int ng = (int)par[0], terms = (int)par[2], shift = (int)par[4];
    const float hz = par[6 + HB_NUFFT_MAX_TERMS], bl2 = par[7 + HB_NUFFT_MAX_TERMS];
    const double cyc = free_cycles64(pos64(p), p, q);
    const float T = (float)L * q.dt;
    float dcyc = 0.0f;  // dL/dcyc
    // 1. the jump
    if (q.flags & (HB_XY_DEAD | HB_PULSE_SPOIL)) {
        lx = ly = 0.0f;
    } else {
        double cl = cyc * (double)L;
        cl -= rint(cl);
        float s2, c2;
        sincospi(-2.0f * (float)cl, &s2, &c2);
        const float d2 = expm1x(-T * p.r2), e2 = 1.0f + d2;
        const float nx = c2 * m0.x - s2 * m0.y, ny = s2 * m0.x + c2 * m0.y;
        acc.v[A_R2] -= T * e2 * (lx * nx + ly * ny);
        dcyc -= HB_TWO_PI * (float)L * e2 * (ly * nx - lx * ny);
        const float ax = lx + lx * d2, ay = ly + ly * d2;
        lx = c2 * ax + s2 * ay;
        ly = c2 * ay - s2 * ax;
    }
    {
        const float d1 = expm1x(-T * p.r1);
        acc.v[A_PD] -= lz * d1;
        acc.v[A_R1] -= lz * (m0.z - p.pd) * T * (1.0f + d1);
        lz = lz + lz * d1;
    }
    // 2. the samples
    if constexpr (NC > 0) {
        const double fr = cyc - floor(cyc);
        const double xg = fr * ng;
        double ph = fr * shift;
        ph -= rint(ph);
        float sp, cp;
        sincospi(2.0f * (float)ph, &sp, &cp);  // e^{+2 pi i ph}
        const float fre = m0.x * cp + m0.y * sp, fim = m0.x * sp - m0.y * cp;  // conj(m0) e^{2 pi i ph}
        const int* win = ad.nu_win + ((size_t)kind * gridDim.x + blockIdx.x) * 2;
        const int wb = win[0], W = win[1];
        const float z0 = (float)(((double)wb - xg) * (double)hz);
        const float* G = ad.nu_grid + (size_t)run * terms * ng * (2 * NC);
        const float beta = bl2 * 0.6931471805599453f;
        float Qr = 0.0f, Qi = 0.0f, Rr = 0.0f, Ri = 0.0f, Xr = 0.0f, Xi = 0.0f;
        float lcf = 0.0f;
        for (int lc = 0; lc < W; ++lc, lcf += 1.0f) {
            const float z = fmaf(lcf, hz, z0);
            const float u = fmaf(-z, z, 1.0f);
            const bool in = u > 0.0f;
            const float su = __builtin_amdgcn_sqrtf(in ? u : 1.0f);
            const float kv = in ? __builtin_amdgcn_exp2f(fmaf(bl2, su, -bl2)) : 0.0f;
            const float kd = in ? -kv * beta * z / su : 0.0f;  // d psi / dz
            int cell = (wb + lc) % ng;
            if (cell < 0) cell += ng;
            for (int m = 0; m < terms; ++m) {
                const float tau = par[6 + m];
                const float e = __builtin_amdgcn_exp2f(-(tau * 1.4426950408889634f) * p.r2);
                const float* g = G + ((size_t)m * ng + cell) * (2 * NC);
                const float w = kv * e;
                float Pr = 0.0f, Pi = 0.0f;
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    const float g_r = g[2 * c], g_i = g[2 * c + 1];
                    Pr += sr[c] * g_r + si[c] * g_i;  // conj(S_c) g
                    Pi += sr[c] * g_i - si[c] * g_r;
                    const float wgr = w * g_r, wgi = w * g_i;
                    gr[c] += fre * wgr - fim * wgi;
                    gi[c] += fre * wgi + fim * wgr;
                }
                Qr += w * Pr, Qi += w * Pi;
                const float wt = -tau * w;
                Rr += wt * Pr, Ri += wt * Pi;
                const float wd = kd * e;
                Xr += wd * Pr, Xi += wd * Pi;
            }
        }
        lx += cp * Qr - sp * Qi;
        ly += cp * Qi + sp * Qr;
        acc.v[A_R2] += m0.x * (cp * Rr - sp * Ri) + m0.y * (cp * Ri + sp * Rr);
        const float a2 = HB_TWO_PI * (float)shift, b2 = -(float)ng * hz;  // d/dfr: 2 pi i shift Q + (-ng hz) X
        const float Fr = -a2 * Qi + b2 * Xr, Fi = a2 * Qr + b2 * Xi;
        dcyc += m0.x * (cp * Fr - sp * Fi) + m0.y * (cp * Fi + sp * Fr);
    }
    const float gb = dcyc * (float)q.dt64;  // dL/dbz
    acc.v[A_PX] += gb * q.g_x;
    acc.v[A_PY] += gb * q.g_y;
    acc.v[A_PZ] += gb * q.g_z;
    acc.v[A_DB0] += gb;
}

// NU: the table has NUFFT runs (their reverse raises the register pressure of the unrolled block loop, so the other
// tables take the variant without it).
// The read-only inputs come in again as noalias parameters (the same pointers as in cs_in and ad_in): no store of the
// kernel can change what they point to, so the loads at wave-uniform addresses (the records, h rows, NUFFT tables)
// go through the scalar cache instead of one vector load per wave and a readfirstlane per value.
template <int NC, bool SEG, bool NU>
__global__ __launch_bounds__(kBlockA) __attribute__((amdgpu_waves_per_eu(NC <= 8 ? 2 : 1))) void adjoint_kernel(
    hb_case cs_in, hb_adjoint ad_in, int s0, int s1, const hb_segment* __restrict__ segs,
    const float* __restrict__ spins, const float* __restrict__ pos_lo, const float* __restrict__ maps,
    const float* __restrict__ ckpt, const float* __restrict__ h, const float* __restrict__ sens,
    const float* __restrict__ maps_d, const float* __restrict__ nu_params, const int32_t* __restrict__ nu_win,
    const float* __restrict__ nu_grid) {
    hb_case cs = cs_in;
    cs.segs = segs, cs.spins = spins, cs.pos_lo = pos_lo, cs.maps = maps;
    hb_adjoint ad = ad_in;
    ad.ckpt = ckpt, ad.h = h, ad.sens = sens, ad.maps_d = maps_d;
    ad.nu_params = nu_params, ad.nu_win = nu_win, ad.nu_grid = nu_grid;
    __shared__ float segbuf[2][kWavesA][8][8];  // alternate buffers: one barrier per workgroup sum
    __shared__ double accs[(8 + 2 * NC) * kBlockA];  // float64 accumulators: slot a of lane tid at a * kBlockA + tid
    __shared__ double seeds_l[12 * kBlockA];          // a single reused pulse's seeds, same layout
    int sbuf = 0;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = blockIdx.x * kBlockA + tid;
    const int np = cs.n_pad;
    const bool coil_grads = NC > 0 && ad.coil_grads;
    const bool seeds_in_lds = cs.n_pulses == 1 && ad.seeds;  // uniform
    if (seeds_in_lds) {
#pragma unroll
        for (int a = 0; a < 12; ++a) seeds_l[a * kBlockA + tid] = ad.seeds[(size_t)a * np + j];
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) accs[a * kBlockA + tid] = ad.gs[(size_t)a * np + j];
    if (coil_grads) {
#pragma unroll
        for (int c = 0; c < 2 * NC; ++c) accs[(8 + c) * kBlockA + tid] = ad.gc[(size_t)c * np + j];
    }
    const Spin p = load_spin(cs.spins, cs.pos_lo, np, j);
    const float b1sq = p.b1 * p.b1;
    const bool rf_only = SEG && ad.seg_grads == 2;

    // 1. replay the interval, r8[b] = state before segment s0 + 8 b (low parts folded)
    {
        State m = load_state(ad.ckpt, np, j);
        Cache k = {1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        bool free_ok = false, rf_ok = false;
        int s_start = s0;
        if (cs.segs[s0].flags & HB_PULSE_BODY) {  // inside an occurrence or a run begun in an earlier interval
            const hb_segment& hq = cs.segs[cs.segs[s0].pad[3]];
            const int e = min(cs.segs[s0].pad[3] + ((hq.flags & HB_PULSE) ? hq.pad[2] : hq.pad[3]), s1);
            r8_after_span(ad.r8, np, j, s0, s1, s0 - 1, e, m);
            s_start = e;
        }
        for (int s = s_start; s < s1; ++s) {
            if (((s - s0) & 7) == 0) {
                float* r = ad.r8 + (size_t)((s - s0) >> 3) * 3 * np;
                r[j] = m.mx, r[np + j] = m.my, r[2 * np + j] = m.mz + m.lz;
            }
            const hb_segment q = cs.segs[s];
            if (q.flags & HB_PULSE) {  // a reused pulse: its map, then the state stands over the occurrence
                apply_map(cs.maps + (size_t)q.pad[1] * 12 * np + j, np, q.w_rf_x, q.w_rf_y, m);
                if (q.flags & HB_PULSE_SPOIL) m.mx = m.my = 0.0f;
                const int e = min(s + q.pad[2], s1);
                r8_after_span(ad.r8, np, j, s0, s1, s, e, m);
                s = e - 1;
                continue;
            }
            if (NU && (q.flags & HB_NUFFT)) {  // a readout run: its jump (its samples add nothing to the state)
                run_jump(q, p, free_cycles64(pos64(p), p, q), expm1x(-((float)q.pad[3] * q.dt) * p.r1), m);
                const int e = min(s + q.pad[3], s1);
                r8_after_span(ad.r8, np, j, s0, s1, s, e, m);
                s = e - 1;
                continue;
            }
            if (q.flags & HB_TRAIN) {  // an RF train: its own loop, the next record in flight (forward_kernel's)
                const int L = min(q.pad[3], s1 - s);
                RfRec cur = load_rf(&cs.segs[s]);
                for (int i = 1; i <= L; ++i) {
                    const RfRec nxt = load_rf(&cs.segs[s + (i < L ? i : 0)]);
                    if (i > 1 && ((s + i - 1 - s0) & 7) == 0) {
                        float* r = ad.r8 + (size_t)((s + i - 1 - s0) >> 3) * 3 * np;
                        r[j] = m.mx, r[np + j] = m.my, r[2 * np + j] = m.mz + m.lz;
                    }
                    rf_step(cur, live_flags(cur.flags, true, rf_ok, 0, HB_SAME_RF_DT), p, b1sq, m, k);
                    rf_ok = true;
                    cur = nxt;
                }
                s += L - 1;
                continue;
            }
            const int f = live_flags(q.flags, free_ok, rf_ok, HB_SAME_FREE, HB_SAME_RF_DT);
            step(q, f, p, b1sq, m, k);
            if (f & HB_RF) rf_ok = true; else free_ok = true;
            if (f & HB_SPOIL) m.mx = m.my = 0.0f;
        }
    }

    // 2. reverse sweep
    float lx = ad.lam[j], ly = ad.lam[np + j], lz = ad.lam[2 * np + j];
    float sr[NC > 0 ? NC : 1], si[NC > 0 ? NC : 1], gr[NC > 0 ? NC : 1], gi[NC > 0 ? NC : 1];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        sr[c] = ad.sens[(2 * c) * (size_t)np + j];
        si[c] = ad.sens[(2 * c + 1) * (size_t)np + j];
        gr[c] = gi[c] = 0.0f;
    }
    Acc acc;
#pragma unroll
    for (int a = 0; a < 8; ++a) acc.v[a] = 0.0f;
    RevCache rc = {1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    bool rfree_ok = false, rrf_ok = false;  // reverse caches (hints: same as the next segment of the kind)
    // a record that adds to the accumulators was reversed since the last flush (else they hold exact zeros and the
    // flush is skipped: the float64 sums are the same bit for bit, and spans of pulses and runs cost no HBM traffic)
    bool dirty = false;
    const int n8 = (s1 - s0 + 7) >> 3;
    for (int b = n8 - 1; b >= 0; --b) {
        const int sb = s0 + 8 * b;
        const int nseg = min(8, s1 - sb);
        // a block inside one reused pulse's occurrence is the identity: only its (zero) segment gradients
        const hb_segment& qf = cs.segs[sb];
        const hb_segment& ql = cs.segs[sb + nseg - 1];
        bool flush_point = (sb & 63) == 0;
        if ((qf.flags & ql.flags & HB_PULSE_BODY) && qf.pad[3] == ql.pad[3]) {
            // so is every block from the first one after the record's head down to this one: all skipped at once
            const int head = qf.pad[3];
            const int b_min = head < s0 ? 0 : (head + 1 - s0 + 7) >> 3;
            const int lo = s0 + 8 * b_min;
            // (their gseg rows are not written: hb_seg_reduce, given the records, sums bodies as 0)
            flush_point = (sb & ~63) >= lo;  // one of the skipped blocks starts at a multiple of 64
            b = b_min;
        } else {
        // replay the block with its states in registers: before[i] before segment sb + i, after[i] after its step
        Stash before[8];
        StashXY after[8];
        {
            const float* r = ad.r8 + (size_t)b * 3 * np;
            State m;
            m.mx = r[j], m.my = r[np + j], m.mz = r[2 * np + j], m.lz = 0.0f;
            Cache k = {1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            bool free_ok = false, rf_ok = false;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                if (i < nseg) {
                    before[i] = {m.mx, m.my, m.mz + m.lz};
                    const hb_segment q = cs.segs[sb + i];
                    if (q.flags & HB_PULSE_BODY) {
                        after[i] = {m.mx, m.my};
                        continue;
                    }
                    if (q.flags & HB_PULSE) {
                        apply_map(cs.maps + (size_t)q.pad[1] * 12 * np + j, np, q.w_rf_x, q.w_rf_y, m);
                        after[i] = {m.mx, m.my};
                        if (q.flags & HB_PULSE_SPOIL) m.mx = m.my = 0.0f;
                        continue;
                    }
                    if (NU && (q.flags & HB_NUFFT)) {
                        run_jump(q, p, free_cycles64(pos64(p), p, q), expm1x(-((float)q.pad[3] * q.dt) * p.r1), m);
                        after[i] = {m.mx, m.my};
                        continue;
                    }
                    const int f = live_flags(q.flags, free_ok, rf_ok, HB_SAME_FREE, HB_SAME_RF_DT);
                    step(q, f, p, b1sq, m, k);
                    if (f & HB_RF) rf_ok = true; else free_ok = true;
                    after[i] = {m.mx, m.my};
                    if (f & HB_SPOIL) m.mx = m.my = 0.0f;
                }
            }
        }
        float sg[N_SEG_GRADS];
        f4 segd = {0.0f, 0.0f, 0.0f, 0.0f};
        bool seg_block = !rf_only;  // the block's segment gradients are formed (rf_only: it holds an RF record)
#pragma unroll
        for (int i = 7; i >= 0; --i) {
            if (i < nseg) {
                const int s = sb + i;
                const hb_segment q = cs.segs[s];
                if (q.flags & HB_PULSE_BODY) continue;  // the identity, no gradients
                if constexpr (NU) {
                    if (q.flags & HB_NUFFT) {
                        nufft_reverse<NC>(q, ad, p, before[i], lx, ly, lz, acc, sr, si, gr, gi);
                        dirty = true;
                        continue;
                    }
                }
                if (q.flags & HB_PULSE) {
                    if (q.flags & HB_PULSE_SPOIL) lx = ly = 0.0f;
                    pulse_reverse<SEG>(q, cs, ad, j, before[i], after[i], lx, ly, lz, sg, seeds_l, seeds_in_lds);
                    if (SEG) seg_add(segd, sg, i, lane);
                    seg_block = true;
                    continue;
                }
                const int f = live_flags(q.flags, rfree_ok, rrf_ok, HB_NEXT_FREE, HB_NEXT_RF_DT);
                dirty = true;  // a stepped record adds to the accumulators (a pulse's map adds only to the seeds)
                if (f & HB_SPOIL) lx = ly = 0.0f;
                if (NC > 0 && (f & HB_ADC)) {
                    const float* hrow = ad.h + (size_t)q.pad[0] * 2 * NC;
                    const float ax = after[i].x, ay = after[i].y;
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        const float hr = hrow[2 * c], hi = hrow[2 * c + 1];
                        lx += hr * sr[c] + hi * si[c];
                        ly += hi * sr[c] - hr * si[c];
                        gr[c] += hr * ax + hi * ay;
                        gi[c] += hi * ax - hr * ay;
                    }
                }
#pragma unroll
                for (int g = 0; g < N_SEG_GRADS; ++g) sg[g] = 0.0f;
                // the reverse hints name the next segment of the kind: map them onto the step's bits
                const int fr = (f & (HB_RF | HB_ADC | HB_SPOIL)) | ((f & HB_NEXT_FREE) ? HB_SAME_FREE : 0) |
                               ((f & HB_NEXT_RF_DT) ? HB_SAME_RF_DT : 0);
                reverse_step<SEG>(q, fr, p, before[i].x, before[i].y, before[i].z, lx, ly, lz, acc, sg, rc);
                if (f & HB_RF) rrf_ok = true; else rfree_ok = true;
                if (SEG && (!rf_only || (f & HB_RF))) {
                    seg_add(segd, sg, i, lane);
                    seg_block = true;
                }
            }
        }
        if (SEG && !seg_block) {  // rf_only and no RF record here: zeros, without the workgroup sum
            if (tid < 8 * nseg) ad.gseg[((size_t)blockIdx.x * ad.k_max + (sb - s0)) * 8 + tid] = 0.0f;
        } else if (SEG) {
            // the other buffer than the last sum's: a wave may still be reading that one (a wave reaches the next
            // write to this one only after every wave has passed the next sum's barrier, i.e. done reading)
            seg_store(segd, segbuf[sbuf][wave], lane);
            __syncthreads();
            if (tid < 8 * nseg) {
                const int i = tid >> 3, g = tid & 7;
                float v = 0.0f;
                if (g < N_SEG_GRADS) {
                    v = segbuf[sbuf][0][i][g];
#pragma unroll
                    for (int w = 1; w < kWavesA; ++w) v += segbuf[sbuf][w][i][g];
                }
                ad.gseg[((size_t)blockIdx.x * ad.k_max + (sb - s0 + i)) * 8 + g] = v;
            }
            sbuf ^= 1;
        }
        }
        if (flush_point && dirty) {  // every 64 segments: into the float64 accumulators
            dirty = false;
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                accs[a * kBlockA + tid] += (double)acc.v[a];
                acc.v[a] = 0.0f;
            }
            if (coil_grads) {
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    accs[(8 + 2 * c) * kBlockA + tid] += (double)gr[c];
                    accs[(9 + 2 * c) * kBlockA + tid] += (double)gi[c];
                    gr[c] = gi[c] = 0.0f;
                }
            }
        }
    }
    // an interval that does not start at a multiple of 64 leaves a remainder (only possible for the first one)
#pragma unroll
    for (int a = 0; a < 8; ++a) ad.gs[(size_t)a * np + j] = accs[a * kBlockA + tid] + (double)acc.v[a];
    if (coil_grads) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            ad.gc[(size_t)(2 * c) * np + j] = accs[(8 + 2 * c) * kBlockA + tid] + (double)gr[c];
            ad.gc[(size_t)(2 * c + 1) * np + j] = accs[(9 + 2 * c) * kBlockA + tid] + (double)gi[c];
        }
    }
    if (seeds_in_lds) {
#pragma unroll
        for (int a = 0; a < 12; ++a) ad.seeds[(size_t)a * np + j] = seeds_l[a * kBlockA + tid];
    }
    ad.lam[j] = lx, ad.lam[np + j] = ly, ad.lam[2 * np + j] = lz;
}

// seg[s0 + k][g] = sum over workgroups of gseg[b][k][g], float64, in a fixed order: the workgroups in chunks of
// kSegChunk summed in order by one lane per (k, g) (consecutive lanes read consecutive floats), then the chunks in
// order. With the records (segs non-NULL), a body of a reused pulse's occurrence or of a NUFFT run (HB_PULSE_BODY:
// the identity, whose gradients are 0 and whose gseg rows the adjoint kernel leaves unwritten) sums to 0 unread.
constexpr int kSegChunk = 256;

__global__ __launch_bounds__(256) void seg_chunks_kernel(const float* __restrict__ gseg, int n_blocks, int k_max,
                                                         int len, const hb_segment* __restrict__ segs, int s0,
                                                         double* __restrict__ work) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 8 * len) return;
    const size_t stride = (size_t)k_max * 8;
    double acc = 0.0;
    if (!segs || !(segs[s0 + (i >> 3)].flags & HB_PULSE_BODY)) {
        const int b0 = blockIdx.y * kSegChunk, b1 = min(n_blocks, b0 + kSegChunk);
#pragma unroll 16
        for (int b = b0; b < b1; ++b) acc += (double)gseg[b * stride + i];
    }
    work[blockIdx.y * stride + i] = acc;
}

__global__ __launch_bounds__(256) void seg_final_kernel(const double* __restrict__ work, int chunks, int k_max,
                                                        int s0, int len, double* __restrict__ seg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 8 * len) return;
    double acc = 0.0;
    for (int c = 0; c < chunks; ++c) acc += work[(size_t)c * k_max * 8 + i];
    seg[(size_t)s0 * 8 + i] = acc;
}

template <int NC>
static hipError_t launch_adjoint(const hb_case* cs, const hb_adjoint* ad, int s0, int s1, hipStream_t st) {
    const dim3 grid(cs->n_pad / kBlockA), block(kBlockA);
    const bool nu = ad->nu_params != nullptr;
    auto kernel = ad->seg_grads ? (nu ? adjoint_kernel<NC, true, true> : adjoint_kernel<NC, true, false>)
                                : (nu ? adjoint_kernel<NC, false, true> : adjoint_kernel<NC, false, false>);
    hipLaunchKernelGGL(kernel, grid, block, 0, st, *cs, *ad, s0, s1, cs->segs, cs->spins, cs->pos_lo, cs->maps,
                       ad->ckpt, ad->h, ad->sens, ad->maps_d, ad->nu_params, ad->nu_win, ad->nu_grid);
    return hipGetLastError();
}

}  // namespace hb

extern "C" int hb_adjoint_interval(const hb_case* cs, const hb_adjoint* ad, int32_t s0, int32_t s1, void* stream) {
    using namespace hb;
    if (!cs || !ad || cs->n_pad <= 0 || cs->n_pad % kBlockA) return hipErrorInvalidValue;
    if (s0 < 0 || s1 < s0 || s1 > cs->n_segments || s1 - s0 > ad->k_max) return hipErrorInvalidValue;
    if (ad->nu_params && (!ad->nu_win || (ad->nc > 0 && !ad->nu_grid))) return hipErrorInvalidValue;
    if (s1 == s0) return hipSuccess;
    hipStream_t st = (hipStream_t)stream;
    switch (ad->nc) {
        case 0: return launch_adjoint<0>(cs, ad, s0, s1, st);
        case 1: return launch_adjoint<1>(cs, ad, s0, s1, st);
        case 2: return launch_adjoint<2>(cs, ad, s0, s1, st);
        case 4: return launch_adjoint<4>(cs, ad, s0, s1, st);
        case 8: return launch_adjoint<8>(cs, ad, s0, s1, st);
        case 16: return launch_adjoint<16>(cs, ad, s0, s1, st);
        default: return hipErrorInvalidValue;
    }
}

extern "C" int64_t hb_seg_reduce_workspace(int32_t n_blocks, int32_t k_max) {
    if (n_blocks < 1 || k_max < 0) return 0;
    return (int64_t)((n_blocks + hb::kSegChunk - 1) / hb::kSegChunk) * k_max * 8;
}

extern "C" int hb_seg_reduce(const float* gseg, int32_t n_blocks, int32_t k_max, int32_t s0, int32_t len,
                             double* work, double* seg, const hb_segment* segs, void* stream) {
    using namespace hb;
    if (n_blocks < 1 || len < 0 || len > k_max || s0 < 0 || !work) return hipErrorInvalidValue;
    if (len == 0) return hipSuccess;
    const int chunks = (n_blocks + kSegChunk - 1) / kSegChunk;
    hipLaunchKernelGGL(seg_chunks_kernel, dim3((8 * len + 255) / 256, chunks), dim3(256), 0, (hipStream_t)stream,
                       gseg, n_blocks, k_max, len, segs, s0, work);
    hipLaunchKernelGGL(seg_final_kernel, dim3((8 * len + 255) / 256), dim3(256), 0, (hipStream_t)stream, work,
                       chunks, k_max, s0, len, seg);
    return hipGetLastError();
}
