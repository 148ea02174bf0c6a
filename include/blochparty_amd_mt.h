/* blochparty_amd: the full-variant forward path, a bound pool and / or the scanner's concomitant fields.
 *
 * Replaces the reference's V = 1 / "_full" GPU variant (engines/gpu/driver.py:520, shaders/common.metal:113-208 and
 * :563-601): for Phantom.bound with the rate model (step_mt and lineshape, exchange.py:209-256) the free pool keeps
 * its Bloch dynamics, the bound pool only Mzb (equilibrium f PD), the two exchange longitudinal magnetization, and RF
 * saturates the bound pool at W = (2 pi b1 |rf|)^2 G(f_rf - bz - dw); for a Scanner's concomitant fields (fields.py,
 * engines/core.py:115-121, 190-217) every segment's offset gains the spin's concomitant field. Conventions as
 * blochparty_amd.h: every pointer is device memory the caller owns, every entry point enqueues on `stream` and returns
 * a hipError_t (0 = success). */
#ifndef BLOCHPARTY_AMD_MT_H
#define BLOCHPARTY_AMD_MT_H

#include <stdint.h>

#include "blochparty_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Segment flag: RF that only the bound pool feels (Sequence.bound_only): the water precesses as in a free segment
 * while the bound pool is saturated by the segment's RF. */
#define HB_BOUND_ONLY 262144

#define HB_LINE_LORENTZIAN 1
#define HB_LINE_SUPERLORENTZIAN 2

/* Concomitant-field modes (the reference's SpinFields.mode). */
#define HB_CONC_NONE 0
#define HB_CONC_LEADING 1 /* conc rows c_0..c_5: f_c = c . (gx^2, gy^2, gz^2, gx gy, gx gz, gy gz) [Hz] */
#define HB_CONC_EXACT 2   /* conc rows ax (3), ay (3), f0: f_c = |B| - f0 - lin, (Bx, By) = (ax . g, ay . g) */

/* A case's bound pool and concomitant fields. pool: (5, n_pad) f, k [Hz], R1b [1/s], T2b [s], dw [Hz], or NULL
 * without a bound pool (then mb, rf_freq and table are unused); mb: (2, n_pad) Mzb and its two-sum low part, in / out;
 * sens: (2 n_coils, n_pad) Re S_0, Im S_0, Re S_1, ...; rf_freq: (n_segments,) each segment's RF frequency [Hz] (a
 * pulse turning as exp(-i 2 pi f t) is resonant at offset f); table: n_table values of ln F(x) at ln x = lx0 + i dlx
 * (the super-Lorentzian integral, G = T2b F(2 pi max(|delta|, cutoff) T2b), 0 past the table's end), NULL for the
 * Lorentzian; conc: (7, n_pad) the spins' concomitant coefficients for conc_mode (HB_CONC_*; rows past the mode's
 * unused), NULL for HB_CONC_NONE. */
typedef struct hb_mt {
    const float* pool;
    float* mb;
    const float* sens;
    const float* rf_freq;
    const float* table;
    const float* conc;
    float lx0, dlx, cutoff;
    int32_t n_table, lineshape; /* HB_LINE_* (ignored without a bound pool) */
    int32_t conc_mode;          /* HB_CONC_* */
} hb_mt;

/* The full-variant event loop (a bound pool, concomitant fields, bound-only RF: any of them) over segments
 * [s0, s1) of plain records (no runs, trains, merged runs, reused pulses or NUFFT runs in the table), one lane per
 * spin: every ADC sample's coil sums per workgroup of 256 spins go to partial[(block * rows + r) * 16 * n_tiles +
 * 2 c + {re, im}], r = the segment's pad[0] - row0 (< rows), as hb_forward's, for hb_signal_reduce. Replaces the
 * reference's bloch_chunk_mma with V = 1 (a rate bound pool and / or concomitant fields). */
int hb_forward_mt(const hb_case* cs, const hb_mt* mt, int32_t s0, int32_t s1, float* partial, int32_t rows,
                  int32_t row0, void* stream);

/* The full variant's reverse sweep of one checkpoint interval [s0, s1) of plain records (replaces the reference's V = 1
 * replay and adjoint, shaders/adjoint_common.metal:66-203, per interval of engines/gpu/adjoint.py:179-220; last
 * interval first). One lane per spin, float64: the interval replayed from its checkpoint with the state before every
 * segment kept in traj, then reversed segment by segment.
 * ckpt: (6, n_pad) the forward's state at s0 (hb_case.state's layout); ckpt_mb: (2, n_pad) Mzb and its low part there
 * (with a bound pool; else unused); traj: (s1 - s0, 4, n_pad) float32 scratch; lam: (4, n_pad) dL/d(Mx, My, Mz, Mzb)
 * after s1 in, before s0 out; gs: (HB_FULL_GRADS, n_pad) float64 per-spin gradients, added to, rows px py pz (encoding
 * positions), db0, pd, R1, R2, b1, the 7 concomitant coefficients (hb_mt.conc's rows), then the bound pool's f, k,
 * R1b, T2b, dw (the reference's N_SPIN_GRADS rows, adjoint.py:42); gc: (2 nc, n_pad) float64 coil gradients (Re, Im),
 * added to (with coil_grads); gseg: (n_pad / 256, k_max, 8) each workgroup's sums per segment of dL/d(dt, rf_x, rf_y,
 * g_x, g_y, g_z, b0_offset, rf_freq) (with seg_grads; every row of the interval written), for hb_seg_reduce;
 * h: (n_adc, nc, 2) dL/dsignal turned by exp(+i psi); sens: (2 nc, n_pad) Re S_0, Im S_0, ... (nc <= 16; more coils
 * take one sweep per tile of 16, the sweep being linear in its sources). */
typedef struct hb_full_adjoint {
    const float* ckpt;
    const float* ckpt_mb;
    float* traj;
    float* lam;
    double* gs;
    double* gc;
    float* gseg;
    const float* h;
    const float* sens;
    int32_t nc, k_max, coil_grads, seg_grads;
} hb_full_adjoint;

#define HB_FULL_GRADS 20

int hb_adjoint_full(const hb_case* cs, const hb_mt* mt, const hb_full_adjoint* ad, int32_t s0, int32_t s1,
                    void* stream);

#ifdef __cplusplus
}
#endif

#endif
