/* blochparty_amd: the bound-pool (magnetization transfer, rate model) forward path.
 *
 * Replaces the reference's V = 1 GPU variant for Phantom.bound with the rate model (step_mt and lineshape of
 * src/blochparty/engines/gpu/shaders/common.metal:113-208, exchange.py:209-256): the free pool keeps its Bloch
 * dynamics, the bound pool only Mzb (equilibrium f PD), the two exchange longitudinal magnetization, and RF saturates
 * the bound pool at W = (2 pi b1 |rf|)^2 G(f_rf - bz - dw). Conventions as blochparty_amd.h: every pointer is device
 * memory the caller owns, every entry point enqueues on `stream` and returns a hipError_t (0 = success). */
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

/* The bound pool of a case. pool: (5, n_pad) f, k [Hz], R1b [1/s], T2b [s], dw [Hz]; mb: (2, n_pad) Mzb and its
 * two-sum low part, in / out; sens: (2 n_coils, n_pad) Re S_0, Im S_0, Re S_1, ...; rf_freq: (n_segments,) each
 * segment's RF frequency [Hz] (a pulse turning as exp(-i 2 pi f t) is resonant at offset f); table: n_table values of
 * ln F(x) at ln x = lx0 + i dlx (the super-Lorentzian integral, G = T2b F(2 pi max(|delta|, cutoff) T2b), 0 past the
 * table's end), NULL for the Lorentzian. */
typedef struct hb_mt {
    const float* pool;
    float* mb;
    const float* sens;
    const float* rf_freq;
    const float* table;
    float lx0, dlx, cutoff;
    int32_t n_table, lineshape; /* HB_LINE_* */
    int32_t reserved;           /* 0 */
} hb_mt;

/* The bound-pool event loop over segments [s0, s1) of plain records (no runs, trains, merged runs, reused pulses or
 * NUFFT runs in the table), one lane per spin: every ADC sample's coil sums per workgroup of 256 spins go to
 * partial[(block * rows + r) * 16 * n_tiles + 2 c + {re, im}], r = the segment's pad[0] - row0 (< rows), as
 * hb_forward's, for hb_signal_reduce. Replaces the reference's bloch_chunk_mma with V = 1 and a rate bound pool. */
int hb_forward_mt(const hb_case* cs, const hb_mt* mt, int32_t s0, int32_t s1, float* partial, int32_t rows,
                  int32_t row0, void* stream);

#ifdef __cplusplus
}
#endif

#endif
