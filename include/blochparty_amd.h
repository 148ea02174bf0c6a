/*
 * blochparty_amd: C ABI of the MI355X (gfx950) engine behind blochparty's simulate(..., engine="gpu").
 *
 * Every entry point takes plain device pointers, sizes and a hipStream_t (passed as void*), enqueues its kernels on
 * that stream, allocates nothing and returns a hipError_t as int (0 = success; hb_error_string() names the code).
 * The caller owns every buffer. Buffers are float32 structure-of-arrays over n_pad spins (n_pad a multiple of 256;
 * padding spins carry zero PD, zero state and zero sensitivities, so they add nothing).
 *
 * The reference binds its GPU kernels from Python through NVRTC + the CUDA driver API
 * (src/blochparty/engines/gpu/cuda/rtc.py:116-126: every argument a device pointer, the launch on PyTorch's current
 * stream, errors raised as RuntimeError). Each entry point below names the kernel or host routine it replaces.
 */
#ifndef BLOCHPARTY_AMD_H
#define BLOCHPARTY_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Segment flags: the reference's flag bits (src/blochparty/engines/gpu/driver.py:375-382), cache hints and the marks of
 * reused pulses, readout runs and RF trains. */
#define HB_RF 1           /* an RF segment (Strang step)                                                   */
#define HB_ADC 2          /* sample the signal at the end of the segment                                  */
#define HB_SPOIL 4        /* zero the transverse magnetization at the end of the segment                  */
#define HB_SAME_FREE 8    /* free segment equal (dt, gradient, b0_offset) to the previous free one here    */
#define HB_SAME_RF_DT 16  /* RF segment with the previous RF segment's dt (reuse its relaxation factors)    */
#define HB_NEXT_FREE 32   /* free segment equal to the next free one in the interval (reverse sweep cache)   */
#define HB_NEXT_RF_DT 64  /* RF segment with the next RF segment's dt in the interval (reverse sweep cache)   */
#define HB_PULSE 128      /* first segment of a reused pulse: apply map pad[1] turned by psi, skip pad[2] segs */
#define HB_PULSE_SPOIL 256 /* the reused pulse's last segment is spoiled                                       */
#define HB_RUN 512        /* first of pad[3] identical free ADC segments: step Mxy per sample, Mz once        */
#define HB_SAME_RUN 1024  /* a readout run (HB_RUN or HB_NUFFT) with the previous one's dt and length (reuse its
                             Mz factor)                                                                         */
#define HB_TRAIN 2048     /* first of pad[3] consecutive RF segments without ADC or spoil: stepped in one loop */
#define HB_MERGED 4096    /* first of pad[3] free segments merged into it (summed dt and moments); only the last
                             may be sampled (HB_ADC: the sample is taken after the merged step)                 */
#define HB_SPOILED_FREE 8192 /* free record without ADC whose Mxy never reaches a sample (spoiled at its end, or
                             starting at Mxy = 0): Mz relaxes over its dt, Mxy is zeroed                        */
#define HB_NUFFT 16384    /* first of pad[3] readout samples taken as a NUFFT: slot pad[1], kind pad[2], index
                             among all runs pad[4]; a spoil at the run's end is HB_PULSE_SPOIL                  */
#define HB_XY_DEAD 32768  /* a NUFFT run whose Mxy at its end is never read (the next record is HB_SPOILED_FREE):
                             the run's jump relaxes Mz only and zeroes Mxy                                     */
#define HB_SAME_FREE_DT 65536 /* free segment with the previous free one's dt here (reuse its relaxation factors,
                             also where the phase differs)                                                    */
#define HB_PULSE_BODY 131072 /* a segment of a reused pulse's occurrence after its first (the map at the first one
                             covers it): the identity, in tables with checkpoints (the adjoint's); pad[3] the
                             occurrence's first segment                                                       */
#define HB_NUFFT_MAX_TERMS 8
#define HB_NUFFT_WMAX 32 /* widest window of grid cells one workgroup of 256 spins may span (engines/nufft.py WMAX) */

/* One segment of the event list, 128 bytes (the reference's segs row + flags + rows, driver.py:357-384). psi is the
 * receiver phase wrapped to [0, 2 pi) in float64 before the cast (driver.py:357). The float64 block carries the
 * duration, the gradient moment g dt and the b0_offset area, from which free segments take their phase in float64
 * (k . r reduced mod 1 cycle per segment). The w block holds the segment's rotation-vector factors, formed in float64
 * on the host: w = -2 pi dt, w_rf = w rf, w_g = w g, w_b0 = w b0_offset, w_rf2 = |w_rf|^2 (a spin's RF rotation
 * vector is (b1 w_rf, w_g . r + w db0 + w_b0)). pad[0] is the ADC row of an ADC segment (the reference's adc_row);
 * a reused pulse's first segment (HB_PULSE) holds the pulse id in pad[1], its length in pad[2], the occurrence's
 * RF phase in psi (driver.py:592-596) and its cosine and sine in w_rf_x, w_rf_y (that segment is never stepped); a readout run's (HB_RUN), an RF train's (HB_TRAIN) or a merged free run's
 * (HB_MERGED) first segment its length in pad[3]. */
typedef struct hb_segment {
    float dt, rf_x, rf_y, g_x, g_y, g_z, b0, psi;
    double dt64, m_x, m_y, m_z, b0_area;
    int32_t flags, pad[5];
    float w, w_rf_x, w_rf_y, w_g_x, w_g_y, w_g_z, w_b0, w_rf2;
} hb_segment;

/* Device buffers of one case (driver.py:320-433). spins: (8, n_pad) px py pz pd r1 r2 db0 b1; pos_lo: (3, n_pad)
 * the float64 positions' remainders (pos = hi + lo) for the float64 k.r phase of free segments; sens_a: coil
 * sensitivities in the matrix-core operand order, (n_pad / 64, n_tiles, 16, 64) floats, n_tiles = ceil(C / 8), rows
 * of a tile Re S_0, Im S_0, Re S_1, Im S_1, ... state: (6, n_pad) mx my mz and their low parts (the reference's M and L
 * buffers); the kernels carry a two-sum low part for mz only and fold the transverse ones in when they load. */
typedef struct hb_case {
    const float* spins;
    const float* pos_lo;
    const float* sens_a;
    const hb_segment* segs;
    float* state;
    const float* maps; /* (n_pulses, 12, n_pad) maps of reused pulses (hb_pulse_maps), or NULL */
    int32_t n, n_pad, n_coils, n_tiles, n_segments, n_pulses;
    /* hb_forward: nonzero promises that no record of [s0, s1) is stepped as an RF segment (every pulse there is
     * applied as a map), so the kernel may leave the RF step out; 0 is always correct. Ignored elsewhere. */
    int32_t rf_free;
    /* hb_forward: powers of two (0 reads as 1) with |sensitivity| * sens_scale <= 1 and |M| * x_scale <= 1 for every
     * magnetization the case can reach (|M| <= max(PD, |m_init|)): the coil contraction works on these scaled values
     * (as float16 hi + lo pairs for up to 16 coils) and undoes the scales exactly. */
    float sens_scale;
    float x_scale;
    int32_t reserved; /* 0 */
} hb_case;

/* NUFFT readouts of a dispatch (readouts="nufft"; the reference's nu buffer and grid rows, driver.py:549-580).
 * params: (8 + HB_NUFFT_MAX_TERMS,) float32 ng, width, terms, beta, shift (= n / 2 + 1), cols, then the decay nodes
 * tau_m [s], then 2 / width and beta log2(e); win: (kinds, n_pad / 256, 2) every workgroup's window start (grid cell, may be negative) and width
 * (cells) per run kind; part: (runs_max, n_pad / 256, cols, 16 n_tiles) the workgroups' window columns (column =
 * local cell * terms + term) of the dispatch's runs, by slot (each slot's columns one dense stretch). */
typedef struct hb_nufft {
    const float* params;
    const int32_t* win;
    float* part;
    int32_t runs_max, cols;
    int32_t terms;    /* = params[2], 1..HB_NUFFT_MAX_TERMS (selects the kernel variant) */
    int32_t reserved; /* 0 */
} hb_nufft;

/* Forward event loop over segments [s0, s1) (replaces bloch_chunk_mma, cuda/kernels.cu.in:2-185, one launch per
 * dispatch of driver.py:658-698). Every ADC sample's coil sums per workgroup of 256 spins go to
 * partial[(block * rows + r) * 16 * n_tiles + 2 c + {0: re, 1: im}], r the sample's index among this dispatch's
 * samples that are not in NUFFT runs (< rows), block = spin / 256; no atomics. nu: NULL, or the dispatch's NUFFT
 * runs (HB_NUFFT records), whose window columns go to nu->part. A dispatch whose first segment is HB_PULSE_BODY starts
 * after that occurrence (its map was applied at the occurrence's first segment, in an earlier dispatch). */
int hb_forward(const hb_case* cs, int32_t s0, int32_t s1, float* partial, int32_t rows, const hb_nufft* nu,
               void* stream);

/* Deterministic cross-block sum (replaces the float atomicAdd of cuda/common.cuh:77-89): for r < rows,
 * sig[((row0 + r) * n_coils + c) * 2 + {re, im}] = exp(-i psi[row0 + r]) * sum over blocks, in float64 in a fixed
 * order (blocks in chunks of 256, each summed in block order, then the chunks in order). psi: (n_adc,) float64.
 * work: hb_signal_reduce_workspace(n_blocks, rows, n_tiles) doubles. */
int64_t hb_signal_reduce_workspace(int32_t n_blocks, int32_t rows, int32_t n_tiles);
int hb_signal_reduce(const float* partial, int32_t n_blocks, int32_t rows, int32_t n_tiles, int32_t n_coils,
                     const double* psi, int32_t row0, const int32_t* row_map, double* work, double* sig,
                     void* stream);

/* hb_signal_reduce without a workspace argument, for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc); no counterpart in the reference. It calls no hipMalloc or
 * hipFree: the chunk sums go to a fixed device array of HB_NOALLOC_SCRATCH_DOUBLES doubles (16 MiB) that each GPU
 * the library runs on holds until the process ends, its address resolved once per device on the first call there.
 * Same arguments and results as hb_signal_reduce (the same kernels, bit for bit). hipErrorInvalidValue when
 * hb_signal_reduce_workspace(n_blocks, rows, n_tiles) exceeds HB_NOALLOC_SCRATCH_DOUBLES (e.g. more than 8,192
 * rows of one tile with 4,096 blocks). Calls on one GPU share the array: calls that overlap there (other streams,
 * other threads) race. */
#define HB_NOALLOC_SCRATCH_DOUBLES (1 << 21)
int hb_signal_reduce_noalloc(const float* partial, int32_t n_blocks, int32_t rows, int32_t n_tiles,
                             int32_t n_coils, const double* psi, int32_t row0, const int32_t* row_map, double* sig,
                             void* stream);

/* The NUFFT grids of a dispatch's runs, float64 (replaces the float atomics of the reference's grid rows): for slot
 * r < n_slots (run run_of[r], kind kind_of[run]), term m, cell u < ng and value v < 16 n_tiles,
 * grid[((r * terms + m) * ng + u) * vals + v] = sum over the entries (block b, local cell lc) of cell u's list, in
 * order, of part[((r * n_blocks + b) * cols + lc * terms + m) * vals + v]. csr_off: (kinds, ng + 1) list offsets
 * (from csr_base[kind]); csr: (entries, 2); n_blocks = n_pad / 256. */
int hb_nufft_grid(const float* part, int32_t n_blocks, int32_t runs_max, int32_t cols, int32_t n_tiles,
                  int32_t n_slots, const int32_t* run_of, const int32_t* kind_of, int32_t ng, int32_t terms,
                  const int32_t* csr_off, const int32_t* csr, const int32_t* csr_base, double* grid, void* stream);

/* The adjoint of hb_nufft_signal for every run (the reference's NUFFT readouts are forward only): for run r < n_runs,
 * term m, cell u < ng and coil c < nc, G[((r * terms + m) * ng + u) * 2 nc + 2 c + {re, im}] = sum over samples
 * i < n of lw[m][i] conj(tw[i][u]) h[(row_of[r] + i) * nc + c] (h: (n_adc, nc, 2) dL/dsignal turned by e^{+i psi},
 * as hb_adjoint's), float64 sums stored as float32. */
int hb_nufft_adjoint_grid(const float* h, int32_t nc, const int32_t* row_of, int32_t n_runs, int32_t n, int32_t ng,
                          int32_t terms, const double* tw, const double* lw, float* G, void* stream);

/* The NUFFT readouts' samples from the grids (replaces driver.py:703-725's FFT and recombination): for slot r,
 * sample i < n and coil c, sig[(row_of[run] + i) * n_coils + c] = exp(-i psi) sum_m lw[m][i] sum_u
 * grid[r][m][u][c] tw[i][u], float64. tw: (n, ng, 2) e^{-2 pi i j u / ng} at the modes j = i - n / 2; lw: (terms, n)
 * Lagrange weights over the kernel's transform; row_of: (runs,) each run's first signal row. */
int hb_nufft_signal(const double* grid, int32_t n_slots, const int32_t* run_of, const int32_t* row_of, int32_t n,
                    int32_t ng, int32_t terms, int32_t n_tiles, int32_t n_coils, const double* tw, const double* lw,
                    const double* psi, double* sig, void* stream);

/* Per-spin affine maps of reused pulses (replaces pulse_maps, cuda/pulse.cu.in:3-32, launched at driver.py:600-613):
 * pulse p's K substeps composed in float64 into m -> m + D m + v; maps[(p * 12 + r) * n_pad + spin] holds D's rows
 * (r = 0..8, D[i][k] at 3 i + k) and v (r = 9..11). table: (rows, 8) float64 dt, rf_x, rf_y, g_x, g_y, g_z,
 * b0_offset, 0 of every pulse's first occurrence; index: (n_pulses, 2) its (first row, length). spin_list: NULL
 * (every spin), or the n_list spins (< n_pad) to compose, one lane each (the others' maps are left as they are:
 * the surrogate's judge and fallback compose a few spins without paying for the rest). */
int hb_pulse_maps(const hb_case* cs, const double* table, const int32_t* index, int32_t n_pulses,
                  const int32_t* spin_list, int32_t n_list, float* maps, void* stream);

/* Gradients through reused pulses' maps (no counterpart in the reference, whose reused pulses are forward only,
 * driver.py:221-225). A map's dependence on the spin's parameters (px, py, pz, db0, pd, r1, r2, b1) is taken by
 * central differences of the float64 composition of hb_pulse_maps (steps: (n_pulses, 8) float64 step per parameter
 * and pulse, chosen by the caller so that the pulse's rotation moves by ~1e-4 rad: truncation ~1e-9 relative).
 * hb_pulse_map_grads: gs[a * n_pad + spin] += sum over pulses p of <seeds_p, d map_p / d theta_a>, seeds: (n_pulses,
 * 12, n_pad) float64 in the maps' row order (the adjoint's sums of lambda m^T over D's entries and of lambda over
 * v's, in the pulse's frame); gs as hb_adjoint's. hb_pulse_map_dlnb1: maps_d (the maps' layout) = b1 d map / d b1
 * (the derivative along the pulse's amplitude: a pulse scaled by (1 + e) maps as one with b1 (1 + e)), relative
 * step rel. */
int hb_pulse_map_grads(const hb_case* cs, const double* table, const int32_t* index, int32_t n_pulses,
                       const double* steps, const double* seeds, double* gs, void* stream);
int hb_pulse_map_dlnb1(const hb_case* cs, const double* table, const int32_t* index, int32_t n_pulses, double rel,
                       float* maps_d, void* stream);

/* ---- The learned RF surrogate (no counterpart in the reference: it stands where the reference composes a reused
 * pulse per spin, cuda/pulse.cu.in:3-32). A pulse is cut into HB_SUR_BLOCKS blocks; a network
 * HB_MLP_INPUTS -> 64 -> 64 -> 4 M (tanh) maps a spin's offset during the pulse f = df + g . r [Hz] and its B1 scale
 * times the pulse's peak b [Hz] to the blocks' rotations as unit quaternions (w, x, y, z) in the frame turning with
 * the off-resonance precession; the spin's map is composed from them in float64 with the relaxation put back half a
 * block before and after each rotation. The network's inputs are features of (f, b): (f - c_f) / h_f,
 * (b - c_b) / h_b, and sin, cos of 2 pi f q S for q = 1 .. HB_MLP_FOURIER (S = the mean block length over
 * 2 HB_MLP_FOURIER: frequencies inside a block's half-width, the terms a block's rotation is made of to first order
 * in the RF). The network learns each block's quaternion less its first-order part: the small-tip rotation
 * sum_k Rz(-phi_k) w_k (w_k = -2 pi dt_k b rf_k the substeps' transverse rotation vectors, phi_k the precession
 * angle from the block's middle to substep k's), linear in b and, per unit b, a function of f alone, tabulated per
 * network over its box's f range (HB_MLP_FIRST_GRID points) and interpolated linearly. The relaxation, put back
 * half a block before and after each rotation, gets that split's correction (first order in the rates, second in the
 * RF) from the blocks' time moments H_j = sum_k s_k Rz(-phi_k) w_k (s_k the substep's time from the block's middle)
 * and G_j = int v v^T ds - (tau/2) v(tau) v(tau)^T (v(s) the rotation vector accumulated up to s), tabulated beside
 * F_j.
 * Float32 parameters, float32 arithmetic. */
#define HB_MLP_HIDDEN 64
#define HB_SUR_BLOCKS 10
#define HB_MLP_FOURIER 4
#define HB_MLP_INPUTS (2 + 2 * HB_MLP_FOURIER)
#define HB_MLP_PARAMS (HB_MLP_HIDDEN * HB_MLP_INPUTS + HB_MLP_HIDDEN + HB_MLP_HIDDEN * HB_MLP_HIDDEN + HB_MLP_HIDDEN + \
                       4 * HB_SUR_BLOCKS * HB_MLP_HIDDEN + 4 * HB_SUR_BLOCKS)
#define HB_MLP_NORM_HEAD 6
#define HB_MLP_FIRST_GRID 2049

/* A network. params: (HB_MLP_PARAMS,) W1 (64, HB_MLP_INPUTS), b1 (64), W2 (64, 64), b2 (64), W3 (n_out, 64),
 * b3 (n_out), 16-byte aligned for the training entry points (hb_mlp_grad, hb_mlp_train_step(s)); adam_m, adam_v: Adam's moments (training only); norm: (HB_MLP_NORM_HEAD + 2 n_out,) c_f, h_f, c_b,
 * h_b (the input box's centres and half-widths), S [s] (the features' base time), s_b (an input b times s_b is b in
 * Hz: 1 for the training inputs, the pulse's peak for a spin's B1 scale), then the outputs' means (n_out) and scales
 * (n_out): the network is trained on (y - mean) / scale; first: (HB_SUR_BLOCKS, n_grid, 8) the blocks' first-order
 * rotation vectors F_j (x, y) and first time moments H_j (x, y) per Hz of b and second moments G_j (xx, xy, yy) per
 * Hz^2 of b, then 0, at n_grid offsets spaced evenly over [c_f - h_f, c_f + h_f] (hb_sur_first_order), or NULL for a
 * network without a first-order part (and no split correction). hidden must be HB_MLP_HIDDEN and n_out 4 HB_SUR_BLOCKS. */
typedef struct hb_mlp {
    float* params;
    float* adam_m;
    float* adam_v;
    const float* norm;
    int32_t hidden, n_out;
    const float* first;
    int32_t n_grid, reserved; /* reserved: 0 */
} hb_mlp;

/* One reused pulse for hb_surrogate_maps: blocks: (HB_SUR_BLOCKS, 2) float64 each block's length tau [s] and
 * b0_offset area [Hz s]; g: the pulse's (constant) gradient [Hz/m]; pid: the pulse's index in maps. */
typedef struct hb_sur_pulse {
    const double* blocks;
    double g[3];
    int32_t pid, pad;
} hb_sur_pulse;

/* Training data: for n inputs x (n, 2) (f [Hz], b [Hz]: the RF columns of table are multiplied by b), the exact
 * interaction-frame block quaternions y (n, 4 M), products of the substeps' rotations in float64. table as
 * hb_pulse_maps' (one pulse, rows from 0); bounds: (M + 1,) the blocks' first substeps and the end. */
int hb_spinor_blocks(const double* table, const int32_t* bounds, const float* x, int32_t n, float* y, void* stream);

/* The first-order table of a pulse (table, bounds as hb_spinor_blocks'): first[(j * n_grid + i) * 8 + {0..7}]
 * block j's F_j (x, y), H_j (x, y), G_j (xx, xy, yy) and 0 at offset f_lo + i df, sums over its substeps in
 * float64. */
int hb_sur_first_order(const double* table, const int32_t* bounds, double f_lo, double df, int32_t n_grid,
                       float* first, void* stream);

/* q (n, 4 M): the first-order block quaternions exp(b F_j(f) / 2) (w, x, y, 0) at inputs x (n, 2) (f, b in the
 * network's units), F_j interpolated in net->first. */
int hb_sur_first_quats(const hb_mlp* net, const float* x, int32_t n, float* q, void* stream);

/* Floats of workspace hb_mlp_grad and hb_mlp_train_step need for a batch of n. */
int64_t hb_mlp_workspace(int32_t n);

/* grad (HB_MLP_PARAMS,): the gradient of the mean squared error between the network's outputs on the features of
 * x (n, 2) and (y - mean) / scale, over the n samples and n_out outputs; the sums of 64-sample groups added in a fixed
 * order. */
int hb_mlp_grad(const hb_mlp* net, const float* x, const float* y, int32_t n, float* work, float* grad, void* stream);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected for step = 1, 2, ...) on that gradient. */
int hb_mlp_train_step(const hb_mlp* net, const float* x, const float* y, int32_t n, float* work, int32_t step,
                      float lr, void* stream);

/* n_steps Adam steps on consecutive batches of x and y (step i reads rows [i batch, (i + 1) batch)), the step size
 * on a cosine from lr0 to lr1 over total_steps (these being steps step0 + 1 .. step0 + n_steps). */
int hb_mlp_train_steps(const hb_mlp* net, const float* x, const float* y, int32_t batch, int32_t n_steps, float* work,
                       int32_t step0, int32_t total_steps, float lr0, float lr1, void* stream);

/* y (n, n_out): the network's de-normalised outputs mean + scale * net(features of x (n, 2)) (without the
 * first-order part). */
int hb_mlp_eval(const hb_mlp* net, const float* x, int32_t n, float* y, void* stream);

/* Maps of pulse sp->pid for every spin from the network, its block quaternions the first-order part plus the
 * network's outputs, normalised (the maps format of hb_pulse_maps), and outside[spin] = 1
 * for spins (< n) whose inputs lie outside the trained box (their maps are to be composed exactly, hb_pulse_maps
 * with their list), else 0. */
int hb_surrogate_maps(const hb_case* cs, const hb_mlp* net, const hb_sur_pulse* sp, float* maps, int32_t* outside,
                      void* stream);

/* One interval of the adjoint (replaces bloch_replay + bloch_adjoint_c{NC}, cuda/replay.cu.in:3-33,
 * cuda/adjoint.cu.in:5-121; one call per interval of adjoint.py:179-220, last interval first). ckpt: (6, n_pad)
 * the state before the interval (the forward's, hi + lo); r8: scratch of (ceil(k_max / 8), 3, n_pad) floats;
 * lam: (3, n_pad) dL/dm after the interval in, before it out; gs: (8, n_pad) float64 accumulators of dL/d(px, py, pz,
 * db0, pd, r1, r2, b1); gc: (2 nc, n_pad) float64 accumulators of dL/d(Re S_c, Im S_c) (when coil_grads); gseg:
 * (n_pad / 256, k_max, 8) per-workgroup sums of dL/d(dt, rf_x, rf_y, g_x, g_y, g_z, b0_offset, -) of the interval's
 * segments (seg_grads 1; seg_grads 2: only the rf_x, rf_y slots of RF and HB_PULSE segments are wanted, every other
 * value is written as 0 without being formed); h: (n_adc, nc, 2) dL/d(signal) turned by exp(+i psi); sens: (2 nc,
 * n_pad) Re S_c,
 * Im S_c. nc in {0, 1, 2, 4, 8, 16} (coils padded with zeros). The interval length is at most k_max.
 * Reused pulses (HB_PULSE records, cs->maps; HB_PULSE_BODY records are the identity): lambda goes back through the
 * occurrence's map, seeds (n_pulses, 12, n_pad) float64 accumulate per spin, in the pulse's frame (turned by -psi),
 * lambda m^T (rows 0..8, as D's entries) and lambda (rows 9..11, as v's) with lambda after the map and m before it
 * (their parameter gradients follow by hb_pulse_map_grads); with seg_grads, the occurrence's first segment's rf_x,
 * rf_y slots get the sums over spins of dL/d(ln amplitude) (from maps_d, hb_pulse_map_dlnb1) and dL/d(RF phase) of
 * the occurrence, its other slots 0. maps_d and seeds may be NULL when the table has no HB_PULSE record.
 * NUFFT readout runs (HB_NUFFT records: pad[2] the run's kind, pad[3] its length L, pad[4] its index r among the
 * runs; its other segments HB_PULSE_BODY): the run's jump reversed, then the transpose of its spreading: each spin
 * interpolates nu_grid[((r * terms + m) * ng + u) * 2 nc + 2 c + {re, im}] (hb_nufft_adjoint_grid) over its
 * workgroup's window of cells (nu_win, as hb_nufft's win) with the forward's kernel and decay terms (nu_params, as
 * hb_nufft's params), for lambda, the coil gradients and the R2, position and db0 gradients; no segment gradients
 * (the caller asks for none of a run's dt, gradient or b0_offset). nu_params, nu_win, nu_grid: NULL without runs;
 * nu_params and nu_win (and nu_grid when nc > 0) must be set when the table has HB_NUFFT records. */
typedef struct hb_adjoint {
    const float* ckpt;
    float* r8;
    float* lam;
    double* gs;
    double* gc;
    float* gseg;
    const float* h;
    const float* sens;
    const float* maps_d;
    double* seeds;
    const float* nu_params;
    const int32_t* nu_win;
    const float* nu_grid;
    int32_t nc, seg_grads, coil_grads, k_max;
} hb_adjoint;

int hb_adjoint_interval(const hb_case* cs, const hb_adjoint* ad, int32_t s0, int32_t s1, void* stream);

/* Fixed-order sum of an interval's per-workgroup segment gradients (replaces seg_reduce, cuda/seg_reduce.cu:3-11):
 * seg[(s0 + k) * 8 + g] = sum over b < n_blocks of gseg[(b * k_max + k) * 8 + g], float64, for k < len (the
 * workgroups in chunks of 256, each summed in order, then the chunks in order). work:
 * hb_seg_reduce_workspace(n_blocks, k_max) doubles. segs: the case's records after hb_adjoint_interval (whose
 * HB_PULSE_BODY segments sum to 0 without their gseg rows, which that kernel does not write), or NULL after
 * hb_adjoint_full (every row written). */
int64_t hb_seg_reduce_workspace(int32_t n_blocks, int32_t k_max);
int hb_seg_reduce(const float* gseg, int32_t n_blocks, int32_t k_max, int32_t s0, int32_t len, double* work,
                  double* seg, const hb_segment* segs, void* stream);

/* Downstream of the signal (replaces the torch FFT calls of recon.py:17-38 for complex64 data on the GPU).
 * hb_ifftc_axis: src viewed as (outer, L, inner) complex64 values (re, im interleaved), into dst (which may be src)
 * along the middle axis: y[m] = (-1)^(m + L/2) / L * sum_l (-1)^l x[l] e^(2 pi i l m / L) = fftshift(ifft(ifftshift(x))) for even L, the
 * centred inverse DFT with the backward normalisation (recon.py:17-22 ifftnc, one call per axis). L a power of two,
 * 2..4096; outer >= 0, inner >= 1.
 * hb_rss: out[v] = sqrt(sum_c |x[c][v]|^2) for x (n_coils, per_coil) complex64, the coils summed in order in float32
 * (recon.py:36-37 rss over dim 0). */
int hb_ifftc_axis(const float* src, float* dst, int64_t outer, int32_t L, int64_t inner, void* stream);
int hb_rss(const float* x, int32_t n_coils, int64_t per_coil, float* out, void* stream);

/* Name of a return code. */
const char* hb_error_string(int code);

/* ABI version (bumped on any change of the structs or signatures above). */
int hb_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif
