// Forward kernels: the event loop with the coil contraction on the matrix cores, and the deterministic cross-block
// signal sum. Replaces the reference's bloch_chunk_mma (src/blochparty/engines/gpu/cuda/kernels.cu.in:2-185) and
// its float atomicAdd flush (cuda/common.cuh:77-117).
//
// Launch: one lane per isochromat, workgroups of 256 lanes (4 waves of 64). Every lane walks the same segments
// [s0, s1), so segment records are uniform loads (scalar cache) and the control flow never diverges.
//
// Coil contraction, per wave: each ADC sample leaves Mxy of the wave's 64 spins; eight samples are staged in LDS as
// the columns of B (64 x 16: Re Mxy of samples 0..7, then Im), and D = A B on the matrix cores with A (16 x 64) the
// wave's sensitivities of one tile of 8 coils, rows Re S_0, Im S_0, Re S_1, Im S_1, ..., held in registers in the
// operand order. With the K index of step u at lane group g being spin 16 g + u, lane l supplies
// A[l % 16][16 (l / 16) + u] and B[16 (l / 16) + u][l % 16] and receives D[4 (l / 16) + r][l % 16], r = 0..3. The
// complex sums are then
//     P_c(t) = D[2c][t] - D[2c + 1][8 + t],   Q_c(t) = D[2c][8 + t] + D[2c + 1][t]   (signal before demodulation),
// the partner column 8 lanes away in the same row of 16 (a DPP rotation). The four waves' sums are added in LDS in a
// fixed order and written as this workgroup's partial row; hb_signal_reduce adds the workgroups in float64, in
// order. No atomics: the signal is the same bit for bit on every run.
//
// Precision of the product: for up to two tiles (16 coils) both operands are split into float16 pairs, x = hi + lo
// (x scaled into [-1, 1] by powers of two), and D = Ah Bh + Ah Bl + Al Bh on v_mfma_f32_16x16x32_f16: float16
// products are exact in the float32 accumulator, and the pairs hold each operand to 2^-24 relative (2^-25 absolute
// below 1/4), so D is as close to the exact sum as a float32 product would be, at six 16-cycle instructions per tile
// and flush in place of sixteen 32-cycle v_mfma_f32_16x16x4f32. More tiles use the float32 instruction, with the
// sensitivities read from memory.
#include <mutex>

#include "bloch_device.h"

namespace hb {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef float f2 __attribute__((ext_vector_type(2)));

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kStgStride = 68;  // floats per staged column row: 64 lanes + 4 (spreads the b128 reads over banks)
constexpr int kStgFloats = 16 * kStgStride;
constexpr int kMaxTiles = 16;   // 128 coils: LDS = 4 x 16 x 68 + 4 x 8 x 256 floats = 49 KB

// LDS: every wave's staging block, then two buffers of the waves' coil sums (flushes alternate between them).
__host__ __device__ constexpr int lds_floats(int nt) { return kWaves * kStgFloats + 2 * kWaves * 8 * 16 * nt; }

// x of lane (jj + 8) mod 16 of this lane's row of 16 (DPP row_ror:8).
__device__ __forceinline__ float dpp_ror8(float x) {
    return __uint_as_float(__builtin_amdgcn_update_dpp(0u, __float_as_uint(x), 0x128, 0xf, 0xf, false));
}

// x = hi + lo in float16 for eight floats: hi rounded to nearest, lo the rounded remainder (exact in float32). For
// |x| <= 1 the pair holds x to 2^-24 |x| or 2^-25, whichever is larger (lo may be subnormal: the matrix cores take
// float16 subnormals as they are).
__device__ __forceinline__ void split8(const float* x, h8& hi, h8& lo) {
#pragma unroll
    for (int e = 0; e < 8; e += 2) {  // in pairs: one v_cvt_pk_f16_f32 per pair of hi and of lo parts
        const h2 h = __builtin_convertvector((f2){x[e], x[e + 1]}, h2);
        // x - float(h) in one v_fma_mix_f32 per value, the float16 operand taken from its half of the packed pair
        // (the compiler forms a conversion and a subtraction); exact, float16 subnormals included
        const unsigned hu = __builtin_bit_cast(unsigned, h);
        float r0, r1;
        asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r0) : "v"(hu), "v"(x[e]));
        asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r1) : "v"(hu), "v"(x[e + 1]));
        const h2 l = __builtin_convertvector((f2){r0, r1}, h2);
        hi[e] = h[0], hi[e + 1] = h[1];
        lo[e] = l[0], lo[e + 1] = l[1];
    }
}

// The wave's sensitivities as matrix-core operands. NT > 0: each tile's 16 values per lane (spins 16 g + u,
// u = 0..15, of row l % 16) scaled by the case's sens_scale and split into float16 hi and lo parts, as the A operands
// of two v_mfma_f32_16x16x32_f16 (u = 0..7 and 8..15); NT = 0: read from memory in float32 at every flush.
// out_scale undoes the sensitivities' and the staged values' scales (powers of two: exact).
template <int NT>
struct Sens {
    h8 hi[NT > 0 ? NT : 1][2], lo[NT > 0 ? NT : 1][2];
    const float* sa;
    float x_scale, out_scale;
};

// The coil sums of the nb (<= 8) staged samples of every wave, added over the workgroup, into the partial rows
// [row, row + nb) of this workgroup. Every lane of the workgroup calls it. One barrier: a wave reads only its own
// staging block (a wave-scope fence orders its lanes' writes before the reads), and the waves' sums alternate
// between two buffers, so a flush's writes never meet the previous flush's reads (every wave has passed this
// flush's barrier, after those reads, before it can write the buffer again two flushes later).
template <int NT>
__device__ __forceinline__ void flush_to(float* stg, float* wred2, int& par, const Sens<NT>& S, int nt,
                                         int n_out_cols, float* __restrict__ out, int k_first = 8,
                                         float* __restrict__ out_rest = nullptr, int k_skip = 0) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, jj = lane & 15;
    const int vals = 16 * nt;
    float* wred = wred2 + par * (kWaves * 8 * vals);
    par ^= 1;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float b[16];
    {
        const f4* src = reinterpret_cast<const f4*>(stg + jj * kStgStride + 16 * g);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const f4 x = src[v];
            b[4 * v] = x[0], b[4 * v + 1] = x[1], b[4 * v + 2] = x[2], b[4 * v + 3] = x[3];
        }
    }
    float* mine = wred + wave * 8 * vals;
    h8 bh[2], bl[2];
    if constexpr (NT > 0) {
        split8(b, bh[0], bl[0]);
        split8(b + 8, bh[1], bl[1]);
    }
    for (int t = 0; t < nt; ++t) {
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        if constexpr (NT > 0) {
            // A B = Ah Bh + Ah Bl + Al Bh (+ Al Bl, under 2^-22 of the product: dropped), small terms first; each
            // product of float16 values is exact in the float32 accumulator
#pragma unroll
            for (int i = 0; i < 2; ++i) acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(S.lo[t][i], bh[i], acc, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 2; ++i) acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(S.hi[t][i], bl[i], acc, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 2; ++i) acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(S.hi[t][i], bh[i], acc, 0, 0, 0);
        } else {
            const float* at = S.sa + t * 1024;
#pragma unroll
            for (int u = 0; u < 16; ++u) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(at[u * 64], b[u], acc, 0, 0, 0);
        }
        // rows 4 g .. 4 g + 3 of lane (g, jj) are Re S_c0, Im S_c0, Re S_c1, Im S_c1 (c0 = 2 g, c1 = 2 g + 1) at
        // column jj; the partner column jj +- 8 sits in the same row of 16 lanes: DPP row_ror:8. Lanes jj < 8 form
        // coil c0's (P, Q) at sample jj, lanes jj >= 8 coil c1's at sample jj - 8.
        const float r0 = dpp_ror8(acc[0]), r1 = dpp_ror8(acc[1]), r2 = dpp_ror8(acc[2]), r3 = dpp_ror8(acc[3]);
        const bool first = jj < 8;
        const f2 pq = first ? (f2){acc[0] - r1, r0 + acc[1]} : (f2){r2 - acc[3], acc[2] + r3};
        // this wave's sums in the output order [sample][tile][coil][re, im]: one 8-byte store per lane
        *reinterpret_cast<f2*>(mine + (jj & 7) * vals + 16 * t + 2 * (2 * g + (jj >> 3))) = pq;
    }
    __syncthreads();
    // columns [0, k_first) to out, [k_first, k_first + k_skip) dropped, the rest to out_rest (output value
    // v = sample * vals + 16 tile + 2 coil + {0: re, 1: im})
    const int n_out = n_out_cols * vals, split = k_first * vals, resume = (k_first + k_skip) * vals;
    for (int v = tid; v < n_out; v += kBlock) {
        float s = wred[v];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) s += wred[w * 8 * vals + v];
        s *= S.out_scale;
        if (v < split)
            out[v] = s;
        else if (v >= resume)
            out_rest[v - resume] = s;
    }
}

// The staged samples into the partial rows [row, row + nb) of this workgroup (never past the buffer's rows).
template <int NT>
__device__ __forceinline__ void flush(float* stg, float* wred2, int& par, const Sens<NT>& S, int nt, int nb, int row,
                                      int rows, float* __restrict__ partial) {
    const int n_out = row + nb <= rows ? nb : (rows > row ? rows - row : 0);
    flush_to<NT>(stg, wred2, par, S, nt, n_out, partial + ((size_t)blockIdx.x * rows + row) * (16 * nt));
}

// A flush of nb staged columns of which the first `pending` are samples (rows [row, row + pending) of the partial
// rows, row advanced) and the rest grid columns (from grid_out).
template <int NT>
__device__ __forceinline__ void flush_cols(float* stg, float* wred, int& par, const Sens<NT>& S, int nt, int nb,
                                           int pending, int& row, int rows, float* __restrict__ partial,
                                           float* __restrict__ grid_out) {
    const int vals = 16 * nt;
    if (pending == 0) {
        flush_to<NT>(stg, wred, par, S, nt, nb, grid_out);
        return;
    }
    const int keep = row + pending <= rows ? pending : (rows > row ? rows - row : 0);  // never past the rows
    flush_to<NT>(stg, wred, par, S, nt, nb, partial + ((size_t)blockIdx.x * rows + row) * vals, keep, grid_out,
                 pending - keep);
    row += pending;
}

// One NUFFT readout run of L samples (record q, the run's first segment; pad[1] its slot in the dispatch, pad[2] its
// kind, pad[3] L). The spin's weight (mx + i my) e^{-2 pi i q (L // 2 + 1)} times the kernel psi at each cell of the
// workgroup's window and e^{-R2 tau_m} for each decay term is staged column by column (column = local cell * terms +
// term) and contracted with the coils like samples; the workgroup's window of columns goes to its slot of the grid
// partials. Then the run is one exact free step of L dt (phase L q, relaxation over L dt), and a spoil at its end.
// Samples still staged (nb of them, rows [row, row + nb)) share the first flush with the grid columns.
template <int NT, int NUT>
__device__ __forceinline__ void nufft_run(const hb_case& cs, const hb_nufft& nu, const hb_segment& q, const Spin& p,
                                          const Pos64& r64, const float (&em)[NUT], float& e1_run, State& m,
                                          float* stg, float* wred, int& par, const Sens<NT>& S, int nt, int lane,
                                          int& nb_io, int& row, int rows, float* __restrict__ partial) {
    const int L = q.pad[3], slot = q.pad[1], kind = q.pad[2];
    const int ng = (int)nu.params[0], terms = (int)nu.params[2], shift = (int)nu.params[4];
    const float h = nu.params[6 + HB_NUFFT_MAX_TERMS];       // 2 / width: z per grid cell
    const float bl2 = nu.params[7 + HB_NUFFT_MAX_TERMS];     // beta log2(e)
    const int vals = 16 * nt;
    const double cyc = free_cycles64(r64, p, q);
    const double fr = cyc - floor(cyc);
    const double xg = fr * ng;  // float64: at x ~ ng a float32 position is 1e-5 of a cycle off at the last sample
    double ph = fr * shift;
    ph -= rint(ph);
    float sp, cp;
    sincospi(-2.0f * (float)ph, &sp, &cp);
    const float cr = (m.mx * cp - m.my * sp) * S.x_scale, ci = (m.mx * sp + m.my * cp) * S.x_scale;
    // the weight (times the staged values' scale) times each decay term e^{-R2 tau_m}
    float wr[NUT], wi[NUT];
#pragma unroll
    for (int t = 0; t < NUT; ++t) wr[t] = em[t] * cr, wi[t] = em[t] * ci;
    const int* win = nu.win + ((size_t)kind * gridDim.x + blockIdx.x) * 2;
    const int wb = win[0], W = win[1];
    float* out = nu.part + ((size_t)slot * gridDim.x + blockIdx.x) * nu.cols * vals;
    int nb = nb_io, col = 0, pending = nb_io;  // the staged samples come first in the first flush
    nb_io = 0;
    // z = (cell - x) 2 / width: at the window's first cell from the float64 position (x ~ ng would leave a float32
    // z 1e-6 off), then by float32 steps (|z| error under 3e-7 over a 32-cell window)
    const float z0 = (float)(((double)wb - xg) * (double)h);
    float lcf = 0.0f;  // lc as a float
    for (int lc = 0; lc < W; ++lc, lcf += 1.0f) {
        const float z = fmaf(lcf, h, z0);
        const float u = fmaf(-z, z, 1.0f);
        // the ES kernel e^{beta (sqrt(1 - z^2) - 1)} = 2^{beta log2(e) (sqrt(u) - 1)}, zero outside |z| < 1, on the
        // hardware square root and exp2 (about 1 ulp each; the argument's rounding, |arg| < 20, leaves the
        // kernel value within 2e-6 relative)
        const float kv = u > 0.0f ? __builtin_amdgcn_exp2f(fmaf(bl2, __builtin_amdgcn_sqrtf(u), -bl2)) : 0.0f;
#pragma unroll
        for (int t = 0; t < NUT; ++t) {
            if (t < terms) {
                stg[nb * kStgStride + lane] = kv * wr[t];
                stg[(8 + nb) * kStgStride + lane] = kv * wi[t];
                if (++nb == 8) {
                    flush_cols<NT>(stg, wred, par, S, nt, 8, pending, row, rows, partial, out + (size_t)col * vals);
                    col += 8 - pending;
                    pending = 0;
                    nb = 0;
                }
            }
        }
    }
    if (nb > 0) {
        flush_cols<NT>(stg, wred, par, S, nt, nb, pending, row, rows, partial, out + (size_t)col * vals);
        pending = 0;
    }
    // the jump over the run: phase L cyc, relaxation over L dt (Mz only when the run's Mxy is never read)
    if (!(q.flags & HB_SAME_RUN)) e1_run = expm1x(-((float)L * q.dt) * p.r1);
    run_jump(q, p, cyc, e1_run, m);
}

// NT: coil tiles (0: any, the sensitivities read from memory per flush); NUT: NUFFT runs with up to NUT decay terms
// (0: none; the terms' weights are held in registers); RF: RF segments are stepped (false when every pulse of the
// dispatch is applied as a map: the RF step is then left out, and its registers).
// The read-only inputs and the NUFFT partials come in again as noalias parameters (the same pointers as in cs_in and
// nu_in): no store of the kernel can change the records, so they load through the scalar cache instead of one vector
// load per wave and a readfirstlane per value.
template <int NT, int NUT, bool RF>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(NT == 2 ? 3 : 5))) void forward_kernel(
    hb_case cs_in, int s0, int s1, float* __restrict__ partial, int rows, hb_nufft nu_in,
    const hb_segment* __restrict__ segs, const float* __restrict__ spins, const float* __restrict__ pos_lo,
    const float* __restrict__ sens_a, const float* __restrict__ maps, const float* __restrict__ nu_params,
    const int32_t* __restrict__ nu_win, float* __restrict__ nu_part) {
    hb_case cs = cs_in;
    cs.segs = segs, cs.spins = spins, cs.pos_lo = pos_lo, cs.sens_a = sens_a, cs.maps = maps;
    hb_nufft nu = nu_in;
    nu.params = nu_params, nu.win = nu_win, nu.part = nu_part;
    extern __shared__ float lds[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = blockIdx.x * kBlock + tid;
    const int np = cs.n_pad;
    const int nt = NT > 0 ? NT : cs.n_tiles;
    float* stg = lds + wave * kStgFloats;
    float* wred = lds + kWaves * kStgFloats;
    const Spin p = load_spin(cs.spins, cs.pos_lo, np, j);
    const Pos64 r64 = pos64(p);
    const float b1sq = p.b1 * p.b1;
    State m = load_state(cs.state, np, j);
    Sens<NT> S;
    S.sa = cs.sens_a + (size_t)(blockIdx.x * kWaves + wave) * nt * 1024 + lane;
    S.x_scale = cs.x_scale > 0.0f ? cs.x_scale : 1.0f;
    const float sens_scale = cs.sens_scale > 0.0f ? cs.sens_scale : 1.0f;
    if constexpr (NT > 0) {
        S.out_scale = 1.0f / (S.x_scale * sens_scale);
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                float a[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) a[e] = S.sa[t * 1024 + (8 * i + e) * 64] * sens_scale;
                split8(a, S.hi[t][i], S.lo[t][i]);
            }
    } else {
        S.out_scale = 1.0f / S.x_scale;
    }
    Cache k = {1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float e1_run = 0.0f;  // expm1(-L dt R1) of the last readout run (per sample or NUFFT)
    // the NUFFT decay terms e^{-R2 tau_m}: the spin's, the same for every run
    float em[NUT > 0 ? NUT : 1];
    if constexpr (NUT > 0) {
        const int terms = (int)nu.params[2];
#pragma unroll
        for (int t = 0; t < NUT; ++t)
            em[t] = t < terms ? __builtin_amdgcn_exp2f(-(nu.params[6 + t] * 1.4426950408889634f) * p.r2) : 0.0f;
    }
    int nb = 0, row = 0, par = 0;
    // a dispatch that starts inside a reused pulse's occurrence or a NUFFT run (checkpoint intervals do not avoid
    // them): the record was applied at its first segment, in the dispatch before
    int s_start = s0;
    if (cs.segs[s0].flags & HB_PULSE_BODY) {
        const int head = cs.segs[s0].pad[3];
        const hb_segment& hq = cs.segs[head];
        s_start = head + ((hq.flags & HB_PULSE) ? hq.pad[2] : hq.pad[3]);
    }
    for (int s = s_start; s < s1; ++s) {
        const hb_segment q = cs.segs[s];
        const int f = q.flags;
        if constexpr (NUT > 0) {
            if (f & HB_NUFFT) {  // a readout run as a type-1 NUFFT: spread onto the grid window, jump the run
                nufft_run<NT, NUT>(cs, nu, q, p, r64, em, e1_run, m, stg, wred, par, S, nt, lane, nb, row, rows,
                                   partial);
                s += q.pad[3] - 1;
                continue;
            }
        }
        if (RF && (f & HB_TRAIN)) {  // RF substeps of a shaped pulse, in a loop of their own (only RF state live)
            const int L = q.pad[3];
            RfRec cur = load_rf(&cs.segs[s]);
            for (int i = 1; i <= L; ++i) {
                const RfRec nxt = load_rf(&cs.segs[s + (i < L ? i : 0)]);  // in flight while this one is stepped
                rf_step(cur, cur.flags, p, b1sq, m, k);
                cur = nxt;
            }
            s += L - 1;
            continue;
        }
        if (f & HB_RUN) {  // L identical free ADC segments: Mxy stepped and sampled per segment, Mz relaxed once
            const int L = q.pad[3];
            if (!(f & HB_SAME_FREE)) {
                sincospi(-2.0f * free_cycles(r64, p, q), &k.s, &k.c);
                if (!(f & HB_SAME_FREE_DT)) {
                    k.e1 = expm1x(-q.dt * p.r1);
                    k.e2 = expm1x(-q.dt * p.r2);
                }
            }
            if (!(f & HB_SAME_RUN)) e1_run = expm1x(-(float)L * q.dt * p.r1);
            for (int i = 0; i < L; ++i) {
                const float nx = k.c * m.mx - k.s * m.my, ny = k.s * m.mx + k.c * m.my;
                m.mx = nx + nx * k.e2;
                m.my = ny + ny * k.e2;
                stg[nb * kStgStride + lane] = m.mx * S.x_scale;
                stg[(8 + nb) * kStgStride + lane] = m.my * S.x_scale;
                if (++nb == 8) {
                    flush<NT>(stg, wred, par, S, nt, 8, row, rows, partial);
                    row += 8;
                    nb = 0;
                }
            }
            two_sum(m.mz, m.lz, (m.mz - p.pd) * e1_run);
            s += L - 1;
            continue;
        }
        if (f & HB_PULSE) {  // a reused pulse: m += Rz(psi) (D Rz(-psi) m + v), then skip its segments
            // cos, sin of the occurrence's RF phase (host float64)
            apply_map(cs.maps + (size_t)q.pad[1] * 12 * np + j, np, q.w_rf_x, q.w_rf_y, m);
            if (f & HB_PULSE_SPOIL) m.mx = m.my = 0.0f;
            s += q.pad[2] - 1;
            continue;
        }
        if (f & HB_SPOILED_FREE) {  // only Mz survives the spoil: relax it, zero Mxy
            two_sum(m.mz, m.lz, (m.mz - p.pd) * expm1x(-q.dt * p.r1));
            m.mx = m.my = 0.0f;
            if (f & HB_MERGED) s += q.pad[3] - 1;
            continue;
        }
        if constexpr (RF)
            step(q, f, p, r64, b1sq, m, k);
        else
            free_step(q, f, p, r64, m, k);
        if (f & HB_MERGED) s += q.pad[3] - 1;
        if (f & HB_ADC) {
            stg[nb * kStgStride + lane] = m.mx * S.x_scale;
            stg[(8 + nb) * kStgStride + lane] = m.my * S.x_scale;
            if (++nb == 8) {
                flush<NT>(stg, wred, par, S, nt, 8, row, rows, partial);
                row += 8;
                nb = 0;
            }
        }
        if (f & HB_SPOIL) m.mx = m.my = 0.0f;
    }
    if (nb > 0) flush<NT>(stg, wred, par, S, nt, nb, row, rows, partial);
    store_state(cs.state, np, j, m);
}

// The cross-block sum in two passes, both in a fixed order (deterministic): the blocks in chunks of kRedChunk,
// summed in float64 in block order by one lane per partial value (consecutive lanes read consecutive floats); then
// the chunks in order, demodulated: sig[row0 + r][c] = exp(-i psi) sum_chunks sum_b partial[b][r][c].
constexpr int kRedChunk = 256;

__global__ __launch_bounds__(256) void reduce_chunks_kernel(const float* __restrict__ partial, int n_blocks,
                                                            int per_block, double* __restrict__ work) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= per_block) return;
    const int b0 = blockIdx.y * kRedChunk, b1 = min(n_blocks, b0 + kRedChunk);
    const float* src = partial + i;
    double acc = 0.0;
#pragma unroll 16
    for (int b = b0; b < b1; ++b) acc += (double)src[(size_t)b * per_block];
    work[(size_t)blockIdx.y * per_block + i] = acc;
}

__global__ __launch_bounds__(256) void reduce_final_kernel(const double* __restrict__ work, int chunks, int rows,
                                                           int vals, int n_coils, const double* __restrict__ psi,
                                                           int row0, const int* __restrict__ row_map,
                                                           double* __restrict__ sig) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= rows * n_coils) return;
    const int r = idx / n_coils, c = idx - r * n_coils;
    const size_t stride = (size_t)rows * vals;
    const double* src = work + (size_t)r * vals + 2 * c;
    double re = 0.0, im = 0.0;
    for (int k = 0; k < chunks; ++k) {
        re += src[k * stride];
        im += src[k * stride + 1];
    }
    const int dst = row_map ? row_map[row0 + r] : row0 + r;
    double sp, cp;
    sincos(psi[dst], &sp, &cp);
    double* out = sig + ((size_t)dst * n_coils + c) * 2;
    out[0] = re * cp + im * sp;
    out[1] = im * cp - re * sp;
}

}  // namespace hb

namespace hb {
template <int NUT, bool RF>
static void launch_nt(const hb_case& cs, int s0, int s1, float* partial, int rows, const hb_nufft& nu, dim3 grid,
                      dim3 block, size_t lds, hipStream_t st) {
    auto kernel = cs.n_tiles == 1   ? forward_kernel<1, NUT, RF>
                  : cs.n_tiles == 2 ? forward_kernel<2, NUT, RF>
                                    : forward_kernel<0, NUT, RF>;
    hipLaunchKernelGGL(kernel, grid, block, lds, st, cs, s0, s1, partial, rows, nu, cs.segs, cs.spins, cs.pos_lo,
                       cs.sens_a, cs.maps, nu.params, nu.win, nu.part);
}

template <int NUT>
static void launch_forward(const hb_case& cs, int s0, int s1, float* partial, int rows, const hb_nufft& nu,
                           dim3 grid, dim3 block, size_t lds, hipStream_t st) {
    if (cs.rf_free)
        launch_nt<NUT, false>(cs, s0, s1, partial, rows, nu, grid, block, lds, st);
    else
        launch_nt<NUT, true>(cs, s0, s1, partial, rows, nu, grid, block, lds, st);
}
}  // namespace hb

extern "C" int hb_forward(const hb_case* cs, int32_t s0, int32_t s1, float* partial, int32_t rows,
                          const hb_nufft* nu, void* stream) {
    using namespace hb;
    if (!cs || cs->n_pad <= 0 || cs->n_pad % kBlock || cs->n < 0 || cs->n > cs->n_pad) return hipErrorInvalidValue;
    if (cs->n_tiles < 1 || cs->n_tiles > kMaxTiles || 8 * cs->n_tiles < cs->n_coils) return hipErrorInvalidValue;
    if (s0 < 0 || s1 < s0 || s1 > cs->n_segments || rows < 0) return hipErrorInvalidValue;
    if (cs->n_pulses < 0 || (cs->n_pulses > 0 && !cs->maps)) return hipErrorInvalidValue;
    if (nu && (!nu->params || !nu->win || !nu->part || nu->runs_max < 1 || nu->cols < 1 || nu->terms < 1 ||
               nu->terms > HB_NUFFT_MAX_TERMS))
        return hipErrorInvalidValue;
    if (s1 == s0) return hipSuccess;
    const dim3 grid(cs->n_pad / kBlock), block(kBlock);
    const size_t lds = sizeof(float) * lds_floats(cs->n_tiles);
    hipStream_t st = (hipStream_t)stream;
    const hb_nufft none = {nullptr, nullptr, nullptr, 0, 0, 0, 0};
    const hb_nufft arg = nu ? *nu : none;
    if (!nu)
        launch_forward<0>(*cs, s0, s1, partial, rows, arg, grid, block, lds, st);
    else if (nu->terms <= 4)
        launch_forward<4>(*cs, s0, s1, partial, rows, arg, grid, block, lds, st);
    else
        launch_forward<8>(*cs, s0, s1, partial, rows, arg, grid, block, lds, st);
    return hipGetLastError();
}

extern "C" int64_t hb_signal_reduce_workspace(int32_t n_blocks, int32_t rows, int32_t n_tiles) {
    if (n_blocks < 1 || rows < 0 || n_tiles < 1) return 0;
    return (int64_t)((n_blocks + hb::kRedChunk - 1) / hb::kRedChunk) * rows * 16 * n_tiles;
}

extern "C" int hb_signal_reduce(const float* partial, int32_t n_blocks, int32_t rows, int32_t n_tiles,
                                int32_t n_coils, const double* psi, int32_t row0, const int32_t* row_map,
                                double* work, double* sig, void* stream) {
    using namespace hb;
    if (rows < 0 || n_blocks < 1 || n_tiles < 1 || n_tiles > kMaxTiles || n_coils < 1 || 8 * n_tiles < n_coils ||
        row0 < 0 || !work)
        return hipErrorInvalidValue;
    if (rows == 0) return hipSuccess;
    const int vals = 16 * n_tiles, per_block = rows * vals, chunks = (n_blocks + kRedChunk - 1) / kRedChunk;
    hipLaunchKernelGGL(reduce_chunks_kernel, dim3((per_block + 255) / 256, chunks), dim3(256), 0,
                       (hipStream_t)stream, partial, n_blocks, per_block, work);
    const int total = rows * n_coils;
    hipLaunchKernelGGL(reduce_final_kernel, dim3((total + 255) / 256), dim3(256), 0, (hipStream_t)stream, work, chunks,
                       rows, vals, n_coils, psi, row0, row_map, sig);
    return hipGetLastError();
}

// ---- hb_signal_reduce_noalloc: hb_signal_reduce with its float64 chunk sums in a fixed device array.
//
// `scratch` is a file-scope device array that each GPU the library runs on holds until the process ends. Host code
// never passes its name to a kernel (in host code the name is a host-side stand-in, not a device address): the
// first call on a device resolves the array's address on that device with hipGetSymbolAddress, and every launch
// uses that address. Calls on one GPU share its array; the chunk sums are written in full before they are read,
// so nothing needs clearing, but calls that overlap on one GPU (other streams, other threads) race.
static __device__ __align__(256) double scratch[HB_NOALLOC_SCRATCH_DOUBLES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
double* scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch` (resolved once per device), or nullptr with *err set.
double* device_scratch(hipError_t* err) {
    int dev = 0;
    *err = hipGetDevice(&dev);
    if (*err != hipSuccess) return nullptr;
    if (dev < 0 || dev >= kMaxDevices) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        *err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (*err != hipSuccess) return nullptr;
        scratch_ptr[dev] = static_cast<double*>(p);
    }
    return scratch_ptr[dev];
}
}  // namespace

extern "C" int hb_signal_reduce_noalloc(const float* partial, int32_t n_blocks, int32_t rows, int32_t n_tiles,
                                        int32_t n_coils, const double* psi, int32_t row0, const int32_t* row_map,
                                        double* sig, void* stream) {
    using namespace hb;
    if (rows < 0 || n_blocks < 1 || n_tiles < 1 || n_tiles > kMaxTiles || n_coils < 1 || 8 * n_tiles < n_coils ||
        row0 < 0)
        return hipErrorInvalidValue;
    if (hb_signal_reduce_workspace(n_blocks, rows, n_tiles) > HB_NOALLOC_SCRATCH_DOUBLES)
        return hipErrorInvalidValue;  // beyond the fixed array: use hb_signal_reduce with a workspace
    hipError_t err = hipSuccess;
    double* work = device_scratch(&err);
    if (!work) return err;
    return hb_signal_reduce(partial, n_blocks, rows, n_tiles, n_coils, psi, row0, row_map, work, sig, stream);
}

extern "C" const char* hb_error_string(int code) { return hipGetErrorString((hipError_t)code); }

extern "C" int hb_abi_version(void) { return 25; }
