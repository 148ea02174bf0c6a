// The full variant's adjoint: the reverse sweep of one checkpoint interval with a rate bound pool, a scanner's
// concomitant fields and / or bound-only RF. Replaces the reference's V = 1 replay and adjoint kernels
// (src/blochparty/engines/gpu/shaders/adjoint_common.metal:66-203, driven per interval by engines/gpu/adjoint.py:
// 179-220).
//
// One lane per isochromat, float64 throughout (the physics and its reverse mode in full_device.h):
//   1. replay the interval [s0, s1) from its checkpoint, the state before every segment into traj (float32, 16 B per
//      spin and segment);
//   2. reverse, last segment first: zero the transverse adjoint at a spoil, add an ADC sample's source
//      sum_c h_c conj(S_c) and its coil gradients h_c conj(Mxy), then reverse the step from its state before
//      (recomputed intermediates), its spin gradients summed in float64 registers;
//   3. per-segment gradients (dt, rf, g, b0_offset, the RF frequency): summed over each wave by a float64
//      reduce-scatter and over the workgroup's 4 waves in LDS in a fixed order, one partial per workgroup and segment,
//      for hb_seg_reduce.
// Factors that depend only on a segment's length (decays, the unsaturated exchange) are kept while consecutive segments
// of the kind have the same dt, in both sweeps.
// Coil gradients sit in LDS for the launch (each lane its own slots), read from and written back to HBM once.
#include "bloch_device.h"
#include "full_device.h"

#include "../../include/blochparty_amd_mt.h"

namespace hb {
namespace {

constexpr int kBlockF = 256;
constexpr int kWavesF = kBlockF / 64;
constexpr int kMaxCoilsF = 16;
static_assert(full::N_G == HB_FULL_GRADS, "per-spin gradient rows");

// The wave's sums of the N_S = 8 per-segment gradients in a fixed order: over lane bits 5, 4, 3 each step keeps half
// the values (a lane with the bit set the upper half) and adds the partner's copy of them, then lane bits 2, 1, 0 by a
// butterfly; gradient i ends on lanes 8 i .. 8 i + 7 (10 shuffles instead of 48).
__device__ __forceinline__ double seg_scatter_sums(double (&v)[full::N_S], int lane) {
    static_assert(full::N_S == 8, "three halving steps");
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const bool up = lane & 32;
        const double send = up ? v[k] : v[k + 4], keep = up ? v[k + 4] : v[k];
        v[k] = keep + __shfl_xor(send, 32, 64);
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const bool up = lane & 16;
        const double send = up ? v[k] : v[k + 2], keep = up ? v[k + 2] : v[k];
        v[k] = keep + __shfl_xor(send, 16, 64);
    }
    {
        const bool up = lane & 8;
        const double send = up ? v[0] : v[1], keep = up ? v[1] : v[0];
        v[0] = keep + __shfl_xor(send, 8, 64);
    }
    double r = v[0];
#pragma unroll
    for (int off = 4; off > 0; off >>= 1) r += __shfl_xor(r, off, 64);
    return r;
}

__device__ __forceinline__ full::Seg seg_of(const hb_segment& r, const hb_mt& mt, int s) {
    full::Seg q;
    q.dt = r.dt64;
    q.rfx = r.rf_x, q.rfy = r.rf_y;
    q.gx = r.g_x, q.gy = r.g_y, q.gz = r.g_z, q.b0 = r.b0;
    q.frf = mt.pool ? (double)mt.rf_freq[s] : 0.0;
    q.mx = r.m_x, q.my = r.m_y, q.mz = r.m_z, q.b0a = r.b0_area;
    q.flags = r.flags;
    return q;
}

// The read-only inputs come in again as noalias parameters (the same pointers as in the structs): no store of the
// kernel can change what they point to, so the records and the h rows load through the scalar cache.
template <bool SEG>
__global__ __launch_bounds__(kBlockF) void adjoint_full_kernel(
    hb_case cs_in, hb_mt mt_in, hb_full_adjoint ad_in, int s0, int s1, const hb_segment* __restrict__ segs,
    const float* __restrict__ spins, const float* __restrict__ pos_lo, const float* __restrict__ pool_p,
    const float* __restrict__ rf_freq, const float* __restrict__ table, const float* __restrict__ conc,
    const float* __restrict__ ckpt, const float* __restrict__ ckpt_mb, const float* __restrict__ h,
    const float* __restrict__ sens) {
    hb_case cs = cs_in;
    cs.segs = segs, cs.spins = spins, cs.pos_lo = pos_lo;
    hb_mt mt = mt_in;
    mt.pool = pool_p, mt.rf_freq = rf_freq, mt.table = table, mt.conc = conc;
    hb_full_adjoint ad = ad_in;
    ad.ckpt = ckpt, ad.ckpt_mb = ckpt_mb, ad.h = h, ad.sens = sens;
    __shared__ double gcs[2 * kMaxCoilsF * kBlockF];       // coil gradients: slot a of lane tid at a * kBlockF + tid
    __shared__ double red[2][kWavesF][full::N_S];          // alternate buffers: one barrier per segment
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int j = blockIdx.x * kBlockF + tid;
    const int np = cs.n_pad;
    const int nc = ad.nc;
    const bool coil = nc > 0 && ad.coil_grads;  // uniform
    const bool pool = mt.pool != nullptr;       // uniform

    const Spin sp = load_spin(cs.spins, cs.pos_lo, np, j);
    full::Params p;
    p.px = (double)sp.px + (double)sp.lx, p.py = (double)sp.py + (double)sp.ly, p.pz = (double)sp.pz + (double)sp.lz;
    p.pd = sp.pd, p.r1 = sp.r1, p.r2 = sp.r2, p.db0 = sp.db0, p.b1 = sp.b1;
    p.f = pool ? mt.pool[j] : 0.0, p.k = pool ? mt.pool[np + j] : 0.0;
    p.r1b = pool ? mt.pool[2 * np + j] : 1.0, p.t2b = pool ? mt.pool[3 * np + j] : 1.0;
    p.dw = pool ? mt.pool[4 * np + j] : 0.0;
#pragma unroll
    for (int i = 0; i < 7; ++i) p.cc[i] = mt.conc_mode != HB_CONC_NONE ? (double)mt.conc[(size_t)i * np + j] : 0.0;
    full::Model md;
    md.pool = pool;
    md.conc = mt.conc_mode;
    md.line = {mt.table, mt.lx0, mt.dlx, mt.cutoff, mt.n_table, mt.lineshape};
    if (coil) {
        for (int c = 0; c < 2 * nc; ++c) gcs[c * kBlockF + tid] = ad.gc[(size_t)c * np + j];
    }

    // 1. replay
    {
        full::Cache ck = full::no_cache();
        full::St m;
        m.x = (double)ad.ckpt[j] + (double)ad.ckpt[3 * np + j];
        m.y = (double)ad.ckpt[np + j] + (double)ad.ckpt[4 * np + j];
        m.z = (double)ad.ckpt[2 * np + j] + (double)ad.ckpt[5 * np + j];
        m.b = pool ? (double)ad.ckpt_mb[j] + (double)ad.ckpt_mb[np + j] : 0.0;
        for (int s = s0; s < s1; ++s) {
            float* tr = ad.traj + (size_t)(s - s0) * 4 * np;
            tr[j] = (float)m.x, tr[np + j] = (float)m.y, tr[2 * np + j] = (float)m.z, tr[3 * np + j] = (float)m.b;
            const hb_segment r = cs.segs[s];
            m = full::step(md, p, seg_of(r, mt, s), m, ck).after;
            if (r.flags & HB_SPOIL) m.x = m.y = 0.0;
        }
    }

    // 2. reverse sweep
    full::St l = {ad.lam[j], ad.lam[np + j], ad.lam[2 * np + j], ad.lam[3 * np + j]};
    double g[full::N_G];
#pragma unroll
    for (int a = 0; a < full::N_G; ++a) g[a] = 0.0;
    int rbuf = 0;
    full::Cache ck = full::no_cache();
    for (int s = s1 - 1; s >= s0; --s) {
        const hb_segment r = cs.segs[s];
        const full::Seg q = seg_of(r, mt, s);
        const float* tr = ad.traj + (size_t)(s - s0) * 4 * np;
        const full::St m0 = {tr[j], tr[np + j], tr[2 * np + j], tr[3 * np + j]};
        const full::Step st = full::step(md, p, q, m0, ck);
        if (r.flags & HB_SPOIL) l.x = l.y = 0.0;
        if (nc > 0 && (r.flags & HB_ADC)) {
            const float* hrow = ad.h + (size_t)r.pad[0] * 2 * nc;
            const double ax = st.after.x, ay = st.after.y;
            for (int c = 0; c < nc; ++c) {
                const double hr = hrow[2 * c], hi = hrow[2 * c + 1];
                const double sr = ad.sens[(size_t)(2 * c) * np + j], si = ad.sens[(size_t)(2 * c + 1) * np + j];
                l.x += hr * sr + hi * si;
                l.y += hi * sr - hr * si;
                if (coil) {
                    gcs[(2 * c) * kBlockF + tid] += hr * ax + hi * ay;
                    gcs[(2 * c + 1) * kBlockF + tid] += hi * ax - hr * ay;
                }
            }
        }
        double sg[full::N_S];
#pragma unroll
        for (int a = 0; a < full::N_S; ++a) sg[a] = 0.0;
        full::step_reverse(md, p, q, m0, st, l, g, sg);
        if (SEG) {  // the workgroup's sums of this segment's gradients (uniform: every lane takes it)
            const double v = seg_scatter_sums(sg, lane);
            if ((lane & 7) == 0) red[rbuf][wave][lane >> 3] = v;
            // the other buffer than the last segment's: a wave writes this one again only after every wave has
            // passed the next segment's barrier, i.e. after the reads below
            __syncthreads();
            if (tid < full::N_S) {
                double t = red[rbuf][0][tid];
#pragma unroll
                for (int w = 1; w < kWavesF; ++w) t += red[rbuf][w][tid];
                ad.gseg[((size_t)blockIdx.x * ad.k_max + (s - s0)) * 8 + tid] = (float)t;
            }
            rbuf ^= 1;
        }
    }
    ad.lam[j] = (float)l.x, ad.lam[np + j] = (float)l.y, ad.lam[2 * np + j] = (float)l.z;
    ad.lam[3 * np + j] = (float)l.b;
#pragma unroll
    for (int a = 0; a < full::N_G; ++a) ad.gs[(size_t)a * np + j] += g[a];
    if (coil) {
        for (int c = 0; c < 2 * nc; ++c) ad.gc[(size_t)c * np + j] = gcs[c * kBlockF + tid];
    }
}

}  // namespace
}  // namespace hb

extern "C" int hb_adjoint_full(const hb_case* cs, const hb_mt* mt, const hb_full_adjoint* ad, int32_t s0, int32_t s1,
                               void* stream) {
    using namespace hb;
    if (!cs || !mt || !ad || !ad->ckpt || !ad->traj || !ad->lam || !ad->gs) return hipErrorInvalidValue;
    if (cs->n_pad <= 0 || cs->n_pad % kBlockF || cs->n < 0 || cs->n > cs->n_pad) return hipErrorInvalidValue;
    if (s0 < 0 || s1 < s0 || s1 > cs->n_segments || s1 - s0 > ad->k_max) return hipErrorInvalidValue;
    if (ad->nc < 0 || ad->nc > kMaxCoilsF || (ad->nc > 0 && (!ad->h || !ad->sens))) return hipErrorInvalidValue;
    if (ad->coil_grads && ad->nc > 0 && !ad->gc) return hipErrorInvalidValue;
    if (ad->seg_grads && !ad->gseg) return hipErrorInvalidValue;
    if (mt->pool && (!mt->rf_freq || !ad->ckpt_mb)) return hipErrorInvalidValue;
    if (mt->pool && mt->lineshape == HB_LINE_SUPERLORENTZIAN && (!mt->table || mt->n_table < 2 || !(mt->dlx > 0.0f)))
        return hipErrorInvalidValue;
    if (mt->pool && mt->lineshape != HB_LINE_SUPERLORENTZIAN && mt->lineshape != HB_LINE_LORENTZIAN)
        return hipErrorInvalidValue;
    if (mt->conc_mode < HB_CONC_NONE || mt->conc_mode > HB_CONC_EXACT || (mt->conc_mode && !mt->conc))
        return hipErrorInvalidValue;
    if (s1 == s0) return hipSuccess;
    const dim3 grid(cs->n_pad / kBlockF), block(kBlockF);
    hipLaunchKernelGGL(ad->seg_grads ? adjoint_full_kernel<true> : adjoint_full_kernel<false>, grid, block, 0,
                       (hipStream_t)stream, *cs, *mt, *ad, s0, s1, cs->segs, cs->spins, cs->pos_lo, mt->pool,
                       mt->rf_freq, mt->table, mt->conc, ad->ckpt, ad->ckpt_mb, ad->h, ad->sens);
    return hipGetLastError();
}
