// NUFFT readouts after the forward kernel's spreading (readouts="nufft"; replaces the reference's float atomics
// into the grid rows and its host-side FFT and recombination, src/blochparty/engines/gpu/driver.py:549-580,
// 703-725).
//
// hb_nufft_grid: every grid cell of every run slot is the sum, in a fixed order, of the window columns of the
// workgroups whose windows hold the cell (the lists are built on the host from the windows): float64, no atomics,
// so the grids are the same bit for bit on every run.
// hb_nufft_signal: the readout's samples are the grids' DFT at the n modes j = i - n / 2 (the modes the samples
// are), divided by the kernel's transform and recombined over the decay terms by their Lagrange weights (folded
// into lw), then demodulated by the sample's receiver phase; float64.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/blochparty_amd.h"

namespace hb {

constexpr int kGridBlock = 256;
constexpr int kBatch = 8;  // list entries per step of a lane's gather

// One workgroup per (cell u, slot r): the cell's list entries (block b, local cell lc) are taken by G groups of
// terms x vals / 4 lanes, group k the entries k, k + G, ... (each entry's terms x vals floats are contiguous in part,
// vals a multiple of 16: one 16-byte load per lane, the group's loads one coalesced read), each lane summing its four
// values in float64; then the G groups' sums are added in order. With 12 lanes per entry (3 terms, 8 coils) 21 groups
// share the cell's ~200 entries, so each lane's chain of dependent list and value loads is ~10 entries long. Fixed
// order throughout: the grid is the same bit for bit on every run.
// Workgroups go round robin over the 8 XCDs, so workgroup i works on item (i mod 8) * ~T/8 + i / 8 of the T (slot,
// cell) items: each XCD takes one contiguous stretch of cells, whose entries share cache lines of part in its L2.
__global__ __launch_bounds__(kGridBlock) void nufft_grid_kernel(const float* __restrict__ part, int n_blocks, int cols,
                                                                int vals, const int* __restrict__ run_of,
                                                                const int* __restrict__ kind_of, int ng, int terms,
                                                                const int* __restrict__ csr_off,
                                                                const int* __restrict__ csr,
                                                                const int* __restrict__ csr_base,
                                                                double* __restrict__ grid) {
    __shared__ double sums[4 * kGridBlock];
    const int T = gridDim.x, xcd = blockIdx.x % 8, per = T / 8, extra = T % 8;
    const int w = xcd * per + min(xcd, extra) + blockIdx.x / 8;
    const int u = w % ng, r = w / ng, tid = threadIdx.x;
    const int vt4 = terms * vals / 4;                          // float4s per entry
    const int span = vt4 >= kGridBlock ? kGridBlock : vt4;     // lanes per entry group
    const int G = kGridBlock / span;                           // entry groups
    const int kind = kind_of[run_of[r]];
    const int* off = csr_off + (size_t)kind * (ng + 1);
    const int* ent = csr + 2 * (size_t)csr_base[kind];
    const int e0 = off[u], e1 = off[u + 1];
    for (int q0 = 0; q0 < vt4; q0 += span) {  // float4 chunks (one unless terms x vals > 1024)
        const int k = tid / span, q = q0 + tid % span;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        if (k < G && q < vt4) {
            // kBatch entries at a time: their list reads, then their value reads, are all in flight together (the
            // sum still takes them in entry order; an entry past the end adds an exact 0)
            for (int e = e0 + k; e < e1; e += kBatch * G) {
                int2 be[kBatch];
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    const int ee = e + j * G;
                    be[j] = ee < e1 ? reinterpret_cast<const int2*>(ent)[ee] : make_int2(-1, 0);
                }
                float4 x[kBatch];
#pragma unroll
                for (int j = 0; j < kBatch; ++j)
                    x[j] = be[j].x >= 0 ? reinterpret_cast<const float4*>(
                                              part + ((size_t)r * n_blocks + be[j].x) * cols * vals +
                                              (size_t)be[j].y * terms * vals)[q]
                                        : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
                for (int j = 0; j < kBatch; ++j) {
                    a0 += (double)x[j].x, a1 += (double)x[j].y;
                    a2 += (double)x[j].z, a3 += (double)x[j].w;
                }
            }
        }
        sums[4 * tid] = a0, sums[4 * tid + 1] = a1, sums[4 * tid + 2] = a2, sums[4 * tid + 3] = a3;
        __syncthreads();
        const int count = 4 * min(span, vt4 - q0);  // values of this chunk
        for (int qq = tid; qq < count; qq += kGridBlock) {
            double t = sums[qq];
            for (int g = 1; g < G; ++g) t += sums[4 * g * span + qq];
            const int qv = 4 * q0 + qq, m = qv / vals, v = qv % vals;
            grid[((size_t)(r * terms + m) * ng + u) * vals + v] = t;
        }
        __syncthreads();
    }
}

constexpr int kSigParts = 32;  // lanes per coil, each summing the cells u = p, p + 32, ...

// One workgroup per (sample i, slot r): lanes (coil c of a group of 8, part p) sum the DFT over their cells and
// every decay term, weighted; the 32 parts are added in order; then the receiver phase is taken off.
__global__ __launch_bounds__(256) void nufft_signal_kernel(const double* __restrict__ grid,
                                                           const int* __restrict__ run_of,
                                                           const int* __restrict__ row_of, int n, int ng, int terms,
                                                           int vals, int n_coils, const double* __restrict__ tw,
                                                           const double* __restrict__ lw,
                                                           const double* __restrict__ psi, double* __restrict__ sig) {
    __shared__ double red[2][256];
    const int i = blockIdx.x, r = blockIdx.y, tid = threadIdx.x;
    const int cl = tid % 8, p = tid / 8;
    const double* twi = tw + (size_t)i * ng * 2;
    const int row = row_of[run_of[r]] + i;
    double sp, cp;
    sincos(psi[row], &sp, &cp);
    for (int c0 = 0; c0 < n_coils; c0 += 8) {
        const int c = c0 + cl;
        double re = 0.0, im = 0.0;
        if (c < n_coils) {
            for (int m = 0; m < terms; ++m) {
                const double* g = grid + ((size_t)(r * terms + m) * ng) * vals + 2 * c;
                double fr = 0.0, fi = 0.0;
                for (int u = p; u < ng; u += kSigParts) {
                    const double gr = g[(size_t)u * vals], gi = g[(size_t)u * vals + 1];
                    const double tr = twi[2 * u], ti = twi[2 * u + 1];
                    fr += gr * tr - gi * ti;
                    fi += gr * ti + gi * tr;
                }
                const double wgt = lw[(size_t)m * n + i];
                re += wgt * fr;
                im += wgt * fi;
            }
        }
        red[0][tid] = re;
        red[1][tid] = im;
        __syncthreads();
        if (tid < 8 && c < n_coils) {
            double sr = 0.0, si = 0.0;
            for (int q = 0; q < kSigParts; ++q) sr += red[0][q * 8 + tid], si += red[1][q * 8 + tid];
            double* out = sig + ((size_t)row * n_coils + c) * 2;
            out[0] = sr * cp + si * sp;
            out[1] = si * cp - sr * sp;
        }
        __syncthreads();
    }
}

// One lane per (cell u, coil c) of run blockIdx.z, term blockIdx.y: the sum over the run's samples in float64.
__global__ __launch_bounds__(256) void nufft_adjoint_grid_kernel(const float* __restrict__ h, int nc,
                                                                 const int* __restrict__ row_of, int n, int ng,
                                                                 int terms, const double* __restrict__ tw,
                                                                 const double* __restrict__ lw,
                                                                 float* __restrict__ G) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ng * nc) return;
    const int u = k / nc, c = k % nc, m = blockIdx.y, r = blockIdx.z;
    const float* hr = h + (size_t)row_of[r] * nc * 2 + 2 * c;
    const double* lm = lw + (size_t)m * n;
    double re = 0.0, im = 0.0;
    for (int i = 0; i < n; ++i) {
        const double a = lm[i], tr = tw[((size_t)i * ng + u) * 2], ti = tw[((size_t)i * ng + u) * 2 + 1];
        const double xr = hr[(size_t)i * nc * 2], xi = hr[(size_t)i * nc * 2 + 1];
        re += a * (tr * xr + ti * xi);  // conj(tw) h
        im += a * (tr * xi - ti * xr);
    }
    float* out = G + (((size_t)r * terms + m) * ng + u) * 2 * nc + 2 * c;
    out[0] = (float)re;
    out[1] = (float)im;
}

}  // namespace hb

extern "C" int hb_nufft_adjoint_grid(const float* h, int32_t nc, const int32_t* row_of, int32_t n_runs, int32_t n,
                                     int32_t ng, int32_t terms, const double* tw, const double* lw, float* G,
                                     void* stream) {
    if (!h || !row_of || !tw || !lw || !G || nc < 1 || n_runs < 0 || n_runs > 65535 || n < 1 || ng < 2 ||
        terms < 1 || terms > HB_NUFFT_MAX_TERMS)
        return hipErrorInvalidValue;
    if (n_runs == 0) return hipSuccess;
    hipLaunchKernelGGL(hb::nufft_adjoint_grid_kernel, dim3((ng * nc + 255) / 256, terms, n_runs), dim3(256), 0,
                       (hipStream_t)stream, h, nc, row_of, n, ng, terms, tw, lw, G);
    return hipGetLastError();
}

extern "C" int hb_nufft_grid(const float* part, int32_t n_blocks, int32_t runs_max, int32_t cols, int32_t n_tiles,
                             int32_t n_slots, const int32_t* run_of, const int32_t* kind_of, int32_t ng,
                             int32_t terms, const int32_t* csr_off, const int32_t* csr, const int32_t* csr_base,
                             double* grid, void* stream) {
    if (!part || !run_of || !kind_of || !csr_off || !csr || !csr_base || !grid || n_blocks < 1 || runs_max < 1 ||
        cols < 1 || n_tiles < 1 || n_slots < 0 || n_slots > runs_max || ng < 2 || terms < 1 ||
        terms > HB_NUFFT_MAX_TERMS || (int64_t)ng * n_slots > INT32_MAX)
        return hipErrorInvalidValue;
    if (n_slots == 0) return hipSuccess;
    hipLaunchKernelGGL(hb::nufft_grid_kernel, dim3(ng * n_slots), dim3(hb::kGridBlock), 0, (hipStream_t)stream, part,
                       n_blocks, cols, 16 * n_tiles, run_of, kind_of, ng, terms, csr_off, csr, csr_base, grid);
    return hipGetLastError();
}

extern "C" int hb_nufft_signal(const double* grid, int32_t n_slots, const int32_t* run_of, const int32_t* row_of,
                               int32_t n, int32_t ng, int32_t terms, int32_t n_tiles, int32_t n_coils,
                               const double* tw, const double* lw, const double* psi, double* sig, void* stream) {
    if (!grid || !run_of || !row_of || !tw || !lw || !psi || !sig || n_slots < 0 || n < 1 || ng < 2 || terms < 1 ||
        terms > HB_NUFFT_MAX_TERMS || n_tiles < 1 || n_coils < 1 || n_coils > 8 * n_tiles)
        return hipErrorInvalidValue;
    if (n_slots == 0) return hipSuccess;
    hipLaunchKernelGGL(hb::nufft_signal_kernel, dim3(n, n_slots), dim3(256), 0, (hipStream_t)stream, grid, run_of,
                       row_of, n, ng, terms, 16 * n_tiles, n_coils, tw, lw, psi, sig);
    return hipGetLastError();
}
