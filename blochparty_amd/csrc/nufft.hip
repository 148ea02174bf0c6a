// NUFFT readouts after the forward kernel's spreading (readouts="nufft"; replaces the reference's float atomics
// into the grid rows and its host-side FFT and recombination, src/blochparty/engines/gpu/driver.py:549-580,
// 703-725).
//
// hb_nufft_grid: every grid cell of every run slot is the sum, in workgroup order, of the window columns of the
// workgroups whose windows hold the cell (the lists are built on the host from the windows): one lane per (slot,
// term, cell, value), float64, no atomics, so the grids are the same bit for bit on every run.
// hb_nufft_signal: the readout's samples are the grids' DFT at the n modes j = i - n / 2 (the modes the samples
// are), divided by the kernel's transform and recombined over the decay terms by their Lagrange weights (folded
// into lw), then demodulated by the sample's receiver phase: one lane per (slot, sample, coil), float64.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/blochparty_amd.h"

namespace hb {

__global__ __launch_bounds__(256) void nufft_grid_kernel(const float* __restrict__ part, int runs_max, int cols,
                                                         int vals, int n_slots, const int* __restrict__ run_of,
                                                         const int* __restrict__ kind_of, int ng, int terms,
                                                         const int* __restrict__ csr_off,
                                                         const int* __restrict__ csr,
                                                         const int* __restrict__ csr_base,
                                                         double* __restrict__ grid) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)n_slots * terms * ng * vals;
    if (idx >= total) return;
    const int v = (int)(idx % vals);
    const int u = (int)((idx / vals) % ng);
    const int m = (int)((idx / ((size_t)vals * ng)) % terms);
    const int r = (int)(idx / ((size_t)vals * ng * terms));
    const int kind = kind_of[run_of[r]];
    const int* off = csr_off + (size_t)kind * (ng + 1);
    const int* ent = csr + 2 * (size_t)csr_base[kind];
    double acc = 0.0;
    for (int e = off[u]; e < off[u + 1]; ++e) {
        const int b = ent[2 * e], lc = ent[2 * e + 1];
        acc += (double)part[(((size_t)b * runs_max + r) * cols + lc * terms + m) * vals + v];
    }
    grid[idx] = acc;
}

__global__ __launch_bounds__(256) void nufft_signal_kernel(const double* __restrict__ grid, int n_slots,
                                                           const int* __restrict__ run_of,
                                                           const int* __restrict__ row_of, int n, int ng, int terms,
                                                           int vals, int n_coils, const double* __restrict__ tw,
                                                           const double* __restrict__ lw,
                                                           const double* __restrict__ psi, double* __restrict__ sig) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_slots * n * n_coils) return;
    const int c = idx % n_coils, i = (idx / n_coils) % n, r = idx / (n_coils * n);
    const double* twi = tw + (size_t)i * ng * 2;
    double re = 0.0, im = 0.0;
    for (int m = 0; m < terms; ++m) {
        const double* g = grid + ((size_t)(r * terms + m) * ng) * vals + 2 * c;
        double fr = 0.0, fi = 0.0;
        for (int u = 0; u < ng; ++u) {
            const double gr = g[(size_t)u * vals], gi = g[(size_t)u * vals + 1];
            const double cr = twi[2 * u], ci = twi[2 * u + 1];
            fr += gr * cr - gi * ci;
            fi += gr * ci + gi * cr;
        }
        const double wgt = lw[(size_t)m * n + i];
        re += wgt * fr;
        im += wgt * fi;
    }
    const int row = row_of[run_of[r]] + i;
    double sp, cp;
    sincos(psi[row], &sp, &cp);
    double* out = sig + ((size_t)row * n_coils + c) * 2;
    out[0] = re * cp + im * sp;
    out[1] = im * cp - re * sp;
}

}  // namespace hb

extern "C" int hb_nufft_grid(const float* part, int32_t runs_max, int32_t cols, int32_t n_tiles, int32_t n_slots,
                             const int32_t* run_of, const int32_t* kind_of, int32_t ng, int32_t terms,
                             const int32_t* csr_off, const int32_t* csr, const int32_t* csr_base, double* grid,
                             void* stream) {
    if (!part || !run_of || !kind_of || !csr_off || !csr || !csr_base || !grid || runs_max < 1 || cols < 1 ||
        n_tiles < 1 || n_slots < 0 || n_slots > runs_max || ng < 2 || terms < 1 || terms > HB_NUFFT_MAX_TERMS)
        return hipErrorInvalidValue;
    if (n_slots == 0) return hipSuccess;
    const size_t total = (size_t)n_slots * terms * ng * 16 * n_tiles;
    hipLaunchKernelGGL(hb::nufft_grid_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, part, runs_max, cols, 16 * n_tiles, n_slots, run_of, kind_of, ng, terms,
                       csr_off, csr, csr_base, grid);
    return hipGetLastError();
}

extern "C" int hb_nufft_signal(const double* grid, int32_t n_slots, const int32_t* run_of, const int32_t* row_of,
                               int32_t n, int32_t ng, int32_t terms, int32_t n_tiles, int32_t n_coils,
                               const double* tw, const double* lw, const double* psi, double* sig, void* stream) {
    if (!grid || !run_of || !row_of || !tw || !lw || !psi || !sig || n_slots < 0 || n < 1 || ng < 2 || terms < 1 ||
        terms > HB_NUFFT_MAX_TERMS || n_tiles < 1 || n_coils < 1 || n_coils > 8 * n_tiles)
        return hipErrorInvalidValue;
    if (n_slots == 0) return hipSuccess;
    const int total = n_slots * n * n_coils;
    hipLaunchKernelGGL(hb::nufft_signal_kernel, dim3((total + 255) / 256), dim3(256), 0, (hipStream_t)stream, grid,
                       n_slots, run_of, row_of, n, ng, terms, 16 * n_tiles, n_coils, tw, lw, psi, sig);
    return hipGetLastError();
}
