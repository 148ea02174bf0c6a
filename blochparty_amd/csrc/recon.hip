// Downstream of the signal: the centred inverse DFT of Cartesian k-space along one axis, and the root-sum-of-squares
// over coils. Replace the torch FFT calls of the reference's recon.py:17-38 (ifftnc = fftshift(ifftn(ifftshift(k))),
// rss) for complex64 data on the GPU.
//
// hb_ifftc_axis: src viewed as (outer, L, inner) complex64 values (re, im interleaved), transformed along the middle
// axis into dst (dst = src: in place; each tile is in LDS before any of it is stored):
//   y[m] = (-1)^(m + L/2) / L * sum_l (-1)^l x[l] e^(2 pi i l m / L),
// which for even L is fftshift(ifft(ifftshift(x))) with the backward normalisation (the shifts become the two sign
// patterns). A 2D or 3D ifftnc is one call per axis (the sign patterns factor over the axes).
// One workgroup per tile of TI lines along the axis (TI consecutive rows when inner = 1, TI consecutive inner indices
// otherwise, so that every load and store is coalesced; TI a power of two): the tile is loaded into LDS with the input
// signs, transformed by a Stockham FFT (radix-4 stages, and one radix-2 stage when log2 L is odd; each stage read
// into registers, a barrier, written to its new places, a barrier; the twiddles from a table of L/2 computed once per
// workgroup), and stored with the output signs and 1/L. All index arithmetic is shifts and masks.
// hb_rss: out[v] = sqrt(sum_c |x[c][v]|^2), the coils summed in order in float32.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/blochparty_amd.h"

namespace hb {
namespace {

constexpr int kFftThreads = 256;
constexpr int kFftTile = 4096;                             // complex values per tile: 32 KB of LDS
constexpr int kFftPerThread = kFftTile / 2 / kFftThreads;  // butterflies per thread and stage, at most

// lgL = log2 L, lgT = log2 TI (TI lines per tile, a power of two; lines past the data are zeros, transformed and
// never stored). LDS (dynamic): the tile, then the L/2 twiddles e^(2 pi i t / L).
__global__ __launch_bounds__(kFftThreads) void ifftc_axis_kernel(const float2* src, float2* dst, int lgL,
                                                                 int64_t inner, int64_t outer, int lgT,
                                                                 int64_t tiles_inner) {
    extern __shared__ float2 lds[];
    const int tid = threadIdx.x;
    const int L = 1 << lgL, TI = 1 << lgT, half = L >> 1, n = L << lgT;
    float2* buf = lds;
    float2* tw = lds + n;
    for (int t = tid; t < half; t += kFftThreads) {
        float s, c;
        sincospif((float)(2 * t) / (float)L, &s, &c);
        tw[t] = make_float2(c, s);
    }
    const bool rows = inner == 1;
    // the tile: element (l, t) of line t at buf[(l << ls) + (t << lt)] and in global memory at
    // gbase + l * gl + t * gt
    int64_t gbase;
    int n_lines;
    if (rows) {
        const int64_t o0 = (int64_t)blockIdx.x << lgT;
        n_lines = (int)min((int64_t)TI, outer - o0);
        gbase = o0 * L;
    } else {
        const int64_t o = blockIdx.x / tiles_inner, i0 = (blockIdx.x % tiles_inner) << lgT;
        n_lines = (int)min((int64_t)TI, inner - i0);
        gbase = o * L * inner + i0;
    }
    const int ls = rows ? 0 : lgT, lt = rows ? lgL : 0;
    const int64_t gl = rows ? 1 : inner, gt = rows ? L : 1;
    // load, with the input signs (-1)^l
    for (int e = tid; e < n; e += kFftThreads) {
        const int l = rows ? e & (L - 1) : e >> lgT, t = rows ? e >> lgL : e & (TI - 1);
        float2 v = make_float2(0.0f, 0.0f);
        if (t < n_lines) v = src[gbase + l * gl + t * gt];
        if (l & 1) v.x = -v.x, v.y = -v.y;
        buf[(l << ls) + (t << lt)] = v;
    }
    __syncthreads();
    // e^(2 pi i t / L) for t < L, from the half table
    auto W = [&](int t) {
        const float2 w = tw[t & (half - 1)];
        return t < half ? w : make_float2(-w.x, -w.y);
    };
    // radix-4 stages while four fit (sub-transforms of ns values grow 4x), then one radix-2 stage if log2 L is odd
    int lgns = 0;
    for (; lgns + 2 <= lgL; lgns += 2) {
        const int ns = 1 << lgns, lgq = lgL - 2, q4 = 1 << lgq, n_bf = q4 << lgT, sh = lgL - lgns - 2;
        float2 v[kFftPerThread / 2][4];
        int to[kFftPerThread / 2];
#pragma unroll
        for (int r = 0; r < kFftPerThread / 2; ++r) {
            const int idx = tid + r * kFftThreads;
            if (idx < n_bf) {
                const int j = rows ? idx & (q4 - 1) : idx >> lgT, t = rows ? idx >> lgq : idx & (TI - 1);
                const int k = j & (ns - 1);
#pragma unroll
                for (int c = 0; c < 4; ++c) v[r][c] = buf[((j + (c << lgq)) << ls) + (t << lt)];
#pragma unroll
                for (int c = 1; c < 4; ++c) {  // v_c e^(2 pi i k c / (4 ns))
                    const float2 w = W((k * c) << sh), x = v[r][c];
                    v[r][c] = make_float2(x.x * w.x - x.y * w.y, x.x * w.y + x.y * w.x);
                }
                // the inverse 4-point DFT: V_s = sum_c v_c i^(c s)
                const float2 a = v[r][0], b = v[r][1], c2 = v[r][2], d = v[r][3];
                const float2 apc = make_float2(a.x + c2.x, a.y + c2.y), amc = make_float2(a.x - c2.x, a.y - c2.y);
                const float2 bpd = make_float2(b.x + d.x, b.y + d.y), bmd = make_float2(b.x - d.x, b.y - d.y);
                v[r][0] = make_float2(apc.x + bpd.x, apc.y + bpd.y);
                v[r][1] = make_float2(amc.x - bmd.y, amc.y + bmd.x);  // a + i b - c - i d
                v[r][2] = make_float2(apc.x - bpd.x, apc.y - bpd.y);
                v[r][3] = make_float2(amc.x + bmd.y, amc.y - bmd.x);  // a - i b - c + i d
                to[r] = ((((j - k) << 2) + k) << ls) + (t << lt);
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kFftPerThread / 2; ++r) {
            if (tid + r * kFftThreads < n_bf) {
#pragma unroll
                for (int c = 0; c < 4; ++c) buf[to[r] + ((c << lgns) << ls)] = v[r][c];
            }
        }
        __syncthreads();
    }
    if (lgns < lgL) {  // the last stage, radix 2: ns = L / 2, so k = j and the twiddle is e^(2 pi i j / L)
        const int n_bf = half << lgT, lgh = lgL - 1;
        float2 va[kFftPerThread], vb[kFftPerThread];
        int to[kFftPerThread];
#pragma unroll
        for (int r = 0; r < kFftPerThread; ++r) {
            const int idx = tid + r * kFftThreads;
            if (idx < n_bf) {
                const int j = rows ? idx & (half - 1) : idx >> lgT, t = rows ? idx >> lgh : idx & (TI - 1);
                const float2 a = buf[(j << ls) + (t << lt)], b = buf[((j + half) << ls) + (t << lt)];
                const float2 w = tw[j];
                const float2 bw = make_float2(b.x * w.x - b.y * w.y, b.x * w.y + b.y * w.x);
                va[r] = make_float2(a.x + bw.x, a.y + bw.y);
                vb[r] = make_float2(a.x - bw.x, a.y - bw.y);
                to[r] = (j << ls) + (t << lt);
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kFftPerThread; ++r) {
            if (tid + r * kFftThreads < n_bf) {
                buf[to[r]] = va[r];
                buf[to[r] + (half << ls)] = vb[r];
            }
        }
        __syncthreads();
    }
    // store, with the output signs (-1)^(m + L/2) and 1/L
    const float scale = 1.0f / (float)L;
    for (int e = tid; e < n; e += kFftThreads) {
        const int m = rows ? e & (L - 1) : e >> lgT, t = rows ? e >> lgL : e & (TI - 1);
        if (t < n_lines) {
            float2 v = buf[(m << ls) + (t << lt)];
            const float f = ((m + half) & 1) ? -scale : scale;
            dst[gbase + m * gl + t * gt] = make_float2(v.x * f, v.y * f);
        }
    }
}

__global__ __launch_bounds__(256) void rss_kernel(const float2* __restrict__ x, int n_coils, int64_t per_coil,
                                                  float* __restrict__ out) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= per_coil) return;
    float acc = 0.0f;
    for (int c = 0; c < n_coils; ++c) {
        const float2 a = x[(int64_t)c * per_coil + v];
        acc += a.x * a.x + a.y * a.y;
    }
    out[v] = sqrtf(acc);
}

}  // namespace
}  // namespace hb

extern "C" int hb_ifftc_axis(const float* src, float* dst, int64_t outer, int32_t L, int64_t inner, void* stream) {
    using namespace hb;
    if (!src || !dst || outer < 0 || inner < 1 || L < 2 || L > kFftTile || (L & (L - 1))) return hipErrorInvalidValue;
    if (outer == 0) return hipSuccess;
    int lgL = 0;
    while ((1 << lgL) < L) ++lgL;
    // lines per tile: a power of two, at most kFftTile / L and no more than the lines there are (rounded up)
    const int64_t lines = inner == 1 ? outer : inner;
    int lgT = 0;
    while ((1 << (lgT + 1)) * L <= kFftTile && ((int64_t)1 << lgT) < lines) ++lgT;
    const int64_t TI = (int64_t)1 << lgT;
    int64_t tiles, tiles_inner = 1;
    if (inner == 1) {
        tiles = (outer + TI - 1) / TI;
    } else {
        tiles_inner = (inner + TI - 1) / TI;
        tiles = outer * tiles_inner;
    }
    if (tiles > INT32_MAX) return hipErrorInvalidValue;
    const size_t lds = ((size_t)L * TI + L / 2) * sizeof(float2);
    hipLaunchKernelGGL(ifftc_axis_kernel, dim3((unsigned)tiles), dim3(kFftThreads), lds, (hipStream_t)stream,
                       reinterpret_cast<const float2*>(src), reinterpret_cast<float2*>(dst), lgL, inner, outer, lgT,
                       tiles_inner);
    return hipGetLastError();
}

extern "C" int hb_rss(const float* x, int32_t n_coils, int64_t per_coil, float* out, void* stream) {
    if (!x || !out || n_coils < 1 || per_coil < 0) return hipErrorInvalidValue;
    if (per_coil == 0) return hipSuccess;
    const int64_t blocks = (per_coil + 255) / 256;
    if (blocks > INT32_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(hb::rss_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       reinterpret_cast<const float2*>(x), n_coils, per_coil, out);
    return hipGetLastError();
}
