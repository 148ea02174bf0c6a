// The learned RF surrogate (no counterpart in the reference; it stands where the reference composes a reused pulse
// per spin, src/blochparty/engines/gpu/cuda/pulse.cu.in:3-32).
//
// A reused pulse of K substeps is cut into M = HB_SUR_BLOCKS blocks. Each block's pure rotation is a unit quaternion
// taken in the frame that turns with the off-resonance precession: Q~_j = Rz(-theta_j/2) Q_j Rz(-theta_j/2),
// theta_j = -2 pi (f tau_j + B_j) the block's free-precession angle (f the spin's offset, tau_j the block's length,
// B_j its b0_offset area), so that Q~_j varies slowly with f. A network 10 -> 64 -> 64 -> 4 M (tanh) maps features
// of a spin's (f, b) (f = df + g . r its offset during the pulse [Hz], b its B1 scale times the pulse's peak [Hz]) to
// the M quaternions. The features are f and b scaled into the trained box and sin, cos of 2 pi f q S, q = 1..4,
// S = the mean block length / 8: to first order in the RF, block j's rotation is the sum over its substeps of
// rf_k e^{i 2 pi f (t_k - t_j)}, t_k - t_j within half a block of the block's middle, and these frequencies span
// that range. The network does not learn that first-order part itself: a block's quaternion is the first-order
// rotation exp(b F_j(f) / 2), F_j tabulated per network over its box's f range (first_order_kernel) and
// interpolated, plus the network's output (the higher orders, two orders of magnitude smaller in x and y). A spin's
// pulse map m -> m + D m + v is then composed in float64 with the relaxation put back exactly, half a block before
// and after each rotation: a Strang split of the pulse into M blocks.
//
// First order: with v_k = -2 pi dt_k (b rf_k, f + off_k) substep k's rotation vector, Phi_k the z angle up to the
// middle of substep k and theta the block's, Q = Rz(theta) prod_k Rz(-Phi_k) exp(v_k,xy) Rz(Phi_k) (each substep
// split symmetrically about its z part, exact to first order), so to first order in the RF
// Q~ = Rz(-theta/2) Q Rz(-theta/2) = exp(F), F = sum_k Rz(theta/2 - Phi_k) v_k,xy.
//
// The split's correction. To first order in the relaxation rates (Duhamel's formula; generator Rg = diag(R2, R2, R1),
// which commutes with Rz), the block's exact map differs from relax(tau/2) U relax(tau/2) by U times the error of the
// trapezoidal rule on int_0^tau V(s)^T Rg V(s) ds (V(s) the rotation from the block's start to s in the frame turning
// with the precession; its part along z commutes with Rg), and on the offset by U times that error on
// int V(s)^T z ds R1 PD. With v(s) = V's rotation vector in the block's middle frame, to second order in v:
// linear part Rz(theta/2) Q~ (R2 - R1) (S(H) - T(G)) Rz(theta/2), offset Rz(theta/2) Q~ R1 PD (H_y, -H_x, -tr(G) / 2),
// where H = sum_k s_k u_k is the first time moment of F's terms u_k (s_k from the block's middle), G =
// int v v^T ds - (tau/2) v(tau) v(tau)^T the 2 x 2 second moment of the path, S(H) the symmetric matrix with
// (0, 2) = (2, 0) = H_y, (1, 2) = (2, 1) = -H_x, and T(G) = [[-G_yy, G_xy, 0], [G_xy, -G_xx, 0], [0, 0, tr G]]. It
// matters where a block's precession angle is large (spins far from the pulse's band, whose magnetization nutates to
// and fro within a block); H and G are tabulated beside F.
//
// Kernels: the exact spinor products that make the training data (float64); the network's training step (forward,
// backward and weight gradients of 64 samples per workgroup as matrix products on the matrix cores, one partial per
// workgroup summed in a fixed order, Adam); its evaluation; and the per-spin maps, with spins outside the trained
// input box flagged for the exact composition (hb_pulse_maps' mask).
#include "bloch_device.h"

namespace hb {

constexpr int kH = HB_MLP_HIDDEN;
constexpr int kM = HB_SUR_BLOCKS;
constexpr int kOut = 4 * kM;
constexpr int kIn = HB_MLP_INPUTS, kFourier = HB_MLP_FOURIER, kNormHead = HB_MLP_NORM_HEAD;
constexpr int kRaw = 2;  // floats per raw input: f [Hz], b [Hz]
constexpr int kW1 = 0, kB1 = kW1 + kH * kIn, kW2 = kB1 + kH, kB2 = kW2 + kH * kH, kW3 = kB2 + kH,
              kB3 = kW3 + kOut * kH, kParams = kB3 + kOut;
static_assert(kParams == HB_MLP_PARAMS, "parameter layout");

struct Quat {
    double w, x, y, z;
};

__device__ __forceinline__ Quat qmul(const Quat& p, const Quat& q) {
    return {p.w * q.w - p.x * q.x - p.y * q.y - p.z * q.z, p.w * q.x + p.x * q.w + p.y * q.z - p.z * q.y,
            p.w * q.y - p.x * q.z + p.y * q.w + p.z * q.x, p.w * q.z + p.x * q.y - p.y * q.x + p.z * q.w};
}

// Right-handed rotation by |w| about w.
__device__ __forceinline__ Quat rot_quat(double wx, double wy, double wz) {
    const double phi = sqrt(wx * wx + wy * wy + wz * wz);
    double s, c;
    sincos(0.5 * phi, &s, &c);
    const double k = phi > 0.0 ? s / phi : 0.5;
    return {c, k * wx, k * wy, k * wz};
}

__device__ __forceinline__ Quat rot_z(double angle) {
    double s, c;
    sincos(0.5 * angle, &s, &c);
    return {c, 0.0, 0.0, s};
}

// ---- training data: exact interaction-frame block quaternions (float64) ----

__global__ __launch_bounds__(256) void spinor_blocks_kernel(const double* __restrict__ table,
                                                            const int* __restrict__ bounds, const float* __restrict__ x,
                                                            int n, float* __restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double f0 = x[kRaw * i], b1 = x[kRaw * i + 1];
    const double w0 = -2.0 * M_PI;
    for (int j = 0; j < kM; ++j) {
        Quat q = {1.0, 0.0, 0.0, 0.0};
        double theta = 0.0;
        for (int k = bounds[j]; k < bounds[j + 1]; ++k) {
            const double* r = table + (size_t)k * 8;
            const double w = w0 * r[0], f = f0 + r[6];
            q = qmul(rot_quat(w * b1 * r[1], w * b1 * r[2], w * f), q);
            theta += w * f;
        }
        const Quat h = rot_z(-0.5 * theta);
        q = qmul(h, qmul(q, h));
        float* out = y + (size_t)i * kOut + 4 * j;
        out[0] = (float)q.w, out[1] = (float)q.x, out[2] = (float)q.y, out[3] = (float)q.z;
    }
}

// ---- the first-order part ----

// first[(j * n_grid + i) * 8 + {0..6}] = F_j (x, y), H_j (x, y) per Hz of b and G_j (xx, xy, yy) per Hz^2 of b at
// f_lo + i df, float64 sums over the block's substeps (v(s) growing linearly within each substep); entry 7 is 0.
__global__ __launch_bounds__(256) void first_order_kernel(const double* __restrict__ table,
                                                          const int* __restrict__ bounds, double f_lo, double df,
                                                          int n_grid, float* __restrict__ first) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= kM * n_grid) return;
    const int j = t / n_grid, i = t - j * n_grid;
    const double f = f_lo + i * df;
    double z_total = 0.0;  // sum dt (f + off): theta = -2 pi z_total
    for (int k = bounds[j]; k < bounds[j + 1]; ++k) z_total += table[(size_t)k * 8] * (f + table[(size_t)k * 8 + 6]);
    double tau = 0.0;
    for (int k = bounds[j]; k < bounds[j + 1]; ++k) tau += table[(size_t)k * 8];
    double z = 0.0, time = 0.0, rx = 0.0, ry = 0.0, hx = 0.0, hy = 0.0, gxx = 0.0, gxy = 0.0, gyy = 0.0;
    for (int k = bounds[j]; k < bounds[j + 1]; ++k) {
        const double* r = table + (size_t)k * 8;
        const double dt = r[0], zk = dt * (f + r[6]);
        const double phi = -2.0 * M_PI * (z + 0.5 * zk - 0.5 * z_total);  // Phi_k - theta / 2
        double sn, cs;
        sincos(phi, &sn, &cs);
        const double ax = -2.0 * M_PI * dt * r[1], ay = -2.0 * M_PI * dt * r[2];
        const double ux = cs * ax + sn * ay, uy = cs * ay - sn * ax;  // Rz(-phi) (ax, ay)
        const double sk = time + 0.5 * dt - 0.5 * tau;               // substep middle - block middle
        hx += sk * ux, hy += sk * uy;
        gxx += dt * (rx * rx + rx * ux + ux * ux / 3.0);  // int over the substep of v v^T, v = (r + a u), a 0 -> 1
        gxy += dt * (rx * ry + 0.5 * (rx * uy + ux * ry) + ux * uy / 3.0);
        gyy += dt * (ry * ry + ry * uy + uy * uy / 3.0);
        rx += ux, ry += uy;
        z += zk;
        time += dt;
    }
    float* o = first + (size_t)t * 8;
    o[0] = (float)rx, o[1] = (float)ry, o[2] = (float)hx, o[3] = (float)hy;
    o[4] = (float)(gxx - 0.5 * tau * rx * rx), o[5] = (float)(gxy - 0.5 * tau * rx * ry);
    o[6] = (float)(gyy - 0.5 * tau * ry * ry), o[7] = 0.0f;
}

// Block j's table entries at f (F_j, H_j per Hz of b, G_j per Hz^2 of b), interpolated linearly on the grid over
// [c_f - h_f, c_f + h_f] (clamped at its ends).
template <int NV>
__device__ __forceinline__ void first_lookup(const float* __restrict__ first, int n_grid,
                                             const float* __restrict__ norm, float f, int j, float (&v)[NV]) {
    const float lo = norm[0] - norm[1], step = 2.0f * norm[1] / (float)(n_grid - 1);
    const float pos = fminf(fmaxf((f - lo) / step, 0.0f), (float)(n_grid - 1));
    const int i = min((int)pos, n_grid - 2);
    const float a = pos - (float)i;
    const float* g = first + ((size_t)j * n_grid + i) * 8;
#pragma unroll
    for (int c = 0; c < NV; ++c) v[c] = fmaf(a, g[8 + c] - g[c], g[c]);
}

// Block j's first-order quaternion at (f, b [Hz]): the rotation by b F_j. Zero without a table (the network then
// learns the whole quaternion).
__device__ __forceinline__ void first_quat(const float* __restrict__ first, int n_grid, const float* __restrict__ norm,
                                           float f, float b, int j, float (&q)[4]) {
    if (!first) {
        q[0] = q[1] = q[2] = q[3] = 0.0f;
        return;
    }
    float v[2];
    first_lookup(first, n_grid, norm, f, j, v);
    const float rx = b * v[0], ry = b * v[1];
    const float ang = sqrtf(rx * rx + ry * ry);
    float sn, cs;
    sincosf(0.5f * ang, &sn, &cs);
    const float k = ang > 0.0f ? sn / ang : 0.5f;
    q[0] = cs, q[1] = k * rx, q[2] = k * ry, q[3] = 0.0f;
}

__global__ __launch_bounds__(256) void first_quats_kernel(const float* __restrict__ first, int n_grid,
                                                          const float* __restrict__ norm, const float* __restrict__ x,
                                                          int n, float* __restrict__ q) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float f = x[kRaw * s], b = x[kRaw * s + 1] * norm[5];
    for (int j = 0; j < kM; ++j) {
        float v[4];
        first_quat(first, n_grid, norm, f, b, j, v);
        float* o = q + (size_t)s * kOut + 4 * j;
        o[0] = v[0], o[1] = v[1], o[2] = v[2], o[3] = v[3];
    }
}

// ---- the network ----

// Parameters staged in LDS once per workgroup; every lane reads the same address (a broadcast).
__device__ __forceinline__ void stage_params(float* __restrict__ lds, const float* __restrict__ p) {
    for (int i = threadIdx.x; i < kParams; i += blockDim.x) lds[i] = p[i];
    __syncthreads();
}

__device__ __forceinline__ void layer1(const float* __restrict__ w, const float (&x)[kIn], float (&a1)[kH]) {
#pragma unroll
    for (int i = 0; i < kH; ++i) {
        float acc = w[kB1 + i];
#pragma unroll
        for (int c = 0; c < kIn; ++c) acc += w[kW1 + kIn * i + c] * x[c];
        a1[i] = tanhf(acc);
    }
}

__device__ __forceinline__ void layer2(const float* __restrict__ w, const float (&a1)[kH], float (&a2)[kH]) {
#pragma unroll
    for (int j = 0; j < kH; ++j) {
        const float4* row = reinterpret_cast<const float4*>(w + kW2 + kH * j);
        float acc = w[kB2 + j];
#pragma unroll
        for (int i = 0; i < kH / 4; ++i) {
            const float4 r = row[i];
            acc += r.x * a1[4 * i] + r.y * a1[4 * i + 1] + r.z * a1[4 * i + 2] + r.w * a1[4 * i + 3];
        }
        a2[j] = tanhf(acc);
    }
}

__device__ __forceinline__ float layer3(const float* __restrict__ w, const float (&a2)[kH], int k) {
    const float4* row = reinterpret_cast<const float4*>(w + kW3 + kH * k);
    float acc = w[kB3 + k];
#pragma unroll
    for (int i = 0; i < kH / 4; ++i) {
        const float4 r = row[i];
        acc += r.x * a2[4 * i] + r.y * a2[4 * i + 1] + r.z * a2[4 * i + 2] + r.w * a2[4 * i + 3];
    }
    return acc;
}

// The network's inputs from a spin's (f, b): norm = c_f, h_f, c_b, h_b, S. The features' phases 2 pi f q S are taken
// in turns by sincospif (|2 f q S| stays within a few turns over a box, where its float32 result is accurate).
__device__ __forceinline__ void features(const float* __restrict__ norm, float f, float b, float (&x)[kIn]) {
    x[0] = (f - norm[0]) / norm[1];
    x[1] = (b - norm[2]) / norm[3];
    const float t = 2.0f * f * norm[4];
#pragma unroll
    for (int q = 1; q <= kFourier; ++q) sincospif(t * (float)q, &x[2 * q], &x[2 * q + 1]);
}

typedef float f4 __attribute__((ext_vector_type(4)));

// ---- the training gradient on the matrix cores ----
// One workgroup of 4 waves per kGroup = 64 samples. Every layer is a matrix product on v_mfma_f32_16x16x4f32 with the
// hidden units as rows and the samples as columns (lane l supplies A[l % 16][k0 + l / 16] and B[k0 + l / 16][l % 16]
// and receives OUT[4 (l / 16) + r][l % 16], r = 0..3):
//   phase 1, each wave on its 16 samples: Z1 = W1a X (W1a = [W1 | b1], X's row kIn the ones), A1 = tanh Z1,
//     A2 = tanh(W2 A1 + b2), out = W3 A2 + b3, D3 = 2 (out - t) / (n n_out) (t the normalised target, 0 for samples
//     past n), D2 = (W3^T D3) (1 - A2^2), D1 = (W2^T D2) (1 - A1^2);
//   phase 2, the workgroup's 64 samples as the inner dimension: dW1a = D1 X^T, dW2 = D2 A1^T, dW3 = D3 A2^T (32 tiles of
//     16 x 16, 8 per wave), the biases b2, b3 as in-order row sums of D2, D3;
// one partial per workgroup, summed over the workgroups in order by mlp_adam_kernel.
// LDS: the parameters with padded rows and the activations / deltas as [row][sample], rows kLd = 68 floats apart
// (68 = 4 mod 64 banks: phase 2's reads of 16 rows x 4 consecutive samples hit 64 distinct banks).
constexpr int kGroup = 64;              // samples per workgroup
constexpr int kLd = 68;                 // LDS row stride (floats)
constexpr int kInP = 16, kOutP = 48;    // inputs + ones, outputs, padded to whole tiles
constexpr int kW1s = 12;                // W1a row: W1 (kIn), b1, 0
constexpr int L_W1 = 0, L_W2 = L_W1 + kH * kW1s, L_B2 = L_W2 + kH * kLd, L_W3 = L_B2 + kH, L_B3 = L_W3 + kOutP * kLd,
              L_PARAMS = L_B3 + kOutP;
static_assert(kIn + 2 <= kW1s, "W1a row");

// The parameters into the padded LDS layout: every thread's loads (float4, 8 per thread) are issued before its first
// store, then each value goes to its row and column; the padding the kernel reads (W1a's last column, W3's and b3's
// rows past n_out) is zeroed.
__device__ __forceinline__ int padded_index(int i) {
    if (i < kB1) return L_W1 + (i / kIn) * kW1s + i % kIn;
    if (i < kW2) return L_W1 + (i - kB1) * kW1s + kIn;
    if (i < kB2) return L_W2 + ((i - kW2) / kH) * kLd + (i - kW2) % kH;
    if (i < kW3) return L_B2 + i - kB2;
    if (i < kB3) return L_W3 + ((i - kW3) / kH) * kLd + (i - kW3) % kH;
    return L_B3 + i - kB3;
}

static_assert(kParams % 4 == 0, "float4 staging");

__device__ __forceinline__ void stage_params_padded(float* __restrict__ lw, const float* __restrict__ p) {
    constexpr int kVec = kParams / 4, kPer = (kVec + 255) / 256;
    const float4* p4 = reinterpret_cast<const float4*>(p);
    float4 v[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int q = threadIdx.x + 256 * j;
        v[j] = q < kVec ? p4[q] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    for (int i = threadIdx.x; i < kH; i += blockDim.x) lw[L_W1 + i * kW1s + kIn + 1] = 0.0f;
    for (int i = threadIdx.x; i < (kOutP - kOut) * kLd; i += blockDim.x) lw[L_W3 + kOut * kLd + i] = 0.0f;
    for (int i = threadIdx.x; i < kOutP - kOut; i += blockDim.x) lw[L_B3 + kOut + i] = 0.0f;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int q = threadIdx.x + 256 * j;
        if (q < kVec) {
            lw[padded_index(4 * q)] = v[j].x;
            lw[padded_index(4 * q + 1)] = v[j].y;
            lw[padded_index(4 * q + 2)] = v[j].z;
            lw[padded_index(4 * q + 3)] = v[j].w;
        }
    }
}

__device__ __forceinline__ void mfma(f4& acc, float a, float b) {
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
}

__global__ __launch_bounds__(256) void mlp_grad_mfma_kernel(const float* __restrict__ params,
                                                            const float* __restrict__ norm,
                                                            const float* __restrict__ x, const float* __restrict__ y,
                                                            int n, float* __restrict__ partial) {
    __shared__ __align__(16) float lw[L_PARAMS];
    __shared__ float X[kInP * kLd], A1[kH * kLd], A2[kH * kLd], D1[kH * kLd], D2[kH * kLd], D3[kOutP * kLd];
    stage_params_padded(lw, params);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, kq = lane >> 4;
    const int col = 16 * wave + m;  // this lane's sample column in phase 1 (the B operand's and the output's)
    const int s = blockIdx.x * kGroup + col;
    const bool valid = s < n;
    if (kq == 0) {  // the wave's 16 samples' inputs, one lane each
        float xi[kIn];
        if (valid) {
            features(norm, x[kRaw * s], x[kRaw * s + 1], xi);
        } else {
#pragma unroll
            for (int c = 0; c < kIn; ++c) xi[c] = 0.0f;
        }
#pragma unroll
        for (int c = 0; c < kInP; ++c) X[c * kLd + col] = c < kIn ? xi[c] : (c == kIn && valid) ? 1.0f : 0.0f;
    }
    __syncthreads();
    // phase 1: forward
    // (each layer's output tiles are independent accumulation chains, interleaved in the inner loop)
    float a1v[4][4], a2v[4][4];
    f4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int u = 0; u < kInP / 4; ++u) {
        const int k = 4 * u + kq;
        const float b = X[k * kLd + col];
#pragma unroll
        for (int t = 0; t < 4; ++t) mfma(acc[t], k < kW1s ? lw[L_W1 + (16 * t + m) * kW1s + k] : 0.0f, b);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            a1v[t][r] = tanhf(acc[t][r]);
            A1[(16 * t + 4 * kq + r) * kLd + col] = a1v[t][r];
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[t][r] = lw[L_B2 + 16 * t + 4 * kq + r];
    }
#pragma unroll
    for (int u = 0; u < kH / 4; ++u) {
        const int k = 4 * u + kq;
        const float b = A1[k * kLd + col];
#pragma unroll
        for (int t = 0; t < 4; ++t) mfma(acc[t], lw[L_W2 + (16 * t + m) * kLd + k], b);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            a2v[t][r] = tanhf(acc[t][r]);
            A2[(16 * t + 4 * kq + r) * kLd + col] = a2v[t][r];
        }
    }
    __syncthreads();
    const float scale = 2.0f / ((float)n * (float)kOut);
    const float* mu = norm + kNormHead;
    const float* sd = mu + kOut;
#pragma unroll
    for (int t = 0; t < kOutP / 16; ++t) {
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[t][r] = lw[L_B3 + 16 * t + 4 * kq + r];
    }
#pragma unroll
    for (int u = 0; u < kH / 4; ++u) {
        const int k = 4 * u + kq;
        const float b = A2[k * kLd + col];
#pragma unroll
        for (int t = 0; t < kOutP / 16; ++t) mfma(acc[t], lw[L_W3 + (16 * t + m) * kLd + k], b);
    }
#pragma unroll
    for (int t = 0; t < kOutP / 16; ++t) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int o = 16 * t + 4 * kq + r;
            float d = 0.0f;
            if (valid && o < kOut) d = scale * (acc[t][r] - (y[(size_t)s * kOut + o] - mu[o]) / sd[o]);
            D3[o * kLd + col] = d;
        }
    }
    __syncthreads();
    // phase 1: backward
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int u = 0; u < kOutP / 4; ++u) {
        const int k = 4 * u + kq;
        const float b = D3[k * kLd + col];
#pragma unroll
        for (int t = 0; t < 4; ++t) mfma(acc[t], lw[L_W3 + k * kLd + 16 * t + m], b);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
            D2[(16 * t + 4 * kq + r) * kLd + col] = acc[t][r] * (1.0f - a2v[t][r] * a2v[t][r]);
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int u = 0; u < kH / 4; ++u) {
        const int k = 4 * u + kq;
        const float b = D2[k * kLd + col];
#pragma unroll
        for (int t = 0; t < 4; ++t) mfma(acc[t], lw[L_W2 + k * kLd + 16 * t + m], b);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
            D1[(16 * t + 4 * kq + r) * kLd + col] = acc[t][r] * (1.0f - a1v[t][r] * a1v[t][r]);
    }
    __syncthreads();
    // phase 2: the weight gradients over the workgroup's samples (tile i of 32: layer 1 rows 4, layer 2 4 x 4,
    // layer 3 3 x 4), then the biases b2, b3
    // (wave w takes tiles w, w + 4, ...: two at a time, as two interleaved accumulation chains)
    float* out = partial + (size_t)blockIdx.x * kParams;
    for (int i0 = wave; i0 < 32; i0 += 8) {
        const float* Dm[2];
        const float* Am[2];
        int rt[2], ct[2], layer[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int i = i0 + 4 * h;
            if (i < 4) layer[h] = 0, rt[h] = i, ct[h] = 0, Dm[h] = D1, Am[h] = X;
            else if (i < 20) layer[h] = 1, rt[h] = (i - 4) >> 2, ct[h] = (i - 4) & 3, Dm[h] = D2, Am[h] = A1;
            else layer[h] = 2, rt[h] = (i - 20) >> 2, ct[h] = (i - 20) & 3, Dm[h] = D3, Am[h] = A2;
            acc[h] = f4{0.0f, 0.0f, 0.0f, 0.0f};
        }
#pragma unroll
        for (int u = 0; u < kGroup / 4; ++u) {
            const int k = 4 * u + kq;
#pragma unroll
            for (int h = 0; h < 2; ++h)
                mfma(acc[h], Dm[h][(16 * rt[h] + m) * kLd + k], Am[h][(16 * ct[h] + m) * kLd + k]);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * rt[h] + 4 * kq + r, c = 16 * ct[h] + m;
                if (layer[h] == 0) {
                    if (c < kIn) out[kW1 + row * kIn + c] = acc[h][r];
                    else if (c == kIn) out[kB1 + row] = acc[h][r];
                } else if (layer[h] == 1) {
                    out[kW2 + row * kH + c] = acc[h][r];
                } else if (row < kOut) {
                    out[kW3 + row * kH + c] = acc[h][r];
                }
            }
        }
    }
    if (tid < kH + kOut) {
        const float* Dm = tid < kH ? D2 + tid * kLd : D3 + (tid - kH) * kLd;
        float acc = 0.0f;
        for (int k = 0; k < kGroup; ++k) acc += Dm[k];
        out[tid < kH ? kB2 + tid : kB3 + tid - kH] = acc;
    }
}

// grad = the workgroups' partials summed in a fixed order (64 parameters per workgroup: wave w sums the partials
// c = w mod 4 in order, then the 4 waves' sums are added in order); then Adam (bias corrections c1 = 1/(1 - beta1^t),
// c2 = 1/(1 - beta2^t) from the host) unless grad_out is set, which receives the gradient instead.
constexpr int kAdamParams = 64;  // parameters per workgroup of mlp_adam_kernel

__global__ __launch_bounds__(256) void mlp_adam_kernel(const float* __restrict__ partial, int chunks,
                                                       float* __restrict__ params, float* __restrict__ m,
                                                       float* __restrict__ v, float lr, float c1, float c2,
                                                       float* __restrict__ grad_out) {
    __shared__ float part[4][kAdamParams];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blockIdx.x * kAdamParams + lane;
    float g = 0.0f;
    if (i < kParams) {
        for (int c = wave; c < chunks; c += 4) g += partial[(size_t)c * kParams + i];
    }
    part[wave][lane] = g;
    __syncthreads();
    if (wave != 0 || i >= kParams) return;
    g = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
    if (grad_out) {
        grad_out[i] = g;
        return;
    }
    const float mi = 0.9f * m[i] + 0.1f * g, vi = 0.999f * v[i] + 0.001f * g * g;
    m[i] = mi, v[i] = vi;
    params[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + 1e-8f);
}

// Outputs of the network, de-normalised: y[s][k] = mu[k] + sd[k] net(x^)[k].
__global__ __launch_bounds__(256) void mlp_eval_kernel(const float* __restrict__ params,
                                                       const float* __restrict__ norm, const float* __restrict__ x,
                                                       int n, float* __restrict__ y) {
    __shared__ __align__(16) float w[kParams];
    stage_params(w, params);
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float xi[kIn], a1[kH], a2[kH];
    features(norm, x[kRaw * s], x[kRaw * s + 1], xi);
    layer1(w, xi, a1);
    layer2(w, a1, a2);
    for (int k = 0; k < kOut; ++k)
        y[(size_t)s * kOut + k] = norm[kNormHead + k] + norm[kNormHead + kOut + k] * layer3(w, a2, k);
}

// ---- per-spin maps from the network ----

// Affine map (A, c) <- step (A, c) for the relaxation over t: A rows scaled by (e2, e2, e1), c likewise plus the
// recovery pd (1 - e1) on z.
__device__ __forceinline__ void relax_map(double (&A)[3][3], double (&c)[3], double e1, double e2, double pd) {
#pragma unroll
    for (int k = 0; k < 3; ++k) A[0][k] *= e2, A[1][k] *= e2, A[2][k] *= e1;
    c[0] *= e2, c[1] *= e2, c[2] = c[2] * e1 + pd * (1.0 - e1);
}

__device__ __forceinline__ void rotate_map(double (&A)[3][3], double (&c)[3], const Quat& q) {
    const double R[3][3] = {{1 - 2 * (q.y * q.y + q.z * q.z), 2 * (q.x * q.y - q.w * q.z), 2 * (q.x * q.z + q.w * q.y)},
                            {2 * (q.x * q.y + q.w * q.z), 1 - 2 * (q.x * q.x + q.z * q.z), 2 * (q.y * q.z - q.w * q.x)},
                            {2 * (q.x * q.z - q.w * q.y), 2 * (q.y * q.z + q.w * q.x), 1 - 2 * (q.x * q.x + q.y * q.y)}};
    double B[3][3], d[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int k = 0; k < 3; ++k) B[i][k] = R[i][0] * A[0][k] + R[i][1] * A[1][k] + R[i][2] * A[2][k];
        d[i] = R[i][0] * c[0] + R[i][1] * c[1] + R[i][2] * c[2];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int k = 0; k < 3; ++k) A[i][k] = B[i][k];
        c[i] = d[i];
    }
}

// v <- the rotation of unit quaternion q applied to v.
__device__ __forceinline__ void qrot(const Quat& q, double (&v)[3]) {
    const double tx = 2.0 * (q.y * v[2] - q.z * v[1]), ty = 2.0 * (q.z * v[0] - q.x * v[2]),
                 tz = 2.0 * (q.x * v[1] - q.y * v[0]);
    v[0] += q.w * tx + (q.y * tz - q.z * ty);
    v[1] += q.w * ty + (q.z * tx - q.x * tz);
    v[2] += q.w * tz + (q.x * ty - q.y * tx);
}

// The split's correction of one block (see the top of the file): A += Rz(theta/2) Q~ (R2 - R1) (S(H) - T(G))
// Rz(theta/2) A0, c += the same on c0 + Rz(theta/2) Q~ R1 PD (H_y, -H_x, -tr(G) / 2), with (A0, c0) the map before the
// block and e = (H_x, H_y, G_xx, G_xy, G_yy) for this spin.
__device__ __forceinline__ void split_correction(double (&A)[3][3], double (&c)[3], const double (&A0)[3][3],
                                                 const double (&c0)[3], const Quat& h, const Quat& q,
                                                 const double (&e)[5], double r1, double r2, double pd) {
    const double k = r2 - r1, hx = e[0], hy = e[1], gxx = e[2], gxy = e[3], gyy = e[4];
#pragma unroll
    for (int col = 0; col < 4; ++col) {
        double u[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) u[i] = col < 3 ? A0[i][col] : c0[i];
        qrot(h, u);
        double v[3] = {k * (hy * u[2] + gyy * u[0] - gxy * u[1]), k * (-hx * u[2] - gxy * u[0] + gxx * u[1]),
                       k * (hy * u[0] - hx * u[1] - (gxx + gyy) * u[2])};
        if (col == 3) v[0] += r1 * pd * hy, v[1] -= r1 * pd * hx, v[2] -= 0.5 * r1 * pd * (gxx + gyy);
        qrot(q, v);
        qrot(h, v);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            if (col < 3)
                A[i][col] += v[i];
            else
                c[i] += v[i];
        }
    }
}

__global__ __launch_bounds__(256) void surrogate_maps_kernel(hb_case cs, const float* __restrict__ params,
                                                             const float* __restrict__ norm,
                                                             const float* __restrict__ first, int n_grid,
                                                             hb_sur_pulse sp, float* __restrict__ maps,
                                                             int* __restrict__ outside) {
    __shared__ __align__(16) float w[kParams];
    stage_params(w, params);
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int np = cs.n_pad;
    const Spin p = load_spin(cs.spins, cs.pos_lo, np, j);
    const double u = ((double)p.px + (double)p.lx) * sp.g[0] + ((double)p.py + (double)p.ly) * sp.g[1] +
                     ((double)p.pz + (double)p.lz) * sp.g[2];
    const double f = (double)p.db0 + u;
    float xi[kIn];
    features(norm, (float)f, p.b1, xi);
    const bool out = fabsf(xi[0]) > 1.0f || fabsf(xi[1]) > 1.0f;
    outside[j] = (j < cs.n && out) ? 1 : 0;
    float a1[kH], a2[kH];
    layer1(w, xi, a1);
    layer2(w, a1, a2);
    double A[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}}, c[3] = {0.0, 0.0, 0.0};
    const double pd = p.pd, r1 = p.r1, r2 = p.r2;
    for (int b = 0; b < kM; ++b) {
        Quat q;
        const float* mu = norm + kNormHead + 4 * b;
        const float* sd = mu + kOut;
        float q1[4];
        first_quat(first, n_grid, norm, (float)f, p.b1 * norm[5], b, q1);
        q.w = q1[0] + (mu[0] + sd[0] * layer3(w, a2, 4 * b));
        q.x = q1[1] + (mu[1] + sd[1] * layer3(w, a2, 4 * b + 1));
        q.y = q1[2] + (mu[2] + sd[2] * layer3(w, a2, 4 * b + 2));
        q.z = q1[3] + (mu[3] + sd[3] * layer3(w, a2, 4 * b + 3));
        const double inv = 1.0 / sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
        q.w *= inv, q.x *= inv, q.y *= inv, q.z *= inv;
        const double tau = sp.blocks[2 * b], area = sp.blocks[2 * b + 1];
        const Quat h = rot_z(-M_PI * (f * tau + area));  // Rz(theta / 2), theta = -2 pi (f tau + B)
        const double e1 = exp(-0.5 * tau * r1), e2 = exp(-0.5 * tau * r2);
        double A0[3][3], c0[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
            for (int k = 0; k < 3; ++k) A0[i][k] = A[i][k];
            c0[i] = c[i];
        }
        relax_map(A, c, e1, e2, pd);
        rotate_map(A, c, qmul(h, qmul(q, h)));
        relax_map(A, c, e1, e2, pd);
        if (first) {
            float v[7];
            first_lookup(first, n_grid, norm, (float)f, b, v);
            const double bh = (double)p.b1 * norm[5];
            const double e[5] = {bh * v[2], bh * v[3], bh * bh * v[4], bh * bh * v[5], bh * bh * v[6]};
            split_correction(A, c, A0, c0, h, q, e, r1, r2, pd);
        }
    }
    float* o = maps + (size_t)sp.pid * 12 * np + j;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) o[(3 * i + k) * (size_t)np] = (float)(A[i][k] - (i == k ? 1.0 : 0.0));
#pragma unroll
    for (int i = 0; i < 3; ++i) o[(9 + i) * (size_t)np] = (float)c[i];
}

int launch_train_grad(const hb_mlp* net, const float* x, const float* y, int32_t n, float* work, void* stream) {
    hipLaunchKernelGGL(mlp_grad_mfma_kernel, dim3((n + kGroup - 1) / kGroup), dim3(256), 0, (hipStream_t)stream,
                       net->params, net->norm, x, y, n, work);
    return hipGetLastError();
}

}  // namespace hb

using namespace hb;

static int check_net(const hb_mlp* net) {
    if (!net || !net->params || !net->norm || net->hidden != kH || net->n_out != kOut) return hipErrorInvalidValue;
    if (net->first && net->n_grid < 2) return hipErrorInvalidValue;
    return hipSuccess;
}

// The training kernel reads the parameters as float4.
static int check_train_net(const hb_mlp* net) {
    if (int e = check_net(net)) return e;
    return reinterpret_cast<uintptr_t>(net->params) % 16 ? hipErrorInvalidValue : hipSuccess;
}

extern "C" int64_t hb_mlp_workspace(int32_t n) {
    if (n < 1) return 0;
    return (int64_t)((n + kGroup - 1) / kGroup) * kParams;  // one gradient partial per workgroup of kGroup samples
}

extern "C" int hb_spinor_blocks(const double* table, const int32_t* bounds, const float* x, int32_t n, float* y,
                                void* stream) {
    if (!table || !bounds || !x || !y || n < 0) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(spinor_blocks_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, table, bounds,
                       x, n, y);
    return hipGetLastError();
}

extern "C" int hb_mlp_grad(const hb_mlp* net, const float* x, const float* y, int32_t n, float* work, float* grad,
                           void* stream) {
    if (int e = check_train_net(net)) return e;
    if (!x || !y || !work || !grad || n < 1) return hipErrorInvalidValue;
    if (int e = launch_train_grad(net, x, y, n, work, stream)) return e;
    const int chunks = (n + kGroup - 1) / kGroup;
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((kParams + kAdamParams - 1) / kAdamParams), dim3(256), 0, (hipStream_t)stream, work,
                       chunks, net->params, net->adam_m, net->adam_v, 0.0f, 0.0f, 0.0f, grad);
    return hipGetLastError();
}

extern "C" int hb_mlp_train_step(const hb_mlp* net, const float* x, const float* y, int32_t n, float* work,
                                 int32_t step, float lr, void* stream) {
    if (int e = check_train_net(net)) return e;
    if (!net->adam_m || !net->adam_v || !x || !y || !work || n < 1 || step < 1) return hipErrorInvalidValue;
    if (int e = launch_train_grad(net, x, y, n, work, stream)) return e;
    const int chunks = (n + kGroup - 1) / kGroup;
    const float c1 = (float)(1.0 / (1.0 - pow(0.9, step))), c2 = (float)(1.0 / (1.0 - pow(0.999, step)));
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((kParams + kAdamParams - 1) / kAdamParams), dim3(256), 0, (hipStream_t)stream, work,
                       chunks, net->params, net->adam_m, net->adam_v, lr, c1, c2, (float*)nullptr);
    return hipGetLastError();
}

extern "C" int hb_mlp_train_steps(const hb_mlp* net, const float* x, const float* y, int32_t batch, int32_t n_steps,
                                  float* work, int32_t step0, int32_t total_steps, float lr0, float lr1,
                                  void* stream) {
    if (int e = check_train_net(net)) return e;
    if (!net->adam_m || !net->adam_v || !x || !y || !work || batch < 1 || n_steps < 0 || step0 < 0 ||
        total_steps < step0 + n_steps)
        return hipErrorInvalidValue;
    const int chunks = (batch + kGroup - 1) / kGroup;
    float* partial = work;
    for (int i = 0; i < n_steps; ++i) {
        const int t = step0 + i + 1;
        const double lr = lr1 + 0.5 * (lr0 - lr1) * (1.0 + cos(M_PI * (t - 1) / (double)total_steps));
        const float c1 = (float)(1.0 / (1.0 - pow(0.9, t))), c2 = (float)(1.0 / (1.0 - pow(0.999, t)));
        if (int e = launch_train_grad(net, x + (size_t)i * batch * kRaw, y + (size_t)i * batch * kOut, batch, work,
                                      stream))
            return e;
        hipLaunchKernelGGL(mlp_adam_kernel, dim3((kParams + kAdamParams - 1) / kAdamParams), dim3(256), 0, (hipStream_t)stream, partial,
                           chunks, net->params, net->adam_m, net->adam_v, (float)lr, c1, c2, (float*)nullptr);
    }
    return hipGetLastError();
}

extern "C" int hb_mlp_eval(const hb_mlp* net, const float* x, int32_t n, float* y, void* stream) {
    if (int e = check_net(net)) return e;
    if (!x || !y || n < 0) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_eval_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, net->params,
                       net->norm, x, n, y);
    return hipGetLastError();
}

extern "C" int hb_surrogate_maps(const hb_case* cs, const hb_mlp* net, const hb_sur_pulse* sp, float* maps,
                                 int32_t* outside, void* stream) {
    if (int e = check_net(net)) return e;
    if (!cs || !sp || !sp->blocks || !maps || !outside || cs->n_pad <= 0 || cs->n_pad % 256 || sp->pid < 0 ||
        sp->pid >= cs->n_pulses)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(surrogate_maps_kernel, dim3(cs->n_pad / 256), dim3(256), 0, (hipStream_t)stream, *cs,
                       net->params, net->norm, net->first, net->n_grid, *sp, maps, outside);
    return hipGetLastError();
}

extern "C" int hb_sur_first_order(const double* table, const int32_t* bounds, double f_lo, double df, int32_t n_grid,
                                  float* first, void* stream) {
    if (!table || !bounds || !first || n_grid < 2 || !(df > 0.0)) return hipErrorInvalidValue;
    const int total = kM * n_grid;
    hipLaunchKernelGGL(first_order_kernel, dim3((total + 255) / 256), dim3(256), 0, (hipStream_t)stream, table, bounds,
                       f_lo, df, n_grid, first);
    return hipGetLastError();
}

extern "C" int hb_sur_first_quats(const hb_mlp* net, const float* x, int32_t n, float* q, void* stream) {
    if (int e = check_net(net)) return e;
    if (!x || !q || n < 0) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(first_quats_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, net->first,
                       net->n_grid, net->norm, x, n, q);
    return hipGetLastError();
}
