// The learned RF surrogate (no counterpart in the reference; it stands where the reference composes a reused pulse
// per spin, src/blochparty/engines/gpu/cuda/pulse.cu.in:3-32).
//
// A reused pulse of K substeps is cut into M = HB_SUR_BLOCKS blocks. Each block's pure rotation is a unit quaternion
// taken in the frame that turns with the off-resonance precession: Q~_j = Rz(-theta_j/2) Q_j Rz(-theta_j/2),
// theta_j = -2 pi (f tau_j + B_j) the block's free-precession angle (f the spin's offset, tau_j the block's length,
// B_j its b0_offset area), so that Q~_j varies slowly with f. A network 3 -> 64 -> 64 -> 4 M (tanh) maps a spin's
// (df, u, b1) (df its off-resonance [Hz], u = g . r its gradient offset during the pulse [Hz], b1 its B1 scale) to
// the M quaternions. A spin's pulse map m -> m + D m + v is then composed in float64 with the relaxation put back
// exactly, half a block before and after each rotation: a Strang split of the pulse into M blocks.
//
// Kernels: the exact spinor products that make the training data (float64); the network's training step (a fused
// per-sample forward + backward, per-chunk weight gradients on the matrix cores summed in a fixed order, Adam); its
// evaluation; and the
// per-spin maps, with spins outside the trained input box flagged for the exact composition (hb_pulse_maps' mask).
#include "bloch_device.h"

namespace hb {

constexpr int kH = HB_MLP_HIDDEN;
constexpr int kM = HB_SUR_BLOCKS;
constexpr int kOut = 4 * kM;
constexpr int kW1 = 0, kB1 = kW1 + kH * 3, kW2 = kB1 + kH, kB2 = kW2 + kH * kH, kW3 = kB2 + kH,
              kB3 = kW3 + kOut * kH, kParams = kB3 + kOut;
static_assert(kParams == HB_MLP_PARAMS, "parameter layout");
constexpr int kChunk = 512;  // samples per weight-gradient partial

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
    const double df = x[3 * i], u = x[3 * i + 1], b1 = x[3 * i + 2];
    const double w0 = -2.0 * M_PI;
    for (int j = 0; j < kM; ++j) {
        Quat q = {1.0, 0.0, 0.0, 0.0};
        double theta = 0.0;
        for (int k = bounds[j]; k < bounds[j + 1]; ++k) {
            const double* r = table + (size_t)k * 8;
            const double w = w0 * r[0], f = df + u + r[6];
            q = qmul(rot_quat(w * b1 * r[1], w * b1 * r[2], w * f), q);
            theta += w * f;
        }
        const Quat h = rot_z(-0.5 * theta);
        q = qmul(h, qmul(q, h));
        float* out = y + (size_t)i * kOut + 4 * j;
        out[0] = (float)q.w, out[1] = (float)q.x, out[2] = (float)q.y, out[3] = (float)q.z;
    }
}

// ---- the network ----

// Parameters staged in LDS once per workgroup; every lane reads the same address (a broadcast).
__device__ __forceinline__ void stage_params(float* __restrict__ lds, const float* __restrict__ p) {
    for (int i = threadIdx.x; i < kParams; i += blockDim.x) lds[i] = p[i];
    __syncthreads();
}

__device__ __forceinline__ void layer1(const float* __restrict__ w, const float (&x)[3], float (&a1)[kH]) {
#pragma unroll
    for (int i = 0; i < kH; ++i)
        a1[i] = tanhf(w[kB1 + i] + w[kW1 + 3 * i] * x[0] + w[kW1 + 3 * i + 1] * x[1] + w[kW1 + 3 * i + 2] * x[2]);
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

__device__ __forceinline__ void normalise_input(const float* __restrict__ norm, float df, float u, float b1,
                                                float (&x)[3]) {
    x[0] = (df - norm[0]) / norm[3];
    x[1] = (u - norm[1]) / norm[4];
    x[2] = (b1 - norm[2]) / norm[5];
}

// One sample per lane: forward, the loss's output deltas d3 = 2 (out - t) / (n n_out) (t the normalised target),
// and backpropagation to d2, d1. Stores the layer inputs (with a ones row for the biases) and the deltas as (rows, n)
// arrays for the weight-gradient kernel: ws = [A0 (4, n) | A1 (65, n) | A2 (65, n) | D1 (64, n) | D2 (64, n) |
// D3 (n_out, n)].
__global__ __launch_bounds__(256) void mlp_fwd_bwd_kernel(const float* __restrict__ params,
                                                          const float* __restrict__ norm, const float* __restrict__ x,
                                                          const float* __restrict__ y, int n, float* __restrict__ ws) {
    __shared__ __align__(16) float w[kParams];
    stage_params(w, params);
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float* A0 = ws;
    float* A1 = A0 + (size_t)4 * n;
    float* A2 = A1 + (size_t)(kH + 1) * n;
    float* D1 = A2 + (size_t)(kH + 1) * n;
    float* D2 = D1 + (size_t)kH * n;
    float* D3 = D2 + (size_t)kH * n;
    float xi[3];
    normalise_input(norm, x[3 * s], x[3 * s + 1], x[3 * s + 2], xi);
#pragma unroll
    for (int c = 0; c < 3; ++c) A0[(size_t)c * n + s] = xi[c];
    A0[(size_t)3 * n + s] = 1.0f;
    float a1[kH], a2[kH], g2[kH];
    layer1(w, xi, a1);
    layer2(w, a1, a2);
#pragma unroll
    for (int j = 0; j < kH; ++j) {
        A1[(size_t)j * n + s] = a1[j];
        A2[(size_t)j * n + s] = a2[j];
        g2[j] = 0.0f;
    }
    A1[(size_t)kH * n + s] = 1.0f;
    A2[(size_t)kH * n + s] = 1.0f;
    const float scale = 2.0f / ((float)n * (float)kOut);
    const float* mu = norm + 6;
    const float* sd = mu + kOut;
    for (int k = 0; k < kOut; ++k) {
        const float t = (y[(size_t)s * kOut + k] - mu[k]) / sd[k];
        const float d3 = scale * (layer3(w, a2, k) - t);
        D3[(size_t)k * n + s] = d3;
        const float4* row = reinterpret_cast<const float4*>(w + kW3 + kH * k);
#pragma unroll
        for (int i = 0; i < kH / 4; ++i) {
            const float4 r = row[i];
            g2[4 * i] += r.x * d3, g2[4 * i + 1] += r.y * d3, g2[4 * i + 2] += r.z * d3, g2[4 * i + 3] += r.w * d3;
        }
    }
#pragma unroll
    for (int j = 0; j < kH; ++j) {
        g2[j] *= 1.0f - a2[j] * a2[j];
        D2[(size_t)j * n + s] = g2[j];
    }
#pragma unroll
    for (int i = 0; i < kH; ++i) {
        float acc = 0.0f;
#pragma unroll
        for (int j = 0; j < kH; ++j) acc += w[kW2 + kH * j + i] * g2[j];
        D1[(size_t)i * n + s] = acc * (1.0f - a1[i] * a1[i]);
    }
}

typedef float f4 __attribute__((ext_vector_type(4)));

// Weight gradients as matrix products on the matrix cores: for each layer, dW|db (rows x (k_in + 1)) =
// D (rows x n) A^T (n x (k_in + 1)), A's last row the ones of the bias. One wave per 16 x 16 output tile and chunk of
// kChunk samples, v_mfma_f32_16x16x4f32 over the chunk's samples 4 at a time (lane l supplies D[row m][s] and
// A[col m][s] with m = l % 16, s = 4 u + l / 16; it receives D_out[4 (l / 16) + r][l % 16]); the tile's entries go to
// partial[chunk][parameter], summed over the chunks in order by mlp_adam_kernel. Tiles: layer 1 4 x 1, layer 2
// 4 x 5, layer 3 3 x 5 (rows and columns past the layer read as zeros).
constexpr int kTiles = 4 * 1 + 4 * 5 + 3 * 5;

__device__ __forceinline__ int param_index(int layer, int r, int c, int k_in) {
    const int w_off = layer == 0 ? kW1 : layer == 1 ? kW2 : kW3, b_off = layer == 0 ? kB1 : layer == 1 ? kB2 : kB3;
    return c < k_in ? w_off + r * k_in + c : b_off + r;
}

__global__ __launch_bounds__(64) void mlp_wgrad_kernel(const float* __restrict__ ws, int n,
                                                       float* __restrict__ partial) {
    const int lane = threadIdx.x, t = blockIdx.x, chunk = blockIdx.y;
    const float* A0 = ws;
    const float* A1 = A0 + (size_t)4 * n;
    const float* A2 = A1 + (size_t)(kH + 1) * n;
    const float* D1 = A2 + (size_t)(kH + 1) * n;
    const float* D2 = D1 + (size_t)kH * n;
    const float* D3 = D2 + (size_t)kH * n;
    int layer, rt, ct;
    if (t < 4) layer = 0, rt = t, ct = 0;
    else if (t < 24) layer = 1, rt = (t - 4) / 5, ct = (t - 4) % 5;
    else layer = 2, rt = (t - 24) / 5, ct = (t - 24) % 5;
    const float* D = layer == 0 ? D1 : layer == 1 ? D2 : D3;
    const float* A = layer == 0 ? A0 : layer == 1 ? A1 : A2;
    const int rows = layer == 2 ? kOut : kH, k_in = layer == 0 ? 3 : kH, cols = k_in + 1;
    const int m = lane & 15, kq = lane >> 4;
    const int row = rt * 16 + m, col = ct * 16 + m;
    const float* dr = D + (size_t)(row < rows ? row : 0) * n;
    const float* ar = A + (size_t)(col < cols ? col : 0) * n;
    const int s0 = chunk * kChunk, s1 = min(n, s0 + kChunk);
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int s = s0 + kq; s < s0 + kChunk; s += 4) {
        const bool ok = s < s1;
        const float a = (ok && row < rows) ? dr[s] : 0.0f;
        const float b = (ok && col < cols) ? ar[s] : 0.0f;
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int orow = rt * 16 + 4 * kq + r, ocol = ct * 16 + m;
        if (orow < rows && ocol < cols) partial[(size_t)chunk * kParams + param_index(layer, orow, ocol, k_in)] = acc[r];
    }
}

// grad = the chunks' partials summed in chunk order; then Adam (bias corrections c1 = 1/(1 - beta1^t),
// c2 = 1/(1 - beta2^t) from the host) unless grad_out is set, which receives the gradient instead.
__global__ __launch_bounds__(256) void mlp_adam_kernel(const float* __restrict__ partial, int chunks,
                                                       float* __restrict__ params, float* __restrict__ m,
                                                       float* __restrict__ v, float lr, float c1, float c2,
                                                       float* __restrict__ grad_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    float g = 0.0f;
    for (int c = 0; c < chunks; ++c) g += partial[(size_t)c * kParams + i];
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
    float xi[3], a1[kH], a2[kH];
    normalise_input(norm, x[3 * s], x[3 * s + 1], x[3 * s + 2], xi);
    layer1(w, xi, a1);
    layer2(w, a1, a2);
    for (int k = 0; k < kOut; ++k) y[(size_t)s * kOut + k] = norm[6 + k] + norm[6 + kOut + k] * layer3(w, a2, k);
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

__global__ __launch_bounds__(256) void surrogate_maps_kernel(hb_case cs, const float* __restrict__ params,
                                                             const float* __restrict__ norm, hb_sur_pulse sp,
                                                             float* __restrict__ maps, int* __restrict__ outside) {
    __shared__ __align__(16) float w[kParams];
    stage_params(w, params);
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int np = cs.n_pad;
    const Spin p = load_spin(cs.spins, cs.pos_lo, np, j);
    const double u = ((double)p.px + (double)p.lx) * sp.g[0] + ((double)p.py + (double)p.ly) * sp.g[1] +
                     ((double)p.pz + (double)p.lz) * sp.g[2];
    const double f = (double)p.db0 + u;
    float xi[3];
    normalise_input(norm, p.db0, (float)u, p.b1, xi);
    const bool out = fabsf(xi[0]) > 1.0f || fabsf(xi[1]) > 1.0f || fabsf(xi[2]) > 1.0f;
    outside[j] = (j < cs.n && out) ? 1 : 0;
    float a1[kH], a2[kH];
    layer1(w, xi, a1);
    layer2(w, a1, a2);
    double A[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}}, c[3] = {0.0, 0.0, 0.0};
    const double pd = p.pd, r1 = p.r1, r2 = p.r2;
    for (int b = 0; b < kM; ++b) {
        Quat q;
        q.w = norm[6 + 4 * b] + norm[6 + kOut + 4 * b] * layer3(w, a2, 4 * b);
        q.x = norm[7 + 4 * b] + norm[6 + kOut + 4 * b + 1] * layer3(w, a2, 4 * b + 1);
        q.y = norm[8 + 4 * b] + norm[6 + kOut + 4 * b + 2] * layer3(w, a2, 4 * b + 2);
        q.z = norm[9 + 4 * b] + norm[6 + kOut + 4 * b + 3] * layer3(w, a2, 4 * b + 3);
        const double inv = 1.0 / sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
        q.w *= inv, q.x *= inv, q.y *= inv, q.z *= inv;
        const double tau = sp.blocks[2 * b], area = sp.blocks[2 * b + 1];
        const Quat h = rot_z(-M_PI * (f * tau + area));  // Rz(theta / 2), theta = -2 pi (f tau + B)
        const double e1 = exp(-0.5 * tau * r1), e2 = exp(-0.5 * tau * r2);
        relax_map(A, c, e1, e2, pd);
        rotate_map(A, c, qmul(h, qmul(q, h)));
        relax_map(A, c, e1, e2, pd);
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
    hipStream_t st = (hipStream_t)stream;
    const int chunks = (n + kChunk - 1) / kChunk;
    // one wave per workgroup: a batch of 8,192 samples is 128 workgroups (the kernel runs at one wave per SIMD)
    hipLaunchKernelGGL(mlp_fwd_bwd_kernel, dim3((n + 63) / 64), dim3(64), 0, st, net->params, net->norm, x, y, n,
                       work);
    float* partial = work + (size_t)(4 + 2 * (kH + 1) + 2 * kH + kOut) * n;
    hipLaunchKernelGGL(mlp_wgrad_kernel, dim3(kTiles, chunks), dim3(64), 0, st, work, n, partial);
    return hipGetLastError();
}

}  // namespace hb

using namespace hb;

static int check_net(const hb_mlp* net) {
    if (!net || !net->params || !net->norm || net->hidden != kH || net->n_out != kOut) return hipErrorInvalidValue;
    return hipSuccess;
}

extern "C" int64_t hb_mlp_workspace(int32_t n) {
    if (n < 1) return 0;
    const int64_t chunks = (n + kChunk - 1) / kChunk;
    return (int64_t)(4 + 2 * (kH + 1) + 2 * kH + kOut) * n + chunks * kParams;
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
    if (int e = check_net(net)) return e;
    if (!x || !y || !work || !grad || n < 1) return hipErrorInvalidValue;
    if (int e = launch_train_grad(net, x, y, n, work, stream)) return e;
    const int chunks = (n + kChunk - 1) / kChunk;
    float* partial = work + hb_mlp_workspace(n) - (int64_t)chunks * kParams;
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, (hipStream_t)stream, partial,
                       chunks, net->params, net->adam_m, net->adam_v, 0.0f, 0.0f, 0.0f, grad);
    return hipGetLastError();
}

extern "C" int hb_mlp_train_step(const hb_mlp* net, const float* x, const float* y, int32_t n, float* work,
                                 int32_t step, float lr, void* stream) {
    if (int e = check_net(net)) return e;
    if (!net->adam_m || !net->adam_v || !x || !y || !work || n < 1 || step < 1) return hipErrorInvalidValue;
    if (int e = launch_train_grad(net, x, y, n, work, stream)) return e;
    const int chunks = (n + kChunk - 1) / kChunk;
    float* partial = work + hb_mlp_workspace(n) - (int64_t)chunks * kParams;
    const float c1 = (float)(1.0 / (1.0 - pow(0.9, step))), c2 = (float)(1.0 / (1.0 - pow(0.999, step)));
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, (hipStream_t)stream, partial,
                       chunks, net->params, net->adam_m, net->adam_v, lr, c1, c2, (float*)nullptr);
    return hipGetLastError();
}

extern "C" int hb_mlp_train_steps(const hb_mlp* net, const float* x, const float* y, int32_t batch, int32_t n_steps,
                                  float* work, int32_t step0, int32_t total_steps, float lr0, float lr1,
                                  void* stream) {
    if (int e = check_net(net)) return e;
    if (!net->adam_m || !net->adam_v || !x || !y || !work || batch < 1 || n_steps < 0 || step0 < 0 ||
        total_steps < step0 + n_steps)
        return hipErrorInvalidValue;
    const int chunks = (batch + kChunk - 1) / kChunk;
    float* partial = work + hb_mlp_workspace(batch) - (int64_t)chunks * kParams;
    for (int i = 0; i < n_steps; ++i) {
        const int t = step0 + i + 1;
        const double lr = lr1 + 0.5 * (lr0 - lr1) * (1.0 + cos(M_PI * (t - 1) / (double)total_steps));
        const float c1 = (float)(1.0 / (1.0 - pow(0.9, t))), c2 = (float)(1.0 / (1.0 - pow(0.999, t)));
        if (int e = launch_train_grad(net, x + (size_t)i * batch * 3, y + (size_t)i * batch * kOut, batch, work,
                                      stream))
            return e;
        hipLaunchKernelGGL(mlp_adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, (hipStream_t)stream, partial,
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
                       net->params, net->norm, *sp, maps, outside);
    return hipGetLastError();
}
