// Learned regrid predictor: a 3-layer convolutional network over one leaf
// block's phase field and velocity that predicts whether the block's region
// is refined (channel 0) or derefined (channel 1) k steps ahead. No
// counterpart in the reference; the plain gradient criterion
// (Refine_Derefine.f90:18-54, :327-352) stays the judge.
//
// Network, per block sample: input 3 x 18 x 18 (C with its first ghost ring;
// Ux, Uy scaled by kUScale with a zero ring), conv1 3 -> 16 (3x3, valid) +
// ReLU, conv2 16 -> 16 (3x3, zero padding) + ReLU, conv3 16 -> 2 (3x3, zero
// padding), then the maximum over the 16 x 16 pixels of each channel: two
// logits. Weights BF16 on the MFMA (v_mfma_f32_16x16x16_bf16, FP32
// accumulate), master weights FP32 (Adam).
//
// Forward: one workgroup of 256 threads per sample; each convolution is an
// implicit GEMM, one 16 x 16 MFMA tile = one image row (16 pixels) x 16
// output channels, K = taps x input channels in slices of 16. Activations
// live in LDS as BF16. Backward: the maximum routes each logit's gradient to
// one pixel, so the gradients of conv3, conv2 and conv1 are non-zero only in
// the 3 x 3, 5 x 5 and 7 x 7 neighbourhoods of those two pixels; they are
// computed directly on the lanes and added to the FP32 gradient with atomics.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "amrlbm_internal.h"

namespace amrlbm {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round to nearest even, FP32 -> BF16 bits (finite inputs)
__device__ __forceinline__ unsigned short bf16_bits(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_float(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

// parameter offsets in the flat FP32 vector (CNN_NPARAM values)
constexpr int W1 = 0, B1 = W1 + 16 * 27, W2 = B1 + 16, B2 = W2 + 16 * 144, W3 = B2 + 16, B3 = W3 + 2 * 144;
static_assert(B3 + 2 == CNN_NPARAM, "parameter layout");

// LDS image of one sample: BF16 activations with a zero ring (18 x 18 each)
struct CnnLds {
    unsigned short x[3][18][18];
    unsigned short a1[16][18][18];
    unsigned short a2[16][18][18];
    float z3[2][16][16];
    unsigned short w1[32][16];    // [k = tap * 3 + cin (27 used)][cout]
    unsigned short w2[144][16];   // [k = tap * 16 + cin][cout]
    unsigned short w3[144][16];   // [k][cout], cout 0..1 used
    float bias[34];               // b1 (16), b2 (16), b3 (2)
    float red[2][256];
    int arg[2];
};

// conv forward for the rows of this wave: layer 1 (valid, Cin 3), 2 or 3 ('same', Cin 16)
template <int LAYER>
__device__ __forceinline__ void conv_rows(CnnLds &s, int wave, int lane)
{
    const int i = lane & 15, kb = lane >> 4;   // A row (pixel x) / B column (cout), K block
    for (int r = 0; r < 4; r++) {
        const int y = wave * 4 + r;
        f32x4 d = {0.f, 0.f, 0.f, 0.f};
        if (LAYER == 1) {
#pragma unroll
            for (int half = 0; half < 2; half++) {
                bf16x4 a, b;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int k = 16 * half + 4 * kb + e;
                    const int tap = k / 3, cin = k - 3 * tap;
                    a[e] = k < 27 ? (short)s.x[cin][y + tap / 3][i + tap % 3] : (short)0;
                    b[e] = (short)s.w1[k][i];
                }
                d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, d, 0, 0, 0);
            }
        } else {
            const unsigned short(*in)[18][18] = LAYER == 2 ? s.a1 : s.a2;
            const unsigned short(*w)[16] = LAYER == 2 ? s.w2 : s.w3;
#pragma unroll
            for (int tap = 0; tap < 9; tap++) {
                bf16x4 a, b;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int cin = 4 * kb + e;
                    a[e] = (short)in[cin][y + tap / 3][i + tap % 3];
                    b[e] = (short)w[tap * 16 + cin][i];
                }
                d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, d, 0, 0, 0);
            }
        }
        // D[pixel x = 4 kb + e][cout = i]
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int x = 4 * kb + e, c = i;
            if (LAYER == 1) {
                const float v = d[e] + s.bias[c];
                s.a1[c][y + 1][x + 1] = bf16_bits(v > 0.f ? v : 0.f);
            } else if (LAYER == 2) {
                const float v = d[e] + s.bias[16 + c];
                s.a2[c][y + 1][x + 1] = bf16_bits(v > 0.f ? v : 0.f);
            } else if (c < 2) {
                s.z3[c][y][x] = d[e] + s.bias[32 + c];
            }
        }
    }
}

// Forward pass of n samples to logits (n x 2, when logits is set). Sample k is
// row idx[k] (k when idx is null) of X (rows of 3 x 18 x 18 FP32) and Y. With
// Y (rows of 2 targets in {0, 1}) also the weighted logistic loss and its
// gradient, added into grad (CNN_NPARAM) and loss[0]. pos_weight multiplies
// the loss of positive targets; every sample's loss is divided by n.
__global__ __launch_bounds__(256) void k_cnn(const float *__restrict__ X, const int *__restrict__ idx, int n,
                                            const float *__restrict__ wts, float *__restrict__ logits,
                                            const float *__restrict__ Y, float pos_weight, float *__restrict__ grad,
                                            float *__restrict__ loss)
{
    const int smp = blockIdx.x, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    if (smp >= n) return;
    const size_t row = idx ? (size_t)idx[smp] : (size_t)smp;
    __shared__ CnnLds s;
    // weights to BF16, K-major per output channel
    for (int t = tid; t < 32 * 16; t += 256) {
        const int k = t >> 4, c = t & 15;
        s.w1[k][c] = k < 27 ? bf16_bits(wts[W1 + c * 27 + k]) : (unsigned short)0;
    }
    for (int t = tid; t < 144 * 16; t += 256) {
        const int k = t >> 4, c = t & 15;
        s.w2[k][c] = bf16_bits(wts[W2 + c * 144 + k]);
        s.w3[k][c] = c < 2 ? bf16_bits(wts[W3 + c * 144 + k]) : (unsigned short)0;
    }
    if (tid < 16) s.bias[tid] = wts[B1 + tid];
    else if (tid < 32) s.bias[tid] = wts[B2 + tid - 16];
    else if (tid < 34) s.bias[tid] = wts[B3 + tid - 32];
    const float *x = X + row * 3 * 18 * 18;
    for (int t = tid; t < 3 * 18 * 18; t += 256) (&s.x[0][0][0])[t] = bf16_bits(x[t]);
    for (int t = tid; t < 16 * 18 * 18; t += 256) {
        (&s.a1[0][0][0])[t] = 0;
        (&s.a2[0][0][0])[t] = 0;
    }
    __syncthreads();
    conv_rows<1>(s, wave, lane);
    __syncthreads();
    conv_rows<2>(s, wave, lane);
    __syncthreads();
    conv_rows<3>(s, wave, lane);
    __syncthreads();
    // maximum over the pixels of each channel, first pixel on ties
    const int ch = tid >> 7, p0 = tid & 127;
    {
        float a = s.z3[ch][p0 >> 4][p0 & 15], b = s.z3[ch][(p0 + 128) >> 4][(p0 + 128) & 15];
        int ia = p0, ib = p0 + 128;
        float v = a;
        int iv = ia;
        if (b > a) {
            v = b;
            iv = ib;
        }
        s.red[ch][p0] = v;
        (reinterpret_cast<int *>(s.red[ch]))[128 + p0] = iv;
    }
    __syncthreads();
    for (int w = 64; w > 0; w >>= 1) {
        if (p0 < w) {
            float a = s.red[ch][p0], b = s.red[ch][p0 + w];
            int ia = (reinterpret_cast<int *>(s.red[ch]))[128 + p0], ib = (reinterpret_cast<int *>(s.red[ch]))[128 + p0 + w];
            if (b > a || (b == a && ib < ia)) {
                s.red[ch][p0] = b;
                (reinterpret_cast<int *>(s.red[ch]))[128 + p0] = ib;
            }
        }
        __syncthreads();
    }
    if (tid < 2) {
        if (logits) logits[(size_t)smp * 2 + tid] = s.red[tid][0];
        s.arg[tid] = (reinterpret_cast<int *>(s.red[tid]))[128];
    }
    if (!Y) return;
    __syncthreads();
    // weighted logistic loss and dL/dlogit
    __shared__ float dl[2];
    if (tid < 2) {
        const float z = s.red[tid][0], y = Y[row * 2 + tid];
        const float p = 1.f / (1.f + __expf(-z));
        const float wy = y > 0.5f ? pos_weight : 1.f;
        // -[w y log p + (1 - y) log(1 - p)], stable in z
        const float lp = z > 0.f ? -__logf(1.f + __expf(-z)) : z - __logf(1.f + __expf(z));   // log p
        const float lq = lp - z;                                                              // log(1 - p)
        atomicAdd(loss, -(wy * y * lp + (1.f - y) * lq) / (float)n);
        dl[tid] = (wy * y * (p - 1.f) + (1.f - y) * p) / (float)n;
    }
    __syncthreads();
    // conv3: dW3[c][k] = dl[c] * A2pad[cin][py + dy][px + dx], db3[c] = dl[c]
    for (int t = tid; t < 2 * 144; t += 256) {
        const int c = t / 144, k = t - 144 * c, tap = k >> 4, cin = k & 15;
        const int p = s.arg[c], py = p >> 4, px = p & 15;
        atomicAdd(&grad[W3 + c * 144 + k], dl[c] * bf16_float(s.a2[cin][py + tap / 3][px + tap % 3]));
    }
    if (tid < 2) atomicAdd(&grad[B3 + tid], dl[tid]);
    // dZ2 on the 3 x 3 neighbourhoods of the two pixels (7 x 7 window around
    // each, to reuse for conv2's input gradient): dA2[cin][q] = sum_c dl[c] W3[c][cin][tap]
    // for q = p_c + tap - 1, masked by A2 > 0
    __shared__ float dz2[2][16][9];   // [c][cin][tap]: gradient at pixel p_c + (tap - 4 offset)
    for (int t = tid; t < 2 * 16 * 9; t += 256) {
        const int c = t / 144, r = t - 144 * c, cin = r / 9, tap = r - 9 * cin;
        const int p = s.arg[c], qy = (p >> 4) + tap / 3 - 1, qx = (p & 15) + tap % 3 - 1;
        float v = 0.f;
        if (qy >= 0 && qy < 16 && qx >= 0 && qx < 16 && s.a2[cin][qy + 1][qx + 1] != 0)
            v = dl[c] * bf16_float(s.w3[tap * 16 + cin][c]);
        dz2[c][cin][tap] = v;
    }
    __syncthreads();
    // conv2 weights: dW2[cout][tap2*16 + cin] = sum over q of dZ2[cout][q] A1pad[cin][q + tap2]
    for (int t = tid; t < 16 * 144; t += 256) {
        const int cout = t / 144, k = t - 144 * cout, tap2 = k >> 4, cin = k & 15;
        float acc = 0.f;
        for (int c = 0; c < 2; c++) {
            const int p = s.arg[c];
            for (int tap = 0; tap < 9; tap++) {
                const float g = dz2[c][cout][tap];
                if (g == 0.f) continue;
                const int qy = (p >> 4) + tap / 3 - 1, qx = (p & 15) + tap % 3 - 1;
                acc += g * bf16_float(s.a1[cin][qy + tap2 / 3][qx + tap2 % 3]);
            }
        }
        if (acc != 0.f) atomicAdd(&grad[W2 + cout * 144 + k], acc);
    }
    if (tid < 16) {
        float acc = 0.f;
        for (int c = 0; c < 2; c++)
            for (int tap = 0; tap < 9; tap++) acc += dz2[c][tid][tap];
        if (acc != 0.f) atomicAdd(&grad[B2 + tid], acc);
    }
    // dZ1 on the 5 x 5 neighbourhoods: dA1[cin][r] = sum_{cout, tap2} dZ2[cout][q] W2[cout][cin][tap2],
    // q = r - tap2 + 1; r = p_c + (a - 2, b - 2), a, b in 0..4
    __shared__ float dz1[2][16][25];
    for (int t = tid; t < 2 * 16 * 25; t += 256) {
        const int c = t / 400, r0 = t - 400 * c, cin = r0 / 25, o = r0 - 25 * cin;
        const int p = s.arg[c], ry = (p >> 4) + o / 5 - 2, rx = (p & 15) + o % 5 - 2;
        float v = 0.f;
        if (ry >= 0 && ry < 16 && rx >= 0 && rx < 16 && s.a1[cin][ry + 1][rx + 1] != 0) {
            for (int tap2 = 0; tap2 < 9; tap2++) {
                // q = r - (tap2 offset): its position relative to p_c in the 3 x 3 window
                const int qy = o / 5 - 2 - (tap2 / 3 - 1), qx = o % 5 - 2 - (tap2 % 3 - 1);
                if (qy < -1 || qy > 1 || qx < -1 || qx > 1) continue;
                const int tap = (qy + 1) * 3 + (qx + 1);
                for (int cout = 0; cout < 16; cout++) {
                    const float g = dz2[c][cout][tap];
                    if (g != 0.f) v += g * bf16_float(s.w2[tap2 * 16 + cin][cout]);
                }
            }
        }
        dz1[c][cin][o] = v;
    }
    __syncthreads();
    // conv1 weights: dW1[cout][tap1*3 + cin] = sum over r of dZ1[cout][r] X[cin][r + tap1]
    for (int t = tid; t < 16 * 27; t += 256) {
        const int cout = t / 27, k = t - 27 * cout, tap1 = k / 3, cin = k - 3 * tap1;
        float acc = 0.f;
        for (int c = 0; c < 2; c++) {
            const int p = s.arg[c];
            for (int o = 0; o < 25; o++) {
                const float g = dz1[c][cout][o];
                if (g == 0.f) continue;
                const int ry = (p >> 4) + o / 5 - 2, rx = (p & 15) + o % 5 - 2;
                acc += g * bf16_float(s.x[cin][ry + tap1 / 3][rx + tap1 % 3]);
            }
        }
        if (acc != 0.f) atomicAdd(&grad[W1 + cout * 27 + k], acc);
    }
    if (tid < 16) {
        float acc = 0.f;
        for (int c = 0; c < 2; c++)
            for (int o = 0; o < 25; o++) acc += dz1[c][tid][o];
        if (acc != 0.f) atomicAdd(&grad[B1 + tid], acc);
    }
}

// Adam on the FP32 master weights; clears the gradient
__global__ void k_cnn_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v, float *__restrict__ g,
                           int np, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    g[i] = 0.f;
}

// Network input of every leaf in list (n entries): the signed distance to
// the interface that the phase field's tanh profile C = 1/2 + 1/2 tanh(2 d /
// W) gives, d = (W / 4) ln(C / (1 - C)) cells of the block's
// level, clamped to +-20 and divided by 16, on x, y in 0..17 (a field that
// BF16 resolves, where C itself reads 1 for every liquid cell within a few
// widths of the interface); Ux and Uy times kUScale on the interior, zero on
// the ring. W is the mesh's, or in an ensemble mesh the block's instance's.
template <class T>
__global__ __launch_bounds__(256) void k_cnn_features(DevMeshT<T> m, const int *__restrict__ list, int n,
                                                      float *__restrict__ X, double W)
{
    const int k = blockIdx.x, tid = threadIdx.x;
    if (k >= n) return;
    const int slot = list[k];
    const T *C = m.C + (size_t)slot * CS;
    const T *PU = m.PU + (size_t)slot * PUB;
    const bool dev = std::is_same<T, float>::value;   // FP32 storage: C - 1
    if (m.phs) W = m.phs[m.blk[slot].inst].W;
    float *x = X + (size_t)k * 3 * 18 * 18;
    for (int t = tid; t < 18 * 18; t += 256) {
        const int yy = t / 18, xx = t - 18 * yy;
        const double v = (double)C[(yy + 1) * CW + (xx + 1)];
        const double c = fmax(dev ? v + 1.0 : v, 1e-12), omc = fmax(dev ? -v : 1.0 - v, 1e-12);
        const double d = 0.25 * W * (log(c) - log(omc));
        x[t] = (float)(fmin(fmax(d, -20.0), 20.0) * (1.0 / 16.0));
        const bool in = xx >= 1 && xx <= NX && yy >= 1 && yy <= NY;
        x[18 * 18 + t] = in ? (float)PU[PS + (xx - 1) + NX * (yy - 1)] * kUScale : 0.f;
        x[2 * 18 * 18 + t] = in ? (float)PU[2 * PS + (xx - 1) + NX * (yy - 1)] * kUScale : 0.f;
    }
}

void launch_cnn(const float *X, const int *idx, int n, const float *w, float *logits, const float *Y,
                float pos_weight, float *grad, float *loss, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_cnn, dim3(n), dim3(256), 0, s, X, idx, n, w, logits, Y, pos_weight, grad, loss);
}

void launch_cnn_adam(float *w, float *m, float *v, float *g, float lr, int t, hipStream_t s)
{
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_cnn_adam, dim3((CNN_NPARAM + 255) / 256), dim3(256), 0, s, w, m, v, g, CNN_NPARAM, lr, b1, b2,
                       1e-8f, c1, c2);
}

void launch_cnn_features(const DevMesh &m, const int *list, int n, float *X, double W, hipStream_t s)
{
    if (n <= 0) return;
    if (m.f32)
        hipLaunchKernelGGL(k_cnn_features<float>, dim3(n), dim3(256), 0, s, typed<float>(m), list, n, X, W);
    else
        hipLaunchKernelGGL(k_cnn_features<double>, dim3(n), dim3(256), 0, s, typed<double>(m), list, n, X, W);
}

}  // namespace amrlbm
