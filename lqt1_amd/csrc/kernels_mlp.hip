// Surrogate MLP kernels: BF16 operands on the gfx950 matrix cores
// (v_mfma_f32_32x32x16_bf16) with FP32 accumulation, GELU epilogues, the
// loss and its gradient, bias reductions and Adam on FP32 master weights.
// The reference has no learned component; these kernels train and evaluate
// the dose-sweep surrogate on strands produced by the strand kernel.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp.hpp"

namespace lqt {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
    // d/dx [x Phi(x)] = Phi(x) + x phi(x)
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * __expf(-0.5f * x * x);
}

// One wavefront per 32x32 tile of C = A . B^T (A: M x K, B: N x K, both
// row-major bf16 with K contiguous; K a multiple of 16, M and N multiples
// of 32).  Lane l loads row l%32 of its operand, k-half l/32 (8 bf16 = 16
// bytes); accumulator j of lane l is C[8*(j/4) + 4*(l/32) + j%4][l%32].
__global__ __launch_bounds__(64) void gemm_bt_kernel(const __bf16* __restrict__ A, int lda,
                                                      const __bf16* __restrict__ B, int ldb, int K, GemmEpi e) {
    const int tm = blockIdx.x * 32, tn = blockIdx.y * 32;
    const int l = threadIdx.x;
    const int r = l & 31, kh = (l >> 5) * 8;
    const __bf16* pa = A + (int64_t)(tm + r) * lda + kh;
    const __bf16* pb = B + (int64_t)(tn + r) * ldb + kh;
    f32x16 acc = {};
    for (int k0 = 0; k0 < K; k0 += 16) {
        const bf16x8 a = *(const bf16x8*)(pa + k0);
        const bf16x8 b = *(const bf16x8*)(pb + k0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    const int col = tn + r;
    const float bias = e.bias ? e.bias[col] : 0.0f;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int row = tm + 8 * (j >> 2) + 4 * (l >> 5) + (j & 3);
        float v = acc[j] + bias;
        if (e.c32) e.c32[(int64_t)row * e.ldc32 + col] = v;
        if (e.pre) v *= gelu_grad(e.pre[(int64_t)row * e.ldpre + col]);
        float o = e.act ? gelu(v) : v;
        if (col >= e.n_valid || row >= e.m_valid) o = 0.0f;
        if (e.cb) e.cb[(int64_t)row * e.ldcb + col] = (__bf16)o;
        if (e.cbt) e.cbt[(int64_t)col * e.ldcbt + row] = (__bf16)o;
    }
}

// Mean loss and its gradient w.r.t. the raw outputs: squared error on the
// first n_reg outputs (weighted by mask), binary cross-entropy with logits
// on the next n_bce outputs.  Writes dOut (bf16, row-major and transposed)
// and adds the batch loss / M to *loss.
__global__ __launch_bounds__(256) void loss_kernel(const float* __restrict__ out, int ldo, const float* __restrict__ y,
                                                   const float* __restrict__ mask, int ny, int n_reg, int n_bce, int M,
                                                   int m_valid, __bf16* __restrict__ dout, int ldd,
                                                   __bf16* __restrict__ doutt, int lddt, float* __restrict__ loss) {
    __shared__ float red[256];
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    float L = 0.0f;
    if (m < M) {
        for (int j = 0; j < ldd; ++j) {
            float g = 0.0f;
            if (m < m_valid && j < n_reg + n_bce) {
                const float o = out[(int64_t)m * ldo + j];
                const float t = y[(int64_t)m * ny + j];
                const float w = mask ? mask[(int64_t)m * ny + j] : 1.0f;
                if (j < n_reg) {
                    const float d = o - t;
                    L += w * d * d;
                    g = 2.0f * w * d;
                } else {
                    // log(1 + e^o) - t o, stable form
                    L += w * (fmaxf(o, 0.0f) - o * t + log1pf(__expf(-fabsf(o))));
                    g = w * (1.0f / (1.0f + __expf(-o)) - t);
                }
                g /= (float)m_valid;
            }
            dout[(int64_t)m * ldd + j] = (__bf16)g;
            doutt[(int64_t)j * lddt + m] = (__bf16)g;
        }
    }
    red[threadIdx.x] = L;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0] / (float)m_valid);
}

// db[n] = sum over m of G^T[n][m] (bf16 rows of length M).
__global__ __launch_bounds__(256) void rowsum_kernel(const __bf16* __restrict__ gt, int ld, int M, float* __restrict__ db) {
    __shared__ float red[256];
    const int n = blockIdx.x;
    float s = 0.0f;
    for (int m = threadIdx.x; m < M; m += blockDim.x) s += (float)gt[(int64_t)n * ld + m];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[n] = red[0];
}

// Adam on FP32 masters (g = gradient), then refresh the bf16 weight copy
// W[rows x ld] and its transpose Wt[ld_t x rows] for weights (is_w), or
// just the masters for biases.  step[0] is the 1-based step count.
__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                                                   float* __restrict__ mo, float* __restrict__ ve, int n,
                                                   const int* __restrict__ step, float lr, float b1, float b2,
                                                   float eps, int cols, __bf16* __restrict__ wb, int ldw,
                                                   __bf16* __restrict__ wt, int ldt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float t = (float)(*step);
    const float gi = g[i];
    const float m = b1 * mo[i] + (1.0f - b1) * gi;
    const float v = b2 * ve[i] + (1.0f - b2) * gi * gi;
    mo[i] = m;
    ve[i] = v;
    const float mh = m / (1.0f - powf(b1, t));
    const float vh = v / (1.0f - powf(b2, t));
    const float w = p[i] - lr * mh / (sqrtf(vh) + eps);
    p[i] = w;
    if (wb) {
        const int r = i / cols, c = i % cols;
        wb[(int64_t)r * ldw + c] = (__bf16)w;
        wt[(int64_t)c * ldt + r] = (__bf16)w;
    }
}

// ctr[0]: Adam step t (1-based, kept across calls); ctr[1]: batch index
// within the current training call.
__global__ void step_kernel(int* ctr, float* loss_log, const float* loss) {
    if (threadIdx.x == 0) {
        loss_log[ctr[1]] = *loss;
        ctr[0] += 1;
        ctr[1] += 1;
    }
}

__global__ void set_int_kernel(int* p, int value) {
    if (threadIdx.x == 0) *p = value;
}

// Rows of X (n_in floats each) selected by idx[off + i] into the padded bf16
// batch (row-major and transposed) and the matching targets and masks.
__global__ __launch_bounds__(256) void gather_kernel(const float* __restrict__ X, int n_in, const float* __restrict__ Y,
                                                     const float* __restrict__ Mk, int ny, const int* __restrict__ idx,
                                                     const int* __restrict__ bstep, int B, int m_valid,
                                                     __bf16* __restrict__ xb, int ldx, __bf16* __restrict__ xt, int ldt,
                                                     float* __restrict__ yb, float* __restrict__ mb,
                                                     float* __restrict__ loss) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *loss = 0.0f;
    if (i >= B) return;
    const int64_t off = (int64_t)(*bstep) * m_valid;
    const bool ok = i < m_valid;
    const int src = ok ? idx[off + i] : 0;
    for (int j = 0; j < ldx; ++j) {
        const float v = (ok && j < n_in) ? X[(int64_t)src * n_in + j] : 0.0f;
        xb[(int64_t)i * ldx + j] = (__bf16)v;
        xt[(int64_t)j * ldt + i] = (__bf16)v;
    }
    for (int j = 0; j < ny; ++j) {
        yb[(int64_t)i * ny + j] = ok ? Y[(int64_t)src * ny + j] : 0.0f;
        mb[(int64_t)i * ny + j] = ok ? (Mk ? Mk[(int64_t)src * ny + j] : 1.0f) : 0.0f;
    }
}

// Inference input: n rows of n_in floats into the padded bf16 batch.
__global__ __launch_bounds__(256) void pack_kernel(const float* __restrict__ X, int n_in, int n, __bf16* __restrict__ xb,
                                                   int ldx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int j = 0; j < ldx; ++j) xb[(int64_t)i * ldx + j] = (__bf16)(j < n_in ? X[(int64_t)i * n_in + j] : 0.0f);
}

// Network outputs: the first n_out of the ldo padded columns of n rows.
__global__ __launch_bounds__(256) void unpack_kernel(const float* __restrict__ out, int ldo, int n, int n_out,
                                                     float* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n * n_out) return;
    const int64_t r = i / n_out, c = i % n_out;
    y[i] = out[r * ldo + c];
}

void launch_gemm_bt(const __bf16* A, int lda, const __bf16* B, int ldb, int M, int N, int K, const GemmEpi& e,
                    hipStream_t st) {
    dim3 grid(M / 32, N / 32);
    gemm_bt_kernel<<<grid, 64, 0, st>>>(A, lda, B, ldb, K, e);
}

void launch_mlp_loss(const float* out, int ldo, const float* y, const float* mask, int ny, int n_reg, int n_bce, int M,
                     int m_valid, __bf16* dout, int ldd, __bf16* doutt, int lddt, float* loss, hipStream_t st) {
    loss_kernel<<<(M + 255) / 256, 256, 0, st>>>(out, ldo, y, mask, ny, n_reg, n_bce, M, m_valid, dout, ldd, doutt, lddt,
                                                 loss);
}

void launch_rowsum(const __bf16* gt, int ld, int M, int N, float* db, hipStream_t st) {
    rowsum_kernel<<<N, 256, 0, st>>>(gt, ld, M, db);
}

void launch_adam(float* p, const float* g, float* m, float* v, int n, const int* step, float lr, float b1, float b2,
                 float eps, int cols, __bf16* wb, int ldw, __bf16* wt, int ldt, hipStream_t st) {
    adam_kernel<<<(n + 255) / 256, 256, 0, st>>>(p, g, m, v, n, step, lr, b1, b2, eps, cols, wb, ldw, wt, ldt);
}

void launch_mlp_step(int* ctr, float* loss_log, const float* loss, hipStream_t st) {
    step_kernel<<<1, 64, 0, st>>>(ctr, loss_log, loss);
}

void launch_set_int(int* p, int value, hipStream_t st) { set_int_kernel<<<1, 64, 0, st>>>(p, value); }

void launch_gather(const float* X, int n_in, const float* Y, const float* Mk, int ny, const int* idx, const int* bstep,
                   int B, int m_valid, __bf16* xb, int ldx, __bf16* xt, int ldt, float* yb, float* mb, float* loss,
                   hipStream_t st) {
    gather_kernel<<<(B + 255) / 256, 256, 0, st>>>(X, n_in, Y, Mk, ny, idx, bstep, B, m_valid, xb, ldx, xt, ldt, yb, mb,
                                                   loss);
}

void launch_pack(const float* X, int n_in, int n, __bf16* xb, int ldx, hipStream_t st) {
    pack_kernel<<<(n + 255) / 256, 256, 0, st>>>(X, n_in, n, xb, ldx);
}

void launch_unpack(const float* out, int ldo, int n, int n_out, float* y, hipStream_t st) {
    const int64_t cnt = (int64_t)n * n_out;
    unpack_kernel<<<(int)((cnt + 255) / 256), 256, 0, st>>>(out, ldo, n, n_out, y);
}

}  // namespace lqt
