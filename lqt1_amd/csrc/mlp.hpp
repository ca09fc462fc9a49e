// Declarations of the surrogate MLP kernels (kernels_mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace lqt {

struct GemmEpi {
    const float* bias;  // [N] added to every row, or null
    int act;            // 1: GELU on the output
    const float* pre;   // backward: multiply by GELU'(pre[row][col]), or null
    int ldpre;
    float* c32;         // fp32 output (before the activation), or null
    int ldc32;
    __bf16* cb;         // bf16 output (after the activation), or null
    int ldcb;
    __bf16* cbt;        // bf16 transposed output, or null
    int ldcbt;
    int n_valid;        // columns >= n_valid are stored as 0 in the bf16 outputs
    int m_valid;        // rows >= m_valid likewise
};

void launch_gemm_bt(const __bf16* A, int lda, const __bf16* B, int ldb, int M, int N, int K, const GemmEpi& e,
                    hipStream_t st);
void launch_mlp_loss(const float* out, int ldo, const float* y, const float* mask, int ny, int n_reg, int n_bce, int M,
                     int m_valid, __bf16* dout, int ldd, __bf16* doutt, int lddt, float* loss, hipStream_t st);
void launch_rowsum(const __bf16* gt, int ld, int M, int N, float* db, hipStream_t st);
void launch_adam(float* p, const float* g, float* m, float* v, int n, const int* step, float lr, float b1, float b2,
                 float eps, int cols, __bf16* wb, int ldw, __bf16* wt, int ldt, hipStream_t st);
void launch_mlp_step(int* ctr, float* loss_log, const float* loss, hipStream_t st);
void launch_gather(const float* X, int n_in, const float* Y, const float* Mk, int ny, const int* idx, const int* bstep,
                   int B, int m_valid, __bf16* xb, int ldx, __bf16* xt, int ldt, float* yb, float* mb, float* loss,
                   hipStream_t st);
void launch_pack(const float* X, int n_in, int n, __bf16* xb, int ldx, hipStream_t st);
void launch_unpack(const float* out, int ldo, int n, int n_out, float* y, hipStream_t st);
void launch_set_int(int* p, int value, hipStream_t st);

}  // namespace lqt
