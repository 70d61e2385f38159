/*
 * gsp_testing.h — test hooks of libgsp_hip.so (not part of the drop-in surface;
 * no counterpart in the reference).
 */
#ifndef GSP_TESTING_H
#define GSP_TESTING_H

#include "gsp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Runs one 16x16x4 MFMA (FP64 when fp32 == 0, else FP32) twice and copies the
 * 4 accumulator values of each of the 64 lanes to out[2][64][4]:
 *   setup 0: A operand = lane + 1, B operand = 1 (outputs are row sums of A)
 *   setup 1: A operand = 1, B operand = lane + 1 (outputs are column sums of B)
 * so a test can confirm the operand and accumulator lane layouts the LUSIM
 * GEMM relies on. fp32 == 2 runs the BF16 16x16x32 MFMA of the surrogate
 * kernels in three setups, out[3][64][4]: A[i][k] = i + 1 with B = 1; A = 1
 * with B[k][j] = j + 1; and a single product per output at lane group 0,
 * item 0 of both operands (the two operands agree on the reduction index). */
int32_t gsp_selftest_mfma(gsp_ctx* ctx, int32_t fp32, double* out);

/* Plain device-to-device copy of `bytes` (a multiple of 16) with a streaming
 * kernel, queued on the context stream: the reference point for calibrating
 * hardware traffic counters against a kernel of known traffic. */
int32_t gsp_selftest_copy(gsp_ctx* ctx, void* dst, const void* src, int64_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* GSP_TESTING_H */
