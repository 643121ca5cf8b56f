/*
 * pcavalue.h -- C ABI of the value network kernels (BF16 MFMA, FP32
 * accumulate) that order the guided search of the PCA-16 superoptimizer.
 *
 * The network scores a search state (a partial program's per-test register
 * file) by how likely it is to reach a solution within the depth budget;
 * it only orders the guided search, the proof of optimality stays with the
 * exhaustive path (include/pcagpu.h).  The reference's learned part is a
 * Python/PyTorch model (python/model.py:16-84, python/train.py); there is no
 * C interface to mirror, so these entry points are this library's own:
 * the building blocks of its forward and backward passes and its optimizer
 * step, so that a caller (pca_amd/value.py) trains and runs the network on
 * the GPU with these kernels only.
 *
 * Conventions: every pointer is a device pointer; bf16 tensors are passed as
 * their 16-bit patterns (uint16_t); matrices are row-major with an explicit
 * leading dimension; `stream` is a hipStream_t (NULL: the default stream).
 * Functions return 0 or a negative PCAGPU_E* code (include/pcagpu.h) and
 * only enqueue work (no synchronisation), except the memory helpers, which
 * give callers without a GPU framework of their own device buffers from
 * this library's HIP runtime.
 */
#ifndef PCAVALUE_H
#define PCAVALUE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* C (M x N) = op(A) (M x K) * op(B) (K x N), BF16 inputs on MFMA with FP32
 * accumulation.  a_trans: A is stored K x M (element (m, k) at A[k*lda+m]),
 * else M x K (A[m*lda+k]); b_trans: B is stored N x K (B[n*ldb+k]), else
 * K x N (B[k*ldb+n]).  Epilogue, in order: + bias[n] (bias may be NULL),
 * ReLU if relu; then C32[m*ldc+n] = v (or += v if accumulate) when C32 is
 * not NULL, and C16[m*ldc16+n] = bf16(v) when C16 is not NULL.  Long-K
 * products (split K) and pcav_colsum use one workspace per device: such
 * calls on different streams of one device must not overlap. */
int pcav_gemm(int M, int N, int K, const uint16_t *A, int lda, int a_trans, const uint16_t *B, int ldb,
              int b_trans, const float *bias, int relu, float *C32, int ldc, int accumulate, uint16_t *C16,
              int ldc16, void *stream);

/* Per-test input features of search states: rec holds, per state b and test
 * t < 32, the 16 words of pcagpu_state_records (input0, input1, target,
 * r0..r7, flags, ...); X[(b*32+t)*180 + f] = bf16 of feature bit f: the
 * 16 bits of input0, of input1, of the target, of r0..r7 (least significant
 * first), then flag bits Z, N, C, V -- the per-test layout of the reference's
 * features (python/search_eval.py:206-251, python/dataset.py:60-63). */
int pcav_features(int B, const uint16_t *rec, uint16_t *X, void *stream);

/* ReLU backward through the stored post-ReLU activation: grad[i] = 0 where
 * act16[i] is not > 0 (in place), and out16[i] = bf16(grad[i]), i < n. */
int pcav_relu_backward(int64_t n, const uint16_t *act16, float *grad, uint16_t *out16, void *stream);

/* Mean over the first nvalid[b] of each state's T per-test rows, then the
 * state's depth embedding: Z[b*(C+E) + c] = bf16(mean_t H[(b*T+t)*C + c])
 * for c < C, Z[b*(C+E) + C + e] = bf16(emb[depth[b]*E + e]) for e < E. */
int pcav_pool_forward(int B, int T, int C, int E, const uint16_t *H, const uint8_t *nvalid, const int32_t *depth,
                      const float *emb, uint16_t *Z, void *stream);

/* Backward of pcav_pool_forward: dH[(b*T+t)*C + c] = dZ[b*(C+E)+c] /
 * nvalid[b] for t < nvalid[b] (0 for padded rows); demb[depth[b]*E + e] +=
 * dZ[b*(C+E) + C + e]. */
int pcav_pool_backward(int B, int T, int C, int E, const float *dZ, const uint8_t *nvalid, const int32_t *depth,
                       float *dH, float *demb, void *stream);

/* out[n] = sum over m < M of G[m*ldg + n] (bias gradients), n < N, ldg >= N;
 * out is overwritten.  Partial sums go through the device workspace
 * pcav_gemm's split K uses (calls on different streams must not overlap). */
int pcav_colsum(int M, int N, const float *G, int ldg, float *out, void *stream);

/* Binary cross-entropy with logits over B states: loss[0] = mean of
 * softplus(y) - label*y; dlogit[b] = (sigmoid(y_b) - label_b) / B. */
int pcav_bce(int B, const float *logit, const float *label, float *dlogit, float *loss, void *stream);
/* The reach + remaining-steps loss of python/train.py:44-60 on head outputs
 * y [B x nh]: column 0 the reach logit (BCE against label, batch mean),
 * columns 1..nh-1 logits over 0..nh-2 remaining steps (cross-entropy against
 * remaining, over the n_pos rows with label > 0.5, their mean, weight w);
 * dy [B x nh] gets the gradient, loss[0] the total. */
int pcav_reach_dist_loss(int B, int nh, const float *y, const float *label, const int32_t *remaining, int n_pos,
                         float w, float *dy, float *loss, void *stream);

/* One Adam step on n FP32 master weights p (moments m, v; step >= 1, bias
 * corrected), then p16 = bf16(p) when p16 is not NULL. */
int pcav_adam(int64_t n, float *p, const float *g, float *m, float *v, uint16_t *p16, float lr, float beta1,
              float beta2, float eps, int step, void *stream);

/* y16[i] = bf16(x[i]) (round to nearest even), i < n. */
int pcav_to_bf16(int64_t n, const float *x, uint16_t *y16, void *stream);

/* Device memory on device `device` (made current): allocation (NULL on
 * failure), release, synchronous copies and clears, and a device-wide
 * synchronisation. */
void *pcav_malloc(int device, size_t bytes);
void pcav_free(void *p);
int pcav_copy_to_device(void *dst, const void *src, size_t bytes);
int pcav_copy_to_host(void *dst, const void *src, size_t bytes);
int pcav_clear(void *dst, size_t bytes);
int pcav_synchronize(void);

#ifdef __cplusplus
}
#endif

#endif /* PCAVALUE_H */
