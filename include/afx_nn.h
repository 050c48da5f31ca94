/* afx_nn.h — C ABI of the learned warm start (3-D convolutional encoder-decoder surrogate).
 *
 * No entry point here has a counterpart in the reference: the reference has no learned or
 * surrogate code (a search of the tree for neural / surrogate / learned code finds none). The
 * surrogate maps log10 conductivity and the tidal source current per cell to a cell-wise vector
 * field, which is averaged onto edges as the initial guess X of the COCG solve (the X argument
 * of COCG13_single, src/solver/COCG13_single.f90:9-36, is the initial guess on entry).
 *
 * Conventions: all pointers are device pointers; bf16 values are passed as uint16_t bit
 * patterns (round-to-nearest-even); activations are channels-last [batch][z][y][x][c];
 * `stream` is a hipStream_t (NULL = the default stream). Every call returns 0 or an error code
 * and leaves a message for afx_nn_last_error().
 */
#ifndef AFX_NN_H
#define AFX_NN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AFX_NN_EINVAL 101  /* invalid shape, leading dimension or alignment */
#define AFX_NN_EHIP (-300) /* a HIP runtime call failed */

const char *afx_nn_last_error(void);

/* C[m][n] = sum_k A[m][k] * B[n][k] (+ bias[n % bias_period], bias_period <= 0 meaning N)
 * (ReLU if relu != 0), bf16 operands, FP32
 * accumulation on v_mfma_f32_32x32x16_bf16. K % 16 == 0 (pad operands with zeros), lda/ldb
 * multiples of 8, operands 16-byte aligned. Outputs: C (f32, ldc) and/or Cb (bf16, ldcb).
 * splits > 1 splits K; the partial sums go to `partial` (splits*M*N floats) and are reduced
 * into C in a fixed order (bias/ReLU/Cb are not applied in that mode). */
int afx_nn_gemm_nt(const uint16_t *A, int lda, const uint16_t *B, int ldb, int M, int N, int K, float *C, int ldc,
                   uint16_t *Cb, int ldcb, const float *bias, int bias_period, int relu, int splits, float *partial,
                   void *stream);

/* 3x3x3 patches (padding 1): cols[v][tap*C + c] = X[v + offset(tap)][c], tap = (dz*3+dy)*3+dx,
 * offset = (dz-1, dy-1, dx-1), or its negation when flip != 0; zero outside the grid and for
 * k >= 27*C up to the padded row length Kp. */
int afx_nn_im2col3(const uint16_t *X, uint16_t *cols, int nb, int D, int H, int W, int C, int Kp, int flip,
                   void *stream);
/* transposed patches colsT[tap*C + c][v], rows of length Vp >= nb*D*H*W (padding zero). */
int afx_nn_im2col3_t(const uint16_t *X, uint16_t *colsT, int nb, int D, int H, int W, int C, int Kp, int Vp,
                     void *stream);
/* W [Co][27*Ci] (row stride ldw) -> Wt [Ci][Kp] with Wt[ci][tap*Co + co] = W[co][tap*Ci + ci]. */
int afx_nn_weight_t3(const uint16_t *W, int ldw, uint16_t *Wt, int Co, int Ci, int Kp, void *stream);
/* space-to-depth by 2: X [nb][D][H][W][C] <-> Y [nb][D/2][H/2][W/2][8C], channel s*C + c with
 * s = (z%2*2 + y%2)*2 + x%2, elements of esize bytes (2: bf16, 4: f32; C*esize % 16 == 0).
 * inverse == 0 writes Y from X, else X from Y. */
int afx_nn_space_to_depth(void *X, void *Y, int nb, int D, int H, int W, int C, int esize, int inverse,
                          void *stream);
/* Y[c][r] = bf16(X[r][c]) for r < R, c < Cc, X f32 (esize 4) or bf16 (esize 2) with row stride
 * ldx; zero for R <= r < Rp. */
int afx_nn_transpose(const void *X, int esize, int ldx, int R, int Cc, uint16_t *Y, int Rp, void *stream);
int afx_nn_cast_f2b(const float *x, uint16_t *y, long long n, void *stream);
/* dP = dY * (Y > 0) with Y the bf16 ReLU output; dPb (optional) its bf16 copy. */
int afx_nn_relu_bwd(const float *dY, const uint16_t *Y, float *dP, uint16_t *dPb, long long n, void *stream);
/* out[c] (+)= sum_r X[r][c], X [R][Cc] f32. */
/* out[c] (+)= sum_r X[r][c], X [R][Cc] f32; part: workspace of ceil(R / 256) * Cc floats */
int afx_nn_colsum(const float *X, int R, int Cc, float *out, int accumulate, float *part, void *stream);
int afx_nn_add_bf16(const uint16_t *a, const uint16_t *b, uint16_t *y, long long n, void *stream);
int afx_nn_add_f32(const float *a, const float *b, float *y, long long n, void *stream);
/* loss = sum over rows with mask != 0 (mask may be NULL) of |out[r][0:Cc] - tgt[r]|^2 * inv_count,
 * out [rows][ld], tgt [rows][Cc]; grad [rows][ld] = d loss / d out (zero beyond Cc).
 * partial holds ceil(rows*ld/256) floats. */
int afx_nn_mse(const float *out, int ld, const float *tgt, const uint8_t *mask, long long rows, int Cc,
               float inv_count, float *grad, float *partial, float *loss, void *stream);
/* Adam step `step` (>= 1) on FP32 master weights w with moments m, v; wb (optional) receives
 * the bf16 copy. */
int afx_nn_adam(float *w, const float *g, float *m, float *v, uint16_t *wb, long long n, float lr, float b1, float b2,
                float eps, int step, void *stream);
/* x[e] (complex, interleaved) = scale * mean over the cells c incident to edge e (CSR eptr/ecell)
 * of F[c] . t_e, F[c] = 6 floats (Re, Im of a 3-vector), t_e = evec[3e..3e+2]; x[e] = 0 where
 * zero[e] != 0 (Dirichlet edges; zero may be NULL). */
int afx_nn_cells_to_edges(const int *eptr, const int *ecell, const double *evec, const float *F, int ldf,
                          double scale, const uint8_t *zero, int ne, double *x, void *stream);
/* F[c] = least-squares constant field over the cell's edges (CSR cptr/cedge): minimises
 * sum_e |F . t_e - x[e]|^2, divided by scale. */
int afx_nn_edges_to_cells(const int *cptr, const int *cedge, const double *evec, const double *x, double scale,
                          int nc, float *F, int ldf, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* AFX_NN_H */
