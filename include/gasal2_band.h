/* gasal2_band.h -- learned band proposal for long pairs (device kernels).
 *
 * No counterpart in the reference: its only band is one fixed tile band for
 * the whole batch (src/kernels/banded.h:10-139, k_band in Parameters,
 * src/args_parser.h:37).  Here a small network proposes a diagonal band per
 * pair from k-mer evidence; the banded DP runs inside it
 * (gasal2_run_banded, gasal2_ext.h) and the full DP is the fallback.
 *
 * All pointers are device pointers; `stream` is a hipStream_t.  Functions
 * return 0 or a hipError_t code.  Diagonals are d = query column - target row.
 */
#ifndef GASAL2_BAND_H
#define GASAL2_BAND_H

#include <stdint.h>

#define LB_NFEAT 68   /* 64 fine diagonal bins + 4 length features */
#define LB_HIDDEN 256

#ifdef __cplusplus
extern "C" {
#endif

/* Per pair: LB_NFEAT features (fractions of k=11 hits in 64 diagonal bins of
 * width binw[k] around anchor[k], the fullest of 256 coarse bins over
 * [-Lt, Lq]; a hit at query i, target j has diagonal i - j; log2(Lq+1)/14, log2(Lt+1)/14, (Lt-Lq)/(Lt+Lq), hits/min(Lq,Lt)).
 * pq/pt: packed batches (8 bases per word, base 0 in bits 31:28); qoff/toff in
 * bases (multiples of 8). */
int32_t gasal2_lb_features(const uint32_t *pq, const uint32_t *pt, const uint32_t *qoff, const uint32_t *toff,
                           const uint32_t *qlen, const uint32_t *tlen, uint32_t n, float *feat, int32_t *anchor,
                           int32_t *binw, void *stream);

/* Dense layer Y[B][N] = act(X[B][K] W[N][K]^T + b): X fp32 rounded to bf16,
 * W bf16 (K even), fp32 accumulation (v_dot2c_f32_bf16), act = ReLU if relu. */
int32_t gasal2_lb_dense_fwd(const float *X, uint32_t B, uint32_t K, const uint16_t *W, const float *bias, uint32_t N,
                            float *Y, int32_t relu, void *stream);

/* Backward of the layer above: dZ = dY * act'(Y); dX = dZ W (bf16 W; skipped
 * when dX is NULL); dW = dZ^T X (fp32 X); db = column sums of dZ. */
int32_t gasal2_lb_dense_bwd(const float *dY, const float *Y, int32_t relu, const float *X, uint32_t B, uint32_t K,
                            const uint16_t *W, uint32_t N, float *dZ, float *dX, float *dW, float *db, void *stream);

/* Weighted MSE over 2 outputs: loss += sum_b wt_b |y_b - t_b|^2 / wsum,
 * dY = 2 wt_b (y_b - t_b) / wsum.  loss must be zeroed by the caller. */
int32_t gasal2_lb_mse(const float *Yp, const float *Tg, const float *wt, uint32_t B, float wsum, float *dY, float *loss,
                      void *stream);

/* Adam step (bias-corrected, step >= 1) on fp32 master weights; when wbf is
 * not NULL also refreshes their bf16 copy. */
int32_t gasal2_lb_adam(float *w, float *m, float *v, const float *g, uint32_t cnt, float lr, float b1, float b2, float eps,
                       int32_t step, uint16_t *wbf, void *stream);
int32_t gasal2_lb_to_bf16(const float *w, uint16_t *out, uint32_t cnt, void *stream);

/* Band from the network output Y[n][2] (centre offset, log half-width in fine
 * bins): [floor(c - h), ceil(c + h)], c = anchor + binw*y0,
 * h = binw*exp(y1)*mul + add.  pre[k] = 16 when the fine bins the band meets
 * hold less than min_share of the pair's k-mer hits (send it to the full DP). */
int32_t gasal2_lb_band_from_net(const float *Y, const float *feat, const int32_t *anchor, const int32_t *binw,
                                uint32_t n, float mul, float add, float min_share, int32_t *dlo, int32_t *dhi,
                                uint8_t *pre, void *stream);

/* Non-learned band: centre of the fullest fine bin, half-width
 * half + frac*min(Lq, Lt); pre as above. */
int32_t gasal2_lb_band_peak(const float *feat, const int32_t *anchor, const int32_t *binw, const uint32_t *qlen,
                            const uint32_t *tlen, uint32_t n, float half, float frac, float min_share, int32_t *dlo,
                            int32_t *dhi, uint8_t *pre, void *stream);

/* Diagonal range of each pair's traceback path from its CIGAR (per-pair
 * slots of cap bytes, reverse order) and end cell (labels for training). */
int32_t gasal2_lb_path_diag(const uint8_t *cigar, uint32_t cap, const uint32_t *nops, const int32_t *qend,
                            const int32_t *tend, uint32_t n, int32_t *dmin, int32_t *dmax, void *stream);

#ifdef __cplusplus
}
#endif

#endif
