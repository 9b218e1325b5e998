// band.hip -- learned band proposal for long pairs: k-mer diagonal features,
// a bf16 MLP with fp32 accumulation (forward, backward, Adam), the band it
// proposes, and the path-diagonal labels it is trained on.
//
// No counterpart in the reference (its banded kernel takes one fixed tile band
// for the whole batch, src/kernels/banded.h:10-139).  Everything here is on
// the device; the entry points take device pointers and a stream.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include "gasal2_band.h"

namespace {

constexpr int LB_K = 11;            // k-mer length
constexpr int LB_SLOTS = 8192;      // LDS hash slots (query k-mers, <= 4096 inserted)
constexpr int LB_COARSE = 256;      // coarse diagonal bins over [-Lq, Lt]
constexpr int LB_FINE = 64;         // fine bins around the anchor (the network's input)
constexpr uint32_t LB_EMPTY = 0xFFFFFFFFu;

__device__ __forceinline__ int32_t regs8(uint32_t len) { return (int32_t)((len >> 3) + ((len & 7) ? 1 : 0)); }

// 2-bit code of a nibble (A=1 -> 0, C=3 -> 1, T=4 -> 2, G=7 -> 3); other bases
// (N, IUPAC) break k-mers.
__device__ __forceinline__ bool base2(uint32_t nib, uint32_t &c) {
    c = (nib >> 1) & 3u;
    return nib == 1u || nib == 3u || nib == 4u || nib == 7u;
}

__device__ __forceinline__ uint32_t nib_at(const uint32_t *W, int32_t i) { return (W[i >> 3] >> (28 - 4 * (i & 7))) & 15u; }

__device__ __forceinline__ bool kmer_at(const uint32_t *W, int32_t i, uint32_t &code) {
    code = 0;
#pragma unroll
    for (int t = 0; t < LB_K; t++) {
        uint32_t c;
        if (!base2(nib_at(W, i + t), c)) return false;
        code = (code << 2) | c;
    }
    return true;
}

__device__ __forceinline__ uint32_t slot_of(uint32_t code) { return (code * 2654435761u) >> 19; }   // 13 bits

// One block per pair.  Query k-mers (every stride-th position, at most 4096;
// a repeated k-mer keeps its smallest position) go into an LDS hash table;
// every target k-mer is looked up and each hit (query i, target j) adds its
// diagonal d = i - j (query column - target row, the DP's convention) to a
// coarse histogram over [-Lt, Lq].  The
// anchor is the centre of the fullest coarse bin (lowest on ties); a second
// scan bins hits into LB_FINE bins of width binw around it.
// Features: 64 fine-bin fractions of all hits, log2(Lq+1)/14, log2(Lt+1)/14,
// (Lt-Lq)/(Lt+Lq), hits/min(Lq,Lt).
__global__ __launch_bounds__(256) void k_features(const uint32_t *__restrict__ pq, const uint32_t *__restrict__ pt,
                                                  const uint32_t *__restrict__ qoff, const uint32_t *__restrict__ toff,
                                                  const uint32_t *__restrict__ qlen, const uint32_t *__restrict__ tlen,
                                                  uint32_t n, float *__restrict__ feat, int32_t *__restrict__ anchor,
                                                  int32_t *__restrict__ binw_out) {
    __shared__ uint32_t keys[LB_SLOTS];
    __shared__ uint32_t pos[LB_SLOTS];
    __shared__ uint32_t coarse[LB_COARSE];
    __shared__ uint32_t fine[LB_FINE];
    __shared__ uint32_t total;
    __shared__ int32_t s_anchor;
    const uint32_t t = threadIdx.x;
    for (uint32_t pr = blockIdx.x; pr < n; pr += gridDim.x) {
        const int32_t Lq = (int32_t)qlen[pr], Lt = (int32_t)tlen[pr];
        const uint32_t *Q = pq + (qoff[pr] >> 3), *T = pt + (toff[pr] >> 3);
        for (uint32_t s = t; s < LB_SLOTS; s += 256) { keys[s] = LB_EMPTY; pos[s] = 0xFFFFFFFFu; }
        if (t < LB_COARSE) coarse[t] = 0;
        if (t < LB_FINE) fine[t] = 0;
        if (t == 0) total = 0;
        __syncthreads();
        const int32_t nq = Lq - LB_K + 1, nt = Lt - LB_K + 1;
        const int32_t stride = nq > 4096 ? (nq + 4095) / 4096 : 1;
        for (int32_t i = (int32_t)t * stride; i < nq; i += 256 * stride) {
            uint32_t code;
            if (!kmer_at(Q, i, code)) continue;
            uint32_t s = slot_of(code);
            for (int probe = 0; probe < LB_SLOTS; probe++) {
                const uint32_t old = atomicCAS(&keys[s], LB_EMPTY, code);
                if (old == LB_EMPTY || old == code) {   // repeated k-mer: the smallest position wins
                    atomicMin(&pos[s], (uint32_t)i);
                    break;
                }
                s = (s + 1) & (LB_SLOTS - 1);
            }
        }
        __syncthreads();
        const int32_t span = Lq + Lt + 1;
        for (int pass = 0; pass < 2; pass++) {
            const int32_t a = s_anchor;
            const int32_t bw = span > 4096 ? (span + 1023) / 1024 : 4;
            for (int32_t j = (int32_t)t; j < nt; j += 256) {
                uint32_t code;
                if (!kmer_at(T, j, code)) continue;
                uint32_t s = slot_of(code);
                int32_t i = -1;
                for (int probe = 0; probe < LB_SLOTS; probe++) {
                    const uint32_t k = keys[s];
                    if (k == LB_EMPTY) break;
                    if (k == code) { i = (int32_t)pos[s]; break; }
                    s = (s + 1) & (LB_SLOTS - 1);
                }
                if (i < 0) continue;
                const int32_t dgl = i - j;
                if (pass == 0) {
                    const int32_t cb = (int32_t)(((int64_t)(dgl + Lt) * LB_COARSE) / span);
                    atomicAdd(&coarse[cb], 1u);
                    atomicAdd(&total, 1u);
                } else {
                    const int32_t off = dgl - a;
                    const int32_t fb = (off >= 0 ? off / bw : -((-off + bw - 1) / bw)) + LB_FINE / 2;
                    if (fb >= 0 && fb < LB_FINE) atomicAdd(&fine[fb], 1u);
                }
            }
            __syncthreads();
            if (pass == 0 && t == 0) {
                uint32_t best = 0, bi = 0;
                for (uint32_t b = 0; b < LB_COARSE; b++)
                    if (coarse[b] > best) { best = coarse[b]; bi = b; }
                // centre of the coarse bin, in diagonal units
                s_anchor = (int32_t)(((2 * (int64_t)bi + 1) * span) / (2 * LB_COARSE)) - Lt;
            }
            __syncthreads();
        }
        const float inv = 1.0f / (float)(total > 0 ? total : 1u);
        float *f = feat + (size_t)pr * LB_NFEAT;
        if (t < LB_FINE) f[t] = (float)fine[t] * inv;
        if (t == 0) {
            f[64] = log2f((float)Lq + 1.0f) / 14.0f;
            f[65] = log2f((float)Lt + 1.0f) / 14.0f;
            f[66] = (float)(Lt - Lq) / (float)(Lt + Lq > 0 ? Lt + Lq : 1);
            f[67] = (float)total / (float)((Lq < Lt ? Lq : Lt) + 1);
            anchor[pr] = s_anchor;
            binw_out[pr] = span > 4096 ? (span + 1023) / 1024 : 4;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------- MLP ----
// Dense layer Y = act(X W^T + b): X fp32 [B][K] (rounded to bf16 for the
// product), W bf16 [N][K] (K even), b fp32 [N], fp32 accumulation with
// v_dot2c_f32_bf16.  One thread per output element.
typedef __bf16 bf2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint16_t f2bf(float x) {   // round to nearest even
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf2_t pack_bf2(uint32_t lo_hi) {
    bf2_t r;
    __builtin_memcpy(&r, &lo_hi, 4);
    return r;
}

__global__ __launch_bounds__(256) void k_dense_fwd(const float *__restrict__ X, uint32_t B, uint32_t K,
                                                   const uint16_t *__restrict__ W, const float *__restrict__ bias, uint32_t N,
                                                   float *__restrict__ Y, int relu) {
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (uint64_t)B * N) return;
    const uint32_t b = (uint32_t)(id / N), o = (uint32_t)(id % N);
    const float *x = X + (size_t)b * K;
    const uint32_t *w2 = reinterpret_cast<const uint32_t *>(W + (size_t)o * K);
    float acc = bias[o];
    for (uint32_t k = 0; k < K; k += 2) {
        const uint32_t xp = (uint32_t)f2bf(x[k]) | ((uint32_t)f2bf(x[k + 1]) << 16);
        acc = __builtin_amdgcn_fdot2_f32_bf16(pack_bf2(xp), pack_bf2(w2[k >> 1]), acc, false);
    }
    Y[id] = relu ? fmaxf(acc, 0.0f) : acc;
}

// dZ = dY * act'(Y) in place (ReLU: Y > 0).
__global__ __launch_bounds__(256) void k_relu_bwd(float *__restrict__ dY, const float *__restrict__ Y, uint64_t cnt) {
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < cnt && !(Y[id] > 0.0f)) dY[id] = 0.0f;
}

// dX[b][k] = sum_n dZ[b][n] W[n][k]  (bf16 weights as in the forward)
__global__ __launch_bounds__(256) void k_dense_bwd_x(const float *__restrict__ dZ, uint32_t B, uint32_t N,
                                                     const uint16_t *__restrict__ W, uint32_t K, float *__restrict__ dX) {
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (uint64_t)B * K) return;
    const uint32_t b = (uint32_t)(id / K), k = (uint32_t)(id % K);
    const float *dz = dZ + (size_t)b * N;
    float acc = 0.0f;
    for (uint32_t o = 0; o < N; o++) acc = fmaf(dz[o], bf2f(W[(size_t)o * K + k]), acc);
    dX[id] = acc;
}

// dW[n][k] = sum_b dZ[b][n] X[b][k];  db[n] = sum_b dZ[b][n]  (fp32 X)
__global__ __launch_bounds__(256) void k_dense_bwd_w(const float *__restrict__ dZ, const float *__restrict__ X, uint32_t B,
                                                     uint32_t N, uint32_t K, float *__restrict__ dW, float *__restrict__ db) {
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (uint64_t)N * (K + 1)) return;
    const uint32_t o = (uint32_t)(id / (K + 1)), k = (uint32_t)(id % (K + 1));
    float acc = 0.0f;
    if (k < K) {
        for (uint32_t b = 0; b < B; b++) acc = fmaf(dZ[(size_t)b * N + o], X[(size_t)b * K + k], acc);
        dW[(size_t)o * K + k] = acc;
    } else {
        for (uint32_t b = 0; b < B; b++) acc += dZ[(size_t)b * N + o];
        db[o] = acc;
    }
}

// Weighted mean squared error over the two outputs:
// L = sum_b w_b sum_o (y - t)^2 / sum_b w_b;  dY = 2 w_b (y - t) / sum w.
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ Yp, const float *__restrict__ Tg,
                                             const float *__restrict__ wt, uint32_t B, float inv_wsum,
                                             float *__restrict__ dY, float *__restrict__ loss) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (b < B) {
        const float w = wt[b];
        for (int o = 0; o < 2; o++) {
            const float e = Yp[2 * b + o] - Tg[2 * b + o];
            dY[2 * b + o] = 2.0f * w * e * inv_wsum;
            l += w * e * e * inv_wsum;
        }
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
    if ((threadIdx.x & 63) == 0 && l != 0.0f) atomicAdd(loss, l);
}

__global__ __launch_bounds__(256) void k_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                                              const float *__restrict__ g, uint32_t cnt, float lr, float b1, float b2,
                                              float eps, float c1, float c2, uint16_t *__restrict__ wbf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    w[i] = wi;
    if (wbf) wbf[i] = f2bf(wi);
}

__global__ __launch_bounds__(256) void k_to_bf16(const float *__restrict__ w, uint16_t *__restrict__ o, uint32_t cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) o[i] = f2bf(w[i]);
}

// Share of a pair's k-mer hits in fine bins that meet [lo, hi]; bin b holds
// diagonals [anchor + (b - 32) * binw, anchor + (b - 31) * binw).
__device__ __forceinline__ float hit_share(const float *f, int32_t anchor, int32_t bw, int32_t lo, int32_t hi) {
    float in = 0.0f;
    for (int b = 0; b < LB_FINE; b++) {
        const int32_t b0 = anchor + (b - LB_FINE / 2) * bw, b1 = b0 + bw - 1;
        if (b1 >= lo && b0 <= hi) in += f[b];
    }
    return in;
}

// Band from the network output (centre offset and log half-width in fine-bin
// units): centre = anchor + binw * y0, half = binw * exp(y1) * mul + add.
// pre[k] = 16 when less than min_share of the pair's k-mer hits lie in bins
// the band meets (the band misses the evidence: full DP instead).
__global__ __launch_bounds__(256) void k_band_from_net(const float *__restrict__ Y, const float *__restrict__ feat,
                                                       const int32_t *__restrict__ anchor,
                                                       const int32_t *__restrict__ binw, uint32_t n, float mul, float add,
                                                       float min_share, int32_t *__restrict__ dlo,
                                                       int32_t *__restrict__ dhi, uint8_t *__restrict__ pre) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float bw = (float)binw[k];
    const float c = (float)anchor[k] + bw * Y[2 * k];
    const float h = bw * expf(fminf(Y[2 * k + 1], 20.0f)) * mul + add;
    const int32_t lo = (int32_t)floorf(c - h), hi = (int32_t)ceilf(c + h);
    dlo[k] = lo;
    dhi[k] = hi;
    pre[k] = hit_share(feat + (size_t)k * LB_NFEAT, anchor[k], binw[k], lo, hi) < min_share ? 16 : 0;
}

// Non-learned band: centre of the fullest fine bin (lowest on ties), fixed
// half-width half + frac * min(Lq, Lt).
__global__ __launch_bounds__(256) void k_band_peak(const float *__restrict__ feat, const int32_t *__restrict__ anchor,
                                                   const int32_t *__restrict__ binw, const uint32_t *__restrict__ qlen,
                                                   const uint32_t *__restrict__ tlen, uint32_t n, float half, float frac,
                                                   float min_share, int32_t *__restrict__ dlo, int32_t *__restrict__ dhi,
                                                   uint8_t *__restrict__ pre) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float *f = feat + (size_t)k * LB_NFEAT;
    int bi = LB_FINE / 2;
    float best = -1.0f;
    for (int b = 0; b < LB_FINE; b++)
        if (f[b] > best) { best = f[b]; bi = b; }
    const float c = (float)anchor[k] + (float)binw[k] * ((float)(bi - LB_FINE / 2) + 0.5f);
    const float m = (float)(qlen[k] < tlen[k] ? qlen[k] : tlen[k]);
    const float h = half + frac * m;
    const int32_t lo = (int32_t)floorf(c - h), hi = (int32_t)ceilf(c + h);
    dlo[k] = lo;
    dhi[k] = hi;
    pre[k] = hit_share(f, anchor[k], binw[k], lo, hi) < min_share ? 16 : 0;
}

// Diagonal range (d = query column - target row) of each pair's alignment path,
// walked back from its end over the CIGAR (reverse order, one byte per run:
// op | count << 2, 0 M, 1 X, 2 D (target), 3 I (query)).
__global__ __launch_bounds__(256) void k_path_diag(const uint8_t *__restrict__ cigar, uint32_t cap,
                                                   const uint32_t *__restrict__ nops, const int32_t *__restrict__ qend,
                                                   const int32_t *__restrict__ tend, uint32_t n, int32_t *__restrict__ dmin,
                                                   int32_t *__restrict__ dmax) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int32_t d = qend[k] - tend[k];
    int32_t lo = d, hi = d;
    const uint8_t *c = cigar + (size_t)k * cap;
    const uint32_t m = nops[k] < cap ? nops[k] : cap;
    for (uint32_t r = 0; r < m; r++) {
        const uint32_t op = c[r] & 3u, cnt = c[r] >> 2;
        if (op == 2u) d += (int32_t)cnt;          // target base consumed alone: d = j - i grows
        else if (op == 3u) d -= (int32_t)cnt;     // query base consumed alone
        lo = d < lo ? d : lo;
        hi = d > hi ? d : hi;
    }
    dmin[k] = lo;
    dmax[k] = hi;
}

inline uint32_t nblk(uint64_t cnt) { return (uint32_t)((cnt + 255) / 256); }

}  // namespace

extern "C" int32_t gasal2_lb_features(const uint32_t *pq, const uint32_t *pt, const uint32_t *qoff, const uint32_t *toff,
                                      const uint32_t *qlen, const uint32_t *tlen, uint32_t n, float *feat, int32_t *anchor,
                                      int32_t *binw, void *stream) {
    if (n == 0) return 0;
    const uint32_t g = n < 65535u ? n : 65535u;
    k_features<<<g, 256, 0, (hipStream_t)stream>>>(pq, pt, qoff, toff, qlen, tlen, n, feat, anchor, binw);
    return (int32_t)hipGetLastError();
}

extern "C" int32_t gasal2_lb_dense_fwd(const float *X, uint32_t B, uint32_t K, const uint16_t *W, const float *bias, uint32_t N,
                                       float *Y, int32_t relu, void *stream) {
    if (K & 1u) return (int32_t)hipErrorInvalidValue;
    if ((uint64_t)B * N == 0) return 0;
    k_dense_fwd<<<nblk((uint64_t)B * N), 256, 0, (hipStream_t)stream>>>(X, B, K, W, bias, N, Y, relu);
    return (int32_t)hipGetLastError();
}

extern "C" int32_t gasal2_lb_dense_bwd(const float *dY, const float *Y, int32_t relu, const float *X, uint32_t B, uint32_t K,
                                       const uint16_t *W, uint32_t N, float *dZ, float *dX, float *dW, float *db, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    const uint64_t cnt = (uint64_t)B * N;
    if (cnt == 0) return 0;
    hipError_t e = hipMemcpyAsync(dZ, dY, cnt * 4, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return (int32_t)e;
    if (relu) k_relu_bwd<<<nblk(cnt), 256, 0, s>>>(dZ, Y, cnt);
    if (dX) k_dense_bwd_x<<<nblk((uint64_t)B * K), 256, 0, s>>>(dZ, B, N, W, K, dX);
    k_dense_bwd_w<<<nblk((uint64_t)N * (K + 1)), 256, 0, s>>>(dZ, X, B, N, K, dW, db);
    return (int32_t)hipGetLastError();
}

extern "C" int32_t gasal2_lb_mse(const float *Yp, const float *Tg, const float *wt, uint32_t B, float wsum, float *dY,
                                 float *loss, void *stream) {
    if (B == 0) return 0;
    k_mse<<<nblk(B), 256, 0, (hipStream_t)stream>>>(Yp, Tg, wt, B, wsum > 0.0f ? 1.0f / wsum : 0.0f, dY, loss);
    return (int32_t)hipGetLastError();
}

extern "C" int32_t gasal2_lb_adam(float *w, float *m, float *v, const float *g, uint32_t cnt, float lr, float b1, float b2,
                                  float eps, int32_t step, uint16_t *wbf, void *stream) {
    if (cnt == 0) return 0;
    const float c1 = 1.0f / (1.0f - powf(b1, (float)step)), c2 = 1.0f / (1.0f - powf(b2, (float)step));
    k_adam<<<nblk(cnt), 256, 0, (hipStream_t)stream>>>(w, m, v, g, cnt, lr, b1, b2, eps, c1, c2, wbf);
    return (int32_t)hipGetLastError();
}

extern "C" int32_t gasal2_lb_to_bf16(const float *w, uint16_t *out, uint32_t cnt, void *stream) {
    if (cnt == 0) return 0;
    k_to_bf16<<<nblk(cnt), 256, 0, (hipStream_t)stream>>>(w, out, cnt);
    return (int32_t)hipGetLastError();
}

extern "C" int32_t gasal2_lb_band_from_net(const float *Y, const float *feat, const int32_t *anchor, const int32_t *binw,
                                           uint32_t n, float mul, float add, float min_share, int32_t *dlo, int32_t *dhi,
                                           uint8_t *pre, void *stream) {
    if (n == 0) return 0;
    k_band_from_net<<<nblk(n), 256, 0, (hipStream_t)stream>>>(Y, feat, anchor, binw, n, mul, add, min_share, dlo, dhi, pre);
    return (int32_t)hipGetLastError();
}

extern "C" int32_t gasal2_lb_band_peak(const float *feat, const int32_t *anchor, const int32_t *binw, const uint32_t *qlen,
                                       const uint32_t *tlen, uint32_t n, float half, float frac, float min_share,
                                       int32_t *dlo, int32_t *dhi, uint8_t *pre, void *stream) {
    if (n == 0) return 0;
    k_band_peak<<<nblk(n), 256, 0, (hipStream_t)stream>>>(feat, anchor, binw, qlen, tlen, n, half, frac, min_share, dlo,
                                                          dhi, pre);
    return (int32_t)hipGetLastError();
}

extern "C" int32_t gasal2_lb_path_diag(const uint8_t *cigar, uint32_t cap, const uint32_t *nops, const int32_t *qend,
                                       const int32_t *tend, uint32_t n, int32_t *dmin, int32_t *dmax, void *stream) {
    if (n == 0) return 0;
    k_path_diag<<<nblk(n), 256, 0, (hipStream_t)stream>>>(cigar, cap, nops, qend, tend, n, dmin, dmax);
    return (int32_t)hipGetLastError();
}
