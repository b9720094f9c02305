// mlp.hip -- refinement-predictor MLP: BF16 MFMA GEMMs with FP32
// accumulation for the forward pass, the backward pass and Adam, all on the
// GPU.  C ABI in include/dacti_mlp.h.
//
// Every layer dimension is padded to a multiple of 16 with zero weights, so
// every GEMM is a grid of 16x16 output tiles; one wavefront computes one tile
// with v_mfma_f32_16x16x16_bf16 over 16-deep slices of the reduction.  The
// operands are addressed through strides, so the forward product
// (activations x W^T), the input gradient (dZ x W) and the weight gradient
// (dZ^T x activations) share one kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/dacti_mlp.h"

namespace dmlp {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// BF16 from FP32: round to nearest, ties to even, on the FP32 bit pattern
// (a NaN stays a quiet NaN).
__host__ __device__ inline uint16_t to_bf16(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t) ((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t) (u >> 16);
}

inline int pad16(int x) { return (x + 15) / 16 * 16; }

// C(m, n) = sum_k A(m, k) B(k, n), A(m, k) = A[m*sam + k*sak], B(k, n) =
// B[k*sbk + n*sbn] (BF16 bits); M, N, K multiples of 16.  Epilogue: + bias[n]
// (if given), ReLU (if relu), zero where mask[m*ldc + n] <= 0 (if mask: the
// derivative of the ReLU whose output is mask), FP32 store to C[m*ldc + n],
// BF16 copy to Cb.
// Lane layout of the 16x16x16 BF16 MFMA: lane l holds A(l%16, 4(l/16)+j) and
// B(4(l/16)+j, l%16) for j < 4, and returns C(4(l/16)+i, l%16) for i < 4.
// One wavefront per output tile (k_gemm_ks below splits long reductions).
// A lane's 4 operand values of one 16-deep slice: one 8-byte load when the
// reduction index is contiguous (stride 1; every padded dimension is a
// multiple of 16, so the address is 8-byte aligned), else 4 loads
template <bool VEC>
__device__ __forceinline__ s4 load_k4(const uint16_t *p, int64_t sk) {
    s4 v;
    if (VEC) {
        v = *(const s4 *) p;
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = (short) p[(int64_t) j * sk];
    }
    return v;
}

template <bool AV, bool BV>
__global__ void __launch_bounds__(256) k_gemm(const uint16_t *__restrict__ A, int64_t sam, int64_t sak,
                                              const uint16_t *__restrict__ B, int64_t sbk, int64_t sbn, int M, int N,
                                              int K, float *__restrict__ C, int64_t ldc, uint16_t *__restrict__ Cb,
                                              const float *__restrict__ bias, int relu,
                                              const float *__restrict__ mask) {
    const int wave = (int) ((blockIdx.x * (int64_t) blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    const int tn = N / 16, tiles = (M / 16) * tn;
    if (wave >= tiles) return;
    const int m0 = (wave / tn) * 16, n0 = (wave % tn) * 16;
    const int r = lane & 15, kq = (lane >> 4) * 4;
    const int kl = K;
    f4 c = {0.0f, 0.0f, 0.0f, 0.0f};
    const uint16_t *pa = A + (int64_t) (m0 + r) * sam + (int64_t) kq * sak;
    const uint16_t *pb = B + (int64_t) kq * sbk + (int64_t) (n0 + r) * sbn;
    // the next slice's operands are loaded while the current one multiplies
    s4 a = load_k4<AV>(pa, sak), b = load_k4<BV>(pb, sbk);
    for (int k0 = 0; k0 < kl; k0 += 16) {
        s4 an = a, bn = b;
        if (k0 + 16 < kl) {
            an = load_k4<AV>(pa + (int64_t) (k0 + 16) * sak, sak);
            bn = load_k4<BV>(pb + (int64_t) (k0 + 16) * sbk, sbk);
        }
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
        a = an;
        b = bn;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int m = m0 + kq + i, n = n0 + r;
        float v = c[i];
        if (bias) v += bias[n];
        if (relu) v = v > 0.0f ? v : 0.0f;
        if (mask) v = mask[(int64_t) m * ldc + n] > 0.0f ? v : 0.0f;
        C[(int64_t) m * ldc + n] = v;
        if (Cb) Cb[(int64_t) m * ldc + n] = to_bf16(v);
    }
}

// The weight gradient's long reduction (K = batch rows) split over the ks
// waves of one block (blockDim = 64 ks, one 16x16 output tile per block):
// wave s reduces the s-th range of K, and wave 0 adds the partial tiles in
// s order, the order of a serial sum of the ranges.  No epilogue.
template <bool AV, bool BV>
__global__ void __launch_bounds__(1024) k_gemm_ks(const uint16_t *__restrict__ A, int64_t sam, int64_t sak,
                                                  const uint16_t *__restrict__ B, int64_t sbk, int64_t sbn, int M,
                                                  int N, int K, float *__restrict__ C, int64_t ldc) {
    __shared__ f4 part[16][64];
    const int ks = (int) (blockDim.x >> 6), s = (int) (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int tn = N / 16, tile = (int) blockIdx.x;
    if (tile >= (M / 16) * tn) return;
    const int m0 = (tile / tn) * 16, n0 = (tile % tn) * 16;
    const int r = lane & 15, kq = (lane >> 4) * 4;
    const int kl = K / ks, kb = s * kl;
    f4 c = {0.0f, 0.0f, 0.0f, 0.0f};
    const uint16_t *pa = A + (int64_t) (m0 + r) * sam + (int64_t) (kb + kq) * sak;
    const uint16_t *pb = B + (int64_t) (kb + kq) * sbk + (int64_t) (n0 + r) * sbn;
    s4 a = load_k4<AV>(pa, sak), b = load_k4<BV>(pb, sbk);
    for (int k0 = 0; k0 < kl; k0 += 16) {
        s4 an = a, bn = b;
        if (k0 + 16 < kl) {
            an = load_k4<AV>(pa + (int64_t) (k0 + 16) * sak, sak);
            bn = load_k4<BV>(pb + (int64_t) (k0 + 16) * sbk, sbk);
        }
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
        a = an;
        b = bn;
    }
    part[s][lane] = c;
    __syncthreads();
    if (s != 0) return;
    f4 v = part[0][lane];
    for (int q = 1; q < ks; ++q) {
        const f4 w = part[q][lane];
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] += w[i];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) C[(int64_t) (m0 + kq + i) * ldc + n0 + r] = v[i];
}

// launch k_gemm with the load form of each operand
inline void launch_gemm(unsigned blocks, hipStream_t s, const uint16_t *A, int64_t sam, int64_t sak,
                        const uint16_t *B, int64_t sbk, int64_t sbn, int M, int N, int K, float *C, int64_t ldc,
                        uint16_t *Cb, const float *bias, int relu, const float *mask) {
    if (sak == 1 && sbk == 1)
        k_gemm<true, true><<<blocks, 256, 0, s>>>(A, sam, sak, B, sbk, sbn, M, N, K, C, ldc, Cb, bias, relu, mask);
    else if (sak == 1)
        k_gemm<true, false><<<blocks, 256, 0, s>>>(A, sam, sak, B, sbk, sbn, M, N, K, C, ldc, Cb, bias, relu, mask);
    else if (sbk == 1)
        k_gemm<false, true><<<blocks, 256, 0, s>>>(A, sam, sak, B, sbk, sbn, M, N, K, C, ldc, Cb, bias, relu, mask);
    else
        k_gemm<false, false><<<blocks, 256, 0, s>>>(A, sam, sak, B, sbk, sbn, M, N, K, C, ldc, Cb, bias, relu, mask);
}

__global__ void k_to_bf16(const float *__restrict__ x, uint16_t *__restrict__ y, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x;
    if (i < n) y[i] = to_bf16(x[i]);
}

// training steps replayed from a graph find their batch here: batch b of
// the epoch takes permutation entries [b * batch_rows, ...) and the Adam
// corrections c12[2b], c12[2b + 1]
struct TrainCtl {
    int64_t batch;
};

// rows of the padded BF16 input: row r <- sample idx[r] (r < nb), zero
// beyond; with ctl, idx is advanced by ctl->batch * nb
__global__ void k_gather(const float *__restrict__ x, int64_t ldx, int nin, const int32_t *__restrict__ idx, int nb,
                         int rows, uint16_t *__restrict__ xb, int ldb, const int32_t *__restrict__ y,
                         int32_t *__restrict__ yb, const TrainCtl *__restrict__ ctl) {
    const int r = blockIdx.x, c = threadIdx.x;
    if (r >= rows) return;
    if (ctl) idx += ctl->batch * nb;
    const int64_t s = r < nb ? idx[r] : -1;
    for (int k = c; k < ldb; k += blockDim.x) xb[(int64_t) r * ldb + k] = (s >= 0 && k < nin) ? to_bf16(x[s * ldx + k]) : 0;
    if (c == 0) yb[r] = (s >= 0 && y) ? y[s] : -1;
}

// softmax cross-entropy over the first nout columns; dZ = (softmax - onehot) / nb
__global__ void k_softmax_ce(const float *__restrict__ z, int ldz, int nout, const int32_t *__restrict__ yb, int rows,
                             float inv_nb, float *__restrict__ dz, uint16_t *__restrict__ dz16,
                             float *__restrict__ loss) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int y = yb[r];
    const float *zr = z + (int64_t) r * ldz;
    float *dr = dz + (int64_t) r * ldz;
    uint16_t *d16 = dz16 + (int64_t) r * ldz;
    if (y < 0) {
        for (int c = 0; c < ldz; ++c) {
            dr[c] = 0.0f;
            d16[c] = 0;
        }
        loss[r] = 0.0f;
        return;
    }
    float mx = zr[0];
    for (int c = 1; c < nout; ++c) mx = fmaxf(mx, zr[c]);
    float s = 0.0f;
    for (int c = 0; c < nout; ++c) s += expf(zr[c] - mx);
    loss[r] = (mx + logf(s)) - zr[y];
    for (int c = 0; c < ldz; ++c) {
        float g = 0.0f;
        if (c < nout) g = (expf(zr[c] - mx) / s - (c == y ? 1.0f : 0.0f)) * inv_nb;
        dr[c] = g;
        d16[c] = to_bf16(g);
    }
}

// out[o] = sum over the rows of dz[r*ld + o], for every layer's bias in one
// launch.  A block takes 16 columns
// with 64 lanes each: lane y adds rows y, y+64, ... (coalesced across the
// columns) and the first lane of a column adds the 64 partial sums in lane
// order, so the result does not depend on timing.
// one layer's bias gradient in the all-layers column sum: its dZ, leading
// dimension (= columns) and output, and its first block
struct ColsumT {
    const float *dz;
    int ld;
    float *out;
    int blk0;
};

__global__ void __launch_bounds__(1024) k_colsum(const ColsumT *__restrict__ tab, int nl, int rows) {
    __shared__ float part[64][17];
    int l = 0;
    while (l + 1 < nl && tab[l + 1].blk0 <= (int) blockIdx.x) ++l;
    const ColsumT &t = tab[l];
    const int x = threadIdx.x & 15, y = threadIdx.x >> 4;
    const int o = ((int) blockIdx.x - t.blk0) * 16 + x;
    float s = 0.0f;
    if (o < t.ld)
        for (int r = y; r < rows; r += 64) s += t.dz[(int64_t) r * t.ld + o];
    part[y][x] = s;
    __syncthreads();
    if (y == 0 && o < t.ld) {
        float v = part[0][x];
        for (int k = 1; k < 64; ++k) v += part[k][x];
        t.out[o] = v;
    }
}

// one parameter tensor of the Adam step (flat elements [off, off + n))
struct AdamT {
    float *w, *g, *m, *v;
    uint16_t *w16;  // BF16 copy for the next forward product (weights only)
    int64_t n, off;
};

// Adam over every weight and bias tensor in one launch, refreshing the BF16
// weight copies; the bias corrections c1, c2 come from the host, or with
// ctl from c12 at the current batch
__global__ void k_adam_all(const AdamT *__restrict__ tab, int nt, int64_t total, float lr, float c1, float c2,
                           const TrainCtl *__restrict__ ctl, const float *__restrict__ c12) {
    const int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x;
    if (i >= total) return;
    int lo = 0, hi = nt - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[mid].off <= i) lo = mid;
        else hi = mid - 1;
    }
    const AdamT &t = tab[lo];
    const int64_t j = i - t.off;
    if (ctl) {
        c1 = c12[2 * ctl->batch];
        c2 = c12[2 * ctl->batch + 1];
    }
    const float gi = t.g[j];
    const float mi = 0.9f * t.m[j] + 0.1f * gi;
    const float vi = 0.999f * t.v[j] + 0.001f * gi * gi;
    t.m[j] = mi;
    t.v[j] = vi;
    const float w = t.w[j] - lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
    t.w[j] = w;
    if (t.w16) t.w16[j] = to_bf16(w);
}

// per-sample losses of a batch scattered to their sample slots (with ctl:
// the current batch's slots)
__global__ void k_scatter_loss(const float *__restrict__ lb, const int32_t *__restrict__ idx, int nb,
                               float *__restrict__ all, const TrainCtl *__restrict__ ctl) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (ctl) idx += ctl->batch * nb;
    if (r < nb) all[idx[r]] = lb[r];
}

// the last kernel of a replayed training step
__global__ void k_next_batch(TrainCtl *ctl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) ctl->batch += 1;
}

inline unsigned nblk(int64_t n, int b = 256) { return (unsigned) ((n + b - 1) / b); }

}  // namespace dmlp

using namespace dmlp;

struct dacti_mlp {
    int n_in = 0, H = 0, NL = 0, n_out = 0, L = 0;
    std::vector<int> ir, orr, ip, op;  // real and padded in/out widths per layer
    std::vector<float *> W, b, gW, gb, mW, vW, mb, vb;
    std::vector<uint16_t *> W16;
    int cap = 0;  // rows of the batch buffers
    uint16_t *X16 = nullptr;
    std::vector<float *> Z;
    std::vector<uint16_t *> A16;
    float *lossb = nullptr;
    // each layer's dZ (FP32 and BF16): the masked input gradient of layer
    // l is written into layer l-1's while layer l's is read, and the bias
    // gradients of all layers are column sums of them, taken at the end
    std::vector<float *> dZl;
    std::vector<uint16_t *> dZ16l;
    ColsumT *d_cs = nullptr;  // per-layer column-sum table (device)
    int cs_blocks = 0;
    int32_t *yb = nullptr;
    AdamT *d_adam = nullptr;  // every parameter tensor, for k_adam_all
    int n_adam = 0;
    int64_t adam_total = 0;
    TrainCtl *d_ctl = nullptr;
    float *d_c12 = nullptr;  // Adam corrections of one epoch's replayed steps
    int64_t c12_cap = 0;
    hipStream_t s = nullptr;
    int dev = 0;  // device the network lives on
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double train_ms = 0, train_flops = 0, flops = 0;
    int64_t t = 0;  // Adam steps
    std::string err;
};

namespace {

thread_local std::string g_err;

#define MCHK(m, call)                                                        \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) {                                              \
            (m)->err = std::string(#call) + ": " + hipGetErrorString(e_);    \
            return -2;                                                       \
        }                                                                    \
    } while (0)

template <typename T>
int dalloc(dacti_mlp *m, T **p, size_t n) {
    MCHK(m, hipMalloc((void **) p, sizeof(T) * std::max<size_t>(n, 1)));
    MCHK(m, hipMemset(*p, 0, sizeof(T) * std::max<size_t>(n, 1)));
    return 0;
}

int ensure_rows(dacti_mlp *m, int rows) {
    if (rows <= m->cap) return 0;
    MCHK(m, hipStreamSynchronize(m->s));
    auto fr = [](void *p) {
        if (p) (void) hipFree(p);
    };
    fr(m->X16), fr(m->yb), fr(m->lossb), fr(m->d_cs);
    m->d_cs = nullptr;
    for (auto p : m->Z) fr(p);
    for (auto p : m->A16) fr(p);
    for (auto p : m->dZl) fr(p);
    for (auto p : m->dZ16l) fr(p);
    const int cap = pad16(rows);
    int maxw = 16;
    for (int l = 0; l < m->L; ++l) maxw = std::max({maxw, m->ip[l], m->op[l]});
    int rc;
    if ((rc = dalloc(m, &m->X16, (size_t) cap * m->ip[0])) || (rc = dalloc(m, &m->yb, (size_t) cap)) ||
        (rc = dalloc(m, &m->lossb, (size_t) cap)))
        return rc;
    m->Z.assign(m->L, nullptr);
    m->A16.assign(m->L, nullptr);
    m->dZl.assign(m->L, nullptr);
    m->dZ16l.assign(m->L, nullptr);
    std::vector<ColsumT> cs;
    int blk = 0;
    for (int l = 0; l < m->L; ++l) {
        if ((rc = dalloc(m, &m->Z[l], (size_t) cap * m->op[l])) || (rc = dalloc(m, &m->A16[l], (size_t) cap * m->op[l])) ||
            (rc = dalloc(m, &m->dZl[l], (size_t) cap * m->op[l])) || (rc = dalloc(m, &m->dZ16l[l], (size_t) cap * m->op[l])))
            return rc;
        cs.push_back(ColsumT{m->dZl[l], m->op[l], m->gb[l], blk});
        blk += (m->op[l] + 15) / 16;
    }
    m->cs_blocks = blk;
    MCHK(m, hipMalloc((void **) &m->d_cs, sizeof(ColsumT) * cs.size()));
    MCHK(m, hipMemcpy(m->d_cs, cs.data(), sizeof(ColsumT) * cs.size(), hipMemcpyHostToDevice));
    m->cap = cap;
    return 0;
}

// A long reduction over few output tiles (the weight gradients: K = batch
// rows) is split so that enough wavefronts run; otherwise one wave per tile.
int gemm(dacti_mlp *m, const uint16_t *A, int64_t sam, int64_t sak, const uint16_t *B, int64_t sbk, int64_t sbn, int M,
         int N, int K, float *C, int64_t ldc, uint16_t *Cb, const float *bias, int relu, const float *mask = nullptr) {
    const int64_t tiles = (int64_t) (M / 16) * (N / 16);
    int ks = 1;
    if (!bias && !relu && !Cb && !mask)
        while (ks < 16 && tiles * ks < 2048 && K / (2 * ks) >= 64 && (K / 16) % (2 * ks) == 0) ks *= 2;
    if (ks > 1) {
        const unsigned nb = (unsigned) tiles, nt = 64u * (unsigned) ks;
        if (sak == 1 && sbk == 1)
            k_gemm_ks<true, true><<<nb, nt, 0, m->s>>>(A, sam, sak, B, sbk, sbn, M, N, K, C, ldc);
        else if (sak == 1)
            k_gemm_ks<true, false><<<nb, nt, 0, m->s>>>(A, sam, sak, B, sbk, sbn, M, N, K, C, ldc);
        else if (sbk == 1)
            k_gemm_ks<false, true><<<nb, nt, 0, m->s>>>(A, sam, sak, B, sbk, sbn, M, N, K, C, ldc);
        else
            k_gemm_ks<false, false><<<nb, nt, 0, m->s>>>(A, sam, sak, B, sbk, sbn, M, N, K, C, ldc);
    } else {
        launch_gemm(nblk(tiles * 64), m->s, A, sam, sak, B, sbk, sbn, M, N, K, C, ldc, Cb, bias, relu, mask);
    }
    m->flops += 2.0 * M * N * K;
    MCHK(m, hipGetLastError());
    return 0;
}

// forward over `rows` rows already in X16
int forward(dacti_mlp *m, int rows) {
    const uint16_t *a = m->X16;
    int lda = m->ip[0];
    for (int l = 0; l < m->L; ++l) {
        const bool last = l == m->L - 1;
        int rc = gemm(m, a, lda, 1, m->W16[l], 1, m->ip[l], rows, m->op[l], m->ip[l], m->Z[l], m->op[l], m->A16[l],
                      m->b[l], last ? 0 : 1);
        if (rc) return rc;
        a = m->A16[l];
        lda = m->op[l];
    }
    return 0;
}

// gradients of the mean loss of nb labelled rows (of `rows`) into gW, gb
int backward(dacti_mlp *m, int rows, int nb) {
    const int L = m->L;
    k_softmax_ce<<<nblk(rows), 256, 0, m->s>>>(m->Z[L - 1], m->op[L - 1], m->n_out, m->yb, rows, 1.0f / (float) nb,
                                              m->dZl[L - 1], m->dZ16l[L - 1], m->lossb);
    MCHK(m, hipGetLastError());
    for (int l = L - 1; l >= 0; --l) {
        const uint16_t *aprev = l > 0 ? m->A16[l - 1] : m->X16;
        const int ldp = m->ip[l], ldz = m->op[l];
        // gW_l(o, i) = sum_r dZ(r, o) A_prev(r, i)
        int rc = gemm(m, m->dZ16l[l], 1, ldz, aprev, ldp, 1, m->op[l], m->ip[l], rows, m->gW[l], m->ip[l], nullptr,
                      nullptr, 0);
        if (rc) return rc;
        if (l > 0) {
            // dZ_{l-1}(r, i) = sum_o dZ_l(r, o) W(o, i), masked by the ReLU of layer l-1
            rc = gemm(m, m->dZ16l[l], ldz, 1, m->W16[l], m->ip[l], 1, rows, m->ip[l], m->op[l], m->dZl[l - 1], m->ip[l],
                      m->dZ16l[l - 1], nullptr, 0, m->Z[l - 1]);
            if (rc) return rc;
        }
    }
    // gb_l(o) = sum_r dZ_l(r, o), every layer
    k_colsum<<<(unsigned) m->cs_blocks, 1024, 0, m->s>>>(m->d_cs, L, rows);
    MCHK(m, hipGetLastError());
    return 0;
}

int refresh_bf16(dacti_mlp *m) {
    for (int l = 0; l < m->L; ++l) {
        const int64_t n = (int64_t) m->op[l] * m->ip[l];
        k_to_bf16<<<nblk(n), 256, 0, m->s>>>(m->W[l], m->W16[l], n);
    }
    MCHK(m, hipGetLastError());
    return 0;
}

int64_t n_params(const dacti_mlp *m) {
    int64_t n = 0;
    for (int l = 0; l < m->L; ++l) n += (int64_t) m->orr[l] * m->ir[l] + m->orr[l];
    return n;
}

// flat parameters <-> padded device arrays
int put_params(dacti_mlp *m, const float *p) {
    for (int l = 0; l < m->L; ++l) {
        std::vector<float> w((size_t) m->op[l] * m->ip[l], 0.0f), bb((size_t) m->op[l], 0.0f);
        for (int o = 0; o < m->orr[l]; ++o)
            for (int i = 0; i < m->ir[l]; ++i) w[(size_t) o * m->ip[l] + i] = *p++;
        for (int o = 0; o < m->orr[l]; ++o) bb[(size_t) o] = *p++;
        MCHK(m, hipMemcpy(m->W[l], w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice));
        MCHK(m, hipMemcpy(m->b[l], bb.data(), sizeof(float) * bb.size(), hipMemcpyHostToDevice));
    }
    return refresh_bf16(m);
}

int get_padded(dacti_mlp *m, const std::vector<float *> &W, const std::vector<float *> &b, float *p) {
    MCHK(m, hipStreamSynchronize(m->s));
    for (int l = 0; l < m->L; ++l) {
        std::vector<float> w((size_t) m->op[l] * m->ip[l]), bb((size_t) m->op[l]);
        MCHK(m, hipMemcpy(w.data(), W[l], sizeof(float) * w.size(), hipMemcpyDeviceToHost));
        MCHK(m, hipMemcpy(bb.data(), b[l], sizeof(float) * bb.size(), hipMemcpyDeviceToHost));
        for (int o = 0; o < m->orr[l]; ++o)
            for (int i = 0; i < m->ir[l]; ++i) *p++ = w[(size_t) o * m->ip[l] + i];
        for (int o = 0; o < m->orr[l]; ++o) *p++ = bb[(size_t) o];
    }
    return 0;
}

void destroy(dacti_mlp *m) {
    if (!m) return;
    if (m->s) (void) hipStreamSynchronize(m->s);
    for (auto *v : {&m->W, &m->b, &m->gW, &m->gb, &m->mW, &m->vW, &m->mb, &m->vb})
        for (float *p : *v)
            if (p) (void) hipFree(p);
    for (auto p : m->W16)
        if (p) (void) hipFree(p);
    for (auto p : m->Z)
        if (p) (void) hipFree(p);
    for (auto p : m->A16)
        if (p) (void) hipFree(p);
    for (auto p : m->dZl)
        if (p) (void) hipFree(p);
    for (auto p : m->dZ16l)
        if (p) (void) hipFree(p);
    for (void *p : {(void *) m->X16, (void *) m->yb, (void *) m->lossb, (void *) m->d_adam, (void *) m->d_ctl,
                    (void *) m->d_c12, (void *) m->d_cs})
        if (p) (void) hipFree(p);
    if (m->e0) (void) hipEventDestroy(m->e0);
    if (m->e1) (void) hipEventDestroy(m->e1);
    if (m->s) (void) hipStreamDestroy(m->s);
    delete m;
}

// upload n samples to the device (FP32) and labels
int upload_data(dacti_mlp *m, const float *x, const int32_t *y, int64_t n, float **dx, int32_t **dy) {
    *dx = nullptr;
    *dy = nullptr;
    MCHK(m, hipMalloc((void **) dx, sizeof(float) * (size_t) std::max<int64_t>(n * m->n_in, 1)));
    MCHK(m, hipMemcpy(*dx, x, sizeof(float) * (size_t) (n * m->n_in), hipMemcpyHostToDevice));
    if (y) {
        MCHK(m, hipMalloc((void **) dy, sizeof(int32_t) * (size_t) std::max<int64_t>(n, 1)));
        MCHK(m, hipMemcpy(*dy, y, sizeof(int32_t) * (size_t) n, hipMemcpyHostToDevice));
    }
    return 0;
}

// batch rows from the device index array idx (nb entries)
int load_rows(dacti_mlp *m, const float *dx, const int32_t *dy, const int32_t *idx, int nb, int rows,
              const TrainCtl *ctl = nullptr) {
    k_gather<<<(unsigned) rows, 64, 0, m->s>>>(dx, m->n_in, m->n_in, idx, nb, rows, m->X16, m->ip[0], dy, m->yb, ctl);
    MCHK(m, hipGetLastError());
    return 0;
}

// a device copy of host indices (synchronous)
int upload_idx(dacti_mlp *m, const std::vector<int32_t> &h, int32_t **d) {
    *d = nullptr;
    MCHK(m, hipMalloc((void **) d, sizeof(int32_t) * std::max<size_t>(h.size(), 1)));
    MCHK(m, hipMemcpy(*d, h.data(), sizeof(int32_t) * h.size(), hipMemcpyHostToDevice));
    return 0;
}

}  // namespace


// ------------------------------------------------------------------
// Inference that allocates nothing (dacti_mlp_forward_noalloc).  Its
// activations live in a fixed device array, `scratch`, of
// NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.  In host code
// the array's name is only a stand-in: the device address is looked up once
// per device with hipGetSymbolAddress and kept in g_scratch_dev[dev].
namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(32) << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;
constexpr int NOALLOC_MAX_ROWS = 8192;

__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
unsigned char *g_scratch_dev[NOALLOC_MAX_DEVICES] = {};

// device address of this device's scratch (current device = m->dev)
int scratch_ptr(dacti_mlp *m, unsigned char **out) {
    if (m->dev < 0 || m->dev >= NOALLOC_MAX_DEVICES) {
        m->err = "dacti_mlp_forward_noalloc: device id out of range";
        return -1;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[m->dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            m->err = std::string("dacti_mlp_forward_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e);
            return -2;
        }
        g_scratch_dev[m->dev] = (unsigned char *) p;
    }
    *out = g_scratch_dev[m->dev];
    return 0;
}

// padded BF16 rows: row r < nb <- x[r], zero beyond nb and beyond nin
__global__ void k_stage_rows(const float *__restrict__ x, int nin, int nb, int rows, uint16_t *__restrict__ xb,
                             int ldb) {
    const int r = blockIdx.x;
    if (r >= rows) return;
    for (int k = threadIdx.x; k < ldb; k += blockDim.x)
        xb[(int64_t) r * ldb + k] = (r < nb && k < nin) ? to_bf16(x[(int64_t) r * nin + k]) : 0;
}

// logits (nb x nout, dense) from the padded last-layer output
__global__ void k_copy_logits(const float *__restrict__ z, int ldz, int nout, int nb, float *__restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x;
    if (i >= (int64_t) nb * nout) return;
    out[i] = z[(i / nout) * ldz + i % nout];
}

}  // namespace

extern "C" {

int dacti_mlp_create(int32_t n_in, int32_t n_hidden, int32_t n_layers, int32_t n_out, uint64_t seed,
                     dacti_mlp **out) {
    if (!out || n_in < 1 || n_hidden < 1 || n_layers < 1 || n_out < 2 || n_in > 4096 || n_hidden > 4096 ||
        n_layers > 16 || n_out > 256) {
        g_err = "dacti_mlp_create: invalid sizes";
        return -1;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        g_err = "dacti_mlp_create: no HIP device available";
        return -2;
    }
    dacti_mlp *m = new dacti_mlp();
    if (hipGetDevice(&m->dev) != hipSuccess) {
        delete m;
        g_err = "dacti_mlp_create: no current HIP device";
        return -2;
    }
    m->n_in = n_in;
    m->H = n_hidden;
    m->NL = n_layers;
    m->n_out = n_out;
    m->L = n_layers + 1;
    for (int l = 0; l < m->L; ++l) {
        const int in = l == 0 ? n_in : n_hidden, o = l == m->L - 1 ? n_out : n_hidden;
        m->ir.push_back(in);
        m->orr.push_back(o);
        m->ip.push_back(pad16(in));
        m->op.push_back(pad16(o));
    }
    auto fail = [&](int rc) {
        g_err = m->err;
        destroy(m);
        return rc;
    };
    if (hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&m->e0) != hipSuccess ||
        hipEventCreate(&m->e1) != hipSuccess) {
        m->err = "stream/event creation failed";
        return fail(-2);
    }
    for (auto *v : {&m->W, &m->b, &m->gW, &m->gb, &m->mW, &m->vW, &m->mb, &m->vb}) v->assign(m->L, nullptr);
    m->W16.assign(m->L, nullptr);
    for (int l = 0; l < m->L; ++l) {
        const size_t nw = (size_t) m->op[l] * m->ip[l], nb = (size_t) m->op[l];
        int rc;
        if ((rc = dalloc(m, &m->W[l], nw)) || (rc = dalloc(m, &m->gW[l], nw)) || (rc = dalloc(m, &m->mW[l], nw)) ||
            (rc = dalloc(m, &m->vW[l], nw)) || (rc = dalloc(m, &m->W16[l], nw)) || (rc = dalloc(m, &m->b[l], nb)) ||
            (rc = dalloc(m, &m->gb[l], nb)) || (rc = dalloc(m, &m->mb[l], nb)) || (rc = dalloc(m, &m->vb[l], nb)))
            return fail(rc);
    }
    std::mt19937_64 rng(seed);
    std::vector<float> p((size_t) n_params(m));
    size_t k = 0;
    for (int l = 0; l < m->L; ++l) {
        std::uniform_real_distribution<float> u(-1.0f / std::sqrt((float) m->ir[l]), 1.0f / std::sqrt((float) m->ir[l]));
        for (int i = 0; i < m->orr[l] * m->ir[l]; ++i) p[k++] = u(rng);
        for (int o = 0; o < m->orr[l]; ++o) p[k++] = 0.0f;
    }
    int rc = put_params(m, p.data());
    if (rc) return fail(rc);
    // the Adam table: every weight and bias tensor, flat offsets
    std::vector<AdamT> tab;
    int64_t off = 0;
    for (int l = 0; l < m->L; ++l) {
        const int64_t nw = (int64_t) m->op[l] * m->ip[l];
        tab.push_back(AdamT{m->W[l], m->gW[l], m->mW[l], m->vW[l], m->W16[l], nw, off});
        off += nw;
        tab.push_back(AdamT{m->b[l], m->gb[l], m->mb[l], m->vb[l], nullptr, (int64_t) m->op[l], off});
        off += m->op[l];
    }
    m->n_adam = (int) tab.size();
    m->adam_total = off;
    if (hipMalloc((void **) &m->d_adam, sizeof(AdamT) * tab.size()) != hipSuccess ||
        hipMemcpy(m->d_adam, tab.data(), sizeof(AdamT) * tab.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc((void **) &m->d_ctl, sizeof(TrainCtl)) != hipSuccess) {
        m->err = "optimizer table allocation failed";
        return fail(-2);
    }
    if (hipStreamSynchronize(m->s) != hipSuccess) return fail(-2);
    *out = m;
    return 0;
}

int dacti_mlp_n_params(dacti_mlp *m, int64_t *n) {
    if (!m || !n) return -1;
    *n = n_params(m);
    return 0;
}

int dacti_mlp_get_params(dacti_mlp *m, float *params) {
    if (!m || !params) return -1;
    return get_padded(m, m->W, m->b, params);
}

int dacti_mlp_set_params(dacti_mlp *m, const float *params) {
    if (!m || !params) return -1;
    MCHK(m, hipStreamSynchronize(m->s));
    const int rc = put_params(m, params);
    if (rc) return rc;
    MCHK(m, hipStreamSynchronize(m->s));
    return 0;
}

int dacti_mlp_forward(dacti_mlp *m, const float *x, int64_t n, float *logits) {
    if (!m || (!x && n > 0) || (!logits && n > 0) || n < 0) return -1;
    if (n == 0) return 0;
    const int chunk = 8192;
    int rc = ensure_rows(m, (int) std::min<int64_t>(n, chunk));
    if (rc) return rc;
    float *dx = nullptr;
    int32_t *dy = nullptr;
    rc = upload_data(m, x, nullptr, n, &dx, &dy);
    if (rc) {
        if (dx) (void) hipFree(dx);
        return rc;
    }
    std::vector<int32_t> ih((size_t) n);
    for (int64_t i = 0; i < n; ++i) ih[(size_t) i] = (int32_t) i;
    int32_t *di = nullptr;
    rc = upload_idx(m, ih, &di);
    std::vector<float> z;
    const int L = m->L;
    for (int64_t s0 = 0; s0 < n && rc == 0; s0 += chunk) {
        const int nb = (int) std::min<int64_t>(chunk, n - s0), rows = pad16(nb);
        if ((rc = load_rows(m, dx, nullptr, di + s0, nb, rows)) || (rc = forward(m, rows))) break;
        z.resize((size_t) rows * m->op[L - 1]);
        if (hipMemcpyAsync(z.data(), m->Z[L - 1], sizeof(float) * z.size(), hipMemcpyDeviceToHost, m->s) != hipSuccess ||
            hipStreamSynchronize(m->s) != hipSuccess) {
            m->err = "logits copy failed";
            rc = -2;
            break;
        }
        for (int r = 0; r < nb; ++r)
            for (int c = 0; c < m->n_out; ++c) logits[(s0 + r) * m->n_out + c] = z[(size_t) r * m->op[L - 1] + c];
    }
    (void) hipFree(dx);
    if (di) (void) hipFree(di);
    return rc;
}

int dacti_mlp_forward_noalloc(dacti_mlp *m, const float *x_dev, int64_t n, float *logits_dev) {
    if (!m) return -1;
    if (n < 0 || n > INT32_MAX || (n > 0 && (!x_dev || !logits_dev))) {
        m->err = "dacti_mlp_forward_noalloc: invalid arguments";
        return -1;
    }
    if (n == 0) return 0;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != m->dev) {
        m->err = "dacti_mlp_forward_noalloc: the network's device must be current";
        return -1;
    }
    unsigned char *base = nullptr;
    int rc = scratch_ptr(m, &base);
    if (rc) return rc;
    const int L = m->L;
    int maxw = 16;
    for (int l = 0; l < L; ++l) maxw = std::max(maxw, m->op[l]);
    // per row: BF16 input, FP32 layer output, two BF16 activation buffers
    const size_t row_bytes = 2 * (size_t) m->ip[0] + 4 * (size_t) maxw + 2 * 2 * (size_t) maxw;
    int64_t cap = (int64_t) (NOALLOC_SCRATCH_BYTES / (row_bytes + 64)) / 16 * 16;
    if (cap < 16) {
        m->err = "dacti_mlp_forward_noalloc: the network is too wide for its " +
                 std::to_string(NOALLOC_SCRATCH_BYTES >> 20) + " MiB scratch (" + std::to_string(row_bytes) +
                 " bytes per row, 16 rows needed)";
        return -1;
    }
    cap = std::min<int64_t>(cap, NOALLOC_MAX_ROWS);
    auto align = [](size_t b) { return (b + 255) / 256 * 256; };
    uint16_t *X16 = (uint16_t *) base;
    float *Z = (float *) (base + align(2 * (size_t) cap * m->ip[0]));
    uint16_t *Aping = (uint16_t *) ((unsigned char *) Z + align(4 * (size_t) cap * maxw));
    uint16_t *Apong = (uint16_t *) ((unsigned char *) Aping + align(2 * (size_t) cap * maxw));
    if ((unsigned char *) Apong + 2 * (size_t) cap * maxw > base + NOALLOC_SCRATCH_BYTES) {
        m->err = "dacti_mlp_forward_noalloc: internal scratch layout overflow";
        return -1;
    }
    for (int64_t s0 = 0; s0 < n; s0 += cap) {
        const int nb = (int) std::min<int64_t>(cap, n - s0), rows = pad16(nb);
        k_stage_rows<<<(unsigned) rows, 64, 0, m->s>>>(x_dev + s0 * m->n_in, m->n_in, nb, rows, X16, m->ip[0]);
        const uint16_t *a = X16;
        int lda = m->ip[0];
        for (int l = 0; l < L; ++l) {
            uint16_t *cb = (l % 2 == 0) ? Aping : Apong;
            const int tiles = (rows / 16) * (m->op[l] / 16);
            launch_gemm(nblk((int64_t) tiles * 64), m->s, a, lda, 1, m->W16[l], 1, m->ip[l], rows, m->op[l],
                                                              m->ip[l], Z, m->op[l], cb, m->b[l],
                                                              l == L - 1 ? 0 : 1, nullptr);
            a = cb;
            lda = m->op[l];
        }
        k_copy_logits<<<nblk((int64_t) nb * m->n_out), 256, 0, m->s>>>(Z, m->op[L - 1], m->n_out, nb,
                                                                         logits_dev + s0 * m->n_out);
        MCHK(m, hipGetLastError());
    }
    MCHK(m, hipStreamSynchronize(m->s));
    return 0;
}

int dacti_mlp_loss_grad(dacti_mlp *m, const float *x, const int32_t *y, int64_t n, float *loss, float *grad) {
    if (!m || !x || !y || n < 1 || n > (1 << 20)) return -1;
    int rc = ensure_rows(m, (int) n);
    if (rc) return rc;
    float *dx = nullptr;
    int32_t *dy = nullptr;
    rc = upload_data(m, x, y, n, &dx, &dy);
    std::vector<int32_t> ih((size_t) n);
    for (int64_t r = 0; r < n; ++r) ih[(size_t) r] = (int32_t) r;
    int32_t *di = nullptr;
    if (!rc) rc = upload_idx(m, ih, &di);
    const int rows = pad16((int) n);
    if (!rc) rc = load_rows(m, dx, dy, di, (int) n, rows);
    if (!rc) rc = forward(m, rows);
    if (!rc) rc = backward(m, rows, (int) n);
    if (!rc) {
        std::vector<float> lb((size_t) rows);
        if (hipMemcpyAsync(lb.data(), m->lossb, sizeof(float) * lb.size(), hipMemcpyDeviceToHost, m->s) != hipSuccess ||
            hipStreamSynchronize(m->s) != hipSuccess) {
            m->err = "loss copy failed";
            rc = -2;
        } else if (loss) {
            double s = 0;
            for (int64_t r = 0; r < n; ++r) s += lb[(size_t) r];
            *loss = (float) (s / (double) n);
        }
    }
    if (!rc && grad) rc = get_padded(m, m->gW, m->gb, grad);
    (void) hipStreamSynchronize(m->s);
    if (dx) (void) hipFree(dx);
    if (dy) (void) hipFree(dy);
    if (di) (void) hipFree(di);
    return rc;
}

namespace {

// One training step (batch of nb rows): gather, forward, backward, loss
// scatter, Adam over every tensor.  With ctl the batch comes from the
// device counter (a step replayed from a graph) and the counter advances at
// the end; otherwise idx already points at the batch and c1, c2 are given.
int train_step(dacti_mlp *m, const float *dx, const int32_t *dy, const int32_t *idx, int nb, float *lall, float lr,
               float c1, float c2, const TrainCtl *ctl) {
    const int rows = pad16(nb);
    int rc;
    if ((rc = load_rows(m, dx, dy, idx, nb, rows, ctl)) || (rc = forward(m, rows)) || (rc = backward(m, rows, nb)))
        return rc;
    k_scatter_loss<<<nblk(nb), 256, 0, m->s>>>(m->lossb, idx, nb, lall, ctl);
    k_adam_all<<<nblk(m->adam_total), 256, 0, m->s>>>(m->d_adam, m->n_adam, m->adam_total, lr, c1, c2, ctl,
                                                      m->d_c12);
    if (ctl) k_next_batch<<<1, 64, 0, m->s>>>(m->d_ctl);
    MCHK(m, hipGetLastError());
    return 0;
}

}  // namespace

// Mini-batch Adam.  The full batches of an epoch replay one HIP graph of the
// whole step (captured after the first step, which allocates what the step
// needs); their batch position and Adam corrections come from device
// memory.  The first step of a call and a last partial batch run eagerly.
int dacti_mlp_train(dacti_mlp *m, const float *x, const int32_t *y, int64_t n, int32_t epochs, int32_t batch,
                    float lr, uint64_t seed, float *loss) {
    if (!m || !x || !y || n < 1 || n > INT32_MAX / 2 || epochs < 0 || batch < 16 || (batch % 16) != 0 || batch > 65536) {
        if (m) m->err = "dacti_mlp_train: invalid arguments (batch: a multiple of 16, at most 65536)";
        return -1;
    }
    int rc = ensure_rows(m, batch);
    if (rc) return rc;
    float *dx = nullptr, *lall = nullptr;
    int32_t *dy = nullptr;
    rc = upload_data(m, x, y, n, &dx, &dy);
    if (!rc && hipMalloc((void **) &lall, sizeof(float) * (size_t) n) != hipSuccess) {
        m->err = "loss buffer allocation failed";
        rc = -3;
    }
    std::vector<int32_t> perm((size_t) n);
    for (int64_t i = 0; i < n; ++i) perm[(size_t) i] = (int32_t) i;
    int32_t *dperm = nullptr;
    if (!rc) rc = upload_idx(m, perm, &dperm);
    const int64_t nfull = n / batch;
    if (!rc && 2 * nfull > m->c12_cap) {
        if (m->d_c12) (void) hipFree(m->d_c12);
        m->d_c12 = nullptr;
        if (hipMalloc((void **) &m->d_c12, sizeof(float) * (size_t) (2 * nfull)) != hipSuccess) {
            m->err = "correction table allocation failed";
            rc = -3;
        } else {
            m->c12_cap = 2 * nfull;
        }
    }
    std::mt19937_64 rng(seed);
    std::vector<float> c12((size_t) (2 * nfull));
    hipGraphExec_t gx = nullptr;
    double step_flops = 0;
    m->flops = 0;
    if (!rc && hipEventRecord(m->e0, m->s) != hipSuccess) rc = -2;
    float last = 0.0f;
    auto corr = [](int64_t t, float &c1, float &c2) {
        c1 = 1.0f - std::pow(0.9f, (float) t);
        c2 = 1.0f - std::pow(0.999f, (float) t);
    };
    for (int ep = 0; ep < epochs && !rc; ++ep) {
        for (int64_t i = n - 1; i > 0; --i) std::swap(perm[(size_t) i], perm[(size_t) (rng() % (uint64_t) (i + 1))]);
        // the previous epoch's kernels read dperm and the tables: wait, then replace them
        for (int64_t b = 0; b < nfull; ++b) corr(m->t + b + 1, c12[(size_t) (2 * b)], c12[(size_t) (2 * b + 1)]);
        const TrainCtl zero{0};
        if (hipStreamSynchronize(m->s) != hipSuccess ||
            hipMemcpy(dperm, perm.data(), sizeof(int32_t) * (size_t) n, hipMemcpyHostToDevice) != hipSuccess ||
            (nfull > 0 &&
             hipMemcpy(m->d_c12, c12.data(), sizeof(float) * c12.size(), hipMemcpyHostToDevice) != hipSuccess) ||
            hipMemcpy(m->d_ctl, &zero, sizeof(TrainCtl), hipMemcpyHostToDevice) != hipSuccess) {
            m->err = "permutation upload failed";
            rc = -2;
            break;
        }
        int64_t b0 = 0;
        if (!gx && nfull > 1) {
            // first step eagerly, then capture the step
            float c1, c2;
            corr(m->t + 1, c1, c2);
            if ((rc = train_step(m, dx, dy, dperm, batch, lall, lr, c1, c2, nullptr))) break;
            m->t++;
            const TrainCtl one{1};
            if (hipMemcpyAsync(m->d_ctl, &one, sizeof(TrainCtl), hipMemcpyHostToDevice, m->s) != hipSuccess ||
                hipStreamSynchronize(m->s) != hipSuccess) {
                m->err = "batch counter upload failed";
                rc = -2;
                break;
            }
            const double f0 = m->flops;
            hipGraph_t graph = nullptr;
            if (hipStreamBeginCapture(m->s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                m->err = "training step capture failed";
                rc = -2;
                break;
            }
            rc = train_step(m, dx, dy, dperm, batch, lall, lr, 0.0f, 0.0f, m->d_ctl);
            const hipError_t e = hipStreamEndCapture(m->s, &graph);
            if (rc) break;
            step_flops = m->flops - f0;
            m->flops = f0;
            if (e != hipSuccess || hipGraphInstantiate(&gx, graph, nullptr, nullptr, 0) != hipSuccess) {
                if (graph) (void) hipGraphDestroy(graph);
                gx = nullptr;
                m->err = "training step graph failed";
                rc = -2;
                break;
            }
            (void) hipGraphDestroy(graph);
            b0 = 1;
        }
        for (int64_t b = b0; b < nfull && !rc; ++b) {
            if (gx) {
                if (hipGraphLaunch(gx, m->s) != hipSuccess) {
                    m->err = "training step replay failed";
                    rc = -2;
                    break;
                }
                m->flops += step_flops;
            } else {
                float c1, c2;
                corr(m->t + 1, c1, c2);
                rc = train_step(m, dx, dy, dperm + b * batch, batch, lall, lr, c1, c2, nullptr);
            }
            m->t++;
        }
        if (!rc && nfull * batch < n) {
            const int nb = (int) (n - nfull * batch);
            float c1, c2;
            corr(m->t + 1, c1, c2);
            rc = train_step(m, dx, dy, dperm + nfull * batch, nb, lall, lr, c1, c2, nullptr);
            m->t++;
        }
        if (!rc && ep == epochs - 1) {
            std::vector<float> la((size_t) n);
            if (hipMemcpyAsync(la.data(), lall, sizeof(float) * (size_t) n, hipMemcpyDeviceToHost, m->s) != hipSuccess ||
                hipStreamSynchronize(m->s) != hipSuccess) {
                m->err = "loss copy failed";
                rc = -2;
            } else {
                double s = 0;
                for (float v : la) s += v;
                last = (float) (s / (double) n);
            }
        }
    }
    if (!rc) {
        float ms = 0;
        if (hipEventRecord(m->e1, m->s) != hipSuccess || hipEventSynchronize(m->e1) != hipSuccess ||
            hipEventElapsedTime(&ms, m->e0, m->e1) != hipSuccess) {
            m->err = "timing failed";
            rc = -2;
        }
        m->train_ms = ms;
        m->train_flops = m->flops;
    }
    if (loss) *loss = last;
    (void) hipStreamSynchronize(m->s);
    if (gx) (void) hipGraphExecDestroy(gx);
    if (dx) (void) hipFree(dx);
    if (dy) (void) hipFree(dy);
    if (lall) (void) hipFree(lall);
    if (dperm) (void) hipFree(dperm);
    return rc;
}

int dacti_mlp_train_stats(dacti_mlp *m, double *ms, double *flops) {
    if (!m) return -1;
    if (ms) *ms = m->train_ms;
    if (flops) *flops = m->train_flops;
    return 0;
}

int dacti_mlp_destroy(dacti_mlp *m) {
    destroy(m);
    return 0;
}

const char *dacti_mlp_last_error(dacti_mlp *m) { return m ? m->err.c_str() : g_err.c_str(); }

}  // extern "C"
