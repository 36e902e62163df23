// Learned warm start: an MLP x0 -> (standardized) primal-dual iterate, trained on solved
// iterates.  Hand-written kernels: a bf16 MFMA GEMM (FP32 accumulation) with fused bias / GELU /
// output affine epilogues, the GELU backward, column sums, the MSE loss and Adam.  Weights are kept
// as FP32 masters with bf16 copies that feed the matrix cores.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <math.h>

#include "engine.hpp"

namespace spk {
namespace mlp {

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

__device__ __forceinline__ unsigned short toBf16(float f) {  // round to nearest even
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short) (u >> 16) | ((u & 0xffff) ? 0x40 : 0);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short) (u >> 16);
}
__device__ __forceinline__ float fromBf16(unsigned short h) { return __uint_as_float(((unsigned int) h) << 16); }

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float geluGrad(float x) {
    return 0.5f * (1.f + erff(x * 0.70710678118654752f)) + x * 0.3989422804014327f * expf(-0.5f * x * x);
}

// Epilogue variants of the GEMM.
enum Epi {
    EPI_STORE = 0,       // C = acc (FP32)
    EPI_BIAS_GELU = 1,   // Z = acc + bias (FP32), A = gelu(Z) (bf16)
    EPI_BIAS = 2,        // C = acc + bias (FP32)
    EPI_BIAS_AFFINE = 3, // C = (acc + bias) * scale[col] + shift[col] (FP32, output dtype T via outF/outD)
};

struct GemmArgs {
    const unsigned short *A;  // bf16, element (i, k) at A[i * sai + k * sak]
    const unsigned short *B;  // bf16, element (k, j) at B[k * sbk + j * sbj]
    long sai, sak, sbk, sbj;
    int M, N, K;
    float *C;                 // FP32 output (row-major, ldc)
    long ldc;
    unsigned short *Cb;       // optional bf16 copy of the activation (row-major, ldc)
    const float *bias, *scale, *shift;
    int epi;
};

constexpr int TM = 64, TN = 64, TK = 32, PADK = 8;

// C (M x N) = A (M x K) . B (K x N): 256 threads, a 64 x 64 tile per workgroup, each wave a
// 32 x 32 block of 2 x 2 MFMA 16x16x16 bf16 tiles.  Operands are staged in LDS as [row][k] (A)
// and [col][k] (B), so every lane reads its 4 consecutive k values with one 64-bit LDS load.
// MFMA 16x16x16 lane layout: A[i][k] from lane i + 16 (k / 4) item k % 4; B[k][j] from lane
// j + 16 (k / 4) item k % 4; D[i][j] in lane j + 16 (i / 4) item i % 4.
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
    __shared__ unsigned short As[TM][TK + PADK];
    __shared__ unsigned short Bs[TN][TK + PADK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = blockIdx.y * TM, j0 = blockIdx.x * TN;
    const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
    f32x4 acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.K; k0 += TK) {
        // stage: 64 x 32 of A and 64 x 32 of B (as [col][k]); 8 elements per thread each
        for (int t = tid; t < TM * TK; t += 256) {
            const int r = t / TK, kk = t % TK;
            const int gi = i0 + r, gk = k0 + kk;
            As[r][kk] = (gi < g.M && gk < g.K) ? g.A[(long) gi * g.sai + (long) gk * g.sak] : (unsigned short) 0;
        }
        for (int t = tid; t < TN * TK; t += 256) {
            const int cidx = t % TN, kk = t / TN;
            const int gj = j0 + cidx, gk = k0 + kk;
            Bs[cidx][kk] = (gj < g.N && gk < g.K) ? g.B[(long) gk * g.sbk + (long) gj * g.sbj] : (unsigned short) 0;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < TK; ks += 16) {
            bf16x4 af[2], bfr[2];
#pragma unroll
            for (int a = 0; a < 2; a++) {
                const unsigned short *p = &As[wi + 16 * a + (lane & 15)][ks + 4 * (lane >> 4)];
                af[a] = *reinterpret_cast<const bf16x4 *>(p);
            }
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const unsigned short *p = &Bs[wj + 16 * b + (lane & 15)][ks + 4 * (lane >> 4)];
                bfr[b] = *reinterpret_cast<const bf16x4 *>(p);
            }
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[a], bfr[b], acc[a][b], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int row = i0 + wi + 16 * a + 4 * (lane >> 4) + t;
                const int col = j0 + wj + 16 * b + (lane & 15);
                if (row >= g.M || col >= g.N) continue;
                float v = acc[a][b][t];
                const long o = (long) row * g.ldc + col;
                if (g.epi == EPI_STORE) {
                    g.C[o] = v;
                } else if (g.epi == EPI_BIAS) {
                    g.C[o] = v + g.bias[col];
                } else if (g.epi == EPI_BIAS_GELU) {
                    v += g.bias[col];
                    g.C[o] = v;
                    g.Cb[o] = toBf16(gelu(v));
                } else {
                    g.C[o] = (v + g.bias[col]) * g.scale[col] + g.shift[col];
                }
            }
}

// ---------------------------------------------------------------- CDNA4 GEMM
// C (M x N) = A (M x K) . B (K x N) on v_mfma_f32_16x16x32_bf16 (gfx950): 256 threads, a 128 x 128
// tile per workgroup in k-steps of 64, each wave a 64 x 64 block of 4 x 4 MFMA tiles (64
// accumulator registers).
// Lane layout of the instruction: A[i][k] from lane i + 16 (k / 8) item k % 8, B[k][j] from lane
// j + 16 (k / 8) item k % 8, D[i][j] in lane j + 16 (i / 4) item i % 4.
// Both operands sit in LDS k-contiguous ([row][k] and [col][k], rows padded by 8 elements), so a
// lane's fragment is one 128-bit LDS read.  Global -> LDS staging moves 128-bit vectors along
// whichever dimension is contiguous in memory: AK / BK say the operand is contiguous along k
// (vectors go to LDS as they are), otherwise along its row / column (a vector of 8 rows at one k
// is scattered into 8 LDS rows).  LDS is double buffered: the next k-step's vectors are loaded
// into registers while the matrix cores work on the current one.  With split-K (gridDim.z > 1)
// each z-slice writes its partial product to a workspace, summed by k_splitk_reduce.
constexpr int G2_BM = 128, G2_BN = 128, G2_BK = 64, G2_PAD = 8, G2_LD = G2_BK + G2_PAD;
constexpr int G2_VEC = G2_BM * G2_BK / 8 / 256;  // 8-element vectors per thread and operand

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4v __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

struct G2Stage {
    u32x4 a[G2_VEC], b[G2_VEC];
};

// Thread tid's v-th vector of a 128 x 64 k-step: KC (contiguous along k): row r, k kk..kk+7;
// otherwise rows r..r+7 at k kk, a wave's 64 lanes taking the 64 k of the same rows, so that
// their scattered LDS writes (one row at a time) fall in 32 different banks.
template <bool KC>
__device__ __forceinline__ void g2Slot(int tid, int v, int &r, int &kk) {
    if (KC) {
        r = (tid >> 3) + 32 * v;
        kk = 8 * (tid & 7);
    } else {
        kk = tid & 63;
        r = 8 * ((tid >> 6) + 4 * v);
    }
}

// One operand's 128 x 64 k-step into registers: G2_VEC 8-element vectors per thread.  sRow /
// sK: strides of the row and k dimensions; rowsTotal / K: bounds.
template <bool KC>
__device__ __forceinline__ void g2Load(const unsigned short *P, long sRow, long sK, int row0, int rowsTotal, int k0,
                                       int K, int tid, u32x4 (&out)[G2_VEC]) {
#pragma unroll
    for (int v = 0; v < G2_VEC; v++) {
        int r, kk;
        g2Slot<KC>(tid, v, r, kk);
        const int gr = row0 + r, gk = k0 + kk;
        const bool full = KC ? (gr < rowsTotal && gk + 7 < K) : (gr + 7 < rowsTotal && gk < K);
        if (full) {
            out[v] = *reinterpret_cast<const u32x4 *>(P + (long) gr * sRow + (long) gk * sK);
        } else {
            unsigned short e[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int qr = KC ? gr : gr + q, qk = KC ? gk + q : gk;
                e[q] = (qr < rowsTotal && qk < K) ? P[(long) qr * sRow + (long) qk * sK] : (unsigned short) 0;
            }
#pragma unroll
            for (int q = 0; q < 4; q++) out[v][q] = (unsigned int) e[2 * q] | ((unsigned int) e[2 * q + 1] << 16);
        }
    }
}

template <bool KC>
__device__ __forceinline__ void g2Store(unsigned short *S, int tid, const u32x4 (&in)[G2_VEC]) {
#pragma unroll
    for (int v = 0; v < G2_VEC; v++) {
        int r, kk;
        g2Slot<KC>(tid, v, r, kk);
        if (KC) {
            *reinterpret_cast<u32x4 *>(S + r * G2_LD + kk) = in[v];
        } else {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                S[(r + 2 * q) * G2_LD + kk] = (unsigned short) (in[v][q] & 0xffffu);
                S[(r + 2 * q + 1) * G2_LD + kk] = (unsigned short) (in[v][q] >> 16);
            }
        }
    }
}

template <bool AK, bool BK>
__global__ void __launch_bounds__(256) k_gemm2(GemmArgs g, float *partial) {
    __shared__ __attribute__((aligned(16))) unsigned short As[2][G2_BM * G2_LD];
    __shared__ __attribute__((aligned(16))) unsigned short Bs[2][G2_BN * G2_LD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = blockIdx.y * G2_BM, j0 = blockIdx.x * G2_BN;
    const int wi = (wave >> 1) * 64, wj = (wave & 1) * 64;
    // this z-slice's k range, a multiple of the k-step
    const int kPer = ((g.K + gridDim.z - 1) / gridDim.z + G2_BK - 1) / G2_BK * G2_BK;
    const int kBeg = blockIdx.z * kPer, kEnd = min(g.K, kBeg + kPer);
    f32x4v acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = f32x4v{0.f, 0.f, 0.f, 0.f};
    // A as rows i (stride sai) by k (stride sak); B as columns j (stride sbj) by k (stride sbk)
    G2Stage st;
    const int nk = kEnd > kBeg ? (kEnd - kBeg + G2_BK - 1) / G2_BK : 0;
    if (nk > 0) {
        g2Load<AK>(g.A, g.sai, g.sak, i0, g.M, kBeg, kEnd, tid, st.a);
        g2Load<BK>(g.B, g.sbj, g.sbk, j0, g.N, kBeg, kEnd, tid, st.b);
        g2Store<AK>(As[0], tid, st.a);
        g2Store<BK>(Bs[0], tid, st.b);
    }
    __syncthreads();
    for (int kt = 0; kt < nk; kt++) {
        const int cur = kt & 1;
        if (kt + 1 < nk) {
            const int kn = kBeg + (kt + 1) * G2_BK;
            g2Load<AK>(g.A, g.sai, g.sak, i0, g.M, kn, kEnd, tid, st.a);
            g2Load<BK>(g.B, g.sbj, g.sbk, j0, g.N, kn, kEnd, tid, st.b);
        }
#pragma unroll
        for (int ks = 0; ks < G2_BK; ks += 32) {
            bf16x8 af[4], bfr[4];
#pragma unroll
            for (int a = 0; a < 4; a++)
                af[a] = *reinterpret_cast<const bf16x8 *>(
                    &As[cur][(wi + 16 * a + (lane & 15)) * G2_LD + ks + 8 * (lane >> 4)]);
#pragma unroll
            for (int b = 0; b < 4; b++)
                bfr[b] = *reinterpret_cast<const bf16x8 *>(
                    &Bs[cur][(wj + 16 * b + (lane & 15)) * G2_LD + ks + 8 * (lane >> 4)]);
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        }
        if (kt + 1 < nk) {
            g2Store<AK>(As[cur ^ 1], tid, st.a);
            g2Store<BK>(Bs[cur ^ 1], tid, st.b);
        }
        __syncthreads();
    }
    float *part = gridDim.z > 1 ? partial + (long) blockIdx.z * g.M * g.N : nullptr;
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int row = i0 + wi + 16 * a + 4 * (lane >> 4) + t;
                const int col = j0 + wj + 16 * b + (lane & 15);
                if (row >= g.M || col >= g.N) continue;
                float v = acc[a][b][t];
                if (part) {
                    part[(long) row * g.N + col] = v;
                    continue;
                }
                const long o = (long) row * g.ldc + col;
                if (g.epi == EPI_STORE) {
                    g.C[o] = v;
                } else if (g.epi == EPI_BIAS) {
                    g.C[o] = v + g.bias[col];
                } else if (g.epi == EPI_BIAS_GELU) {
                    v += g.bias[col];
                    g.C[o] = v;
                    g.Cb[o] = toBf16(gelu(v));
                } else {
                    g.C[o] = (v + g.bias[col]) * g.scale[col] + g.shift[col];
                }
            }
}

// C = sum over the split-K slices of the partial products (EPI_STORE), in slice order.
__global__ void k_splitk_reduce(const float *partial, int slices, int M, int N, float *C, long ldc) {
    const long n = (long) M * N;
    for (long e = blockIdx.x * (long) blockDim.x + threadIdx.x; e < n; e += (long) gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int z = 0; z < slices; z++) s += partial[(long) z * n + e];
        C[(e / N) * ldc + e % N] = s;
    }
}

// dZ = dA . gelu'(Z) (FP32) and its bf16 copy.
__global__ void k_gelu_back(const float *dA, const float *Z, float *dZ, unsigned short *dZb, long n) {
    for (long t = blockIdx.x * (long) blockDim.x + threadIdx.x; t < n; t += (long) gridDim.x * blockDim.x) {
        const float v = dA[t] * geluGrad(Z[t]);
        dZ[t] = v;
        dZb[t] = toBf16(v);
    }
}

// db[j] = sum_i dZ[i][j] (rows in order, one thread per column).
// 256 threads per 64 columns: thread (c, q) sums rows q, q + 4, ... of column c, then the four
// partials of a column are added in q order (the same order on every run).
__global__ void __launch_bounds__(256) k_colsum(const float *dZ, int rows, int cols, float *db) {
    __shared__ float part[4][64];
    const int c = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + c;
    float s = 0.f;
    if (j < cols)
        for (int i = q; i < rows; i += 4) s += dZ[(long) i * cols + j];
    part[q][c] = s;
    __syncthreads();
    if (q == 0 && j < cols) db[j] = ((part[0][c] + part[1][c]) + part[2][c]) + part[3][c];
}

// MSE on standardized outputs: dY = 2 (y - t) / count and per-block partial sums of (y - t)^2.
__global__ void k_mse(const float *Y, const float *Tg, float *dY, unsigned short *dYb, long n, float inv,
                      double *partial) {
    __shared__ double red[256];
    double acc = 0.;
    for (long t = blockIdx.x * (long) blockDim.x + threadIdx.x; t < n; t += (long) gridDim.x * blockDim.x) {
        const float d = Y[t] - Tg[t];
        acc += (double) d * d;
        const float g = 2.f * d * inv;
        dY[t] = g;
        dYb[t] = toBf16(g);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Adam (bias-corrected) on the FP32 masters, refreshing the bf16 copies.
__global__ void k_adam(float *w, unsigned short *wb, const float *g, float *m, float *v, long n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
    for (long t = blockIdx.x * (long) blockDim.x + threadIdx.x; t < n; t += (long) gridDim.x * blockDim.x) {
        const float gt = g[t];
        const float mt = b1 * m[t] + (1.f - b1) * gt;
        const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
        m[t] = mt;
        v[t] = vt;
        const float upd = lr * (mt / c1) / (sqrtf(vt / c2) + eps);
        const float wn = w[t] - upd;
        w[t] = wn;
        if (wb) wb[t] = toBf16(wn);
    }
}

__global__ void k_to_bf16(const float *in, unsigned short *out, long n) {
    for (long t = blockIdx.x * (long) blockDim.x + threadIdx.x; t < n; t += (long) gridDim.x * blockDim.x)
        out[t] = toBf16(in[t]);
}

// Minibatch gather: rows idx[r] of X (FP32 -> bf16) and of T (FP32).
__global__ void k_gather(const float *X, int in, const float *Tg, long out, const int *idx, int rows,
                         unsigned short *Xb, float *Tb) {
    const long totalX = (long) rows * in, totalT = (long) rows * out;
    for (long t = blockIdx.x * (long) blockDim.x + threadIdx.x; t < totalX + totalT;
         t += (long) gridDim.x * blockDim.x) {
        if (t < totalX) {
            const long r = t / in, c = t % in;
            Xb[t] = toBf16(X[(long) idx[r] * in + c]);
        } else {
            const long u = t - totalX;
            const long r = u / out, c = u % out;
            Tb[u] = Tg[(long) idx[r] * out + c];
        }
    }
}

// Per-column mean and standard deviation of an N x cols matrix (two passes, rows in order).
__global__ void k_col_stats(const float *Tg, long rows, long cols, float *mean, float *stdv) {
    const long j = blockIdx.x * (long) blockDim.x + threadIdx.x;
    if (j >= cols) return;
    double s = 0.;
    for (long i = 0; i < rows; i++) s += Tg[i * cols + j];
    const double mu = s / (double) rows;
    double q = 0.;
    for (long i = 0; i < rows; i++) {
        const double d = Tg[i * cols + j] - mu;
        q += d * d;
    }
    const double sd = sqrt(q / (double) rows);
    mean[j] = (float) mu;
    stdv[j] = (float) (sd > 1e-6 * (fabs(mu) + 1e-12) && sd > 1e-12 ? sd : 1.0);
}

__global__ void k_standardize(float *Tg, long rows, long cols, const float *mean, const float *stdv) {
    const long total = rows * cols;
    for (long t = blockIdx.x * (long) blockDim.x + threadIdx.x; t < total; t += (long) gridDim.x * blockDim.x) {
        const long j = t % cols;
        Tg[t] = (Tg[t] - mean[j]) / stdv[j];
    }
}

}  // namespace mlp
}  // namespace spk

// ====================================================================== host side (C ABI)
#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/spock_amd.h"
#include "common.hpp"

using spk::check;
using spk::DevBuf;
using spk::fail;
namespace M = spk::mlp;

struct spk_mlp {
    spk_ctx *ctx = nullptr;  // holds a reference
    explicit spk_mlp(spk_ctx *c) : ctx(spk::retain(c)) {}
    ~spk_mlp() {
        if (ctx && ctx->stream) (void) hipStreamSynchronize(ctx->stream);
        spk::releaseCtx(ctx);
    }
    spk_mlp(const spk_mlp &) = delete;
    spk_mlp &operator=(const spk_mlp &) = delete;
    std::vector<int> dims;          // in, hidden..., out
    std::vector<long> wOff, bOff;   // parameter offsets (weights [d_l][d_l+1] row-major, then bias)
    long nParams = 0;
    DevBuf w, wb, g, m, v;          // FP32 masters, bf16 copies, gradients, Adam moments
    long step = 0;
    DevBuf inMean, inStd, outMean, outStd;
    std::vector<float> hInMean, hInStd;
    int cap = 0;                    // rows the workspace holds
    DevBuf X, Y, dY, dYb, dA, dZ, dZb, Tb, idx, partial, Xf;
    DevBuf splitWs;                 // split-K partial products of the GEMM
    std::vector<std::unique_ptr<DevBuf>> Z, A;
    long nTrain = 0;
    DevBuf trX, trT;
    int layers() const { return (int) dims.size() - 1; }
    int in() const { return dims.front(); }
    int out() const { return dims.back(); }
};

namespace {

dim3 grid1(long n) { return dim3((unsigned) std::min<long>(4096, (n + 255) / 256)); }

// 128-bit vectors along an operand's contiguous dimension need that dimension's stride to be 1,
// the other stride a multiple of 8 elements and a 16-byte aligned base.
bool vectorizable(const unsigned short *P, long sRow, long sK) {
    if (((uintptr_t) P & 15) != 0) return false;
    if (sK == 1) return sRow % 8 == 0;
    if (sRow == 1) return sK % 8 == 0;
    return false;
}

void gemm(spk_mlp *m, const unsigned short *A, long sai, long sak, const unsigned short *B, long sbk, long sbj,
          int Mr, int Nc, int K, float *C, long ldc, unsigned short *Cb, const float *bias, int epi,
          const float *scale = nullptr, const float *shift = nullptr) {
    M::GemmArgs a;
    a.A = A; a.B = B; a.sai = sai; a.sak = sak; a.sbk = sbk; a.sbj = sbj;
    a.M = Mr; a.N = Nc; a.K = K; a.C = C; a.ldc = ldc; a.Cb = Cb; a.bias = bias; a.scale = scale; a.shift = shift;
    a.epi = epi;
    if (!vectorizable(A, sai, sak) || !vectorizable(B, sbj, sbk)) {  // e.g. the 50-wide input layer
        const dim3 grid((unsigned) ((Nc + M::TN - 1) / M::TN), (unsigned) ((Mr + M::TM - 1) / M::TM));
        hipLaunchKernelGGL(M::k_gemm, grid, dim3(256), 0, m->ctx->stream, a);
        return;
    }
    const int tiles = ((Nc + M::G2_BN - 1) / M::G2_BN) * ((Mr + M::G2_BM - 1) / M::G2_BM);
    // split K when the output has too few tiles to fill the device and K is long (the output
    // layer's dA = dY W': 8 tiles, K = 52,360): up to 256 workgroups, 16 k-steps each at least
    int split = 1;
    if (epi == M::EPI_STORE && tiles < 128)
        split = std::max(1, std::min(256 / tiles, K / (16 * M::G2_BK)));
    float *part = nullptr;
    if (split > 1) {
        const size_t need = (size_t) split * Mr * Nc * sizeof(float);
        if (m->splitWs.bytes < need) m->splitWs.alloc(need);
        part = m->splitWs.as<float>();
    }
    const dim3 grid((unsigned) ((Nc + M::G2_BN - 1) / M::G2_BN), (unsigned) ((Mr + M::G2_BM - 1) / M::G2_BM),
                    (unsigned) split);
    const bool ak = sak == 1, bk = sbk == 1;
    if (ak && bk) hipLaunchKernelGGL((M::k_gemm2<true, true>), grid, dim3(256), 0, m->ctx->stream, a, part);
    else if (ak) hipLaunchKernelGGL((M::k_gemm2<true, false>), grid, dim3(256), 0, m->ctx->stream, a, part);
    else if (bk) hipLaunchKernelGGL((M::k_gemm2<false, true>), grid, dim3(256), 0, m->ctx->stream, a, part);
    else hipLaunchKernelGGL((M::k_gemm2<false, false>), grid, dim3(256), 0, m->ctx->stream, a, part);
    if (split > 1) {
        const long n = (long) Mr * Nc;
        hipLaunchKernelGGL(M::k_splitk_reduce, dim3((unsigned) std::min<long>(2048, (n + 255) / 256)), dim3(256), 0,
                           m->ctx->stream, part, split, Mr, Nc, C, ldc);
    }
}

void ensureCap(spk_mlp *m, int rows) {
    if (rows <= m->cap) return;
    const int h = *std::max_element(m->dims.begin() + 1, m->dims.end() - 1);
    const long out = m->out(), in = m->in();
    m->X.alloc((size_t) rows * in * 2);
    m->Xf.alloc((size_t) rows * in * 4);
    m->Y.alloc((size_t) rows * out * 4);
    m->dY.alloc((size_t) rows * out * 4);
    m->dYb.alloc((size_t) rows * out * 2);
    m->Tb.alloc((size_t) rows * out * 4);
    m->dA.alloc((size_t) rows * h * 4);
    m->dZ.alloc((size_t) rows * h * 4);
    m->dZb.alloc((size_t) rows * h * 2);
    m->idx.alloc((size_t) rows * 4);
    m->partial.alloc(4096 * 8);
    m->Z.clear();
    m->A.clear();
    for (int l = 0; l + 1 < m->layers(); l++) {
        m->Z.push_back(std::make_unique<DevBuf>());
        m->A.push_back(std::make_unique<DevBuf>());
        m->Z.back()->alloc((size_t) rows * m->dims[l + 1] * 4);
        m->A.back()->alloc((size_t) rows * m->dims[l + 1] * 2);
    }
    m->cap = rows;
}

const unsigned short *wbf(spk_mlp *m, int l) { return m->wb.as<unsigned short>() + m->wOff[l]; }
const float *bias(spk_mlp *m, int l) { return m->w.as<float>() + m->bOff[l]; }

// Forward of `rows` rows whose bf16 inputs are in m->X; Y gets the standardized outputs, or the
// denormalized iterate with `denorm`.
void forward(spk_mlp *m, int rows, bool denorm) {
    const int L = m->layers();
    const unsigned short *prev = m->X.as<unsigned short>();
    long dprev = m->in();
    for (int l = 0; l < L; l++) {
        const int N = m->dims[l + 1];
        if (l + 1 < L) {
            gemm(m, prev, dprev, 1, wbf(m, l), N, 1, rows, N, (int) dprev, m->Z[l]->as<float>(), N,
                 m->A[l]->as<unsigned short>(), bias(m, l), M::EPI_BIAS_GELU);
            prev = m->A[l]->as<unsigned short>();
            dprev = N;
        } else if (denorm) {
            gemm(m, prev, dprev, 1, wbf(m, l), N, 1, rows, N, (int) dprev, m->Y.as<float>(), N, nullptr, bias(m, l),
                 M::EPI_BIAS_AFFINE, m->outStd.as<float>(), m->outMean.as<float>());
        } else {
            gemm(m, prev, dprev, 1, wbf(m, l), N, 1, rows, N, (int) dprev, m->Y.as<float>(), N, nullptr, bias(m, l),
                 M::EPI_BIAS);
        }
    }
    check(hipGetLastError(), "mlp forward");
}

// Gradients of the loss whose output gradient is in m->dY / m->dYb (rows of the last forward).
void backward(spk_mlp *m, int rows) {
    const int L = m->layers();
    hipStream_t st = m->ctx->stream;
    float *g = m->g.as<float>();
    const unsigned short *dCur = m->dYb.as<unsigned short>();
    const float *dCurF = m->dY.as<float>();
    for (int l = L - 1; l >= 0; l--) {
        const int Nl = m->dims[l + 1], Kl = m->dims[l];
        const unsigned short *Ain = l == 0 ? m->X.as<unsigned short>() : m->A[l - 1]->as<unsigned short>();
        // dW_l = A_{l-1}' dZ_l  (Kl x Nl), db_l = colsum dZ_l
        gemm(m, Ain, 1, Kl, dCur, Nl, 1, Kl, Nl, rows, g + m->wOff[l], Nl, nullptr, nullptr, M::EPI_STORE);
        hipLaunchKernelGGL(M::k_colsum, dim3((Nl + 63) / 64), dim3(256), 0, st, dCurF, rows, Nl, g + m->bOff[l]);
        if (l == 0) break;
        // dA_{l-1} = dZ_l W_l'  (rows x Kl), then dZ_{l-1} = dA . gelu'(Z_{l-1})
        gemm(m, dCur, Nl, 1, wbf(m, l), 1, Nl, rows, Kl, Nl, m->dA.as<float>(), Kl, nullptr, nullptr, M::EPI_STORE);
        const long n = (long) rows * Kl;
        hipLaunchKernelGGL(M::k_gelu_back, grid1(n), dim3(256), 0, st, m->dA.as<float>(), m->Z[l - 1]->as<float>(),
                           m->dZ.as<float>(), m->dZb.as<unsigned short>(), n);
        dCur = m->dZb.as<unsigned short>();
        dCurF = m->dZ.as<float>();
    }
    check(hipGetLastError(), "mlp backward");
}

double mseLoss(spk_mlp *m, int rows, const float *target) {
    const long n = (long) rows * m->out();
    const int blocks = (int) std::min<long>(1024, (n + 255) / 256);
    hipLaunchKernelGGL(M::k_mse, dim3(blocks), dim3(256), 0, m->ctx->stream, m->Y.as<float>(), target, m->dY.as<float>(),
                       m->dYb.as<unsigned short>(), n, 1.f / (float) n, m->partial.as<double>());
    check(hipGetLastError(), "mlp loss");
    std::vector<double> part(blocks);
    check(hipMemcpyAsync(part.data(), m->partial.p, blocks * sizeof(double), hipMemcpyDeviceToHost, m->ctx->stream),
          "loss");
    check(hipStreamSynchronize(m->ctx->stream), "sync");
    double s = 0.;
    for (double v : part) s += v;
    return s / (double) n;
}

void adam(spk_mlp *m, float lr, float b1, float b2, float eps) {
    m->step++;
    const float c1 = 1.f - std::pow(b1, (float) m->step), c2 = 1.f - std::pow(b2, (float) m->step);
    hipLaunchKernelGGL(M::k_adam, grid1(m->nParams), dim3(256), 0, m->ctx->stream, m->w.as<float>(),
                       m->wb.as<unsigned short>(), m->g.as<float>(), m->m.as<float>(), m->v.as<float>(), m->nParams, lr,
                       b1, b2, eps, c1, c2);
    check(hipGetLastError(), "mlp adam");
}

// Host rows -> normalized FP32 -> bf16 workspace inputs.
template <typename S>
void uploadInputs(spk_mlp *m, const S *x, int rows) {
    ensureCap(m, rows);
    const int in = m->in();
    std::vector<float> xs((size_t) rows * in);
    for (int r = 0; r < rows; r++)
        for (int c = 0; c < in; c++)
            xs[(size_t) r * in + c] = (float) ((x[(size_t) r * in + c] - m->hInMean[c]) / m->hInStd[c]);
    check(hipMemcpyAsync(m->Xf.p, xs.data(), xs.size() * 4, hipMemcpyHostToDevice, m->ctx->stream), "inputs");
    hipLaunchKernelGGL(M::k_to_bf16, grid1((long) xs.size()), dim3(256), 0, m->ctx->stream, m->Xf.as<float>(),
                       m->X.as<unsigned short>(), (long) xs.size());
    check(hipStreamSynchronize(m->ctx->stream), "sync");
}

void refreshBf16(spk_mlp *m) {
    hipLaunchKernelGGL(M::k_to_bf16, grid1(m->nParams), dim3(256), 0, m->ctx->stream, m->w.as<float>(),
                       m->wb.as<unsigned short>(), m->nParams);
    check(hipGetLastError(), "bf16 copy");
}

}  // namespace

extern "C" {

int spk_mlp_create(spk_ctx *ctx, int in, int hidden, int layers, int out, unsigned long long seed, spk_mlp **res) {
    try {
        if (in <= 0 || hidden <= 0 || layers < 0 || out <= 0) throw std::runtime_error("bad MLP dimensions");
        if (!ctx) throw std::runtime_error("null context");
        auto m = std::make_unique<spk_mlp>(ctx);
        check(hipSetDevice(ctx->device), "hipSetDevice");
        m->dims.push_back(in);
        for (int l = 0; l < layers; l++) m->dims.push_back(hidden);
        m->dims.push_back(out);
        long off = 0;
        for (int l = 0; l < m->layers(); l++) {
            m->wOff.push_back(off);
            off += (long) m->dims[l] * m->dims[l + 1];
            m->bOff.push_back(off);
            off += m->dims[l + 1];
        }
        m->nParams = off;
        // He-uniform initialisation from a seeded generator; zero biases
        std::vector<float> p(off, 0.f);
        std::mt19937_64 gen(seed);
        for (int l = 0; l < m->layers(); l++) {
            const float a = std::sqrt(6.f / (float) m->dims[l]) * (l + 1 == m->layers() ? 0.5f : 1.f);
            std::uniform_real_distribution<float> U(-a, a);
            for (long k = 0; k < (long) m->dims[l] * m->dims[l + 1]; k++) p[m->wOff[l] + k] = U(gen);
        }
        m->w.alloc(off * 4);
        m->wb.alloc(off * 2);
        m->g.alloc(off * 4);
        m->m.alloc(off * 4);
        m->v.alloc(off * 4);
        check(hipMemcpy(m->w.p, p.data(), off * 4, hipMemcpyHostToDevice), "params");
        check(hipMemset(m->g.p, 0, off * 4), "zero");
        check(hipMemset(m->m.p, 0, off * 4), "zero");
        check(hipMemset(m->v.p, 0, off * 4), "zero");
        refreshBf16(m.get());
        m->hInMean.assign(in, 0.f);
        m->hInStd.assign(in, 1.f);
        std::vector<float> zo(out, 0.f), on(out, 1.f);
        m->inMean.alloc(in * 4);
        m->inStd.alloc(in * 4);
        m->outMean.alloc(out * 4);
        m->outStd.alloc(out * 4);
        check(hipMemcpy(m->inMean.p, m->hInMean.data(), in * 4, hipMemcpyHostToDevice), "norm");
        check(hipMemcpy(m->inStd.p, m->hInStd.data(), in * 4, hipMemcpyHostToDevice), "norm");
        check(hipMemcpy(m->outMean.p, zo.data(), out * 4, hipMemcpyHostToDevice), "norm");
        check(hipMemcpy(m->outStd.p, on.data(), out * 4, hipMemcpyHostToDevice), "norm");
        check(hipStreamSynchronize(ctx->stream), "sync");
        *res = m.release();
        return 0;
    } catch (const std::exception &e) {
        return fail(ctx, e);
    }
}

int spk_mlp_destroy(spk_mlp *m) {
    delete m;
    return 0;
}

long spk_mlp_num_params(const spk_mlp *m) { return m ? m->nParams : -1; }

int spk_mlp_get_params(spk_mlp *m, float *p) {
    try {
        check(hipStreamSynchronize(m->ctx->stream), "sync");
        check(hipMemcpy(p, m->w.p, m->nParams * 4, hipMemcpyDeviceToHost), "params");
        return 0;
    } catch (const std::exception &e) {
        return fail(m->ctx, e);
    }
}

int spk_mlp_set_params(spk_mlp *m, const float *p) {
    try {
        check(hipStreamSynchronize(m->ctx->stream), "sync");
        check(hipMemcpy(m->w.p, p, m->nParams * 4, hipMemcpyHostToDevice), "params");
        refreshBf16(m);
        check(hipStreamSynchronize(m->ctx->stream), "sync");
        return 0;
    } catch (const std::exception &e) {
        return fail(m->ctx, e);
    }
}

int spk_mlp_get_grads(spk_mlp *m, float *p) {
    try {
        check(hipStreamSynchronize(m->ctx->stream), "sync");
        check(hipMemcpy(p, m->g.p, m->nParams * 4, hipMemcpyDeviceToHost), "grads");
        return 0;
    } catch (const std::exception &e) {
        return fail(m->ctx, e);
    }
}

int spk_mlp_get_normalization(spk_mlp *m, float *inMean, float *inStd, float *outMean, float *outStd) {
    try {
        check(hipStreamSynchronize(m->ctx->stream), "sync");
        std::copy(m->hInMean.begin(), m->hInMean.end(), inMean);
        std::copy(m->hInStd.begin(), m->hInStd.end(), inStd);
        check(hipMemcpy(outMean, m->outMean.p, m->out() * 4, hipMemcpyDeviceToHost), "norm");
        check(hipMemcpy(outStd, m->outStd.p, m->out() * 4, hipMemcpyDeviceToHost), "norm");
        return 0;
    } catch (const std::exception &e) {
        return fail(m->ctx, e);
    }
}

int spk_mlp_set_normalization(spk_mlp *m, const float *inMean, const float *inStd, const float *outMean,
                              const float *outStd) {
    try {
        check(hipStreamSynchronize(m->ctx->stream), "sync");
        m->hInMean.assign(inMean, inMean + m->in());
        m->hInStd.assign(inStd, inStd + m->in());
        check(hipMemcpy(m->inMean.p, inMean, m->in() * 4, hipMemcpyHostToDevice), "norm");
        check(hipMemcpy(m->inStd.p, inStd, m->in() * 4, hipMemcpyHostToDevice), "norm");
        check(hipMemcpy(m->outMean.p, outMean, m->out() * 4, hipMemcpyHostToDevice), "norm");
        check(hipMemcpy(m->outStd.p, outStd, m->out() * 4, hipMemcpyHostToDevice), "norm");
        return 0;
    } catch (const std::exception &e) {
        return fail(m->ctx, e);
    }
}

int spk_mlp_forward(spk_mlp *m, const float *x, int rows, float *y) {
    try {
        if (rows <= 0) throw std::runtime_error("rows must be positive");
        uploadInputs(m, x, rows);
        forward(m, rows, false);
        if (y) check(hipMemcpyAsync(y, m->Y.p, (size_t) rows * m->out() * 4, hipMemcpyDeviceToHost, m->ctx->stream),
                     "outputs");
        check(hipStreamSynchronize(m->ctx->stream), "sync");
        return 0;
    } catch (const std::exception &e) {
        return fail(m->ctx, e);
    }
}

int spk_mlp_loss_backward(spk_mlp *m, const float *target, int rows, double *loss) {
    try {
        if (rows <= 0 || rows > m->cap) throw std::runtime_error("rows must match the last forward");
        check(hipMemcpyAsync(m->Tb.p, target, (size_t) rows * m->out() * 4, hipMemcpyHostToDevice, m->ctx->stream),
              "targets");
        const double l = mseLoss(m, rows, m->Tb.as<float>());
        backward(m, rows);
        check(hipStreamSynchronize(m->ctx->stream), "sync");
        if (loss) *loss = l;
        return 0;
    } catch (const std::exception &e) {
        return fail(m->ctx, e);
    }
}

int spk_mlp_adam(spk_mlp *m, float lr, float b1, float b2, float eps) {
    try {
        adam(m, lr, b1, b2, eps);
        check(hipStreamSynchronize(m->ctx->stream), "sync");
        return 0;
    } catch (const std::exception &e) {
        return fail(m->ctx, e);
    }
}

// The training set: targets already in trT (N rows), inputs x0 (host, N rows).
static void finishDataset(spk_mlp *m, int N, const double *x0);

int spk_mlp_dataset(spk_mlp *m, spk_batch *solved, const double *x0) {
    try {
        const int N = spk_batch_count(solved);
        const long out = m->out();
        m->trT.alloc((size_t) N * out * 4);
        if (spk_batch_get_iterates_device(solved, m->trT.as<float>(), out)) throw std::runtime_error(m->ctx->err);
        finishDataset(m, N, x0);
        return 0;
    } catch (const std::exception &e) {
        return fail(m->ctx, e);
    }
}

int spk_mlp_dataset_lanes(spk_mlp *m, spk_batch *solved, int count, const int *lanes, const double *x0) {
    try {
        if (count < 1) throw std::runtime_error("an empty training set");
        const long out = m->out();
        m->trT.alloc((size_t) count * out * 4);
        if (spk_batch_get_lanes_device(solved, count, lanes, m->trT.as<float>(), out)) throw std::runtime_error(m->ctx->err);
        finishDataset(m, count, x0);
        return 0;
    } catch (const std::exception &e) {
        return fail(m->ctx, e);
    }
}

static void finishDataset(spk_mlp *m, int N, const double *x0) {
    {
        const int in = m->in();
        const long out = m->out();
        m->nTrain = N;
        hipStream_t st = m->ctx->stream;
        hipLaunchKernelGGL(M::k_col_stats, dim3((unsigned) ((out + 255) / 256)), dim3(256), 0, st, m->trT.as<float>(),
                           (long) N, out, m->outMean.as<float>(), m->outStd.as<float>());
        hipLaunchKernelGGL(M::k_standardize, grid1((long) N * out), dim3(256), 0, st, m->trT.as<float>(), (long) N,
                           out, m->outMean.as<float>(), m->outStd.as<float>());
        check(hipGetLastError(), "dataset");
        // input normalization from the training inputs
        for (int c = 0; c < in; c++) {
            double s = 0., q = 0.;
            for (int r = 0; r < N; r++) s += x0[(size_t) r * in + c];
            const double mu = s / N;
            for (int r = 0; r < N; r++) q += (x0[(size_t) r * in + c] - mu) * (x0[(size_t) r * in + c] - mu);
            const double sd = std::sqrt(q / N);
            m->hInMean[c] = (float) mu;
            m->hInStd[c] = (float) (sd > 1e-12 ? sd : 1.);
        }
        check(hipMemcpy(m->inMean.p, m->hInMean.data(), in * 4, hipMemcpyHostToDevice), "norm");
        check(hipMemcpy(m->inStd.p, m->hInStd.data(), in * 4, hipMemcpyHostToDevice), "norm");
        std::vector<float> xs((size_t) N * in);
        for (int r = 0; r < N; r++)
            for (int c = 0; c < in; c++)
                xs[(size_t) r * in + c] = (float) ((x0[(size_t) r * in + c] - m->hInMean[c]) / m->hInStd[c]);
        m->trX.alloc(xs.size() * 4);
        check(hipMemcpy(m->trX.p, xs.data(), xs.size() * 4, hipMemcpyHostToDevice), "inputs");
        check(hipStreamSynchronize(st), "sync");
    }
}

int spk_mlp_train(spk_mlp *m, int epochs, int minibatch, float lr, unsigned long long seed, double *epochLoss) {
    try {
        if (m->nTrain <= 0) throw std::runtime_error("no training set (spk_mlp_dataset)");
        if (minibatch <= 0) throw std::runtime_error("minibatch must be positive");
        const int mb = (int) std::min<long>(minibatch, m->nTrain);
        ensureCap(m, mb);
        std::vector<int> perm(m->nTrain);
        std::iota(perm.begin(), perm.end(), 0);
        std::mt19937_64 gen(seed);
        hipStream_t st = m->ctx->stream;
        for (int ep = 0; ep < epochs; ep++) {
            std::shuffle(perm.begin(), perm.end(), gen);
            double acc = 0.;
            int nb = 0;
            for (long s0 = 0; s0 + mb <= m->nTrain; s0 += mb) {
                check(hipMemcpyAsync(m->idx.p, perm.data() + s0, mb * 4, hipMemcpyHostToDevice, st), "indices");
                hipLaunchKernelGGL(M::k_gather, grid1((long) mb * (m->in() + m->out())), dim3(256), 0, st,
                                   m->trX.as<float>(), m->in(), m->trT.as<float>(), (long) m->out(), m->idx.as<int>(),
                                   mb, m->X.as<unsigned short>(), m->Tb.as<float>());
                forward(m, mb, false);
                acc += mseLoss(m, mb, m->Tb.as<float>());
                backward(m, mb);
                adam(m, lr, 0.9f, 0.999f, 1e-8f);
                nb++;
            }
            check(hipStreamSynchronize(st), "sync");
            if (epochLoss) epochLoss[ep] = nb ? acc / nb : 0.;
        }
        return 0;
    } catch (const std::exception &e) {
        return fail(m->ctx, e);
    }
}

int spk_mlp_predict(spk_mlp *m, const double *x0, int rows, spk_batch *dst) {
    try {
        if (rows != spk_batch_count(dst)) throw std::runtime_error("rows must equal the batch count");
        uploadInputs(m, x0, rows);
        forward(m, rows, true);
        check(hipStreamSynchronize(m->ctx->stream), "sync");
        if (spk_batch_set_iterates_device(dst, m->Y.as<float>(), m->out())) throw std::runtime_error(m->ctx->err);
        return 0;
    } catch (const std::exception &e) {
        return fail(m->ctx, e);
    }
}

int spk_mlp_predict_lanes(spk_mlp *m, const double *x0, int rows, spk_batch *dst, const int *lanes) {
    try {
        uploadInputs(m, x0, rows);
        forward(m, rows, true);
        check(hipStreamSynchronize(m->ctx->stream), "sync");
        if (spk_batch_refill_rows(dst, rows, lanes, x0, m->Y.as<float>(), m->out())) throw std::runtime_error(m->ctx->err);
        return 0;
    } catch (const std::exception &e) {
        return fail(m->ctx, e);
    }
}

}  // extern "C"
