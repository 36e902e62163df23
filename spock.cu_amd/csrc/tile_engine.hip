// Tile engine: many instances of ONE problem, 16 per workgroup, the per-node products on the
// matrix cores.
//
// Layout (LaneInst with width 16): element e of instance b of every state region is at
// ((b / 16) * rows + e) * 16 + b % 16, so a workgroup's state is one contiguous slab per region
// and four consecutive rows of its 16 instances are 64 consecutive words.
//
// Lane l of a wave serves instance j = l % 16 and row group g = l / 16.  A node's product of a
// factor M (r x c, column major) with the 16 instances' vectors X (c x 16) is a GEMM on
// v_mfma_{f32,f64}_16x16x4: the A operand is a 16 x 4 block of M (lane l holds M[r0 + l%16][k0 + l/16]),
// the B operand four rows of X (lane l holds X[k0 + l/16][l%16], its own instance's entry) and each
// lane receives four rows of its instance (RowMap: rows 4g .. 4g + 3 of each 16-row tile in FP32,
// rows g, g + 4, g + 8, g + 12 in FP64).  The elementwise parts of the operators use the same row
// ownership, so a lane only ever reads back what it wrote itself.  Intermediate vectors a wave
// needs again as a B operand (R~^-1 e, d, u of the tree sweeps) pass through its own LDS slice.
// Per-instance sums are per-lane partials, combined across the four row groups by cross-lane
// shuffles (the same value in all four lanes) and across waves through LDS in wave order.
//
// Operator semantics are those of the other engines and of the reference (src/operator.cuh:41-82,
// src/dynamics.cuh:74-174, src/cache.cuh:650-801); the SuperMann control is the shared lockstep
// solve (lockstep.hpp).  As in the lane engine the dynamics sweep applies the precomputed
// (L L')^-1 and the Anderson least squares is the FP64 Gram solve.
#include "engine.hpp"
#include "lockstep.hpp"

#include <hip/hip_runtime.h>
#include <math.h>

// Launch bound of this compilation: 512 threads (two waves per SIMD, up to 256 registers each) or 1024
// (four waves per SIMD, 128 registers each: more waves to hide memory latency, more spills).  The
// library is built once per bound; the 1024 build's kernels carry the suffix _x1024.
#ifndef SPK_TILE_MAX_THREADS
#define SPK_TILE_MAX_THREADS 512
#endif
#if SPK_TILE_MAX_THREADS == 512
#define SPK_TILE_KERNEL(name) name
#else
#define SPK_TILE_KERNEL_CAT2(a, b) a##_x##b
#define SPK_TILE_KERNEL_CAT(a, b) SPK_TILE_KERNEL_CAT2(a, b)
#define SPK_TILE_KERNEL(name) SPK_TILE_KERNEL_CAT(name, SPK_TILE_MAX_THREADS)
#endif
#ifndef SPK_TILE_KB
#define SPK_TILE_KB 2  // k-steps of 4 columns whose operands are loaded together (2 measured
                       // 2.5% faster than 4 on the 229-node problem; 8 spills)
#endif

namespace spk {
namespace tile {
using namespace ls;

constexpr int W = 16;                                // instances per workgroup
constexpr int MAXW = SPK_TILE_MAX_THREADS / 64;      // waves per workgroup
constexpr int STG = 32 * W;                          // words of one LDS staging slice (32 rows)

template <typename T>
struct AccT;
template <>
struct AccT<float> {
    typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct AccT<double> {
    typedef double type __attribute__((ext_vector_type(4)));
};
template <typename T>
using Acc = typename AccT<T>::type;

__device__ __forceinline__ Acc<float> mfma(float a, float b, Acc<float> c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ Acc<double> mfma(double a, double b, Acc<double> c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

struct Ctx {
    int lane, g, j, wave, nw;
    int gl;       // this lane's instance
    bool writer;  // the row group 0 lane writes the instance's records
};

__device__ __forceinline__ Ctx ctx() {
    Ctx c;
    c.lane = threadIdx.x & 63;
    c.g = c.lane >> 4;
    c.j = c.lane & 15;
    c.wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    c.nw = blockDim.x >> 6;
    c.gl = blockIdx.x * W + c.j;
    c.writer = c.g == 0;
    return c;
}

__device__ __forceinline__ int rowGroup() { return (threadIdx.x >> 4) & 3; }
__device__ __forceinline__ int instLane() { return threadIdx.x & 15; }

// View of a region: v[e] is element e of this lane's instance (per-lane e), a global-memory access.
template <typename T>
struct TV {
    T *p;   // this workgroup's slab of the region
    int j;  // instance within the workgroup
    __device__ __forceinline__ SPK_GLOBAL T &operator[](long e) const { return ((SPK_GLOBAL T *) p)[e * W + j]; }
    __device__ __forceinline__ TV shift(long e0) const { return TV{p + e0 * W, j}; }
};

// A wave's LDS staging slices.
#define SPK_LDS __attribute__((address_space(3)))
template <typename T>
using SP = SPK_LDS T *;

template <typename T>
__device__ __forceinline__ TV<T> tview(T *region, long rows, int j) {
    return TV<T>{region + (long) blockIdx.x * rows * W, j};
}

// Sum (max) over the four lanes of one instance; every one of them receives the same value.
template <typename T>
__device__ __forceinline__ T sumG(T v) {
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}
template <typename T>
__device__ __forceinline__ T maxG(T v) {
    v = fmax(v, __shfl_xor(v, 16));
    v = fmax(v, __shfl_xor(v, 32));
    return v;
}

// Which lane receives which row of a product: v_mfma_f32_16x16x4f32 gives lane (g, j) rows
// 16 t + 4 g + q of tile t in its four accumulator entries q, v_mfma_f64_16x16x4f64 rows
// 16 t + 4 q + g.  Lane g owns those rows of every node block, (r / 4) % 4 == g in FP32 and
// r % 4 == g in FP64.
template <typename T>
struct RowMap;
template <>
struct RowMap<float> {
    __device__ static __forceinline__ int row(int t, int g, int q) { return 16 * t + 4 * g + q; }
};
template <>
struct RowMap<double> {
    __device__ static __forceinline__ int row(int t, int g, int q) { return 16 * t + 4 * q + g; }
};

// The rows r < len this lane owns, f(r) for each.
template <typename T, typename F>
__device__ __forceinline__ void forRows(int len, F f) {
    const int g = rowGroup();
    for (int t = 0; 16 * t < len; t++) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int r = RowMap<T>::row(t, g, q);
            if (r < len) f(r);
        }
    }
}

template <int RT, typename T>
__device__ __forceinline__ void zeroAcc(Acc<T> (&a)[RT]) {
#pragma unroll
    for (int t = 0; t < RT; t++) a[t] = Acc<T>{T(0), T(0), T(0), T(0)};
}

// acc[t] += M[16 t + i][c] x(c) summed over c < cols for the 16 t + i < rows of RT row tiles:
// M is column major with leading dimension ld; x(c) is this lane's instance's entry c.  Every load
// is clamped into the matrix and the vector, so none is conditional (a conditional load would be
// waited for on its own): rows past the end compute outputs nobody reads, and columns past the end
// enter with x multiplied by 0 against finite matrix entries.  The operands of KB k-steps are
// loaded together before their products, so a wave keeps KB (RT + 1) loads in flight.
template <int RT, typename T, typename XF>
__device__ __forceinline__ void mma(Acc<T> (&acc)[RT], CP<T> M, int rows, int ld, int cols, XF x) {
    constexpr int KB = SPK_TILE_KB;
    const int i = instLane(), g = rowGroup();
    int ri[RT];
#pragma unroll
    for (int t = 0; t < RT; t++) ri[t] = min(16 * t + i, rows - 1);
    for (int k0 = 0; k0 < cols; k0 += 4 * KB) {
        T a[KB][RT], b[KB];
#pragma unroll
        for (int s = 0; s < KB; s++) {
            const int c = k0 + 4 * s + g;
            const int cc = min(c, cols - 1);
            b[s] = x(cc) * T(c < cols ? 1 : 0);
            CP<T> col = M + (size_t) cc * ld;
#pragma unroll
            for (int t = 0; t < RT; t++) a[s][t] = col[ri[t]];
        }
        // steps past the end add zero products (b = 0); no early exit, so the loads above are
        // all issued before the first product waits for them
#pragma unroll
        for (int s = 0; s < KB; s++) {
#pragma unroll
            for (int t = 0; t < RT; t++) acc[t] = mfma(a[s][t], b[s], acc[t]);
        }
    }
}

template <int RT, typename T, typename F>
__device__ __forceinline__ void eachOut(const Acc<T> (&acc)[RT], int rows, F f) {
    const int g = rowGroup();
#pragma unroll
    for (int t = 0; t < RT; t++) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int r = RowMap<T>::row(t, g, q);
            if (r < rows) f(r, acc[t][q]);
        }
    }
}

template <int RT, typename T, typename Terms, typename F>
__device__ __forceinline__ void rowGroupProduct(int r0, int rr, Terms &terms, F &out) {
    Acc<T> a[RT];
    zeroAcc<RT, T>(a);
    terms(a, r0, rr);
    eachOut<RT, T>(a, rr, [&](int r, T v) { out(r0 + r, v); });
}

// out(r, y_r) for every row r < rows this lane owns, y = the sum of the products terms(acc, r0, rr)
// accumulates (calls of mma on the row block r0 .. r0 + rr of each term's matrix), 64 rows at a time.
template <typename T, typename Terms, typename F>
__device__ __forceinline__ void gemv(int rows, Terms terms, F out) {
    for (int r0 = 0; r0 < rows; r0 += 64) {
        const int rr = min(64, rows - r0);
        if (rr > 48) rowGroupProduct<4, T>(r0, rr, terms, out);
        else if (rr > 32) rowGroupProduct<3, T>(r0, rr, terms, out);
        else if (rr > 16) rowGroupProduct<2, T>(r0, rr, terms, out);
        else rowGroupProduct<1, T>(r0, rr, terms, out);
    }
}

template <typename T>
__device__ __forceinline__ T projBox(T v, T lb, T ub) {  // src/definitions.cu:38-52
    const int lower = (v <= lb), upper = (v >= ub);
    const T lbi = isinf(lb) ? T(0) : T(lower) * lb;
    const T ubi = isinf(ub) ? T(0) : T(upper) * ub;
    return v * T(1 - lower) * T(1 - upper) + lbi + ubi;
}

// ------------------------------------------------------------------ L* per node
// src/operator.cuh:63-82; src/costs.cuh:162-195 (quadratic), :418-458 (quadratic+linear);
// src/constraints.cuh:171-191 (rectangle), :265-286 (polyhedron).
// emit(kind, k, value): kind 0 = x (k < nx), 1 = u (k < nu), 2 = y (k < yD), 3 = s, 4 = t; each
// entry of this lane's instance is emitted by exactly one lane.
template <typename T, typename E>
__device__ __forceinline__ void adjNode(LP<T> P, const TV<T> &d, int n, E emit) {
    const int nx = P.nx, nu = P.nu, nxu = nx + nu, d4 = nxu + 2;
    const bool first = rowGroup() == 0;
    if (n < P.nl) {
        const T II = d[P.oII + n];
        if (first) emit(3, 0, II);
        CP<T> bn = tab(P.b, n);
        forRows<T>(P.yD, [&](int k) { emit(2, k, d[P.oI + (long) n * P.yD + k] - bn[k] * II); });
        const long b3 = P.oIII + (long) n * P.con3;
        const int c0 = cst(P.chFrom)[n], nc = cst(P.numCh)[n];
        for (int part = 0; part < 2; part++) {  // x rows, then u rows
            const int rows = part == 0 ? nx : nu;
            gemv<T>(
                rows,
                [&](auto &acc, int r0, int rr) {
                    if (P.ncon == CON_POLY_DEV)
                        mma(acc, tab(P.nGT, n) + part * nx + r0, rr, nxu, P.con3, [&](int cc) { return d[b3 + cc]; });
                    if (P.ncost == COST_LIN_DEV) return;  // elementwise below
                    for (int ci = 0; ci < nc; ci++) {  // children accumulate (src/costs.cuh:172-177)
                        const long iv = P.oIV + (long) (c0 + ci) * d4;
                        if (part == 0)
                            mma(acc, tab(P.sQ, c0 + ci) + r0, rr, nx, nx, [&](int cc) { return d[iv + cc]; });
                        else
                            mma(acc, tab(P.sR, c0 + ci) + r0, rr, nu, nu, [&](int cc) { return d[iv + nx + cc]; });
                    }
                },
                [&](int r, T v) {
                    if (P.ncon == CON_RECT_DEV) v += d[b3 + part * nx + r];
                    if (P.ncost == COST_LIN_DEV) {  // q_j sum_j, r_j sum_j (src/costs.cuh:274-285)
                        for (int ci = 0; ci < nc; ci++) {
                            const int j = c0 + ci;
                            v += (part == 0 ? tab(P.q, j) : tab(P.r, j))[r] * d[P.oIV + j];
                        }
                    }
                    if (P.ncost == COST_QUADLIN_DEV) {
                        for (int ci = 0; ci < nc; ci++) {
                            const int j = c0 + ci;
                            const long iv = P.oIV + (long) j * d4;
                            const T a = d[iv + nxu], bb = d[iv + nxu + 1];
                            const T qr = (part == 0 ? tab(P.q, j) : tab(P.r, j))[r];
                            v = v - T(0.5) * qr * a - T(0.5) * qr * bb;
                        }
                    }
                    emit(part, r, v);
                });
        }
    } else {
        const int l = n - P.nl;
        const long b5 = P.oV + (long) l * P.con5;
        const bool lin = P.lcost == COST_LIN_DEV;  // block VI = (sum of every leaf; set of every leaf)
        const long vi = P.oVI + (long) l * (lin ? 1 : nx + 2);
        const T a = lin ? T(0) : d[vi + nx], bb = lin ? T(0) : d[vi + nx + 1];
        gemv<T>(
            nx,
            [&](auto &acc, int r0, int rr) {
                if (P.lcon == CON_POLY_DEV)
                    mma(acc, tab(P.lGT, l) + r0, rr, nx, P.con5, [&](int cc) { return d[b5 + cc]; });
                if (!lin) mma(acc, tab(P.lsQ, l) + r0, rr, nx, nx, [&](int cc) { return d[vi + cc]; });
            },
            [&](int r, T v) {
                if (P.lcon == CON_RECT_DEV) v += d[b5 + r];
                if (lin) v += tab(P.lq, l)[r] * d[vi];  // q sum_l (src/costs.cuh:289-292)
                if (P.lcost == COST_QUADLIN_DEV) {
                    const T qr = tab(P.lq, l)[r];
                    v = v - T(0.5) * qr * a - T(0.5) * qr * bb;
                }
                emit(0, r, v);
            });
        if (first) emit(3, 0, lin ? d[P.oVI + P.nlf + l] : (a + bb) * T(0.5));
    }
    if (first) {
        if (P.ncost == COST_LIN_DEV) {
            emit(4, 0, d[P.oIV + P.N + n]);  // t <- set, every node (src/costs.cuh:286)
        } else if (n >= 1) {
            const long iv = P.oIV + (long) n * d4;
            emit(4, 0, (d[iv + nxu] + d[iv + nxu + 1]) * T(0.5));
        } else {
            emit(4, 0, T(0));
        }
    }
}

// ------------------------------------------------------------------ L per node
// src/operator.cuh:41-60; src/costs.cuh:126-160, :379-416; src/constraints.cuh:158-169, :251-263.
// Handler: elemI(n, k, v), elemII(n, v) (row group 0 only), elemBox(e, v, lb, ub) for III and V,
// and for the SOC blocks IV and VI socBegin(), socPart(e, v, trans) for the owned rows of all but
// the last two entries, then socEnd(base, n1, n2, h, trans) for a block of n1 + n2 + 2 entries
// whose last two both carry h.
template <typename T, typename PV, typename H>
__device__ __forceinline__ void opNode(LP<T> P, int n, PV pv, H &h) {
    const int nx = P.nx, nu = P.nu, nxu = nx + nu;
    const int g = rowGroup();
    if (n < P.nl) {
        CP<T> bn = tab(P.b, n);
        T by = T(0);
        forRows<T>(P.yD, [&](int k) {
            const T yk = pv(P.oY + (long) n * P.yD + k);
            by += bn[k] * yk;
            h.elemI(n, k, yk);
        });
        by = sumG(by);
        if (g == 0) h.elemII(n, pv(P.oS + n) - by);
        const long base = P.oIII + (long) n * P.con3;
        if (P.ncon == CON_RECT_DEV) {
            CP<T> lb = tab(P.nLB, n), ub = tab(P.nUB, n);
            forRows<T>(nx, [&](int k) { h.elemBox(base + k, pv(P.oX + (long) n * nx + k), lb[k], ub[k]); });
            forRows<T>(nu, [&](int k) { h.elemBox(base + nx + k, pv(P.oU + (long) n * nu + k), lb[nx + k], ub[nx + k]); });
        } else if (P.ncon == CON_POLY_DEV) {
            CP<T> G = tab(P.nG, n), lb = tab(P.nLB, n), ub = tab(P.nUB, n);
            gemv<T>(
                P.con3,
                [&](auto &acc, int r0, int rr) {
                    mma(acc, G + r0, rr, P.con3, nxu, [&](int cc) {
                        return cc < nx ? pv(P.oX + (long) n * nx + cc) : pv(P.oU + (long) n * nu + cc - nx);
                    });
                },
                [&](int r, T v) { h.elemBox(base + r, v, lb[r], ub[r]); });
        }
    }
    if (P.ncost == COST_LIN_DEV) {  // IV: (q_n' x_anc + r_n' u_anc, t_n), node 0 (0, t_0) (src/costs.cuh:261-267)
        T sum = T(0);
        if (n >= 1) {
            const int a = cst(P.anc)[n];
            CP<T> qn = tab(P.q, n), rn = tab(P.r, n);
            T pq = T(0), pr = T(0);
            forRows<T>(nx, [&](int k) { pq += qn[k] * pv(P.oX + (long) a * nx + k); });
            forRows<T>(nu, [&](int k) { pr += rn[k] * pv(P.oU + (long) a * nu + k); });
            sum = sumG(pq) + sumG(pr);
        }
        if (g == 0) h.lin(P.oIV + n, P.oIV + P.N + n, sum, pv(P.oT + n));
    } else {  // IV
        const long base = P.oIV + (long) n * (nxu + 2);
        CP<T> tr = tab(P.trans, n);
        h.socBegin();
        if (n >= 1) {
            const int a = cst(P.anc)[n];
            gemv<T>(
                nx,
                [&](auto &acc, int r0, int rr) {
                    mma(acc, tab(P.sQ, n) + r0, rr, nx, nx, [&](int cc) { return pv(P.oX + (long) a * nx + cc); });
                },
                [&](int r, T v) { h.socPart(base + r, v, tr[r]); });
            gemv<T>(
                nu,
                [&](auto &acc, int r0, int rr) {
                    mma(acc, tab(P.sR, n) + r0, rr, nu, nu, [&](int cc) { return pv(P.oU + (long) a * nu + cc); });
                },
                [&](int r, T v) { h.socPart(base + nx + r, v, tr[nx + r]); });
            T hv = pv(P.oT + n) * T(0.5);
            if (P.ncost == COST_QUADLIN_DEV) {
                CP<T> qn = tab(P.q, n), rn = tab(P.r, n);
                T pq = T(0), pr = T(0);
                forRows<T>(nx, [&](int k) { pq += qn[k] * pv(P.oX + (long) a * nx + k); });
                forRows<T>(nu, [&](int k) { pr += rn[k] * pv(P.oU + (long) a * nu + k); });
                pq = sumG(pq);
                pr = sumG(pr);
                hv = hv + T(-0.5) * pq;
                hv = hv + T(-0.5) * pr;
            }
            h.socEnd(base, nx, nu, hv, tr);
        } else {
            forRows<T>(nx, [&](int r) { h.socPart(base + r, T(0), tr[r]); });
            forRows<T>(nu, [&](int r) { h.socPart(base + nx + r, T(0), tr[nx + r]); });
            h.socEnd(base, nx, nu, T(0), tr);
        }
    }
    if (n >= P.nl) {  // V and VI
        const int l = n - P.nl;
        const long b5 = P.oV + (long) l * P.con5;
        if (P.lcon == CON_RECT_DEV) {
            CP<T> lb = tab(P.lLB, l), ub = tab(P.lUB, l);
            forRows<T>(nx, [&](int k) { h.elemBox(b5 + k, pv(P.oX + (long) n * nx + k), lb[k], ub[k]); });
        } else if (P.lcon == CON_POLY_DEV) {
            CP<T> lb = tab(P.lLB, l), ub = tab(P.lUB, l);
            gemv<T>(
                P.con5,
                [&](auto &acc, int r0, int rr) {
                    mma(acc, tab(P.lG, l) + r0, rr, P.con5, nx, [&](int cc) { return pv(P.oX + (long) n * nx + cc); });
                },
                [&](int r, T v) { h.elemBox(b5 + r, v, lb[r], ub[r]); });
        }
        if (P.lcost == COST_LIN_DEV) {  // VI: (q' x_l, s_l) (src/costs.cuh:269-272)
            CP<T> lq = tab(P.lq, l);
            T pq = T(0);
            forRows<T>(nx, [&](int k) { pq += lq[k] * pv(P.oX + (long) n * nx + k); });
            pq = sumG(pq);
            if (g == 0) h.lin(P.oVI + l, P.oVI + P.nlf + l, pq, pv(P.oS + n));
            return;
        }
        const long base = P.oVI + (long) l * (nx + 2);
        CP<T> tr = tab(P.ltrans, l);
        h.socBegin();
        gemv<T>(
            nx,
            [&](auto &acc, int r0, int rr) {
                mma(acc, tab(P.lsQ, l) + r0, rr, nx, nx, [&](int cc) { return pv(P.oX + (long) n * nx + cc); });
            },
            [&](int r, T v) { h.socPart(base + r, v, tr[r]); });
        T hs = pv(P.oS + n) * T(0.5);
        if (P.lcost == COST_QUADLIN_DEV) {
            CP<T> lq = tab(P.lq, l);
            T pq = T(0);
            forRows<T>(nx, [&](int k) { pq += lq[k] * pv(P.oX + (long) n * nx + k); });
            pq = sumG(pq);
            hs = hs + T(-0.5) * pq;
        }
        h.socEnd(base, nx, 0, hs, tr);
    }
}

// Dual prox of the CP step (src/cache.cuh:770-789): w = (L gamma + z_d)/gamma + c and
// cand_d = gamma (w - Pi(w)), Pi the masked NNOC (src/risks.cuh:99-106), NNOC (src/cones.cuh:134-137),
// box (src/constraints.cuh:70-74) or SOC (src/projections.cuh:102-128).  SOC blocks stage w in
// cand_d (each lane its own rows), accumulate |w|^2 and rescale in socEnd.
template <typename T>
struct Prox {
    LP<T> P;
    TV<T> zd, cd;
    T ss;
    __device__ __forceinline__ T wOf(T Lv, long e, T tr) const { return (Lv * P.gamma + zd[e]) * P.gammaRecip + tr; }
    __device__ __forceinline__ void elemI(int n, int k, T v) {
        const long e = P.oI + (long) n * P.yD + k;
        const T w = wOf(v, e, T(0));
        const int idx = (k < 2 * cst(P.numCh)[n]) && (w < T(0));
        cd[e] = (w - T(1 - idx) * w) * P.gamma;
    }
    __device__ __forceinline__ void elemII(int n, T v) {
        const long e = P.oII + n;
        const T w = wOf(v, e, T(0));
        cd[e] = (w - fmax(T(0), w)) * P.gamma;
    }
    __device__ __forceinline__ void elemBox(long e, T v, T lb, T ub) {
        const T w = wOf(v, e, T(0));
        cd[e] = (w - projBox(w, lb, ub)) * P.gamma;
    }
    __device__ __forceinline__ void socBegin() { ss = T(0); }
    __device__ __forceinline__ void socPart(long e, T v, T tr) {
        const T w = wOf(v, e, tr);
        cd[e] = w;
        ss += w * w;
    }
    __device__ __forceinline__ void socEnd(long base, int n1, int n2, T hv, CP<T> tr) {
        const int dim = n1 + n2 + 2;
        const long k1 = base + dim - 2, k2 = base + dim - 1;
        const T w1 = wOf(hv, k1, tr[dim - 2]);
        const T w2 = wOf(hv, k2, tr[dim - 1]);
        const T sq = sumG(ss) + w1 * w1;
        const T nrm = sqrt(sq), tl = w2;
        const int i1 = nrm <= tl;
        const int i2 = (nrm <= -tl) && !i1;
        const int i3 = !i1 && !i2;
        const T sc = (nrm + tl) / (T(2) * nrm + T(1 - i3));
        const T mult = T(1 - i2) * (T(i3) * sc + T(1 - i3));
        auto rescale = [&](long e) {
            const T w = cd[e];
            cd[e] = (w - w * mult) * P.gamma;
        };
        forRows<T>(n1, [&](int r) { rescale(base + r); });
        forRows<T>(n2, [&](int r) { rescale(base + n1 + r); });
        if (rowGroup() == 0) {
            const T p1 = w1 * mult;
            const int i1b = (1 - i2) * (1 - i3);
            const T p2 = T(i1b) * (w2 * T(1 - i2)) + T((1 - i1b) * (1 - i2)) * (T(i3) * (sc * nrm - T(1)) + T(1));
            cd[k1] = (w1 - p1) * P.gamma;
            cd[k2] = (w2 - p2) * P.gamma;
        }
    }
    __device__ __forceinline__ void lin(long eSum, long eSet, T sum, T set) {
        const T ws = wOf(sum, eSum, T(0)), wt = wOf(set, eSet, T(0));
        cd[eSum] = (ws - projBelow(ws, wt)) * P.gamma;
        cd[eSet] = (wt - wt) * P.gamma;
    }
};

// ell(y) = [L* y_d; L y_p] reductions: sum x.(y - gamma ell) and max |y/gamma - ell|
// (src/cache.cuh:722-734, 917-931), per lane over its own entries.
template <typename T>
struct Red {
    LP<T> P;
    TV<T> x, y;
    T dot, mx;
    __device__ __forceinline__ void acc(long e, T ell) {
        const T yk = y[e];
        dot += x[e] * (ell * (-P.gamma) + yk);
        mx = fmax(mx, fabs(yk * P.gammaRecip - ell));
    }
    __device__ __forceinline__ void elemI(int n, int k, T v) { acc(P.sPrim + P.oI + (long) n * P.yD + k, v); }
    __device__ __forceinline__ void elemII(int n, T v) { acc(P.sPrim + P.oII + n, v); }
    __device__ __forceinline__ void elemBox(long e, T v, T, T) { acc(P.sPrim + e, v); }
    __device__ __forceinline__ void socBegin() {}
    __device__ __forceinline__ void socPart(long e, T v, T) { acc(P.sPrim + e, v); }
    __device__ __forceinline__ void socEnd(long base, int n1, int n2, T hv, CP<T>) {
        if (rowGroup() != 0) return;
        const int dim = n1 + n2 + 2;
        acc(P.sPrim + base + dim - 2, hv);
        acc(P.sPrim + base + dim - 1, hv);
    }
    __device__ __forceinline__ void lin(long eSum, long eSet, T sum, T set) {
        acc(P.sPrim + eSum, sum);
        acc(P.sPrim + eSet, set);
    }
};

template <typename T>
struct Store {
    LP<T> P;
    TV<T> out;  // dual part
    __device__ __forceinline__ void elemI(int n, int k, T v) { out[P.oI + (long) n * P.yD + k] = v; }
    __device__ __forceinline__ void elemII(int n, T v) { out[P.oII + n] = v; }
    __device__ __forceinline__ void elemBox(long e, T v, T, T) { out[e] = v; }
    __device__ __forceinline__ void socBegin() {}
    __device__ __forceinline__ void socPart(long e, T v, T) { out[e] = v; }
    __device__ __forceinline__ void socEnd(long base, int n1, int n2, T hv, CP<T>) {
        if (rowGroup() != 0) return;
        const int dim = n1 + n2 + 2;
        out[base + dim - 2] = hv;
        out[base + dim - 1] = hv;
    }
    __device__ __forceinline__ void lin(long eSum, long eSet, T sum, T set) {
        out[eSum] = sum;
        out[eSet] = set;
    }
};

template <typename T>
__device__ __forceinline__ long primIndex(LP<T> P, int n, int kind, int k) {
    if (kind == 0) return P.oX + (long) n * P.nx + k;
    if (kind == 1) return P.oU + (long) n * P.nu + k;
    if (kind == 2) return P.oY + (long) n * P.yD + k;
    if (kind == 3) return P.oS + n;
    return P.oT + n;
}

// ------------------------------------------------------------------ T phases
// modifyPrimal + proxRootS (src/cache.cuh:748-753, 639-642); leaves seed q = -xbar (src/dynamics.cuh:79-81).
template <typename T>
__device__ __forceinline__ void phaseAdj(LP<T> P, const Ctx &c, const TV<T> &z, const TV<T> &cand,
                                         const TV<T> &qd) {
    const T g = -P.gamma;
    const TV<T> zd = z.shift(P.sPrim);
    for (int n = c.wave; n < P.N; n += c.nw) {
        adjNode(P, zd, n, [&](int kind, int k, T v) {
            const long e = primIndex(P, n, kind, k);
            T val = v * g + z[e];
            if (kind == 3 && n == 0) val -= P.gamma;
            cand[e] = val;
            if (kind == 0 && n >= P.nl) qd[(long) n * P.nx + k] = -val;
        });
    }
}

// Backward sweep step for node i (src/dynamics.cuh:85-139) with R~^-1 applied as one product.
// stg: this wave's two LDS slices of 32 rows x 16 instances.
template <typename T>
__device__ __forceinline__ void dynBackward(LP<T> P, const TV<T> &cand, const TV<T> &qd, const TV<T> &dd, int i,
                                            int stage, SP<T> stg) {
    const int nx = P.nx, nu = P.nu;  // nu <= 32 (checked on the host)
    const int c0 = cst(P.chFrom)[i], nc = cst(P.numCh)[i];
    const int jj = instLane();
    SP<T> e1 = stg, d1 = stg + STG;
    // affine dynamics: the children's q enter as q_j + P_j e_j (src/dynamics.cuh:93-97)
    const bool affine = P.affine != 0;
    auto qv = [&](int j, int cc) { return affine ? qd[(long) j * nx + cc] + tab(P.Pe, j)[cc] : qd[(long) j * nx + cc]; };
    // e = ubar - sum_j B_j' q_j
    gemv<T>(
        nu,
        [&](auto &acc, int r0, int rr) {
            for (int ci = 0; ci < nc; ci++)
                mma(acc, tab(P.BTr, c0 + ci) + r0, rr, nu, nx, [&](int cc) { return qv(c0 + ci, cc); });
        },
        [&](int r, T v) { e1[r * W + jj] = -v + cand[P.oU + (long) i * nu + r]; });
    __builtin_amdgcn_wave_barrier();
    // d = R~^-1 e
    gemv<T>(
        nu,
        [&](auto &acc, int r0, int rr) {
            mma(acc, tab(P.RtInv, i) + r0, rr, nu, nu, [&](int cc) { return e1[cc * W + jj]; });
        },
        [&](int r, T v) {
            dd[(long) i * nu + r] = v;
            d1[r * W + jj] = v;
        });
    __builtin_amdgcn_wave_barrier();
    if (stage == 0) return;
    // q_i = sum_j (APB_j d_i + ABKTr_j q_j) + K_i'(d_i - u_i) - x_i
    gemv<T>(
        nx,
        [&](auto &acc, int r0, int rr) {
            for (int ci = 0; ci < nc; ci++) {
                const int j = c0 + ci;
                mma(acc, tab(P.APB, j) + r0, rr, nx, nu, [&](int cc) { return d1[cc * W + jj]; });
                mma(acc, tab(P.ABKTr, j) + r0, rr, nx, nx, [&](int cc) { return qv(j, cc); });
            }
            mma(acc, tab(P.KTr, i) + r0, rr, nx, nu,
                [&](int cc) { return d1[cc * W + jj] - cand[P.oU + (long) i * nu + cc]; });
        },
        [&](int r, T v) { qd[(long) i * nx + r] = v - cand[P.oX + (long) i * nx + r]; });
}

// Kernel projection for nonleaf node i (src/cache.cuh:650-676); padded child slots enter as 0.
// In place: every entry is read (as a B operand) before any result is written.
template <typename T>
__device__ __forceinline__ void kernelProj(LP<T> P, const TV<T> &cand, int i) {
    const int yD = P.yD, E = P.E, nc = cst(P.numCh)[i], c0 = cst(P.chFrom)[i];
    auto vin = [&](int k) -> T {
        if (k < yD) return cand[P.oY + (long) i * yD + k];
        if (k < yD + E) return (k - yD < nc) ? cand[P.oT + c0 + (k - yD)] : T(0);
        return (k - yD - E < nc) ? cand[P.oS + c0 + (k - yD - E)] : T(0);
    };
    gemv<T>(
        P.nD, [&](auto &acc, int r0, int rr) { mma(acc, tab(P.NN, i) + r0, rr, P.nD, P.nD, vin); },
        [&](int k, T v) {
            if (k < yD) cand[P.oY + (long) i * yD + k] = v;
            else if (k < yD + E) {
                if (k - yD < nc) cand[P.oT + c0 + (k - yD)] = v;
            } else if (k - yD - E < nc) {
                cand[P.oS + c0 + (k - yD - E)] = v;
            }
        });
}

// Forward rollout for node i (src/dynamics.cuh:144-173); affine dynamics add e_j to each child.
template <typename T>
__device__ __forceinline__ void dynForward(LP<T> P, const TV<T> &cand, const TV<T> &x0s, const TV<T> &dd, int i,
                                           SP<T> stg) {
    const int nx = P.nx, nu = P.nu;
    const int jj = instLane();
    SP<T> u1 = stg;
    // x_0 is the (scaled) initial state; other nodes' x were written by their parent's step
    auto xv = [&](int cc) { return i == 0 ? x0s[cc] : cand[P.oX + (long) i * nx + cc]; };
    if (i == 0) forRows<T>(nx, [&](int r) { cand[P.oX + r] = x0s[r]; });
    gemv<T>(
        nu, [&](auto &acc, int r0, int rr) { mma(acc, tab(P.K, i) + r0, rr, nu, nx, xv); },
        [&](int r, T v) {
            const T u = v + dd[(long) i * nu + r];
            cand[P.oU + (long) i * nu + r] = u;
            u1[r * W + jj] = u;
        });
    __builtin_amdgcn_wave_barrier();
    const int c0 = cst(P.chFrom)[i], nc = cst(P.numCh)[i];
    for (int ci = 0; ci < nc; ci++) {
        const int j = c0 + ci;
        gemv<T>(
            nx,
            [&](auto &acc, int r0, int rr) {
                mma(acc, tab(P.AB, j) + r0, rr, nx, nx + nu,
                    [&](int cc) { return cc < nx ? xv(cc) : u1[(cc - nx) * W + jj]; });
            },
            [&](int r, T v) {
                if (P.affine) v += tab(P.e, j)[r];
                cand[P.oX + (long) j * nx + r] = v;
            });
    }
    __builtin_amdgcn_wave_barrier();
}

template <typename T>
__device__ __forceinline__ void phaseProj(LP<T> P, const Ctx &c, const TV<T> &cand, const TV<T> &qd, const TV<T> &dd,
                                          const TV<T> &x0s, bool doDyn, bool doKernel, SP<T> stg) {
    if (doDyn) {
        for (int stage = P.S - 2; stage >= 0; stage--) {
            const int f = cst(P.stageFrom)[stage], nStage = cst(P.stageTo)[stage] - f + 1;
            const int extra = (doKernel && stage == P.S - 2) ? P.nl : 0;
            for (int task = c.wave; task < nStage + extra; task += c.nw) {
                if (task < nStage) dynBackward(P, cand, qd, dd, f + task, stage, stg);
                else kernelProj(P, cand, task - nStage);
            }
            __syncthreads();
        }
        for (int stage = 0; stage <= P.S - 2; stage++) {
            const int f = cst(P.stageFrom)[stage], nStage = cst(P.stageTo)[stage] - f + 1;
            for (int task = c.wave; task < nStage; task += c.nw) dynForward(P, cand, x0s, dd, f + task, stg);
            __syncthreads();
        }
    } else if (doKernel) {
        for (int task = c.wave; task < P.nl; task += c.nw) kernelProj(P, cand, task);
        __syncthreads();
    }
}

// modifyDual + proximalDual (src/cache.cuh:770-789) on 2 xbar - z_p.
template <typename T>
__device__ __forceinline__ void phaseDual(LP<T> P, const Ctx &c, const TV<T> &z, const TV<T> &cand) {
    Prox<T> h{P, z.shift(P.sPrim), cand.shift(P.sPrim), T(0)};
    auto pv = [&](long e) { return cand[e] * T(2) - z[e]; };
    for (int n = c.wave; n < P.N; n += c.nw) opNode(P, n, pv, h);
}

template <typename T>
__device__ __forceinline__ SP<T> waveStage() {
    __shared__ __align__(16) T stage[MAXW * 2 * STG];
    return (SP<T>) (stage + (threadIdx.x >> 6) * 2 * STG);
}

// The CP operator T (src/cache.cuh:796-801): cand = T(z).
template <typename T>
__device__ __noinline__ void applyT(const LaneInst<T> *Pp, T *z, T *cand, long *phaseAcc) {
    LaneClock clk{wall_clock64(), phaseAcc};
    LP<T> P = *cst(uniformPtr(Pp));
    z = uniformPtr(z);
    cand = uniformPtr(cand);
    const Ctx c = ctx();
    const TV<T> zv = tview(z, P.sIter, c.j), cv = tview(cand, P.sIter, c.j);
    const TV<T> qd = tview(P.qd, (long) P.nx * P.N, c.j), dd = tview(P.dd, (long) P.nu * P.nl, c.j);
    const TV<T> x0s = tview(const_cast<T *>(P.x0s), P.nx, c.j);
    phaseAdj(P, c, zv, cv, qd);
    __syncthreads();
    clk.mark(PH_ADJ);
    phaseProj(P, c, cv, qd, dd, x0s, true, true, waveStage<T>());
    clk.mark(PH_PROJ);
    phaseDual(P, c, zv, cv);
    __syncthreads();
    clk.mark(PH_DUAL);
}

// Per-instance combination: the four row-group lanes, then the waves in wave order (LDS).
__device__ __forceinline__ double tileSum(double v, double *lds, const Ctx &c) {
    v = sumG(v);
    __syncthreads();
    if (c.g == 0) lds[c.wave * W + c.j] = v;
    __syncthreads();
    double s = 0.;
    for (int w = 0; w < c.nw; w++) s += lds[w * W + c.j];
    return s;
}

__device__ __forceinline__ double tileMax(double v, double *lds, const Ctx &c) {
    v = maxG(v);
    __syncthreads();
    if (c.g == 0) lds[c.wave * W + c.j] = v;
    __syncthreads();
    double s = 0.;
    for (int w = 0; w < c.nw; w++) s = fmax(s, lds[w * W + c.j]);
    return s;
}

// sum x.(y - gamma ell(y)) and max |y/gamma - ell(y)| per instance.
template <typename T>
__device__ __noinline__ DotMax<T> reduceEll(const LaneInst<T> *Pp, T *x, T *y, double *lds) {
    LP<T> P = *cst(uniformPtr(Pp));
    x = uniformPtr(x);
    y = uniformPtr(y);
    const Ctx c = ctx();
    const TV<T> xv = tview(x, P.sIter, c.j), yv = tview(y, P.sIter, c.j);
    const TV<T> yd = yv.shift(P.sPrim);
    Red<T> h{P, xv, yv, T(0), T(0)};
    T dot = T(0), mx = T(0);
    T cross = T(0);
    auto pv = [&](long e) { return yv[e]; };
    for (int n = c.wave; n < P.N; n += c.nw) {
        adjNode(P, yd, n, [&](int kind, int k, T v) {
            const long e = primIndex(P, n, kind, k);
            const T yk = yv[e], xk = xv[e];
            dot += xk * (v * (-P.gamma) + yk);
            cross += xk * v;
            mx = fmax(mx, fabs(yk * P.gammaRecip - v));
        });
        opNode(P, n, pv, h);
    }
    DotMax<T> r;
    r.dot = (T) tileSum((double) (dot + h.dot), lds, c);
    r.mx = (T) tileMax((double) fmax(mx, h.mx), lds, c);
    r.cross = (T) tileSum((double) cross, lds, c);
    return r;
}

// c1 = sum_p x_p . (L* x_d),  c2 = sum_p x_p . (L* y_d) + (L* x_d) . y_p (see the lane engine).
template <typename T>
__device__ __noinline__ Cross2<T> adjDots(const LaneInst<T> *Pp, T *x, T *y, double *lds, bool withY) {
    LP<T> P = *cst(uniformPtr(Pp));
    x = uniformPtr(x);
    y = uniformPtr(y);
    const Ctx c = ctx();
    const TV<T> xv = tview(x, P.sIter, c.j), yv = tview(y, P.sIter, c.j);
    const TV<T> xd = xv.shift(P.sPrim), yd = yv.shift(P.sPrim);
    T c1 = T(0), c2 = T(0);
    for (int n = c.wave; n < P.N; n += c.nw) {
        adjNode(P, xd, n, [&](int kind, int k, T v) {
            const long e = primIndex(P, n, kind, k);
            c1 += xv[e] * v;
            if (withY) c2 += v * yv[e];
        });
        if (withY) adjNode(P, yd, n, [&](int kind, int k, T v) { c2 += xv[primIndex(P, n, kind, k)] * v; });
    }
    Cross2<T> r;
    r.c1 = (T) tileSum((double) c1, lds, c);
    r.c2 = withY ? (T) tileSum((double) c2, lds, c) : T(0);
    return r;
}

// Iterate-wide pass: wave w takes a contiguous range of elements (a multiple of 4 long); lane
// (g, j) visits elements e0 + 4 s + g of instance j, so each access of the wave moves 4 rows x 16
// instances = 64 consecutive words, with U accesses in flight per lane.
template <int U, typename Ld, typename St>
__device__ __forceinline__ void pass(long n, const Ctx &c, Ld ld, St st) {
    using Vals = decltype(ld(0L));
    const long per = (((n + c.nw - 1) / c.nw) + 3) & ~3L;
    const long e0 = (long) c.wave * per, e1 = min(n, e0 + per);
    for (long b = e0; b < e1; b += 4 * U) {
        Vals v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const long e = b + 4 * u + c.g;
            if (e < e1) v[u] = ld(e);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const long e = b + 4 * u + c.g;
            if (e < e1) st(e, v[u]);
        }
    }
}

// The tile engine as a policy of the lockstep solve.
struct Policy {
    template <typename T>
    using View = TV<T>;
    __device__ static Ctx ctx() { return tile::ctx(); }
    template <typename T>
    __device__ static TV<T> view(T *region, long rows, const Ctx &c) { return tview(region, rows, c.j); }
    template <int U, typename Ld, typename St>
    __device__ static void pass(long n, const Ctx &c, Ld ld, St st) { tile::pass<U>(n, c, ld, st); }
    __device__ static double sum(double v, double *lds, const Ctx &c) { return tileSum(v, lds, c); }
    __device__ static double max(double v, double *lds, const Ctx &c) { return tileMax(v, lds, c); }
    template <typename T>
    __device__ static void applyT(const LaneInst<T> *P, T *z, T *cand, long *phaseAcc) { tile::applyT(P, z, cand, phaseAcc); }
    template <typename T>
    __device__ static DotMax<T> reduceEll(const LaneInst<T> *P, T *x, T *y, double *lds) { return tile::reduceEll(P, x, y, lds); }
    template <typename T>
    __device__ static Cross2<T> adjDots(const LaneInst<T> *P, T *x, T *y, double *lds, bool withY) {
        return tile::adjDots(P, x, y, lds, withY);
    }
};

}  // namespace tile

template <typename T>
__global__ void __launch_bounds__(SPK_TILE_MAX_THREADS) SPK_TILE_KERNEL(k_tile_solve)(const LaneInst<T> *__restrict__ P, Out *outs) {
    __shared__ double lds[tile::MAXW * tile::W];
    __shared__ long phaseAcc[NUM_PHASES];
    lockstepSolve<T, tile::Policy>(P, outs, lds, phaseAcc);
}

// Single operators on every instance (parity tests): z = zio, cand = work[1].
template <typename T>
__global__ void __launch_bounds__(SPK_TILE_MAX_THREADS) SPK_TILE_KERNEL(k_tile_debug)(const LaneInst<T> *__restrict__ Pp, Out *outs,
                                                                     int op) {
    using namespace tile;
    __shared__ double lds[MAXW * W];
    LP<T> P = *cst(Pp);
    const Ctx c = ctx();
    T *z = P.zio, *cand = P.work[1];
    const TV<T> zv = tview(z, P.sIter, c.j), cv = tview(cand, P.sIter, c.j);
    if (op == 0) {  // L
        Store<T> h{P, cv.shift(P.sPrim)};
        auto pv = [&](long e) { return zv[e]; };
        for (int n = c.wave; n < P.N; n += c.nw) opNode(P, n, pv, h);
    } else if (op == 1) {  // L*
        const TV<T> zd = zv.shift(P.sPrim);
        for (int n = c.wave; n < P.N; n += c.nw)
            adjNode(P, zd, n, [&](int kind, int k, T v) { cv[primIndex(P, n, kind, k)] = v; });
    } else if (op == 2 || op == 3) {  // dynamics / kernel projection of the primal part
        tile::pass<16>(P.sPrim, c, [&](long e) { return zv[e]; }, [&](long e, T v) { cv[e] = v; });
        __syncthreads();
        const TV<T> qd = tview(P.qd, (long) P.nx * P.N, c.j), dd = tview(P.dd, (long) P.nu * P.nl, c.j);
        const TV<T> x0s = tview(const_cast<T *>(P.x0s), P.nx, c.j);
        if (op == 2)
            for (int n = c.wave; n < P.N; n += c.nw)
                if (n >= P.nl) forRows<T>(P.nx, [&](int r) { qd[(long) n * P.nx + r] = -cv[P.oX + (long) n * P.nx + r]; });
        __syncthreads();
        phaseProj(P, c, cv, qd, dd, x0s, op == 2, op == 3, waveStage<T>());
    } else if (op == 4) {  // T
        __shared__ long phaseAcc[NUM_PHASES];
        applyT(Pp, z, cand, phaseAcc);
    } else if (op == 5) {  // ell reductions
        const DotMax<T> r = reduceEll(Pp, z, cand, lds);
        if (c.wave == 0 && c.writer && c.gl < P.count) {
            outs[c.gl].w = (double) r.dot;
            outs[c.gl].err = (double) r.mx;
        }
    }
}

template __global__ void SPK_TILE_KERNEL(k_tile_solve)<float>(const LaneInst<float> *, Out *);
template __global__ void SPK_TILE_KERNEL(k_tile_solve)<double>(const LaneInst<double> *, Out *);
template __global__ void SPK_TILE_KERNEL(k_tile_debug)<float>(const LaneInst<float> *, Out *, int);
template __global__ void SPK_TILE_KERNEL(k_tile_debug)<double>(const LaneInst<double> *, Out *, int);

}  // namespace spk
