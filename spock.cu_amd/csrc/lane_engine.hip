// Lane engine: many instances of ONE problem solved together, lanes over instances.
//
// Layout (LaneInst in engine.hpp): element e of instance b of every state vector is at
// base[e * B + b], so each memory instruction of a wave moves 64 consecutive instances, and the
// factor tables are read at wave-uniform addresses (scalar loads, the operand of every FMA is a
// scalar register).  A workgroup owns 64 instances; its waves split the node tasks of each tree
// stage and the element ranges of the iterate-wide passes.  Per-instance norms are per-lane
// sums, combined across waves through LDS in a fixed order, so results do not depend on timing.
// SuperMann (src/cache.cuh:1008-1132) runs in lockstep outer iterations with per-lane decisions:
// the buffer roles rotate uniformly where every lane takes the same step and per-lane masked
// copies replace the rotations where lanes branch (K0 vs line search, K1/K2/K3).
//
// Operator semantics are those of the single-instance engine and of the reference
// (src/operator.cuh:41-82, src/dynamics.cuh:74-174, src/cache.cuh:650-801).  Deviations: the
// dynamics sweep always applies the precomputed (L L')^-1 and the Anderson least squares is the
// FP64 Gram/Cholesky solve (the same arithmetic as the single-instance engine's Gram option).
#include "engine.hpp"
#include "lockstep.hpp"

#include <hip/hip_runtime.h>
#include <math.h>

#ifndef SPK_LANE_MAX_THREADS
#define SPK_LANE_MAX_THREADS 1024
#endif

// Matrix-vector schedule of this compilation: 0 reads the factor entries with scalar loads; 4, 8 or 16
// stage them through LDS in chunks of that many columns.  The library is built once per schedule;
// the kernels of a build carry its suffix (k_lane_solve, k_lane_solve_x8, ...), and the autotuner
// of the host side picks one per batch.
#ifndef SPK_MV_XB
#define SPK_MV_XB 0
#endif
#if SPK_MV_XB == 0
#define SPK_LANE_KERNEL(name) name
#else
#define SPK_LANE_KERNEL_CAT2(a, b) a##_x##b
#define SPK_LANE_KERNEL_CAT(a, b) SPK_LANE_KERNEL_CAT2(a, b)
#define SPK_LANE_KERNEL(name) SPK_LANE_KERNEL_CAT(name, SPK_MV_XB)
#endif

namespace spk {
namespace lane {
using namespace ls;

struct Ctx {
    int lane, wave, nw, tid;
    int gl;       // this lane's instance index
    bool writer;  // every lane writes its own instance's records
};

__device__ __forceinline__ Ctx ctx() {
    Ctx c;
    c.tid = threadIdx.x;
    c.lane = c.tid & 63;
    c.wave = __builtin_amdgcn_readfirstlane(c.tid >> 6);
    c.nw = blockDim.x >> 6;
    c.gl = blockIdx.x * 64 + c.lane;
    c.writer = true;
    return c;
}

// Lane view of a region in the block-interleaved layout (each 64-instance block holds its rows
// x 64 words contiguously, so a workgroup streams its own contiguous slab): v[e] is element e of
// this lane's instance.  The row pointer p + 64 e is wave-uniform; the lane is the only per-lane
// part of the address.
template <typename T>
struct LV {
    T *p;     // this workgroup's block of the region
    int off;  // lane
    __device__ __forceinline__ SPK_GLOBAL T &operator[](long e) const { return ((SPK_GLOBAL T *) (p + e * 64))[off]; }
    __device__ __forceinline__ LV shift(long e0) const { return LV{p + e0 * 64, off}; }
};

template <typename T>
__device__ __forceinline__ LV<T> lview(T *region, long rows, int lane) {
    return LV<T>{region + (long) blockIdx.x * rows * 64, lane};
}

template <typename T>
struct RowChunk {
    static constexpr int v = 16;  // result rows held in registers per chunk
};

template <typename T, int NY>
__device__ __forceinline__ void zeroY(T (&y)[NY]) {
#pragma unroll
    for (int i = 0; i < NY; i++) y[i] = T(0);
}

#if SPK_MV_XB == 0
// y[r] += sum_c M[r + c * ld] * x(c) for r < RR: M is wave-uniform, x(c) the lane's own value.
// XB vector loads are issued before their FMAs; CB columns of the matrix at a time stay in
// scalar registers (CB * RR <= 64).  Columns past the end are clamped to the last one and enter
// with x = 0, so every load stays inside the table (which carries 64 elements of padding for the
// RR rows that may run past a matrix's last column).
template <int RR, typename T, int NY, typename XF>
__device__ __forceinline__ void mvFixed(T (&y)[NY], CP<T> M, int ld, int cols, XF x) {
    static_assert(RR <= NY, "row chunk");
    constexpr int XB = 16;
    constexpr int CB = RR >= 16 ? 4 : 8;
    for (int c0 = 0; c0 < cols; c0 += XB) {
        T xv[XB];
#pragma unroll
        for (int q = 0; q < XB; q++) {
            const int cc = min(c0 + q, cols - 1);
            const T v = x(cc);
            xv[q] = (c0 + q < cols) ? v : T(0);
        }
#pragma unroll
        for (int q0 = 0; q0 < XB; q0 += CB) {
            if (c0 + q0 >= cols) break;
#pragma unroll
            for (int q = q0; q < q0 + CB; q++) {
                CP<T> m = M + (size_t) min(c0 + q, cols - 1) * ld;
#pragma unroll
                for (int r = 0; r < RR; r++) y[r] += m[r] * xv[q];
            }
        }
    }
}

#else
// Per-wave LDS staging area of the matrix-vector products: XB columns x 16 rows of one matrix.
constexpr int MV_XB = SPK_MV_XB;
__shared__ __align__(16) double laneStage[(SPK_LANE_MAX_THREADS / 64) * MV_XB * 16];

// y[r] += sum_c M[r + c * ld] * x(c) for r < RR: M is wave-uniform, x(c) the lane's own value.
// The matrix streams through the wave's LDS slice in chunks of XB columns: each lane fetches 4
// of the chunk's XB x 16 entries with vector loads (memory order, so the next chunk's loads are
// in flight while the current chunk is consumed), the wave writes them to LDS, and every FMA
// reads its matrix entry as an LDS broadcast.  (Scalar loads of the entries complete out of
// order, so each use would wait for every outstanding one.)  Columns past the end are clamped to
// the last one and enter with x = 0, and rows past RR are not read, so every load stays inside
// the table (which carries 64 elements of padding for the RR rows that may run past a matrix's
// last column).  The accumulation order over columns is that of a plain column loop.
template <int RR, typename T, int NY, typename XF>
__device__ __forceinline__ void mvFixed(T (&y)[NY], CP<T> M, int ld, int cols, XF x) {
    static_assert(RR <= NY && RR % 4 == 0, "row chunk");
    constexpr int XB = MV_XB;
    constexpr int PER = XB * 16 / 64;  // entries each lane stages per chunk
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    T *st = reinterpret_cast<T *>(laneStage) + w * (XB * 16);
    const int q = lane / (16 / PER), r4 = (lane % (16 / PER)) * PER;  // rows r4.. of column q
    const SPK_GLOBAL T *Mg = (const SPK_GLOBAL T *) (unsigned long long) M;
    T nxt[PER];
    auto fetch = [&](int c0) {
        const SPK_GLOBAL T *col = Mg + (size_t) min(c0 + q, cols - 1) * ld;
#pragma unroll
        for (int k = 0; k < PER; k++) nxt[k] = (r4 + k < RR) ? col[r4 + k] : T(0);
    };
    fetch(0);
    for (int c0 = 0; c0 < cols; c0 += XB) {
        T xv[XB];
#pragma unroll
        for (int qq = 0; qq < XB; qq++) {
            const int cc = min(c0 + qq, cols - 1);
            const T v = x(cc);
            xv[qq] = (c0 + qq < cols) ? v : T(0);
        }
        // the previous chunk's LDS reads were issued before these writes (LDS is in order per wave)
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int k = 0; k < PER; k++) st[q * 16 + r4 + k] = nxt[k];
        __builtin_amdgcn_wave_barrier();
        if (c0 + XB < cols) fetch(c0 + XB);
#pragma unroll
        for (int qq = 0; qq < XB; qq++) {
            if (c0 + qq >= cols) break;
#pragma unroll
            for (int r = 0; r < RR; r++) y[r] += st[qq * 16 + r] * xv[qq];
        }
    }
    __builtin_amdgcn_wave_barrier();
}

#endif

// rows <= NY; the row count picks a fixed-size variant so every scalar load is unconditional.
template <typename T, int NY, typename XF>
__device__ __forceinline__ void mv(T (&y)[NY], CP<T> M, int rows, int ld, int cols, XF x) {
    static_assert(NY == 16, "row chunk");
    if (rows <= 0 || cols <= 0) return;
    if (rows <= 8) mvFixed<8>(y, M, ld, cols, x);
    else mvFixed<16>(y, M, ld, cols, x);
}

// Up to 32 result rows held at once (two 16-row blocks), for the in-place updates of d and of
// the kernel projection, which read every input row before any output row is written.
template <typename T, typename XF>
__device__ __forceinline__ void mv32(T (&y0)[16], T (&y1)[16], CP<T> M, int rows, int ld, int cols, XF x) {
    mv(y0, M, min(rows, 16), ld, cols, x);
    if (rows > 16) mv(y1, M + 16, rows - 16, ld, cols, x);
}

template <typename T, int NY, typename F>
__device__ __forceinline__ void each(const T (&y)[NY], int rows, F f) {
#pragma unroll
    for (int r = 0; r < NY; r++)
        if (r < rows) f(r, y[r]);
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
// emit(kind, k, value): kind 0 = x (k < nx), 1 = u (k < nu), 2 = y (k < yD), 3 = s, 4 = t.
template <typename T, typename E>
__device__ __forceinline__ void adjNode(LP<T> P, const LV<T> &d, int n, E emit) {
    constexpr int NY = RowChunk<T>::v;
    const int nx = P.nx, nu = P.nu, nxu = nx + nu, d4 = nxu + 2;
    T y[NY];
    if (n < P.nl) {
        const T II = d[P.oII + n];
        emit(3, 0, II);
        CP<T> bn = tab(P.b, n);
        for (int k = 0; k < P.yD; k++) emit(2, k, d[P.oI + (long) n * P.yD + k] - bn[k] * II);
        const long b3 = P.oIII + (long) n * P.con3;
        const int c0 = cst(P.chFrom)[n], nc = cst(P.numCh)[n];
        for (int part = 0; part < 2; part++) {  // x rows, then u rows
            const int rows = part == 0 ? nx : nu;
            for (int r0 = 0; r0 < rows; r0 += NY) {
                const int rr = min(NY, rows - r0);
                zeroY(y);
                if (P.ncon == CON_RECT_DEV) {
#pragma unroll
                    for (int r = 0; r < NY; r++)
                        if (r < rr) y[r] = d[b3 + part * nx + r0 + r];
                } else if (P.ncon == CON_POLY_DEV) {
                    mv(y, tab(P.nGT, n) + part * nx + r0, rr, nxu, P.con3, [&](int cc) { return d[b3 + cc]; });
                }
                for (int ci = 0; ci < nc; ci++) {  // children accumulate in order (src/costs.cuh:172-177)
                    const int j = c0 + ci;
                    if (P.ncost == COST_LIN_DEV) {  // q_j sum_j, r_j sum_j (src/costs.cuh:274-285)
                        const T sj = d[P.oIV + j];
                        CP<T> gj = (part == 0 ? tab(P.q, j) : tab(P.r, j)) + r0;
#pragma unroll
                        for (int r = 0; r < NY; r++)
                            if (r < rr) y[r] += gj[r] * sj;
                        continue;
                    }
                    const long iv = P.oIV + (long) j * d4;
                    if (part == 0) mv(y, tab(P.sQ, j) + r0, rr, nx, nx, [&](int cc) { return d[iv + cc]; });
                    else mv(y, tab(P.sR, j) + r0, rr, nu, nu, [&](int cc) { return d[iv + nx + cc]; });
                    if (P.ncost == COST_QUADLIN_DEV) {
                        const T a = d[iv + nxu], bb = d[iv + nxu + 1];
                        CP<T> qj = (part == 0 ? tab(P.q, j) : tab(P.r, j)) + r0;
#pragma unroll
                        for (int r = 0; r < NY; r++)
                            if (r < rr) y[r] = y[r] - T(0.5) * qj[r] * a - T(0.5) * qj[r] * bb;
                    }
                }
                each(y, rr, [&](int r, T v) { emit(part, r0 + r, v); });
            }
        }
    } else {
        const int l = n - P.nl;
        const long b5 = P.oV + (long) l * P.con5;
        const bool lin = P.lcost == COST_LIN_DEV;  // block VI = (sum of every leaf; set of every leaf)
        const long vi = P.oVI + (long) l * (lin ? 1 : nx + 2);
        const T a = lin ? T(0) : d[vi + nx], bb = lin ? T(0) : d[vi + nx + 1];
        for (int r0 = 0; r0 < nx; r0 += NY) {
            const int rr = min(NY, nx - r0);
            zeroY(y);
            if (P.lcon == CON_RECT_DEV) {
#pragma unroll
                for (int r = 0; r < NY; r++)
                    if (r < rr) y[r] = d[b5 + r0 + r];
            } else if (P.lcon == CON_POLY_DEV) {
                mv(y, tab(P.lGT, l) + r0, rr, nx, P.con5, [&](int cc) { return d[b5 + cc]; });
            }
            if (lin) {  // q sum_l (src/costs.cuh:289-292)
                const T sl = d[vi];
                CP<T> lq = tab(P.lq, l) + r0;
#pragma unroll
                for (int r = 0; r < NY; r++)
                    if (r < rr) y[r] += lq[r] * sl;
            } else {
                mv(y, tab(P.lsQ, l) + r0, rr, nx, nx, [&](int cc) { return d[vi + cc]; });
            }
            if (P.lcost == COST_QUADLIN_DEV) {
                CP<T> lq = tab(P.lq, l) + r0;
#pragma unroll
                for (int r = 0; r < NY; r++)
                    if (r < rr) y[r] = y[r] - T(0.5) * lq[r] * a - T(0.5) * lq[r] * bb;
            }
            each(y, rr, [&](int r, T v) { emit(0, r0 + r, v); });
        }
        emit(3, 0, lin ? d[P.oVI + P.nlf + l] : (a + bb) * T(0.5));
    }
    if (P.ncost == COST_LIN_DEV) {
        emit(4, 0, d[P.oIV + P.N + n]);  // t <- set, every node (src/costs.cuh:286)
    } else if (n >= 1) {
        const long iv = P.oIV + (long) n * d4;
        emit(4, 0, (d[iv + nxu] + d[iv + nxu + 1]) * T(0.5));
    } else {
        emit(4, 0, T(0));
    }
}

// ------------------------------------------------------------------ L per node
// src/operator.cuh:41-60; src/costs.cuh:126-160, :379-416; src/constraints.cuh:158-169, :251-263.
// Handler: elemI(n, k, v), elemII(n, v), elemBox(e, v, lb, ub) for III and V, and for the SOC
// blocks IV and VI socBegin(), socPart(e, v, trans) for all but the last two entries, then
// socEnd(base, dim, h, trans) for the two entries that both carry h.
template <typename T, typename PV, typename H>
__device__ __forceinline__ void opNode(LP<T> P, int n, PV pv, H &h) {
    constexpr int NY = RowChunk<T>::v;
    const int nx = P.nx, nu = P.nu, nxu = nx + nu;
    T y[NY];
    if (n < P.nl) {
        CP<T> bn = tab(P.b, n);
        T by = T(0);
        for (int k = 0; k < P.yD; k++) {
            const T yk = pv(P.oY + (long) n * P.yD + k);
            by += bn[k] * yk;
            h.elemI(n, k, yk);
        }
        h.elemII(n, pv(P.oS + n) - by);
        const long base = P.oIII + (long) n * P.con3;
        if (P.ncon == CON_RECT_DEV) {
            CP<T> lb = tab(P.nLB, n), ub = tab(P.nUB, n);
            for (int k = 0; k < nx; k++) h.elemBox(base + k, pv(P.oX + (long) n * nx + k), lb[k], ub[k]);
            for (int k = 0; k < nu; k++) h.elemBox(base + nx + k, pv(P.oU + (long) n * nu + k), lb[nx + k], ub[nx + k]);
        } else if (P.ncon == CON_POLY_DEV) {
            CP<T> G = tab(P.nG, n), lb = tab(P.nLB, n), ub = tab(P.nUB, n);
            for (int r0 = 0; r0 < P.con3; r0 += NY) {
                const int rr = min(NY, P.con3 - r0);
                zeroY(y);
                mv(y, G + r0, rr, P.con3, nx, [&](int cc) { return pv(P.oX + (long) n * nx + cc); });
                mv(y, G + (size_t) P.con3 * nx + r0, rr, P.con3, nu, [&](int cc) { return pv(P.oU + (long) n * nu + cc); });
                each(y, rr, [&](int r, T v) { h.elemBox(base + r0 + r, v, lb[r0 + r], ub[r0 + r]); });
            }
        }
    }
    if (P.ncost == COST_LIN_DEV) {  // IV: (q_n' x_anc + r_n' u_anc, t_n), node 0 (0, t_0) (src/costs.cuh:261-267)
        T sum = T(0);
        if (n >= 1) {
            const int a = cst(P.anc)[n];
            CP<T> qn = tab(P.q, n), rn = tab(P.r, n);
            T pq = T(0), pr = T(0);
            for (int k = 0; k < nx; k++) pq += qn[k] * pv(P.oX + (long) a * nx + k);
            for (int k = 0; k < nu; k++) pr += rn[k] * pv(P.oU + (long) a * nu + k);
            sum = pq + pr;
        }
        h.lin(P.oIV + n, P.oIV + P.N + n, sum, pv(P.oT + n));
    } else {  // IV
        const long base = P.oIV + (long) n * (nxu + 2);
        CP<T> tr = tab(P.trans, n);
        h.socBegin();
        if (n >= 1) {
            const int a = cst(P.anc)[n];
            for (int r0 = 0; r0 < nx; r0 += NY) {
                const int rr = min(NY, nx - r0);
                zeroY(y);
                mv(y, tab(P.sQ, n) + r0, rr, nx, nx, [&](int cc) { return pv(P.oX + (long) a * nx + cc); });
                each(y, rr, [&](int r, T v) { h.socPart(base + r0 + r, v, tr[r0 + r]); });
            }
            for (int r0 = 0; r0 < nu; r0 += NY) {
                const int rr = min(NY, nu - r0);
                zeroY(y);
                mv(y, tab(P.sR, n) + r0, rr, nu, nu, [&](int cc) { return pv(P.oU + (long) a * nu + cc); });
                each(y, rr, [&](int r, T v) { h.socPart(base + nx + r0 + r, v, tr[nx + r0 + r]); });
            }
            T hv = pv(P.oT + n) * T(0.5);
            if (P.ncost == COST_QUADLIN_DEV) {
                CP<T> qn = tab(P.q, n), rn = tab(P.r, n);
                T pq = T(0), pr = T(0);
                for (int k = 0; k < nx; k++) pq += qn[k] * pv(P.oX + (long) a * nx + k);
                for (int k = 0; k < nu; k++) pr += rn[k] * pv(P.oU + (long) a * nu + k);
                hv = hv + T(-0.5) * pq;
                hv = hv + T(-0.5) * pr;
            }
            h.socEnd(base, nxu + 2, hv, tr);
        } else {
            for (int k = 0; k < nxu; k++) h.socPart(base + k, T(0), tr[k]);
            h.socEnd(base, nxu + 2, T(0), tr);
        }
    }
    if (n >= P.nl) {  // V and VI
        const int l = n - P.nl;
        const long b5 = P.oV + (long) l * P.con5;
        if (P.lcon == CON_RECT_DEV) {
            CP<T> lb = tab(P.lLB, l), ub = tab(P.lUB, l);
            for (int k = 0; k < nx; k++) h.elemBox(b5 + k, pv(P.oX + (long) n * nx + k), lb[k], ub[k]);
        } else if (P.lcon == CON_POLY_DEV) {
            CP<T> lb = tab(P.lLB, l), ub = tab(P.lUB, l);
            for (int r0 = 0; r0 < P.con5; r0 += NY) {
                const int rr = min(NY, P.con5 - r0);
                zeroY(y);
                mv(y, tab(P.lG, l) + r0, rr, P.con5, nx, [&](int cc) { return pv(P.oX + (long) n * nx + cc); });
                each(y, rr, [&](int r, T v) { h.elemBox(b5 + r0 + r, v, lb[r0 + r], ub[r0 + r]); });
            }
        }
        if (P.lcost == COST_LIN_DEV) {  // VI: (q' x_l, s_l) (src/costs.cuh:269-272)
            CP<T> lq = tab(P.lq, l);
            T pq = T(0);
            for (int k = 0; k < nx; k++) pq += lq[k] * pv(P.oX + (long) n * nx + k);
            h.lin(P.oVI + l, P.oVI + P.nlf + l, pq, pv(P.oS + n));
            return;
        }
        const long base = P.oVI + (long) l * (nx + 2);
        CP<T> tr = tab(P.ltrans, l);
        h.socBegin();
        for (int r0 = 0; r0 < nx; r0 += NY) {
            const int rr = min(NY, nx - r0);
            zeroY(y);
            mv(y, tab(P.lsQ, l) + r0, rr, nx, nx, [&](int cc) { return pv(P.oX + (long) n * nx + cc); });
            each(y, rr, [&](int r, T v) { h.socPart(base + r0 + r, v, tr[r0 + r]); });
        }
        T hs = pv(P.oS + n) * T(0.5);
        if (P.lcost == COST_QUADLIN_DEV) {
            CP<T> lq = tab(P.lq, l);
            T pq = T(0);
            for (int k = 0; k < nx; k++) pq += lq[k] * pv(P.oX + (long) n * nx + k);
            hs = hs + T(-0.5) * pq;
        }
        h.socEnd(base, nx + 2, hs, tr);
    }
}

// Dual prox of the CP step (src/cache.cuh:770-789): w = (L gamma + z_d)/gamma + c and
// cand_d = gamma (w - Pi(w)), Pi the masked NNOC (src/risks.cuh:99-106), NNOC (src/cones.cuh:134-137),
// box (src/constraints.cuh:70-74) or SOC (src/projections.cuh:102-128).  SOC blocks stage w in
// cand_d, accumulate |w|^2 and rescale in socEnd.
template <typename T>
struct Prox {
    LP<T> P;
    LV<T> zd, cd;
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
    __device__ __forceinline__ void socEnd(long base, int dim, T hv, CP<T> tr) {
        const long k1 = base + dim - 2, k2 = base + dim - 1;
        const T w1 = wOf(hv, k1, tr[dim - 2]);
        const T w2 = wOf(hv, k2, tr[dim - 1]);
        const T sq = ss + w1 * w1;
        const T nrm = sqrt(sq), tl = w2;
        const int i1 = nrm <= tl;
        const int i2 = (nrm <= -tl) && !i1;
        const int i3 = !i1 && !i2;
        const T sc = (nrm + tl) / (T(2) * nrm + T(1 - i3));
        const T mult = T(1 - i2) * (T(i3) * sc + T(1 - i3));
        for (int k = 0; k < dim - 2; k++) {
            const T w = cd[base + k];
            cd[base + k] = (w - w * mult) * P.gamma;
        }
        const T p1 = w1 * mult;
        const int i1b = (1 - i2) * (1 - i3);
        const T p2 = T(i1b) * (w2 * T(1 - i2)) + T((1 - i1b) * (1 - i2)) * (T(i3) * (sc * nrm - T(1)) + T(1));
        cd[k1] = (w1 - p1) * P.gamma;
        cd[k2] = (w2 - p2) * P.gamma;
    }
    __device__ __forceinline__ void lin(long eSum, long eSet, T sum, T set) {
        const T ws = wOf(sum, eSum, T(0)), wt = wOf(set, eSet, T(0));
        cd[eSum] = (ws - projBelow(ws, wt)) * P.gamma;
        cd[eSet] = (wt - wt) * P.gamma;
    }
};

// ell(y) = [L* y_d; L y_p] reductions: sum x.(y - gamma ell) and max |y/gamma - ell|
// (src/cache.cuh:722-734, 917-931).
template <typename T>
struct Red {
    LP<T> P;
    LV<T> x, y;
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
    __device__ __forceinline__ void socEnd(long base, int dim, T hv, CP<T>) {
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
    LV<T> out;  // dual part
    __device__ __forceinline__ void elemI(int n, int k, T v) { out[P.oI + (long) n * P.yD + k] = v; }
    __device__ __forceinline__ void elemII(int n, T v) { out[P.oII + n] = v; }
    __device__ __forceinline__ void elemBox(long e, T v, T, T) { out[e] = v; }
    __device__ __forceinline__ void socBegin() {}
    __device__ __forceinline__ void socPart(long e, T v, T) { out[e] = v; }
    __device__ __forceinline__ void socEnd(long base, int dim, T hv, CP<T>) {
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
__device__ __forceinline__ void phaseAdj(LP<T> P, const Ctx &c, const LV<T> &z, const LV<T> &cand,
                                         const LV<T> &qd) {
    const T g = -P.gamma;
    const LV<T> zd = z.shift(P.sPrim);
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

// Backward sweep step for node i (src/dynamics.cuh:85-139) with R~^-1 applied as one matvec.
template <typename T>
__device__ __forceinline__ void dynBackward(LP<T> P, const LV<T> &cand, const LV<T> &qd,
                                            const LV<T> &dd, int i, int stage) {
    constexpr int NY = RowChunk<T>::v;
    const int nx = P.nx, nu = P.nu;  // nu <= 32 (checked on the host)
    const int c0 = cst(P.chFrom)[i], nc = cst(P.numCh)[i];
    T y0[NY], y1[NY];
    // affine dynamics: the children's q enter as q_j + P_j e_j (src/dynamics.cuh:93-97)
    const bool affine = P.affine != 0;
    auto qv = [&](int j, int cc) { return affine ? qd[(long) j * nx + cc] + tab(P.Pe, j)[cc] : qd[(long) j * nx + cc]; };
    // e = ubar - sum_j B_j' q_j, stored in dd, then d = R~^-1 e in place
    zeroY(y0);
    zeroY(y1);
    for (int ci = 0; ci < nc; ci++) {
        const int j = c0 + ci;
        mv32(y0, y1, tab(P.BTr, j), nu, nu, nx, [&](int cc) { return qv(j, cc); });
    }
    auto storeE = [&](int r, T v) { dd[(long) i * nu + r] = -v + cand[P.oU + (long) i * nu + r]; };
    each(y0, min(nu, 16), storeE);
    each(y1, nu - 16, [&](int r, T v) { storeE(16 + r, v); });
    zeroY(y0);
    zeroY(y1);
    mv32(y0, y1, tab(P.RtInv, i), nu, nu, nu, [&](int cc) { return dd[(long) i * nu + cc]; });
    each(y0, min(nu, 16), [&](int r, T v) { dd[(long) i * nu + r] = v; });
    each(y1, nu - 16, [&](int r, T v) { dd[(long) i * nu + 16 + r] = v; });
    if (stage == 0) return;
    // q_i = sum_j (APB_j d_i + ABKTr_j q_j) + K_i'(d_i - u_i) - x_i
    T y[NY];
    for (int r0 = 0; r0 < nx; r0 += NY) {
        const int rr = min(NY, nx - r0);
        zeroY(y);
        for (int ci = 0; ci < nc; ci++) {
            const int j = c0 + ci;
            mv(y, tab(P.APB, j) + r0, rr, nx, nu, [&](int cc) { return dd[(long) i * nu + cc]; });
            mv(y, tab(P.ABKTr, j) + r0, rr, nx, nx, [&](int cc) { return qv(j, cc); });
        }
        mv(y, tab(P.KTr, i) + r0, rr, nx, nu,
           [&](int cc) { return dd[(long) i * nu + cc] - cand[P.oU + (long) i * nu + cc]; });
        each(y, rr, [&](int r, T v) { qd[(long) i * nx + r0 + r] = v - cand[P.oX + (long) i * nx + r0 + r]; });
    }
}

// Kernel projection for nonleaf node i (src/cache.cuh:650-676); padded child slots enter as 0.
template <typename T>
__device__ __forceinline__ void kernelProj(LP<T> P, const LV<T> &cand, int i) {
    constexpr int NY = RowChunk<T>::v;
    const int yD = P.yD, E = P.E, nc = cst(P.numCh)[i], c0 = cst(P.chFrom)[i];
    auto vin = [&](int k) -> T {
        if (k < yD) return cand[P.oY + (long) i * yD + k];
        if (k < yD + E) return (k - yD < nc) ? cand[P.oT + c0 + (k - yD)] : T(0);
        return (k - yD - E < nc) ? cand[P.oS + c0 + (k - yD - E)] : T(0);
    };
    T y0[NY], y1[NY];
    zeroY(y0);
    zeroY(y1);
    mv32(y0, y1, tab(P.NN, i), P.nD, P.nD, P.nD, vin);  // nD = 4E+1 <= 32 (checked on the host)
    auto put = [&](int k, T v) {
        if (k < yD) cand[P.oY + (long) i * yD + k] = v;
        else if (k < yD + E) {
            if (k - yD < nc) cand[P.oT + c0 + (k - yD)] = v;
        } else if (k - yD - E < nc) {
            cand[P.oS + c0 + (k - yD - E)] = v;
        }
    };
    each(y0, min(P.nD, 16), put);
    each(y1, P.nD - 16, [&](int k, T v) { put(16 + k, v); });
}

// Forward rollout for node i (src/dynamics.cuh:144-173); affine dynamics add e_j to each child.
template <typename T>
__device__ __forceinline__ void dynForward(LP<T> P, const LV<T> &cand, const LV<T> &x0s,
                                           const LV<T> &dd, int i) {
    constexpr int NY = RowChunk<T>::v;
    const int nx = P.nx, nu = P.nu;
    if (i == 0)
        for (int k = 0; k < nx; k++) cand[P.oX + k] = x0s[k];
    T y[NY];
    for (int r0 = 0; r0 < nu; r0 += NY) {
        const int rr = min(NY, nu - r0);
        zeroY(y);
        mv(y, tab(P.K, i) + r0, rr, nu, nx, [&](int cc) { return cand[P.oX + (long) i * nx + cc]; });
        each(y, rr, [&](int r, T v) { cand[P.oU + (long) i * nu + r0 + r] = v + dd[(long) i * nu + r0 + r]; });
    }
    const int c0 = cst(P.chFrom)[i], nc = cst(P.numCh)[i];
    for (int ci = 0; ci < nc; ci++) {
        const int j = c0 + ci;
        CP<T> AB = tab(P.AB, j);
        for (int r0 = 0; r0 < nx; r0 += NY) {
            const int rr = min(NY, nx - r0);
            zeroY(y);
            mv(y, AB + r0, rr, nx, nx, [&](int cc) { return cand[P.oX + (long) i * nx + cc]; });
            mv(y, AB + (size_t) nx * nx + r0, rr, nx, nu, [&](int cc) { return cand[P.oU + (long) i * nu + cc]; });
            if (P.affine) {
                CP<T> ej = tab(P.e, j) + r0;
                each(y, rr, [&](int r, T v) { cand[P.oX + (long) j * nx + r0 + r] = v + ej[r]; });
            } else {
                each(y, rr, [&](int r, T v) { cand[P.oX + (long) j * nx + r0 + r] = v; });
            }
        }
    }
}

template <typename T>
__device__ __forceinline__ void phaseProj(LP<T> P, const Ctx &c, const LV<T> &cand, const LV<T> &qd,
                                          const LV<T> &dd, const LV<T> &x0s, bool doDyn, bool doKernel) {
    if (doDyn) {
        for (int stage = P.S - 2; stage >= 0; stage--) {
            const int f = cst(P.stageFrom)[stage], nStage = cst(P.stageTo)[stage] - f + 1;
            const int extra = (doKernel && stage == P.S - 2) ? P.nl : 0;
            for (int task = c.wave; task < nStage + extra; task += c.nw) {
                if (task < nStage) dynBackward(P, cand, qd, dd, f + task, stage);
                else kernelProj(P, cand, task - nStage);
            }
            __syncthreads();
        }
        for (int stage = 0; stage <= P.S - 2; stage++) {
            const int f = cst(P.stageFrom)[stage], nStage = cst(P.stageTo)[stage] - f + 1;
            for (int task = c.wave; task < nStage; task += c.nw) dynForward(P, cand, x0s, dd, f + task);
            __syncthreads();
        }
    } else if (doKernel) {
        for (int task = c.wave; task < P.nl; task += c.nw) kernelProj(P, cand, task);
        __syncthreads();
    }
}

// modifyDual + proximalDual (src/cache.cuh:770-789) on 2 xbar - z_p.
template <typename T>
__device__ __forceinline__ void phaseDual(LP<T> P, const Ctx &c, const LV<T> &z, const LV<T> &cand) {
    Prox<T> h{P, z.shift(P.sPrim), cand.shift(P.sPrim), T(0)};
    auto pv = [&](long e) { return cand[e] * T(2) - z[e]; };
    for (int n = c.wave; n < P.N; n += c.nw) opNode(P, n, pv, h);
}


// The CP operator T (src/cache.cuh:796-801): cand = T(z).
template <typename T>
__device__ __noinline__ void applyT(const LaneInst<T> *Pp, T *z, T *cand, long *phaseAcc) {
    LaneClock clk{wall_clock64(), phaseAcc};
    LP<T> P = *cst(uniformPtr(Pp));
    z = uniformPtr(z);
    cand = uniformPtr(cand);
    const Ctx c = ctx();
    const LV<T> zv = lview(z, P.sIter, c.lane), cv = lview(cand, P.sIter, c.lane);
    const LV<T> qd = lview(P.qd, (long) P.nx * P.N, c.lane), dd = lview(P.dd, (long) P.nu * P.nl, c.lane);
    const LV<T> x0s = lview(const_cast<T *>(P.x0s), P.nx, c.lane);
    phaseAdj(P, c, zv, cv, qd);
    __syncthreads();
    clk.mark(PH_ADJ);
    phaseProj(P, c, cv, qd, dd, x0s, true, true);
    clk.mark(PH_PROJ);
    phaseDual(P, c, zv, cv);
    __syncthreads();
    clk.mark(PH_DUAL);
}

// Per-lane combination of per-wave partials through LDS, in wave order.
__device__ __forceinline__ double laneSum(double v, double *lds, const Ctx &c) {
    __syncthreads();
    lds[c.wave * 64 + c.lane] = v;
    __syncthreads();
    double s = 0.;
    for (int w = 0; w < c.nw; w++) s += lds[w * 64 + c.lane];
    return s;
}

__device__ __forceinline__ double laneMax(double v, double *lds, const Ctx &c) {
    __syncthreads();
    lds[c.wave * 64 + c.lane] = v;
    __syncthreads();
    double s = 0.;
    for (int w = 0; w < c.nw; w++) s = fmax(s, lds[w * 64 + c.lane]);
    return s;
}


// sum x.(y - gamma ell(y)) and max |y/gamma - ell(y)| per lane.
template <typename T>
__device__ __noinline__ DotMax<T> reduceEll(const LaneInst<T> *Pp, T *x, T *y, double *lds) {
    LP<T> P = *cst(uniformPtr(Pp));
    x = uniformPtr(x);
    y = uniformPtr(y);
    const Ctx c = ctx();
    const LV<T> xv = lview(x, P.sIter, c.lane), yv = lview(y, P.sIter, c.lane);
    const LV<T> yd = yv.shift(P.sPrim);
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
    r.dot = (T) laneSum((double) (dot + h.dot), lds, c);
    r.mx = (T) laneMax((double) fmax(mx, h.mx), lds, c);
    r.cross = (T) laneSum((double) cross, lds, c);
    return r;
}


// The adjoint-only terms of <x, M x> and <x, M y> (M = [[I, -gamma L*], [-gamma L, I]], and
// x_d' L y_p = (L* x_d)' y_p):  c1 = sum_p x_p . (L* x_d),  c2 = sum_p x_p . (L* y_d) + (L* x_d) . y_p.
template <typename T>
__device__ __noinline__ Cross2<T> adjDots(const LaneInst<T> *Pp, T *x, T *y, double *lds, bool withY) {
    LP<T> P = *cst(uniformPtr(Pp));
    x = uniformPtr(x);
    y = uniformPtr(y);
    const Ctx c = ctx();
    const LV<T> xv = lview(x, P.sIter, c.lane), yv = lview(y, P.sIter, c.lane);
    const LV<T> xd = xv.shift(P.sPrim), yd = yv.shift(P.sPrim);
    T c1 = T(0), c2 = T(0);
    for (int n = c.wave; n < P.N; n += c.nw) {
        adjNode(P, xd, n, [&](int kind, int k, T v) {
            const long e = primIndex(P, n, kind, k);
            c1 += xv[e] * v;
            if (withY) c2 += v * yv[e];
        });
        if (withY)
            adjNode(P, yd, n, [&](int kind, int k, T v) { c2 += xv[primIndex(P, n, kind, k)] * v; });
    }
    Cross2<T> r;
    r.c1 = (T) laneSum((double) c1, lds, c);
    r.c2 = withY ? (T) laneSum((double) c2, lds, c) : T(0);
    return r;
}

// Iterate-wide pass: wave w takes a contiguous range of elements, U loads in flight per lane.
template <int U, typename Ld, typename St>
__device__ __forceinline__ void pass(long n, const Ctx &c, Ld ld, St st) {
    using Vals = decltype(ld(0L));
    const long per = (n + c.nw - 1) / c.nw;
    const long e0 = (long) c.wave * per, e1 = min(n, e0 + per);
    long e = e0;
    for (; e + U <= e1; e += U) {
        Vals v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = ld(e + u);
#pragma unroll
        for (int u = 0; u < U; u++) st(e + u, v[u]);
    }
    for (; e < e1; e++) st(e, ld(e));
}




// The lane engine as a policy of the lockstep solve: lane = instance, 64 per wave.
struct Policy {
    template <typename T>
    using View = LV<T>;
    __device__ static Ctx ctx() { return lane::ctx(); }
    template <typename T>
    __device__ static LV<T> view(T *region, long rows, const Ctx &c) { return lview(region, rows, c.lane); }
    template <int U, typename Ld, typename St>
    __device__ static void pass(long n, const Ctx &c, Ld ld, St st) { lane::pass<U>(n, c, ld, st); }
    __device__ static double sum(double v, double *lds, const Ctx &c) { return laneSum(v, lds, c); }
    __device__ static double max(double v, double *lds, const Ctx &c) { return laneMax(v, lds, c); }
    template <typename T>
    __device__ static void applyT(const LaneInst<T> *P, T *z, T *cand, long *phaseAcc) { lane::applyT(P, z, cand, phaseAcc); }
    template <typename T>
    __device__ static DotMax<T> reduceEll(const LaneInst<T> *P, T *x, T *y, double *lds) { return lane::reduceEll(P, x, y, lds); }
    template <typename T>
    __device__ static Cross2<T> adjDots(const LaneInst<T> *P, T *x, T *y, double *lds, bool withY) {
        return lane::adjDots(P, x, y, lds, withY);
    }
};

}  // namespace lane

template <typename T>
__global__ void __launch_bounds__(SPK_LANE_MAX_THREADS) SPK_LANE_KERNEL(k_lane_solve)(const LaneInst<T> *__restrict__ P, Out *outs) {
    __shared__ double lds[16 * 64];
    __shared__ long phaseAcc[NUM_PHASES];
    lockstepSolve<T, lane::Policy>(P, outs, lds, phaseAcc);
}

// Single operators on every instance (parity tests): z = zio, cand = work[1].
template <typename T>
__global__ void __launch_bounds__(SPK_LANE_MAX_THREADS) SPK_LANE_KERNEL(k_lane_debug)(const LaneInst<T> *__restrict__ Pp,
                                                                                      Out *outs, int op) {
    using namespace lane;
    __shared__ double lds[16 * 64];
    LP<T> P = *cst(Pp);
    const Ctx c = ctx();
    T *z = P.zio, *cand = P.work[1];
    const LV<T> zv = lview(z, P.sIter, c.lane), cv = lview(cand, P.sIter, c.lane);
    if (op == 0) {  // L
        Store<T> h{P, cv.shift(P.sPrim)};
        auto pv = [&](long e) { return zv[e]; };
        for (int n = c.wave; n < P.N; n += c.nw) opNode(P, n, pv, h);
    } else if (op == 1) {  // L*
        const LV<T> zd = zv.shift(P.sPrim);
        for (int n = c.wave; n < P.N; n += c.nw)
            adjNode(P, zd, n, [&](int kind, int k, T v) { cv[primIndex(P, n, kind, k)] = v; });
    } else if (op == 2 || op == 3) {  // dynamics / kernel projection of the primal part
        pass<16>(P.sPrim, c, [&](long e) { return zv[e]; }, [&](long e, T v) { cv[e] = v; });
        __syncthreads();
        const LV<T> qd = lview(P.qd, (long) P.nx * P.N, c.lane), dd = lview(P.dd, (long) P.nu * P.nl, c.lane);
        const LV<T> x0s = lview(const_cast<T *>(P.x0s), P.nx, c.lane);
        if (op == 2)
            for (int n = c.wave; n < P.N; n += c.nw)
                if (n >= P.nl)
                    for (int r = 0; r < P.nx; r++) qd[(long) n * P.nx + r] = -cv[P.oX + (long) n * P.nx + r];
        __syncthreads();
        phaseProj(P, c, cv, qd, dd, x0s, op == 2, op == 3);
    } else if (op == 4) {  // T
        __shared__ long phaseAcc[NUM_PHASES];
        applyT(Pp, z, cand, phaseAcc);
    } else if (op == 5) {  // ell reductions
        const DotMax<T> r = reduceEll(Pp, z, cand, lds);
        if (c.wave == 0 && c.gl < P.count) {
            outs[c.gl].w = (double) r.dot;
            outs[c.gl].err = (double) r.mx;
        }
    }
}

template __global__ void SPK_LANE_KERNEL(k_lane_solve)<float>(const LaneInst<float> *, Out *);
template __global__ void SPK_LANE_KERNEL(k_lane_solve)<double>(const LaneInst<double> *, Out *);
template __global__ void SPK_LANE_KERNEL(k_lane_debug)<float>(const LaneInst<float> *, Out *, int);
template __global__ void SPK_LANE_KERNEL(k_lane_debug)<double>(const LaneInst<double> *, Out *, int);

#if SPK_MV_XB == 0
// [instance][element] rows (leading dimension ld) to the block-interleaved layout of a region of
// n rows and back, with dtype conversion.
template <typename S, typename D>
__global__ void k_interleave(const S *__restrict__ in, long ld, D *__restrict__ out, long n, int count, long w) {
    const long total = n * (long) count;
    for (long t = blockIdx.x * (long) blockDim.x + threadIdx.x; t < total; t += (long) gridDim.x * blockDim.x) {
        const long b = t % count, e = t / count;
        out[laneIndex(n, e, b, w)] = (D) in[b * ld + e];
    }
}

template <typename S, typename D>
__global__ void k_deinterleave(const S *__restrict__ in, long w, D *__restrict__ out, long ld, long n, int count) {
    const long total = n * (long) count;
    for (long t = blockIdx.x * (long) blockDim.x + threadIdx.x; t < total; t += (long) gridDim.x * blockDim.x) {
        const long e = t % n, b = t / n;
        out[b * ld + e] = (D) in[laneIndex(n, e, b, w)];
    }
}

// A region of n rows from blocks of win instances to blocks of wout instances (all padded slots).
template <typename T>
__global__ void k_relayout(const T *__restrict__ in, long win, T *__restrict__ out, long wout, long n, long B) {
    const long total = n * B;
    for (long t = blockIdx.x * (long) blockDim.x + threadIdx.x; t < total; t += (long) gridDim.x * blockDim.x) {
        const long b = t % B, e = t / B;
        out[laneIndex(n, e, b, wout)] = in[laneIndex(n, e, b, win)];
    }
}

template __global__ void k_relayout<float>(const float *, long, float *, long, long, long);
template __global__ void k_relayout<double>(const double *, long, double *, long, long, long);

// Starts new solves in lanes of a resumable (chunked) solve of a shared batch with blocks of w
// instances: lane lanes[i] gets the scaled initial state x0s[i] (nx values) and the starting
// iterate z[i] (n values, leading dimension ldz; z == null: the zero iterate of a cold start), with
// the previous iterate and residual that initialisePrev sets (src/cache.cuh:611-632; a warm start
// takes the iterate and a zero residual, a cold start ones), fresh SuperMann state, and an
// iteration count that starts at its workgroup's next iteration.  The starting iterate goes to the
// buffer the next iteration reads as z: the candidate role, which that iteration accepts, or the
// iterate role itself if the workgroup has not run yet.
template <typename T, typename Z>
__global__ void k_lane_refill(const LaneInst<T> *__restrict__ Pp, const int *__restrict__ lanes, int count,
                              const Z *__restrict__ z, long ldz, const T *__restrict__ x0s, long w) {
    const LaneInst<T> &P = *Pp;
    const long n = P.sIter;
    const long stride = (long) gridDim.x * blockDim.x;
    const long tid = blockIdx.x * (long) blockDim.x + threadIdx.x;
    for (long t = tid; t < n * count; t += stride) {
        const long li = t / n, e = t % n;
        const long gl = lanes[li];
        const LaneGroupSave &gs = P.gsave[gl / w];
        const T v = z ? (T) z[li * ldz + e] : T(0);
        const size_t at = laneIndex(n, e, gl, w);
        P.work[gs.it == 0 ? gs.roles[0] : gs.roles[1]][at] = v;
        P.work[gs.roles[2]][at] = z ? v : T(1);
        P.work[gs.roles[4]][at] = z ? T(0) : T(1);
        P.zio[at] = v;
        P.rio[at] = T(0);
    }
    T *x0 = const_cast<T *>(P.x0s);
    for (long t = tid; t < (long) P.nx * count; t += stride) {
        const long li = t / P.nx, e = t % P.nx;
        x0[laneIndex(P.nx, e, lanes[li], w)] = x0s[li * P.nx + e];
    }
    for (long li = tid; li < count; li += stride) {
        const long gl = lanes[li];
        LaneSave &sv = P.save[gl];
        for (int i = 0; i < 4; i++) {
            sv.gb[i] = 0.;
            for (int j = 0; j < 4; j++) sv.G[i][j] = 0.;
        }
        const double given = P.tolFixed ? (double) P.tolFixed[gl] : 0.;
        sv.tol = given > 0. ? given : 0.;
        sv.initFlag = given > 0. ? 0 : 1;
        sv.lastErr = sv.wLast = NAN;
        sv.zeta = sv.w = sv.wSafe = INFINITY;
        sv.iters = sv.k0 = sv.k1 = sv.k2 = sv.k2bt = sv.k3 = sv.nT = sv.nEll = 0;
        sv.it0 = P.gsave[gl / w].it;
        sv.status = ST_NOTRUN;
        sv.active = 1;
    }
}

// After the refill kernel: the refilled lanes' workgroups run on their next resume.
template <typename T>
__global__ void k_lane_refill_groups(const LaneInst<T> *__restrict__ Pp, const int *__restrict__ lanes, int count, long w) {
    const long stride = (long) gridDim.x * blockDim.x;
    for (long li = blockIdx.x * (long) blockDim.x + threadIdx.x; li < count; li += stride)
        Pp->gsave[lanes[li] / w].finished = 0;
}

template __global__ void k_lane_refill<float, float>(const LaneInst<float> *, const int *, int, const float *, long,
                                                      const float *, long);
template __global__ void k_lane_refill<double, float>(const LaneInst<double> *, const int *, int, const float *, long,
                                                       const double *, long);
template __global__ void k_lane_refill<double, double>(const LaneInst<double> *, const int *, int, const double *,
                                                        long, const double *, long);
template __global__ void k_lane_refill<float, double>(const LaneInst<float> *, const int *, int, const double *, long,
                                                       const float *, long);
// The accepted iterates of the listed lanes (blocks of w instances) as rows of leading dimension ld.
template <typename S, typename D>
__global__ void k_gather_lanes(const S *__restrict__ in, long w, D *__restrict__ out, long ld, long n,
                               const int *__restrict__ lanes, int count) {
    const long total = n * (long) count;
    for (long t = blockIdx.x * (long) blockDim.x + threadIdx.x; t < total; t += (long) gridDim.x * blockDim.x) {
        const long li = t / n, e = t % n;
        out[li * ld + e] = (D) in[laneIndex(n, e, lanes[li], w)];
    }
}

template __global__ void k_gather_lanes<float, float>(const float *, long, float *, long, long, const int *, int);
template __global__ void k_gather_lanes<double, double>(const double *, long, double *, long, long, const int *, int);
template __global__ void k_gather_lanes<double, float>(const double *, long, float *, long, long, const int *, int);
template __global__ void k_lane_refill_groups<float>(const LaneInst<float> *, const int *, int, long);
template __global__ void k_lane_refill_groups<double>(const LaneInst<double> *, const int *, int, long);

template __global__ void k_interleave<float, float>(const float *, long, float *, long, int, long);
template __global__ void k_interleave<float, double>(const float *, long, double *, long, int, long);
template __global__ void k_interleave<double, float>(const double *, long, float *, long, int, long);
template __global__ void k_interleave<double, double>(const double *, long, double *, long, int, long);
template __global__ void k_deinterleave<float, float>(const float *, long, float *, long, long, int);
template __global__ void k_deinterleave<float, double>(const float *, long, double *, long, long, int);
template __global__ void k_deinterleave<double, float>(const double *, long, float *, long, long, int);
template __global__ void k_deinterleave<double, double>(const double *, long, double *, long, long, int);
#endif

}  // namespace spk
