// The lockstep SuperMann solve of many instances of ONE problem (src/cache.cuh:1008-1132), shared
// by the engines of shared batches.  An engine policy E supplies the lane-to-instance mapping and
// the operators:
//   E::ctx()                 wave, nw, gl (this lane's instance) and writer (the lane that writes
//                            the instance's per-instance records; several lanes may serve one
//                            instance and then hold identical per-instance state);
//   E::view<T>(region, rows, c)  v[e] = element e of this lane's instance;
//   E::pass<U>(n, c, ld, st) an iterate-wide pass in which every (element, instance) pair is
//                            visited by exactly one lane;
//   E::sum / E::max          per-instance combination of per-lane partials, in a fixed order;
//   E::applyT, E::reduceEll, E::adjDots  the CP operator and the [L*; L] reductions.
// SuperMann runs in lockstep outer iterations with per-lane decisions: the buffer roles rotate
// uniformly where every lane takes the same step, and per-lane masked copies replace the rotations
// where lanes branch (K0 vs line search, K1/K2/K3).
#pragma once
#include "engine.hpp"

#include <hip/hip_runtime.h>
#include <math.h>

namespace spk {

// A pointer every lane holds the same value of, moved to scalar registers: loads through it
// become scalar loads (needed inside non-inlined functions, whose arguments arrive per lane).
template <typename T>
__device__ __forceinline__ T *uniformPtr(T *p) {
    const unsigned long long v = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned) v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned) (v >> 32));
    return reinterpret_cast<T *>(((unsigned long long) hi << 32) | lo);
}

// Constant address space: the descriptor, the tree arrays and the factor tables are never written
// while a solve runs, and wave-uniform reads through such pointers become scalar loads.
#define SPK_CONST __attribute__((address_space(4)))
#define SPK_GLOBAL __attribute__((address_space(1)))
template <typename T>
using CP = const SPK_CONST T *;
template <typename T>
using LP = const SPK_CONST LaneInst<T> &;

template <typename T>
__device__ __forceinline__ CP<T> cst(const T *p) {
    return (CP<T>) p;
}

// Node n's entry of a per-node table, as a constant-space pointer.
template <typename T>
__device__ __forceinline__ CP<T> tab(const SPK_CONST DTab<T> &t, int n) {
    return cst(t.data) + (size_t) cst(t.idx)[n] * t.rows * t.cols;
}

namespace ls {

constexpr int MAXM = 4;  // Anderson memory supported by the engines of shared batches

// Device wall-clock per phase, accumulated by thread 0 into LDS (uniform control flow, so
// thread 0's view is the workgroup's).
struct LaneClock {
    long last;
    long *acc;
    __device__ __forceinline__ void mark(int ph) {
        const long t = wall_clock64();
        if (threadIdx.x == 0) acc[ph] += t - last;
        last = t;
    }
    __device__ __forceinline__ void skip() { last = wall_clock64(); }
};

template <typename T>
struct DotMax {
    T dot, mx, cross;  // cross = sum over primal elements of x_p . (L* y_d)
};

template <typename T>
struct Cross2 {
    T c1, c2;
};

template <typename T>
struct V2 {
    T a, b;
};
template <typename T>
struct V4 {
    T a, b, c, d;
};
template <typename T, int M>
struct VH {  // two vector values and M history values of one element
    T a, b, h[M];
};
template <typename T, int M>
struct V4H {
    T a, b, c, d, h[M];
};

__device__ __forceinline__ bool anyLane(bool v) { return __syncthreads_or(v ? 1 : 0) != 0; }

// Least squares argmin ||Y g - r|| from the Gram matrix: Cholesky of G + 1e-12 max(diag) I in FP64.
__device__ __forceinline__ bool gramSolve(const double (&G)[MAXM][MAXM], const double (&gb)[MAXM], int m,
                                          double (&g)[MAXM]) {
    double Lm[MAXM][MAXM];
    double dmax = 0.;
#pragma unroll
    for (int i = 0; i < MAXM; i++)
        if (i < m) dmax = fmax(dmax, G[i][i]);
    if (!(dmax > 0.)) return false;
    const double ridge = 1e-12 * dmax;
    bool ok = true;
#pragma unroll
    for (int i = 0; i < MAXM; i++) {
#pragma unroll
        for (int j = 0; j < MAXM; j++) {
            Lm[i][j] = 0.;
            if (i < m && j <= i) {
                double v = G[i][j] + (i == j ? ridge : 0.);
#pragma unroll
                for (int k = 0; k < MAXM; k++)
                    if (k < j) v -= Lm[i][k] * Lm[j][k];
                if (i == j) {
                    ok = ok && (v > 0.);
                    Lm[i][i] = sqrt(fmax(v, 0.));
                } else {
                    Lm[i][j] = v / Lm[j][j];
                }
            }
        }
    }
    if (!ok) return false;
    double y[MAXM];
#pragma unroll
    for (int i = 0; i < MAXM; i++) {
        y[i] = 0.;
        if (i < m) {
            double v = gb[i];
#pragma unroll
            for (int k = 0; k < MAXM; k++)
                if (k < i) v -= Lm[i][k] * y[k];
            y[i] = v / Lm[i][i];
        }
    }
#pragma unroll
    for (int i = MAXM - 1; i >= 0; i--) {
        g[i] = 0.;
        if (i < m) {
            double v = y[i];
#pragma unroll
            for (int k = 0; k < MAXM; k++)
                if (k > i && k < m) v -= Lm[k][i] * g[k];
            g[i] = v / Lm[i][i];
        }
    }
    return true;
}

}  // namespace ls

// ---------------------------------------------------------------- the lockstep solve
template <typename T, typename E>
__device__ void lockstepSolve(const LaneInst<T> *Pp, Out *outs, double *lds, long *phaseAcc) {
    using namespace ls;
    using VW = typename E::template View<T>;
    LP<T> P = *cst(Pp);
    if (threadIdx.x < NUM_PHASES) phaseAcc[threadIdx.x] = 0;
    __syncthreads();
    LaneClock clk{wall_clock64(), phaseAcc};
    const auto c = E::ctx();
    const long n = P.sIter, B = P.B;
    const int gl = c.gl;
    const bool real = gl < P.count;
    const int m = P.andBuff;
    auto V = [&](T *p) { return E::template view<T>(p, n, c); };
    T *Z = P.work[0], *C = P.work[1], *ZP = P.work[2], *RR = P.work[3], *RP = P.work[4], *BK = P.work[5];
    T *C2 = P.work[6], *D = P.work[7], *DS = P.work[8];
    const long t0 = wall_clock64();
    const bool resume = P.resume != 0;
    LaneGroupSave *gs = P.gsave + blockIdx.x;
    long itStart = 0, ticksBefore = 0;
    int headSaved = m - 1;
    // a resume of a workgroup whose solve has ended leaves its results as they are
    if (resume && __builtin_amdgcn_readfirstlane(gs->finished)) return;
    if (resume) {  // roles of the six rotating buffers, ring head, next iteration (uniform)
        T **roles[6] = {&Z, &C, &ZP, &RR, &RP, &BK};
#pragma unroll
        for (int i = 0; i < 6; i++) *roles[i] = P.work[__builtin_amdgcn_readfirstlane(gs->roles[i])];
        headSaved = __builtin_amdgcn_readfirstlane(gs->head);
        const long itv = gs->it;
        itStart = ((long) __builtin_amdgcn_readfirstlane((int) (itv >> 32)) << 32) |
                  (unsigned) __builtin_amdgcn_readfirstlane((int) itv);
        ticksBefore = gs->ticks;
    }
    // initialiseIter / initialisePrev (src/cache.cuh:611-632)
    if (!resume) {
        const VW zio = V(P.zio), rio = V(P.rio), z = V(Z), zp = V(ZP), rp = V(RP);
        const bool spock = P.method == METHOD_SPOCK, warm = P.warm != 0;
        E::template pass<16>(
            n, c, [&](long e) { return V2<T>{zio[e], rio[e]}; },
            [&](long e, const V2<T> &v) {
                z[e] = v.a;
                if (spock) {
                    zp[e] = warm ? v.a : T(1);
                    rp[e] = warm ? v.b : T(1);
                }
            });
    }
    __syncthreads();
    double G[MAXM][MAXM], gb[MAXM];
#pragma unroll
    for (int i = 0; i < MAXM; i++) {
        gb[i] = 0.;
#pragma unroll
        for (int j = 0; j < MAXM; j++) G[i][j] = 0.;
    }
    int head = headSaved;
    bool active = real;
    const T given = (real && P.tolFixed) ? P.tolFixed[gl] : T(0);
    bool initFlag = !(given > T(0));
    int status = ST_NOTRUN;
    long iters = 0;
    T tol = initFlag ? T(0) : given, lastErr = T(NAN), wLast = T(NAN);
    const T c0 = T(0.99), c1 = T(0.99), c2 = T(0.99), beta = T(0.5), sigma = T(0.1), lambda = T(1);
    const T inf = T(INFINITY);
    T zeta = inf, w = inf, wSafe = inf;
    long k0 = 0, k1 = 0, k2 = 0, k2bt = 0, k3 = 0, nT = 0, nEll = 0;
    long it0 = 0;  // the workgroup iteration this lane's solve started at (later than 0 after a refill)
    if (resume) {  // this lane's SuperMann state
        const LaneSave &sv = P.save[gl];
#pragma unroll
        for (int i = 0; i < MAXM; i++) {
            gb[i] = sv.gb[i];
#pragma unroll
            for (int j = 0; j < MAXM; j++) G[i][j] = sv.G[i][j];
        }
        tol = (T) sv.tol;
        lastErr = (T) sv.lastErr;
        wLast = (T) sv.wLast;
        zeta = (T) sv.zeta;
        w = (T) sv.w;
        wSafe = (T) sv.wSafe;
        iters = sv.iters;
        k0 = sv.k0; k1 = sv.k1; k2 = sv.k2; k2bt = sv.k2bt; k3 = sv.k3; nT = sv.nT; nEll = sv.nEll;
        status = sv.status;
        active = sv.active != 0;
        initFlag = sv.initFlag != 0;
        it0 = sv.it0;
    }
    // the iteration cap is only tested at error checks (src/cache.cuh:937-940): a lane's last outer
    // iteration is the first check index at or beyond maxIters of its own count; without a cap a
    // lane still running after 2e6 iterations ends with its current iterate
    const long hardCap =
        P.maxIters > 0 ? (P.maxIters + P.checkPeriod - 1) / P.checkPeriod * P.checkPeriod + 1 : 2000000;
    long it;
    bool chunkExit = false;
    for (it = itStart;; it++) {
        const bool capNow = active && it - it0 >= hardCap;
        if (anyLane(capNow)) {  // the current iterate is the output
            const VW zio = V(P.zio), rio = V(P.rio), z = V(Z), rr = V(RR);
            E::template pass<16>(
                n, c, [&](long e) { return V2<T>{z[e], rr[e]}; },
                [&](long e, const V2<T> &v) {
                    if (capNow) {
                        zio[e] = v.a;
                        rio[e] = v.b;
                    }
                });
            __syncthreads();
        }
        if (capNow) {
            status = ST_OUTOFITER;
            iters = it - it0;
            active = false;
        }
        if (!anyLane(active)) break;
        // uniform: save and leave.  The time slice is thread 0's clock, made uniform by the
        // workgroup-wide vote, so every wave leaves at the same iteration.
        if (P.chunkIters > 0 &&
            (it >= itStart + P.chunkIters ||
             (P.chunkTicks > 0. && it > itStart &&
              anyLane(threadIdx.x == 0 && (double) (wall_clock64() - t0) >= P.chunkTicks)))) {
            chunkExit = true;
            break;
        }
        if (it != 0 && P.method == METHOD_SPOCK) {  // acceptCandidate: z <- cand
            T *t = Z;
            Z = C;
            C = t;
        }
        clk.mark(PH_OTHER);
        E::applyT(Pp, Z, C, phaseAcc);
        clk.skip();
        if (active) nT++;
        T sRR;
        const bool spock = P.method == METHOD_SPOCK;
        if (spock) head = (head + 1) % m;
        if (!spock) {  // r = z - T(z) and ||r||^2
            const VW rr = V(RR), z = V(Z), cc = V(C);
            T acc = T(0);
            E::template pass<16>(n, c, [&](long e) { return V2<T>{z[e], cc[e]}; }, [&](long e, const V2<T> &v) {
                const T d = v.a - v.b;
                rr[e] = d;
                acc += d * d;
            });
            sRR = (T) E::sum((double) acc, lds, c);
        } else {
            // r = z - T(z), ||r||^2 and, in the same pass, the new Anderson column
            // (src/cache.cuh:861-896) with its Gram entries: Y_head = r - r_prev,
            // S_head = (z - z_prev) - Y_head
            double acc[2 * MAXM];
#pragma unroll
            for (int q = 0; q < 2 * MAXM; q++) acc[q] = 0.;
            T accR = T(0);
            const VW Yh = V(P.hY + (long) head * n * B), Sh = V(P.hS + (long) head * n * B);
            const VW rr = V(RR), cc = V(C), rp = V(RP), z = V(Z), zp = V(ZP);
            VW Ys[MAXM];
#pragma unroll
            for (int s2 = 0; s2 < MAXM; s2++) Ys[s2] = V(P.hY + (long) (s2 < m ? s2 : 0) * n * B);
            E::template pass<4>(
                n, c,
                [&](long e) {
                    V4H<T, MAXM> v;
                    v.a = z[e];
                    v.b = cc[e];
                    v.c = rp[e];
                    v.d = zp[e];
#pragma unroll
                    for (int s2 = 0; s2 < MAXM; s2++) v.h[s2] = (s2 < m && s2 != head) ? Ys[s2][e] : T(0);
                    return v;
                },
                [&](long e, const V4H<T, MAXM> &v) {
                    const T r = v.a - v.b;
                    rr[e] = r;
                    accR += r * r;
                    const T dr = r - v.c;
                    Yh[e] = dr;
                    Sh[e] = (v.a - v.d) - dr;
                    const double drd = (double) dr, rd = (double) r;
#pragma unroll
                    for (int s2 = 0; s2 < MAXM; s2++) {
                        if (s2 < m) {
                            const double ys = (s2 == head) ? drd : (double) v.h[s2];
                            acc[s2] += drd * ys;
                            acc[MAXM + s2] += ys * rd;
                        }
                    }
                });
            sRR = (T) E::sum((double) accR, lds, c);
#pragma unroll
            for (int q = 0; q < 2 * MAXM; q++)
                if ((q & (MAXM - 1)) < m) acc[q] = E::sum(acc[q], lds, c);
#pragma unroll
            for (int s2 = 0; s2 < MAXM; s2++) {
                if (s2 < m) {
#pragma unroll
                    for (int hh = 0; hh < MAXM; hh++) {
                        if (hh == head) {
                            G[hh][s2] = acc[s2];
                            G[s2][hh] = acc[s2];
                        }
                    }
                    gb[s2] = acc[MAXM + s2];
                }
            }
        }
        clk.mark(PH_RES);
        // <r, M r> = ||r||^2 - 2 gamma r_p . L* r_d; the full [L*; L] only at the error checks
        const bool checkIt = (it - it0) % P.checkPeriod == 0;
        T mx = T(0), crossRR;
        if (anyLane(active && checkIt)) {
            const DotMax<T> dm = E::reduceEll(Pp, RR, RR, lds);
            mx = dm.mx;
            crossRR = dm.cross;
        } else {
            crossRR = E::adjDots(Pp, RR, RR, lds, false).c1;
        }
        const T dotRR = sRR - T(2) * P.gamma * crossRR;
        if (active) nEll++;
        clk.mark(PH_ELL);
        bool done = false;
        if (active && checkIt) {  // computeError (src/cache.cuh:910-961)
            if (initFlag) {
                tol = fmax(P.tolAbs, P.tolRel * mx);
                initFlag = false;
            } else {
                lastErr = mx;
                int code = ST_NOTRUN;
                if (mx <= tol) code = ST_CONVERGED;
                if (P.maxIters && it - it0 >= P.maxIters) code = ST_OUTOFITER;
                if (P.maxTicks > 0 && (it - it0) % 1000 == 0 &&
                    (double) (wall_clock64() - t0 + ticksBefore) >= P.maxTicks)
                    code = ST_OUTOFTIME;
                if (code != ST_NOTRUN) {
                    status = code;
                    iters = it - it0;
                    done = true;
                }
            }
        }
        if (anyLane(done)) {  // accepted candidate and its residual are the solve's output
            const VW zio = V(P.zio), rio = V(P.rio), cc = V(C), rr = V(RR);
            E::template pass<16>(
                n, c, [&](long e) { return V2<T>{cc[e], rr[e]}; },
                [&](long e, const V2<T> &v) {
                    if (done) {
                        zio[e] = v.a;
                        rio[e] = v.b;
                    }
                });
            __syncthreads();
        }
        if (done) active = false;
        if (P.method == METHOD_CP) {
            T *t = Z;
            Z = C;
            C = t;
            continue;
        }
        {  // backup (src/cache.cuh:804-806)
            T *t = BK;
            BK = C;
            C = t;
        }
        w = sqrt(dotRR);
        if (active) wLast = w;
        clk.mark(PH_ACOL);
        {  // saveToPrev as role moves
            T *t = ZP;
            ZP = Z;
            Z = t;
            t = RP;
            RP = RR;
            RR = t;
        }
        double gs[MAXM];
#pragma unroll
        for (int q = 0; q < MAXM; q++) gs[q] = 0.;
        const bool useLS = (it - it0 >= m) && gramSolve(G, gb, m, gs);
        clk.mark(PH_ALS);
        const bool blind = active && (w <= c0 * zeta);
        {  // d = -S g - r; K0 lanes write cand = z_prev + d (src/cache.cuh:897-898, 1055-1057)
            T gg[MAXM];
            VW Ss[MAXM];
#pragma unroll
            for (int s2 = 0; s2 < MAXM; s2++) {
                gg[s2] = (useLS && s2 < m) ? (T) gs[s2] : T(0);
                Ss[s2] = V(P.hS + (long) (s2 < m ? s2 : 0) * n * B);
            }
            const VW rp = V(RP), zp = V(ZP), dd = V(D), cc = V(C);
            E::template pass<4>(
                n, c,
                [&](long e) {
                    VH<T, MAXM> v;
                    v.a = rp[e];
                    v.b = zp[e];
#pragma unroll
                    for (int s2 = 0; s2 < MAXM; s2++) v.h[s2] = (s2 < m) ? Ss[s2][e] : T(0);
                    return v;
                },
                [&](long e, const VH<T, MAXM> &v) {
                    T sg = T(0);
#pragma unroll
                    for (int s2 = 0; s2 < MAXM; s2++)
                        if (s2 < m) sg += v.h[s2] * gg[s2];
                    const T d = -sg - v.a;
                    if (blind) cc[e] = v.b + d;
                    else dd[e] = d;
                });
        }
        __syncthreads();
        clk.mark(PH_AAPPLY);
        if (blind) {
            zeta = w;
            k0++;
        }
        bool ls = active && !blind;
        T tau = T(1);
        for (int in = 0; in < P.maxInner; in++) {  // line search (src/cache.cuh:1063-1099)
            if (!anyLane(ls)) break;
            {
                const VW d = V(D), zp = V(ZP), ds = V(DS), z = V(Z);
                E::template pass<16>(
                    n, c, [&](long e) { return V2<T>{d[e], zp[e]}; },
                    [&](long e, const V2<T> &v) {
                        if (ls) {
                            const T s = v.a * tau;
                            ds[e] = s;
                            z[e] = v.b + s;
                        }
                    });
            }
            __syncthreads();
            clk.mark(PH_LINE);
            E::applyT(Pp, Z, C2, phaseAcc);
            clk.skip();
            if (ls) nT++;
            T sRRt, sRD;
            {  // r~ = z - T(z), ||r~||^2 and r~ . ds
                const VW rr = V(RR), z = V(Z), cc = V(C2), ds = V(DS);
                T a1 = T(0), a2 = T(0);
                E::template pass<16>(n, c, [&](long e) { return V4<T>{z[e], cc[e], ds[e], T(0)}; },
                        [&](long e, const V4<T> &v) {
                            const T d = v.a - v.b;
                            rr[e] = d;
                            a1 += d * d;
                            a2 += d * v.c;
                        });
                sRRt = (T) E::sum((double) a1, lds, c);
                sRD = (T) E::sum((double) a2, lds, c);
            }
            clk.mark(PH_RES);
            // <r~, M r~> and <r~, M ds> from two adjoints (src/cache.cuh:1066,1080)
            const Cross2<T> cr = E::adjDots(Pp, RR, DS, lds, true);
            const T wTilde = sqrt(sRRt - T(2) * P.gamma * cr.c1);
            const T dotRD = sRD - P.gamma * cr.c2;
            if (ls) nEll++;
            clk.mark(PH_ELL);
            const bool takeK1 = ls && (w <= wSafe) && (wTilde <= c1 * w);
            if (anyLane(takeK1)) {  // K1: cand <- z
                const VW cc = V(C), z = V(Z);
                E::template pass<16>(n, c, [&](long e) { return z[e]; }, [&](long e, T v) {
                    if (takeK1) cc[e] = v;
                });
                __syncthreads();
            }
            if (takeK1) {
                wSafe = wTilde + pow(c2, T(it - it0));
                k1++;
            }
            const bool still = ls && !takeK1;
            const T rho = wTilde * wTilde - dotRD;
            const bool takeK2 = still && (rho >= sigma * wTilde * w);
            if (anyLane(takeK2)) {  // K2: cand <- z_prev - f r
                const T f = lambda * rho / (wTilde * wTilde);
                const VW cc = V(C), rr = V(RR), zp = V(ZP);
                E::template pass<16>(
                    n, c, [&](long e) { return V2<T>{rr[e], zp[e]}; },
                    [&](long e, const V2<T> &v) {
                        if (takeK2) cc[e] = v.b - v.a * f;
                    });
                __syncthreads();
            }
            if (takeK2) k2++;
            const bool back = still && !takeK2;
            if (back) {
                tau *= beta;
                k2bt++;
            }
            const bool takeK3 = back && (in >= P.maxInner - 1);
            if (anyLane(takeK3)) {  // K3: restore the plain CP step
                const VW cc = V(C), bk = V(BK);
                E::template pass<16>(n, c, [&](long e) { return bk[e]; }, [&](long e, T v) {
                    if (takeK3) cc[e] = v;
                });
                __syncthreads();
            }
            if (takeK3) k3++;
            ls = back;
            clk.mark(PH_LINE);
        }
    }
    // A resumable (chunked) launch keeps everything a resume or a refill needs, whether it ended
    // at the chunk's end or because every lane had finished; results report the progress so far.
    if (P.chunkIters > 0) {
        if (threadIdx.x == 0) {
            gs->finished = chunkExit ? 0 : 1;
            T *cur[6] = {Z, C, ZP, RR, RP, BK};
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++)
                    if (cur[i] == P.work[j]) gs->roles[i] = j;
            gs->head = head;
            gs->it = it;
            gs->ticks = ticksBefore + (long) (wall_clock64() - t0);
        }
        if (c.wave == 0 && c.writer) {
            LaneSave &sv = P.save[gl];
            for (int i = 0; i < MAXM; i++) {
                sv.gb[i] = gb[i];
                for (int j = 0; j < MAXM; j++) sv.G[i][j] = G[i][j];
            }
            sv.tol = (double) tol;
            sv.lastErr = (double) lastErr;
            sv.wLast = (double) wLast;
            sv.zeta = (double) zeta;
            sv.w = (double) w;
            sv.wSafe = (double) wSafe;
            sv.iters = active ? it - it0 : iters;
            sv.k0 = k0; sv.k1 = k1; sv.k2 = k2; sv.k2bt = k2bt; sv.k3 = k3; sv.nT = nT; sv.nEll = nEll;
            sv.status = status;
            sv.active = active ? 1 : 0;
            sv.initFlag = initFlag ? 1 : 0;
            sv.it0 = it0;
        }
    } else if (threadIdx.x == 0) {
        gs->finished = 1;
    }
    if (active) iters = it - it0;
    __syncthreads();
    if (c.wave == 0 && c.writer && real) {
        Out &o = outs[gl];
        o.status = status;
        o.iters = iters;
        o.err = (double) lastErr;
        o.tol = (double) tol;
        o.w = (double) wLast;
        o.k0 = k0;
        o.k1 = k1;
        o.k2 = k2;
        o.k2bt = k2bt;
        o.k3 = k3;
        o.nT = nT;
        o.nEll = nEll;
        for (int i = 0; i < NUM_PHASES; i++) o.ticks[i] = phaseAcc[i];
    }
}

}  // namespace spk
