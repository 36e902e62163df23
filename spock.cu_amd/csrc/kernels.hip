// Fused SPOCK engine for gfx950: one workgroup owns one problem instance and runs the
// whole SuperMann/Anderson solve (src/cache.cuh:1008-1132) inside a single launch.
//
// Work decomposition inside the workgroup: a "node task" is handled by one 64-lane
// wavefront with lanes over the rows of the node's blocks (node data are column major,
// so every matrix column is one coalesced load); stage-sequential sweeps are separated
// by workgroup barriers; iterate-wide passes are block-strided; reductions use
// cross-lane shuffles and an LDS combine in a fixed order (bitwise repeatable).
//
// Operator semantics follow the reference (file:line cited per function); outputs the
// reference leaves unwritten (t[0] of L*, the IV block of node 0 of L) are zero here.
#include "engine.hpp"

#include <hip/hip_runtime.h>
#include <math.h>

namespace spk {

// ------------------------------------------------------------------ wave primitives
__device__ __forceinline__ double bcastLane(double v, int src) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_readlane(lo, src);
    hi = __builtin_amdgcn_readlane(hi, src);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ float bcastLane(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

template <typename T>
__device__ __forceinline__ T waveSum(T v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <typename T>
__device__ __forceinline__ T waveMax(T v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

struct Tc {
    int lane, wave, nw, tid, nt;
};

__device__ __forceinline__ Tc threadCtx() {
    Tc c;
    c.tid = threadIdx.x;
    c.nt = blockDim.x;
    c.lane = c.tid & 63;
    c.wave = c.tid >> 6;
    c.nw = c.nt >> 6;
    return c;
}

template <typename T, int K>
__device__ __forceinline__ void zero(T (&a)[K]) {
#pragma unroll
    for (int i = 0; i < K; i++) a[i] = T(0);
}

// v[c] holds element lane + 64 c of a vector of length len (0 beyond)
template <typename T, typename F>
__device__ __forceinline__ void loadV(T (&v)[MAXV], int len, int lane, F f) {
#pragma unroll
    for (int c = 0; c < MAXV; c++) {
        const int k = lane + 64 * c;
        v[c] = (k < len) ? f(k) : T(0);
    }
}

template <typename T>
__device__ __forceinline__ void loadPtr(T (&v)[MAXV], const T *p, int len, int lane) {
    loadV(v, len, lane, [&](int k) { return p[k]; });
}

// acc[rc] (row lane + 64 rc) += sum_{c < cols} M[row + c * ld] * v(c)
template <typename T>
__device__ __forceinline__ void mvAcc(T (&acc)[MAXR], const T *M, int rows, int ld, int cols, const T (&v)[MAXV],
                                      int lane) {
#pragma unroll
    for (int vc = 0; vc < MAXV; vc++) {
        if (64 * vc >= cols) break;
        const int cend = min(64, cols - 64 * vc);
#pragma unroll 4
        for (int cc = 0; cc < cend; cc++) {
            const T xc = bcastLane(v[vc], cc);
            const T *col = M + (size_t) (64 * vc + cc) * ld;
#pragma unroll
            for (int rc = 0; rc < MAXR; rc++) {
                if (64 * rc < rows) {
                    const int r = lane + 64 * rc;
                    if (r < rows) acc[rc] += col[r] * xc;
                }
            }
        }
    }
}

// Sum of K per-thread values over the workgroup, identical result in every thread.
template <typename T, int K>
__device__ __forceinline__ void blockSum(T (&v)[K], T *lds, const Tc &c) {
#pragma unroll
    for (int i = 0; i < K; i++) v[i] = waveSum(v[i]);
    if (c.lane == 0)
#pragma unroll
        for (int i = 0; i < K; i++) lds[i * 32 + c.wave] = v[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < K; i++) {
        T s = T(0);
        for (int w = 0; w < c.nw; w++) s += lds[i * 32 + w];
        v[i] = s;
    }
    __syncthreads();
}

template <typename T>
__device__ __forceinline__ T blockMax(T v, T *lds, const Tc &c) {
    v = waveMax(v);
    if (c.lane == 0) lds[c.wave] = v;
    __syncthreads();
    T s = T(0);
    for (int w = 0; w < c.nw; w++) s = fmax(s, lds[w]);
    __syncthreads();
    return s;
}

// ------------------------------------------------------------------ projections
// Box projection in the arithmetic form of src/definitions.cu:38-52.
template <typename T>
__device__ __forceinline__ T projBox(T v, T lb, T ub) {
    const int lower = (v <= lb), upper = (v >= ub);
    const T lbi = isinf(lb) ? T(0) : T(lower) * lb;
    const T ubi = isinf(ub) ? T(0) : T(upper) * ub;
    return v * T(1 - lower) * T(1 - upper) + lbi + ubi;
}

// ------------------------------------------------------------------ L* per node
// src/operator.cuh:63-82; src/costs.cuh:162-195 (quadratic), :418-458 (quadratic+linear);
// src/constraints.cuh:171-191 (rectangle), :265-286 (polyhedron).
template <typename T>
struct AdjOut {
    T x[MAXR];  // rows nx
    T u[MAXR];  // rows nu (nonleaf only)
    T y;        // lane < yD (nonleaf only)
    T s, t;     // uniform scalars
};

template <typename T>
__device__ __forceinline__ void adjNode(const Inst<T> &P, const T *dual, int n, int lane, AdjOut<T> &o) {
    const int nx = P.nx, nu = P.nu, nxu = nx + nu, d4 = nxu + 2;
    zero(o.x);
    zero(o.u);
    o.y = T(0);
    o.s = T(0);
    o.t = T(0);
    T v[MAXV], w[MAXR];
    if (n < P.nl) {
        const T II = dual[P.oII + n];
        o.s = II;
        if (lane < P.yD) o.y = dual[P.oI + (long) n * P.yD + lane] - P.b.at(n)[lane] * II;
        const long b3 = P.oIII + (long) n * P.con3;
        if (P.ncon == CON_RECT_DEV) {
#pragma unroll
            for (int rc = 0; rc < MAXR; rc++) {
                const int r = lane + 64 * rc;
                if (r < nx) o.x[rc] = dual[b3 + r];
                if (r < nu) o.u[rc] = dual[b3 + nx + r];
            }
        } else if (P.ncon == CON_POLY_DEV) {
            loadPtr(v, dual + b3, P.con3, lane);
            const T *GT = P.nGT.at(n);
            mvAcc(o.x, GT, nx, nxu, P.con3, v, lane);
            mvAcc(o.u, GT + nx, nu, nxu, P.con3, v, lane);
        }
        const int c0 = P.chFrom[n], nc = P.numCh[n];
        for (int ci = 0; ci < nc; ci++) {  // children accumulate in order (src/costs.cuh:172-177)
            const int j = c0 + ci;
            const T *iv = dual + P.oIV + (long) j * d4;
            const T a = iv[nxu], bb = iv[nxu + 1];
            loadPtr(v, iv, nx, lane);
            zero(w);
            mvAcc(w, P.sQ.at(j), nx, nx, nx, v, lane);
            if (P.ncost == COST_QUADLIN_DEV) {
                const T *qj = P.q.at(j);
#pragma unroll
                for (int rc = 0; rc < MAXR; rc++) {
                    const int r = lane + 64 * rc;
                    if (r < nx) w[rc] = w[rc] - T(0.5) * qj[r] * a - T(0.5) * qj[r] * bb;
                }
            }
#pragma unroll
            for (int rc = 0; rc < MAXR; rc++) o.x[rc] += w[rc];
            loadPtr(v, iv + nx, nu, lane);
            zero(w);
            mvAcc(w, P.sR.at(j), nu, nu, nu, v, lane);
            if (P.ncost == COST_QUADLIN_DEV) {
                const T *rj = P.r.at(j);
#pragma unroll
                for (int rc = 0; rc < MAXR; rc++) {
                    const int r = lane + 64 * rc;
                    if (r < nu) w[rc] = w[rc] - T(0.5) * rj[r] * a - T(0.5) * rj[r] * bb;
                }
            }
#pragma unroll
            for (int rc = 0; rc < MAXR; rc++) o.u[rc] += w[rc];
        }
    } else {
        const int l = n - P.nl;
        const long b5 = P.oV + (long) l * P.con5;
        if (P.lcon == CON_RECT_DEV) {
#pragma unroll
            for (int rc = 0; rc < MAXR; rc++) {
                const int r = lane + 64 * rc;
                if (r < nx) o.x[rc] = dual[b5 + r];
            }
        } else if (P.lcon == CON_POLY_DEV) {
            loadPtr(v, dual + b5, P.con5, lane);
            mvAcc(o.x, P.lGT.at(l), nx, nx, P.con5, v, lane);
        }
        const T *vi = dual + P.oVI + (long) l * (nx + 2);
        const T a = vi[nx], bb = vi[nx + 1];
        loadPtr(v, vi, nx, lane);
        zero(w);
        mvAcc(w, P.lsQ.at(l), nx, nx, nx, v, lane);
#pragma unroll
        for (int rc = 0; rc < MAXR; rc++) o.x[rc] += w[rc];
        if (P.lcost == COST_QUADLIN_DEV) {
            const T *lq = P.lq.at(l);
#pragma unroll
            for (int rc = 0; rc < MAXR; rc++) {
                const int r = lane + 64 * rc;
                if (r < nx) o.x[rc] = o.x[rc] - T(0.5) * lq[r] * a - T(0.5) * lq[r] * bb;
            }
        }
        o.s = (a + bb) * T(0.5);
    }
    if (n >= 1) {
        const T *iv = dual + P.oIV + (long) n * d4;
        o.t = (iv[nxu] + iv[nxu + 1]) * T(0.5);
    }
}

// ------------------------------------------------------------------ L per node
// src/operator.cuh:41-60; src/costs.cuh:126-160 (quadratic), :379-416 (quadratic+linear);
// src/constraints.cuh:158-169 (rectangle), :251-263 (polyhedron).
// The handler H receives each dual block of node n as register values:
//   blockI(n, yk (lane < yD), IIval)
//   blockElem(dualIndexBase, vals, rows, lb table entry, ub table entry)  [III and V]
//   blockSoc(dualIndexBase, xs (rows nx), us (rows nu, may be 0 rows), h, transEntry)  [IV and VI]
template <typename T, typename PV, typename H>
__device__ __forceinline__ void opNode(const Inst<T> &P, int n, int lane, PV pv, H &h) {
    const int nx = P.nx, nu = P.nu, nxu = nx + nu;
    T v[MAXV], v2[MAXV], acc[MAXR], acc2[MAXR];
    if (n < P.nl) {
        // I and II
        const T yk = (lane < P.yD) ? pv(P.oY + (long) n * P.yD + lane) : T(0);
        const T bk = (lane < P.yD) ? P.b.at(n)[lane] : T(0);
        const T by = waveSum(bk * yk);
        const T IIval = pv(P.oS + n) - by;
        h.blockI(n, yk, IIval);
        // III
        if (P.ncon == CON_RECT_DEV) {
            zero(acc);
            zero(acc2);
#pragma unroll
            for (int rc = 0; rc < MAXR; rc++) {
                const int r = lane + 64 * rc;
                if (r < nx) acc[rc] = pv(P.oX + (long) n * nx + r);
                if (r < nu) acc2[rc] = pv(P.oU + (long) n * nu + r);
            }
            const long base = P.oIII + (long) n * P.con3;
            h.blockElem(base, acc, nx, P.nLB.at(n), P.nUB.at(n));
            h.blockElem(base + nx, acc2, nu, P.nLB.at(n) + nx, P.nUB.at(n) + nx);
        } else if (P.ncon == CON_POLY_DEV) {
            loadV(v, nx, lane, [&](int k) { return pv(P.oX + (long) n * nx + k); });
            loadV(v2, nu, lane, [&](int k) { return pv(P.oU + (long) n * nu + k); });
            zero(acc);
            const T *G = P.nG.at(n);
            mvAcc(acc, G, P.con3, P.con3, nx, v, lane);
            mvAcc(acc, G + (size_t) P.con3 * nx, P.con3, P.con3, nu, v2, lane);
            h.blockElem(P.oIII + (long) n * P.con3, acc, P.con3, P.nLB.at(n), P.nUB.at(n));
        }
    }
    // IV
    {
        zero(acc);
        zero(acc2);
        T hv = T(0);
        if (n >= 1) {
            const int a = P.anc[n];
            loadV(v, nx, lane, [&](int k) { return pv(P.oX + (long) a * nx + k); });
            loadV(v2, nu, lane, [&](int k) { return pv(P.oU + (long) a * nu + k); });
            mvAcc(acc, P.sQ.at(n), nx, nx, nx, v, lane);
            mvAcc(acc2, P.sR.at(n), nu, nu, nu, v2, lane);
            hv = pv(P.oT + n) * T(0.5);
            if (P.ncost == COST_QUADLIN_DEV) {
                const T *qn = P.q.at(n);
                const T *rn = P.r.at(n);
                T pq = T(0), pr = T(0);
#pragma unroll
                for (int c = 0; c < MAXV; c++) {
                    const int k = lane + 64 * c;
                    if (k < nx) pq += qn[k] * v[c];
                    if (k < nu) pr += rn[k] * v2[c];
                }
                hv = hv + T(-0.5) * waveSum(pq);
                hv = hv + T(-0.5) * waveSum(pr);
            }
        }
        h.blockSoc(P.oIV + (long) n * (nxu + 2), acc, nx, acc2, nu, hv, P.trans.at(n));
    }
    // V and VI
    if (n >= P.nl) {
        const int l = n - P.nl;
        loadV(v, nx, lane, [&](int k) { return pv(P.oX + (long) n * nx + k); });
        if (P.lcon == CON_RECT_DEV) {
            zero(acc);
#pragma unroll
            for (int rc = 0; rc < MAXR; rc++) {
                const int r = lane + 64 * rc;
                if (r < nx) acc[rc] = pv(P.oX + (long) n * nx + r);
            }
            h.blockElem(P.oV + (long) l * P.con5, acc, nx, P.lLB.at(l), P.lUB.at(l));
        } else if (P.lcon == CON_POLY_DEV) {
            zero(acc);
            mvAcc(acc, P.lG.at(l), P.con5, P.con5, nx, v, lane);
            h.blockElem(P.oV + (long) l * P.con5, acc, P.con5, P.lLB.at(l), P.lUB.at(l));
        }
        zero(acc);
        zero(acc2);
        mvAcc(acc, P.lsQ.at(l), nx, nx, nx, v, lane);
        T hs = pv(P.oS + n) * T(0.5);
        if (P.lcost == COST_QUADLIN_DEV) {
            const T *lq = P.lq.at(l);
            T pq = T(0);
#pragma unroll
            for (int c = 0; c < MAXV; c++) {
                const int k = lane + 64 * c;
                if (k < nx) pq += lq[k] * v[c];
            }
            hs = hs + T(-0.5) * waveSum(pq);
        }
        h.blockSoc(P.oVI + (long) l * (nx + 2), acc, nx, acc2, 0, hs, P.ltrans.at(l));
    }
}

// ---------------------------------------------------------------- L handlers
// Raw store of L values into out (dual-local indices).
template <typename T>
struct StoreL {
    const Inst<T> &P;
    T *out;
    int lane;
    __device__ void blockI(int n, T yk, T IIval) {
        if (lane < P.yD) out[P.oI + (long) n * P.yD + lane] = yk;
        if (lane == 0) out[P.oII + n] = IIval;
    }
    __device__ void blockElem(long base, const T (&a)[MAXR], int rows, const T *, const T *) {
#pragma unroll
        for (int rc = 0; rc < MAXR; rc++) {
            const int r = lane + 64 * rc;
            if (r < rows) out[base + r] = a[rc];
        }
    }
    __device__ void blockSoc(long base, const T (&xs)[MAXR], int nxr, const T (&us)[MAXR], int nur, T hv, const T *) {
#pragma unroll
        for (int rc = 0; rc < MAXR; rc++) {
            const int r = lane + 64 * rc;
            if (r < nxr) out[base + r] = xs[rc];
            if (r < nur) out[base + nxr + r] = us[rc];
        }
        if (lane == 0) {
            out[base + nxr + nur] = hv;
            out[base + nxr + nur + 1] = hv;
        }
    }
};

// Dual prox of the CP step (src/cache.cuh:770-789): a = L*gamma + z_d; w = a/gamma + c;
// cand_d = gamma (w - Pi(w)).  Pi: masked NNOC (src/risks.cuh:99-106, definitions.cu:430-443),
// NNOC (src/cones.cuh:134-137), box (src/constraints.cuh:70-74), SOC (src/projections.cuh:102-128).
template <typename T>
struct ProxL {
    const Inst<T> &P;
    const T *zd;
    T *cd;
    int lane;
    __device__ __forceinline__ T wOf(T Lv, long k, T trans) const { return (Lv * P.gamma + zd[k]) * P.gammaRecip + trans; }
    __device__ void blockI(int n, T yk, T IIval) {
        if (lane < P.yD) {
            const long k = P.oI + (long) n * P.yD + lane;
            const T w = wOf(yk, k, T(0));
            const int idx = (lane < 2 * P.numCh[n]) && (w < T(0));
            const T p = T(1 - idx) * w;
            cd[k] = (w - p) * P.gamma;
        }
        if (lane == 0) {
            const long k = P.oII + n;
            const T w = wOf(IIval, k, T(0));
            const T p = fmax(T(0), w);
            cd[k] = (w - p) * P.gamma;
        }
    }
    __device__ void blockElem(long base, const T (&a)[MAXR], int rows, const T *lb, const T *ub) {
#pragma unroll
        for (int rc = 0; rc < MAXR; rc++) {
            const int r = lane + 64 * rc;
            if (r < rows) {
                const long k = base + r;
                const T w = wOf(a[rc], k, T(0));
                const T p = projBox(w, lb[r], ub[r]);
                cd[k] = (w - p) * P.gamma;
            }
        }
    }
    __device__ void blockSoc(long base, const T (&xs)[MAXR], int nxr, const T (&us)[MAXR], int nur, T hv,
                             const T *tr) {
        T wx[MAXR], wu[MAXR];
        T ss = T(0);
#pragma unroll
        for (int rc = 0; rc < MAXR; rc++) {
            const int r = lane + 64 * rc;
            wx[rc] = T(0);
            wu[rc] = T(0);
            if (r < nxr) {
                wx[rc] = wOf(xs[rc], base + r, tr[r]);
                ss += wx[rc] * wx[rc];
            }
            if (r < nur) {
                wu[rc] = wOf(us[rc], base + nxr + r, tr[nxr + r]);
                ss += wu[rc] * wu[rc];
            }
        }
        const long kh = base + nxr + nur;
        const T w1 = wOf(hv, kh, tr[nxr + nur]);
        const T w2 = wOf(hv, kh + 1, tr[nxr + nur + 1]);
        ss = waveSum(ss) + w1 * w1;
        const T nrm = sqrt(ss), tl = w2;
        const int i1 = nrm <= tl;
        const int i2 = (nrm <= -tl) && !i1;
        const int i3 = !i1 && !i2;
        const T sc = (nrm + tl) / (T(2) * nrm + T(1 - i3));
        const T mult = T(1 - i2) * (T(i3) * sc + T(1 - i3));
#pragma unroll
        for (int rc = 0; rc < MAXR; rc++) {
            const int r = lane + 64 * rc;
            if (r < nxr) cd[base + r] = (wx[rc] - wx[rc] * mult) * P.gamma;
            if (r < nur) cd[base + nxr + r] = (wu[rc] - wu[rc] * mult) * P.gamma;
        }
        if (lane == 0) {
            const T p1 = w1 * mult;
            const int i1b = (1 - i2) * (1 - i3);
            const T p2 = T(i1b) * (w2 * T(1 - i2)) + T((1 - i1b) * (1 - i2)) * (T(i3) * (sc * nrm - T(1)) + T(1));
            cd[kh] = (w1 - p1) * P.gamma;
            cd[kh + 1] = (w2 - p2) * P.gamma;
        }
    }
};

// Reduction handler for ell(y) = [L* y_d; L y_p]: accumulates sum x . (y - gamma ell) and
// max |y / gamma - ell| over the dual part (src/cache.cuh:722-734 and :917-931).
template <typename T>
struct ReduceL {
    const Inst<T> &P;
    const T *xd, *yd;  // dual parts of the partner and input vectors
    int lane;
    T dot, mx;
    __device__ __forceinline__ void acc(long k, T ell) {
        const T y = yd[k];
        dot += xd[k] * (ell * (-P.gamma) + y);
        mx = fmax(mx, fabs(y * P.gammaRecip - ell));
    }
    __device__ void blockI(int n, T yk, T IIval) {
        if (lane < P.yD) acc(P.oI + (long) n * P.yD + lane, yk);
        if (lane == 0) acc(P.oII + n, IIval);
    }
    __device__ void blockElem(long base, const T (&a)[MAXR], int rows, const T *, const T *) {
#pragma unroll
        for (int rc = 0; rc < MAXR; rc++) {
            const int r = lane + 64 * rc;
            if (r < rows) acc(base + r, a[rc]);
        }
    }
    __device__ void blockSoc(long base, const T (&xs)[MAXR], int nxr, const T (&us)[MAXR], int nur, T hv, const T *) {
#pragma unroll
        for (int rc = 0; rc < MAXR; rc++) {
            const int r = lane + 64 * rc;
            if (r < nxr) acc(base + r, xs[rc]);
            if (r < nur) acc(base + nxr + r, us[rc]);
        }
        if (lane == 0) {
            acc(base + nxr + nur, hv);
            acc(base + nxr + nur + 1, hv);
        }
    }
};

// ---------------------------------------------------------------- T phases
// modifyPrimal + proxRootS (src/cache.cuh:748-753, 639-642): cand_p = z_p - gamma L*(z_d); s_0 -= gamma.
// Leaf tasks also seed the dynamics sweep with q_leaf = -xbar_leaf (src/dynamics.cuh:79-81).
template <typename T>
__device__ void phaseAdjAxpy(const Inst<T> &P, const Tc &c, const T *z, T *cand) {
    const T *zd = z + P.sPrim;
    const T g = -P.gamma;
    const int nx = P.nx, nu = P.nu;
    for (int n = c.wave; n < P.N; n += c.nw) {
        AdjOut<T> o;
        adjNode(P, zd, n, c.lane, o);
#pragma unroll
        for (int rc = 0; rc < MAXR; rc++) {
            const int r = c.lane + 64 * rc;
            if (r < nx) {
                const long k = P.oX + (long) n * nx + r;
                const T v = o.x[rc] * g + z[k];
                cand[k] = v;
                if (n >= P.nl) P.qd[(long) n * nx + r] = -v;
            }
            if (n < P.nl && r < nu) {
                const long k = P.oU + (long) n * nu + r;
                cand[k] = o.u[rc] * g + z[k];
            }
        }
        if (n < P.nl && c.lane < P.yD) {
            const long k = P.oY + (long) n * P.yD + c.lane;
            cand[k] = o.y * g + z[k];
        }
        if (c.lane == 0) {
            T sv = o.s * g + z[P.oS + n];
            if (n == 0) sv -= P.gamma;
            cand[P.oS + n] = sv;
            cand[P.oT + n] = o.t * g + z[P.oT + n];
        }
    }
}

// Backward sweep step for node i (src/dynamics.cuh:85-139).
template <typename T>
__device__ void dynBackwardNode(const Inst<T> &P, T *cand, int i, int stage, int lane) {
    const int nx = P.nx, nu = P.nu;
    const T *ubar = cand + P.oU + (long) i * nu;
    const T *xbar = cand + P.oX + (long) i * nx;
    T v[MAXV], w[MAXR], acc[MAXR];
    zero(acc);
    const int c0 = P.chFrom[i], nc = P.numCh[i];
    for (int ci = 0; ci < nc; ci++) {
        const int j = c0 + ci;
        loadPtr(v, P.qd + (long) j * nx, nx, lane);
        zero(w);
        mvAcc(w, P.BTr.at(j), nu, nu, nx, v, lane);
#pragma unroll
        for (int rc = 0; rc < MAXR; rc++) acc[rc] = (ci == 0) ? w[rc] : acc[rc] + w[rc];
    }
    // d = R~^-1 (u - sum B'q) with R~ = L L' (potrs semantics)
    T d = (lane < nu) ? (-acc[0] + ubar[lane]) : T(0);
    const T *L = P.cholL.at(i);
    for (int k = 0; k < nu; k++) {
        const T dk = bcastLane(d, k) / L[k + (size_t) k * nu];
        if (lane == k) d = dk;
        else if (lane > k && lane < nu) d -= L[lane + (size_t) k * nu] * dk;
    }
    const T *U = P.cholU.at(i);
    for (int k = nu - 1; k >= 0; k--) {
        const T dk = bcastLane(d, k) / U[k + (size_t) k * nu];
        if (lane == k) d = dk;
        else if (lane < k) d -= U[lane + (size_t) k * nu] * dk;
    }
    if (lane < nu) P.dd[(long) i * nu + lane] = d;
    if (stage == 0) return;
    // q_i = sum_j (APB_j d_i + ABKTr_j q_j) + K_i'(d_i - u_i) - x_i
    T dv[MAXV];
    zero(dv);
    dv[0] = (lane < nu) ? d : T(0);
    zero(acc);
    for (int ci = 0; ci < nc; ci++) {
        const int j = c0 + ci;
        zero(w);
        mvAcc(w, P.APB.at(j), nx, nx, nu, dv, lane);
        loadPtr(v, P.qd + (long) j * nx, nx, lane);
        mvAcc(w, P.ABKTr.at(j), nx, nx, nx, v, lane);
#pragma unroll
        for (int rc = 0; rc < MAXR; rc++) acc[rc] = (ci == 0) ? w[rc] : acc[rc] + w[rc];
    }
    dv[0] = (lane < nu) ? d - ubar[lane] : T(0);
    zero(w);
    mvAcc(w, P.KTr.at(i), nx, nx, nu, dv, lane);
#pragma unroll
    for (int rc = 0; rc < MAXR; rc++) {
        const int r = lane + 64 * rc;
        if (r < nx) P.qd[(long) i * nx + r] = (acc[rc] + w[rc]) - xbar[r];
    }
}

// Kernel projection for nonleaf node i (src/cache.cuh:650-676); padded child slots enter as 0.
template <typename T>
__device__ void kernelNode(const Inst<T> &P, T *cand, int i, int lane) {
    const int yD = P.yD, E = P.E, nc = P.numCh[i], c0 = P.chFrom[i];
    T v = T(0);
    if (lane < yD) v = cand[P.oY + (long) i * yD + lane];
    else if (lane < yD + E) {
        if (lane - yD < nc) v = cand[P.oT + c0 + (lane - yD)];
    } else if (lane < yD + 2 * E) {
        if (lane - yD - E < nc) v = cand[P.oS + c0 + (lane - yD - E)];
    }
    T vin[MAXV], w[MAXR];
    zero(vin);
    vin[0] = v;
    zero(w);
    mvAcc(w, P.NN.at(i), P.nD, P.nD, P.nD, vin, lane);
    if (lane < yD) cand[P.oY + (long) i * yD + lane] = w[0];
    else if (lane < yD + E) {
        if (lane - yD < nc) cand[P.oT + c0 + (lane - yD)] = w[0];
    } else if (lane < yD + 2 * E) {
        if (lane - yD - E < nc) cand[P.oS + c0 + (lane - yD - E)] = w[0];
    }
}

// Forward rollout for node i (src/dynamics.cuh:144-173).
template <typename T>
__device__ void dynForwardNode(const Inst<T> &P, T *cand, int i, int lane) {
    const int nx = P.nx, nu = P.nu;
    T xv[MAXV], uv[MAXV], w[MAXR];
    if (i == 0) {
        loadPtr(xv, P.x0s, nx, lane);
#pragma unroll
        for (int c = 0; c < MAXV; c++) {
            const int k = lane + 64 * c;
            if (k < nx) cand[P.oX + k] = xv[c];
        }
    } else {
        loadPtr(xv, cand + P.oX + (long) i * nx, nx, lane);
    }
    zero(w);
    mvAcc(w, P.K.at(i), nu, nu, nx, xv, lane);
    const T u = (lane < nu) ? w[0] + P.dd[(long) i * nu + lane] : T(0);
    if (lane < nu) cand[P.oU + (long) i * nu + lane] = u;
    zero(uv);
    uv[0] = u;
    const int c0 = P.chFrom[i], nc = P.numCh[i];
    for (int ci = 0; ci < nc; ci++) {
        const int j = c0 + ci;
        const T *AB = P.AB.at(j);
        zero(w);
        mvAcc(w, AB, nx, nx, nx, xv, lane);
        mvAcc(w, AB + (size_t) nx * nx, nx, nx, nu, uv, lane);
#pragma unroll
        for (int rc = 0; rc < MAXR; rc++) {
            const int r = lane + 64 * rc;
            if (r < nx) cand[P.oX + (long) j * nx + r] = w[rc];
        }
    }
}

// Dynamics projection + kernel projection on cand (in place).  The kernel projection touches
// (y, t, s) only, so it shares the first backward stage's barrier interval.
template <typename T>
__device__ void phaseProjections(const Inst<T> &P, const Tc &c, T *cand, bool doDyn, bool doKernel) {
    if (doDyn) {
        for (int stage = P.S - 2; stage >= 0; stage--) {
            const int f = P.stageFrom[stage], nStage = P.stageTo[stage] - f + 1;
            const int extra = (doKernel && stage == P.S - 2) ? P.nl : 0;
            for (int task = c.wave; task < nStage + extra; task += c.nw) {
                if (task < nStage) dynBackwardNode(P, cand, f + task, stage, c.lane);
                else kernelNode(P, cand, task - nStage, c.lane);
            }
            __syncthreads();
        }
        for (int stage = 0; stage <= P.S - 2; stage++) {
            const int f = P.stageFrom[stage], nStage = P.stageTo[stage] - f + 1;
            for (int task = c.wave; task < nStage; task += c.nw) dynForwardNode(P, cand, f + task, c.lane);
            __syncthreads();
        }
    } else if (doKernel) {
        for (int task = c.wave; task < P.nl; task += c.nw) kernelNode(P, cand, task, c.lane);
        __syncthreads();
    }
}

// The CP operator T (src/cache.cuh:796-801): cand = T(z).
template <typename T>
__device__ void applyT(const Inst<T> &P, const Tc &c, const T *z, T *cand) {
    phaseAdjAxpy(P, c, z, cand);
    __syncthreads();
    phaseProjections(P, c, cand, true, true);
    // modifyDual + proximalDual on 2 x-bar - z_p
    ProxL<T> h{P, z + P.sPrim, cand + P.sPrim, c.lane};
    auto pv = [&](long k) { return cand[k] * T(2) - z[k]; };
    for (int n = c.wave; n < P.N; n += c.nw) opNode(P, n, c.lane, pv, h);
    __syncthreads();
}

// sum x.(y - gamma ell(y)) and max |y/gamma - ell(y)| with ell(y) = [L* y_d; L y_p].
template <typename T>
__device__ void reduceEll(const Inst<T> &P, const Tc &c, const T *x, const T *y, T *lds, T &dotOut, T &maxOut) {
    const int nx = P.nx, nu = P.nu;
    T dot = T(0), mx = T(0);
    auto acc = [&](long k, T ell) {
        const T yk = y[k];
        dot += x[k] * (ell * (-P.gamma) + yk);
        mx = fmax(mx, fabs(yk * P.gammaRecip - ell));
    };
    const T *yd = y + P.sPrim;
    ReduceL<T> h{P, x + P.sPrim, yd, c.lane, T(0), T(0)};
    auto pv = [&](long k) { return y[k]; };
    for (int n = c.wave; n < P.N; n += c.nw) {
        AdjOut<T> o;
        adjNode(P, yd, n, c.lane, o);
#pragma unroll
        for (int rc = 0; rc < MAXR; rc++) {
            const int r = c.lane + 64 * rc;
            if (r < nx) acc(P.oX + (long) n * nx + r, o.x[rc]);
            if (n < P.nl && r < nu) acc(P.oU + (long) n * nu + r, o.u[rc]);
        }
        if (n < P.nl && c.lane < P.yD) acc(P.oY + (long) n * P.yD + c.lane, o.y);
        if (c.lane == 0) {
            acc(P.oS + n, o.s);
            acc(P.oT + n, o.t);
        }
        opNode(P, n, c.lane, pv, h);
    }
    T v[1] = {dot + h.dot};
    blockSum(v, lds, c);
    dotOut = v[0];
    maxOut = blockMax(fmax(mx, h.mx), lds, c);
}

// ---------------------------------------------------------------- Anderson direction
// src/cache.cuh:861-904 with the least squares argmin ||Y g - r|| solved by Householder QR
// (geqrf + ormqr + trsm).  Y and S are rings of andBuff columns; logical column 0 is the newest.
template <typename T>
__device__ void andersonDirection(const Inst<T> &P, const Tc &c, long idx, int head, const T *r, T *dir, T *lds) {
    const long n = P.sIter;
    const int m = P.andBuff;
    auto slot = [&](int lc) { return (head - lc + m) % m; };
    if (idx < m) {
        for (long k = c.tid; k < n; k += c.nt) dir[k] = -r[k];
        __syncthreads();
        return;
    }
    T *W = P.W;
    // copy [Y (logical order) | r] and the first column's tail norm
    T sig[1] = {T(0)};
    for (long k = c.tid; k < n; k += c.nt) {
        for (int lc = 0; lc < m; lc++) W[(long) lc * n + k] = P.hY[(long) slot(lc) * n + k];
        W[(long) m * n + k] = r[k];
        if (k >= 1) {
            const T y0 = P.hY[(long) slot(0) * n + k];
            sig[0] += y0 * y0;
        }
    }
    blockSum(sig, lds, c);
    T sigma = sig[0];
    T Rd[8];
    for (int kk = 0; kk < m; kk++) {
        const T alpha = W[(long) kk * n + kk];
        T tau = T(0), beta = alpha, scale = T(0);
        if (sigma != T(0)) {
            const T nrm = sqrt(alpha * alpha + sigma);
            beta = alpha >= T(0) ? -nrm : nrm;
            tau = (beta - alpha) / beta;
            scale = T(1) / (alpha - beta);
        }
        Rd[kk] = beta;
        const T *Wk = W + (long) kk * n;
        if (tau != T(0)) {
            T wj[8];
#pragma unroll
            for (int j = 0; j < 8; j++) wj[j] = T(0);
            for (long k = kk + c.tid; k < n; k += c.nt) {
                const T vk = (k == kk) ? T(1) : Wk[k] * scale;
                for (int j = kk + 1; j <= m; j++) wj[j - kk - 1] += vk * W[(long) j * n + k];
            }
            // block reduce the (m - kk) dot products
            blockSum(wj, lds, c);
            T wt[8];
#pragma unroll
            for (int j = 0; j < 8; j++) wt[j] = wj[j] * tau;
            T sg[1] = {T(0)};
            for (long k = kk + c.tid; k < n; k += c.nt) {
                const T vk = (k == kk) ? T(1) : Wk[k] * scale;
                for (int j = kk + 1; j <= m; j++) {
                    T &e = W[(long) j * n + k];
                    e -= wt[j - kk - 1] * vk;
                    if (j == kk + 1 && k >= kk + 2) sg[0] += e * e;
                }
            }
            blockSum(sg, lds, c);
            sigma = sg[0];
        } else {
            T sg[1] = {T(0)};
            if (kk + 1 < m)
                for (long k = kk + 2 + c.tid; k < n; k += c.nt) {
                    const T e = W[(long) (kk + 1) * n + k];
                    sg[0] += e * e;
                }
            blockSum(sg, lds, c);
            sigma = sg[0];
        }
    }
    // back substitution R g = (Q' r)[0:m]; row kk of R sits in W[j][kk]
    T g[8];
    for (int kk = m - 1; kk >= 0; kk--) {
        T v = W[(long) m * n + kk];
        for (int j = kk + 1; j < m; j++) v -= W[(long) j * n + kk] * g[j];
        g[kk] = v / Rd[kk];
    }
    for (long k = c.tid; k < n; k += c.nt) {
        T sg = T(0);
        for (int lc = 0; lc < m; lc++) sg += P.hS[(long) slot(lc) * n + k] * g[lc];
        dir[k] = -sg - r[k];
    }
    __syncthreads();
}

// ---------------------------------------------------------------- the solve
template <typename T>
__device__ void solveInstance(const Inst<T> &P, Out &out, T *lds, double tickRate) {
    const Tc c = threadCtx();
    const long n = P.sIter;
    T *z = P.z, *cand = P.cand;
    const int m = P.andBuff;
    int head = m - 1;
    const long t0 = wall_clock64();
    long nT = 0, nEll = 0;
    // initialisePrev / initialiseIter (src/cache.cuh:611-632)
    if (P.method == METHOD_SPOCK) {
        for (long k = c.tid; k < n; k += c.nt) {
            if (P.warm) {
                P.zPrev[k] = z[k];
                P.resPrev[k] = P.res[k];
            } else {
                P.zPrev[k] = T(1);
                P.resPrev[k] = T(1);
            }
        }
    }
    __syncthreads();
    bool initFlag = true;
    int status = ST_NOTRUN;
    long iters = 0;
    T tol = T(0), lastErr = T(NAN), wLast = T(NAN);
    const T c0 = T(0.99), c1 = T(0.99), c2 = T(0.99), beta = T(0.5), sigma = T(0.1), lambda = T(1);
    const T inf = T(INFINITY);
    T zeta = inf, w = inf, wSafe = inf;
    long k0 = 0, k1 = 0, k2 = 0, k2bt = 0, k3 = 0;
    const long hardCap = P.maxIters > 0 ? P.maxIters + 1 : 2000000;
    long it = 0;
    for (it = 0; it < hardCap; it++) {
        if (it != 0 && P.method == METHOD_SPOCK) {  // acceptCandidate: z <- cand
            T *tmp = z;
            z = cand;
            cand = tmp;
        }
        applyT(P, c, z, cand);
        nT++;
        // computeError (src/cache.cuh:910-961): r = z - cand, backup = cand
        for (long k = c.tid; k < n; k += c.nt) {
            const T cv = cand[k];
            P.res[k] = z[k] - cv;
            P.backup[k] = cv;
        }
        __syncthreads();
        bool haveDot = false;
        T dotRR = T(0);
        if (it % P.checkPeriod == 0) {
            T mx;
            reduceEll(P, c, P.res, P.res, lds, dotRR, mx);
            nEll++;
            haveDot = true;
            if (initFlag) {
                tol = fmax(P.tolAbs, P.tolRel * mx);
                initFlag = false;
            } else {
                lastErr = mx;
                int code = ST_NOTRUN;
                if (mx <= tol) code = ST_CONVERGED;
                if (P.maxIters && it >= P.maxIters) code = ST_OUTOFITER;
                if (P.maxTicks > 0 && it % 1000 == 0 && (double) (wall_clock64() - t0) >= P.maxTicks) code = ST_OUTOFTIME;
                if (code != ST_NOTRUN) {
                    status = code;
                    iters = it;
                }
            }
        }
        if (status != ST_NOTRUN) {  // acceptCandidate and stop
            T *tmp = z;
            z = cand;
            cand = tmp;
            break;
        }
        if (P.method == METHOD_CP) {  // runCp: accept every candidate (src/cache.cuh:967-1002)
            T *tmp = z;
            z = cand;
            cand = tmp;
            continue;
        }
        // normM(r) (src/cache.cuh:740-742)
        if (!haveDot) {
            T mx;
            reduceEll(P, c, P.res, P.res, lds, dotRR, mx);
            nEll++;
        }
        w = sqrt(dotRR);
        wLast = w;
        // delta iterate/residual, saveToPrev and the new Anderson column
        head = (head + 1) % m;
        {
            T *Yc = P.hY + (long) head * n, *Sc = P.hS + (long) head * n;
            for (long k = c.tid; k < n; k += c.nt) {
                const T rk = P.res[k], zk = z[k];
                const T dr = rk - P.resPrev[k];
                const T dz = zk - P.zPrev[k];
                Yc[k] = dr;
                Sc[k] = dz - dr;
                P.zPrev[k] = zk;
                P.resPrev[k] = rk;
            }
        }
        __syncthreads();
        andersonDirection(P, c, it, head, P.res, P.dir, lds);
        if (w <= c0 * zeta) {  // K0 blind update (src/cache.cuh:1055-1061)
            for (long k = c.tid; k < n; k += c.nt) cand[k] = P.zPrev[k] + P.dir[k];
            __syncthreads();
            zeta = w;
            k0++;
            continue;
        }
        T tau = T(1);
        for (int in = 0; in < P.maxInner; in++) {  // line search (src/cache.cuh:1063-1099)
            for (long k = c.tid; k < n; k += c.nt) {
                const T ds = P.dir[k] * tau;
                P.dirS[k] = ds;
                z[k] = P.zPrev[k] + ds;
            }
            __syncthreads();
            applyT(P, c, z, cand);
            nT++;
            for (long k = c.tid; k < n; k += c.nt) P.res[k] = z[k] - cand[k];
            __syncthreads();
            T dot, mx;
            reduceEll(P, c, P.res, P.res, lds, dot, mx);
            nEll++;
            const T wTilde = sqrt(dot);
            if (w <= wSafe && wTilde <= c1 * w) {  // K1
                for (long k = c.tid; k < n; k += c.nt) cand[k] = z[k];
                __syncthreads();
                wSafe = wTilde + pow(c2, T(it));
                k1++;
                break;
            }
            T dotRD;
            reduceEll(P, c, P.res, P.dirS, lds, dotRD, mx);
            nEll++;
            const T rho = wTilde * wTilde - dotRD;
            if (rho >= sigma * wTilde * w) {  // K2
                const T f = lambda * rho / (wTilde * wTilde);
                for (long k = c.tid; k < n; k += c.nt) {
                    const T rk = P.res[k] * f;
                    P.res[k] = rk;
                    cand[k] = P.zPrev[k] - rk;
                }
                __syncthreads();
                k2++;
                break;
            } else {
                tau *= beta;
                k2bt++;
            }
            if (in >= P.maxInner - 1) {  // K3: restore the plain CP step
                for (long k = c.tid; k < n; k += c.nt) cand[k] = P.backup[k];
                __syncthreads();
                k3++;
            }
        }
    }
    if (status == ST_NOTRUN) {
        status = ST_OUTOFITER;
        iters = it;
    }
    // the accepted iterate must end in P.z
    if (z != P.z) {
        for (long k = c.tid; k < n; k += c.nt) P.z[k] = z[k];
    }
    __syncthreads();
    if (c.tid == 0) {
        out.status = status;
        out.iters = iters;
        out.err = (double) lastErr;
        out.tol = (double) tol;
        out.w = (double) wLast;
        out.k0 = k0;
        out.k1 = k1;
        out.k2 = k2;
        out.k2bt = k2bt;
        out.k3 = k3;
        out.nT = nT;
        out.nEll = nEll;
    }
}

template <typename T>
__global__ void __launch_bounds__(1024) k_solve(const Inst<T> *__restrict__ insts, Out *outs, int count,
                                                double tickRate) {
    __shared__ T lds[8 * 32];
    const int k = blockIdx.x;
    if (k >= count) return;
    solveInstance(insts[k], outs[k], lds, tickRate);
}

// ---------------------------------------------------------------- single-operator kernels
// Debug entry points used by the parity tests: they read P.z and write P.cand.
enum DebugOp { OP_L = 0, OP_LADJ = 1, OP_DYN = 2, OP_KERNEL = 3, OP_T = 4, OP_ELL = 5 };

template <typename T>
__global__ void __launch_bounds__(1024) k_debug(const Inst<T> *__restrict__ insts, Out *outs, int count, int op) {
    __shared__ T lds[8 * 32];
    const int kInst = blockIdx.x;
    if (kInst >= count) return;
    const Inst<T> &P = insts[kInst];
    const Tc c = threadCtx();
    T *z = P.z, *cand = P.cand;
    if (op == OP_L) {
        StoreL<T> h{P, cand + P.sPrim, c.lane};
        auto pv = [&](long k) { return z[k]; };
        for (int n = c.wave; n < P.N; n += c.nw) opNode(P, n, c.lane, pv, h);
    } else if (op == OP_LADJ) {
        const T *zd = z + P.sPrim;
        for (int n = c.wave; n < P.N; n += c.nw) {
            AdjOut<T> o;
            adjNode(P, zd, n, c.lane, o);
            for (int rc = 0; rc < MAXR; rc++) {
                const int r = c.lane + 64 * rc;
                if (r < P.nx) cand[P.oX + (long) n * P.nx + r] = o.x[rc];
                if (n < P.nl && r < P.nu) cand[P.oU + (long) n * P.nu + r] = o.u[rc];
            }
            if (n < P.nl && c.lane < P.yD) cand[P.oY + (long) n * P.yD + c.lane] = o.y;
            if (c.lane == 0) {
                cand[P.oS + n] = o.s;
                cand[P.oT + n] = o.t;
            }
        }
    } else if (op == OP_DYN || op == OP_KERNEL) {
        for (long k = c.tid; k < P.sPrim; k += c.nt) cand[k] = z[k];
        __syncthreads();
        if (op == OP_DYN)
            for (int n = c.wave; n < P.N; n += c.nw)
                if (n >= P.nl)
                    for (int r = c.lane; r < P.nx; r += 64) P.qd[(long) n * P.nx + r] = -cand[P.oX + (long) n * P.nx + r];
        __syncthreads();
        phaseProjections(P, c, cand, op == OP_DYN, op == OP_KERNEL);
    } else if (op == OP_T) {
        applyT(P, c, z, cand);
    } else if (op == OP_ELL) {
        T dot, mx;
        reduceEll(P, c, z, cand, lds, dot, mx);
        if (c.tid == 0) {
            outs[kInst].w = (double) dot;
            outs[kInst].err = (double) mx;
        }
    }
}

// Elementwise conversions used by the FP32-bulk / FP64-polish path.
__global__ void k_f32_to_f64(const float *in, double *out, long n) {
    for (long k = blockIdx.x * (long) blockDim.x + threadIdx.x; k < n; k += (long) gridDim.x * blockDim.x) out[k] = in[k];
}

__global__ void k_f64_to_f32(const double *in, float *out, long n) {
    for (long k = blockIdx.x * (long) blockDim.x + threadIdx.x; k < n; k += (long) gridDim.x * blockDim.x) out[k] = (float) in[k];
}

template <typename T>
__global__ void k_fill(T *p, T v, long n) {
    for (long k = blockIdx.x * (long) blockDim.x + threadIdx.x; k < n; k += (long) gridDim.x * blockDim.x) p[k] = v;
}

// explicit instantiations
template __global__ void k_solve<double>(const Inst<double> *, Out *, int, double);
template __global__ void k_solve<float>(const Inst<float> *, Out *, int, double);
template __global__ void k_debug<double>(const Inst<double> *, Out *, int, int);
template __global__ void k_debug<float>(const Inst<float> *, Out *, int, int);
template __global__ void k_fill<double>(double *, double, long);
template __global__ void k_fill<float>(float *, float, long);

}  // namespace spk
