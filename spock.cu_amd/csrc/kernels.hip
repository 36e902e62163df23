// Fused SPOCK engine for gfx950: one workgroup owns one problem instance and runs the
// whole SuperMann/Anderson solve (src/cache.cuh:1008-1132) inside a single launch.
// Grid mode: for one large tree, every workgroup of a cooperative launch works on the same
// instance; the work distribution below then spans the grid, barriers are grid barriers and
// reductions combine per-workgroup partials (in a fixed order, so every workgroup takes the
// same branch decisions).
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

#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>
#include <math.h>

#ifndef SPK_MAX_THREADS
#define SPK_MAX_THREADS 512
#endif

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

// Work distribution of a thread: tid/nt and wave/nw index the threads and waves that share one
// instance (the workgroup, or the whole grid in grid mode); ltid/lw/lnw are workgroup-local.
struct Tc {
    int lane, wave, nw, tid, nt;
    int ltid, lw, lnw;
    int grid;
    void *gpart;      // grid mode: two buffers of GRID_MAXK x GRID_MAXWG reduction partials
    mutable int seq;  // reductions done so far; selects the partials buffer
};

__device__ __forceinline__ Tc threadCtx(int grid = 0, void *gpart = nullptr) {
    Tc c;
    c.ltid = threadIdx.x;
    c.lw = c.ltid >> 6;
    c.lnw = blockDim.x >> 6;
    c.grid = grid;
    c.gpart = gpart;
    c.seq = 0;
    c.tid = grid ? (int) (blockIdx.x * blockDim.x + threadIdx.x) : c.ltid;
    c.nt = grid ? (int) (gridDim.x * blockDim.x) : (int) blockDim.x;
    c.lane = c.tid & 63;
    c.wave = c.tid >> 6;
    c.nw = c.nt >> 6;
    return c;
}

// Barrier over the threads that share the instance.
__device__ __forceinline__ void sync(const Tc &c) {
    if (c.grid) cooperative_groups::this_grid().sync();
    else __syncthreads();
}

// Grid mode: combine the workgroup values v[0..K) of every workgroup.  Each workgroup publishes its
// values, then every wave reads all partials lane-strided and reduces them with the same shuffle
// tree, so all threads of the grid hold bitwise identical results.  Two partial buffers alternate:
// a workgroup can only start overwriting one after every workgroup passed the barrier of the
// reduction in between, which is after its reads of the same buffer.
template <typename T, int K, bool MAX>
__device__ __forceinline__ void gridCombine(T (&v)[K], const Tc &c) {
    static_assert(K <= GRID_MAXK, "too many values per grid reduction");
    T *part = static_cast<T *>(c.gpart) + (size_t) (c.seq & 1) * GRID_MAXK * GRID_MAXWG;
    c.seq++;
    if (c.ltid == 0)
#pragma unroll
        for (int i = 0; i < K; i++) part[i * GRID_MAXWG + blockIdx.x] = v[i];
    cooperative_groups::this_grid().sync();
    const int nb = gridDim.x;
#pragma unroll
    for (int i = 0; i < K; i++) {
        T a = T(0);
        for (int b = c.lane; b < nb; b += 64) {
            const T p = part[i * GRID_MAXWG + b];
            a = MAX ? fmax(a, p) : a + p;
        }
        v[i] = MAX ? waveMax(a) : waveSum(a);
    }
}

// Per-phase wall-clock accounting; mark() is called right after a workgroup barrier so the
// interval covers the whole phase.
struct PhaseClock {
    long last;
    long acc[NUM_PHASES];
    __device__ void start() {
        last = wall_clock64();
        for (int i = 0; i < NUM_PHASES; i++) acc[i] = 0;
    }
    __device__ __forceinline__ void mark(int ph) {
        const long t = wall_clock64();
        acc[ph] += t - last;
        last = t;
    }
};

template <typename T, int K>
__device__ __forceinline__ void zero(T (&a)[K]) {
#pragma unroll
    for (int i = 0; i < K; i++) a[i] = T(0);
}

// v[c] holds element lane + 64 c of a vector of length len (0 beyond)
template <typename T, int V, typename F>
__device__ __forceinline__ void loadV(T (&v)[V], int len, int lane, F f) {
#pragma unroll
    for (int c = 0; c < V; c++) {
        const int k = lane + 64 * c;
        v[c] = (k < len) ? f(k) : T(0);
    }
}

template <typename T, int V>
__device__ __forceinline__ void loadPtr(T (&v)[V], const T *p, int len, int lane) {
    loadV(v, len, lane, [&](int k) { return p[k]; });
}

// acc[rc] (row lane + 64 rc) += sum_{c < cols} M[row + c * ld] * v(c).
// All columns of a 64-column chunk (32 per block in FP64) are requested before any is
// consumed, so a matvec costs one or two memory round trips instead of one per column;
// out-of-range columns load 0 and v holds 0 beyond its length.
template <typename T, int R, int V>
__device__ __forceinline__ void mvAcc(T (&acc)[R], const T *M, int rows, int ld, int cols, const T (&v)[V], int lane) {
    constexpr int U = (sizeof(T) == 8 ? 32 : 64) / R;
#pragma unroll
    for (int vc = 0; vc < V; vc++) {
        if (64 * vc >= cols) break;
        const int cend = min(64, cols - 64 * vc);
        const T *base = M + (size_t) (64 * vc) * ld;
#pragma unroll 1
        for (int c0 = 0; c0 < cend; c0 += U) {
            T m[U][R];
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int rc = 0; rc < R; rc++) {
                    const int r = lane + 64 * rc;
                    m[u][rc] = (c0 + u < cend && r < rows) ? base[(size_t) (c0 + u) * ld + r] : T(0);
                }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const T xc = bcastLane(v[vc], (c0 + u) & 63);
#pragma unroll
                for (int rc = 0; rc < R; rc++) acc[rc] += m[u][rc] * xc;
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
        for (int i = 0; i < K; i++) lds[i * 32 + c.lw] = v[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < K; i++) {
        T s = T(0);
        for (int w = 0; w < c.lnw; w++) s += lds[i * 32 + w];
        v[i] = s;
    }
    __syncthreads();
    if (c.grid) gridCombine<T, K, false>(v, c);
}

template <typename T>
__device__ __forceinline__ T blockMax(T v, T *lds, const Tc &c) {
    v = waveMax(v);
    if (c.lane == 0) lds[c.lw] = v;
    __syncthreads();
    T s = T(0);
    for (int w = 0; w < c.lnw; w++) s = fmax(s, lds[w]);
    __syncthreads();
    if (c.grid) {
        T a[1] = {s};
        gridCombine<T, 1, true>(a, c);
        s = a[0];
    }
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
template <typename T, int R>
struct AdjOut {
    T x[R];  // rows nx
    T u[R];  // rows nu (nonleaf only)
    T y;        // lane < yD (nonleaf only)
    T s, t;     // uniform scalars
};

template <typename T, int R, int V>
__device__ __forceinline__ void adjNode(const Inst<T> &P, const T *dual, int n, int lane, AdjOut<T, R> &o) {
    const int nx = P.nx, nu = P.nu, nxu = nx + nu, d4 = nxu + 2;
    zero(o.x);
    zero(o.u);
    o.y = T(0);
    o.s = T(0);
    o.t = T(0);
    T v[V], w[R];
    if (n < P.nl) {
        const T II = dual[P.oII + n];
        o.s = II;
        if (lane < P.yD) o.y = dual[P.oI + (long) n * P.yD + lane] - P.b.at(n)[lane] * II;
        const long b3 = P.oIII + (long) n * P.con3;
        if (P.ncon == CON_RECT_DEV) {
#pragma unroll
            for (int rc = 0; rc < R; rc++) {
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
            if (P.ncost == COST_LIN_DEV) {  // x_i += q_j sum_j, u_i += r_j sum_j (src/costs.cuh:274-285)
                const T sj = dual[P.oIV + j];
                const T *qj = P.q.at(j), *rj = P.r.at(j);
#pragma unroll
                for (int rc = 0; rc < R; rc++) {
                    const int r = lane + 64 * rc;
                    if (r < nx) o.x[rc] += qj[r] * sj;
                    if (r < nu) o.u[rc] += rj[r] * sj;
                }
                continue;
            }
            const T *iv = dual + P.oIV + (long) j * d4;
            const T a = iv[nxu], bb = iv[nxu + 1];
            loadPtr(v, iv, nx, lane);
            zero(w);
            mvAcc(w, P.sQ.at(j), nx, nx, nx, v, lane);
            if (P.ncost == COST_QUADLIN_DEV) {
                const T *qj = P.q.at(j);
#pragma unroll
                for (int rc = 0; rc < R; rc++) {
                    const int r = lane + 64 * rc;
                    if (r < nx) w[rc] = w[rc] - T(0.5) * qj[r] * a - T(0.5) * qj[r] * bb;
                }
            }
#pragma unroll
            for (int rc = 0; rc < R; rc++) o.x[rc] += w[rc];
            loadPtr(v, iv + nx, nu, lane);
            zero(w);
            mvAcc(w, P.sR.at(j), nu, nu, nu, v, lane);
            if (P.ncost == COST_QUADLIN_DEV) {
                const T *rj = P.r.at(j);
#pragma unroll
                for (int rc = 0; rc < R; rc++) {
                    const int r = lane + 64 * rc;
                    if (r < nu) w[rc] = w[rc] - T(0.5) * rj[r] * a - T(0.5) * rj[r] * bb;
                }
            }
#pragma unroll
            for (int rc = 0; rc < R; rc++) o.u[rc] += w[rc];
        }
    } else {
        const int l = n - P.nl;
        const long b5 = P.oV + (long) l * P.con5;
        if (P.lcon == CON_RECT_DEV) {
#pragma unroll
            for (int rc = 0; rc < R; rc++) {
                const int r = lane + 64 * rc;
                if (r < nx) o.x[rc] = dual[b5 + r];
            }
        } else if (P.lcon == CON_POLY_DEV) {
            loadPtr(v, dual + b5, P.con5, lane);
            mvAcc(o.x, P.lGT.at(l), nx, nx, P.con5, v, lane);
        }
        if (P.lcost == COST_LIN_DEV) {  // x_l += q sum_l, s_l = set_l (src/costs.cuh:289-292)
            const T sl = dual[P.oVI + l];
            const T *lq = P.lq.at(l);
#pragma unroll
            for (int rc = 0; rc < R; rc++) {
                const int r = lane + 64 * rc;
                if (r < nx) o.x[rc] += lq[r] * sl;
            }
            o.s = dual[P.oVI + P.nlf + l];
        } else {
            const T *vi = dual + P.oVI + (long) l * (nx + 2);
            const T a = vi[nx], bb = vi[nx + 1];
            loadPtr(v, vi, nx, lane);
            zero(w);
            mvAcc(w, P.lsQ.at(l), nx, nx, nx, v, lane);
#pragma unroll
            for (int rc = 0; rc < R; rc++) o.x[rc] += w[rc];
            if (P.lcost == COST_QUADLIN_DEV) {
                const T *lq = P.lq.at(l);
#pragma unroll
                for (int rc = 0; rc < R; rc++) {
                    const int r = lane + 64 * rc;
                    if (r < nx) o.x[rc] = o.x[rc] - T(0.5) * lq[r] * a - T(0.5) * lq[r] * bb;
                }
            }
            o.s = (a + bb) * T(0.5);
        }
    }
    if (P.ncost == COST_LIN_DEV) {
        o.t = dual[P.oIV + P.N + n];  // t <- set, every node (src/costs.cuh:286)
    } else if (n >= 1) {
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
template <typename T, int R, int V, typename PV, typename H>
__device__ __forceinline__ void opNode(const Inst<T> &P, int n, int lane, PV pv, H &h) {
    const int nx = P.nx, nu = P.nu, nxu = nx + nu;
    T v[V], v2[V], acc[R], acc2[R];
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
            for (int rc = 0; rc < R; rc++) {
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
    if (P.ncost == COST_LIN_DEV) {  // (sum_n, set_n) = (q_n' x_anc + r_n' u_anc, t_n), node 0: (0, t_0)
        T sum = T(0);              // (src/costs.cuh:261-267)
        if (n >= 1) {
            const int a = P.anc[n];
            const T *qn = P.q.at(n), *rn = P.r.at(n);
            T pq = T(0), pr = T(0);
#pragma unroll
            for (int c = 0; c < V; c++) {
                const int k = lane + 64 * c;
                if (k < nx) pq += qn[k] * pv(P.oX + (long) a * nx + k);
                if (k < nu) pr += rn[k] * pv(P.oU + (long) a * nu + k);
            }
            sum = waveSum(pq) + waveSum(pr);
        }
        h.blockLin(P.oIV + n, P.oIV + P.N + n, sum, pv(P.oT + n));
    } else {
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
                for (int c = 0; c < V; c++) {
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
            for (int rc = 0; rc < R; rc++) {
                const int r = lane + 64 * rc;
                if (r < nx) acc[rc] = pv(P.oX + (long) n * nx + r);
            }
            h.blockElem(P.oV + (long) l * P.con5, acc, nx, P.lLB.at(l), P.lUB.at(l));
        } else if (P.lcon == CON_POLY_DEV) {
            zero(acc);
            mvAcc(acc, P.lG.at(l), P.con5, P.con5, nx, v, lane);
            h.blockElem(P.oV + (long) l * P.con5, acc, P.con5, P.lLB.at(l), P.lUB.at(l));
        }
        if (P.lcost == COST_LIN_DEV) {  // (q' x_l, s_l) (src/costs.cuh:269-272)
            const T *lq = P.lq.at(l);
            T pq = T(0);
#pragma unroll
            for (int c = 0; c < V; c++) {
                const int k = lane + 64 * c;
                if (k < nx) pq += lq[k] * v[c];
            }
            h.blockLin(P.oVI + l, P.oVI + P.nlf + l, waveSum(pq), pv(P.oS + n));
            return;
        }
        zero(acc);
        zero(acc2);
        mvAcc(acc, P.lsQ.at(l), nx, nx, nx, v, lane);
        T hs = pv(P.oS + n) * T(0.5);
        if (P.lcost == COST_QUADLIN_DEV) {
            const T *lq = P.lq.at(l);
            T pq = T(0);
#pragma unroll
            for (int c = 0; c < V; c++) {
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
template <typename T, int R>
struct StoreL {
    const Inst<T> &P;
    T *out;
    int lane;
    __device__ void blockI(int n, T yk, T IIval) {
        if (lane < P.yD) out[P.oI + (long) n * P.yD + lane] = yk;
        if (lane == 0) out[P.oII + n] = IIval;
    }
    __device__ void blockElem(long base, const T (&a)[R], int rows, const T *, const T *) {
#pragma unroll
        for (int rc = 0; rc < R; rc++) {
            const int r = lane + 64 * rc;
            if (r < rows) out[base + r] = a[rc];
        }
    }
    __device__ void blockSoc(long base, const T (&xs)[R], int nxr, const T (&us)[R], int nur, T hv, const T *) {
#pragma unroll
        for (int rc = 0; rc < R; rc++) {
            const int r = lane + 64 * rc;
            if (r < nxr) out[base + r] = xs[rc];
            if (r < nur) out[base + nxr + r] = us[rc];
        }
        if (lane == 0) {
            out[base + nxr + nur] = hv;
            out[base + nxr + nur + 1] = hv;
        }
    }
    __device__ void blockLin(long eSum, long eSet, T sum, T set) {
        if (lane == 0) {
            out[eSum] = sum;
            out[eSet] = set;
        }
    }
};

// Dual prox of the CP step (src/cache.cuh:770-789): a = L*gamma + z_d; w = a/gamma + c;
// cand_d = gamma (w - Pi(w)).  Pi: masked NNOC (src/risks.cuh:99-106, definitions.cu:430-443),
// NNOC (src/cones.cuh:134-137), box (src/constraints.cuh:70-74), SOC (src/projections.cuh:102-128).
template <typename T, int R>
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
    __device__ void blockElem(long base, const T (&a)[R], int rows, const T *lb, const T *ub) {
#pragma unroll
        for (int rc = 0; rc < R; rc++) {
            const int r = lane + 64 * rc;
            if (r < rows) {
                const long k = base + r;
                const T w = wOf(a[rc], k, T(0));
                const T p = projBox(w, lb[r], ub[r]);
                cd[k] = (w - p) * P.gamma;
            }
        }
    }
    __device__ void blockSoc(long base, const T (&xs)[R], int nxr, const T (&us)[R], int nur, T hv,
                             const T *tr) {
        T wx[R], wu[R];
        T ss = T(0);
#pragma unroll
        for (int rc = 0; rc < R; rc++) {
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
        for (int rc = 0; rc < R; rc++) {
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
    // linear cost: the cost entry clamped to at most the epigraph entry, which is kept
    // (src/costs.cuh:296-301 with k_projectRectangle, lower bound -inf)
    __device__ void blockLin(long eSum, long eSet, T sum, T set) {
        if (lane == 0) {
            const T ws = wOf(sum, eSum, T(0)), wt = wOf(set, eSet, T(0));
            cd[eSum] = (ws - projBelow(ws, wt)) * P.gamma;
            cd[eSet] = (wt - wt) * P.gamma;
        }
    }
};

// Reduction handler for ell(y) = [L* y_d; L y_p]: accumulates sum x . (y - gamma ell) and
// max |y / gamma - ell| over the dual part (src/cache.cuh:722-734 and :917-931).
template <typename T, int R>
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
    __device__ void blockElem(long base, const T (&a)[R], int rows, const T *, const T *) {
#pragma unroll
        for (int rc = 0; rc < R; rc++) {
            const int r = lane + 64 * rc;
            if (r < rows) acc(base + r, a[rc]);
        }
    }
    __device__ void blockSoc(long base, const T (&xs)[R], int nxr, const T (&us)[R], int nur, T hv, const T *) {
#pragma unroll
        for (int rc = 0; rc < R; rc++) {
            const int r = lane + 64 * rc;
            if (r < nxr) acc(base + r, xs[rc]);
            if (r < nur) acc(base + nxr + r, us[rc]);
        }
        if (lane == 0) {
            acc(base + nxr + nur, hv);
            acc(base + nxr + nur + 1, hv);
        }
    }
    __device__ void blockLin(long eSum, long eSet, T sum, T set) {
        if (lane == 0) {
            acc(eSum, sum);
            acc(eSet, set);
        }
    }
};

// ---------------------------------------------------------------- T phases
// modifyPrimal + proxRootS (src/cache.cuh:748-753, 639-642): cand_p = z_p - gamma L*(z_d); s_0 -= gamma.
// Leaf tasks also seed the dynamics sweep with q_leaf = -xbar_leaf (src/dynamics.cuh:79-81).
template <typename T, int R, int V>
__device__ __noinline__ void phaseAdjAxpy(const Inst<T> &P, const Tc &c, const T *z, T *cand) {
    const T *zd = z + P.sPrim;
    const T g = -P.gamma;
    const int nx = P.nx, nu = P.nu;
    for (int n = c.wave; n < P.N; n += c.nw) {
        AdjOut<T, R> o;
        adjNode<T, R, V>(P, zd, n, c.lane, o);
#pragma unroll
        for (int rc = 0; rc < R; rc++) {
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

// Backward sweep step for node i (src/dynamics.cuh:85-139).  With affine dynamics the children's
// q enter as q_j + P_j e_j (src/dynamics.cuh:93-97).
template <typename T, int R, int V>
__device__ __forceinline__ void dynBackwardNode(const Inst<T> &P, T *cand, int i, int stage, int lane) {
    const int nx = P.nx, nu = P.nu;
    const T *ubar = cand + P.oU + (long) i * nu;
    const T *xbar = cand + P.oX + (long) i * nx;
    T v[V], w[R], acc[R];
    zero(acc);
    const int c0 = P.chFrom[i], nc = P.numCh[i];
    auto loadQ = [&](int j) {
        const T *qj = P.qd + (long) j * nx;
        if (P.affine) {
            const T *pe = P.Pe.at(j);
            loadV(v, nx, lane, [&](int k) { return qj[k] + pe[k]; });
        } else {
            loadPtr(v, qj, nx, lane);
        }
    };
    for (int ci = 0; ci < nc; ci++) {
        const int j = c0 + ci;
        loadQ(j);
        zero(w);
        mvAcc(w, P.BTr.at(j), nu, nu, nx, v, lane);
#pragma unroll
        for (int rc = 0; rc < R; rc++) acc[rc] = (ci == 0) ? w[rc] : acc[rc] + w[rc];
    }
    // d = R~^-1 (u - sum B'q) with R~ = L L': one matvec with the precomputed inverse, or the
    // two triangular solves of potrs
    T d = (lane < nu) ? (-acc[0] + ubar[lane]) : T(0);
    if (P.dynInverse) {
        T dvin[V], dout[R];
        zero(dvin);
        dvin[0] = d;
        zero(dout);
        mvAcc(dout, P.RtInv.at(i), nu, nu, nu, dvin, lane);
        d = (lane < nu) ? dout[0] : T(0);
    } else {
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
    }
    if (lane < nu) P.dd[(long) i * nu + lane] = d;
    if (stage == 0) return;
    // q_i = sum_j (APB_j d_i + ABKTr_j q_j) + K_i'(d_i - u_i) - x_i
    T dv[V];
    zero(dv);
    dv[0] = (lane < nu) ? d : T(0);
    zero(acc);
    for (int ci = 0; ci < nc; ci++) {
        const int j = c0 + ci;
        zero(w);
        mvAcc(w, P.APB.at(j), nx, nx, nu, dv, lane);
        loadQ(j);
        mvAcc(w, P.ABKTr.at(j), nx, nx, nx, v, lane);
#pragma unroll
        for (int rc = 0; rc < R; rc++) acc[rc] = (ci == 0) ? w[rc] : acc[rc] + w[rc];
    }
    dv[0] = (lane < nu) ? d - ubar[lane] : T(0);
    zero(w);
    mvAcc(w, P.KTr.at(i), nx, nx, nu, dv, lane);
#pragma unroll
    for (int rc = 0; rc < R; rc++) {
        const int r = lane + 64 * rc;
        if (r < nx) P.qd[(long) i * nx + r] = (acc[rc] + w[rc]) - xbar[r];
    }
}

// Kernel projection for nonleaf node i (src/cache.cuh:650-676); padded child slots enter as 0.
template <typename T, int R, int V>
__device__ __forceinline__ void kernelNode(const Inst<T> &P, T *cand, int i, int lane) {
    const int yD = P.yD, E = P.E, nc = P.numCh[i], c0 = P.chFrom[i];
    T v = T(0);
    if (lane < yD) v = cand[P.oY + (long) i * yD + lane];
    else if (lane < yD + E) {
        if (lane - yD < nc) v = cand[P.oT + c0 + (lane - yD)];
    } else if (lane < yD + 2 * E) {
        if (lane - yD - E < nc) v = cand[P.oS + c0 + (lane - yD - E)];
    }
    T vin[V], w[R];
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

// Forward rollout for node i (src/dynamics.cuh:144-173); affine dynamics add e_j to each child.
template <typename T, int R, int V>
__device__ __forceinline__ void dynForwardNode(const Inst<T> &P, T *cand, int i, int lane) {
    const int nx = P.nx, nu = P.nu;
    T xv[V], uv[V], w[R];
    if (i == 0) {
        loadPtr(xv, P.x0s, nx, lane);
#pragma unroll
        for (int c = 0; c < V; c++) {
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
        const T *ej = P.e.at(j);
#pragma unroll
        for (int rc = 0; rc < R; rc++) {
            const int r = lane + 64 * rc;
            if (r < nx) cand[P.oX + (long) j * nx + r] = P.affine ? w[rc] + ej[r] : w[rc];
        }
    }
}

// Dynamics projection + kernel projection on cand (in place).  The kernel projection touches
// (y, t, s) only, so it shares the first backward stage's barrier interval.
template <typename T, int R, int V>
__device__ __noinline__ void phaseProjections(const Inst<T> &P, const Tc &c, T *cand, bool doDyn, bool doKernel) {
    if (doDyn) {
        for (int stage = P.S - 2; stage >= 0; stage--) {
            const int f = P.stageFrom[stage], nStage = P.stageTo[stage] - f + 1;
            const int extra = (doKernel && stage == P.S - 2) ? P.nl : 0;
            for (int task = c.wave; task < nStage + extra; task += c.nw) {
                if (task < nStage) dynBackwardNode<T, R, V>(P, cand, f + task, stage, c.lane);
                else kernelNode<T, R, V>(P, cand, task - nStage, c.lane);
            }
            sync(c);
        }
        for (int stage = 0; stage <= P.S - 2; stage++) {
            const int f = P.stageFrom[stage], nStage = P.stageTo[stage] - f + 1;
            for (int task = c.wave; task < nStage; task += c.nw) dynForwardNode<T, R, V>(P, cand, f + task, c.lane);
            sync(c);
        }
    } else if (doKernel) {
        for (int task = c.wave; task < P.nl; task += c.nw) kernelNode<T, R, V>(P, cand, task, c.lane);
        sync(c);
    }
}

// modifyDual + proximalDual (src/cache.cuh:770-789) on 2 x-bar - z_p.
template <typename T, int R, int V>
__device__ __noinline__ void phaseDualProx(const Inst<T> &P, const Tc &c, const T *z, T *cand) {
    ProxL<T, R> h{P, z + P.sPrim, cand + P.sPrim, c.lane};
    auto pv = [&](long k) { return cand[k] * T(2) - z[k]; };
    for (int n = c.wave; n < P.N; n += c.nw) opNode<T, R, V>(P, n, c.lane, pv, h);
    sync(c);
}

// The CP operator T (src/cache.cuh:796-801): cand = T(z).
template <typename T, int R, int V>
__device__ __forceinline__ void applyT(const Inst<T> &P, const Tc &c, const T *z, T *cand, PhaseClock *clk = nullptr) {
    phaseAdjAxpy<T, R, V>(P, c, z, cand);
    sync(c);
    if (clk) clk->mark(PH_ADJ);
    phaseProjections<T, R, V>(P, c, cand, true, true);
    if (clk) clk->mark(PH_PROJ);
    phaseDualProx<T, R, V>(P, c, z, cand);
    if (clk) clk->mark(PH_DUAL);
}

// sum x.(y - gamma ell(y)) and max |y/gamma - ell(y)| with ell(y) = [L* y_d; L y_p].
// SPLIT: also the maxima over the primal and the dual part apart, in split[0] and split[1]
// (err1 and err2 of the debug trace).
template <typename T, int R, int V, bool SPLIT = false>
__device__ __noinline__ void reduceEll(const Inst<T> &P, const Tc &c, const T *x, const T *y, T *lds, T &dotOut, T &maxOut,
                                       T *split = nullptr) {
    const int nx = P.nx, nu = P.nu;
    T dot = T(0), mx = T(0);
    auto acc = [&](long k, T ell) {
        const T yk = y[k];
        dot += x[k] * (ell * (-P.gamma) + yk);
        mx = fmax(mx, fabs(yk * P.gammaRecip - ell));
    };
    const T *yd = y + P.sPrim;
    ReduceL<T, R> h{P, x + P.sPrim, yd, c.lane, T(0), T(0)};
    auto pv = [&](long k) { return y[k]; };
    for (int n = c.wave; n < P.N; n += c.nw) {
        AdjOut<T, R> o;
        adjNode<T, R, V>(P, yd, n, c.lane, o);
#pragma unroll
        for (int rc = 0; rc < R; rc++) {
            const int r = c.lane + 64 * rc;
            if (r < nx) acc(P.oX + (long) n * nx + r, o.x[rc]);
            if (n < P.nl && r < nu) acc(P.oU + (long) n * nu + r, o.u[rc]);
        }
        if (n < P.nl && c.lane < P.yD) acc(P.oY + (long) n * P.yD + c.lane, o.y);
        if (c.lane == 0) {
            acc(P.oS + n, o.s);
            acc(P.oT + n, o.t);
        }
        opNode<T, R, V>(P, n, c.lane, pv, h);
    }
    T v[1] = {dot + h.dot};
    blockSum(v, lds, c);
    dotOut = v[0];
    if constexpr (SPLIT) {
        split[0] = blockMax(mx, lds, c);
        split[1] = blockMax(h.mx, lds, c);
        maxOut = fmax(split[0], split[1]);
    } else {
        maxOut = blockMax(fmax(mx, h.mx), lds, c);
    }
}

// ---------------------------------------------------------------- iterate-wide passes
// Every thread issues U independent loads before it stores anything, so a single
// workgroup keeps U * blockDim loads in flight while it streams an iterate.
template <int U, typename Ld, typename St>
__device__ __forceinline__ void stream(long n, const Tc &c, Ld ld, St st) {
    using Vals = decltype(ld(0L));
    const long nt = c.nt;
    long k = c.tid;
    for (; k + (U - 1) * nt < n; k += U * nt) {
        Vals v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = ld(k + u * nt);
#pragma unroll
        for (int u = 0; u < U; u++) st(k + u * nt, v[u]);
    }
    for (; k < n; k += nt) st(k, ld(k));
}

template <typename T>
struct V2 {
    T a, b;
};
template <typename T>
struct V4 {
    T a, b, c, d;
};

// ---------------------------------------------------------------- Anderson direction
// src/cache.cuh:861-904 with the least squares argmin ||Y g - r|| solved by Householder QR
// (geqrf + ormqr + trsm).  Y and S are rings of andBuff columns; logical column 0 is the newest.
template <typename T>
__device__ __noinline__ void andersonHouseholder(const Inst<T> &P, const Tc &c, int head, const T *r, T *lds,
                                                double (&gs)[8]) {
    const long n = P.sIter;
    const int m = P.andBuff;
    auto slot = [&](int lc) { return (head - lc + m) % m; };
    T *W = P.W;
    // copy [Y (logical order) | r] and the first column's tail norm
    const T *Yl[8];
#pragma unroll
    for (int lc = 0; lc < 8; lc++) Yl[lc] = P.hY + (long) slot(lc < m ? lc : 0) * n;
    T sig[1] = {T(0)};
    stream<4>(
        n, c, [&](long k) { return k; },
        [&](long k, long) {
#pragma unroll
            for (int lc = 0; lc < 8; lc++)
                if (lc < m) W[(long) lc * n + k] = Yl[lc][k];
            W[(long) m * n + k] = r[k];
            if (k >= 1) sig[0] += Yl[0][k] * Yl[0][k];
        });
    blockSum(sig, lds, c);
    T sigma = sig[0];
    T Rd[8];
#pragma unroll
    for (int q = 0; q < 8; q++) Rd[q] = T(1);
    for (int kk = 0; kk < m; kk++) {
        const T alpha = W[(long) kk * n + kk];
        T tau = T(0), beta = alpha, scale = T(0);
        if (sigma != T(0)) {
            const T nrm = sqrt(alpha * alpha + sigma);
            beta = alpha >= T(0) ? -nrm : nrm;
            tau = (beta - alpha) / beta;
            scale = T(1) / (alpha - beta);
        }
#pragma unroll
        for (int q = 0; q < 8; q++)
            if (q == kk) Rd[q] = beta;
        const T *Wk = W + (long) kk * n;
        const long len = n - kk;
        if (tau != T(0)) {
            T wj[8];
#pragma unroll
            for (int jj = 0; jj < 8; jj++) wj[jj] = T(0);
            stream<4>(
                len, c, [&](long k) { return k; },
                [&](long k0, long) {
                    const long k = k0 + kk;
                    const T vk = (k == kk) ? T(1) : Wk[k] * scale;
#pragma unroll
                    for (int jj = 0; jj < 8; jj++)
                        if (kk + 1 + jj <= m) wj[jj] += vk * W[(long) (kk + 1 + jj) * n + k];
                });
            blockSum(wj, lds, c);  // the (m - kk) reflector dot products
            T wt[8];
#pragma unroll
            for (int jj = 0; jj < 8; jj++) wt[jj] = wj[jj] * tau;
            T sg[1] = {T(0)};
            stream<4>(
                len, c, [&](long k) { return k; },
                [&](long k0, long) {
                    const long k = k0 + kk;
                    const T vk = (k == kk) ? T(1) : Wk[k] * scale;
#pragma unroll
                    for (int jj = 0; jj < 8; jj++) {
                        if (kk + 1 + jj <= m) {
                            T &e = W[(long) (kk + 1 + jj) * n + k];
                            e -= wt[jj] * vk;
                            if (jj == 0 && k >= kk + 2) sg[0] += e * e;
                        }
                    }
                });
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
    // back substitution R g = (Q' r)[0:m]; row q of R sits in W[j][q]
    T g[8];
#pragma unroll
    for (int q = 7; q >= 0; q--) {
        g[q] = T(0);
        if (q < m) {
            T v = W[(long) m * n + q];
#pragma unroll
            for (int j = q + 1; j < 8; j++)
                if (j < m) v -= W[(long) j * n + q] * g[j];
            g[q] = v / Rd[q];
        }
    }
    // logical column lc sits in slot (head - lc) mod m
#pragma unroll
    for (int lc = 0; lc < 8; lc++)
        if (lc < m) gs[slot(lc)] = (double) g[lc];
    sync(c);
}

// New Anderson column (src/cache.cuh:863-868): Y_h = r - r_prev, S_h = (z - z_prev) - Y_h.
// With the Gram least squares it also accumulates, in FP64, <Y_h, Y_s> for every slot s and
// <Y_s, r> for every slot, and updates the slot-indexed Gram matrix G held in LDS.
template <typename T>
__device__ __noinline__ void andersonColumn(const Inst<T> &P, const Tc &c, int head, const T *Rv, const T *RP,
                                            const T *Z, const T *ZP, bool gram, double *G, double *gb, double *ldsd) {
    const long n = P.sIter;
    const int m = P.andBuff;
    T *Yh = P.hY + (long) head * n, *Sh = P.hS + (long) head * n;
    if (!gram) {
        stream<8>(
            n, c, [&](long k) { return V4<T>{Rv[k], RP[k], Z[k], ZP[k]}; },
            [&](long k, const V4<T> &v) {
                const T dr = v.a - v.b;
                Yh[k] = dr;
                Sh[k] = (v.c - v.d) - dr;
            });
        sync(c);
        return;
    }
    double acc[16];
#pragma unroll
    for (int q = 0; q < 16; q++) acc[q] = 0.;
    const T *Ys[8];
#pragma unroll
    for (int s2 = 0; s2 < 8; s2++) Ys[s2] = P.hY + (long) (s2 < m ? s2 : 0) * n;
    stream<4>(
        n, c, [&](long k) { return V4<T>{Rv[k], RP[k], Z[k], ZP[k]}; },
        [&](long k, const V4<T> &v) {
            const T dr = v.a - v.b;
            Yh[k] = dr;
            Sh[k] = (v.c - v.d) - dr;
            const double drd = (double) dr, rd = (double) v.a;
#pragma unroll
            for (int s2 = 0; s2 < 8; s2++) {
                if (s2 < m) {
                    const double ys = (s2 == head) ? drd : (double) Ys[s2][k];
                    acc[s2] += drd * ys;
                    acc[8 + s2] += ys * rd;
                }
            }
        });
    blockSum(acc, ldsd, c);
    if (c.ltid == 0) {
        for (int s2 = 0; s2 < m; s2++) {
            G[head * 8 + s2] = acc[s2];
            G[s2 * 8 + head] = acc[s2];
            gb[s2] = acc[8 + s2];
        }
    }
    sync(c);
}

// Least squares argmin ||Y g - r|| from the Gram matrix: Cholesky of G + eps * max(diag) I in
// FP64 (uniform in every thread).  Returns false when the factorisation breaks down.
__device__ __forceinline__ bool gramSolve(const double *G, const double *gb, int m, double (&g)[8]) {
    double Lm[8][8];
    double dmax = 0.;
    for (int i = 0; i < m; i++) dmax = fmax(dmax, G[i * 8 + i]);
    if (!(dmax > 0.)) return false;
    const double ridge = 1e-12 * dmax;
#pragma unroll
    for (int i = 0; i < 8; i++) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (i < m && j <= i) {
                double v = G[i * 8 + j] + (i == j ? ridge : 0.);
#pragma unroll
                for (int k = 0; k < 8; k++)
                    if (k < j) v -= Lm[i][k] * Lm[j][k];
                if (i == j) {
                    if (!(v > 0.)) return false;
                    Lm[i][i] = sqrt(v);
                } else {
                    Lm[i][j] = v / Lm[j][j];
                }
            }
        }
    }
    double y[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        y[i] = 0.;
        if (i < m) {
            double v = gb[i];
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (k < i) v -= Lm[i][k] * y[k];
            y[i] = v / Lm[i][i];
        }
    }
#pragma unroll
    for (int i = 7; i >= 0; i--) {
        g[i] = 0.;
        if (i < m) {
            double v = y[i];
#pragma unroll
            for (int k = 0; k < 8; k++)
                if (k > i && k < m) v -= Lm[k][i] * g[k];
            g[i] = v / Lm[i][i];
        }
    }
    return true;
}

// d = -S g - r (src/cache.cuh:897-898); with toCand the K0 update cand = z_prev + d is written
// directly (src/cache.cuh:1055-1057) and d itself is not stored.
template <typename T>
__device__ __noinline__ void andersonApply(const Inst<T> &P, const Tc &c, const double (&gs)[8], bool useLS,
                                           const T *Rv, const T *ZP, T *out, bool toCand) {
    const long n = P.sIter;
    const int m = P.andBuff;
    T gg[8];
    const T *Ss[8];
#pragma unroll
    for (int s2 = 0; s2 < 8; s2++) {
        gg[s2] = (useLS && s2 < m) ? (T) gs[s2] : T(0);
        Ss[s2] = P.hS + (long) (s2 < m ? s2 : 0) * n;
    }
    if (!useLS) {
        stream<8>(
            n, c, [&](long k) { return V2<T>{Rv[k], ZP[k]}; },
            [&](long k, const V2<T> &v) {
                const T d = -v.a;
                out[k] = toCand ? v.b + d : d;
            });
    } else {
        stream<4>(
            n, c, [&](long k) { return V2<T>{Rv[k], ZP[k]}; },
            [&](long k, const V2<T> &v) {
                T sg = T(0);
#pragma unroll
                for (int s2 = 0; s2 < 8; s2++)
                    if (s2 < m) sg += Ss[s2][k] * gg[s2];
                const T d = -sg - v.a;
                out[k] = toCand ? v.b + d : d;
            });
    }
    sync(c);
}

// ---------------------------------------------------------------- debug trace
// Index of the first non-finite entry of v[0, n), or -1.  The index travels through the block's
// T-valued max reduction as n - k, exact while n < 2^24 in FP32 (2^53 in FP64).
template <typename T>
__device__ __noinline__ long firstNonfinite(const T *v, long n, T *lds, const Tc &c) {
    T bad = T(0);
    for (long k = c.tid; k < n; k += c.nt)
        if (!isfinite(v[k])) bad = fmax(bad, T(n - k));
    bad = blockMax(bad, lds, c);
    return bad > T(0) ? n - (long) bad : -1;
}

// One record of the reference's debug caches at an error check (src/cache.cuh:947-959):
//   err1, err2  the max-norms of the primal and dual parts of r/gamma - [L*; L] r   (split)
//   err3        ||z - cand||_2 = ||r||_2
//   err0        the reference computes e <- cand_prim - e on a vector e that starts at zero and
//               is only zeroed by reset() (its L* product goes to another buffer, so it does not
//               enter e), then ||e||_2; restated as computed, with e in P.dbgE
//   callsL      calls to L and L* counted so far: two per application of T (src/cache.cuh:704-716,
//               :750, :773); the host adds the count of earlier solves
template <typename T>
__device__ __noinline__ void debugRecord(const Inst<T> &P, const Tc &c, T *lds, long it, long callsL,
                                         const T (&split)[2], const T *RR, const T *C, long &nDbg) {
    T acc[2] = {T(0), T(0)};
    for (long k = c.tid; k < P.sIter; k += c.nt) acc[0] += RR[k] * RR[k];
    T *E = P.dbgE;
    for (long k = c.tid; k < P.sPrim; k += c.nt) {
        const T e = C[k] - E[k];
        E[k] = e;
        acc[1] += e * e;
    }
    blockSum(acc, lds, c);
    if (nDbg < P.dbgCap && c.tid == 0) {
        double *r = P.dbg + nDbg * DBG_WORDS;
        r[0] = (double) it;
        r[1] = (double) callsL;
        r[2] = sqrt((double) acc[1]);
        r[3] = (double) split[0];
        r[4] = (double) split[1];
        r[5] = sqrt((double) acc[0]);
    }
    nDbg++;
    sync(c);
}

// ---------------------------------------------------------------- the solve
// SuperMann with Anderson directions (src/cache.cuh:1008-1132) and runCp (:967-1002).
// The six iterate buffers rotate by pointer instead of being copied: acceptCandidate,
// backup/restore and saveToPrev are pointer moves.
template <typename T, int R, int V>
__device__ void solveInstance(const Inst<T> &P, Out &out, T *lds, double *ldsd, double tickRate, int grid) {
    const Tc c = threadCtx(grid, P.gpart);
    const long n = P.sIter;
    const int m = P.andBuff;
    const bool gram = P.lsMethod == 1;
    __shared__ double G[64], gb[8];
    T *Z = P.z, *C = P.cand, *ZP = P.zPrev, *RR = P.res, *RP = P.resPrev, *BK = P.backup;
    T *D = P.dir, *DS = P.dirS;
    int head = m - 1;
    const long t0 = wall_clock64();
    long nT = 0, nEll = 0, nDbg = 0, nonfinite = -1;
    PhaseClock clk;
    clk.start();
    // initialisePrev / initialiseIter (src/cache.cuh:611-632)
    if (P.method == METHOD_SPOCK) {
        if (P.warm) {
            stream<8>(
                n, c, [&](long k) { return V2<T>{Z[k], RR[k]}; },
                [&](long k, const V2<T> &v) {
                    ZP[k] = v.a;
                    RP[k] = v.b;
                });
        } else {
            stream<8>(
                n, c, [&](long k) { return 0; },
                [&](long k, int) {
                    ZP[k] = T(1);
                    RP[k] = T(1);
                });
        }
    }
    if (c.ltid < 64) G[c.ltid] = 0.;
    sync(c);
    // a given tolerance (> 0) skips the reference's first-check calibration, so warm starts and
    // polish solves are accepted against the threshold of the cold solve they continue
    bool initFlag = !(P.tolFixed > T(0));
    int status = ST_NOTRUN;
    long iters = 0;
    T tol = initFlag ? T(0) : P.tolFixed, lastErr = T(NAN), wLast = T(NAN);
    const T c0 = T(0.99), c1 = T(0.99), c2 = T(0.99), beta = T(0.5), sigma = T(0.1), lambda = T(1);
    const T inf = T(INFINITY);
    T zeta = inf, w = inf, wSafe = inf;
    long k0 = 0, k1 = 0, k2 = 0, k2bt = 0, k3 = 0;
    // the iteration cap is only tested at error checks (src/cache.cuh:937-940): the last outer
    // iteration is the first check index at or beyond maxIters
    const long hardCap =
        P.maxIters > 0 ? (P.maxIters + P.checkPeriod - 1) / P.checkPeriod * P.checkPeriod + 1 : 2000000;
    long it = 0;
    for (it = 0; it < hardCap; it++) {
        if (it != 0 && P.method == METHOD_SPOCK) {  // acceptCandidate: z <- cand
            T *tmp = Z;
            Z = C;
            C = tmp;
        }
        clk.mark(PH_OTHER);
        applyT<T, R, V>(P, c, Z, C, &clk);
        nT++;
        // computeError (src/cache.cuh:910-961): r = z - cand
        if (P.dbg) {  // isFinite(candidate) (src/cache.cuh:912): stop at the first non-finite entry
            const long bad = firstNonfinite<T>(C, n, lds, c);
            if (bad >= 0) {
                status = ST_NONFINITE;
                iters = it;
                nonfinite = bad;
                break;
            }
        }
        stream<8>(
            n, c, [&](long k) { return V2<T>{Z[k], C[k]}; }, [&](long k, const V2<T> &v) { RR[k] = v.a - v.b; });
        sync(c);
        clk.mark(PH_RES);
        bool haveDot = false;
        T dotRR = T(0);
        if (it % P.checkPeriod == 0) {
            T mx;
            if (P.dbg) {
                T split[2];
                reduceEll<T, R, V, true>(P, c, RR, RR, lds, dotRR, mx, split);
                debugRecord<T>(P, c, lds, it, 2 * nT, split, RR, C, nDbg);
            } else {
                reduceEll<T, R, V>(P, c, RR, RR, lds, dotRR, mx);
            }
            nEll++;
            haveDot = true;
            clk.mark(PH_ELL);
            if (initFlag) {
                tol = fmax(P.tolAbs, P.tolRel * mx);
                initFlag = false;
            } else {
                lastErr = mx;
                int code = ST_NOTRUN;
                if (mx <= tol) code = ST_CONVERGED;
                if (P.maxIters && it >= P.maxIters) code = ST_OUTOFITER;
                if (P.maxTicks > 0 && it % 1000 == 0) {
                    // threads read the clock at different moments: decide on the latest reading
                    const T late = (double) (wall_clock64() - t0) >= P.maxTicks ? T(1) : T(0);
                    if (blockMax(late, lds, c) > T(0)) code = ST_OUTOFTIME;
                }
                if (code != ST_NOTRUN) {
                    status = code;
                    iters = it;
                }
            }
        }
        if (status != ST_NOTRUN || P.method == METHOD_CP) {  // acceptCandidate (and stop)
            T *tmp = Z;
            Z = C;
            C = tmp;
            if (status != ST_NOTRUN) break;
            continue;
        }
        {  // backup (src/cache.cuh:804-806): keep T(z) aside, continue in a fresh buffer
            T *tmp = BK;
            BK = C;
            C = tmp;
        }
        // normM(r) (src/cache.cuh:740-742)
        if (!haveDot) {
            T mx;
            reduceEll<T, R, V>(P, c, RR, RR, lds, dotRR, mx);
            nEll++;
            clk.mark(PH_ELL);
        }
        w = sqrt(dotRR);
        wLast = w;
        // delta iterate/residual, new Anderson column, then saveToPrev as pointer moves
        head = (head + 1) % m;
        andersonColumn<T>(P, c, head, RR, RP, Z, ZP, gram, G, gb, ldsd);
        clk.mark(PH_ACOL);
        {
            T *tmp = ZP;
            ZP = Z;
            Z = tmp;
            tmp = RP;
            RP = RR;
            RR = tmp;
        }
        // updateDirection (src/cache.cuh:861-904); r now lives in RP
        double gs[8];
#pragma unroll
        for (int q = 0; q < 8; q++) gs[q] = 0.;
        bool useLS = false;
        if (it >= m) {
            if (gram) {
                useLS = gramSolve(G, gb, m, gs);
            } else {
                andersonHouseholder<T>(P, c, head, RP, lds, gs);
                useLS = true;
            }
        }
        clk.mark(PH_ALS);
        const bool blind = w <= c0 * zeta;
        andersonApply<T>(P, c, gs, useLS, RP, ZP, blind ? C : D, blind);
        clk.mark(PH_AAPPLY);
        if (blind) {  // K0 blind update (src/cache.cuh:1055-1061)
            zeta = w;
            k0++;
            continue;
        }
        T tau = T(1);
        for (int in = 0; in < P.maxInner; in++) {  // line search (src/cache.cuh:1063-1099)
            stream<8>(
                n, c, [&](long k) { return V2<T>{D[k], ZP[k]}; },
                [&](long k, const V2<T> &v) {
                    const T ds = v.a * tau;
                    DS[k] = ds;
                    Z[k] = v.b + ds;
                });
            sync(c);
            clk.mark(PH_LINE);
            applyT<T, R, V>(P, c, Z, C, &clk);
            nT++;
            stream<8>(
                n, c, [&](long k) { return V2<T>{Z[k], C[k]}; }, [&](long k, const V2<T> &v) { RR[k] = v.a - v.b; });
            sync(c);
            clk.mark(PH_RES);
            T dot, mx;
            reduceEll<T, R, V>(P, c, RR, RR, lds, dot, mx);
            nEll++;
            clk.mark(PH_ELL);
            const T wTilde = sqrt(dot);
            if (w <= wSafe && wTilde <= c1 * w) {  // K1: cand <- z
                T *tmp = C;
                C = Z;
                Z = tmp;
                wSafe = wTilde + pow(c2, T(it));
                k1++;
                break;
            }
            T dotRD;
            reduceEll<T, R, V>(P, c, RR, DS, lds, dotRD, mx);
            nEll++;
            clk.mark(PH_ELL);
            const T rho = wTilde * wTilde - dotRD;
            if (rho >= sigma * wTilde * w) {  // K2
                const T f = lambda * rho / (wTilde * wTilde);
                stream<8>(
                    n, c, [&](long k) { return V2<T>{RR[k], ZP[k]}; },
                    [&](long k, const V2<T> &v) {
                        const T rk = v.a * f;
                        RR[k] = rk;
                        C[k] = v.b - rk;
                    });
                sync(c);
                k2++;
                break;
            } else {
                tau *= beta;
                k2bt++;
            }
            if (in >= P.maxInner - 1) {  // K3: restore the plain CP step
                T *tmp = C;
                C = BK;
                BK = tmp;
                k3++;
            }
        }
    }
    if (status == ST_NOTRUN) {
        status = ST_OUTOFITER;
        iters = it;
    }
    // the accepted iterate must end in P.z and the last residual in P.res; both sources are
    // read before either destination is written, so rotated roles cannot clobber each other
    if (Z != P.z || RR != P.res) {
        T *dz = P.z, *dr = P.res;
        stream<8>(
            n, c, [&](long k) { return V2<T>{Z[k], RR[k]}; },
            [&](long k, const V2<T> &v) {
                dz[k] = v.a;
                dr[k] = v.b;
            });
    }
    sync(c);
    if (c.tid == 0) {
        out.nDbg = nDbg;
        out.nonfinite = nonfinite;
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
        clk.mark(PH_OTHER);
        for (int i = 0; i < NUM_PHASES; i++) out.ticks[i] = clk.acc[i];
    }
}

// Copies the instance descriptor into LDS and, when they fit in the dynamic LDS budget,
// its per-node index arrays and tree arrays, repointing the copy at the LDS arrays.
template <typename T>
__device__ const Inst<T> &stageInstance(const Inst<T> &G, Inst<T> &L, int *pool, int poolInts) {
    const Tc c = threadCtx();
    if (c.tid == 0) L = G;
    __syncthreads();
    if (poolInts <= 0) return L;
    int used = 0;
    auto take = [&](const int *&p, int count) {
        if (!p || count <= 0 || used + count > poolInts) return;
        int *dst = pool + used;
        for (int i = c.tid; i < count; i += c.nt) dst[i] = p[i];
        used += count;
        __syncthreads();
        if (c.tid == 0) p = dst;
        __syncthreads();
    };
    const int N = G.N, nl = G.nl, nlf = G.nlf;
    take(L.anc, N);
    take(L.chFrom, nl);
    take(L.numCh, nl);
    take(L.stageFrom, G.S);
    take(L.stageTo, G.S);
    take(L.BTr.idx, N);
    take(L.AB.idx, N);
    take(L.ABKTr.idx, N);
    take(L.APB.idx, N);
    take(L.sQ.idx, N);
    take(L.sR.idx, N);
    take(L.trans.idx, N);
    take(L.q.idx, N);
    take(L.r.idx, N);
    take(L.K.idx, nl);
    take(L.KTr.idx, nl);
    take(L.RtInv.idx, nl);
    take(L.cholL.idx, nl);
    take(L.cholU.idx, nl);
    take(L.NN.idx, nl);
    take(L.b.idx, nl);
    take(L.nLB.idx, nl);
    take(L.nUB.idx, nl);
    take(L.nG.idx, nl);
    take(L.nGT.idx, nl);
    take(L.lsQ.idx, nlf);
    take(L.lq.idx, nlf);
    take(L.ltrans.idx, nlf);
    take(L.lLB.idx, nlf);
    take(L.lUB.idx, nlf);
    take(L.lG.idx, nlf);
    take(L.lGT.idx, nlf);
    take(L.e.idx, N);
    take(L.Pe.idx, N);
    return L;
}

// grid = 0: workgroup k solves instance k.  grid = 1 (cooperative launch, count = 1): every
// workgroup works on instance 0.
template <typename T, int R, int V>
__global__ void __launch_bounds__(SPK_MAX_THREADS) k_solve(const Inst<T> *__restrict__ insts, Out *outs, int count,
                                                double tickRate, int poolInts, int grid) {
    __shared__ T lds[8 * 32];
    __shared__ double ldsd[16 * 32];
    __shared__ __align__(16) char localBuf[sizeof(Inst<T>)];
    Inst<T> &local = *reinterpret_cast<Inst<T> *>(localBuf);
    extern __shared__ int pool[];
    const int k = grid ? 0 : (int) blockIdx.x;
    if (k >= count) return;
    const Inst<T> &P = stageInstance(insts[k], local, pool, poolInts);
    solveInstance<T, R, V>(P, outs[P.slot], lds, ldsd, tickRate, grid);
}

// ---------------------------------------------------------------- single-operator kernels
// Debug entry points used by the parity tests: they read P.z and write P.cand.
enum DebugOp { OP_L = 0, OP_LADJ = 1, OP_DYN = 2, OP_KERNEL = 3, OP_T = 4, OP_ELL = 5, OP_DUAL = 6 };

template <typename T, int R, int V>
__global__ void __launch_bounds__(SPK_MAX_THREADS) k_debug(const Inst<T> *__restrict__ insts, Out *outs, int count, int op,
                                                int poolInts, int grid) {
    __shared__ T lds[8 * 32];
    __shared__ __align__(16) char localBuf[sizeof(Inst<T>)];
    Inst<T> &local = *reinterpret_cast<Inst<T> *>(localBuf);
    extern __shared__ int pool[];
    const int kInst = grid ? 0 : (int) blockIdx.x;
    if (kInst >= count) return;
    const Inst<T> &P = stageInstance(insts[kInst], local, pool, poolInts);
    const Tc c = threadCtx(grid, P.gpart);
    T *z = P.z, *cand = P.cand;
    if (op == OP_L) {
        StoreL<T, R> h{P, cand + P.sPrim, c.lane};
        auto pv = [&](long k) { return z[k]; };
        for (int n = c.wave; n < P.N; n += c.nw) opNode<T, R, V>(P, n, c.lane, pv, h);
    } else if (op == OP_LADJ) {
        const T *zd = z + P.sPrim;
        for (int n = c.wave; n < P.N; n += c.nw) {
            AdjOut<T, R> o;
            adjNode<T, R, V>(P, zd, n, c.lane, o);
            for (int rc = 0; rc < R; rc++) {
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
        sync(c);
        if (op == OP_DYN)
            for (int n = c.wave; n < P.N; n += c.nw)
                if (n >= P.nl)
                    for (int r = c.lane; r < P.nx; r += 64) P.qd[(long) n * P.nx + r] = -cand[P.oX + (long) n * P.nx + r];
        sync(c);
        phaseProjections<T, R, V>(P, c, cand, op == OP_DYN, op == OP_KERNEL);
    } else if (op == OP_T) {
        applyT<T, R, V>(P, c, z, cand);
    } else if (op == OP_DUAL) {  // Cache::getDual (src/cache.cuh:371-376): dual prox of the given primal
        phaseDualProx<T, R, V>(P, c, z, cand);
    } else if (op == OP_ELL) {
        T dot, mx;
        reduceEll<T, R, V>(P, c, z, cand, lds, dot, mx);
        if (c.tid == 0) {
            outs[P.slot].w = (double) dot;
            outs[P.slot].err = (double) mx;
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

// explicit instantiations: register chunk classes (R, V) = (1,1) for dims <= 64,
// and (4,4) for dims <= 256
#define SPK_INST(T, R, V)                                                                  \
    template __global__ void k_solve<T, R, V>(const Inst<T> *, Out *, int, double, int, int); \
    template __global__ void k_debug<T, R, V>(const Inst<T> *, Out *, int, int, int, int);
SPK_INST(double, 1, 1)
#ifndef SPK_FAST_BUILD
SPK_INST(double, 4, 4)
SPK_INST(float, 1, 1)
SPK_INST(float, 4, 4)
#endif
#undef SPK_INST
template __global__ void k_fill<double>(double *, double, long);
template __global__ void k_fill<float>(float *, float, long);

}  // namespace spk
