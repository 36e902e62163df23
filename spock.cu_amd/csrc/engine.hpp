// Device-side data model of the fused SPOCK engine (shared by the kernels and the host glue).
#pragma once
#include <hip/hip_runtime.h>

namespace spk {

// Per-node factor table on the device: entry for node n starts at data + idx[n] * rows * cols.
template <typename T>
struct DTab {
    const T *data = nullptr;
    const int *idx = nullptr;
    int rows = 0, cols = 0;
    __device__ __forceinline__ const T *at(int n) const { return data + (size_t) idx[n] * rows * cols; }
};

// One solve instance: problem sizes, device tables, state buffers and options.
template <typename T>
struct Inst {
    int N, nl, nlf, S, E, nx, nu, yD, nD, ncost, lcost, ncon, lcon, con3, con5;
    long oU, oX, oY, oT, oS, sPrim, oI, oII, oIII, oIV, oV, oVI, sDual, sIter;
    T gamma, gammaRecip;
    const int *anc, *chFrom, *numCh, *stageFrom, *stageTo;
    DTab<T> BTr, AB, K, KTr, ABKTr, APB, cholL, cholU, RtInv, NN, b, sQ, sR, q, r, trans, lsQ, lq, ltrans;
    DTab<T> nG, nLB, nUB, lG, lLB, lUB, nGT, lGT;
    DTab<T> e, Pe;  // affine dynamics (zero tables when linear)
    int affine;
    // state (iterate-sized unless noted)
    T *z, *zPrev, *cand, *backup, *res, *resPrev, *dir, *dirS;
    T *hS, *hY;    // Anderson histories (iterate and residual differences), andBuff columns each
    T *W;          // Householder scratch, andBuff + 1 columns
    T *qd, *dd;    // dynamics sweep intermediates: nx*N and nu*nl
    const T *x0s;  // scaled initial state (nx)
    // options
    T tolAbs, tolRel;  // already multiplied by min(scaling) when preconditioned
    T tolFixed;        // > 0: accept when the error is <= tolFixed from the first check on
    long maxIters;
    int maxInner, andBuff, checkPeriod, method, warm, lsMethod, dynInverse;
    double maxTicks;  // wall-clock cap in device ticks (0 = none)
    T *gpart;         // grid mode: 2 x GRID_MAXK x GRID_MAXWG reduction partials
    // debug trace (enableDebug, src/cache.cuh:947-959); dbg == nullptr: off.  One record of
    // DBG_WORDS doubles per error check, at most dbgCap; dbgE (sPrim) carries err0's running
    // vector from check to check and from solve to solve
    double *dbg;
    long dbgCap;
    T *dbgE;
    int slot;  // this instance's index in the batch (its result record); descriptors may be reordered
};

// A debug record: outer iteration, calls to L and L* so far, err0, err1, err2, err3.
constexpr int DBG_WORDS = 6;

// Grid mode limits: values per grid reduction and workgroups per cooperative launch.
constexpr int GRID_MAXK = 16;
constexpr int GRID_MAXWG = 1024;

// Block-interleaved layout of a region of `rows` rows: instances are grouped in blocks of w (64
// for the lane engine, 16 for the tile engine) and each block stores rows x w words contiguously,
// element e of instance b at ((b / w) * rows + e) * w + b % w.
__host__ __device__ inline size_t laneIndex(long rows, long e, long b, long w = 64) {
    return ((size_t) (b / w) * rows + e) * w + (b % w);
}

// SuperMann state of one lane kept between the launches of a chunked solve.
struct LaneSave {
    double G[4][4], gb[4];
    double tol, lastErr, wLast, zeta, w, wSafe;
    long iters, k0, k1, k2, k2bt, k3, nT, nEll;
    long it0;  // workgroup iteration at which this lane's current solve started
    int status, active, initFlag, pad;
};

// Per-workgroup state of a chunked solve: buffer roles (indices of z, cand, zPrev, res, resPrev,
// backup into work[0..5]), Anderson ring head, whether the workgroup's solve has ended (its
// results are final and a resume returns at once), next outer iteration and device ticks spent.
struct LaneGroupSave {
    int roles[6];
    int head, finished;
    long it, ticks;
};

// Many instances of ONE problem in the block-interleaved layout: a wave whose lanes are the 64
// instances of one block moves 64 consecutive words per access, a workgroup's state is one
// contiguous slab per region, and the factor tables are read at wave-uniform addresses (scalar
// loads shared by the wave).
template <typename T>
struct LaneInst {
    int N, nl, nlf, S, E, nx, nu, yD, nD, ncost, lcost, ncon, lcon, con3, con5;
    long oU, oX, oY, oT, oS, sPrim, oI, oII, oIII, oIV, oV, oVI, sDual, sIter;
    T gamma, gammaRecip;
    const int *anc, *chFrom, *numCh, *stageFrom, *stageTo;
    DTab<T> BTr, AB, K, KTr, ABKTr, APB, cholL, cholU, RtInv, NN, b, sQ, sR, q, r, trans, lsQ, lq, ltrans;
    DTab<T> nG, nLB, nUB, lG, lLB, lUB, nGT, lGT;
    DTab<T> e, Pe;  // affine dynamics (zero tables when linear)
    int affine;
    long B;     // padded instance count (multiple of 64); a region of r rows holds r * B words
    int count;  // instances in use
    T *work[9];   // rotating roles z, cand, zPrev, res, resPrev, backup + cand2, dir, dirS
    T *zio, *rio; // accepted iterate and its residual (warm-start input, solve output)
    T *hS, *hY;   // Anderson rings, andBuff slots of sIter x B each
    T *qd, *dd;   // dynamics sweep intermediates (nx*N and nu*nl rows)
    const T *x0s; // scaled initial states (nx rows)
    const T *tolFixed;  // per-instance given tolerance (> 0), or nullptr
    T tolAbs, tolRel;
    long maxIters;
    int maxInner, andBuff, checkPeriod, method, warm;
    double maxTicks;
    long chunkIters;          // > 0: stop after this many outer iterations, state saved for a resume
    double chunkTicks;        // > 0 (chunked solves): also stop, after at least one outer iteration, once
                              // the launch has run this many wall-clock ticks
    int resume;               // 1: continue the saved solve
    LaneSave *save;           // B entries
    LaneGroupSave *gsave;     // B / 64 entries
};

// Phases timed inside the solve (wall_clock64 ticks, summed over the solve).
enum Phase { PH_ADJ = 0, PH_PROJ = 1, PH_DUAL = 2, PH_RES = 3, PH_ELL = 4, PH_ACOL = 5, PH_ALS = 6, PH_AAPPLY = 7,
             PH_LINE = 8, PH_OTHER = 9, NUM_PHASES = 10 };

struct Out {
    int status;
    int pad;
    long iters;
    double err, tol, w;
    long k0, k1, k2, k2bt, k3, nT, nEll;
    long ticks[NUM_PHASES];  // device wall-clock ticks spent per phase (thread 0's view)
    long nDbg;               // debug records written by the last solve
    long nonfinite;          // debug: first non-finite entry of the candidate (-1: none)
};

enum Method { METHOD_SPOCK = 0, METHOD_CP = 1 };
enum DevKinds { CON_NO_DEV = 0, CON_RECT_DEV = 1, CON_POLY_DEV = 2, COST_QUAD_DEV = 0, COST_QUADLIN_DEV = 1,
                COST_LIN_DEV = 2 };

// A linear cost's dual "projection" (src/costs.cuh:296-301: k_projectRectangle of
// src/definitions.cu:39-48 with lower bound -inf and the epigraph entry as upper bound): the cost
// entry is clamped to at most the epigraph entry, which is kept.  Not the Euclidean projection
// onto {c <= t}; restated as the reference computes it.
template <typename T>
__host__ __device__ __forceinline__ T projBelow(T c, T t) { return c >= t ? t : c; }
// ST_NONFINITE: debug mode found a non-finite candidate entry and stopped the solve, where the
// reference's isFinite throws (src/cache.cuh:52-62, :912)
enum Status { ST_NOTRUN = -1, ST_CONVERGED = 0, ST_OUTOFITER = 1, ST_OUTOFTIME = 2, ST_NONFINITE = 3 };

constexpr int MAXR = 4;  // row chunks of 64 per wave task (rows <= 256)
constexpr int MAXV = 4;  // vector chunks of 64 held in registers (length <= 256)

}  // namespace spk
