// Host side of the C ABI (include/spock_amd.h): contexts, device-resident problems,
// batches of solve instances and the launches of the fused engine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/spock_amd.h"
#include "engine.hpp"
#include "loader.hpp"
#include "common.hpp"

// The scratch of spk_batch_solve_noalloc: the instance descriptors, Anderson rings and grid-mode
// partials of one per-instance solve, laid out from offset 0 (256-byte aligned regions).  Each
// device holds its own copy until the process ends.  Host code never names the array: it uses the
// address hipGetSymbolAddress returned for the device (scratchOf), taken once per device by
// spk_ctx_create with that device current.  No region needs to start at zero: the descriptors are
// copied in before the launch, an Anderson column enters the least-squares system only once the
// solve has written it (the first m iterations take the plain step), and every grid partial is
// written before the grid barrier that precedes its read.  spk_batch_solve reuses its own rings
// uncleared for the same reason.
static __device__ __align__(256) unsigned char scratch[SPK_NOALLOC_SCRATCH_BYTES];

namespace spk {
template <typename T, int R, int V>
__global__ void k_solve(const Inst<T> *, Out *, int, double, int, int);
template <typename T, int R, int V>
__global__ void k_debug(const Inst<T> *, Out *, int, int, int, int);
__global__ void k_f32_to_f64(const float *, double *, long);
__global__ void k_f64_to_f32(const double *, float *, long);
template <typename T>
__global__ void k_fill(T *, T, long);
template <typename T>
__global__ void k_lane_solve(const LaneInst<T> *, Out *);
template <typename T>
__global__ void k_lane_debug(const LaneInst<T> *, Out *, int);
template <typename T>
__global__ void k_lane_solve_x8(const LaneInst<T> *, Out *);
template <typename T>
__global__ void k_lane_debug_x8(const LaneInst<T> *, Out *, int);
template <typename T>
__global__ void k_tile_solve(const LaneInst<T> *, Out *);
template <typename T>
__global__ void k_tile_debug(const LaneInst<T> *, Out *, int);
template <typename T>
__global__ void k_tile_solve_x1024(const LaneInst<T> *, Out *);
template <typename T>
__global__ void k_tile_debug_x1024(const LaneInst<T> *, Out *, int);
template <typename T>
__global__ void k_relayout(const T *, long, T *, long, long, long);
template <typename T, typename Z>
__global__ void k_lane_refill(const LaneInst<T> *, const int *, int, const Z *, long, const T *, long);
template <typename T>
__global__ void k_lane_refill_groups(const LaneInst<T> *, const int *, int, long);
template <typename S, typename D>
__global__ void k_gather_lanes(const S *, long, D *, long, long, const int *, int);
template <typename S, typename D>
__global__ void k_interleave(const S *, long, D *, long, int, long);
template <typename S, typename D>
__global__ void k_deinterleave(const S *, long, D *, long, long, int);

// One instance's column of a block-interleaved region (`rows` rows, blocks of w instances) copied
// to a contiguous row (toRow 1) or back from one (toRow 0).
template <typename T>
__global__ void k_lane_column(T *region, long rows, long k, long w, T *row, int toRow) {
    for (long e = blockIdx.x * (long) blockDim.x + threadIdx.x; e < rows; e += (long) gridDim.x * blockDim.x) {
        const size_t at = laneIndex(rows, e, k, w);
        if (toRow) row[e] = region[at];
        else region[at] = row[e];
    }
}
}  // namespace spk

using spk::HostProblem;
using spk::Table;

namespace {

using spk::check;
using spk::DevBuf;
using spk::fail;

template <typename T>
struct DevTableSet {  // device copies of one table: data as T, index as int
    std::unique_ptr<DevBuf> data, idx;
    int rows = 0, cols = 0;
};

}  // namespace

struct spk_problem : spk::Counted {
    spk_ctx *ctx = nullptr;  // holds a reference
    explicit spk_problem(spk_ctx *c) : ctx(spk::retain(c)) {}
    ~spk_problem() { spk::releaseCtx(ctx); }
    spk_problem(const spk_problem &) = delete;
    spk_problem &operator=(const spk_problem &) = delete;
    HostProblem hp;
    DevBuf anc, chFrom, numCh, stageFrom, stageTo;
    std::map<std::string, DevTableSet<double>> t64;
    std::map<std::string, DevTableSet<float>> t32;
    bool up64 = false, up32 = false;
};

struct spk_batch {
    spk_ctx *ctx = nullptr;  // holds a reference, as does every entry of heldProbs
    explicit spk_batch(spk_ctx *c) : ctx(spk::retain(c)) {}
    ~spk_batch();
    spk_batch(const spk_batch &) = delete;
    spk_batch &operator=(const spk_batch &) = delete;
    int dtype = SPK_F64;
    int count = 0;
    std::vector<spk_problem *> probs;
    std::vector<spk_problem *> heldProbs;  // the distinct problems of probs, one reference each
    void holdProblems() {
        for (auto *p : probs)
            if (p && std::find(heldProbs.begin(), heldProbs.end(), p) == heldProbs.end())
                heldProbs.push_back(spk::retain(p));
    }
    std::vector<size_t> coreOff, andOff, x0Off;  // element offsets of each instance's regions
    size_t coreElems = 0, andElems = 0, x0Elems = 0;
    int mAlloc = 0;
    DevBuf core, anders, x0buf, dInst, dOut;  // x0buf: scaled initial states (inputs, kept by reset)
    int wgs = 0;    // workgroups per instance (0 or 1: one; see gridWorkgroups)
    DevBuf gpart;   // grid-mode reduction partials (single-instance batches)
    std::vector<char> hInst;  // array of Inst<T>
    std::vector<double> tolFixed;  // per-instance given tolerance (0 = calibrate at the first check)
    // shared-problem batch in the instance-interleaved layout (lane engine)
    bool lanes = false;
    long Bpad = 0;     // count rounded up to 64
    int laneW = 16;    // instances per block of the interleaved layout: 16 (tile engine) or 64 (lane engine)
    DevBuf lbuf, ltol; // interleaved state; per-instance given tolerances
    DevBuf lsave, lgsave;   // solver state kept between the launches of a chunked solve
    spk_options lastOpts{};
    int lastWarm = 0;
    bool resumable = false;
    double chunkSeconds = 0.;  // wall-clock slice of each launch of a chunked solve (0: none)
    size_t lRegion[16] = {};  // element offsets: work[0..8], zio, rio, hS, hY, qd, dd, x0
    static constexpr int LANE_MAXM = 4;
    int laneMv = 17;     // schedule: 16, 17 the tile engine (MFMA; 512- or 1024-thread build); 0, 8 the lane engine
    DevBuf dInstResume;  // the lane instance of the resume launches of a captured chunked solve
    DevBuf rLanes, rX0, rZ;  // staging of a refill: lane indices, scaled initial states, starting iterates
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    int laneBlockSz = 0; // workgroup size chosen by spk_batch_autotune (0: none)
    bool tuned = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool pending = false;
    // debug trace of per-instance batches (spk_batch_set_debug): records, err0's running vectors
    long dbgCap = 0;
    DevBuf dbgRec, dbgE;
    std::vector<size_t> dbgEOff;
    // per-instance batches: workgroup size and workgroup order chosen by spk_batch_autotune_uneven
    // (0 / false: 256 threads, instances in batch order)
    int fusedBlock = 0;
    bool largestFirst = false;
    // per-instance batches: graph/graphExec hold the captured persistent solve launch
    // (spk_batch_solve_graph); any reallocation of the buffers it points to drops it
    bool instGraph = false;
    size_t elSize() const { return dtype == SPK_F64 ? 8 : 4; }
};

namespace {

template <typename T>
std::map<std::string, DevTableSet<T>> &tablesOf(spk_problem *p);
template <>
std::map<std::string, DevTableSet<double>> &tablesOf<double>(spk_problem *p) { return p->t64; }
template <>
std::map<std::string, DevTableSet<float>> &tablesOf<float>(spk_problem *p) { return p->t32; }

void uploadInts(DevBuf &b, const std::vector<int> &v) {
    b.alloc(v.size() * sizeof(int));
    if (!v.empty()) check(hipMemcpy(b.p, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice), "upload ints");
}

template <typename T>
void uploadTable(spk_problem *p, const std::string &name, const Table &t) {
    DevTableSet<T> s;
    s.rows = t.rows;
    s.cols = t.cols;
    std::vector<T> d(t.data.begin(), t.data.end());
    // 64 zero elements of padding: the lane engine reads fixed row blocks past a table's end
    d.resize(d.size() + 64, T(0));
    s.data = std::make_unique<DevBuf>();
    s.data->alloc(d.size() * sizeof(T));
    check(hipMemcpy(s.data->p, d.data(), d.size() * sizeof(T), hipMemcpyHostToDevice), "upload table");
    s.idx = std::make_unique<DevBuf>();
    s.idx->alloc(t.index.size() * sizeof(int));
    if (!t.index.empty())
        check(hipMemcpy(s.idx->p, t.index.data(), t.index.size() * sizeof(int), hipMemcpyHostToDevice), "upload index");
    tablesOf<T>(p)[name] = std::move(s);
}

#define SPK_TABLES(X) \
    X(BTr) X(AB) X(K) X(KTr) X(ABKTr) X(APB) X(cholL) X(cholU) X(RtInv) X(NN) X(b) X(sQ) X(sR) X(q) X(r) X(trans) \
    X(lsQ) X(lq) X(ltrans) X(nG) X(nLB) X(nUB) X(lG) X(lLB) X(lUB) X(nGT) X(lGT) X(e) X(Pe)

template <typename T>
void ensureUploaded(spk_problem *p) {
    bool &done = std::is_same<T, double>::value ? p->up64 : p->up32;
    if (done) return;
    check(hipSetDevice(p->ctx->device), "hipSetDevice");
#define UP(name) uploadTable<T>(p, #name, p->hp.name);
    SPK_TABLES(UP)
#undef UP
    done = true;
}

template <typename T>
spk::DTab<T> dtab(spk_problem *p, const std::string &name) {
    auto &m = tablesOf<T>(p);
    auto it = m.find(name);
    spk::DTab<T> t;
    if (it == m.end()) return t;
    t.data = static_cast<const T *>(it->second.data->p);
    t.idx = static_cast<const int *>(it->second.idx->p);
    t.rows = it->second.rows;
    t.cols = it->second.cols;
    return t;
}

size_t alignUp(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Per-instance element counts of the two state regions.
size_t coreCount(const HostProblem &h) {
    return alignUp(8 * (size_t) h.sIter + (size_t) h.nx * h.N + (size_t) h.nu * h.nl, 64);
}
size_t andCount(const HostProblem &h, int m) { return alignUp((size_t) (3 * m + 1) * h.sIter, 64); }

// Problem sizes, offsets, tree arrays and device tables shared by both instance descriptors.
template <typename T, typename I>
void fillProblem(I &I_, spk_problem *p) {
    const HostProblem &h = p->hp;
    I_.N = h.N; I_.nl = h.nl; I_.nlf = h.nlf; I_.S = h.S; I_.E = h.E; I_.nx = h.nx; I_.nu = h.nu;
    I_.yD = h.yD; I_.nD = h.nD; I_.ncost = h.ncost; I_.lcost = h.lcost; I_.ncon = h.ncon; I_.lcon = h.lcon;
    I_.con3 = h.con3; I_.con5 = h.con5;
    I_.affine = h.affine ? 1 : 0;
    I_.oU = 0; I_.oX = h.sU; I_.oY = I_.oX + h.sX; I_.oT = I_.oY + h.sY; I_.oS = I_.oT + h.sT; I_.sPrim = h.sPrim;
    I_.oI = 0; I_.oII = h.sI; I_.oIII = I_.oII + h.sII; I_.oIV = I_.oIII + h.sIII; I_.oV = I_.oIV + h.sIV;
    I_.oVI = I_.oV + h.sV; I_.sDual = h.sDual; I_.sIter = h.sIter;
    I_.gamma = (T) h.stepSize;
    I_.gammaRecip = (T) (1. / (double) I_.gamma);
    I_.anc = static_cast<const int *>(p->anc.p);
    I_.chFrom = static_cast<const int *>(p->chFrom.p);
    I_.numCh = static_cast<const int *>(p->numCh.p);
    I_.stageFrom = static_cast<const int *>(p->stageFrom.p);
    I_.stageTo = static_cast<const int *>(p->stageTo.p);
#define DT(name) I_.name = dtab<T>(p, #name);
    SPK_TABLES(DT)
#undef DT
}

double tolScaleOf(const HostProblem &h) {
    return h.precond ? *std::min_element(h.scaling.begin(), h.scaling.end()) : 1.;
}

// The instance descriptors of a per-instance solve in b->hInst: Anderson rings of instance k at
// anders + andOff[k], grid-mode partials at gpart.
template <typename T>
void fillInsts(spk_batch *b, const spk_options &o, int warm, T *anders, const std::vector<size_t> &andOff,
               T *gpart) {
    b->hInst.assign(sizeof(spk::Inst<T>) * b->count, 0);
    auto *arr = reinterpret_cast<spk::Inst<T> *>(b->hInst.data());
    T *core = static_cast<T *>(b->core.p);
    T *and_ = anders;
    for (int k = 0; k < b->count; k++) {
        spk_problem *p = b->probs[k];
        const HostProblem &h = p->hp;
        spk::Inst<T> &I = arr[k];
        I.N = h.N; I.nl = h.nl; I.nlf = h.nlf; I.S = h.S; I.E = h.E; I.nx = h.nx; I.nu = h.nu;
        I.yD = h.yD; I.nD = h.nD; I.ncost = h.ncost; I.lcost = h.lcost; I.ncon = h.ncon; I.lcon = h.lcon;
        I.con3 = h.con3; I.con5 = h.con5;
        I.affine = h.affine ? 1 : 0;
        I.oU = 0; I.oX = h.sU; I.oY = I.oX + h.sX; I.oT = I.oY + h.sY; I.oS = I.oT + h.sT; I.sPrim = h.sPrim;
        I.oI = 0; I.oII = h.sI; I.oIII = I.oII + h.sII; I.oIV = I.oIII + h.sIII; I.oV = I.oIV + h.sIV;
        I.oVI = I.oV + h.sV; I.sDual = h.sDual; I.sIter = h.sIter;
        I.gamma = (T) h.stepSize;
        I.gammaRecip = (T) (1. / (double) I.gamma);
        I.anc = static_cast<const int *>(p->anc.p);
        I.chFrom = static_cast<const int *>(p->chFrom.p);
        I.numCh = static_cast<const int *>(p->numCh.p);
        I.stageFrom = static_cast<const int *>(p->stageFrom.p);
        I.stageTo = static_cast<const int *>(p->stageTo.p);
#define DT(name) I.name = dtab<T>(p, #name);
        SPK_TABLES(DT)
#undef DT
        const size_t n = h.sIter;
        T *c = core + b->coreOff[k];
        I.z = c; I.zPrev = c + n; I.cand = c + 2 * n; I.backup = c + 3 * n; I.res = c + 4 * n;
        I.resPrev = c + 5 * n; I.dir = c + 6 * n; I.dirS = c + 7 * n;
        I.qd = c + 8 * n;
        I.dd = I.qd + (size_t) h.nx * h.N;
        I.x0s = static_cast<const T *>(b->x0buf.p) + b->x0Off[k];
        T *a = and_ + andOff[k];
        const int m = o.anderson_buffer;
        I.hS = a; I.hY = a + (size_t) m * n; I.W = a + (size_t) 2 * m * n;
        double tolScale = 1.;
        if (h.precond) tolScale = *std::min_element(h.scaling.begin(), h.scaling.end());
        I.tolAbs = (T) (o.tol_abs * tolScale);
        I.tolRel = (T) (o.tol_rel * tolScale);
        I.tolFixed = b->tolFixed.empty() ? T(0) : (T) b->tolFixed[k];
        I.maxIters = o.max_iters;
        I.maxInner = o.max_inner;
        I.andBuff = m;
        I.checkPeriod = o.check_period > 0 ? o.check_period : 25;
        I.method = o.method;
        I.warm = warm;
        I.lsMethod = o.ls_method;
        I.dynInverse = o.dyn_inverse;
        I.maxTicks = o.max_time_secs > 0 ? o.max_time_secs * b->ctx->tickRateHz : 0.;
        I.gpart = gpart;
        I.slot = k;
        if (b->dbgCap > 0) {
            I.dbg = static_cast<double *>(b->dbgRec.p) + (size_t) k * b->dbgCap * spk::DBG_WORDS;
            I.dbgCap = b->dbgCap;
            I.dbgE = static_cast<T *>(b->dbgE.p) + b->dbgEOff[k];
        }
    }
}

// Workgroup w solves the instance of descriptor w.  Largest first: the descriptors in order of
// decreasing iterate size (ties in batch order), so the longest solves start in the first dispatch
// wave when a batch holds more instances than the device runs at once.
template <typename T>
void orderLargestFirst(spk_batch *b) {
    auto *arr = reinterpret_cast<spk::Inst<T> *>(b->hInst.data());
    std::vector<spk::Inst<T>> tmp(arr, arr + b->count);
    std::stable_sort(tmp.begin(), tmp.end(),
                     [](const spk::Inst<T> &x, const spk::Inst<T> &y) { return x.sIter > y.sIter; });
    std::copy(tmp.begin(), tmp.end(), arr);
}

void dropInstGraph(spk_batch *b) {
    if (!b->instGraph) return;
    if (b->graphExec) (void) hipGraphExecDestroy(b->graphExec);
    if (b->graph) (void) hipGraphDestroy(b->graph);
    b->graphExec = nullptr;
    b->graph = nullptr;
    b->instGraph = false;
}

template <typename T>
void buildInsts(spk_batch *b, const spk_options &o, int warm) {
    dropInstGraph(b);  // the descriptor buffer is reallocated below
    if (b->count == 1 && !b->gpart.p) b->gpart.alloc((size_t) 2 * spk::GRID_MAXK * spk::GRID_MAXWG * sizeof(T));
    fillInsts<T>(b, o, warm, static_cast<T *>(b->anders.p), b->andOff, static_cast<T *>(b->gpart.p));
    if (b->largestFirst && b->count > 1) orderLargestFirst<T>(b);
    b->dInst.alloc(b->hInst.size());
    check(hipMemcpyAsync(b->dInst.p, b->hInst.data(), b->hInst.size(), hipMemcpyHostToDevice, b->ctx->stream),
          "upload instances");
}

void ensureAnderson(spk_batch *b, int m) {
    if (m <= b->mAlloc) return;
    dropInstGraph(b);
    b->andOff.resize(b->count);
    size_t off = 0;
    for (int k = 0; k < b->count; k++) {
        b->andOff[k] = off;
        off += andCount(b->probs[k]->hp, m);
    }
    b->andElems = off;
    b->anders.alloc(off * b->elSize());
    check(hipMemsetAsync(b->anders.p, 0, off * b->elSize(), b->ctx->stream), "zero anderson");
    b->mAlloc = m;
}

template <typename T>
void toDevice(spk_batch *b, T *dst, const double *src, size_t n) {
    std::vector<T> tmp(src, src + n);
    check(hipMemcpy(dst, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice), "copy to device");
}

template <typename T>
void toHost(spk_batch *b, double *dst, const T *src, size_t n) {
    std::vector<T> tmp(n);
    check(hipMemcpy(tmp.data(), src, n * sizeof(T), hipMemcpyDeviceToHost), "copy to host");
    std::copy(tmp.begin(), tmp.end(), dst);
}

template <typename T>
T *corePtr(spk_batch *b, int k) {
    return static_cast<T *>(b->core.p) + b->coreOff[k];
}

void validateProblem(const HostProblem &h) {
    if (h.nu > 64) throw std::runtime_error("numInputs > 64 is not supported by the fused engine");
    if (h.nx > 64 * spk::MAXR || h.nx + h.nu > 64 * spk::MAXV)
        throw std::runtime_error("numStates too large for the fused engine (nx <= 256, nx + nu <= 256)");
    if (h.nD > 64) throw std::runtime_error("numEvents > 15 is not supported by the fused engine");
    if (h.con3 > 64 * spk::MAXR || h.con5 > 64 * spk::MAXR)
        throw std::runtime_error("constraint rows > 256 are not supported by the fused engine");
}

// Smallest register chunk class covering every per-node dimension of the batch.
int chunkClass(const spk_batch *b) {
    int mx = 1;
    for (auto *p : b->probs) {
        const HostProblem &h = p->hp;
        mx = std::max({mx, h.nx, h.nu, h.con3, h.con5, h.nD, h.yD});
    }
    return mx <= 64 ? 1 : 4;
}

// Ints of per-node index and tree arrays staged in LDS by the kernels (0 when they do not
// fit in the budget, then the kernels read them from global memory).
int ldsPoolInts(const spk_batch *b) {
    int mx = 0;
    for (auto *p : b->probs) {
        const HostProblem &h = p->hp;
        const int need = 12 * h.N + 15 * h.nl + 7 * h.nlf + 2 * h.S;
        mx = std::max(mx, need);
    }
    const int budget = 48 * 1024 / 4;
    return mx <= budget ? mx : 0;
}

// Workgroups that cooperate on one instance: 1 for a batch of several instances and by default;
// for a single instance the count set by spk_batch_set_workgroups.  Bounded by what can be
// co-resident, which a cooperative launch requires.
int gridWorkgroups(spk_batch *b, const void *kernel, int bs, size_t shm) {
    if (b->count != 1 || b->wgs <= 1) return 1;
    const int dev = b->ctx->device;
    int coop = 0, cus = 0, perCU = 0;
    check(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev), "device attribute");
    if (!coop) {
        if (b->wgs > 1) throw std::runtime_error("the device does not support cooperative launches");
        return 1;
    }
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "device attribute");
    check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kernel, bs, shm), "occupancy");
    const int most = std::min(spk::GRID_MAXWG, cus * perCU);
    if (most < 1) throw std::runtime_error("the solve kernel does not fit on a compute unit");
    if (b->wgs > most)
        throw std::runtime_error("at most " + std::to_string(most) + " workgroups of " + std::to_string(bs) +
                                 " threads can cooperate on one instance on this device");
    return b->wgs;
}

template <typename K, typename... A>
void launchMaybeGrid(spk_batch *b, K kernel, int bs, size_t shm, A... args) {
    const void *fn = reinterpret_cast<const void *>(kernel);
    const int g = gridWorkgroups(b, fn, bs, shm);
    int grid = g > 1 ? 1 : 0;
    if (!grid) {
        hipLaunchKernelGGL(kernel, dim3(b->count), dim3(bs), shm, b->ctx->stream, args..., grid);
        return;
    }
    void *argv[] = {static_cast<void *>(&args)..., static_cast<void *>(&grid)};
    const hipError_t e = hipLaunchCooperativeKernel(fn, dim3(g), dim3(bs), argv, (unsigned) shm, b->ctx->stream);
    // a refused launch queues nothing: read the error back so a later launch check does not see it
    if (e != hipSuccess) {
        (void) hipGetLastError();
        throw spk::CoopRefused(std::string("cooperative launch: ") + hipGetErrorString(e));
    }
}

// insDev: the device copy of the descriptors (default: b->dInst)
template <typename T>
void launchKernel(spk_batch *b, int bs, bool debug, int op, const void *insDev = nullptr) {
    const int cls = chunkClass(b);
    const int pool = ldsPoolInts(b);
    const size_t shm = (size_t) pool * sizeof(int);
    auto *ins = static_cast<const spk::Inst<T> *>(insDev ? insDev : b->dInst.p);
    auto *outs = static_cast<spk::Out *>(b->dOut.p);
    const int count = b->count;
    const double tick = b->ctx->tickRateHz;
    if (!debug) {
        if (cls == 1) launchMaybeGrid(b, spk::k_solve<T, 1, 1>, bs, shm, ins, outs, count, tick, pool);
        else launchMaybeGrid(b, spk::k_solve<T, 4, 4>, bs, shm, ins, outs, count, tick, pool);
    } else {
        if (cls == 1) launchMaybeGrid(b, spk::k_debug<T, 1, 1>, bs, shm, ins, outs, count, op, pool);
        else launchMaybeGrid(b, spk::k_debug<T, 4, 4>, bs, shm, ins, outs, count, op, pool);
    }
}

// Threads per workgroup of a per-instance solve: the option, else the autotuned size, else 512
// (one workgroup per instance) or 256 (grid mode).  Measured on the 229-node problem (FP64, one
// instance): 7.3 ms per outer iteration with 512 threads against 11.4 with 256; in grid mode with
// 64 workgroups 1.6 ms with 256 against 1.9 with 512.
int solveBlock(const spk_options &o, const spk_batch *b = nullptr) {
    const bool grid = b && b->count == 1 && b->wgs > 1;
    const int bs = o.block_size > 0 ? o.block_size : (b && b->fusedBlock > 0 ? b->fusedBlock : (grid ? 256 : 512));
    return std::max(64, std::min(512, bs / 64 * 64));
}

template <typename T>
void launchSolve(spk_batch *b, const spk_options &o, int warm) {
    buildInsts<T>(b, o, warm);
    const int bs = solveBlock(o, b);
    b->dOut.alloc(sizeof(spk::Out) * b->count);
    check(hipEventRecord(b->ev0, b->ctx->stream), "event");
    launchKernel<T>(b, bs, false, 0);
    check(hipGetLastError(), "launch k_solve");
    check(hipEventRecord(b->ev1, b->ctx->stream), "event");
    b->pending = true;
}

// Per-instance batches: the persistent solve launch captured as a graph with one kernel node and
// launched; spk_batch_replay_graph launches it again after new initial states, iterates or
// tolerances were written into the batch's buffers (no rebuild of the launch, no reallocation).
template <typename T>
void launchInstGraph(spk_batch *b, const spk_options &o, int warm) {
    if (b->count == 1 && b->wgs > 1)
        throw std::runtime_error("grid mode (a cooperative launch) is not captured in a graph; use spk_batch_solve");
    hipStream_t st = b->ctx->stream;
    buildInsts<T>(b, o, warm);  // drops an earlier captured solve
    if (b->dOut.bytes != sizeof(spk::Out) * b->count) b->dOut.alloc(sizeof(spk::Out) * b->count);
    check(hipStreamSynchronize(st), "sync before capture");
    if (b->graphExec) check(hipGraphExecDestroy(b->graphExec), "graph exec destroy");
    if (b->graph) check(hipGraphDestroy(b->graph), "graph destroy");
    b->graphExec = nullptr;
    b->graph = nullptr;
    check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
    launchKernel<T>(b, solveBlock(o, b), false, 0);
    check(hipStreamEndCapture(st, &b->graph), "end capture");
    check(hipGraphInstantiate(&b->graphExec, b->graph, nullptr, nullptr, 0), "graph instantiate");
    b->instGraph = true;
    b->lastOpts = o;
    b->lastWarm = warm;
    check(hipEventRecord(b->ev0, st), "event");
    check(hipGraphLaunch(b->graphExec, st), "graph launch");
    check(hipEventRecord(b->ev1, st), "event");
    b->pending = true;
}

template <typename T>
void replayInstGraph(spk_batch *b) {
    hipStream_t st = b->ctx->stream;
    // the descriptors again, into the same buffer: given tolerances may have changed
    fillInsts<T>(b, b->lastOpts, b->lastWarm, static_cast<T *>(b->anders.p), b->andOff, static_cast<T *>(b->gpart.p));
    if (b->largestFirst && b->count > 1) orderLargestFirst<T>(b);
    check(hipMemcpyAsync(b->dInst.p, b->hInst.data(), b->hInst.size(), hipMemcpyHostToDevice, st), "instances");
    check(hipEventRecord(b->ev0, st), "event");
    check(hipGraphLaunch(b->graphExec, st), "graph launch");
    check(hipEventRecord(b->ev1, st), "event");
    b->pending = true;
}

// ---------------------------------------------------------------- allocation-free solves
// Host side of the scratch array (declared at the top of this file).
constexpr int SCRATCH_DEVICES = 64;
std::mutex scratchMu;
void *scratchAddr[SCRATCH_DEVICES] = {};

void initScratch(int device) {
    if (device < 0 || device >= SCRATCH_DEVICES) return;  // spk_batch_solve_noalloc reports it
    std::lock_guard<std::mutex> lock(scratchMu);
    if (scratchAddr[device]) return;
    void *ptr = nullptr;
    check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    scratchAddr[device] = ptr;
}

char *scratchOf(int device) {
    void *ptr = nullptr;
    if (device >= 0 && device < SCRATCH_DEVICES) {
        std::lock_guard<std::mutex> lock(scratchMu);
        ptr = scratchAddr[device];
    }
    if (!ptr) throw std::runtime_error("no allocation-free scratch on device " + std::to_string(device));
    return static_cast<char *>(ptr);
}

// Bytes of scratch a per-instance solve of b with Anderson buffer m needs, and the element
// offsets of its instances' rings.
template <typename T>
size_t scratchLayout(const spk_batch *b, int m, std::vector<size_t> &andOff, size_t &andAt, size_t &gpartAt) {
    andOff.resize(b->count);
    size_t el = 0;
    for (int k = 0; k < b->count; k++) {
        andOff[k] = el;
        el += andCount(b->probs[k]->hp, m);
    }
    andAt = alignUp(sizeof(spk::Inst<T>) * b->count, 256);
    gpartAt = alignUp(andAt + el * sizeof(T), 256);
    return gpartAt + (b->count == 1 ? (size_t) 2 * spk::GRID_MAXK * spk::GRID_MAXWG * sizeof(T) : 0);
}

template <typename T>
void launchSolveNoAlloc(spk_batch *b, const spk_options &o, int warm) {
    std::vector<size_t> andOff;
    size_t andAt = 0, gpartAt = 0;
    const size_t need = scratchLayout<T>(b, o.anderson_buffer, andOff, andAt, gpartAt);
    if (need > SPK_NOALLOC_SCRATCH_BYTES)
        throw std::runtime_error("the allocation-free solve of this batch needs " + std::to_string(need) +
                                 " bytes of scratch; the limit is " + std::to_string(SPK_NOALLOC_SCRATCH_BYTES) +
                                 " (SPK_NOALLOC_SCRATCH_BYTES)");
    if (b->dOut.bytes < sizeof(spk::Out) * b->count) throw std::runtime_error("the batch has no result records");
    char *base = scratchOf(b->ctx->device);
    fillInsts<T>(b, o, warm, reinterpret_cast<T *>(base + andAt), andOff,
                 b->count == 1 ? reinterpret_cast<T *>(base + gpartAt) : nullptr);
    check(hipMemcpyAsync(base, b->hInst.data(), b->hInst.size(), hipMemcpyHostToDevice, b->ctx->stream),
          "upload instances");
    check(hipEventRecord(b->ev0, b->ctx->stream), "event");
    launchKernel<T>(b, solveBlock(o, b), false, 0, base);
    check(hipGetLastError(), "launch k_solve");
    check(hipEventRecord(b->ev1, b->ctx->stream), "event");
    b->pending = true;
}


// ---------------------------------------------------------------- shared-problem (lane) batches
// Regions of the interleaved state, each (rows x Bpad): work[0..8] and zio, rio (sIter rows),
// the Anderson rings hS, hY (LANE_MAXM slots of sIter rows), qd (nx*N), dd (nu*nl), x0 (nx).
enum LaneRegion { LR_ZIO = 9, LR_RIO = 10, LR_HS = 11, LR_HY = 12, LR_QD = 13, LR_DD = 14, LR_X0 = 15, LR_END = 16 };

size_t laneLayout(spk_batch *b) {
    const HostProblem &h = b->probs[0]->hp;
    const size_t n = h.sIter, B = b->Bpad, m = spk_batch::LANE_MAXM;
    size_t off = 0;
    for (int r = 0; r <= LR_RIO; r++) {
        b->lRegion[r] = off;
        off += n * B;
    }
    b->lRegion[LR_HS] = off;
    off += m * n * B;
    b->lRegion[LR_HY] = off;
    off += m * n * B;
    b->lRegion[LR_QD] = off;
    off += (size_t) h.nx * h.N * B;
    b->lRegion[LR_DD] = off;
    off += (size_t) h.nu * h.nl * B;
    b->lRegion[LR_X0] = off;
    off += (size_t) h.nx * B;
    return off;
}

template <typename T>
T *lanePtr(const spk_batch *b, int region) {
    return static_cast<T *>(b->lbuf.p) + b->lRegion[region];
}

void validateLanes(const HostProblem &h, int) {
    // d = R~^-1 e and the kernel projection are computed in place from up to 32 register rows
    if (h.nu > 32) throw std::runtime_error("shared batches support numInputs <= 32");
    if (h.nD > 32) throw std::runtime_error("shared batches support numEvents <= 7");
}

template <typename T>
void buildLaneInst(spk_batch *b, const spk_options &o, int warm, int resume = 0) {
    spk::LaneInst<T> L;
    std::memset(static_cast<void *>(&L), 0, sizeof(L));
    spk_problem *p = b->probs[0];
    const HostProblem &h = p->hp;
    fillProblem<T>(L, p);
    L.B = b->Bpad;
    L.count = b->count;
    for (int r = 0; r < 9; r++) L.work[r] = lanePtr<T>(b, r);
    L.zio = lanePtr<T>(b, LR_ZIO);
    L.rio = lanePtr<T>(b, LR_RIO);
    L.hS = lanePtr<T>(b, LR_HS);
    L.hY = lanePtr<T>(b, LR_HY);
    L.qd = lanePtr<T>(b, LR_QD);
    L.dd = lanePtr<T>(b, LR_DD);
    L.x0s = lanePtr<T>(b, LR_X0);
    L.tolFixed = b->tolFixed.empty() ? nullptr : static_cast<const T *>(b->ltol.p);
    const double tolScale = tolScaleOf(h);
    L.tolAbs = (T) (o.tol_abs * tolScale);
    L.tolRel = (T) (o.tol_rel * tolScale);
    L.maxIters = o.max_iters;
    L.maxInner = o.max_inner;
    L.andBuff = o.anderson_buffer;
    L.checkPeriod = o.check_period > 0 ? o.check_period : 25;
    L.method = o.method;
    L.warm = warm;
    L.maxTicks = o.max_time_secs > 0 ? o.max_time_secs * b->ctx->tickRateHz : 0.;
    L.chunkIters = o.chunk_iters > 0 ? o.chunk_iters : 0;
    L.chunkTicks = (L.chunkIters > 0 && b->chunkSeconds > 0.) ? b->chunkSeconds * b->ctx->tickRateHz : 0.;
    L.resume = resume;
    L.save = static_cast<spk::LaneSave *>(b->lsave.p);
    L.gsave = static_cast<spk::LaneGroupSave *>(b->lgsave.p);
    b->hInst.assign(sizeof(L), 0);
    std::memcpy(b->hInst.data(), &L, sizeof(L));
    b->dInst.alloc(sizeof(L));
    check(hipMemcpyAsync(b->dInst.p, b->hInst.data(), sizeof(L), hipMemcpyHostToDevice, b->ctx->stream),
          "upload lane instance");
}

// Schedules of shared batches: the lane engine with scalar (0) or LDS-staged (8) factor loads,
// the tile engine built for workgroups of up to 512 (16) or 1024 threads (17).
bool tileSchedule(int mv) { return mv == 16 || mv == 17; }
int scheduleMaxThreads(int mv) { return mv == 16 ? 512 : 1024; }

// Workgroup size of the shared-batch kernels: the caller's, else the autotuner's, else the
// schedule's default (256 for the 512-thread tile build, else its launch bound).
int laneBlock(const spk_batch *b, const spk_options &o) {
    const int most = scheduleMaxThreads(b->laneMv);
    const int bs = o.block_size > 0 ? o.block_size : (b->laneBlockSz > 0 ? b->laneBlockSz : (b->laneMv == 16 ? 256 : most));
    return std::max(64, std::min(most, bs / 64 * 64));
}

template <typename T>
void launchLaneKernel(spk_batch *b, const spk_options &o, const void *inst = nullptr) {
    const dim3 grid((unsigned) (b->Bpad / b->laneW)), block(laneBlock(b, o));
    auto *P = static_cast<const spk::LaneInst<T> *>(inst ? inst : b->dInst.p);
    auto *outs = static_cast<spk::Out *>(b->dOut.p);
    if (b->laneMv == 17) hipLaunchKernelGGL(spk::k_tile_solve_x1024<T>, grid, block, 0, b->ctx->stream, P, outs);
    else if (b->laneMv == 16) hipLaunchKernelGGL(spk::k_tile_solve<T>, grid, block, 0, b->ctx->stream, P, outs);
    else if (b->laneMv == 8) hipLaunchKernelGGL(spk::k_lane_solve_x8<T>, grid, block, 0, b->ctx->stream, P, outs);
    else hipLaunchKernelGGL(spk::k_lane_solve<T>, grid, block, 0, b->ctx->stream, P, outs);
}

// Re-lays the batch's inputs and outputs (initial states, accepted iterates and their residuals)
// out for blocks of w instances and clears the workspace regions; a chunked solve cannot be
// resumed across the change.
template <typename T>
void relayoutRegions(spk_batch *b, int w) {
    const HostProblem &h = b->probs[0]->hp;
    hipStream_t st = b->ctx->stream;
    const int regions[3] = {LR_ZIO, LR_RIO, LR_X0};
    const long rows[3] = {(long) h.sIter, (long) h.sIter, (long) h.nx};
    DevBuf tmp;
    tmp.alloc((size_t) h.sIter * b->Bpad * sizeof(T));
    for (int r = 0; r < 3; r++) {
        T *reg = lanePtr<T>(b, regions[r]);
        const size_t bytes = (size_t) rows[r] * b->Bpad * sizeof(T);
        hipLaunchKernelGGL(spk::k_relayout<T>, dim3(1024), dim3(256), 0, st, reg, (long) b->laneW, tmp.as<T>(), (long) w,
                           rows[r], b->Bpad);
        check(hipGetLastError(), "launch k_relayout");
        check(hipMemcpyAsync(reg, tmp.p, bytes, hipMemcpyDeviceToDevice, st), "relayout copy");
    }
    for (int r = 0; r < LR_X0; r++)  // regions are contiguous in this order, X0 last
        if (r != LR_ZIO && r != LR_RIO)
            check(hipMemsetAsync(lanePtr<T>(b, r), 0, (b->lRegion[r + 1] - b->lRegion[r]) * sizeof(T), st),
                  "clear workspace");
    check(hipStreamSynchronize(st), "sync");
    b->laneW = w;
    b->resumable = false;
}

void setLaneWidth(spk_batch *b, int w) {
    if (b->laneW == w) return;
    check(hipStreamSynchronize(b->ctx->stream), "sync");
    if (b->dtype == SPK_F64) relayoutRegions<double>(b, w);
    else relayoutRegions<float>(b, w);
}

template <typename T>
void launchLaneSolve(spk_batch *b, const spk_options &o, int warm, int resume = 0) {
    buildLaneInst<T>(b, o, warm, resume);
    b->lastOpts = o;
    b->lastWarm = warm;
    b->resumable = o.chunk_iters > 0;
    // kept across launches: a resume of finished workgroups leaves their results in place
    if (b->dOut.bytes != sizeof(spk::Out) * b->count) b->dOut.alloc(sizeof(spk::Out) * b->count);
    check(hipEventRecord(b->ev0, b->ctx->stream), "event");
    launchLaneKernel<T>(b, o);
    check(hipGetLastError(), "launch k_lane_solve");
    check(hipEventRecord(b->ev1, b->ctx->stream), "event");
    b->pending = true;
}

// A chunked solve of a shared batch as one HIP graph: the first launch and launches - 1 resumes,
// with no host round trip between them (workgroups whose solve has ended return at once).
template <typename T>
void launchLaneGraph(spk_batch *b, const spk_options &o, int warm, int launches, bool resumeOnly = false) {
    hipStream_t st = b->ctx->stream;
    buildLaneInst<T>(b, o, warm, 1);
    b->dInstResume.alloc(sizeof(spk::LaneInst<T>));
    check(hipMemcpyAsync(b->dInstResume.p, b->dInst.p, sizeof(spk::LaneInst<T>), hipMemcpyDeviceToDevice, st),
          "resume instance");
    if (!resumeOnly) buildLaneInst<T>(b, o, warm, 0);
    if (b->dOut.bytes != sizeof(spk::Out) * b->count) b->dOut.alloc(sizeof(spk::Out) * b->count);
    check(hipStreamSynchronize(st), "sync before capture");
    if (b->graphExec) check(hipGraphExecDestroy(b->graphExec), "graph exec destroy");
    if (b->graph) check(hipGraphDestroy(b->graph), "graph destroy");
    b->graphExec = nullptr;
    b->graph = nullptr;
    check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
    if (!resumeOnly) launchLaneKernel<T>(b, o);
    for (int i = resumeOnly ? 0 : 1; i < launches; i++) launchLaneKernel<T>(b, o, b->dInstResume.p);
    check(hipStreamEndCapture(st, &b->graph), "end capture");
    check(hipGraphInstantiate(&b->graphExec, b->graph, nullptr, nullptr, 0), "graph instantiate");
    b->lastOpts = o;
    b->lastWarm = warm;
    b->resumable = true;
    check(hipEventRecord(b->ev0, st), "event");
    check(hipGraphLaunch(b->graphExec, st), "graph launch");
    check(hipEventRecord(b->ev1, st), "event");
    b->pending = true;
}

// Host rows (instance-major, n values each) <-> one interleaved region.
template <typename T>
void laneUpload(spk_batch *b, int region, const double *rows, size_t n) {
    std::vector<T> tmp(n * b->Bpad, T(0));
    for (int k = 0; k < b->count; k++)
        for (size_t e = 0; e < n; e++) tmp[spk::laneIndex((long) n, (long) e, k, b->laneW)] = (T) rows[(size_t) k * n + e];
    check(hipMemcpy(lanePtr<T>(b, region), tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
}

template <typename T>
void laneDownload(spk_batch *b, int region, double *rows, size_t n) {
    std::vector<T> tmp(n * b->Bpad);
    check(hipMemcpy(tmp.data(), lanePtr<T>(b, region), tmp.size() * sizeof(T), hipMemcpyDeviceToHost), "download");
    for (int k = 0; k < b->count; k++)
        for (size_t e = 0; e < n; e++) rows[(size_t) k * n + e] = (double) tmp[spk::laneIndex((long) n, (long) e, k, b->laneW)];
}

template <typename T>
void laneColumn(spk_batch *b, int region, int k, double *z, const double *zin, size_t n) {
    // instance k's n rows of the region, through a contiguous device row (a strided copy of
    // n single words costs about 0.4 s for the 52,360 rows of the 229-node problem)
    hipStream_t st = b->ctx->stream;
    DevBuf row;
    row.alloc(n * sizeof(T));
    std::vector<T> tmp(n);
    const int grid = (int) std::min<size_t>(1024, (n + 255) / 256);
    if (zin) {
        for (size_t e = 0; e < n; e++) tmp[e] = (T) zin[e];
        check(hipMemcpyAsync(row.p, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice, st), "set column");
    }
    hipLaunchKernelGGL(spk::k_lane_column<T>, dim3(grid), dim3(256), 0, st, lanePtr<T>(b, region), (long) n, (long) k,
                       (long) b->laneW, row.as<T>(), zin ? 0 : 1);
    check(hipGetLastError(), "launch k_lane_column");
    if (!zin) check(hipMemcpyAsync(tmp.data(), row.p, n * sizeof(T), hipMemcpyDeviceToHost, st), "get column");
    check(hipStreamSynchronize(st), "column");
    if (!zin)
        for (size_t e = 0; e < n; e++) z[e] = (double) tmp[e];
}

// Uniform row stride of a per-instance batch whose instances all share one problem's sizes.
size_t uniformCoreStride(const spk_batch *b) {
    const size_t stride = b->count > 1 ? b->coreOff[1] - b->coreOff[0] : coreCount(b->probs[0]->hp);
    for (int k = 0; k < b->count; k++)
        if (b->coreOff[k] != (size_t) k * stride || b->probs[k]->hp.sIter != b->probs[0]->hp.sIter)
            throw std::runtime_error("copy between layouts needs instances of one problem size");
    return stride;
}

template <typename S, typename D>
void launchInterleave(const spk_batch *src, spk_batch *dst, size_t n) {
    const size_t stride = uniformCoreStride(src);
    hipLaunchKernelGGL((spk::k_interleave<S, D>), dim3(1024), dim3(256), 0, dst->ctx->stream,
                       static_cast<const S *>(src->core.p), (long) stride, lanePtr<D>(dst, LR_ZIO), (long) n,
                       dst->count, (long) dst->laneW);
}

template <typename S, typename D>
void launchDeinterleave(const spk_batch *src, spk_batch *dst, size_t n) {
    const size_t stride = uniformCoreStride(dst);
    hipLaunchKernelGGL((spk::k_deinterleave<S, D>), dim3(1024), dim3(256), 0, dst->ctx->stream,
                       lanePtr<S>(src, LR_ZIO), (long) src->laneW, static_cast<D *>(dst->core.p), (long) stride, (long) n,
                       dst->count);
}

template <typename T>
void fillDevice(spk_batch *b, T *p, size_t n) {
    hipLaunchKernelGGL(spk::k_fill<T>, dim3(1024), dim3(256), 0, b->ctx->stream, p, T(0), (long) n);
}

// What an autotuner changes on a batch while it times: the given tolerances (cleared so every
// timed run does the same fixed iterations) and the launch choices it varies.  Put back as they
// were when a timed run fails; on success the caller restores the tolerances and keeps the choice.
struct TuneGuard {
    spk_batch *b;
    std::vector<double> tol;
    int wgs, mv, block, width;
    bool kept = false;
    explicit TuneGuard(spk_batch *bb)
        : b(bb), tol(bb->tolFixed), wgs(bb->wgs), mv(bb->laneMv), block(bb->laneBlockSz), width(bb->laneW) {
        b->tolFixed.clear();
    }
    void keep() {
        b->tolFixed = tol;
        kept = true;
    }
    ~TuneGuard() {
        if (kept) return;
        b->tolFixed = tol;
        b->wgs = wgs;
        b->laneMv = mv;
        b->laneBlockSz = block;
        if (b->lanes) {
            try {
                setLaneWidth(b, width);
            } catch (const std::exception &) {
                // the layout stays as the failed timed run left it; the error being reported names the cause
            }
        }
    }
};

// Lane indices of a refill or gather, checked on the host and staged on the device.
void stageLanes(spk_batch *b, int count, const int *lanes) {
    if (!b->lanes) throw std::runtime_error("refills and lane gathers apply to shared batches");
    if (count < 1 || !lanes) throw std::runtime_error("no lanes given");
    std::vector<char> seen(b->count, 0);
    for (int i = 0; i < count; i++) {
        if (lanes[i] < 0 || lanes[i] >= b->count) throw std::runtime_error("lane index out of range");
        if (seen[lanes[i]]) throw std::runtime_error("a lane is listed twice");
        seen[lanes[i]] = 1;
    }
    if (b->rLanes.bytes < count * sizeof(int)) b->rLanes.alloc(count * sizeof(int));
    check(hipMemcpyAsync(b->rLanes.p, lanes, count * sizeof(int), hipMemcpyHostToDevice, b->ctx->stream), "lanes");
}

// New solves in the staged lanes of a resumable shared batch, from the zero iterate (z == null) or
// from device rows z (leading dimension ld).
template <typename T, typename Z = float>
void refillLanes(spk_batch *b, int count, const double *x0, const Z *z, long ld) {
    if (!b->resumable) throw std::runtime_error("a refill continues a chunked solve (chunk_iters > 0) of the batch");
    const HostProblem &h = b->probs[0]->hp;
    std::vector<T> xs((size_t) count * h.nx);
    for (int i = 0; i < count; i++)
        for (int k = 0; k < h.nx; k++)
            xs[(size_t) i * h.nx + k] = (T) (x0[(size_t) i * h.nx + k] * (h.precond ? h.scaling[k] : 1.));  // src/cache.cuh:599-601
    hipStream_t st = b->ctx->stream;
    if (b->rX0.bytes < xs.size() * sizeof(T)) b->rX0.alloc(xs.size() * sizeof(T));
    check(hipMemcpyAsync(b->rX0.p, xs.data(), xs.size() * sizeof(T), hipMemcpyHostToDevice, st), "x0");
    buildLaneInst<T>(b, b->lastOpts, b->lastWarm, 1);
    auto *P = static_cast<const spk::LaneInst<T> *>(b->dInst.p);
    const int *lanes = b->rLanes.as<int>();
    hipLaunchKernelGGL((spk::k_lane_refill<T, Z>), dim3(1024), dim3(256), 0, st, P, lanes, count, z, ld,
                       b->rX0.as<T>(), (long) b->laneW);
    check(hipGetLastError(), "launch k_lane_refill");
    hipLaunchKernelGGL(spk::k_lane_refill_groups<T>, dim3(16), dim3(256), 0, st, P, lanes, count, (long) b->laneW);
    check(hipGetLastError(), "launch k_lane_refill_groups");
    check(hipStreamSynchronize(st), "sync");
}

// The accepted iterates of the staged lanes as device rows of the batch's dtype (leading dimension ld).
template <typename T>
void gatherLanes(spk_batch *b, int count, T *dst, long ld) {
    const long n = b->probs[0]->hp.sIter;
    hipLaunchKernelGGL((spk::k_gather_lanes<T, T>), dim3(1024), dim3(256), 0, b->ctx->stream, lanePtr<T>(b, LR_ZIO),
                       (long) b->laneW, dst, ld, n, b->rLanes.as<int>(), count);
    check(hipGetLastError(), "launch k_gather_lanes");
}

}  // namespace

extern "C" {

void spk_default_options(spk_options *o) {
    o->tol_abs = 1e-3;
    o->tol_rel = 0.;
    o->max_time_secs = 0.;
    o->max_iters = 0;
    o->max_inner = 8;
    o->anderson_buffer = 3;
    o->check_period = 25;
    o->method = SPK_SPOCK;
    o->block_size = 0;  // the engine's own default (see spk_options in the header)
    o->ls_method = SPK_LS_HOUSEHOLDER;
    o->dyn_inverse = 1;
    o->chunk_iters = 0;
}

int spk_ctx_create(int device, spk_ctx **out) {
    try {
        auto *c = new spk_ctx();
        c->device = device;
        check(hipSetDevice(device), "hipSetDevice");
        check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
            c->tickRateHz = khz * 1000.0;
        initScratch(device);
        *out = c;
        return 0;
    } catch (const std::exception &e) {
        *out = nullptr;
        return -1;
    }
}

int spk_ctx_destroy(spk_ctx *ctx) {
    spk::releaseCtx(ctx);
    return 0;
}

const char *spk_last_error(spk_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int spk_ctx_stream(spk_ctx *ctx, void **stream) {
    *stream = ctx->stream;
    return 0;
}

int spk_ctx_synchronize(spk_ctx *ctx) {
    try {
        check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        return 0;
    } catch (const std::exception &e) {
        return fail(ctx, e);
    }
}

int spk_problem_load(spk_ctx *ctx, const char *folder, spk_problem **out) {
    *out = nullptr;
    try {
        if (!ctx) throw std::runtime_error("null context");
        auto p = std::make_unique<spk_problem>(ctx);
        p->hp = spk::loadProblem(folder);
        validateProblem(p->hp);
        check(hipSetDevice(ctx->device), "hipSetDevice");
        uploadInts(p->anc, p->hp.anc);
        uploadInts(p->chFrom, p->hp.chFrom);
        uploadInts(p->numCh, p->hp.numCh);
        uploadInts(p->stageFrom, p->hp.stageFrom);
        uploadInts(p->stageTo, p->hp.stageTo);
        *out = p.release();
        return 0;
    } catch (const std::exception &e) {
        return fail(ctx, e);
    }
}

int spk_problem_destroy(spk_problem *p) {
    spk::release(p, [](spk_problem *q) { delete q; });
    return 0;
}

int spk_problem_sizes(const spk_problem *p, long *out) {
    const HostProblem &h = p->hp;
    long v[16] = {h.N, h.nl, h.S, h.E, h.nx, h.nu, h.sPrim, h.sDual, h.sIter, h.precond ? 1 : 0, h.con3, h.con5,
                  h.ABKTr.count(), h.sQ.count(), 0, 0};
    std::memcpy(out, v, sizeof(v));
    return 0;
}

int spk_problem_scaling(const spk_problem *p, double *out) {
    const HostProblem &h = p->hp;
    for (int k = 0; k < h.nx + h.nu; k++) out[k] = h.precond ? h.scaling[k] : 1.;
    return 0;
}

double spk_problem_step_size(const spk_problem *p) { return p->hp.stepSize; }

int spk_batch_create(spk_ctx *ctx, int count, spk_problem *const *problems, int dtype, spk_batch **out) {
    *out = nullptr;
    try {
        if (count <= 0) throw std::runtime_error("batch count must be positive");
        if (dtype != SPK_F64 && dtype != SPK_F32) throw std::runtime_error("dtype must be SPK_F64 or SPK_F32");
        if (!ctx) throw std::runtime_error("null context");
        auto b = std::make_unique<spk_batch>(ctx);
        b->dtype = dtype;
        b->count = count;
        b->probs.assign(problems, problems + count);
        for (auto *p : b->probs)
            if (!p) throw std::runtime_error("null problem in batch");
        b->holdProblems();
        check(hipSetDevice(ctx->device), "hipSetDevice");
        for (auto *p : b->probs) {
            if (dtype == SPK_F64) ensureUploaded<double>(p);
            else ensureUploaded<float>(p);
        }
        b->coreOff.resize(count);
        size_t off = 0;
        for (int k = 0; k < count; k++) {
            b->coreOff[k] = off;
            off += coreCount(b->probs[k]->hp);
        }
        b->coreElems = off;
        b->core.alloc(off * b->elSize());
        check(hipMemset(b->core.p, 0, off * b->elSize()), "zero state");
        b->x0Off.resize(count);
        off = 0;
        for (int k = 0; k < count; k++) {
            b->x0Off[k] = off;
            off += b->probs[k]->hp.nx;
        }
        b->x0Elems = off;
        b->x0buf.alloc(off * b->elSize());
        check(hipMemset(b->x0buf.p, 0, off * b->elSize()), "zero x0");
        check(hipEventCreate(&b->ev0), "event");
        check(hipEventCreate(&b->ev1), "event");
        ensureAnderson(b.get(), 3);
        b->dOut.alloc(sizeof(spk::Out) * count);  // result records (spk_batch_solve_noalloc writes here)
        check(hipStreamSynchronize(ctx->stream), "sync");
        *out = b.release();
        return 0;
    } catch (const std::exception &e) {
        return fail(ctx, e);
    }
}

int spk_batch_create_shared(spk_ctx *ctx, spk_problem *problem, int count, int dtype, spk_batch **out) {
    try {
        if (count <= 0) throw std::runtime_error("batch count must be positive");
        if (!problem) throw std::runtime_error("null problem");
        if (dtype != SPK_F64 && dtype != SPK_F32) throw std::runtime_error("dtype must be SPK_F64 or SPK_F32");
        validateLanes(problem->hp, dtype);
        if (!ctx) throw std::runtime_error("null context");
        auto b = std::make_unique<spk_batch>(ctx);
        b->dtype = dtype;
        b->count = count;
        b->lanes = true;
        b->Bpad = ((long) count + 63) / 64 * 64;
        b->probs.assign(count, problem);
        b->holdProblems();
        check(hipSetDevice(ctx->device), "hipSetDevice");
        if (dtype == SPK_F64) ensureUploaded<double>(problem);
        else ensureUploaded<float>(problem);
        const size_t total = laneLayout(b.get());
        b->lbuf.alloc(total * b->elSize());
        b->lsave.alloc(sizeof(spk::LaneSave) * b->Bpad);
        b->lgsave.alloc(sizeof(spk::LaneGroupSave) * (b->Bpad / 16));  // one per workgroup of either engine
        check(hipMemset(b->lbuf.p, 0, total * b->elSize()), "zero state");
        check(hipEventCreate(&b->ev0), "event");
        check(hipEventCreate(&b->ev1), "event");
        *out = b.release();
        return 0;
    } catch (const std::exception &e) {
        return fail(ctx, e);
    }
}

int spk_batch_is_shared(const spk_batch *b) { return b && b->lanes ? 1 : 0; }

// Waits for the batch's work, frees its graph and events and drops its references; the device
// buffers (members) are freed after this body.
spk_batch::~spk_batch() {
    if (ctx && ctx->stream) (void) hipStreamSynchronize(ctx->stream);
    if (graphExec) (void) hipGraphExecDestroy(graphExec);
    if (graph) (void) hipGraphDestroy(graph);
    if (ev0) (void) hipEventDestroy(ev0);
    if (ev1) (void) hipEventDestroy(ev1);
    for (auto *p : heldProbs) spk::release(p, [](spk_problem *q) { delete q; });
    spk::releaseCtx(ctx);
}

int spk_batch_destroy(spk_batch *b) {
    delete b;
    return 0;
}

int spk_batch_count(const spk_batch *b) { return b->count; }

int spk_batch_set_workgroups(spk_batch *b, int workgroups) {
    try {
        if (workgroups < 0 || workgroups > spk::GRID_MAXWG)
            throw std::runtime_error("workgroups must be in 0.." + std::to_string(spk::GRID_MAXWG));
        if (b->lanes) throw std::runtime_error("shared batches run one workgroup per 64 instances");
        if (workgroups > 1 && b->count != 1)
            throw std::runtime_error("several workgroups per instance need a batch of one instance");
        b->wgs = workgroups;
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

static_assert(spk::DBG_WORDS == SPK_DEBUG_WORDS, "debug record layout");

int spk_batch_set_debug(spk_batch *b, long max_records) {
    try {
        if (max_records < 0) throw std::invalid_argument("spk_batch_set_debug: max_records must be >= 0");
        if (b->lanes)
            throw std::invalid_argument("spk_batch_set_debug: debug traces are recorded by per-instance batches "
                                        "(spk_batch_create), not shared ones");
        check(hipStreamSynchronize(b->ctx->stream), "sync");
        dropInstGraph(b);  // the descriptors point into the debug buffers
        b->dbgCap = max_records;
        if (max_records == 0) {
            b->dbgRec.release();
            b->dbgE.release();
            b->dbgEOff.clear();
            return 0;
        }
        b->dbgRec.alloc((size_t) b->count * max_records * spk::DBG_WORDS * sizeof(double));
        check(hipMemset(b->dbgRec.p, 0, b->dbgRec.bytes), "zero debug records");
        b->dbgEOff.resize(b->count);
        size_t off = 0;
        for (int k = 0; k < b->count; k++) {
            b->dbgEOff[k] = off;
            off += alignUp((size_t) b->probs[k]->hp.sPrim, 64);
        }
        b->dbgE.alloc(off * b->elSize());
        check(hipMemset(b->dbgE.p, 0, b->dbgE.bytes), "zero debug vectors");
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_debug_records(spk_batch *b, int inst, long cap, double *records, long *count, long *nonfinite) {
    try {
        if (b->lanes || b->dbgCap == 0)
            throw std::invalid_argument("spk_batch_debug_records: debug traces are off (spk_batch_set_debug)");
        if (inst < 0 || inst >= b->count) throw std::invalid_argument("spk_batch_debug_records: no such instance");
        if (!b->pending) throw std::invalid_argument("spk_batch_debug_records: no solve has been run on this batch");
        check(hipStreamSynchronize(b->ctx->stream), "sync");
        spk::Out o;
        check(hipMemcpy(&o, static_cast<spk::Out *>(b->dOut.p) + inst, sizeof(o), hipMemcpyDeviceToHost),
              "debug result");
        const long n = std::min(std::min(o.nDbg, b->dbgCap), std::max(cap, 0L));
        if (n > 0)
            check(hipMemcpy(records, static_cast<double *>(b->dbgRec.p) + (size_t) inst * b->dbgCap * spk::DBG_WORDS,
                            (size_t) n * spk::DBG_WORDS * sizeof(double), hipMemcpyDeviceToHost),
                  "debug records");
        if (count) *count = o.nDbg;
        if (nonfinite) *nonfinite = o.nonfinite;
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_reset(spk_batch *b) {
    try {
        check(hipStreamSynchronize(b->ctx->stream), "sync");
        if (b->lanes) {  // everything before the initial states
            check(hipMemset(b->lbuf.p, 0, b->lRegion[LR_X0] * b->elSize()), "reset");
            return 0;
        }
        check(hipMemset(b->core.p, 0, b->coreElems * b->elSize()), "reset");
        if (b->anders.p) check(hipMemset(b->anders.p, 0, b->andElems * b->elSize()), "reset");
        if (b->dbgE.p) check(hipMemset(b->dbgE.p, 0, b->dbgE.bytes), "reset");  // err0's vector (src/cache.cuh:469)
        // the initial states are inputs of the next solve (spk_batch_set_x0) and stay; the
        // reference zeroes its copy because runSpock(x0) uploads x0 again (src/cache.cuh:448,1011)
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_set_x0(spk_batch *b, const double *x0) {
    try {
        check(hipStreamSynchronize(b->ctx->stream), "sync");
        if (b->lanes) {
            const HostProblem &h = b->probs[0]->hp;
            std::vector<double> s(x0, x0 + (size_t) b->count * h.nx);
            if (h.precond)
                for (int k = 0; k < b->count; k++)
                    for (int i = 0; i < h.nx; i++) s[(size_t) k * h.nx + i] *= h.scaling[i];  // src/cache.cuh:599-601
            if (b->dtype == SPK_F64) laneUpload<double>(b, LR_X0, s.data(), h.nx);
            else laneUpload<float>(b, LR_X0, s.data(), h.nx);
            return 0;
        }
        std::vector<double> s(x0, x0 + b->x0Elems);
        for (int k = 0; k < b->count; k++) {
            const HostProblem &h = b->probs[k]->hp;
            if (h.precond)
                for (int i = 0; i < h.nx; i++) s[b->x0Off[k] + i] *= h.scaling[i];  // src/cache.cuh:599-601
        }
        if (b->dtype == SPK_F64) toDevice<double>(b, static_cast<double *>(b->x0buf.p), s.data(), s.size());
        else toDevice<float>(b, static_cast<float *>(b->x0buf.p), s.data(), s.size());
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_set_tolerances(spk_batch *b, const double *tol) {
    try {
        if (!tol) {
            b->tolFixed.clear();
        } else {
            b->tolFixed.assign(tol, tol + b->count);
            for (double t : b->tolFixed)
                if (!(t >= 0.)) throw std::runtime_error("tolerances must be >= 0");
            if (b->lanes) {
                check(hipStreamSynchronize(b->ctx->stream), "sync");
                // the same buffer across calls: launches and captured graphs keep its address
                if (b->ltol.bytes != b->Bpad * b->elSize()) b->ltol.alloc(b->Bpad * b->elSize());
                std::vector<double> t(b->tolFixed);
                t.resize(b->Bpad, 0.);
                if (b->dtype == SPK_F64) toDevice<double>(b, static_cast<double *>(b->ltol.p), t.data(), t.size());
                else toDevice<float>(b, static_cast<float *>(b->ltol.p), t.data(), t.size());
            }
        }
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

// Iterates cross the host boundary in the reference's order of the dual blocks; for
// polyhedronWithIdentity the engines keep the constraint blocks per node (spk::reorderIdentityBlocks).
// `rows` consecutive iterates of problem h in the engines' order: z itself, or a reordered copy in tmp.
const double *engineOrder(const HostProblem &h, const double *z, size_t rows, std::vector<double> &tmp) {
    if (!h.id3 && !h.id5) return z;
    tmp.assign(z, z + rows * h.sIter);
    for (size_t r = 0; r < rows; r++) spk::reorderIdentityBlocks(h, tmp.data() + r * h.sIter, true);
    return tmp.data();
}

void referenceOrder(const HostProblem &h, double *z, size_t rows) {
    if (!h.id3 && !h.id5) return;
    for (size_t r = 0; r < rows; r++) spk::reorderIdentityBlocks(h, z + r * h.sIter, false);
}

int spk_batch_set_iterate(spk_batch *b, int k, const double *z) {
    try {
        if (k < 0 || k >= b->count) throw std::runtime_error("instance index out of range");
        check(hipStreamSynchronize(b->ctx->stream), "sync");
        const size_t n = b->probs[k]->hp.sIter;
        std::vector<double> tmp;
        z = engineOrder(b->probs[k]->hp, z, 1, tmp);
        if (b->lanes) {
            if (b->dtype == SPK_F64) laneColumn<double>(b, LR_ZIO, k, nullptr, z, n);
            else laneColumn<float>(b, LR_ZIO, k, nullptr, z, n);
            return 0;
        }
        if (b->dtype == SPK_F64) toDevice<double>(b, corePtr<double>(b, k), z, n);
        else toDevice<float>(b, corePtr<float>(b, k), z, n);
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_get_iterate(spk_batch *b, int k, double *z) {
    try {
        if (k < 0 || k >= b->count) throw std::runtime_error("instance index out of range");
        check(hipStreamSynchronize(b->ctx->stream), "sync");
        const size_t n = b->probs[k]->hp.sIter;
        if (b->lanes) {
            if (b->dtype == SPK_F64) laneColumn<double>(b, LR_ZIO, k, z, nullptr, n);
            else laneColumn<float>(b, LR_ZIO, k, z, nullptr, n);
        } else if (b->dtype == SPK_F64) {
            toHost<double>(b, z, corePtr<double>(b, k), n);
        } else {
            toHost<float>(b, z, corePtr<float>(b, k), n);
        }
        referenceOrder(b->probs[k]->hp, z, 1);
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_set_iterates(spk_batch *b, const double *z) {
    if (b->lanes) {
        try {
            check(hipStreamSynchronize(b->ctx->stream), "sync");
            std::vector<double> tmp;
            z = engineOrder(b->probs[0]->hp, z, b->count, tmp);
            if (b->dtype == SPK_F64) laneUpload<double>(b, LR_ZIO, z, b->probs[0]->hp.sIter);
            else laneUpload<float>(b, LR_ZIO, z, b->probs[0]->hp.sIter);
            return 0;
        } catch (const std::exception &e) {
            return fail(b->ctx, e);
        }
    }
    size_t pos = 0;
    for (int k = 0; k < b->count; k++) {
        if (spk_batch_set_iterate(b, k, z + pos)) return -1;
        pos += b->probs[k]->hp.sIter;
    }
    return 0;
}

int spk_batch_get_iterates(spk_batch *b, double *z) {
    if (b->lanes) {
        try {
            check(hipStreamSynchronize(b->ctx->stream), "sync");
            if (b->dtype == SPK_F64) laneDownload<double>(b, LR_ZIO, z, b->probs[0]->hp.sIter);
            else laneDownload<float>(b, LR_ZIO, z, b->probs[0]->hp.sIter);
            referenceOrder(b->probs[0]->hp, z, b->count);
            return 0;
        } catch (const std::exception &e) {
            return fail(b->ctx, e);
        }
    }
    size_t pos = 0;
    for (int k = 0; k < b->count; k++) {
        if (spk_batch_get_iterate(b, k, z + pos)) return -1;
        pos += b->probs[k]->hp.sIter;
    }
    return 0;
}

// Device rows (instance-major, FP32, leading dimension ld) <-> the batch's accepted iterates.
int spk_batch_set_iterates_device(spk_batch *b, const float *src, long ld) {
    try {
        check(hipSetDevice(b->ctx->device), "hipSetDevice");
        const size_t n = b->probs[0]->hp.sIter;
        for (int k = 0; k < b->count; k++)
            if ((size_t) b->probs[k]->hp.sIter != n) throw std::runtime_error("device rows need one problem size");
        if (ld < (long) n) throw std::runtime_error("leading dimension smaller than the iterate");
        hipStream_t st = b->ctx->stream;
        if (b->lanes) {
            if (b->dtype == SPK_F64) {
                hipLaunchKernelGGL((spk::k_interleave<float, double>), dim3(1024), dim3(256), 0, st, src, ld,
                                   lanePtr<double>(b, LR_ZIO), (long) n, b->count, (long) b->laneW);
                fillDevice<double>(b, lanePtr<double>(b, LR_RIO), n * b->Bpad);
            } else {
                hipLaunchKernelGGL((spk::k_interleave<float, float>), dim3(1024), dim3(256), 0, st, src, ld,
                                   lanePtr<float>(b, LR_ZIO), (long) n, b->count, (long) b->laneW);
                fillDevice<float>(b, lanePtr<float>(b, LR_RIO), n * b->Bpad);
            }
        } else {
            for (int k = 0; k < b->count; k++) {
                if (b->dtype == SPK_F64) {
                    hipLaunchKernelGGL(spk::k_f32_to_f64, dim3(64), dim3(256), 0, st, src + (size_t) k * ld,
                                       corePtr<double>(b, k), (long) n);
                    fillDevice<double>(b, corePtr<double>(b, k) + 4 * n, n);
                } else {
                    check(hipMemcpyAsync(corePtr<float>(b, k), src + (size_t) k * ld, n * sizeof(float),
                                         hipMemcpyDeviceToDevice, st), "copy");
                    fillDevice<float>(b, corePtr<float>(b, k) + 4 * n, n);
                }
            }
        }
        check(hipGetLastError(), "set iterates (device)");
        check(hipStreamSynchronize(st), "sync");
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_get_iterates_device(spk_batch *b, float *dst, long ld) {
    try {
        check(hipSetDevice(b->ctx->device), "hipSetDevice");
        const size_t n = b->probs[0]->hp.sIter;
        for (int k = 0; k < b->count; k++)
            if ((size_t) b->probs[k]->hp.sIter != n) throw std::runtime_error("device rows need one problem size");
        if (ld < (long) n) throw std::runtime_error("leading dimension smaller than the iterate");
        hipStream_t st = b->ctx->stream;
        check(hipStreamSynchronize(st), "sync");
        if (b->lanes) {
            if (b->dtype == SPK_F64)
                hipLaunchKernelGGL((spk::k_deinterleave<double, float>), dim3(1024), dim3(256), 0, st,
                                   lanePtr<double>(b, LR_ZIO), (long) b->laneW, dst, ld, (long) n, b->count);
            else
                hipLaunchKernelGGL((spk::k_deinterleave<float, float>), dim3(1024), dim3(256), 0, st,
                                   lanePtr<float>(b, LR_ZIO), (long) b->laneW, dst, ld, (long) n, b->count);
        } else {
            for (int k = 0; k < b->count; k++) {
                if (b->dtype == SPK_F64)
                    hipLaunchKernelGGL(spk::k_f64_to_f32, dim3(64), dim3(256), 0, st, corePtr<double>(b, k),
                                       dst + (size_t) k * ld, (long) n);
                else
                    check(hipMemcpyAsync(dst + (size_t) k * ld, corePtr<float>(b, k), n * sizeof(float),
                                         hipMemcpyDeviceToDevice, st), "copy");
            }
        }
        check(hipGetLastError(), "get iterates (device)");
        check(hipStreamSynchronize(st), "sync");
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

void *spk_batch_iterate_ptr(spk_batch *b, int k) {
    if (k < 0 || k >= b->count) return nullptr;
    if (b->lanes)
        return static_cast<char *>(b->lbuf.p) +
               (b->lRegion[LR_ZIO] + spk::laneIndex(b->probs[0]->hp.sIter, 0, k, b->laneW)) * b->elSize();
    return static_cast<char *>(b->core.p) + b->coreOff[k] * b->elSize();
}

int spk_batch_copy_iterates(spk_batch *dst, const spk_batch *src) {
    try {
        if (dst->count != src->count) throw std::runtime_error("batch sizes differ");
        check(hipStreamSynchronize(src->ctx->stream), "sync");
        if (src->lanes || dst->lanes) {
            const size_t n = dst->probs[0]->hp.sIter;
            for (int k = 0; k < dst->count; k++)
                if ((size_t) src->probs[k]->hp.sIter != n || (size_t) dst->probs[k]->hp.sIter != n)
                    throw std::runtime_error("instance sizes differ");
            const bool s64 = src->dtype == SPK_F64, d64 = dst->dtype == SPK_F64;
            if (src->lanes && dst->lanes) {
                if (src->Bpad != dst->Bpad) throw std::runtime_error("batch sizes differ");
                const long len = (long) (n * dst->Bpad);
                if (s64 == d64)
                    check(hipMemcpyAsync(static_cast<char *>(dst->lbuf.p) + dst->lRegion[LR_ZIO] * dst->elSize(),
                                         static_cast<const char *>(src->lbuf.p) + src->lRegion[LR_ZIO] * src->elSize(),
                                         len * dst->elSize(), hipMemcpyDeviceToDevice, dst->ctx->stream), "copy");
                else if (!s64)
                    hipLaunchKernelGGL(spk::k_f32_to_f64, dim3(1024), dim3(256), 0, dst->ctx->stream,
                                       lanePtr<float>(src, LR_ZIO), lanePtr<double>(dst, LR_ZIO), len);
                else
                    hipLaunchKernelGGL(spk::k_f64_to_f32, dim3(1024), dim3(256), 0, dst->ctx->stream,
                                       lanePtr<double>(src, LR_ZIO), lanePtr<float>(dst, LR_ZIO), len);
            } else if (dst->lanes) {
                if (s64 && d64) launchInterleave<double, double>(src, dst, n);
                else if (s64) launchInterleave<double, float>(src, dst, n);
                else if (d64) launchInterleave<float, double>(src, dst, n);
                else launchInterleave<float, float>(src, dst, n);
            } else {
                if (s64 && d64) launchDeinterleave<double, double>(src, dst, n);
                else if (s64) launchDeinterleave<double, float>(src, dst, n);
                else if (d64) launchDeinterleave<float, double>(src, dst, n);
                else launchDeinterleave<float, float>(src, dst, n);
            }
            // fresh residual history for the continued solve
            if (dst->lanes) {
                if (d64) fillDevice<double>(dst, lanePtr<double>(dst, LR_RIO), n * dst->Bpad);
                else fillDevice<float>(dst, lanePtr<float>(dst, LR_RIO), n * dst->Bpad);
            } else {
                for (int k = 0; k < dst->count; k++) {
                    if (d64) fillDevice<double>(dst, corePtr<double>(dst, k) + 4 * n, n);
                    else fillDevice<float>(dst, corePtr<float>(dst, k) + 4 * n, n);
                }
            }
            check(hipGetLastError(), "copy kernels");
            check(hipStreamSynchronize(dst->ctx->stream), "sync");
            return 0;
        }
        for (int k = 0; k < dst->count; k++) {
            const size_t n = dst->probs[k]->hp.sIter;
            if (src->probs[k]->hp.sIter != (long) n) throw std::runtime_error("instance sizes differ");
            void *d = static_cast<char *>(dst->core.p) + dst->coreOff[k] * dst->elSize();
            const void *s = static_cast<const char *>(src->core.p) + src->coreOff[k] * src->elSize();
            if (src->dtype == dst->dtype) {
                check(hipMemcpyAsync(d, s, n * dst->elSize(), hipMemcpyDeviceToDevice, dst->ctx->stream), "copy");
            } else if (src->dtype == SPK_F32) {
                hipLaunchKernelGGL(spk::k_f32_to_f64, dim3(256), dim3(256), 0, dst->ctx->stream,
                                   static_cast<const float *>(s), static_cast<double *>(d), (long) n);
            } else {
                hipLaunchKernelGGL(spk::k_f64_to_f32, dim3(256), dim3(256), 0, dst->ctx->stream,
                                   static_cast<const double *>(s), static_cast<float *>(d), (long) n);
            }
            // fresh residual history for the continued solve
            void *res = static_cast<char *>(d) + 4 * n * dst->elSize();
            if (dst->dtype == SPK_F64)
                hipLaunchKernelGGL(spk::k_fill<double>, dim3(256), dim3(256), 0, dst->ctx->stream,
                                   static_cast<double *>(res), 0.0, (long) n);
            else
                hipLaunchKernelGGL(spk::k_fill<float>, dim3(256), dim3(256), 0, dst->ctx->stream,
                                   static_cast<float *>(res), 0.0f, (long) n);
            check(hipGetLastError(), "copy kernels");
        }
        check(hipStreamSynchronize(dst->ctx->stream), "sync");
        return 0;
    } catch (const std::exception &e) {
        return fail(dst->ctx, e);
    }
}

int spk_batch_solve(spk_batch *b, const spk_options *opts, int warm) {
    try {
        spk_options o;
        if (opts) o = *opts;
        else spk_default_options(&o);
        if (o.anderson_buffer < 1 || o.anderson_buffer > 8) throw std::runtime_error("anderson_buffer must be 1..8");
        if (o.max_inner < 0) throw std::runtime_error("max_inner must be >= 0");
        if (o.ls_method != SPK_LS_HOUSEHOLDER && o.ls_method != SPK_LS_GRAM)
            throw std::runtime_error("ls_method must be SPK_LS_HOUSEHOLDER or SPK_LS_GRAM");
        check(hipSetDevice(b->ctx->device), "hipSetDevice");
        if (!b->lanes && o.chunk_iters > 0) throw std::runtime_error("chunked solves need a shared batch");
        if (b->lanes) {
            if (o.anderson_buffer > spk_batch::LANE_MAXM)
                throw std::runtime_error("shared batches support anderson_buffer <= 4");
            if (b->dtype == SPK_F64) launchLaneSolve<double>(b, o, warm);
            else launchLaneSolve<float>(b, o, warm);
            return 0;
        }
        ensureAnderson(b, o.anderson_buffer);
        if (b->dtype == SPK_F64) launchSolve<double>(b, o, warm);
        else launchSolve<float>(b, o, warm);
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_solve_noalloc(spk_batch *b, const spk_options *opts, int warm) {
    try {
        spk_options o;
        if (opts) o = *opts;
        else spk_default_options(&o);
        if (b->lanes) throw std::runtime_error("spk_batch_solve_noalloc takes per-instance batches (spk_batch_create)");
        if (o.anderson_buffer < 1 || o.anderson_buffer > 8) throw std::runtime_error("anderson_buffer must be 1..8");
        if (o.max_inner < 0) throw std::runtime_error("max_inner must be >= 0");
        if (o.ls_method != SPK_LS_HOUSEHOLDER && o.ls_method != SPK_LS_GRAM)
            throw std::runtime_error("ls_method must be SPK_LS_HOUSEHOLDER or SPK_LS_GRAM");
        if (o.chunk_iters > 0) throw std::runtime_error("chunked solves need a shared batch");
        check(hipSetDevice(b->ctx->device), "hipSetDevice");
        if (b->dtype == SPK_F64) launchSolveNoAlloc<double>(b, o, warm);
        else launchSolveNoAlloc<float>(b, o, warm);
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_refill(spk_batch *b, int count, const int *lanes, const double *x0) {
    try {
        check(hipSetDevice(b->ctx->device), "hipSetDevice");
        check(hipStreamSynchronize(b->ctx->stream), "sync");
        stageLanes(b, count, lanes);
        if (b->dtype == SPK_F64) refillLanes<double, float>(b, count, x0, nullptr, 0);
        else refillLanes<float, float>(b, count, x0, nullptr, 0);
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_refill_from_batch(spk_batch *b, int count, const int *lanes, const double *x0, spk_batch *src,
                                const int *src_lanes) {
    try {
        if (!src || !src->lanes || src->probs[0]->hp.sIter != b->probs[0]->hp.sIter)
            throw std::runtime_error("the source must be a shared batch of the same problem size");
        check(hipSetDevice(b->ctx->device), "hipSetDevice");
        check(hipStreamSynchronize(src->ctx->stream), "sync");
        check(hipStreamSynchronize(b->ctx->stream), "sync");
        // the starting iterates pass through rows of the source's dtype, so an FP64 source loses nothing
        const long n = b->probs[0]->hp.sIter;
        stageLanes(src, count, src_lanes);
        if (b->rZ.bytes < (size_t) count * n * src->elSize()) b->rZ.alloc((size_t) count * n * src->elSize());
        if (src->dtype == SPK_F64) gatherLanes<double>(src, count, b->rZ.as<double>(), n);
        else gatherLanes<float>(src, count, b->rZ.as<float>(), n);
        check(hipStreamSynchronize(src->ctx->stream), "sync");
        stageLanes(b, count, lanes);
        if (b->dtype == SPK_F64) {
            if (src->dtype == SPK_F64) refillLanes<double, double>(b, count, x0, b->rZ.as<double>(), n);
            else refillLanes<double, float>(b, count, x0, b->rZ.as<float>(), n);
        } else {
            if (src->dtype == SPK_F64) refillLanes<float, double>(b, count, x0, b->rZ.as<double>(), n);
            else refillLanes<float, float>(b, count, x0, b->rZ.as<float>(), n);
        }
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_refill_rows(spk_batch *b, int count, const int *lanes, const double *x0, const float *z, long ld) {
    try {
        if (!z || ld < b->probs[0]->hp.sIter) throw std::runtime_error("starting iterates need rows of the iterate size");
        check(hipSetDevice(b->ctx->device), "hipSetDevice");
        check(hipStreamSynchronize(b->ctx->stream), "sync");
        stageLanes(b, count, lanes);
        if (b->dtype == SPK_F64) refillLanes<double>(b, count, x0, z, ld);
        else refillLanes<float>(b, count, x0, z, ld);
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_get_lanes_device(spk_batch *b, int count, const int *lanes, float *dst, long ld) {
    try {
        if (!dst || ld < b->probs[0]->hp.sIter) throw std::runtime_error("rows must hold the iterate size");
        check(hipSetDevice(b->ctx->device), "hipSetDevice");
        check(hipStreamSynchronize(b->ctx->stream), "sync");
        stageLanes(b, count, lanes);
        const long n = b->probs[0]->hp.sIter;
        if (b->dtype == SPK_F64)
            hipLaunchKernelGGL((spk::k_gather_lanes<double, float>), dim3(1024), dim3(256), 0, b->ctx->stream,
                               lanePtr<double>(b, LR_ZIO), (long) b->laneW, dst, ld, n, b->rLanes.as<int>(), count);
        else
            hipLaunchKernelGGL((spk::k_gather_lanes<float, float>), dim3(1024), dim3(256), 0, b->ctx->stream,
                               lanePtr<float>(b, LR_ZIO), (long) b->laneW, dst, ld, n, b->rLanes.as<int>(), count);
        check(hipGetLastError(), "launch k_gather_lanes");
        check(hipStreamSynchronize(b->ctx->stream), "sync");
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_pool_open(spk_batch *b, const spk_options *opts) {
    try {
        if (!b->lanes) throw std::runtime_error("a pool is a shared batch");
        if (!opts || opts->chunk_iters <= 0) throw std::runtime_error("a pool runs in chunks (chunk_iters > 0)");
        if (opts->anderson_buffer < 1 || opts->anderson_buffer > spk_batch::LANE_MAXM)
            throw std::runtime_error("shared batches support anderson_buffer 1..4");
        check(hipSetDevice(b->ctx->device), "hipSetDevice");
        check(hipStreamSynchronize(b->ctx->stream), "sync");
        const int groups = (int) (b->Bpad / 16);
        std::vector<spk::LaneSave> ls(b->Bpad);
        for (auto &v : ls) {
            std::memset(static_cast<void *>(&v), 0, sizeof(v));
            v.status = spk::ST_NOTRUN;
            v.active = 0;
            v.initFlag = 1;
        }
        std::vector<spk::LaneGroupSave> gs(groups);
        for (auto &g : gs) {
            for (int i = 0; i < 6; i++) g.roles[i] = i;
            g.head = opts->anderson_buffer - 1;
            g.finished = 1;
            g.it = 0;
            g.ticks = 0;
        }
        check(hipMemcpy(b->lsave.p, ls.data(), ls.size() * sizeof(spk::LaneSave), hipMemcpyHostToDevice), "pool state");
        check(hipMemcpy(b->lgsave.p, gs.data(), gs.size() * sizeof(spk::LaneGroupSave), hipMemcpyHostToDevice),
              "pool state");
        if (b->dOut.bytes != sizeof(spk::Out) * b->count) b->dOut.alloc(sizeof(spk::Out) * b->count);
        std::vector<spk::Out> outs(b->count);
        for (auto &o : outs) {
            std::memset(static_cast<void *>(&o), 0, sizeof(o));
            o.status = spk::ST_NOTRUN;
        }
        check(hipMemcpy(b->dOut.p, outs.data(), outs.size() * sizeof(spk::Out), hipMemcpyHostToDevice), "pool results");
        b->lastOpts = *opts;
        b->lastWarm = 0;
        b->resumable = true;
        b->pending = false;
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_set_chunk(spk_batch *b, long chunk_iters) {
    try {
        if (!b->lanes || !b->resumable) throw std::runtime_error("no chunked solve of a shared batch");
        if (chunk_iters <= 0) throw std::runtime_error("chunk_iters must be positive");
        b->lastOpts.chunk_iters = chunk_iters;
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_set_chunk_time(spk_batch *b, double seconds) {
    try {
        if (!b->lanes) throw std::runtime_error("a time slice applies to the chunked solves of a shared batch");
        if (!(seconds >= 0.) || !std::isfinite(seconds)) throw std::runtime_error("the time slice must be finite and >= 0");
        b->chunkSeconds = seconds;
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_resume(spk_batch *b) {
    try {
        if (!b->lanes || !b->resumable) throw std::runtime_error("no chunked solve of a shared batch to resume");
        check(hipSetDevice(b->ctx->device), "hipSetDevice");
        check(hipStreamSynchronize(b->ctx->stream), "sync");
        if (b->dtype == SPK_F64) launchLaneSolve<double>(b, b->lastOpts, b->lastWarm, 1);
        else launchLaneSolve<float>(b, b->lastOpts, b->lastWarm, 1);
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_solve_graph(spk_batch *b, const spk_options *opts, int warm, int launches) {
    try {
        if (!b->lanes) {  // the persistent solve launch of a per-instance batch, captured and launched
            spk_options o;
            if (opts) o = *opts;
            else spk_default_options(&o);
            if (launches != 1) throw std::runtime_error("a per-instance batch's solve is one launch: launches must be 1");
            if (o.chunk_iters > 0) throw std::runtime_error("chunked solves need a shared batch");
            if (o.anderson_buffer < 1 || o.anderson_buffer > 8) throw std::runtime_error("anderson_buffer must be 1..8");
            if (o.max_inner < 0) throw std::runtime_error("max_inner must be >= 0");
            if (o.ls_method != SPK_LS_HOUSEHOLDER && o.ls_method != SPK_LS_GRAM)
                throw std::runtime_error("ls_method must be SPK_LS_HOUSEHOLDER or SPK_LS_GRAM");
            check(hipSetDevice(b->ctx->device), "hipSetDevice");
            ensureAnderson(b, o.anderson_buffer);
            if (b->dtype == SPK_F64) launchInstGraph<double>(b, o, warm);
            else launchInstGraph<float>(b, o, warm);
            return 0;
        }
        if (!opts || opts->chunk_iters <= 0) throw std::runtime_error("a graph-captured solve needs chunk_iters > 0");
        if (launches < 1) throw std::runtime_error("launches must be >= 1");
        if (opts->anderson_buffer < 1 || opts->anderson_buffer > spk_batch::LANE_MAXM)
            throw std::runtime_error("shared batches support anderson_buffer 1..4");
        check(hipSetDevice(b->ctx->device), "hipSetDevice");
        if (b->dtype == SPK_F64) launchLaneGraph<double>(b, *opts, warm, launches);
        else launchLaneGraph<float>(b, *opts, warm, launches);
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_replay_graph(spk_batch *b) {
    try {
        if (b->lanes || !b->instGraph || !b->graphExec)
            throw std::runtime_error("no captured solve of a per-instance batch to replay (spk_batch_solve_graph)");
        check(hipSetDevice(b->ctx->device), "hipSetDevice");
        check(hipStreamSynchronize(b->ctx->stream), "sync");
        if (b->dtype == SPK_F64) replayInstGraph<double>(b);
        else replayInstGraph<float>(b);
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_resume_graph(spk_batch *b, int launches) {
    try {
        if (!b->lanes || !b->resumable) throw std::runtime_error("no chunked solve of a shared batch to resume");
        if (launches < 1) throw std::runtime_error("launches must be >= 1");
        check(hipSetDevice(b->ctx->device), "hipSetDevice");
        if (b->dtype == SPK_F64) launchLaneGraph<double>(b, b->lastOpts, b->lastWarm, launches, true);
        else launchLaneGraph<float>(b, b->lastOpts, b->lastWarm, launches, true);
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_set_schedule(spk_batch *b, int mv, int block) {
    try {
        if (!b->lanes) throw std::runtime_error("schedules apply to shared batches");
        if (mv != 0 && mv != 8 && !tileSchedule(mv)) throw std::runtime_error("schedule must be 0, 8, 16 or 17");
        const int most = scheduleMaxThreads(mv);
        if (block != 0 && (block < 64 || block > most || block % 64))
            throw std::runtime_error("block must be 0 or a multiple of 64 in 64.." + std::to_string(most));
        check(hipStreamSynchronize(b->ctx->stream), "sync");
        setLaneWidth(b, tileSchedule(mv) ? 16 : 64);
        b->laneMv = mv;
        b->laneBlockSz = block;
        b->resumable = false;  // a chunked solve continues with the schedule it started with
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_autotune(spk_batch *b, const spk_options *opts, int iters, int *mv, int *block, double *seconds) {
    try {
        if (!b->lanes) throw std::runtime_error("the autotuner tunes shared batches");
        if (iters < 1) throw std::runtime_error("iters must be >= 1");
        spk_options o;
        if (opts) o = *opts;
        else spk_default_options(&o);
        o.tol_abs = o.tol_rel = 0.;  // a fixed number of iterations
        o.max_iters = iters;
        o.chunk_iters = 0;
        o.block_size = 0;
        o.max_time_secs = 0.;
        TuneGuard guard(b);
        const int cand[SPK_AUTOTUNE_CANDIDATES][2] = {{0, 1024}, {8, 1024}, {0, 512}, {8, 512},
                                                       {16, 256}, {16, 512}, {17, 1024}};
        double best = 1e300;
        int bi = 0;
        for (int i = 0; i < SPK_AUTOTUNE_CANDIDATES; i++) {
            setLaneWidth(b, tileSchedule(cand[i][0]) ? 16 : 64);
            b->laneMv = cand[i][0];
            b->laneBlockSz = cand[i][1];
            if (spk_batch_reset(b)) throw std::runtime_error(b->ctx->err);
            if (b->dtype == SPK_F64) launchLaneSolve<double>(b, o, 0);
            else launchLaneSolve<float>(b, o, 0);
            double sec = 0.;
            if (spk_batch_results(b, nullptr, &sec)) throw std::runtime_error(b->ctx->err);
            if (seconds) seconds[i] = sec;
            if (sec < best) {
                best = sec;
                bi = i;
            }
        }
        setLaneWidth(b, tileSchedule(cand[bi][0]) ? 16 : 64);
        b->laneMv = cand[bi][0];
        b->laneBlockSz = cand[bi][1];
        b->tuned = true;
        b->resumable = false;
        guard.keep();
        if (spk_batch_reset(b)) throw std::runtime_error(b->ctx->err);
        if (mv) *mv = b->laneMv;
        if (block) *block = b->laneBlockSz;
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_autotune_uneven(spk_batch *b, const spk_options *opts, int iters, int *block, int *largest_first,
                              double *seconds) {
    try {
        if (b->lanes) throw std::runtime_error("the uneven-batch autotuner tunes per-instance batches");
        if (iters < 1) throw std::runtime_error("iters must be >= 1");
        spk_options o;
        if (opts) o = *opts;
        else spk_default_options(&o);
        o.tol_abs = o.tol_rel = 0.;  // a fixed number of iterations
        o.max_iters = iters;
        o.max_time_secs = 0.;
        o.block_size = 0;
        TuneGuard guard(b);
        const int fb = b->fusedBlock;
        const bool lf = b->largestFirst;
        const int cand[SPK_AUTOTUNE_UNEVEN_CANDIDATES][2] = {{128, 0}, {256, 0}, {512, 0}, {128, 1}, {256, 1}, {512, 1}};
        double best = 1e300;
        int bi = 1;
        try {
            for (int i = 0; i < SPK_AUTOTUNE_UNEVEN_CANDIDATES; i++) {
                b->fusedBlock = cand[i][0];
                b->largestFirst = cand[i][1] != 0;
                if (spk_batch_reset(b)) throw std::runtime_error(b->ctx->err);
                if (spk_batch_solve(b, &o, 0)) throw std::runtime_error(b->ctx->err);
                double sec = 0.;
                if (spk_batch_results(b, nullptr, &sec)) throw std::runtime_error(b->ctx->err);
                if (seconds) seconds[i] = sec;
                if (sec < best) {
                    best = sec;
                    bi = i;
                }
            }
        } catch (...) {
            b->fusedBlock = fb;
            b->largestFirst = lf;
            throw;
        }
        b->fusedBlock = cand[bi][0];
        b->largestFirst = cand[bi][1] != 0;
        guard.keep();
        if (spk_batch_reset(b)) throw std::runtime_error(b->ctx->err);
        if (block) *block = b->fusedBlock;
        if (largest_first) *largest_first = b->largestFirst ? 1 : 0;
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_set_uneven_schedule(spk_batch *b, int block, int largest_first) {
    try {
        if (b->lanes) throw std::runtime_error("uneven-batch schedules apply to per-instance batches");
        if (block != 0 && (block < 64 || block > 512 || block % 64))
            throw std::runtime_error("block must be 0 or a multiple of 64 in 64..512");
        check(hipStreamSynchronize(b->ctx->stream), "sync");
        b->fusedBlock = block;
        b->largestFirst = largest_first != 0;
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_autotune_workgroups(spk_batch *b, const spk_options *opts, int iters, int *workgroups,
                                  double *seconds) {
    try {
        if (b->lanes || b->count != 1) throw std::runtime_error("the workgroup autotuner tunes one-instance batches");
        if (iters < 1) throw std::runtime_error("iters must be >= 1");
        spk_options o;
        if (opts) o = *opts;
        else spk_default_options(&o);
        o.tol_abs = o.tol_rel = 0.;  // a fixed number of iterations
        o.max_iters = iters;
        o.max_time_secs = 0.;
        TuneGuard guard(b);
        const int cand[5] = {1, 8, 32, 64, 128};
        double best = 1e300;
        int bw = 1;
        for (int i = 0; i < 5; i++) {
            b->wgs = cand[i];
            if (spk_batch_reset(b)) throw std::runtime_error(b->ctx->err);
            if (spk_batch_solve(b, &o, 0)) throw std::runtime_error(b->ctx->err);
            double sec = 0.;
            if (spk_batch_results(b, nullptr, &sec)) throw std::runtime_error(b->ctx->err);
            if (seconds) seconds[i] = sec;
            if (sec < best) {
                best = sec;
                bw = cand[i];
            }
        }
        b->wgs = bw;
        guard.keep();
        if (spk_batch_reset(b)) throw std::runtime_error(b->ctx->err);
        if (workgroups) *workgroups = bw;
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_results(spk_batch *b, spk_result *out, double *kernelSeconds) {
    try {
        check(hipStreamSynchronize(b->ctx->stream), "sync after solve");
        if (!b->pending) throw std::runtime_error("no solve has been run on this batch");
        if (kernelSeconds) {
            float ms = 0.f;
            check(hipEventElapsedTime(&ms, b->ev0, b->ev1), "event time");
            *kernelSeconds = ms * 1e-3;
        }
        if (out) {
            std::vector<spk::Out> tmp(b->count);
            check(hipMemcpy(tmp.data(), b->dOut.p, sizeof(spk::Out) * b->count, hipMemcpyDeviceToHost), "results");
            for (int k = 0; k < b->count; k++) {
                out[k].status = tmp[k].status;
                out[k].iters = tmp[k].iters;
                out[k].err = tmp[k].err;
                out[k].tol = tmp[k].tol;
                out[k].w = tmp[k].w;
                out[k].k0 = tmp[k].k0;
                out[k].k1 = tmp[k].k1;
                out[k].k2 = tmp[k].k2;
                out[k].k2bt = tmp[k].k2bt;
                out[k].k3 = tmp[k].k3;
                out[k].num_T = tmp[k].nT;
                out[k].num_ell = tmp[k].nEll;
                for (int ph = 0; ph < spk::NUM_PHASES; ph++)
                    out[k].phase_seconds[ph] = tmp[k].ticks[ph] / b->ctx->tickRateHz;
            }
        }
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

int spk_batch_debug_op(spk_batch *b, int op, const double *in, double *out, double *scalars) {
    try {
        if (op < SPK_OP_L || op > SPK_OP_DUAL) throw std::runtime_error("unknown debug op");
        if (op == SPK_OP_DUAL && b->lanes) throw std::runtime_error("the DUAL op needs a per-instance batch");
        check(hipSetDevice(b->ctx->device), "hipSetDevice");
        check(hipStreamSynchronize(b->ctx->stream), "sync");
        if (b->lanes) {
            const size_t n = b->probs[0]->hp.sIter;
            const bool f64 = b->dtype == SPK_F64;
            std::vector<double> tin, tout;
            const double *inE = engineOrder(b->probs[0]->hp, in, b->count, tin);
            const double *outE = engineOrder(b->probs[0]->hp, out, b->count, tout);
            if (f64) {
                laneUpload<double>(b, LR_ZIO, inE, n);
                laneUpload<double>(b, 1, outE, n);
            } else {
                laneUpload<float>(b, LR_ZIO, inE, n);
                laneUpload<float>(b, 1, outE, n);
            }
            spk_options o;
            spk_default_options(&o);
            b->dOut.alloc(sizeof(spk::Out) * b->count);
            check(hipMemset(b->dOut.p, 0, sizeof(spk::Out) * b->count), "zero out");
            const dim3 grid((unsigned) (b->Bpad / b->laneW)), block(laneBlock(b, o));
            auto *outs = static_cast<spk::Out *>(b->dOut.p);
            const bool tile = b->laneW == 16;
            if (f64) {
                buildLaneInst<double>(b, o, 0);
                auto *P = static_cast<const spk::LaneInst<double> *>(b->dInst.p);
                if (b->laneMv == 17) hipLaunchKernelGGL(spk::k_tile_debug_x1024<double>, grid, block, 0, b->ctx->stream, P, outs, op);
                else if (tile) hipLaunchKernelGGL(spk::k_tile_debug<double>, grid, block, 0, b->ctx->stream, P, outs, op);
                else if (b->laneMv == 8) hipLaunchKernelGGL(spk::k_lane_debug_x8<double>, grid, block, 0, b->ctx->stream, P, outs, op);
                else hipLaunchKernelGGL(spk::k_lane_debug<double>, grid, block, 0, b->ctx->stream, P, outs, op);
            } else {
                buildLaneInst<float>(b, o, 0);
                auto *P = static_cast<const spk::LaneInst<float> *>(b->dInst.p);
                if (b->laneMv == 17) hipLaunchKernelGGL(spk::k_tile_debug_x1024<float>, grid, block, 0, b->ctx->stream, P, outs, op);
                else if (tile) hipLaunchKernelGGL(spk::k_tile_debug<float>, grid, block, 0, b->ctx->stream, P, outs, op);
                else if (b->laneMv == 8) hipLaunchKernelGGL(spk::k_lane_debug_x8<float>, grid, block, 0, b->ctx->stream, P, outs, op);
                else hipLaunchKernelGGL(spk::k_lane_debug<float>, grid, block, 0, b->ctx->stream, P, outs, op);
            }
            check(hipGetLastError(), "launch k_lane_debug");
            check(hipStreamSynchronize(b->ctx->stream), "sync debug op");
            if (f64) laneDownload<double>(b, 1, out, n);
            else laneDownload<float>(b, 1, out, n);
            referenceOrder(b->probs[0]->hp, out, b->count);
            std::vector<spk::Out> res(b->count);
            check(hipMemcpy(res.data(), b->dOut.p, sizeof(spk::Out) * b->count, hipMemcpyDeviceToHost), "results");
            if (scalars)
                for (int k = 0; k < b->count; k++) {
                    scalars[2 * k] = res[k].w;
                    scalars[2 * k + 1] = res[k].err;
                }
            return 0;
        }
        size_t pos = 0;
        for (int k = 0; k < b->count; k++) {
            const size_t n = b->probs[k]->hp.sIter;
            std::vector<double> tin, tout;
            const double *inE = engineOrder(b->probs[k]->hp, in + pos, 1, tin);
            const double *outE = engineOrder(b->probs[k]->hp, out + pos, 1, tout);
            if (b->dtype == SPK_F64) {
                toDevice<double>(b, corePtr<double>(b, k), inE, n);
                toDevice<double>(b, corePtr<double>(b, k) + 2 * n, outE, n);
            } else {
                toDevice<float>(b, corePtr<float>(b, k), inE, n);
                toDevice<float>(b, corePtr<float>(b, k) + 2 * n, outE, n);
            }
            pos += n;
        }
        spk_options o;
        spk_default_options(&o);
        ensureAnderson(b, o.anderson_buffer);
        b->dOut.alloc(sizeof(spk::Out) * b->count);
        check(hipMemset(b->dOut.p, 0, sizeof(spk::Out) * b->count), "zero out");
        if (b->dtype == SPK_F64) {
            buildInsts<double>(b, o, 0);
            launchKernel<double>(b, 256, true, op);
        } else {
            buildInsts<float>(b, o, 0);
            launchKernel<float>(b, 256, true, op);
        }
        check(hipGetLastError(), "launch k_debug");
        check(hipStreamSynchronize(b->ctx->stream), "sync debug op");
        pos = 0;
        std::vector<spk::Out> res(b->count);
        check(hipMemcpy(res.data(), b->dOut.p, sizeof(spk::Out) * b->count, hipMemcpyDeviceToHost), "results");
        for (int k = 0; k < b->count; k++) {
            const size_t n = b->probs[k]->hp.sIter;
            if (b->dtype == SPK_F64) toHost<double>(b, out + pos, corePtr<double>(b, k) + 2 * n, n);
            else toHost<float>(b, out + pos, corePtr<float>(b, k) + 2 * n, n);
            referenceOrder(b->probs[k]->hp, out + pos, 1);
            if (scalars) {
                scalars[2 * k] = res[k].w;
                scalars[2 * k + 1] = res[k].err;
            }
            pos += n;
        }
        return 0;
    } catch (const std::exception &e) {
        return fail(b->ctx, e);
    }
}

}  // extern "C"
