// C++ drop-in for the reference's solver classes, over the C ABI of libspock_amd.so.
//
// The class and method names, argument meanings and error behaviour follow the reference so
// that its example drivers compile against this header with the include swapped:
//   ScenarioTree<T>(path)                       src/tree.cuh:102-179
//   ProblemData<T>(tree)                        src/problem.cuh:126-153
//   CacheBuilder<T>(tree, data).tol(...)...     src/cache.cuh:1154-1244 (defaults: tolAbs 1e-3,
//                                               tolRel 0, maxTime 0, maxIters 0, maxInner 8, m 3)
//   Cache<T>::runSpock / runCp (x0, previous)   src/cache.cuh:353-355, 967-1132
//   status / solveIter / solveTime / states / inputs / solution / solutionPrim / solutionDual /
//   getDual / reset / sizePrimal / sizeDual     src/cache.cuh:357-412, 436-479
//   CacheBuilder<T>::build / make_unique        src/cache.cuh:1215-1244
//   enum CacheStatus and toString               src/cache.cuh:68-88
//   CacheBuilder::enableDebug, the debug caches  src/cache.cuh:947-959, 1088-1130, 1204-1209,
//   and printToJson (log/cacheSp.json, ...Cp)    :1274-1311
// Errors are std::invalid_argument with the library's message, as the reference throws.
// T = double solves in FP64, T = float in FP32 (the reference instantiates both,
// src/definitions.cu:24-443).  solution(), solutionPrim() and solutionDual() return a DTensor<T>&
// and getDual takes and returns DTensor<T>, as the reference does (src/cache.cuh:371-382); here
// DTensor<T> is a host copy (the iterate stays on the device inside the library), refreshed by
// each call, with the reference's download(std::vector<T>&) and numEl(), and it converts to
// std::vector<T>.
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <iostream>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "spock_amd.h"

namespace spock_amd {

enum CacheStatus { notRun = -1, converged = 0, outOfIter = 1, outOfTime = 2 };

// The reference's status strings, which its drivers print (src/cache.cuh:75-88).
static inline const char *toString(int status) {
    switch (status) {
        case notRun: return "Cache has not been run!";
        case converged: return "Converged!";
        case outOfIter: return "Out of iterations!";
        case outOfTime: return "Out of time!";
        default: return "Unknown status.";
    }
}

namespace detail {
inline spk_ctx *context() {  // one context (and HIP stream) per host thread
    thread_local struct Holder {
        spk_ctx *c = nullptr;
        Holder() {
            if (spk_ctx_create(0, &c) != 0) throw std::invalid_argument("[spock_amd] no HIP device");
        }
        ~Holder() { spk_ctx_destroy(c); }
    } h;
    return h.c;
}
inline void check(int rc, const char *what) {
    if (rc != 0) throw std::invalid_argument(std::string("[") + what + "] " + spk_last_error(context()));
}
}  // namespace detail

// Reads a .bt tensor file (three uint64 dimensions, then the data in column-major order; the
// element type is fixed by the file suffix: _d double, _f float, _u uint64) into a flat vector,
// as DTensor<T>::parseFromFile does for the reference's drivers (examples/random/main.cu:48-50).
template <typename T>
std::vector<T> parseFromFile(const std::string &file) {
    std::ifstream in(file, std::ios::binary);
    if (!in) throw std::invalid_argument("[parseFromFile] cannot open " + file);
    uint64_t dims[3];
    in.read(reinterpret_cast<char *>(dims), sizeof(dims));
    const size_t n = (size_t) (dims[0] * dims[1] * dims[2]);
    std::vector<T> out(n);
    const std::string sfx = file.size() >= 5 ? file.substr(file.size() - 5) : "";
    if (sfx == "_d.bt") {
        std::vector<double> v(n);
        in.read(reinterpret_cast<char *>(v.data()), n * sizeof(double));
        for (size_t k = 0; k < n; k++) out[k] = (T) v[k];
    } else if (sfx == "_f.bt") {
        std::vector<float> v(n);
        in.read(reinterpret_cast<char *>(v.data()), n * sizeof(float));
        for (size_t k = 0; k < n; k++) out[k] = (T) v[k];
    } else {
        std::vector<uint64_t> v(n);
        in.read(reinterpret_cast<char *>(v.data()), n * sizeof(uint64_t));
        for (size_t k = 0; k < n; k++) out[k] = (T) v[k];
    }
    if (!in) throw std::invalid_argument("[parseFromFile] short file " + file);
    return out;
}

// The part of GPUtils' DTensor<T> that the reference's callers use: DTensor<T>::parseFromFile(path)
// then download(vector) (examples/*/main.cu), and the tensors Cache::solution*() and getDual
// return.  Here a host copy in column-major order; it converts to std::vector<T>.
template <typename T>
class DTensor {
public:
    DTensor() = default;
    explicit DTensor(std::vector<T> data) : m_data(std::move(data)) {}
    static DTensor parseFromFile(const std::string &file) { return DTensor(spock_amd::parseFromFile<T>(file)); }
    void download(std::vector<T> &out) const { out.assign(m_data.begin(), m_data.end()); }
    void upload(const std::vector<T> &in) { m_data.assign(in.begin(), in.end()); }
    size_t numEl() const { return m_data.size(); }
    size_t numRows() const { return m_data.size(); }
    size_t numCols() const { return 1; }
    size_t numMats() const { return 1; }
    operator std::vector<T>() const { return m_data; }
    const std::vector<T> &values() const { return m_data; }

private:
    std::vector<T> m_data;
};

template <typename T>
class ScenarioTree {
public:
    explicit ScenarioTree(std::string path = "./data/") : m_path(std::move(path)) {
        if (!m_path.empty() && m_path.back() != '/') m_path += '/';
        std::ifstream in(m_path + "data.json");
        if (!in) throw std::invalid_argument("[ScenarioTree] no data.json in " + m_path);
        std::stringstream ss;
        ss << in.rdbuf();
        m_json = ss.str();
    }
    const std::string &path() const { return m_path; }
    std::string fpFileExt() const { return sizeof(T) == 8 ? "_d.bt" : "_f.bt"; }
    size_t numStates() const { return field("numStates"); }
    size_t numInputs() const { return field("numInputs"); }
    size_t numStatesAndInputs() const { return numStates() + numInputs(); }  // src/tree.cuh:121
    size_t numNodes() const { return field("numNodes"); }
    size_t numNonleafNodes() const { return field("numNonleafNodes"); }
    size_t numStages() const { return field("numStages"); }
    size_t numEvents() const { return field("numEvents"); }
    // the largest number of children of any node of each stage, 0 for the last (src/tree.cuh:171-178)
    std::vector<size_t> childMax() const {
        const std::vector<size_t> nc = spock_amd::parseFromFile<size_t>(m_path + "numChildren_u.bt");
        const std::vector<size_t> from = spock_amd::parseFromFile<size_t>(m_path + "stageFrom_u.bt");
        const std::vector<size_t> to = spock_amd::parseFromFile<size_t>(m_path + "stageTo_u.bt");
        std::vector<size_t> out(numStages(), 0);
        for (size_t st = 0; st + 1 < numStages() && st < from.size() && st < to.size(); st++)
            for (size_t n = from[st]; n <= to[st] && n < nc.size(); n++) out[st] = std::max(out[st], nc[n]);
        return out;
    }

private:
    size_t field(const char *key) const {
        const size_t k = m_json.find(std::string("\"") + key + "\"");
        if (k == std::string::npos) throw std::invalid_argument(std::string("[ScenarioTree] missing ") + key);
        return (size_t) std::stoul(m_json.substr(m_json.find(':', k) + 1));
    }
    std::string m_path, m_json;
};

template <typename T>
class ProblemData {
public:
    explicit ProblemData(ScenarioTree<T> &tree) {
        detail::check(spk_problem_load(detail::context(), tree.path().c_str(), &m_p), "ProblemData");
        long s[16];
        spk_problem_sizes(m_p, s);
        m_numNodes = s[0];
        m_numNonleaf = s[1];
        m_numStates = s[4];
        m_numInputs = s[5];
        m_sizePrim = s[6];
        m_sizeDual = s[7];
        m_scaling.resize(m_numStates + m_numInputs);
        spk_problem_scaling(m_p, m_scaling.data());
    }
    ~ProblemData() { spk_problem_destroy(m_p); }
    ProblemData(const ProblemData &) = delete;
    ProblemData &operator=(const ProblemData &) = delete;
    spk_problem *handle() const { return m_p; }
    size_t numStates() const { return m_numStates; }
    size_t numInputs() const { return m_numInputs; }
    size_t numNodes() const { return m_numNodes; }
    size_t numNonleafNodes() const { return m_numNonleaf; }
    size_t sizePrimal() const { return m_sizePrim; }
    size_t sizeDual() const { return m_sizeDual; }
    const std::vector<double> &scaling() const { return m_scaling; }
    T stepSize() const { return (T) spk_problem_step_size(m_p); }

private:
    spk_problem *m_p = nullptr;
    size_t m_numNodes = 0, m_numNonleaf = 0, m_numStates = 0, m_numInputs = 0, m_sizePrim = 0, m_sizeDual = 0;
    std::vector<double> m_scaling;
};

template <typename T>
class Cache {
public:
    // Iterate sizes for which the automatic choice spreads one instance over kAutoGridWorkgroups
    // cooperating workgroups (grid mode): the range of the 229-node, 52,360-element problem this
    // has been run and measured on (2.0 ms per outer iteration with 32 workgroups, 2.4 with 64,
    // 8.9 with one).  Elsewhere one workgroup solves the instance unless workgroups(n) asks
    // otherwise.
    static constexpr size_t kAutoGridMin = 50000, kAutoGridMax = 100000;
    static constexpr int kAutoGridWorkgroups = 32;

    // workgroups: how many workgroups solve the instance together (0: automatic, see above)
    // debug: CacheBuilder::enableDebug (src/cache.cuh:1204-1209, 312-318)
    Cache(ScenarioTree<T> &tree, ProblemData<T> &data, const spk_options &opts, int workgroups = 0,
          bool debug = false)
        : m_data(&data), m_opts(opts), m_debug(debug) {
        spk_problem *p = data.handle();
        detail::check(spk_batch_create(detail::context(), 1, &p, sizeof(T) == 8 ? SPK_F64 : SPK_F32, &m_b),
                      "Cache");
        if (m_debug) {
            const size_t len = (size_t) std::max<long>(m_opts.max_iters, 0);
            m_cacheCallsToL.assign(len, 0);
            m_cacheError0.assign(len, T(0));
            m_cacheError1.assign(len, T(0));
            m_cacheError2.assign(len, T(0));
            m_cacheError3.assign(len, T(0));
            const long period = m_opts.check_period > 0 ? m_opts.check_period : 25;
            m_records = m_opts.max_iters / period + 2;
            detail::check(spk_batch_set_debug(m_b, m_records), "Cache");
            m_horizon = tree.numStages() - 1;
            const std::vector<size_t> cm = tree.childMax();
            m_branchFactor = (size_t) (std::find(cm.begin(), cm.end(), (size_t) 1) - cm.begin());
            m_numEvents = tree.numEvents();
        }
        if (workgroups == 0) {
            const size_t n = data.sizePrimal() + data.sizeDual();
            workgroups = (n >= kAutoGridMin && n <= kAutoGridMax) ? kAutoGridWorkgroups : 1;
            m_autoGrid = workgroups > 1;
        }
        m_workgroups = workgroups;
        detail::check(spk_batch_set_workgroups(m_b, workgroups), "Cache");
    }
    ~Cache() { spk_batch_destroy(m_b); }
    Cache(Cache &&o) noexcept
        : m_data(o.m_data), m_opts(o.m_opts), m_debug(o.m_debug), m_b(o.m_b), m_autoGrid(o.m_autoGrid),
          m_workgroups(o.m_workgroups), m_gridFallback(o.m_gridFallback), m_exit(o.m_exit), m_iters(o.m_iters),
          m_time(o.m_time), m_tol(o.m_tol), m_records(o.m_records), m_callsToL(o.m_callsToL),
          m_horizon(o.m_horizon), m_branchFactor(o.m_branchFactor), m_numEvents(o.m_numEvents),
          m_cacheCallsToL(std::move(o.m_cacheCallsToL)), m_cacheError0(std::move(o.m_cacheError0)),
          m_cacheError1(std::move(o.m_cacheError1)), m_cacheError2(std::move(o.m_cacheError2)),
          m_cacheError3(std::move(o.m_cacheError3)) {
        o.m_b = nullptr;
    }
    Cache(const Cache &) = delete;

    int runSpock(std::vector<T> &x0, std::vector<T> *previousSolution = nullptr) {
        return run(SPK_SPOCK, x0, previousSolution);
    }
    int runCp(std::vector<T> &x0, std::vector<T> *previousSolution = nullptr) { return run(SPK_CP, x0, previousSolution); }

    int status() const { return m_exit; }
    size_t solveIter() const { return m_iters; }
    T solveTime() const { return m_time; }
    size_t sizePrimal() const { return m_data->sizePrimal(); }
    size_t sizeDual() const { return m_data->sizeDual(); }

    // The accepted iterate [prim | dual] and its parts (src/cache.cuh:378-382), copied from the
    // device by each call.
    DTensor<T> &solution() {
        m_solution.upload(iterate());
        return m_solution;
    }
    DTensor<T> &solutionPrim() {
        std::vector<T> z = iterate();
        z.resize(sizePrimal());
        m_solutionPrim.upload(z);
        return m_solutionPrim;
    }
    DTensor<T> &solutionDual() {
        const std::vector<T> z = iterate();
        m_solutionDual.upload(std::vector<T>(z.begin() + sizePrimal(), z.end()));
        return m_solutionDual;
    }

    // Dual point of the CP step from the current iterate with `primal` in place of the primal prox
    // (src/cache.cuh:371-376): modifyDual + proximalDual.
    DTensor<T> getDual(DTensor<T> &primal) { return DTensor<T>(getDual(primal.values())); }
    std::vector<T> getDual(const std::vector<T> &primal) {
        if (primal.size() != sizePrimal()) throw std::invalid_argument("[getDual] primal has the wrong size");
        std::vector<double> z(sizePrimal() + sizeDual()), out(z.size(), 0.);
        detail::check(spk_batch_get_iterate(m_b, 0, z.data()), "getDual");
        std::copy(primal.begin(), primal.end(), out.begin());
        detail::check(spk_batch_debug_op(m_b, SPK_OP_DUAL, z.data(), out.data(), nullptr), "getDual");
        if (m_debug) m_callsToL += 1;  // modifyDual's L(true) (src/cache.cuh:771-776)
        return std::vector<T>(out.begin() + sizePrimal(), out.end());
    }

    // Workgroups the last solve ran with, and whether a refused cooperative launch made the
    // automatic grid mode fall back to one workgroup (no counterpart in the reference).
    int workgroups() const { return m_workgroups; }
    bool gridFallback() const { return m_gridFallback; }

    // states x (numNodes x numStates) and inputs u (numNonleafNodes x numInputs) of the solution,
    // undoing the preconditioning (src/cache.cuh:386-412)
    std::vector<T> states() {
        const std::vector<T> z = iterate();
        const size_t nx = m_data->numStates(), nu = m_data->numInputs(), oX = m_data->numNonleafNodes() * nu;
        std::vector<T> x(z.begin() + oX, z.begin() + oX + m_data->numNodes() * nx);
        for (size_t k = 0; k < x.size(); k++) x[k] /= (T) m_data->scaling()[k % nx];
        return x;
    }
    std::vector<T> inputs() {
        const std::vector<T> z = iterate();
        const size_t nx = m_data->numStates(), nu = m_data->numInputs();
        std::vector<T> u(z.begin(), z.begin() + m_data->numNonleafNodes() * nu);
        for (size_t k = 0; k < u.size(); k++) u[k] /= (T) m_data->scaling()[nx + k % nu];
        return u;
    }

    // zeroes the iterates and, as the reference, the callsL, err0, err1 and err2 caches but not err3
    // nor the running count of calls to L (src/cache.cuh:437-478)
    void reset() {
        detail::check(spk_batch_reset(m_b), "reset");
        m_exit = notRun;
        std::fill(m_cacheCallsToL.begin(), m_cacheCallsToL.end(), 0);
        std::fill(m_cacheError0.begin(), m_cacheError0.end(), T(0));
        std::fill(m_cacheError1.begin(), m_cacheError1.end(), T(0));
        std::fill(m_cacheError2.begin(), m_cacheError2.end(), T(0));
    }

    // the debug caches, one entry per outer iteration (zero where no error check ran)
    const std::vector<size_t> &cacheCallsToL() const { return m_cacheCallsToL; }
    const std::vector<T> &cacheError0() const { return m_cacheError0; }
    const std::vector<T> &cacheError1() const { return m_cacheError1; }
    const std::vector<T> &cacheError2() const { return m_cacheError2; }
    const std::vector<T> &cacheError3() const { return m_cacheError3; }

    // log/cache<file>.json under the working directory (src/cache.cuh:1274-1311)
    void printToJson(const std::string &file) {
        namespace fs = std::filesystem;
        const fs::path path = fs::current_path() / "log" / ("cache" + file + ".json");
        fs::create_directories(path.parent_path());
        std::ofstream f(path);
        if (!f.is_open()) throw std::runtime_error("[Cache::printToJson] Failed to open file for writing.\n");
        auto num = [&](double v) {
            if (!std::isfinite(v)) { f << "null"; return; }
            char buf[40];
            std::snprintf(buf, sizeof(buf), "%.17g", v);
            f << buf;
        };
        auto arr = [&](const char *name, const auto &v) {
            f << ",\"" << name << "\":[";
            for (size_t k = 0; k < v.size(); k++) {
                if (k) f << ",";
                num((double) v[k]);
            }
            f << "]";
        };
        f << "{\"maxIters\":" << m_opts.max_iters << ",\"tol\":";
        num((double) m_tol);
        f << ",\"horizon\":" << m_horizon << ",\"branchFactor\":" << m_branchFactor << ",\"numEvents\":"
          << m_numEvents << ",\"numStates\":" << m_data->numStates() << ",\"numInputs\":" << m_data->numInputs()
          << ",\"sizeCache\":" << m_opts.max_iters;
        arr("callsL", m_cacheCallsToL);
        arr("err0", m_cacheError0);
        arr("err1", m_cacheError1);
        arr("err2", m_cacheError2);
        arr("err3", m_cacheError3);
        arr("states", states());
        arr("inputs", inputs());
        f << "}";
        f.close();
        if (!f.good()) throw std::runtime_error("[Cache::printToJson] Can't write the JSON string to the file.\n");
    }

private:
    std::vector<T> iterate() {
        std::vector<double> z(sizePrimal() + sizeDual());
        detail::check(spk_batch_get_iterate(m_b, 0, z.data()), "solution");
        return std::vector<T>(z.begin(), z.end());
    }

    int run(int method, std::vector<T> &x0, std::vector<T> *prev) {
        if (x0.size() != m_data->numStates())
            throw std::invalid_argument("[initialiseState] Error initialising state: problem setup for " +
                                        std::to_string(m_data->numStates()) + " but given " +
                                        std::to_string(x0.size()) + " states");
        const auto t0 = std::chrono::high_resolution_clock::now();
        std::vector<double> xd(x0.begin(), x0.end());
        detail::check(spk_batch_set_x0(m_b, xd.data()), "initialiseState");
        if (prev) {
            std::vector<double> zd(prev->begin(), prev->end());
            detail::check(spk_batch_set_iterate(m_b, 0, zd.data()), "initialisePrev");
        }
        spk_options o = m_opts;
        o.method = method;
        int rc = spk_batch_solve(m_b, &o, prev ? 1 : 0);
        if (rc == SPK_ERR_COOPERATIVE_LAUNCH && m_autoGrid) {
            // the automatic grid mode needs a cooperative launch, which the device can refuse (it
            // did while another live process had made cooperative launches); the refused launch
            // queued nothing, and from then on one workgroup solves the instance
            m_autoGrid = false;
            m_gridFallback = true;
            m_workgroups = 1;
            detail::check(spk_batch_set_workgroups(m_b, 1), "Cache");
            rc = spk_batch_solve(m_b, &o, prev ? 1 : 0);
        }
        detail::check(rc, method == SPK_SPOCK ? "runSpock" : "runCp");
        spk_result r;
        detail::check(spk_batch_results(m_b, &r, nullptr), "results");
        m_exit = r.status;
        m_iters = (size_t) r.iters;
        m_tol = (T) r.tol;
        m_time = (T) std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
        if (m_debug) debugReport(method, r);
        return m_exit;
    }

    // the debug caches from the solve's records, then the reference's JSON and console summary
    // (src/cache.cuh:947-959, 988-1001, 1101-1130)
    void debugReport(int method, const spk_result &r) {
        std::vector<double> rec((size_t) m_records * SPK_DEBUG_WORDS);
        long made = 0, bad = -1;
        detail::check(spk_batch_debug_records(m_b, 0, m_records, rec.data(), &made, &bad), "debug");
        const size_t len = m_cacheError0.size();
        for (long i = 0; i < std::min(made, m_records); i++) {
            const double *q = &rec[(size_t) i * SPK_DEBUG_WORDS];
            const size_t idx = (size_t) q[0];
            if (idx >= len) continue;  // the check at maxIters itself (the reference writes past its caches)
            m_cacheCallsToL[idx] = m_callsToL + (size_t) q[1];
            m_cacheError0[idx] = (T) q[2];
            m_cacheError1[idx] = (T) q[3];
            m_cacheError2[idx] = (T) q[4];
            m_cacheError3[idx] = (T) q[5];
        }
        m_callsToL += 2 * (size_t) r.num_T;
        if (r.status == SPK_NONFINITE)
            throw std::invalid_argument("[isFinite] DTensor has a non-finite entry at (" + std::to_string(bad) +
                                        ").\n");
        printToJson(method == SPK_SPOCK ? "Sp" : "Cp");
        if (method == SPK_CP) {
            if (m_exit == converged)
                std::cout << "\nConverged in " << m_iters << " iterations, to a tolerance of " << m_tol << "\n";
            else if (m_exit == outOfIter)
                std::cout << "\nOut of iterations (" << m_opts.max_iters << " iters).\n";
            else if (m_exit == outOfTime)
                std::cout << "\nOut of time (" << m_opts.max_time_secs << " secs).\n";
            else
                std::cout << "\nExited.\n";
            return;
        }
        const std::string counts = ", K1: " + std::to_string(r.k1) + ", K2: " + std::to_string(r.k2) +
                                   ", bt: " + std::to_string(r.k2bt) + ", K3: " + std::to_string(r.k3) + "].\n";
        if (m_exit == converged)
            std::cout << "\nConverged in " << m_iters << " iters and " << m_time << " secs, to a tolerance of "
                      << m_tol << ", [K0: " << r.k0 << counts;
        else if (m_exit == outOfIter)
            std::cout << "\nMax iterations (" << m_opts.max_iters << ") reached [K0: " << r.k0 << counts;
        else if (m_exit == outOfTime)
            std::cout << "\nMax time (" << m_opts.max_time_secs << "s) reached [K0: " << r.k0 << counts;
        else
            std::cout << "\nExited.\n";
    }

    ProblemData<T> *m_data;
    spk_options m_opts;
    bool m_debug = false;
    spk_batch *m_b = nullptr;
    bool m_autoGrid = false;  // grid mode chosen by size, not asked for
    int m_workgroups = 1;
    bool m_gridFallback = false;
    DTensor<T> m_solution, m_solutionPrim, m_solutionDual;
    int m_exit = notRun;
    size_t m_iters = 0;
    T m_time = 0;
    T m_tol = 0;
    // debug caches; m_callsToL runs over every solve of this cache, as the reference's counter,
    // which nothing resets (src/cache.cuh:111)
    long m_records = 0;
    size_t m_callsToL = 0;
    size_t m_horizon = 0, m_branchFactor = 0, m_numEvents = 0;
    std::vector<size_t> m_cacheCallsToL;
    std::vector<T> m_cacheError0, m_cacheError1, m_cacheError2, m_cacheError3;
};

template <typename T>
class CacheBuilder {
public:
    CacheBuilder(ScenarioTree<T> &tree, ProblemData<T> &data) : m_tree(tree), m_data(data) {
        spk_default_options(&m_opts);  // tolAbs 1e-3, tolRel 0, maxInner 8, Anderson 3 (src/cache.cuh:1154-1163)
    }
    CacheBuilder &toleranceAbsolute(T tol) { m_opts.tol_abs = tol; return *this; }
    CacheBuilder &toleranceRelative(T tol) { m_opts.tol_rel = tol; return *this; }
    CacheBuilder &tol(T tol) { m_opts.tol_abs = tol; m_opts.tol_rel = tol; return *this; }
    CacheBuilder &maxTimeSecs(T time) { m_opts.max_time_secs = time; return *this; }
    CacheBuilder &maxIters(size_t iters) { m_opts.max_iters = (long) iters; return *this; }
    CacheBuilder &maxItersInner(size_t iters) { m_opts.max_inner = (int) iters; return *this; }
    CacheBuilder &andersonBuffer(size_t buffer) { m_opts.anderson_buffer = (int) buffer; return *this; }
    // debug caches and JSON dump; with no iteration cap set yet, caps at 1,000 (src/cache.cuh:1204-1209)
    CacheBuilder &enableDebug(bool enable = true) {
        if (enable && m_opts.max_iters == 0) m_opts.max_iters = 1000;
        m_debug = enable;
        return *this;
    }
    // workgroups per solve (no counterpart in the reference): 0 automatic, 1, or n > 1 (grid mode)
    CacheBuilder &workgroups(int n) { m_wgs = n; return *this; }
    Cache<T> build() { return Cache<T>(m_tree, m_data, m_opts, m_wgs, m_debug); }
    std::unique_ptr<Cache<T>> make_unique() {
        return std::make_unique<Cache<T>>(m_tree, m_data, m_opts, m_wgs, m_debug);
    }

private:
    ScenarioTree<T> &m_tree;
    ProblemData<T> &m_data;
    spk_options m_opts;
    int m_wgs = 0;
    bool m_debug = false;
};

}  // namespace spock_amd
