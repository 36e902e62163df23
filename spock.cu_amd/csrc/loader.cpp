// Problem folder loader: data.json + .bt tensors -> HostProblem with deduplicated tables.
#include "loader.hpp"

#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace spk {

namespace {

struct JsonCursor {
    const std::string &s;
    size_t i = 0;
    explicit JsonCursor(const std::string &str) : s(str) {}
    void ws() {
        while (i < s.size() && std::isspace((unsigned char) s[i])) i++;
    }
    [[noreturn]] void fail(const char *what) const {
        throw std::runtime_error(std::string("data.json: ") + what + " at offset " + std::to_string(i));
    }
    std::string str() {
        ws();
        if (i >= s.size() || s[i] != '"') fail("expected string");
        i++;
        std::string out;
        while (i < s.size() && s[i] != '"') {
            if (s[i] == '\\' && i + 1 < s.size()) i++;
            out += s[i++];
        }
        if (i >= s.size()) fail("unterminated string");
        i++;
        return out;
    }
    void value(const std::string &key, std::map<std::string, std::string> &out) {
        ws();
        if (i >= s.size()) fail("unexpected end");
        char c = s[i];
        if (c == '{') {
            object(key, out);
        } else if (c == '"') {
            out[key] = str();
        } else if (c == '[') {
            int depth = 0;
            size_t st = i;
            do {
                if (s[i] == '[') depth++;
                if (s[i] == ']') depth--;
                i++;
            } while (i < s.size() && depth > 0);
            out[key] = s.substr(st, i - st);
        } else {
            size_t st = i;
            while (i < s.size() && s[i] != ',' && s[i] != '}' && !std::isspace((unsigned char) s[i])) i++;
            out[key] = s.substr(st, i - st);
        }
    }
    void object(const std::string &prefix, std::map<std::string, std::string> &out) {
        ws();
        if (s[i] != '{') fail("expected object");
        i++;
        ws();
        if (i < s.size() && s[i] == '}') { i++; return; }
        while (true) {
            std::string k = str();
            ws();
            if (i >= s.size() || s[i] != ':') fail("expected ':'");
            i++;
            value(prefix.empty() ? k : prefix + "." + k, out);
            ws();
            if (i < s.size() && s[i] == ',') { i++; continue; }
            if (i < s.size() && s[i] == '}') { i++; return; }
            fail("expected ',' or '}'");
        }
    }
};

std::string need(const std::map<std::string, std::string> &m, const std::string &k) {
    auto it = m.find(k);
    if (it == m.end()) throw std::runtime_error("data.json: missing key '" + k + "'");
    return it->second;
}

int needInt(const std::map<std::string, std::string> &m, const std::string &k) { return std::stoi(need(m, k)); }

std::vector<double> readFp(const std::string &dir, const std::string &name, bool required = true,
                           std::vector<uint64_t> *dimsOut = nullptr) {
    std::vector<uint64_t> dims;
    std::vector<double> out;
    if (readTensor(dir + name + "_d.bt", 'd', dims, out) || readTensor(dir + name + "_f.bt", 'f', dims, out)) {
        if (dimsOut) *dimsOut = dims;
        return out;
    }
    if (required) throw std::runtime_error("missing tensor file " + dir + name + "_{d,f}.bt");
    return out;
}

std::vector<int> readInts(const std::string &dir, const std::string &name) {
    std::vector<uint64_t> dims;
    std::vector<double> tmp;
    std::ifstream f(dir + name + "_u.bt", std::ios::binary);
    if (!f) throw std::runtime_error("missing tensor file " + dir + name + "_u.bt");
    uint64_t d[3];
    f.read(reinterpret_cast<char *>(d), 24);
    if (!f) throw std::runtime_error("truncated header in " + name);
    size_t n = d[0] * d[1] * d[2];
    std::vector<uint64_t> raw(n);
    f.read(reinterpret_cast<char *>(raw.data()), (std::streamsize) (n * 8));
    if ((size_t) f.gcount() != n * 8) throw std::runtime_error("truncated data in " + name);
    std::vector<int> out(n);
    for (size_t k = 0; k < n; k++) out[k] = (int) (int64_t) raw[k];
    return out;
}

void expectSize(const std::vector<double> &v, size_t n, const std::string &name) {
    if (v.size() != n)
        throw std::runtime_error("tensor " + name + " has " + std::to_string(v.size()) + " elements, expected " +
                                 std::to_string(n));
}

// replicate a single per-node vector/matrix `count` times then dedup (always one entry)
Table single(const std::vector<double> &m, int rows, int cols, int count) {
    Table t;
    t.rows = rows;
    t.cols = cols;
    t.data = m;
    t.index.assign(count, 0);
    return t;
}

// (L L')^-1 for every distinct lower Cholesky factor, in FP64: L^-1 by forward substitution,
// then L^-T L^-1.  Used by the dynamics sweep as one matvec instead of two triangular solves.
Table cholInverseTable(const Table &L) {
    Table t;
    t.rows = L.rows;
    t.cols = L.cols;
    t.index = L.index;
    const int n = L.rows;
    const int c = L.count();
    t.data.assign((size_t) c * n * n, 0.);
    std::vector<double> Li((size_t) n * n);
    for (int k = 0; k < c; k++) {
        const double *Lk = L.data.data() + (size_t) k * n * n;
        std::fill(Li.begin(), Li.end(), 0.);
        for (int col = 0; col < n; col++) {  // solve L x = e_col
            for (int i = 0; i < n; i++) {
                double v = (i == col) ? 1. : 0.;
                for (int j = 0; j < i; j++) v -= Lk[i + (size_t) j * n] * Li[j + (size_t) col * n];
                Li[i + (size_t) col * n] = v / Lk[i + (size_t) i * n];
            }
        }
        double *out = t.data.data() + (size_t) k * n * n;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                double v = 0.;
                for (int r = 0; r < n; r++) v += Li[r + (size_t) i * n] * Li[r + (size_t) j * n];
                out[i + (size_t) j * n] = v;
            }
    }
    return t;
}

Table transposeTable(const Table &in) {
    Table t;
    t.rows = in.cols;
    t.cols = in.rows;
    t.index = in.index;
    t.data.resize(in.data.size());
    int c = in.count();
    for (int k = 0; k < c; k++)
        for (int i = 0; i < in.rows; i++)
            for (int j = 0; j < in.cols; j++)
                t.data[(size_t) k * in.rows * in.cols + j + (size_t) i * in.cols] =
                    in.data[(size_t) k * in.rows * in.cols + i + (size_t) j * in.rows];
    return t;
}

}  // namespace

std::map<std::string, std::string> parseJsonFlat(const std::string &text) {
    std::map<std::string, std::string> out;
    JsonCursor c(text);
    c.object("", out);
    return out;
}

bool readTensor(const std::string &path, char kind, std::vector<uint64_t> &dims, std::vector<double> &out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    dims.assign(3, 0);
    f.read(reinterpret_cast<char *>(dims.data()), 24);
    if (!f) throw std::runtime_error("truncated header in " + path);
    size_t n = dims[0] * dims[1] * dims[2];
    out.resize(n);
    if (kind == 'd') {
        f.read(reinterpret_cast<char *>(out.data()), (std::streamsize) (n * 8));
        if ((size_t) f.gcount() != n * 8) throw std::runtime_error("truncated data in " + path);
    } else if (kind == 'f') {
        std::vector<float> tmp(n);
        f.read(reinterpret_cast<char *>(tmp.data()), (std::streamsize) (n * 4));
        if ((size_t) f.gcount() != n * 4) throw std::runtime_error("truncated data in " + path);
        for (size_t k = 0; k < n; k++) out[k] = tmp[k];
    } else {
        throw std::runtime_error("unsupported tensor kind");
    }
    return true;
}

Table dedup(const std::vector<double> &stack, int rows, int cols, int count) {
    Table t;
    t.rows = rows;
    t.cols = cols;
    t.index.resize(count);
    const size_t sz = (size_t) rows * cols;
    std::unordered_map<std::string, int> seen;
    for (int k = 0; k < count; k++) {
        std::string key(reinterpret_cast<const char *>(stack.data() + k * sz), sz * sizeof(double));
        auto it = seen.find(key);
        if (it != seen.end()) {
            t.index[k] = it->second;
        } else {
            int id = (int) (t.data.size() / (sz ? sz : 1));
            seen.emplace(std::move(key), id);
            t.data.insert(t.data.end(), stack.begin() + k * sz, stack.begin() + (k + 1) * sz);
            t.index[k] = id;
        }
    }
    return t;
}

// PolyhedronWithIdentity (src/constraints.cuh:299-435): ilb <= z <= iub and glb <= Gamma z <= gub
// on the same vector z (dimension nz), for every node of its part.  The engines run it as the
// polyhedron whose matrix is [I; Gamma] (column-major, rows = nz + rows of Gamma) and whose bounds
// are [ilb; glb], [iub; gub]: the same operator, adjoint and projection, with each node's dual
// block ordered [identity part; Gamma part].  The reference orders the whole dual block as the
// identity parts of every node, then the Gamma parts; iterates cross the host boundary in that
// order (reorderIdentityBlocks).  No bound is relaxed at the root (unlike Rectangle's,
// src/constraints.cuh:142-148).
static void stackIdentity(const std::string &dir, const std::string &prefix, const std::vector<double> &G, int nz,
                   int nodes, int &rows, Table &M, Table &LB, Table &UB) {
    auto ilb = readFp(dir, prefix + "ILB"), iub = readFp(dir, prefix + "IUB");
    auto glb = readFp(dir, prefix + "GLB"), gub = readFp(dir, prefix + "GUB");
    expectSize(ilb, (size_t) nz, prefix + "ILB");
    expectSize(iub, (size_t) nz, prefix + "IUB");
    const int pg = (int) glb.size();
    expectSize(gub, (size_t) pg, prefix + "GUB");
    expectSize(G, (size_t) pg * nz, prefix + "Gamma");
    rows = nz + pg;
    std::vector<double> m((size_t) rows * nz, 0.);
    for (int c = 0; c < nz; c++) {
        m[(size_t) c * rows + c] = 1.;
        for (int r = 0; r < pg; r++) m[(size_t) c * rows + nz + r] = G[(size_t) c * pg + r];
    }
    std::vector<double> lb = ilb, ub = iub;
    lb.insert(lb.end(), glb.begin(), glb.end());
    ub.insert(ub.end(), gub.begin(), gub.end());
    M = single(m, rows, nz, nodes);
    LB = single(lb, rows, 1, nodes);
    UB = single(ub, rows, 1, nodes);
}

void reorderIdentityBlocks(const HostProblem &h, double *z, bool toEngine) {
    auto one = [&](long off, int rows, int id, int nodes) {
        if (!id) return;
        const std::vector<double> tmp(z + off, z + off + (size_t) rows * nodes);
        const int pg = rows - id;
        for (int i = 0; i < nodes; i++)
            for (int k = 0; k < rows; k++) {
                // engine position i * rows + k <-> reference position: the identity part of node i
                // (k < id), else row k - id of its Gamma part
                const size_t eng = (size_t) i * rows + k;
                const size_t ref = k < id ? (size_t) i * id + k : (size_t) nodes * id + (size_t) i * pg + (k - id);
                if (toEngine) z[off + eng] = tmp[ref];
                else z[off + ref] = tmp[eng];
            }
    };
    one(h.sPrim + h.sI + h.sII, h.con3, h.id3, h.nl);
    one(h.sPrim + h.sI + h.sII + h.sIII + h.sIV, h.con5, h.id5, h.nlf);
}

HostProblem loadProblem(const std::string &folderIn) {
    std::string dir = folderIn;
    if (!dir.empty() && dir.back() != '/') dir += '/';
    std::ifstream jf(dir + "data.json");
    if (!jf) throw std::runtime_error("cannot open " + dir + "data.json");
    std::stringstream ss;
    ss << jf.rdbuf();
    auto js = parseJsonFlat(ss.str());
    HostProblem p;
    p.E = needInt(js, "numEvents");
    p.nl = needInt(js, "numNonleafNodes");
    p.N = needInt(js, "numNodes");
    p.S = needInt(js, "numStages");
    p.nx = needInt(js, "numStates");
    p.nu = needInt(js, "numInputs");
    p.nlf = p.N - p.nl;
    p.yD = 2 * p.E + 1;
    p.nD = needInt(js, "rowsNNtr");
    if (p.nD != 4 * p.E + 1) throw std::runtime_error("data.json: rowsNNtr must equal 4*numEvents+1");
    p.stepSize = std::stod(need(js, "stepSize"));
    p.precond = need(js, "preconditioned") == "true";
    std::string dyn = need(js, "dynamics.type");
    if (dyn != "linear" && dyn != "affine")
        throw std::runtime_error("dynamics type '" + dyn + "' is not supported (linear, affine)");
    p.affine = dyn == "affine";
    auto costKind = [](const std::string &s) {
        if (s == "quadratic") return (int) COST_QUAD;
        if (s == "quadraticPlusLinear") return (int) COST_QUADLIN;
        if (s == "linear") return (int) COST_LIN;
        throw std::runtime_error("cost type '" + s + "' is not supported (quadratic, linear, quadraticPlusLinear)");
    };
    auto conKind = [](const std::string &s) {
        if (s == "no") return (int) CON_NO;
        if (s == "rectangle") return (int) CON_RECT;
        if (s == "polyhedron") return (int) CON_POLY;
        if (s == "polyhedronWithIdentity") return (int) CON_PWI;
        throw std::runtime_error("constraint type '" + s +
                                 "' is not supported (no, rectangle, polyhedron, polyhedronWithIdentity)");
    };
    p.ncost = costKind(need(js, "cost.nonleaf"));
    p.lcost = costKind(need(js, "cost.leaf"));
    p.ncon = conKind(need(js, "constraint.nonleaf"));
    p.lcon = conKind(need(js, "constraint.leaf"));
    // polyhedronWithIdentity runs as the polyhedron with matrix [I; Gamma] (see stackIdentity)
    const bool pwiN = p.ncon == CON_PWI, pwiL = p.lcon == CON_PWI;
    if (pwiN) p.ncon = CON_POLY;
    if (pwiL) p.lcon = CON_POLY;
    if (need(js, "risk.type") != "avar") throw std::runtime_error("risk type must be avar");
    auto al = js.find("risk.alpha");
    if (al != js.end()) p.alpha = std::stod(al->second);

    const int N = p.N, nl = p.nl, nlf = p.nlf, nx = p.nx, nu = p.nu;
    // tree (src/tree.cuh:124-178)
    p.anc = readInts(dir, "ancestors");
    p.events = readInts(dir, "events");
    p.chFrom = readInts(dir, "childrenFrom");
    p.numCh = readInts(dir, "numChildren");
    p.stageFrom = readInts(dir, "stageFrom");
    p.stageTo = readInts(dir, "stageTo");
    if ((int) p.anc.size() != N || (int) p.chFrom.size() != nl || (int) p.numCh.size() != nl ||
        (int) p.stageFrom.size() != p.S || (int) p.stageTo.size() != p.S)
        throw std::runtime_error("tree files do not match data.json sizes");
    for (int j = 1; j < N; j++)
        if (p.anc[j] < 0 || p.anc[j] >= nl) throw std::runtime_error("ancestor index out of range");
    for (int i = 0; i < nl; i++)
        if (p.numCh[i] < 1 || p.numCh[i] > p.E || p.chFrom[i] < 1 || p.chFrom[i] + p.numCh[i] > N)
            throw std::runtime_error("children range out of bounds");
    for (int s = 0; s < p.S; s++)
        if (p.stageFrom[s] < 0 || p.stageTo[s] >= N || p.stageFrom[s] > p.stageTo[s])
            throw std::runtime_error("stage range out of bounds");
    p.prob = readFp(dir, "probabilities", false);
    p.condProb = readFp(dir, "conditionalProbabilities", false);
    if (p.precond) {
        p.scaling = readFp(dir, "scaling");
        expectSize(p.scaling, (size_t) nx + nu, "scaling");
    }
    // dynamics (src/dynamics.cuh:40-72)
    auto stackTable = [&](const std::string &name, int rows, int cols, int count) {
        auto v = readFp(dir, name);
        expectSize(v, (size_t) rows * cols * count, name);
        return dedup(v, rows, cols, count);
    };
    p.BTr = stackTable("dynamics_BTr", nu, nx, N);
    p.AB = stackTable("dynamics_AB", nx, nx + nu, N);
    p.K = stackTable("dynamics_K", nu, nx, nl);
    p.KTr = transposeTable(p.K);
    p.ABKTr = stackTable("dynamics_ABKTr", nx, nx, N);
    p.APB = stackTable("dynamics_APB", nx, nu, N);
    p.cholL = stackTable("dynamics_lowCholesky", nu, nu, nl);
    p.cholU = transposeTable(p.cholL);
    p.RtInv = cholInverseTable(p.cholL);
    if (p.affine) {  // src/dynamics.cuh:229-237: e from dynamics_c, P e formed once at load
        auto e = readFp(dir, "dynamics_c");
        expectSize(e, (size_t) nx * N, "dynamics_c");
        auto P = readFp(dir, "dynamics_P");
        expectSize(P, (size_t) nx * nx * N, "dynamics_P");
        std::vector<double> pe((size_t) nx * N, 0.);
        for (int n = 0; n < N; n++)
            for (int c = 0; c < nx; c++)
                for (int r = 0; r < nx; r++)
                    pe[(size_t) n * nx + r] += P[((size_t) n * nx + c) * nx + r] * e[(size_t) n * nx + c];
        p.e = dedup(e, nx, 1, N);
        p.Pe = dedup(pe, nx, 1, N);
    } else {
        p.e = single(std::vector<double>(nx, 0.), nx, 1, N);
        p.Pe = single(std::vector<double>(nx, 0.), nx, 1, N);
    }
    // risk (src/risks.cuh:26-33)
    p.NN = stackTable("risk_NNtr", p.nD, p.nD, nl);
    p.b = stackTable("risk_b", p.yD, 1, nl);
    // costs (src/costs.cuh:85-110, 331-368): nonleaf per node, leaf single matrix replicated.  A
    // linear cost (src/costs.cuh:228-250) reads only its gradient [q' r'] per node; its q and r
    // parts fill the q and r tables and the quadratic tables are zero.
    if (p.ncost == COST_LIN) {
        auto g = readFp(dir, "nonleafCost_gradient");
        expectSize(g, (size_t) (nx + nu) * N, "nonleafCost_gradient");
        std::vector<double> qv((size_t) nx * N), rv((size_t) nu * N);
        for (int n = 0; n < N; n++) {
            for (int k = 0; k < nx; k++) qv[(size_t) n * nx + k] = g[(size_t) n * (nx + nu) + k];
            for (int k = 0; k < nu; k++) rv[(size_t) n * nu + k] = g[(size_t) n * (nx + nu) + nx + k];
        }
        p.q = dedup(qv, nx, 1, N);
        p.r = dedup(rv, nu, 1, N);
        p.sQ = single(std::vector<double>((size_t) nx * nx, 0.), nx, nx, N);
        p.sR = single(std::vector<double>((size_t) nu * nu, 0.), nu, nu, N);
        p.trans = single(std::vector<double>((size_t) nx + nu + 2, 0.), nx + nu + 2, 1, N);
    } else {
        p.sQ = stackTable("nonleafCost_sqrtQ", nx, nx, N);
        p.sR = stackTable("nonleafCost_sqrtR", nu, nu, N);
        p.trans = stackTable("nonleafCost_translation", nx + nu + 2, 1, N);
    }
    if (p.ncost == COST_QUADLIN) {
        p.q = stackTable("nonleafCost_q", nx, 1, N);
        p.r = stackTable("nonleafCost_r", nu, 1, N);
    } else if (p.ncost == COST_QUAD) {
        p.q = single(std::vector<double>(nx, 0.), nx, 1, N);
        p.r = single(std::vector<double>(nu, 0.), nu, 1, N);
    }
    if (p.lcost == COST_LIN) {
        auto g = readFp(dir, "leafCost_gradient");
        expectSize(g, (size_t) nx, "leafCost_gradient");
        p.lq = single(g, nx, 1, nlf);
        p.lsQ = single(std::vector<double>((size_t) nx * nx, 0.), nx, nx, nlf);
        p.ltrans = single(std::vector<double>((size_t) nx + 2, 0.), nx + 2, 1, nlf);
    } else {
        auto m = readFp(dir, "leafCost_sqrtQ");
        expectSize(m, (size_t) nx * nx, "leafCost_sqrtQ");
        p.lsQ = single(m, nx, nx, nlf);
        auto tr = readFp(dir, "leafCost_translation");
        expectSize(tr, (size_t) nx + 2, "leafCost_translation");
        p.ltrans = single(tr, nx + 2, 1, nlf);
        if (p.lcost == COST_QUADLIN) {
            auto lq = readFp(dir, "leafCost_q");
            expectSize(lq, (size_t) nx, "leafCost_q");
            p.lq = single(lq, nx, 1, nlf);
        } else {
            p.lq = single(std::vector<double>(nx, 0.), nx, 1, nlf);
        }
    }
    // constraints (src/constraints.cuh:120-150, 203-245)
    if (pwiN) {
        auto G = readFp(dir, "nonleafConstraintGamma");
        p.id3 = nx + nu;
        stackIdentity(dir, "nonleafConstraint", G, p.id3, nl, p.con3, p.nG, p.nLB, p.nUB);
        p.nGT = transposeTable(p.nG);
    } else if (p.ncon == CON_RECT) {
        auto lb = readFp(dir, "nonleafConstraintILB"), ub = readFp(dir, "nonleafConstraintIUB");
        expectSize(lb, (size_t) nx + nu, "nonleafConstraintILB");
        expectSize(ub, (size_t) nx + nu, "nonleafConstraintIUB");
        p.con3 = nx + nu;
        // node 0 has its state bounds relaxed to +-inf (src/constraints.cuh:142-148)
        std::vector<double> lb0 = lb, ub0 = ub;
        for (int k = 0; k < nx; k++) { lb0[k] = -INFINITY; ub0[k] = INFINITY; }
        p.nLB.rows = p.nUB.rows = p.con3;
        p.nLB.cols = p.nUB.cols = 1;
        p.nLB.data = lb0; p.nLB.data.insert(p.nLB.data.end(), lb.begin(), lb.end());
        p.nUB.data = ub0; p.nUB.data.insert(p.nUB.data.end(), ub.begin(), ub.end());
        p.nLB.index.assign(nl, 1); p.nLB.index[0] = 0;
        p.nUB.index = p.nLB.index;
    } else if (p.ncon == CON_POLY) {
        std::vector<uint64_t> gd;
        auto G = readFp(dir, "nonleafConstraintGamma", true, &gd);
        auto lb = readFp(dir, "nonleafConstraintGLB"), ub = readFp(dir, "nonleafConstraintGUB");
        p.con3 = (int) lb.size();
        expectSize(G, (size_t) p.con3 * (nx + nu), "nonleafConstraintGamma");
        expectSize(ub, (size_t) p.con3, "nonleafConstraintGUB");
        p.nG = single(G, p.con3, nx + nu, nl);
        p.nGT = transposeTable(p.nG);
        p.nLB = single(lb, p.con3, 1, nl);
        p.nUB = single(ub, p.con3, 1, nl);
    }
    if (pwiL) {
        auto G = readFp(dir, "leafConstraintGamma");
        p.id5 = nx;
        stackIdentity(dir, "leafConstraint", G, p.id5, nlf, p.con5, p.lG, p.lLB, p.lUB);
        p.lGT = transposeTable(p.lG);
    } else if (p.lcon == CON_RECT) {
        auto lb = readFp(dir, "leafConstraintILB"), ub = readFp(dir, "leafConstraintIUB");
        expectSize(lb, (size_t) nx, "leafConstraintILB");
        expectSize(ub, (size_t) nx, "leafConstraintIUB");
        p.con5 = nx;
        p.lLB = single(lb, nx, 1, nlf);
        p.lUB = single(ub, nx, 1, nlf);
    } else if (p.lcon == CON_POLY) {
        auto G = readFp(dir, "leafConstraintGamma");
        auto lb = readFp(dir, "leafConstraintGLB"), ub = readFp(dir, "leafConstraintGUB");
        p.con5 = (int) lb.size();
        expectSize(G, (size_t) p.con5 * nx, "leafConstraintGamma");
        expectSize(ub, (size_t) p.con5, "leafConstraintGUB");
        p.lG = single(G, p.con5, nx, nlf);
        p.lGT = transposeTable(p.lG);
        p.lLB = single(lb, p.con5, 1, nlf);
        p.lUB = single(ub, p.con5, 1, nlf);
    }
    // iterate layout (src/cache.cuh:482-581)
    p.sU = (long) nu * nl; p.sX = (long) nx * N; p.sY = (long) p.yD * nl; p.sT = N; p.sS = N;
    p.sPrim = p.sU + p.sX + p.sY + p.sT + p.sS;
    // a linear cost's block is (cost of every node; epigraph variable of every node)
    p.sI = (long) p.yD * nl; p.sII = nl; p.sIII = (long) p.con3 * nl;
    p.sIV = (long) (p.ncost == COST_LIN ? 2 : nx + nu + 2) * N;
    p.sV = (long) p.con5 * nlf; p.sVI = (long) (p.lcost == COST_LIN ? 2 : nx + 2) * nlf;
    p.sDual = p.sI + p.sII + p.sIII + p.sIV + p.sV + p.sVI;
    p.sIter = p.sPrim + p.sDual;
    return p;
}

}  // namespace spk
