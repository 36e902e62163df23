// FP64 CPU restatement of the reference SPOCK iteration.  TEST INFRASTRUCTURE ONLY:
// it is the checker for the HIP solver (tests, smoke(), bench.py's cpu_baseline),
// never part of the product path.
//
// Each function cites the reference file:line it restates.  Node data are column
// major per node and node-major across nodes (src/definitions.cu:11-14).  Operator
// outputs that the reference leaves unwritten (t[0] of L*, the IV block of node 0
// of L) are defined as zero here; from a zero start the reference keeps them at
// zero too (src/costs.cuh:136,179-183), so iterates agree.  The polyhedron-with-identity
// constraint's L* is the plain adjoint.  The reference's also adds G' (G part) into the
// identity part of its input in place (src/constraints.cuh:418-433), but that input is always
// the work iterate's dual, which the next L rewrites before anything reads it
// (src/cache.cuh:725-730, 748-751, 916-920), so iterates agree.
//
// Parity pins for each function live in tests/test_oracle.py (dense operator
// matrix, KKT solve, closed-form projectors, textbook cone projections, lstsq); the
// whole iteration is pinned there by an independent solve of the same conic program
// (scipy SLSQP on the generator's own operator) reaching the oracle's optimal cost.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <string>
#include <vector>

namespace {

using Vec = std::vector<double>;

// ----------------------------------------------------------------- file reading
// .bt layout: 3 x uint64 dims then data (tests/julia/modelFactory/src/modelFactory.jl:29-79)
bool readBt(const std::string &path, std::vector<uint64_t> &dims, std::vector<char> &raw, size_t elSize) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    dims.assign(3, 0);
    f.read(reinterpret_cast<char *>(dims.data()), 24);
    size_t n = dims[0] * dims[1] * dims[2];
    raw.resize(n * elSize);
    f.read(raw.data(), (std::streamsize) raw.size());
    return (bool) f || f.gcount() == (std::streamsize) raw.size();
}

Vec readFp(const std::string &dir, const std::string &name, bool required = true) {
    std::vector<uint64_t> dims;
    std::vector<char> raw;
    Vec out;
    if (readBt(dir + name + "_d.bt", dims, raw, 8)) {
        out.resize(raw.size() / 8);
        std::memcpy(out.data(), raw.data(), raw.size());
    } else if (readBt(dir + name + "_f.bt", dims, raw, 4)) {
        out.resize(raw.size() / 4);
        for (size_t i = 0; i < out.size(); i++) {
            float v;
            std::memcpy(&v, raw.data() + 4 * i, 4);
            out[i] = v;
        }
    } else if (required) {
        throw std::runtime_error("missing tensor " + name);
    }
    return out;
}

std::vector<long> readU(const std::string &dir, const std::string &name) {
    std::vector<uint64_t> dims;
    std::vector<char> raw;
    if (!readBt(dir + name + "_u.bt", dims, raw, 8)) throw std::runtime_error("missing tensor " + name);
    std::vector<long> out(raw.size() / 8);
    for (size_t i = 0; i < out.size(); i++) {
        uint64_t v;
        std::memcpy(&v, raw.data() + 8 * i, 8);
        out[i] = (long) (int64_t) v;
    }
    return out;
}

// y (+)= M x, M rows x cols column major
inline void gemv(const double *M, int rows, int cols, const double *x, double *y, bool acc) {
    for (int r = 0; r < rows; r++) {
        double s = 0.;
        for (int c = 0; c < cols; c++) s += M[r + (size_t) c * rows] * x[c];
        y[r] = acc ? y[r] + s : s;
    }
}

// y (+)= M' x
inline void gemvT(const double *M, int rows, int cols, const double *x, double *y, bool acc) {
    for (int c = 0; c < cols; c++) {
        double s = 0.;
        for (int r = 0; r < rows; r++) s += M[r + (size_t) c * rows] * x[r];
        y[c] = acc ? y[c] + s : s;
    }
}

enum { CON_NO = 0, CON_RECT = 1, CON_POLY = 2, CON_PWI = 3 };
enum { COST_QUAD = 0, COST_QUADLIN = 1, COST_LIN = 2 };

// ------------------------------------------------------------------- problem
struct Problem {
    int N, nl, nlf, S, E, nx, nu, yD, nD;
    int ncost, lcost, ncon, lcon;
    int con3, con5;  // per-node sizes of dual blocks III and V
    int id3 = 0, id5 = 0;  // polyhedron with identity: size of the identity part per node
    bool precond;
    double gamma, gammaRecip;
    std::vector<long> anc, chFrom, numCh, stageFrom, stageTo;
    Vec BTr, AB, K, ABKTr, APB, lowChol, NNtr, b, scaling;
    bool affine = false;
    Vec e, Pe;  // affine dynamics: constant e_j and P_j e_j per node (src/dynamics.cuh:229-237)
    Vec sqrtQ, sqrtR, trans, q, r, lsqrtQ, ltrans, lq;
    Vec grad, lgrad;  // linear costs: the gradient row [q' r'] per node (nonleaf), q' (leaf)
    int d4 = 0, d6 = 0;  // entries per node of dual blocks IV and VI
    Vec nLB, nUB, nG, lLB, lUB, lG;  // per-node bounds (replicated) and matrices
    size_t sU, sX, sY, sT, sS, sPrim, sI, sII, sIII, sIV, sV, sVI, sDual, sIter;
    size_t oU, oX, oY, oT, oS;                  // offsets inside prim
    size_t oI, oII, oIII, oIV, oV, oVI;         // offsets inside dual

    // src/tree.cuh:102-179, src/problem.cuh:126-153 and the component constructors
    void load(const std::string &dirIn, const long *ints, double stepSize) {
        std::string dir = dirIn;
        if (!dir.empty() && dir.back() != '/') dir += '/';
        N = (int) ints[0]; nl = (int) ints[1]; S = (int) ints[2]; E = (int) ints[3];
        nx = (int) ints[4]; nu = (int) ints[5];
        ncost = (int) ints[6]; lcost = (int) ints[7]; ncon = (int) ints[8]; lcon = (int) ints[9];
        precond = ints[10] != 0;
        nlf = N - nl;
        yD = 2 * E + 1;
        nD = 4 * E + 1;
        gamma = stepSize;
        gammaRecip = 1. / stepSize;
        anc = readU(dir, "ancestors");
        chFrom = readU(dir, "childrenFrom");
        numCh = readU(dir, "numChildren");
        stageFrom = readU(dir, "stageFrom");
        stageTo = readU(dir, "stageTo");
        BTr = readFp(dir, "dynamics_BTr");
        AB = readFp(dir, "dynamics_AB");
        K = readFp(dir, "dynamics_K");
        ABKTr = readFp(dir, "dynamics_ABKTr");
        APB = readFp(dir, "dynamics_APB");
        lowChol = readFp(dir, "dynamics_lowCholesky");
        affine = ints[11] != 0;
        if (affine) {  // src/dynamics.cuh:232-236: e from dynamics_c, Pe = P e per node
            e = readFp(dir, "dynamics_c");
            const Vec P = readFp(dir, "dynamics_P");
            Pe.assign((size_t) nx * N, 0.);
            for (int n = 0; n < N; n++)
                gemv(P.data() + (size_t) n * nx * nx, nx, nx, e.data() + (size_t) n * nx, Pe.data() + (size_t) n * nx,
                     false);
        }
        NNtr = readFp(dir, "risk_NNtr");
        b = readFp(dir, "risk_b");
        if (precond) scaling = readFp(dir, "scaling");
        // src/costs.cuh:228-250: a linear cost reads only its gradient
        if (ncost == COST_LIN) grad = readFp(dir, "nonleafCost_gradient");
        else {
            sqrtQ = readFp(dir, "nonleafCost_sqrtQ");
            sqrtR = readFp(dir, "nonleafCost_sqrtR");
            trans = readFp(dir, "nonleafCost_translation");
        }
        if (lcost == COST_LIN) lgrad = readFp(dir, "leafCost_gradient");
        else {
            lsqrtQ = readFp(dir, "leafCost_sqrtQ");
            ltrans = readFp(dir, "leafCost_translation");
        }
        if (ncost == COST_QUADLIN) { q = readFp(dir, "nonleafCost_q"); r = readFp(dir, "nonleafCost_r"); }
        else { q.assign((size_t) nx * N, 0.); r.assign((size_t) nu * N, 0.); }
        if (lcost == COST_QUADLIN) lq = readFp(dir, "leafCost_q");
        else lq.assign(nx, 0.);
        // Constraints (src/constraints.cuh:120-191, 203-286); rectangle root-state bounds are
        // relaxed to +-inf (src/constraints.cuh:142-148).
        int nxu = nx + nu;
        con3 = 0;
        if (ncon == CON_RECT) {
            Vec lb = readFp(dir, "nonleafConstraintILB"), ub = readFp(dir, "nonleafConstraintIUB");
            con3 = nxu;
            nLB.resize((size_t) con3 * nl); nUB.resize((size_t) con3 * nl);
            for (int i = 0; i < nl; i++)
                for (int k = 0; k < con3; k++) {
                    nLB[(size_t) i * con3 + k] = lb[k];
                    nUB[(size_t) i * con3 + k] = ub[k];
                }
            for (int k = 0; k < nx; k++) { nLB[k] = -INFINITY; nUB[k] = INFINITY; }
        } else if (ncon == CON_POLY) {
            nG = readFp(dir, "nonleafConstraintGamma");
            Vec lb = readFp(dir, "nonleafConstraintGLB"), ub = readFp(dir, "nonleafConstraintGUB");
            con3 = (int) lb.size();
            nLB.resize((size_t) con3 * nl); nUB.resize((size_t) con3 * nl);
            for (int i = 0; i < nl; i++)
                for (int k = 0; k < con3; k++) {
                    nLB[(size_t) i * con3 + k] = lb[k];
                    nUB[(size_t) i * con3 + k] = ub[k];
                }
        } else if (ncon == CON_PWI) {
            // src/constraints.cuh:316-357: bounds [ilb of every node; glb of every node], the
            // root state bounds kept
            nG = readFp(dir, "nonleafConstraintGamma");
            splitBounds(readFp(dir, "nonleafConstraintILB"), readFp(dir, "nonleafConstraintIUB"),
                        readFp(dir, "nonleafConstraintGLB"), readFp(dir, "nonleafConstraintGUB"), nl, nxu, id3, con3,
                        nLB, nUB);
        }
        con5 = 0;
        if (lcon == CON_RECT) {
            Vec lb = readFp(dir, "leafConstraintILB"), ub = readFp(dir, "leafConstraintIUB");
            con5 = nx;
            lLB.resize((size_t) con5 * nlf); lUB.resize((size_t) con5 * nlf);
            for (int i = 0; i < nlf; i++)
                for (int k = 0; k < con5; k++) {
                    lLB[(size_t) i * con5 + k] = lb[k];
                    lUB[(size_t) i * con5 + k] = ub[k];
                }
        } else if (lcon == CON_POLY) {
            lG = readFp(dir, "leafConstraintGamma");
            Vec lb = readFp(dir, "leafConstraintGLB"), ub = readFp(dir, "leafConstraintGUB");
            con5 = (int) lb.size();
            lLB.resize((size_t) con5 * nlf); lUB.resize((size_t) con5 * nlf);
            for (int i = 0; i < nlf; i++)
                for (int k = 0; k < con5; k++) {
                    lLB[(size_t) i * con5 + k] = lb[k];
                    lUB[(size_t) i * con5 + k] = ub[k];
                }
        } else if (lcon == CON_PWI) {
            lG = readFp(dir, "leafConstraintGamma");
            splitBounds(readFp(dir, "leafConstraintILB"), readFp(dir, "leafConstraintIUB"),
                        readFp(dir, "leafConstraintGLB"), readFp(dir, "leafConstraintGUB"), nlf, nx, id5, con5,
                        lLB, lUB);
        }
        // Iterate layout (src/cache.cuh:482-581)
        sU = (size_t) nu * nl; sX = (size_t) nx * N; sY = (size_t) yD * nl; sT = N; sS = N;
        sPrim = sU + sX + sY + sT + sS;
        // a linear cost's block is [cost of every node; epigraph variable of every node]
        // (src/costs.cuh:232, 248, 254-258)
        d4 = ncost == COST_LIN ? 2 : nxu + 2;
        d6 = lcost == COST_LIN ? 2 : nx + 2;
        sI = (size_t) yD * nl; sII = nl; sIII = (size_t) con3 * nl; sIV = (size_t) d4 * N;
        sV = (size_t) con5 * nlf; sVI = (size_t) d6 * nlf;
        sDual = sI + sII + sIII + sIV + sV + sVI;
        sIter = sPrim + sDual;
        oU = 0; oX = oU + sU; oY = oX + sX; oT = oY + sY; oS = oT + sT;
        oI = 0; oII = oI + sI; oIII = oII + sII; oIV = oIII + sIII; oV = oIV + sIV; oVI = oV + sV;
    }

    // Polyhedron with identity over n nodes: the identity parts of all nodes come first, then
    // the G parts (src/constraints.cuh:340-373); bounds laid out the same way.
    static void splitBounds(const Vec &ilb, const Vec &iub, const Vec &glb, const Vec &gub, int n, int width,
                            int &idPart, int &perNode, Vec &lo, Vec &hi) {
        if ((int) ilb.size() != width || iub.size() != ilb.size() || gub.size() != glb.size())
            throw std::runtime_error("polyhedron with identity: bound sizes do not match");
        idPart = width;
        perNode = width + (int) glb.size();
        lo.resize((size_t) perNode * n);
        hi.resize((size_t) perNode * n);
        const size_t g0 = (size_t) width * n, rows = glb.size();
        for (int i = 0; i < n; i++) {
            for (int k = 0; k < width; k++) {
                lo[(size_t) i * width + k] = ilb[k];
                hi[(size_t) i * width + k] = iub[k];
            }
            for (size_t k = 0; k < rows; k++) {
                lo[g0 + (size_t) i * rows + k] = glb[k];
                hi[g0 + (size_t) i * rows + k] = gub[k];
            }
        }
    }
};

// ------------------------------------------------------------------ operator L
// src/operator.cuh:41-60 with src/costs.cuh:126-160 (quadratic), :379-416 (quadratic+linear),
// src/constraints.cuh:158-169 (rectangle), :251-263 (polyhedron)
void opL(const Problem &p, const double *prim, double *dual) {
    const int nx = p.nx, nu = p.nu, nxu = nx + nu, yD = p.yD;
    const double *u = prim + p.oU, *x = prim + p.oX, *y = prim + p.oY, *t = prim + p.oT, *s = prim + p.oS;
    double *I = dual + p.oI, *II = dual + p.oII, *III = dual + p.oIII, *IV = dual + p.oIV, *V = dual + p.oV,
        *VI = dual + p.oVI;
    for (int i = 0; i < p.nl; i++) {
        const double *yi = y + (size_t) i * yD;
        for (int k = 0; k < yD; k++) I[(size_t) i * yD + k] = yi[k];
        double by = 0.;
        for (int k = 0; k < yD; k++) by += p.b[(size_t) i * yD + k] * yi[k];
        II[i] = s[i] - by;
        double *iii = III + (size_t) i * p.con3;
        if (p.ncon == CON_RECT) {
            for (int k = 0; k < nx; k++) iii[k] = x[(size_t) i * nx + k];
            for (int k = 0; k < nu; k++) iii[nx + k] = u[(size_t) i * nu + k];
        } else if (p.ncon == CON_POLY) {
            double xu[1024];
            for (int k = 0; k < nx; k++) xu[k] = x[(size_t) i * nx + k];
            for (int k = 0; k < nu; k++) xu[nx + k] = u[(size_t) i * nu + k];
            gemv(p.nG.data(), p.con3, nxu, xu, iii, false);
        } else if (p.ncon == CON_PWI) {  // src/constraints.cuh:386-390
            const int rows = p.con3 - p.id3;
            double *idPart = III + (size_t) i * p.id3, *gPart = III + (size_t) p.nl * p.id3 + (size_t) i * rows;
            for (int k = 0; k < nx; k++) idPart[k] = x[(size_t) i * nx + k];
            for (int k = 0; k < nu; k++) idPart[nx + k] = u[(size_t) i * nu + k];
            gemv(p.nG.data(), rows, nxu, idPart, gPart, false);
        }
    }
    if (p.ncost == COST_LIN) {  // src/costs.cuh:261-267: sum_j = g_j [x_anc; u_anc] (node 0: 0), set = t
        IV[0] = 0.;
        for (int j = 1; j < p.N; j++) {
            const long a = p.anc[j];
            const double *g = p.grad.data() + (size_t) j * nxu;
            double acc = 0.;
            for (int k = 0; k < nx; k++) acc += g[k] * x[(size_t) a * nx + k];
            for (int k = 0; k < nu; k++) acc += g[nx + k] * u[(size_t) a * nu + k];
            IV[j] = acc;
        }
        for (int j = 0; j < p.N; j++) IV[p.N + j] = t[j];
    } else {
        const int d4 = nxu + 2;
        for (int k = 0; k < d4; k++) IV[k] = 0.;
        for (int j = 1; j < p.N; j++) {
            const long a = p.anc[j];
            double *iv = IV + (size_t) j * d4;
            const double *xa = x + (size_t) a * nx, *ua = u + (size_t) a * nu;
            gemv(p.sqrtQ.data() + (size_t) j * nx * nx, nx, nx, xa, iv, false);
            gemv(p.sqrtR.data() + (size_t) j * nu * nu, nu, nu, ua, iv + nx, false);
            double half = t[j] * 0.5;
            if (p.ncost == COST_QUADLIN) {
                double qx = 0., ru = 0.;
                for (int k = 0; k < nx; k++) qx += p.q[(size_t) j * nx + k] * xa[k];
                for (int k = 0; k < nu; k++) ru += p.r[(size_t) j * nu + k] * ua[k];
                half = half - 0.5 * qx - 0.5 * ru;
            }
            iv[nxu] = half;
            iv[nxu + 1] = half;
        }
    }
    for (int l = 0; l < p.nlf; l++) {
        const double *xl = x + (size_t) (p.nl + l) * nx;
        double *v = V + (size_t) l * p.con5;
        if (p.lcon == CON_RECT) for (int k = 0; k < nx; k++) v[k] = xl[k];
        else if (p.lcon == CON_POLY) gemv(p.lG.data(), p.con5, nx, xl, v, false);
        else if (p.lcon == CON_PWI) {  // src/constraints.cuh:398-401
            const int rows = p.con5 - p.id5;
            double *idPart = V + (size_t) l * p.id5, *gPart = V + (size_t) p.nlf * p.id5 + (size_t) l * rows;
            for (int k = 0; k < nx; k++) idPart[k] = xl[k];
            gemv(p.lG.data(), rows, nx, xl, gPart, false);
        }
        if (p.lcost == COST_LIN) {  // src/costs.cuh:269-272
            double acc = 0.;
            for (int k = 0; k < nx; k++) acc += p.lgrad[k] * xl[k];
            VI[l] = acc;
            VI[p.nlf + l] = s[p.nl + l];
            continue;
        }
        double *vi = VI + (size_t) l * (nx + 2);
        gemv(p.lsqrtQ.data(), nx, nx, xl, vi, false);
        double half = s[p.nl + l] * 0.5;
        if (p.lcost == COST_QUADLIN) {
            double qx = 0.;
            for (int k = 0; k < nx; k++) qx += p.lq[k] * xl[k];
            half = half - 0.5 * qx;
        }
        vi[nx] = half;
        vi[nx + 1] = half;
    }
}

// ------------------------------------------------------------------ operator L*
// src/operator.cuh:63-82 with src/costs.cuh:162-195, :418-458 and src/constraints.cuh:171-191, :265-286
void opLadj(const Problem &p, const double *dual, double *prim) {
    const int nx = p.nx, nu = p.nu, nxu = nx + nu, yD = p.yD, d4 = p.d4;
    double *u = prim + p.oU, *x = prim + p.oX, *y = prim + p.oY, *t = prim + p.oT, *s = prim + p.oS;
    const double *I = dual + p.oI, *II = dual + p.oII, *III = dual + p.oIII, *IV = dual + p.oIV,
        *V = dual + p.oV, *VI = dual + p.oVI;
    for (int i = 0; i < p.nl; i++) {
        s[i] = II[i];
        for (int k = 0; k < yD; k++) y[(size_t) i * yD + k] = I[(size_t) i * yD + k] - p.b[(size_t) i * yD + k] * II[i];
        double *xi = x + (size_t) i * nx, *ui = u + (size_t) i * nu;
        const double *iii = III + (size_t) i * p.con3;
        if (p.ncon == CON_RECT) {
            for (int k = 0; k < nx; k++) xi[k] = iii[k];
            for (int k = 0; k < nu; k++) ui[k] = iii[nx + k];
        } else if (p.ncon == CON_POLY) {
            double xu[1024];
            gemvT(p.nG.data(), p.con3, nxu, iii, xu, false);
            for (int k = 0; k < nx; k++) xi[k] = xu[k];
            for (int k = 0; k < nu; k++) ui[k] = xu[nx + k];
        } else if (p.ncon == CON_PWI) {
            // [x; u] = identity part + G' (G part), the adjoint of the operator above.  The
            // reference adds G' (G part) into its input's identity part in place
            // (src/constraints.cuh:418-422); this adjoint leaves its input unchanged.
            const int rows = p.con3 - p.id3;
            const double *idPart = III + (size_t) i * p.id3, *gPart = III + (size_t) p.nl * p.id3 + (size_t) i * rows;
            double xu[1024];
            gemvT(p.nG.data(), rows, nxu, gPart, xu, false);
            for (int k = 0; k < nx; k++) xi[k] = idPart[k] + xu[k];
            for (int k = 0; k < nu; k++) ui[k] = idPart[nx + k] + xu[nx + k];
        } else {
            for (int k = 0; k < nx; k++) xi[k] = 0.;
            for (int k = 0; k < nu; k++) ui[k] = 0.;
        }
        // children accumulate in child order (src/costs.cuh:172-177; linear: :274-285)
        for (long c = 0; c < p.numCh[i]; c++) {
            const long j = p.chFrom[i] + c;
            if (p.ncost == COST_LIN) {
                const double *g = p.grad.data() + (size_t) j * nxu;
                for (int k = 0; k < nx; k++) xi[k] += g[k] * IV[j];
                for (int k = 0; k < nu; k++) ui[k] += g[nx + k] * IV[j];
                continue;
            }
            const double *iv = IV + (size_t) j * d4;
            double w[1024];
            gemv(p.sqrtQ.data() + (size_t) j * nx * nx, nx, nx, iv, w, false);
            double cd = iv[nxu] + iv[nxu + 1];
            if (p.ncost == COST_QUADLIN)
                for (int k = 0; k < nx; k++) w[k] = w[k] - 0.5 * p.q[(size_t) j * nx + k] * iv[nxu] - 0.5 * p.q[(size_t) j * nx + k] * iv[nxu + 1];
            for (int k = 0; k < nx; k++) xi[k] += w[k];
            gemv(p.sqrtR.data() + (size_t) j * nu * nu, nu, nu, iv + nx, w, false);
            if (p.ncost == COST_QUADLIN)
                for (int k = 0; k < nu; k++) w[k] = w[k] - 0.5 * p.r[(size_t) j * nu + k] * iv[nxu] - 0.5 * p.r[(size_t) j * nu + k] * iv[nxu + 1];
            for (int k = 0; k < nu; k++) ui[k] += w[k];
            (void) cd;
        }
    }
    if (p.ncost == COST_LIN) {  // t <- set, every node (src/costs.cuh:286)
        for (int j = 0; j < p.N; j++) t[j] = IV[p.N + j];
    } else {
        t[0] = 0.;
        for (int j = 1; j < p.N; j++) {
            const double *iv = IV + (size_t) j * d4;
            t[j] = (iv[nxu] + iv[nxu + 1]) * 0.5;
        }
    }
    for (int l = 0; l < p.nlf; l++) {
        double *xl = x + (size_t) (p.nl + l) * nx;
        const double *v = V + (size_t) l * p.con5;
        if (p.lcon == CON_RECT) for (int k = 0; k < nx; k++) xl[k] = v[k];
        else if (p.lcon == CON_POLY) gemvT(p.lG.data(), p.con5, nx, v, xl, false);
        else if (p.lcon == CON_PWI) {  // src/constraints.cuh:430-433, without the in-place update
            const int rows = p.con5 - p.id5;
            const double *idPart = V + (size_t) l * p.id5, *gPart = V + (size_t) p.nlf * p.id5 + (size_t) l * rows;
            gemvT(p.lG.data(), rows, nx, gPart, xl, false);
            for (int k = 0; k < nx; k++) xl[k] += idPart[k];
        } else for (int k = 0; k < nx; k++) xl[k] = 0.;
        if (p.lcost == COST_LIN) {  // src/costs.cuh:289-292
            for (int k = 0; k < nx; k++) xl[k] += p.lgrad[k] * VI[l];
            s[p.nl + l] = VI[p.nlf + l];
            continue;
        }
        const double *vi = VI + (size_t) l * (nx + 2);
        double w[1024];
        gemv(p.lsqrtQ.data(), nx, nx, vi, w, false);
        for (int k = 0; k < nx; k++) xl[k] += w[k];
        if (p.lcost == COST_QUADLIN)
            for (int k = 0; k < nx; k++) xl[k] = xl[k] - 0.5 * p.lq[k] * vi[nx] - 0.5 * p.lq[k] * vi[nx + 1];
        s[p.nl + l] = (vi[nx] + vi[nx + 1]) * 0.5;
    }
}

// --------------------------------------------------------- dynamics projection
// src/dynamics.cuh:74-174: backward Riccati-like sweep over stages, then forward rollout.
// x: nx*N (in: x-bar, out: projected), u: nu*nl (in: u-bar, out: projected)
void dynProject(const Problem &p, const double *x0, double *x, double *u) {
    const int nx = p.nx, nu = p.nu;
    Vec q((size_t) nx * p.N, 0.), d((size_t) nu * p.nl, 0.), ubar(u, u + (size_t) nu * p.nl), xbar(x, x + (size_t) nx * p.N);
    for (int l = p.nl; l < p.N; l++)
        for (int k = 0; k < nx; k++) q[(size_t) l * nx + k] = -xbar[(size_t) l * nx + k];
    double w[1024], acc[1024];
    for (int stage = p.S - 2; stage >= 0; stage--) {
        for (long i = p.stageFrom[stage]; i <= p.stageTo[stage]; i++) {
            // affine dynamics: the children's q carry + P_j e_j from here on (src/dynamics.cuh:93-97)
            if (p.affine)
                for (long c = 0; c < p.numCh[i]; c++) {
                    const long j = p.chFrom[i] + c;
                    for (int k = 0; k < nx; k++) q[(size_t) j * nx + k] += p.Pe[(size_t) j * nx + k];
                }
            // d_i = R~^-1 (u_i - sum_j B_j' q_j)
            for (int k = 0; k < nu; k++) acc[k] = 0.;
            for (long c = 0; c < p.numCh[i]; c++) {
                const long j = p.chFrom[i] + c;
                gemv(p.BTr.data() + (size_t) j * nu * nx, nu, nx, q.data() + (size_t) j * nx, w, false);
                for (int k = 0; k < nu; k++) acc[k] = (c == 0) ? w[k] : acc[k] + w[k];
            }
            double *di = d.data() + (size_t) i * nu;
            for (int k = 0; k < nu; k++) di[k] = -acc[k] + ubar[(size_t) i * nu + k];
            const double *L = p.lowChol.data() + (size_t) i * nu * nu;
            for (int k = 0; k < nu; k++) {  // L y = d
                double v = di[k];
                for (int m = 0; m < k; m++) v -= L[k + (size_t) m * nu] * di[m];
                di[k] = v / L[k + (size_t) k * nu];
            }
            for (int k = nu - 1; k >= 0; k--) {  // L' x = y
                double v = di[k];
                for (int m = k + 1; m < nu; m++) v -= L[m + (size_t) k * nu] * di[m];
                di[k] = v / L[k + (size_t) k * nu];
            }
            if (stage == 0) continue;
            // q_i = sum_j (APB_j d_i + ABKTr_j q_j) + K_i'(d_i - u_i) - x_i
            for (int k = 0; k < nx; k++) acc[k] = 0.;
            for (long c = 0; c < p.numCh[i]; c++) {
                const long j = p.chFrom[i] + c;
                gemv(p.APB.data() + (size_t) j * nx * nu, nx, nu, di, w, false);
                gemv(p.ABKTr.data() + (size_t) j * nx * nx, nx, nx, q.data() + (size_t) j * nx, w, true);
                for (int k = 0; k < nx; k++) acc[k] = (c == 0) ? w[k] : acc[k] + w[k];
            }
            double du[1024];
            for (int k = 0; k < nu; k++) du[k] = di[k] - ubar[(size_t) i * nu + k];
            gemvT(p.K.data() + (size_t) i * nu * nx, nu, nx, du, w, false);
            for (int k = 0; k < nx; k++) q[(size_t) i * nx + k] = (acc[k] + w[k]) - xbar[(size_t) i * nx + k];
        }
    }
    for (int k = 0; k < nx; k++) x[k] = x0[k];
    for (int stage = 0; stage < p.S - 1; stage++) {
        for (long i = p.stageFrom[stage]; i <= p.stageTo[stage]; i++) {
            double *ui = u + (size_t) i * nu;
            gemv(p.K.data() + (size_t) i * nu * nx, nu, nx, x + (size_t) i * nx, ui, false);
            for (int k = 0; k < nu; k++) ui[k] += d[(size_t) i * nu + k];
            double xu[1024];
            for (int k = 0; k < nx; k++) xu[k] = x[(size_t) i * nx + k];
            for (int k = 0; k < nu; k++) xu[nx + k] = ui[k];
            for (long c = 0; c < p.numCh[i]; c++) {
                const long j = p.chFrom[i] + c;
                gemv(p.AB.data() + (size_t) j * nx * (nx + nu), nx, nx + nu, xu, x + (size_t) j * nx, false);
                if (p.affine)  // src/dynamics.cuh:168-171
                    for (int k = 0; k < nx; k++) x[(size_t) j * nx + k] += p.e[(size_t) j * nx + k];
            }
        }
    }
}

// ------------------------------------------------------------ kernel projection
// src/cache.cuh:650-676: v_i = [y_i; t_ch(i); s_ch(i)] <- NN'_i v_i for every nonleaf node;
// padded child slots enter as zero.
void kernelProject(const Problem &p, double *y, double *t, double *s) {
    const int yD = p.yD, nD = p.nD, E = p.E;
    double v[256], w[256];
    for (int i = 0; i < p.nl; i++) {
        for (int k = 0; k < nD; k++) v[k] = 0.;
        for (int k = 0; k < yD; k++) v[k] = y[(size_t) i * yD + k];
        for (long c = 0; c < p.numCh[i]; c++) {
            v[yD + c] = t[p.chFrom[i] + c];
            v[yD + E + c] = s[p.chFrom[i] + c];
        }
        gemv(p.NNtr.data() + (size_t) i * nD * nD, nD, nD, v, w, false);
        for (int k = 0; k < yD; k++) y[(size_t) i * yD + k] = w[k];
        for (long c = 0; c < p.numCh[i]; c++) {
            t[p.chFrom[i] + c] = w[yD + c];
            s[p.chFrom[i] + c] = w[yD + E + c];
        }
    }
}

// ------------------------------------------------------------------ projections
// Box projection as computed by src/definitions.cu:38-52 (including its arithmetic form).
inline double projBox(double v, double lb, double ub) {
    int lower = (v <= lb), upper = (v >= ub);
    double lbi = std::isinf(lb) ? 0. : lower * lb;
    double ubi = std::isinf(ub) ? 0. : upper * ub;
    return v * (1 - lower) * (1 - upper) + lbi + ubi;
}

// SOC projection of one cone (last entry is the scalar), src/definitions.cu:301-428
void projSoc(double *v, int dim) {
    double ss = 0.;
    for (int k = 0; k < dim - 1; k++) ss += v[k] * v[k];
    double nrm = std::sqrt(ss), tl = v[dim - 1];
    int i1 = nrm <= tl;
    int i2 = (nrm <= -tl) && !i1;
    int i3 = !i1 && !i2;
    double sc = (nrm + tl) / (2. * nrm + (1 - i3));
    for (int k = 0; k < dim; k++) v[k] *= (1 - i2);
    double mult = i3 * sc + 1 - i3;
    for (int k = 0; k < dim - 1; k++) v[k] *= mult;
    int i1b = (1 - i2) * (1 - i3);
    v[dim - 1] = i1b * v[dim - 1] + (1 - i1b) * (1 - i2) * (i3 * (sc * nrm - 1) + 1);
}

// A linear cost's "projection" (src/costs.cuh:296-301 with k_projectRectangle,
// src/definitions.cu:39-48, lower bound -inf, upper bound the epigraph variable): the cost entry
// is clamped to at most the epigraph entry, which stays as it is.  This is not the Euclidean
// projection onto {(c, t) : c <= t} (that one would move both to their mean); it is restated as
// the reference computes it.
double projBelow(double c, double t) { return c >= t ? t : c; }

// Dual prox on the dual workspace w (already scaled by 1/gamma and translated):
// w <- Pi_C(w) block by block, src/cache.cuh:685-698
void projectDual(const Problem &p, double *w) {
    const int yD = p.yD;
    double *I = w + p.oI, *II = w + p.oII, *III = w + p.oIII, *IV = w + p.oIV, *V = w + p.oV, *VI = w + p.oVI;
    for (int i = 0; i < p.nl; i++)  // AVaR: max(0, .) on the first 2*#children entries (src/risks.cuh:81-106)
        for (int k = 0; k < yD; k++) {
            double &e = I[(size_t) i * yD + k];
            int idx = (k < 2 * p.numCh[i]) && (e < 0.);
            e = (1 - idx) * e;
        }
    for (int i = 0; i < p.nl; i++) II[i] = std::max(0., II[i]);  // src/cones.cuh:134-137
    for (size_t k = 0; k < p.sIII; k++) III[k] = projBox(III[k], p.nLB[k], p.nUB[k]);
    if (p.ncost == COST_LIN) {
        for (int j = 0; j < p.N; j++) IV[j] = projBelow(IV[j], IV[p.N + j]);
    } else {
        const int d4 = p.nx + p.nu + 2;
        for (int j = 0; j < p.N; j++) projSoc(IV + (size_t) j * d4, d4);
    }
    for (size_t k = 0; k < p.sV; k++) V[k] = projBox(V[k], p.lLB[k], p.lUB[k]);
    if (p.lcost == COST_LIN) {
        for (int l = 0; l < p.nlf; l++) VI[l] = projBelow(VI[l], VI[p.nlf + l]);
    } else {
        for (int l = 0; l < p.nlf; l++) projSoc(VI + (size_t) l * (p.nx + 2), p.nx + 2);
    }
}

void translateDual(const Problem &p, double *w) {  // src/costs.cuh:197-199, 460-462; linear: none (:294)
    const int d4 = p.nx + p.nu + 2, d6 = p.nx + 2;
    if (p.ncost != COST_LIN)
        for (int j = 0; j < p.N; j++)
            for (int k = 0; k < d4; k++) w[p.oIV + (size_t) j * d4 + k] += p.trans[(size_t) j * d4 + k];
    if (p.lcost != COST_LIN)
        for (int l = 0; l < p.nlf; l++)
            for (int k = 0; k < d6; k++) w[p.oVI + (size_t) l * d6 + k] += p.ltrans[k];
}

// ------------------------------------------------------------------ CP operator T
// src/cache.cuh:748-801.  z -> cand = T(z).  x0s is the (scaled) initial state.
void opT(const Problem &p, const double *x0s, const double *z, double *cand) {
    const double g = p.gamma;
    Vec ws(p.sIter);
    double *wp = ws.data(), *wd = ws.data() + p.sPrim;
    const double *zp = z, *zd = z + p.sPrim;
    // modifyPrimal: x-bar = z_p - gamma L*(z_d)
    opLadj(p, zd, wp);
    for (size_t k = 0; k < p.sPrim; k++) wp[k] = wp[k] * (-g) + zp[k];
    // proximalPrimal
    wp[p.oS] -= g;  // proxRootS (src/cache.cuh:639-642)
    dynProject(p, x0s, wp + p.oX, wp + p.oU);
    kernelProject(p, wp + p.oY, wp + p.oT, wp + p.oS);
    for (size_t k = 0; k < p.sPrim; k++) cand[k] = wp[k];
    // modifyDual: gamma L(2 x-bar - z_p) + z_d
    for (size_t k = 0; k < p.sPrim; k++) wp[k] = wp[k] * 2. - zp[k];
    opL(p, wp, wd);
    for (size_t k = 0; k < p.sDual; k++) wd[k] = wd[k] * g + zd[k];
    // proximalDual: cand_d = gamma (w - Pi(w)), w = ws/gamma + translation
    for (size_t k = 0; k < p.sDual; k++) wd[k] *= p.gammaRecip;
    translateDual(p, wd);
    double *cd = cand + p.sPrim;
    for (size_t k = 0; k < p.sDual; k++) cd[k] = wd[k];
    projectDual(p, wd);
    for (size_t k = 0; k < p.sDual; k++) cd[k] = (cd[k] - wd[k]) * g;
}

// ell(v) = [L* v_d ; L v_p]
void ell(const Problem &p, const double *v, double *out) {
    opLadj(p, v + p.sPrim, out);
    opL(p, v, out + p.sPrim);
}

// dotM(x, y) = x' (y - gamma ell(y))   (src/cache.cuh:722-734)
double dotM(const Problem &p, const double *x, const double *y) {
    Vec l(p.sIter);
    ell(p, y, l.data());
    double acc = 0.;
    for (size_t k = 0; k < p.sIter; k++) acc += x[k] * (l[k] * (-p.gamma) + y[k]);
    return acc;
}

// <x, M y> through adjoints only: x'y - gamma (x_p' L* y_d + (L* x_d)' y_p), equal to dotM since
// x_d' L y_p = (L* x_d)' y_p.  The shared-batch GPU engine computes SuperMann's norms this way.
double dotMAdj(const Problem &p, const double *x, const double *y) {
    Vec ax(p.sPrim), ay(p.sPrim);
    opLadj(p, y + p.sPrim, ay.data());
    opLadj(p, x + p.sPrim, ax.data());
    double a = 0., b = 0.;
    for (size_t k = 0; k < p.sIter; k++) a += x[k] * y[k];
    for (size_t k = 0; k < p.sPrim; k++) b += x[k] * ay[k] + ax[k] * y[k];
    return a - p.gamma * b;
}

// -------------------------------------------------------- Householder least squares
// argmin ||Y g - r|| for Y n x m (column major), via Householder QR (geqrf + ormqr + trsm).
bool lstsqHouseholder(const double *Yin, size_t n, int m, const double *rin, double *g) {
    Vec Y(Yin, Yin + n * m), r(rin, rin + n);
    std::vector<double> tau(m);
    for (int k = 0; k < m; k++) {
        double *col = Y.data() + (size_t) k * n;
        double alpha = col[k], sig = 0.;
        for (size_t i = k + 1; i < n; i++) sig += col[i] * col[i];
        if (sig == 0.) { tau[k] = 0.; continue; }
        double nrm = std::sqrt(alpha * alpha + sig);
        double beta = alpha >= 0. ? -nrm : nrm;
        tau[k] = (beta - alpha) / beta;
        double sc = 1. / (alpha - beta);
        for (size_t i = k + 1; i < n; i++) col[i] *= sc;
        col[k] = beta;
        for (int jj = k + 1; jj <= m; jj++) {  // jj == m is the right-hand side
            double *cj = (jj < m) ? Y.data() + (size_t) jj * n : r.data();
            double w = cj[k];
            for (size_t i = k + 1; i < n; i++) w += col[i] * cj[i];
            w *= tau[k];
            cj[k] -= w;
            for (size_t i = k + 1; i < n; i++) cj[i] -= w * col[i];
        }
    }
    for (int k = m - 1; k >= 0; k--) {
        double v = r[k];
        for (int jj = k + 1; jj < m; jj++) v -= Y[(size_t) jj * n + k] * g[jj];
        g[k] = v / Y[(size_t) k * n + k];
    }
    return true;
}

// argmin ||Y g - r|| from the normal equations: Cholesky of G = Y'Y + 1e-12 max(diag G) I.  The
// GPU engines' alternative to the Householder solve (their Gram entries are accumulated as the
// columns arrive); returns false when the factorisation breaks down, and the direction is then -r.
bool lstsqGram(const double *Y, size_t n, int m, const double *r, double *g) {
    std::vector<double> G((size_t) m * m, 0.), b(m, 0.), L((size_t) m * m, 0.), y(m, 0.);
    for (int i = 0; i < m; i++) {
        for (int j = 0; j <= i; j++) {
            double acc = 0.;
            for (size_t k = 0; k < n; k++) acc += Y[(size_t) i * n + k] * Y[(size_t) j * n + k];
            G[(size_t) i * m + j] = G[(size_t) j * m + i] = acc;
        }
        double acc = 0.;
        for (size_t k = 0; k < n; k++) acc += Y[(size_t) i * n + k] * r[k];
        b[i] = acc;
    }
    double dmax = 0.;
    for (int i = 0; i < m; i++) dmax = std::max(dmax, G[(size_t) i * m + i]);
    if (!(dmax > 0.)) return false;
    const double ridge = 1e-12 * dmax;
    for (int i = 0; i < m; i++)
        for (int j = 0; j <= i; j++) {
            double v = G[(size_t) i * m + j] + (i == j ? ridge : 0.);
            for (int k = 0; k < j; k++) v -= L[(size_t) i * m + k] * L[(size_t) j * m + k];
            if (i == j) {
                if (!(v > 0.)) return false;
                L[(size_t) i * m + i] = std::sqrt(v);
            } else {
                L[(size_t) i * m + j] = v / L[(size_t) j * m + j];
            }
        }
    for (int i = 0; i < m; i++) {
        double v = b[i];
        for (int k = 0; k < i; k++) v -= L[(size_t) i * m + k] * y[k];
        y[i] = v / L[(size_t) i * m + i];
    }
    for (int i = m - 1; i >= 0; i--) {
        double v = y[i];
        for (int k = i + 1; k < m; k++) v -= L[(size_t) k * m + i] * g[k];
        g[i] = v / L[(size_t) i * m + i];
    }
    return true;
}

// ------------------------------------------------------------------ solver
struct Opts {
    double tolAbs, tolRel, maxTime;
    long maxIters, maxInner, andBuff, checkPeriod;
    long lsGram = 0;        // 1: Gram/Cholesky least squares instead of Householder
    double tolFixed = 0.;   // > 0: accept against this tolerance from the first check on
    long adjDot = 0;        // 1: SuperMann's norms and inner products through dotMAdj
};

struct Trace {  // per outer iteration: branch taken and the residual norm
    std::vector<int> branch;
    std::vector<double> w, err;
};

struct Solver {
    const Problem &p;
    Opts o;
    Vec z, zPrev, cand, backup, r, rPrev, dir, dirS, S, Y, x0s;
    double tol = 0., tolAbs, tolRel;
    long iters = 0;
    double lastErr = NAN;
    int status = -1;
    long k0 = 0, k1 = 0, k2 = 0, k2bt = 0, k3 = 0;
    Trace trace;
    // debug caches (src/cache.cuh:947-959): on when errAcc is set.  nT counts applications of T
    // (each calls L and L* once, the calls the reference counts, src/cache.cuh:704-716, :750, :773)
    long nT = 0;
    double *errAcc = nullptr;     // err0's running vector e (sPrim), kept by the caller across solves
    std::vector<double> records;  // {idx, callsL, err0, err1, err2, err3} per error check
    long nonfinite = -1;

    Solver(const Problem &pp, const Opts &oo) : p(pp), o(oo) {
        size_t n = p.sIter;
        z.assign(n, 0.); zPrev.assign(n, 0.); cand.assign(n, 0.); backup.assign(n, 0.);
        r.assign(n, 0.); rPrev.assign(n, 0.); dir.assign(n, 0.); dirS.assign(n, 0.);
        S.assign(n * o.andBuff, 0.); Y.assign(n * o.andBuff, 0.);
        tolAbs = o.tolAbs; tolRel = o.tolRel;
        if (p.precond) {  // src/cache.cuh:302-306
            double m = *std::min_element(p.scaling.begin(), p.scaling.end());
            tolAbs *= m; tolRel *= m;
        }
    }

    void setState(const double *x0) {  // src/cache.cuh:592-605
        x0s.assign(x0, x0 + p.nx);
        if (p.precond) for (int k = 0; k < p.nx; k++) x0s[k] = x0[k] * p.scaling[k];
    }

    // computeError (src/cache.cuh:910-961); returns true when the solve ends
    bool computeError(long idx, bool &initFlag, const std::chrono::steady_clock::time_point &t0, Vec &ellR) {
        size_t n = p.sIter;
        if (errAcc) {  // isFinite(candidate) (src/cache.cuh:52-62, :912): the reference throws
            for (size_t k = 0; k < n; k++)
                if (!std::isfinite(cand[k])) {
                    nonfinite = (long) k;
                    status = 3;
                    iters = idx;
                    return true;
                }
        }
        for (size_t k = 0; k < n; k++) r[k] = z[k] - cand[k];
        if (idx % o.checkPeriod != 0) return false;
        ell(p, r.data(), ellR.data());
        double mx = 0.;
        for (size_t k = 0; k < n; k++) mx = std::max(mx, std::fabs(r[k] * p.gammaRecip - ellR[k]));
        if (errAcc) {
            // err1/err2: max-norms of the primal and dual parts of r/gamma - ell(r); err3 = ||z - cand||;
            // err0: e <- cand_prim - e, ||e|| (the reference's m_d_err is never given the L* product
            // computed beside it, so e alternates over the candidates since the last reset)
            double e1 = 0., e2 = 0., s3 = 0., s0 = 0.;
            for (size_t k = 0; k < n; k++) {
                const double v = std::fabs(r[k] * p.gammaRecip - ellR[k]);
                if (k < p.sPrim) e1 = std::max(e1, v); else e2 = std::max(e2, v);
                s3 += r[k] * r[k];
            }
            for (size_t k = 0; k < p.sPrim; k++) {
                errAcc[k] = cand[k] - errAcc[k];
                s0 += errAcc[k] * errAcc[k];
            }
            const double rec[] = {(double) idx, (double) (2 * nT), std::sqrt(s0), e1, e2, std::sqrt(s3)};
            records.insert(records.end(), rec, rec + 6);
        }
        if (initFlag && o.tolFixed > 0.) {
            tol = o.tolFixed;
            initFlag = false;
        }
        if (initFlag) {
            tol = std::max(tolAbs, tolRel * mx);
            initFlag = false;
            return false;
        }
        lastErr = mx;
        int code = -1;
        if (mx <= tol) code = 0;
        if (o.maxIters && idx >= o.maxIters) code = 1;
        if (o.maxTime > 0 && idx % 1000 == 0) {
            double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (el >= o.maxTime) code = 2;
        }
        if (code != -1) { status = code; iters = idx; return true; }
        return false;
    }

    // src/cache.cuh:861-904
    void updateDirection(long idx) {
        size_t n = p.sIter;
        int m = (int) o.andBuff;
        for (int c = m - 1; c >= 1; c--) {
            std::copy(S.begin() + (size_t) (c - 1) * n, S.begin() + (size_t) c * n, S.begin() + (size_t) c * n);
            std::copy(Y.begin() + (size_t) (c - 1) * n, Y.begin() + (size_t) c * n, Y.begin() + (size_t) c * n);
        }
        for (size_t k = 0; k < n; k++) {
            double dr = r[k] - rPrev[k];
            double dz = z[k] - zPrev[k];
            Y[k] = dr;
            S[k] = dz - dr;
        }
        std::vector<double> g(m);
        const bool useLS = idx >= m && (o.lsGram ? lstsqGram(Y.data(), n, m, r.data(), g.data())
                                                 : lstsqHouseholder(Y.data(), n, m, r.data(), g.data()));
        if (useLS) {
            for (size_t k = 0; k < n; k++) {
                double sg = 0.;
                for (int c = 0; c < m; c++) sg += S[(size_t) c * n + k] * g[c];
                dir[k] = -sg - r[k];
            }
        } else {
            for (size_t k = 0; k < n; k++) dir[k] = -r[k];
        }
    }

    int runSpock(const double *x0, const double *warm, bool recordTrace) {
        auto t0 = std::chrono::steady_clock::now();
        size_t n = p.sIter;
        setState(x0);
        if (warm) {  // initialisePrev with a previous solution (src/cache.cuh:624-632)
            std::copy(warm, warm + n, z.begin());
            zPrev = z;
            rPrev = r;
        } else {
            std::fill(zPrev.begin(), zPrev.end(), 1.);
            std::fill(rPrev.begin(), rPrev.end(), 1.);
        }
        bool initFlag = true;
        status = -1;
        const double c0 = .99, c1 = .99, c2 = .99, beta = .5, sigma = .1, lambda = 1.;
        double zeta = INFINITY, w = INFINITY, wSafe = INFINITY, wTilde = INFINITY;
        Vec ellR(n), work(n);
        const long hardCap = 10000000;
        for (long it = 0; it < hardCap; it++) {
            if (it != 0) z = cand;
            opT(p, x0s.data(), z.data(), cand.data());
            nT++;
            if (computeError(it, initFlag, t0, ellR)) {
                if (status != 3) z = cand;
                break;
            }
            backup = cand;
            // normM(r): w = sqrt(r' (r - gamma ell(r)))  (src/cache.cuh:740-742)
            if (o.adjDot) {
                w = std::sqrt(dotMAdj(p, r.data(), r.data()));
            } else {
                ell(p, r.data(), ellR.data());
                double acc = 0.;
                for (size_t k = 0; k < n; k++) acc += r[k] * (ellR[k] * (-p.gamma) + r[k]);
                w = std::sqrt(acc);
            }
            updateDirection(it);
            zPrev = z;
            rPrev = r;
            if (recordTrace) { trace.w.push_back(w); trace.err.push_back(lastErr); }
            if (w <= c0 * zeta) {  // K0 (src/cache.cuh:1055-1061)
                for (size_t k = 0; k < n; k++) cand[k] = zPrev[k] + dir[k];
                zeta = w;
                k0++;
                if (recordTrace) trace.branch.push_back(0);
                continue;
            }
            double tau = 1.;
            int br = 3;
            for (long in = 0; in < o.maxInner; in++) {  // line search (src/cache.cuh:1063-1099)
                for (size_t k = 0; k < n; k++) { dirS[k] = dir[k] * tau; z[k] = zPrev[k] + dirS[k]; }
                opT(p, x0s.data(), z.data(), cand.data());
                nT++;
                for (size_t k = 0; k < n; k++) r[k] = z[k] - cand[k];
                wTilde = std::sqrt(o.adjDot ? dotMAdj(p, r.data(), r.data()) : dotM(p, r.data(), r.data()));
                if (w <= wSafe && wTilde <= c1 * w) {  // K1
                    cand = z;
                    wSafe = wTilde + std::pow(c2, (double) it);
                    k1++;
                    br = 1;
                    break;
                }
                double rho = wTilde * wTilde - (o.adjDot ? dotMAdj(p, r.data(), dirS.data())
                                                         : dotM(p, r.data(), dirS.data()));
                if (rho >= sigma * wTilde * w) {  // K2
                    double f = lambda * rho / (wTilde * wTilde);
                    for (size_t k = 0; k < n; k++) { r[k] *= f; cand[k] = zPrev[k] - r[k]; }
                    k2++;
                    br = 2;
                    break;
                } else {
                    tau *= beta;
                    k2bt++;
                }
                if (in >= o.maxInner - 1) { cand = backup; k3++; }
            }
            if (recordTrace) trace.branch.push_back(br);
        }
        return status;
    }

    int runCp(const double *x0, const double *warm) {  // src/cache.cuh:967-1002
        auto t0 = std::chrono::steady_clock::now();
        setState(x0);
        if (warm) std::copy(warm, warm + p.sIter, z.begin());
        bool initFlag = true;
        status = -1;
        Vec ellR(p.sIter);
        for (long it = 0; it < 10000000; it++) {
            opT(p, x0s.data(), z.data(), cand.data());
            nT++;
            bool done = computeError(it, initFlag, t0, ellR);
            if (!(done && status == 3)) z = cand;
            if (done) break;
        }
        return status;
    }
};

}  // namespace

// ====================================================================== C ABI
extern "C" {

void *orc_load(const char *path, const long *ints, double stepSize, char *errBuf, int errLen) {
    try {
        auto *p = new Problem();
        p->load(path, ints, stepSize);
        return p;
    } catch (const std::exception &e) {
        if (errBuf && errLen > 0) std::snprintf(errBuf, errLen, "%s", e.what());
        return nullptr;
    }
}

void orc_free(void *h) { delete static_cast<Problem *>(h); }

void orc_sizes(void *h, long *out) {
    auto &p = *static_cast<Problem *>(h);
    long v[] = {(long) p.sPrim, (long) p.sDual, (long) p.sIter, (long) p.sU, (long) p.sX, (long) p.sY,
                (long) p.sT, (long) p.sS, (long) p.sI, (long) p.sII, (long) p.sIII, (long) p.sIV,
                (long) p.sV, (long) p.sVI};
    std::memcpy(out, v, sizeof(v));
}

void orc_L(void *h, const double *prim, double *dual) { opL(*static_cast<Problem *>(h), prim, dual); }

void orc_Ladj(void *h, const double *dual, double *prim) { opLadj(*static_cast<Problem *>(h), dual, prim); }

void orc_dyn_project(void *h, const double *x0, double *x, double *u) {
    dynProject(*static_cast<Problem *>(h), x0, x, u);
}

void orc_kernel_project(void *h, double *prim) {
    auto &p = *static_cast<Problem *>(h);
    kernelProject(p, prim + p.oY, prim + p.oT, prim + p.oS);
}

void orc_project_dual(void *h, double *w) { projectDual(*static_cast<Problem *>(h), w); }

void orc_soc(double *v, int dim) { projSoc(v, dim); }

double orc_box(double v, double lb, double ub) { return projBox(v, lb, ub); }

void orc_T(void *h, const double *x0s, const double *z, double *cand) {
    opT(*static_cast<Problem *>(h), x0s, z, cand);
}

double orc_dotM(void *h, const double *x, const double *y) { return dotM(*static_cast<Problem *>(h), x, y); }

void orc_lstsq(const double *Y, long n, int m, const double *r, double *g) { lstsqHouseholder(Y, (size_t) n, m, r, g); }

int orc_lstsq_gram(const double *Y, long n, int m, const double *r, double *g) {
    return lstsqGram(Y, (size_t) n, m, r, g) ? 1 : 0;
}

// method 0 = SPOCK (runSpock), 1 = CP (runCp).  dopts = {tolAbs, tolRel, maxTime, tolFixed};
// iopts = {maxIters, maxInner, andBuff, checkPeriod, lsGram, adjDot}.  out = {iters, status, k0, k1, k2, k2bt, k3}
// outD = {lastErr, tol}.  z_out receives the final iterate; trace arrays (nullable) get per-iteration
// branch codes and residual norms up to traceLen entries.
int orc_solve(void *h, int method, const double *x0, const double *warm, const double *dopts, const long *iopts,
              double *zOut, long *out, double *outD, int *traceBranch, double *traceW, long traceLen) {
    auto &p = *static_cast<Problem *>(h);
    Opts o{dopts[0], dopts[1], dopts[2], iopts[0], iopts[1], iopts[2], iopts[3] > 0 ? iopts[3] : 25, iopts[4],
           dopts[3], iopts[5]};
    Solver s(p, o);
    bool tr = traceBranch || traceW;
    int st = method == 0 ? s.runSpock(x0, warm, tr) : s.runCp(x0, warm);
    std::copy(s.z.begin(), s.z.end(), zOut);
    if (out) {
        long v[] = {s.iters, (long) st, s.k0, s.k1, s.k2, s.k2bt, s.k3};
        std::memcpy(out, v, sizeof(v));
    }
    if (outD) { outD[0] = s.lastErr; outD[1] = s.tol; }
    if (tr) {
        long n = std::min<long>(traceLen, (long) s.trace.branch.size());
        for (long i = 0; i < n; i++) {
            if (traceBranch) traceBranch[i] = s.trace.branch[i];
            if (traceW) traceW[i] = s.trace.w[i];
        }
    }
    return st;
}

// orc_solve with the debug caches on (CacheBuilder::enableDebug, src/cache.cuh:947-959):
// errAcc (sPrim, in/out) is err0's running vector, zero after a reset; records (cap x 6) receive
// {idx, callsL, err0, err1, err2, err3} per error check; info = {records made, first non-finite
// candidate entry or -1, applications of T}.  Status 3: a non-finite candidate entry stopped the solve.
int orc_solve_debug(void *h, int method, const double *x0, const double *warm, const double *dopts,
                    const long *iopts, double *zOut, long *out, double *outD, double *errAcc, double *records,
                    long cap, long *info) {
    auto &p = *static_cast<Problem *>(h);
    Opts o{dopts[0], dopts[1], dopts[2], iopts[0], iopts[1], iopts[2], iopts[3] > 0 ? iopts[3] : 25, iopts[4],
           dopts[3], iopts[5]};
    Solver s(p, o);
    s.errAcc = errAcc;
    int st = method == 0 ? s.runSpock(x0, warm, false) : s.runCp(x0, warm);
    std::copy(s.z.begin(), s.z.end(), zOut);
    if (out) {
        long v[] = {s.iters, (long) st, s.k0, s.k1, s.k2, s.k2bt, s.k3};
        std::memcpy(out, v, sizeof(v));
    }
    if (outD) { outD[0] = s.lastErr; outD[1] = s.tol; }
    const long made = (long) s.records.size() / 6;
    std::copy(s.records.begin(), s.records.begin() + 6 * std::min(made, cap), records);
    info[0] = made;
    info[1] = s.nonfinite;
    info[2] = s.nT;
    return st;
}

}  // extern "C"
