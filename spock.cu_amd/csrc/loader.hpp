// Host-side loader for the problem file set (data.json + *.bt) and the packing of
// per-node tensors into deduplicated tables.
//
// File set and JSON keys: src/tree.cuh:102-179, src/problem.cuh:126-153 and the
// component constructors (src/dynamics.cuh:40-72, src/costs.cuh:85-110, 331-368,
// src/constraints.cuh:120-150, 203-245, src/risks.cuh:26-33, 81-97).
//
// Per-node tensors in the file set repeat the same matrix for every node that shares
// an event (the Python generator writes one copy per node).  They are packed here as
// a table of distinct matrices plus an int32 index per node, which keeps the solve's
// working set small enough to stay resident in L2/MALL.
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace spk {

enum CostKind { COST_QUAD = 0, COST_QUADLIN = 1, COST_LIN = 2 };
enum ConKind { CON_NO = 0, CON_RECT = 1, CON_POLY = 2, CON_PWI = 3 };  // CON_PWI: as read; loaded as CON_POLY

struct Table {                // distinct matrices (rows x cols, column major) + per-node index
    int rows = 0, cols = 0;
    std::vector<double> data;  // count * rows * cols
    std::vector<int> index;    // one entry per node of the owning range
    int count() const { return (rows > 0 && cols > 0) ? (int) (data.size() / ((size_t) rows * cols)) : 0; }
};

struct HostProblem {
    // sizes (data.json)
    int N = 0, nl = 0, nlf = 0, S = 0, E = 0, nx = 0, nu = 0, yD = 0, nD = 0;
    int ncost = COST_QUAD, lcost = COST_QUAD, ncon = CON_NO, lcon = CON_NO;
    int con3 = 0, con5 = 0;   // per-node sizes of dual blocks III and V
    int id3 = 0, id5 = 0;     // polyhedronWithIdentity: rows of the identity part (0 otherwise)
    bool precond = false;
    double stepSize = 0., alpha = 0.;
    std::vector<double> scaling;
    // tree
    std::vector<int> anc, chFrom, numCh, stageFrom, stageTo, events;
    std::vector<double> prob, condProb;
    // dynamics factors
    Table BTr, AB, K, KTr, ABKTr, APB, cholL, cholU, RtInv, P;  // RtInv = (L L')^-1
    bool affine = false;
    Table e, Pe;  // affine dynamics: e_j and P_j e_j per node (zero tables for linear dynamics)
    // risk
    Table NN, b;
    // costs (nonleaf tables are indexed per node 0..N-1, leaf tables per leaf)
    Table sQ, sR, q, r, trans, lsQ, lq, ltrans;
    // constraints: matrices (polyhedron) and bounds (rect/poly); nonleaf per nl, leaf per leaf
    Table nG, nGT, nLB, nUB, lG, lGT, lLB, lUB;
    // iterate layout (src/cache.cuh:482-581)
    long sU = 0, sX = 0, sY = 0, sT = 0, sS = 0, sPrim = 0;
    long sI = 0, sII = 0, sIII = 0, sIV = 0, sV = 0, sVI = 0, sDual = 0, sIter = 0;
};

// Reads a problem folder.  Throws std::runtime_error with a descriptive message on failure.
HostProblem loadProblem(const std::string &folder);

// polyhedronWithIdentity: one iterate z (size_iterate values) from the reference's order of the
// constraint dual blocks to the engines' per-node order (toEngine) or back.  No-op otherwise.
void reorderIdentityBlocks(const HostProblem &h, double *z, bool toEngine);

// Minimal JSON reader for data.json: flattens nested objects into dotted keys.
std::map<std::string, std::string> parseJsonFlat(const std::string &text);

// Reads a .bt tensor with the given element kind suffix ('d', 'f' or 'u').
bool readTensor(const std::string &path, char kind, std::vector<uint64_t> &dims, std::vector<double> &out);

// Groups per-node matrices into distinct entries (bytewise equality).
Table dedup(const std::vector<double> &stack, int rows, int cols, int count);

}  // namespace spk
