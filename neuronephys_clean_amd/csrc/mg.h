// mg.h - geometric multigrid cycle for the variable-coefficient 7-point
// backward-Euler operator (mg.hip).
#pragma once
#include "grid_dev.h"
#include "internal.h"

namespace nnx {

constexpr int kMgMaxLevels = 12;
// levels with at most this many cells (and every level below them) are
// cycled by one workgroup in one launch, always as a V-cycle
constexpr int kMgSubCells = 4096;

// One step of a cycle, executed either as a kernel launch (large levels) or
// by the one-workgroup kernel (small levels).  `level` is the level the
// operator of the step belongs to (the fine level for RESTRICT/PROLONG).
enum MgOpType { MG_CHEB = 0, MG_JAC, MG_RESTRICT, MG_RESID, MG_PROLONG, MG_ADD };
struct MgOp {
    int type, level;
    const float* rhs;   // CHEB/JAC/RESTRICT: right-hand side
    const float* xin;   // CHEB/JAC: iterate (NULL = 0); RESTRICT/RESID: x; PROLONG/ADD: what is added
    float* xout;        // CHEB/JAC: new iterate; RESTRICT: coarse rhs; RESID: b (in place); PROLONG/ADD: target
    float* dvec;        // CHEB: direction
    const float* dinv;  // CHEB/JAC
    float c1, c2;       // CHEB: d = c1 d + c2 D^-1 r; JAC: c1 = weight
};

// One hierarchy: level 0 is the caller's grid; level l+1 halves every
// dimension of level l.  Buffers are context scratch.
struct Mg {
    int nlev = 0, nu = 3, coarse_sweeps = 16, gamma = 2;  // gamma: 1 V-cycle, 2 W-cycle
    int smoother = 0;               // NNX_MG_CHEBYSHEV or NNX_MG_JACOBI
    float w = 0.8f;                 // damped-Jacobi weight (and the coarsest level's)
    float c1[16], c2[16];           // Chebyshev sweep coefficients (sweep s: d = c1 d + c2 D^-1 r)
    GridK g[kMgMaxLevels];
    float* dinv[kMgMaxLevels];
    float* d[kMgMaxLevels];         // Chebyshev direction
    float* b[kMgMaxLevels];         // levels >= 1: restricted residual
    float* x[2][kMgMaxLevels];      // per visit set: iterate and its ping-pong partner
    float* t[2][kMgMaxLevels];
    int sub = -1;                   // first level cycled in one workgroup (-1: none)
    const MgOp* sub_ops[2] = {nullptr, nullptr};  // device op lists of a visit of level `sub`, per set
    int sub_nops[2] = {0, 0};
    float* sub_res[2] = {nullptr, nullptr};       // where each of those visits leaves its result
    const GridK* d_gk = nullptr;    // device copy of g[]
};

// Builds the hierarchy of a variable-coefficient 7-point grid (coarse face
// coefficients, diagonals, buffers, the small levels' op lists); nu =
// smoothing sweeps per level, gamma = visits of each coarser level (1 V,
// 2 W).
nnx_status mg_setup(nnx_ctx* c, const GridK& g0, int nu, int smoother, int gamma, Mg* m);
// z = M(r): one cycle from a zero initial guess.
nnx_status mg_vcycle(nnx_ctx* c, Mg& m, const float* r, float* z);

}  // namespace nnx
