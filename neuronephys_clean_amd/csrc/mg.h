// mg.h - geometric multigrid V-cycle for the variable-coefficient 7-point
// backward-Euler operator (mg.hip).
#pragma once
#include "grid_dev.h"
#include "internal.h"

namespace nnx {

constexpr int kMgMaxLevels = 12;

// One hierarchy: level 0 is the caller's grid; level l+1 halves every
// dimension of level l.  Buffers are context scratch.
struct Mg {
    int nlev = 0, nu = 3, coarse_sweeps = 16;
    int smoother = 0;               // NNX_MG_CHEBYSHEV or NNX_MG_JACOBI
    float w = 0.8f;                 // damped-Jacobi weight (and the coarsest level's)
    float c1[16], c2[16];           // Chebyshev sweep coefficients (sweep s: d = c1 d + c2 D^-1 r)
    float* d[kMgMaxLevels];         // Chebyshev direction per level
    GridK g[kMgMaxLevels];
    float* dinv[kMgMaxLevels];
    float* x[kMgMaxLevels];         // levels >= 1: the level's correction
    float* t[kMgMaxLevels];         // ping-pong partner of x
    float* b[kMgMaxLevels];         // levels >= 1: restricted residual
};

// Builds the hierarchy of a variable-coefficient 7-point grid (coarse face
// coefficients, diagonals, buffers); nu = smoothing sweeps per level.
nnx_status mg_setup(nnx_ctx* c, const GridK& g0, int nu, int smoother, Mg* m);
// z = V(r): one V-cycle from a zero initial guess.
nnx_status mg_vcycle(nnx_ctx* c, Mg& m, const float* r, float* z);

}  // namespace nnx
