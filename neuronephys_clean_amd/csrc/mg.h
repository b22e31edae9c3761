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
    int nlev = 0, nu = 4, coarse_sweeps = 16;
    float w = 0.8f;                 // damped-Jacobi weight
    GridK g[kMgMaxLevels];
    float* dinv[kMgMaxLevels];
    float* x[kMgMaxLevels];         // levels >= 1: the level's correction
    float* t[kMgMaxLevels];         // ping-pong partner of x
    float* b[kMgMaxLevels];         // levels >= 1: restricted residual
};

// Builds the hierarchy of a variable-coefficient 7-point grid (coarse face
// coefficients, diagonals, buffers); nu = smoothing sweeps per level.
nnx_status mg_setup(nnx_ctx* c, const GridK& g0, int nu, Mg* m);
// z = V(r): one V-cycle from a zero initial guess.
nnx_status mg_vcycle(nnx_ctx* c, Mg& m, const float* r, float* z);

}  // namespace nnx
