// Launch interface of the discrete adjoint (capi_adjoint.cpp <-> kernels/adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "launch.hpp"

namespace mh {

struct AdjGrid {
    int32_t xc, yc, ng, nx, ny, ncells;
    int32_t bc[4];  // xs, xe, ys, ye
    double gamma;
    double dtdx;    // dt / dx of the step (frozen)
};

struct AdjScratch {
    double* w;       // primitives [4][cells]
    double* lam_rr;  // cotangents of the right-face reconstruction [4][cells]
    double* lam_rl;  // ... of the left-face reconstruction
    double* gx;      // stencil contributions [5][4][cells], x sweep
    double* gy;      // ... y sweep
};

// out = J_L(q)^T mu on the interior (ghosts zero); q ghost-filled.
hipError_t launch_adjoint_jt(const Schemes& sch, const AdjGrid& g, const double* q, const double* mu, double* out,
                             const AdjScratch& s, hipStream_t st);
// y = a*x + b*y on the interior, 0 on ghosts.
hipError_t launch_adjoint_axpby(const AdjGrid& g, double a, const double* x, double b, double* y, hipStream_t st);
// unit probes of the hand VJPs against dual-number Jacobians
hipError_t launch_probe_vjp(int what, int n, const double* in, const double* cot, double* out, double gamma,
                            hipStream_t st);

}  // namespace mh
