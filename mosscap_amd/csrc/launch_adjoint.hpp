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
    double dt;
    int32_t gravity;  // a gravity source term is registered
    double gx, gy;
    // the stage inputs' ghosts are not filled: ghost cells are read from
    // their sources (every face periodic or zerograd; xper/yper: periodic)
    int32_t virt, xper, yper;
};

// One reverse stage of the stage output a_qn*Q^n + a_qs*Q_s + a_l*L(Q_s).
struct AdjStage {
    const double* q;   // Q_s, ghost-filled unless AdjGrid::virt
    const double* mu;  // cotangent of the stage output (interior; ghost entries zero)
    double* lamx;      // scratch [4][cells]
    double* nxt;       // out: a_qs*mu + a_l*J_L(Q_s)^T mu (interior; ghosts folded and zero)
    double* acc;       // cotangent of Q^n through the stages' Q^n terms (interior; ghosts zero)
    double a_qn, a_qs, a_l;
    // 0: acc += a_qn*mu; 1 (the step's last stage, reversed first): acc = a_qn*mu;
    // 2 (stage 0, reversed last): nxt += acc, giving dJ/dQ^n
    int32_t acc_mode;
};

hipError_t launch_adjoint_stage(const Schemes& sch, const AdjGrid& g, const AdjStage& st, hipStream_t stream);

// One tangent-linear stage (kernels/tangent.hip): values and derivatives of
// Q^(s) and Q^n (ghost-filled) to those of Q^(s+1) on the interior cells.
struct TanStage {
    const double* qv;  // Q^(s) values and derivatives
    const double* qd;
    const double* nv;  // Q^n
    const double* nd;
    double* ov;        // Q^(s+1)
    double* od;
    int32_t xc, yc, ng, nx, ny;
    int32_t kind;      // StageKind
    double gamma, dt, dx;
    int32_t gravity;
    double gx, gy;
};
hipError_t launch_tangent_stage(const Schemes& sch, const TanStage& a, hipStream_t st);
// y = a*x + b*y on the interior, 0 on ghosts.
hipError_t launch_adjoint_axpby(const AdjGrid& g, double a, const double* x, double b, double* y, hipStream_t st);
// unit probes of the hand VJPs against dual-number Jacobians
hipError_t launch_probe_vjp(int what, int n, const double* in, const double* cot, double* out, double gamma,
                            hipStream_t st);

}  // namespace mh
