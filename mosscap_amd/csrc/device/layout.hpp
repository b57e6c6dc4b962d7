// Variable index maps, scheme enums and POD layouts of the hydro state.
//
// Index conventions follow Mosscap's State.hpp:10-60 and Types.hpp:66-77 so a
// caller's arrays need no reshuffling: conserved/primitive fields are stored
// [var][k][j][i] (x fastest), ghost cells included; Rho=0, the velocity or
// momentum along axis a is 1+a, Pres/Ene is 1+NumDim; unused axes map to 1024.
#pragma once

#include <cstdint>

namespace mh {

enum class Reconstruction : int32_t { Fog = 0, Muscl, Ppm, Weno5Z };              // Reconstruct.hpp:10-15
enum class SlopeLimiter : int32_t { VanLeer = 0, MonotonizedCentral, Minmod };    // Reconstruct.hpp:22-26
enum class RiemannSolver : int32_t { Rusanov = 0, Hll, Hllc };                     // Riemann.hpp:10-14
enum class TimeStepScheme : int32_t { Rk2 = 0, SspRk3, SspRk4 };                   // TimeStepping.hpp:10-14
enum class BoundaryType : int32_t {                                                // State.hpp:62-70
    Wall = 0,
    Periodic,
    Symmetric,
    SymmetricOutflowDiode,
    ZeroGrad,
    Constant,
    UserFn
};

template <int NumDim = 1>
struct Prim {  // State.hpp:10-17
    static constexpr int32_t Rho = 0;
    static constexpr int32_t Vx = 1;
    static constexpr int32_t Vy = NumDim > 1 ? 2 : 1024;
    static constexpr int32_t Vz = NumDim > 2 ? 3 : 1024;
    static constexpr int32_t Pres = 1 + NumDim;
};

template <int NumDim = 1>
struct Cons {  // State.hpp:19-26
    static constexpr int32_t Rho = 0;
    static constexpr int32_t MomX = 1;
    static constexpr int32_t MomY = NumDim > 1 ? 2 : 1024;
    static constexpr int32_t MomZ = NumDim > 2 ? 3 : 1024;
    static constexpr int32_t Ene = 1 + NumDim;
};

template <int NumDim>
constexpr int N_HYDRO_VARS = 2 + NumDim;  // State.hpp:28-29

template <int Axis, int NumDim>
constexpr int Velocity() {  // State.hpp:40-49
    return Axis == 0 ? Prim<NumDim>::Vx : Axis == 1 ? Prim<NumDim>::Vy : Prim<NumDim>::Vz;
}

template <int Axis, int NumDim>
constexpr int Momentum() {  // State.hpp:51-60
    return Axis == 0 ? Cons<NumDim>::MomX : Axis == 1 ? Cons<NumDim>::MomY : Cons<NumDim>::MomZ;
}

struct CellIndex {  // Types.hpp:66-77
    int32_t i;
    int32_t j;
    int32_t k;
};

// Minimum ghost cells per scheme, Reconstruct.hpp:47-61.
constexpr int min_ghost_cells(Reconstruction r) {
    return r == Reconstruction::Fog ? 1 : r == Reconstruction::Muscl ? 2 : 3;
}

constexpr int num_stages(TimeStepScheme s) {
    return s == TimeStepScheme::Rk2 ? 2 : s == TimeStepScheme::SspRk3 ? 3 : 4;
}

}  // namespace mh
