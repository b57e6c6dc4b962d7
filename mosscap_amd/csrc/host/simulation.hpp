// Host side of the drop-in: the Simulation with its operator slots
// (Simulation.hpp:17-62), the setup modules (SetupModules.hpp:9-262), the
// problem registry (ProblemGenerator.hpp:12-53) and the driver loop
// (main.cpp:50-73). The slots are bound to the C ABI of the HIP library
// (mosscap_hip.h); the conserved state lives on the device inside the mh_ctx
// and `state.Q` is the host copy used to set initial conditions and to write
// output.
#pragma once

#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/mosscap_hip.h"
#include "yaml_lite.hpp"

namespace mh::host {

using Config = std::shared_ptr<YNode>;

enum class Reconstruction { Fog = 0, Muscl, Ppm, Weno5Z };
enum class SlopeLimiter { VanLeer = 0, MonotonizedCentral, Minmod };
enum class RiemannSolver { Rusanov = 0, Hll, Hllc };
enum class TimeStepScheme { Rk2 = 0, SspRk3, SspRk4 };
enum class BoundaryType { Wall = 0, Periodic, Symmetric, SymmetricOutflowDiode, ZeroGrad, Constant, UserFn };

// Reconstruct.hpp:47-61.
inline int min_ghost_cells(Reconstruction r) {
    return r == Reconstruction::Fog ? 1 : r == Reconstruction::Muscl ? 2 : 3;
}

inline int num_hydro_vars(int num_dim) { return num_dim + 2; }

struct GridSize {
    int32_t xc = 0, yc = 0, zc = 0, ng = 0;
};
struct GridLoc {
    double x = 0.0, y = 0.0, z = 0.0;
};
struct Boundaries {
    BoundaryType xs = BoundaryType::Wall, xe = BoundaryType::Wall, ys = BoundaryType::Wall, ye = BoundaryType::Wall,
                 zs = BoundaryType::Wall, ze = BoundaryType::Wall;
    double xs_const[5] = {}, xe_const[5] = {}, ys_const[5] = {}, ye_const[5] = {}, zs_const[5] = {},
           ze_const[5] = {};
};

struct State {
    GridSize sz;
    double dx = 0.0;
    GridLoc loc;
    Boundaries boundaries;
    int32_t num_tracers = 0;
    int nvar = 0;
    std::vector<double> Q;  // host copy, [var][k][j][i], ghosts included

    // State::get_pos (State.hpp:123-134), same grouping.
    double pos_x(int i) const { return i * dx + (-(sz.ng - 0.5) * dx) + loc.x; }
    double pos_y(int j) const { return j * dx + (-(sz.ng - 0.5) * dx) + loc.y; }
    double pos_z(int k) const { return k * dx + (-(sz.ng - 0.5) * dx) + loc.z; }
    int64_t cells() const { return (int64_t)sz.xc * sz.yc * sz.zc; }
    double& q(int var, int k, int j, int i) { return Q[(((int64_t)var * sz.zc + k) * sz.yc + j) * sz.xc + i]; }
};

struct NumericalSchemes {
    Reconstruction reconstruction = Reconstruction::Muscl;
    SlopeLimiter slope_limit = SlopeLimiter::MonotonizedCentral;
    RiemannSolver riemann_solver = RiemannSolver::Hll;
    TimeStepScheme time_stepper = TimeStepScheme::SspRk3;
};

// Eos (Eos.hpp:27-47): ideal gas, or the analytic LTE hydrogen EOS whose
// ionisation fraction and temperature fields are written with the output.
// The tabulated LTE hydrogen EOS table (TabulatedLteH.hpp:10-28), read from
// eos.table_path by read_lte_table (table_nc.cpp).
struct LteTable {
    double log_eint_rho_start = 0.0, log_eint_rho_step = 0.0;
    int n_eint_rho = 0;
    double log_rho_start = 0.0, log_rho_step = 0.0;
    int n_rho = 0;
    std::vector<double> y_table, log_T_table;  // [n_rho][n_eint_rho]
};
LteTable read_lte_table(const std::string& path);

struct Eos {
    double gamma = 1.4;
    bool is_constant = true;
    double avg_mass = 1.0;
    bool include_ionisation_energy = false;
    bool tabulated = false;                // eos.type = tabulatedlteh
    LteTable table;
    std::vector<double> y_space, T_space;  // LTE fields of the last update_eos
};

// OutputConfig (Output.hpp), read by setup_output (SetupModules.hpp:179-196).
struct OutputConfig {
    std::string problem_name;
    std::string filename;
    bool single_file = true;
    double delta = 0.1;
    int output_count = 0;
    double prev_output_time = -1.0;
    struct {
        bool conserved = true, primitive = false, fluxes = false, source = false;
    } variables;
    std::string directory = ".";
};

// Options of the HIP library (no counterpart in the reference), from the
// optional "hip" section of the configuration.
struct DeviceOptions {
    int precision_form = MH_FORM_FP64;
    int span_rows = 0;
    double rho_floor = 0.0, p_floor = 0.0;
};

struct Simulation;
using SimFn = std::function<void(const Simulation&)>;

// The user_bc slot runs inside the library's ghost fills; the library's hook
// finds the Simulation through this holder, which every host operator points
// at the Simulation it was called with (a Simulation may move after setup).
struct HookHolder {
    const Simulation* sim = nullptr;
};

struct Simulation {
    int num_dim = 0;
    int64_t current_step = 0;
    double max_cfl = 0.8;
    double time = 0.0;
    double max_time = 1.0;
    int64_t max_steps = -1;  // timestep.max_steps; < 0 means no limit
    double dt = 0.0;
    Eos eos;
    State state;
    NumericalSchemes scheme;
    OutputConfig out_cfg;
    DeviceOptions device;
    // registered source terms (Simulation::compute_source_terms): gravity only
    bool gravity = false;
    double grav[3] = {0.0, 0.0, 0.0};
    std::shared_ptr<mh_ctx> ctx;  // device state and operators
    std::shared_ptr<HookHolder> hook = std::make_shared<HookHolder>();

    // operator slots (Simulation.hpp:52-58)
    std::function<double(const Simulation&)> compute_dt;
    std::function<void(Simulation&, double)> time_step;
    SimFn user_bc;
    std::function<bool(Simulation&)> write_output;

    mh_ctx* handle() const {
        if (!ctx) throw std::runtime_error("the simulation has no device context (setup_sim not run)");
        hook->sim = this;
        return ctx.get();
    }
};

// Throws std::runtime_error with the library's message for a failed C ABI call.
void check(const Simulation& sim, int rc, const char* what);

// ---- setup (SetupModules.hpp) ----
void setup_grid(Simulation& sim, const Config& config);
void setup_boundaries(Simulation& sim, const Config& config);
void setup_hydro_fns(Simulation& sim, const Config& config);
void setup_timestepper(Simulation& sim, const Config& config);
void setup_eos(Simulation& sim, const Config& config);
void setup_output(Simulation& sim, const Config& config);
void setup_problem(Simulation& sim, const Config& config);
void setup_device_options(Simulation& sim, const Config& config);
// Host part of setup_sim: every module above, ending with the problem's
// initial state in state.Q. No device work.
Simulation setup_sim_host(const Config& config);
// Creates the device context, uploads Q, binds the operator slots, fills the
// ghosts and writes the initial output (SetupModules.hpp:232-262).
void attach_device(Simulation& sim, bool write_initial_output = true);
Simulation setup_sim(const Config& config);

// ---- operators bound to the device ----
void fill_bcs(const Simulation& sim);           // Boundaries.hpp:132-149
double compute_dt(const Simulation& sim);       // Hydro.hpp:16
void download_state(Simulation& sim);           // device Q -> state.Q
std::vector<double> primitive_state(const Simulation& sim);  // cons_to_prim of state.Q, all cells

// ---- output (Output.cpp:201-278) ----
bool write_output(Simulation& sim);

// ---- problem registry (ProblemGenerator.hpp:12-53) ----
template <class T>
struct DispatchFactory {
    using Callable = std::function<void(T&, const Config&)>;
    std::map<std::string, Callable> registry;
    void dispatch(const std::string& identifier, T& arg, const Config& cfg) const {
        auto it = registry.find(identifier);
        if (it != registry.end()) return it->second(arg, cfg);
        throw std::runtime_error("No method for " + identifier + " found.");
    }
    bool register_method(const std::string& identifier, const Callable& fn) {
        if (registry.count(identifier)) return false;
        registry[identifier] = fn;
        return true;
    }
};
DispatchFactory<Simulation>& get_problem_generator();

// ---- driver (main.cpp:50-73) ----
struct DriverResult {
    int64_t steps = 0;
    double time = 0.0;
    int outputs = 0;
};
DriverResult run_driver(Simulation& sim, bool verbose = true);

// setup_gravity (Gravity.cpp:43-70): registers the gravity source with
// sources.gravity.{x,y,z} (defaults -1, 0, 0); `forced` components (x, y, z
// order, NaN = read from the configuration) take precedence.
void setup_gravity(Simulation& sim, const Config& config, const double (&forced)[3]);

// splitmix64 counter generator for seeded initial conditions.
uint64_t splitmix64(uint64_t x);
double uniform01(uint64_t seed, uint64_t counter);

}  // namespace mh::host
