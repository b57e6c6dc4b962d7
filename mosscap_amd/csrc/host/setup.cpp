// Setup modules and operator binding (SetupModules.hpp:9-262, Hydro.cpp:244-323,
// TimeStepping.cpp:95-111) for the host side of the drop-in.
#include <cstdio>
#include <cstring>

#include "config.hpp"
#include "simulation.hpp"

namespace mh::host {

void check(const Simulation& sim, int rc, const char* what) {
    if (rc == MH_OK) return;
    const char* msg = mh_last_error(sim.ctx ? sim.ctx.get() : nullptr);
    throw std::runtime_error(std::string(what) + ": " + (msg && *msg ? msg : "error code " + std::to_string(rc)));
}

// setup_grid, SetupModules.hpp:9-64.
void setup_grid(Simulation& sim, const Config& config) {
    auto& sz = sim.state.sz;
    sz.ng = get_or<int>(config, "grid.num_ghost", 3);
    sz.xc = get_or<int>(config, "grid.x", 256) + 2 * sz.ng;
    sz.yc = 1;
    sz.zc = 1;
    if (sim.num_dim > 1) sz.yc = get_or<int>(config, "grid.y", 256) + 2 * sz.ng;
    if (sim.num_dim > 2) sz.zc = get_or<int>(config, "grid.z", 256) + 2 * sz.ng;
    if (sz.ng < 0 || sz.xc - 2 * sz.ng < 1 || sz.yc < 1 || sz.zc < 1)
        throw std::runtime_error("grid sizes must be positive");

    const YNode* grid = config->get("grid");
    if (grid && grid->get("dx") && !grid->get("dx")->is_null()) {
        // The reference reads grid.dx into a shadowing local and leaves the
        // grid spacing unset (SetupModules.hpp:23-28); the value given is used here.
        const double dx = get_or<double>(config, "grid.dx", -1.0);
        if (dx < 0.0) throw std::runtime_error("grid.dx cannot be negative.");
        sim.state.dx = dx;
    } else {
        const double x_dim = get_or<double>(config, "grid.x_dim", 1.0);
        sim.state.dx = x_dim / (sz.xc - 2 * sz.ng);
    }
    sim.state.loc.x = get_or<double>(config, "grid.x_start", 0.0);
    sim.state.loc.y = get_or<double>(config, "grid.y_start", 0.0);
    sim.state.loc.z = get_or<double>(config, "grid.z_start", 0.0);

    if (get_or<bool>(config, "dex.enable", false))
        throw std::runtime_error("dex.enable: the radiative-transfer coupling is not supported by this library");
    sim.state.num_tracers = get_or<int>(config, "simulation.n_extra_fields", 0);
    if (sim.state.num_tracers < 0) throw std::runtime_error("simulation.n_extra_fields cannot be negative");
    sim.state.nvar = num_hydro_vars(sim.num_dim) + sim.state.num_tracers;
    sim.state.Q.assign((size_t)sim.state.nvar * sim.state.cells(), 0.0);
}

// setup_boundaries, SetupModules.hpp:66-136.
void setup_boundaries(Simulation& sim, const Config& config) {
    auto& bound = sim.state.boundaries;
    auto set_boundary = [&](BoundaryType& out, const std::string& face) {
        const std::string s = get_or<std::string>(config, "boundary." + face, "wall");
        out = find_associated_enum<BoundaryType>(BoundaryTypeName, 7, s, "Mosscap::BoundaryType");
    };
    set_boundary(bound.xs, "xs");
    set_boundary(bound.xe, "xe");
    if (sim.num_dim > 1) {
        set_boundary(bound.ys, "ys");
        set_boundary(bound.ye, "ye");
    }
    if (sim.num_dim > 2) {
        set_boundary(bound.zs, "zs");
        set_boundary(bound.ze, "ze");
    }

    // Constant faces: a 5-entry list in 3-D conserved order (rho, m_x, m_y,
    // m_z, E) mapped onto the grid's own variable order, or the string
    // "problem_supplied" (the problem generator fills the values).
    auto load_constant = [&](BoundaryType b, double* arr, const std::string& face) {
        if (b != BoundaryType::Constant) return;
        const YNode* bnode = config->get("boundary");
        const YNode* node = bnode ? bnode->get(face + "_const") : nullptr;
        if (!node || node->is_null()) return;
        if (node->kind == YNode::Seq) {
            auto at = [&](int i) {
                if (i >= (int)node->seq.size() || node->seq[i]->kind != YNode::Scalar)
                    throw std::runtime_error("boundary." + face + "_const needs 5 values (rho, m_x, m_y, m_z, E)");
                double v = 0.0;
                if (!convert_scalar(node->seq[i]->scalar, v))
                    throw std::runtime_error("Unable to convert boundary." + face + "_const entry \"" +
                                             node->seq[i]->scalar + "\" to type \"double\".");
                return v;
            };
            if (sim.num_dim == 1) {
                arr[0] = at(0);
                arr[1] = at(1);
                arr[2] = at(4);
            } else if (sim.num_dim == 2) {
                arr[0] = at(0);
                arr[1] = at(1);
                arr[2] = at(2);
                arr[3] = at(4);
            } else {
                for (int i = 0; i < 5; ++i) arr[i] = at(i);
            }
        } else {
            const std::string vals = get_or<std::string>(config, "boundary." + face + "_const", "xxx");
            if (vals == "xxx")
                throw std::runtime_error("Provide constant boundary values for " + face +
                                         "_const, or set to problem_supplied and fill in your problem setup function");
        }
    };
    load_constant(bound.xs, bound.xs_const, "xs");
    load_constant(bound.xe, bound.xe_const, "xe");
    if (sim.num_dim > 1) {
        load_constant(bound.ys, bound.ys_const, "ys");
        load_constant(bound.ye, bound.ye_const, "ye");
    }
    if (sim.num_dim > 2) {
        load_constant(bound.zs, bound.zs_const, "zs");
        load_constant(bound.ze, bound.ze_const, "ze");
    }
}

// setup_hydro_fns, SetupModules.hpp:138-156.
void setup_hydro_fns(Simulation& sim, const Config& config) {
    auto& scheme = sim.scheme;
    const std::string recon = get_or<std::string>(config, "scheme.reconstruction", "muscl");
    scheme.reconstruction =
        find_associated_enum<Reconstruction>(ReconstructionName, 4, recon, "Mosscap::Reconstruction");
    const int min_ghost = min_ghost_cells(scheme.reconstruction);
    if (sim.state.sz.ng < min_ghost) {
        throw std::runtime_error(std::string("Insufficient ghost cells for reconstruction scheme ") +
                                 ReconstructionName[(int)scheme.reconstruction] + ", need minimum of " +
                                 std::to_string(min_ghost) + ".");
    }
    const std::string sl = get_or<std::string>(config, "scheme.slope_limiter", "monotonizedcentral");
    scheme.slope_limit = find_associated_enum<SlopeLimiter>(SlopeLimiterName, 3, sl, "Mosscap::SlopeLimiter");
    const std::string rs = get_or<std::string>(config, "scheme.riemann_solver", "hll");
    scheme.riemann_solver = find_associated_enum<RiemannSolver>(RiemannSolverName, 3, rs, "Mosscap::RiemannSolver");
}

// setup_timestepper, SetupModules.hpp:158-168; timestep.max_steps is new.
void setup_timestepper(Simulation& sim, const Config& config) {
    const std::string ts = get_or<std::string>(config, "timestep.scheme", "ssprk3");
    sim.scheme.time_stepper = find_associated_enum<TimeStepScheme>(TimeStepName, 3, ts, "Mosscap::TimeStepScheme");
    sim.max_cfl = get_or<double>(config, "timestep.max_cfl", 0.8);
    sim.max_time = get_or<double>(config, "timestep.max_time", 1.0);
    sim.max_steps = get_or<long long>(config, "timestep.max_steps", -1);
}

// Eos::init, Eos.cpp:12-88: the ideal gas and the analytic and tabulated LTE
// hydrogen EOS (the table is read here, as TabulatedLteH::init does at setup).
void setup_eos(Simulation& sim, const Config& config) {
    static const char* const EosTypeName[] = {"ideal", "analyticlteh", "tabulatedlteh", "dexpressure"};
    const std::string type = get_or<std::string>(config, "eos.type", "ideal");
    const int t = find_associated_enum<int>(EosTypeName, 4, type, "Mosscap::EosType");
    sim.eos.avg_mass = get_or<double>(config, "eos.avg_mass", 1.0);
    if (t == 0) {
        sim.eos.gamma = get_or<double>(config, "eos.gamma", 1.4);
        (void)get_or<double>(config, "eos.ion_frac", 1.0);
        sim.eos.is_constant = true;
    } else if (t == 1) {
        sim.eos.gamma = get_or<double>(config, "eos.gamma", 5.0 / 3.0);
        sim.eos.include_ionisation_energy = get_or<bool>(config, "eos.include_ionisation_energy", false);
        sim.eos.is_constant = false;
    } else if (t == 2) {
        sim.eos.gamma = get_or<double>(config, "eos.gamma", 5.0 / 3.0);
        const std::string table_path = get_or<std::string>(config, "eos.table_path", "mosscap_lte_h_tables.nc");
        sim.eos.table = read_lte_table(table_path);
        sim.eos.tabulated = true;
        sim.eos.is_constant = false;
    } else {
        throw std::runtime_error("eos.type " + type + ": the DexRT equation of state is not supported by this library");
    }
}

// update_eos (Hydro.cpp:325-328) for the analytic LTE EOS: fields of the state
// the last stage started from (which = 1) or of the current state (which = 0).
static void update_eos_fields(Simulation& sim, int which) {
    if (sim.eos.is_constant) return;
    const size_t n = (size_t)sim.state.cells();
    sim.eos.y_space.resize(n);
    sim.eos.T_space.resize(n);
    if (sim.eos.tabulated) {  // the fields the device's update_eos kept
        check(sim, mh_tabulated_eos_fields(sim.handle(), sim.eos.y_space.data(), sim.eos.T_space.data()),
              "update_eos");
        return;
    }
    check(sim, mh_eos_fields(sim.handle(), which, sim.eos.avg_mass, sim.eos.include_ionisation_energy ? 1 : 0,
                             sim.eos.y_space.data(), sim.eos.T_space.data()),
          "update_eos");
}

// setup_output, SetupModules.hpp:179-196.
void setup_output(Simulation& sim, const Config& config) {
    auto& o = sim.out_cfg;
    o.problem_name = get_or<std::string>(config, "problem.name", "circular_explosion");
    o.filename = get_or<std::string>(config, "output.name", "output_" + o.problem_name);
    o.single_file = get_or<bool>(config, "output.single_file", true);
    o.delta = get_or<double>(config, "output.delta_t", 0.1);
    o.output_count = 0;
    o.prev_output_time = -1.0;
    o.variables.conserved = get_or<bool>(config, "output.variables.conserved", true);
    o.variables.primitive = get_or<bool>(config, "output.variables.primitive", false);
    o.variables.fluxes = get_or<bool>(config, "output.variables.fluxes", false);
    o.variables.source = get_or<bool>(config, "output.variables.source", false);
    if (o.variables.fluxes || o.variables.source)
        throw std::runtime_error(
            "output.variables.fluxes / source: the fused stage kernel keeps no flux or source arrays to write");
    if (!sim.write_output) sim.write_output = write_output;
}

void setup_problem(Simulation& sim, const Config& config) {
    const std::string name = get_or<std::string>(config, "problem.name", "circular_explosion");
    get_problem_generator().dispatch(name, sim, config);
}

// The "hip" section (no counterpart in the reference): precision form and
// kernel options of the HIP library.
void setup_device_options(Simulation& sim, const Config& config) {
    static const char* const FormName[] = {"fp64", "fp32c"};
    const std::string form = get_or<std::string>(config, "hip.precision_form", "fp64");
    sim.device.precision_form = find_associated_enum<int>(FormName, 2, form, "precision_form");
    sim.device.span_rows = get_or<int>(config, "hip.span_rows", 0);
    sim.device.rho_floor = get_or<double>(config, "hip.rho_floor", 0.0);
    sim.device.p_floor = get_or<double>(config, "hip.p_floor", 0.0);
}

void setup_gravity(Simulation& sim, const Config& config, const double (&forced)[3]) {
    if (sim.gravity) throw std::runtime_error("Source \"gravity\" already registered.");
    const double dflt[3] = {-1.0, 0.0, 0.0};
    const char* key[3] = {"sources.gravity.x", "sources.gravity.y", "sources.gravity.z"};
    for (int a = 0; a < 3; ++a)
        sim.grav[a] = forced[a] == forced[a] ? forced[a] : get_or<double>(config, key[a], dflt[a]);
    sim.gravity = true;
}

Simulation setup_sim_host(const Config& config) {
    const int num_dim = get_or<int>(config, "simulation.num_dim", 0);
    if (num_dim < 1 || num_dim > 3)
        throw std::runtime_error("simulation.num_dim = " + std::to_string(num_dim) + ", must be in range [1, 3]");
    Simulation sim{};
    sim.num_dim = num_dim;
    setup_grid(sim, config);
    setup_boundaries(sim, config);
    setup_hydro_fns(sim, config);
    setup_timestepper(sim, config);
    setup_eos(sim, config);
    setup_output(sim, config);
    setup_device_options(sim, config);
    setup_problem(sim, config);
    return sim;
}

namespace {

int user_bc_trampoline(mh_ctx*, void* user) {
    auto* h = static_cast<HookHolder*>(user);
    if (!h || !h->sim || !h->sim->user_bc) return 1;
    try {
        h->sim->user_bc(*h->sim);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "user_bc: %s\n", e.what());
        return 1;
    }
    return 0;
}

}  // namespace

void fill_bcs(const Simulation& sim) {
    check(sim, mh_fill_bcs(sim.handle()), "fill_bcs");
}

double compute_dt(const Simulation& sim) { return sim.compute_dt(sim); }

void download_state(Simulation& sim) {
    check(sim, mh_download(sim.handle(), sim.state.Q.data()), "download");
}

// cons_to_prim (Eos.hpp:61-79) on the host copy, every cell, same grouping.
std::vector<double> primitive_state(const Simulation& sim) {
    const int nd = sim.num_dim;
    const int64_t n = sim.state.cells();
    const double* Q = sim.state.Q.data();
    std::vector<double> W(sim.state.Q.size());
    const double gamma = sim.eos.gamma;
    for (int64_t c = 0; c < n; ++c) {
        const double rho = Q[c];
        const double vx = Q[n + c] / rho;
        double v2 = vx * vx;
        W[c] = rho;
        W[n + c] = vx;
        if (nd > 1) {
            const double vy = Q[2 * n + c] / rho;
            W[2 * n + c] = vy;
            v2 += vy * vy;
        }
        if (nd > 2) {
            const double vz = Q[3 * n + c] / rho;
            W[3 * n + c] = vz;
            v2 += vz * vz;
        }
        const double e_kin = 0.5 * rho * v2;
        W[(1 + nd) * n + c] = (gamma - 1.0) * (Q[(1 + nd) * n + c] - e_kin);
    }
    for (int v = 2 + nd; v < sim.state.nvar; ++v)  // tracers: W = Q / rho (Hydro.cpp:27-33)
        for (int64_t c = 0; c < n; ++c) W[v * n + c] = Q[v * n + c] * (1.0 / Q[c]);
    return W;
}

void attach_device(Simulation& sim, bool write_initial_output) {
    mh_grid_size sz{sim.state.sz.xc, sim.state.sz.yc, sim.state.sz.zc, sim.state.sz.ng};
    mh_grid_loc loc{sim.state.loc.x, sim.state.loc.y, sim.state.loc.z};
    mh_boundaries b{};
    const Boundaries& B = sim.state.boundaries;
    b.xs = (int32_t)B.xs;
    b.xe = (int32_t)B.xe;
    b.ys = (int32_t)B.ys;
    b.ye = (int32_t)B.ye;
    b.zs = (int32_t)B.zs;
    b.ze = (int32_t)B.ze;
    std::memcpy(b.xs_const, B.xs_const, sizeof b.xs_const);
    std::memcpy(b.xe_const, B.xe_const, sizeof b.xe_const);
    std::memcpy(b.ys_const, B.ys_const, sizeof b.ys_const);
    std::memcpy(b.ye_const, B.ye_const, sizeof b.ye_const);
    std::memcpy(b.zs_const, B.zs_const, sizeof b.zs_const);
    std::memcpy(b.ze_const, B.ze_const, sizeof b.ze_const);
    mh_schemes sch{(int32_t)sim.scheme.reconstruction, (int32_t)sim.scheme.slope_limit,
                   (int32_t)sim.scheme.riemann_solver, (int32_t)sim.scheme.time_stepper};
    mh_options opt{};
    opt.precision_form = sim.device.precision_form;
    opt.span_rows = sim.device.span_rows;
    opt.rho_floor = (float)sim.device.rho_floor;
    opt.p_floor = (float)sim.device.p_floor;
    mh_ctx* raw = nullptr;
    const int rc = mh_create(&sz, sim.state.dx, &loc, &b, sim.num_dim, sim.state.num_tracers, &sch, sim.eos.gamma,
                             &opt, &raw);
    if (rc != MH_OK) {
        const char* msg = mh_last_error(nullptr);
        throw std::runtime_error(std::string("mh_create: ") + (msg ? msg : "failed"));
    }
    sim.ctx = std::shared_ptr<mh_ctx>(raw, [](mh_ctx* c) { mh_destroy(c); });
    check(sim, mh_upload(raw, sim.state.Q.data()), "upload");
    if (sim.eos.tabulated) {
        const LteTable& t = sim.eos.table;
        const mh_lte_table mt{t.log_eint_rho_start, t.log_eint_rho_step, t.n_eint_rho, t.log_rho_start,
                              t.log_rho_step,      t.n_rho,             t.y_table.data(), t.log_T_table.data()};
        check(sim, mh_set_tabulated_eos(raw, &mt), "tabulated EOS");
        check(sim, mh_tabulated_eos_update(raw), "update_eos");
    }
    // setup_sim runs update_eos on the problem's state before filling ghosts (SetupModules.hpp:250-255)
    update_eos_fields(sim, 0);

    // select_hydro_fns / select_timestepper: the operator slots call the library.
    sim.compute_dt = [](const Simulation& s) {
        double dt = 0.0;
        check(s, mh_compute_dt(s.handle(), s.max_cfl, s.time, s.max_time, s.out_cfg.prev_output_time,
                               s.out_cfg.delta, &dt),
              "compute_dt");
        return dt;
    };
    sim.time_step = [](Simulation& s, double dt) {
        check(s, mh_time_step(s.handle(), dt), "time_step");
        update_eos_fields(s, 1);
        s.dt = dt;
        s.time += dt;  // TimeStepping.cpp:188-189
        s.current_step += 1;
    };
    if (sim.user_bc) {
        check(sim, mh_set_user_bc(raw, user_bc_trampoline, sim.hook.get()), "user_bc");
    }
    if (sim.gravity) {
        check(sim, mh_set_gravity(raw, 1, sim.grav[0], sim.grav[1], sim.grav[2]), "gravity");
    }
    fill_bcs(sim);
    if (write_initial_output) sim.write_output(sim);
}

Simulation setup_sim(const Config& config) {
    Simulation sim = setup_sim_host(config);
    attach_device(sim, true);
    return sim;
}

DriverResult run_driver(Simulation& sim, bool verbose) {
    DriverResult r;
    fill_bcs(sim);
    while (sim.time < sim.max_time && (sim.max_steps < 0 || sim.current_step < sim.max_steps)) {
        const double dt = compute_dt(sim);
        sim.time_step(sim, dt);
        if (sim.time >= sim.out_cfg.prev_output_time + sim.out_cfg.delta) {
            sim.write_output(sim);
            r.outputs += 1;
            if (verbose) std::printf("t = %.3f, dt = %.3e, iter = %lld\n", sim.time, dt, (long long)sim.current_step);
        }
    }
    sim.write_output(sim);
    r.outputs += 1;
    r.steps = sim.current_step;
    r.time = sim.time;
    return r;
}

}  // namespace mh::host
