// Problem registry and the initial-condition generators of the workloads:
// shock_tube_2d (ShockTube2d.cpp:10-49) and the seeded kh_band and sedov
// (2-D and 3-D) setups. Every cell, ghosts included, is set on the host in FP64 with
// prim_to_cons's grouping (Eos.hpp:81-99); the driver fills ghosts after.
#include <array>
#include <cmath>

#include "config.hpp"
#include "simulation.hpp"

namespace mh::host {

DispatchFactory<Simulation>& get_problem_generator() {
    static DispatchFactory<Simulation> factory;
    return factory;
}

uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double uniform01(uint64_t seed, uint64_t counter) {
    return (double)(splitmix64(seed + counter) >> 11) * (1.0 / 9007199254740992.0);
}

namespace {

// prim_to_cons (Eos.hpp:81-99) into cell (j, i) of a 2-D state.
void set_prim_2d(State& s, double gamma, int j, int i, double rho, double vx, double vy, double p) {
    s.q(0, 0, j, i) = rho;
    s.q(1, 0, j, i) = rho * vx;
    double v2 = vx * vx;
    s.q(2, 0, j, i) = rho * vy;
    v2 += vy * vy;
    const double e_kin = 0.5 * rho * v2;
    const double e_int = p / (gamma - 1.0);
    s.q(3, 0, j, i) = e_int + e_kin;
}

// prim_to_cons (Eos.hpp:81-99) into cell (k, j, i) of a 3-D state at rest.
void set_rest_3d(State& s, double gamma, int k, int j, int i, double rho, double p) {
    s.q(0, k, j, i) = rho;
    for (int a = 1; a <= 3; ++a) s.q(a, k, j, i) = rho * 0.0;
    s.q(4, k, j, i) = p / (gamma - 1.0) + 0.5 * rho * 0.0;
}

void require_2d(const Simulation& sim, const char* name) {
    if (sim.num_dim != 2) throw std::runtime_error(std::string(name) + " only handles 2d problems");
}

// ShockTube2d.cpp:10-49: (rho, p) = (1, 1) left of x = 0.5, (0.125, 0.1) right.
void config_shock_tube_2d(Simulation& sim, const Config& config) {
    require_2d(sim, "shock_tube_2d");
    State& s = sim.state;
    for (int j = 0; j < s.sz.yc; ++j)
        for (int i = 0; i < s.sz.xc; ++i) {
            const bool left = s.pos_x(i) < 0.5;
            set_prim_2d(s, sim.eos.gamma, j, i, left ? 1.0 : 0.125, 0.0, 0.0, left ? 1.0 : 0.1);
        }
    sim.max_time = get_or<double>(config, "timestep.max_time", 0.2);
}

// 1-D Riemann problem on the grid (no counterpart in the reference's
// registry): primitive (rho, vx, p) = problem.left for x < problem.interface,
// problem.right elsewhere; defaults are Sod's states at x = 0.5.
void config_shock_tube_1d(Simulation& sim, const Config& config) {
    if (sim.num_dim != 1) throw std::runtime_error("shock_tube_1d only handles 1d problems");
    auto state_of = [&](const char* key, double r, double v, double p) {
        double w[3] = {r, v, p};
        const YNode* n = find_node(config.get(), std::string("problem.") + key);
        if (n && n->kind == YNode::Seq) {
            if (n->seq.size() != 3) throw std::runtime_error(std::string("problem.") + key + " needs [rho, vx, p]");
            for (int k = 0; k < 3; ++k)
                if (n->seq[k]->kind != YNode::Scalar || !convert_scalar(n->seq[k]->scalar, w[k]))
                    throw std::runtime_error(std::string("problem.") + key + " needs [rho, vx, p]");
        }
        return std::array<double, 3>{w[0], w[1], w[2]};
    };
    const auto L = state_of("left", 1.0, 0.0, 1.0);
    const auto R = state_of("right", 0.125, 0.0, 0.1);
    const double x0 = get_or<double>(config, "problem.interface", 0.5);
    State& s = sim.state;
    const double gamma = sim.eos.gamma;
    for (int i = 0; i < s.sz.xc; ++i) {
        const auto& w = s.pos_x(i) < x0 ? L : R;
        s.q(0, 0, 0, i) = w[0];
        s.q(1, 0, 0, i) = w[0] * w[1];
        const double v2 = w[1] * w[1];
        const double e_kin = 0.5 * w[0] * v2;
        const double e_int = w[2] / (gamma - 1.0);
        s.q(2, 0, 0, i) = e_int + e_kin;
    }
}

// Kelvin-Helmholtz shear band on [0,1]^2: p = 2.5; |y - 0.5| < 0.25 has
// rho = 2, vx = +0.5, elsewhere rho = 1, vx = -0.5; vx and vy get
// U(-amp, amp) from the per-cell counters 2*(j*xc + i) and 2*(j*xc + i) + 1
// of splitmix64(seed + counter). Tracer fields are a dye of mass fraction 1
// inside the band. Other problems leave tracer densities at zero.
void config_kh_band(Simulation& sim, const Config& config) {
    require_2d(sim, "kh_band");
    const uint64_t seed = (uint64_t)get_or<long long>(config, "problem.seed", 20261202LL);
    const double amp = get_or<double>(config, "problem.amplitude", 0.01);
    State& s = sim.state;
    for (int j = 0; j < s.sz.yc; ++j) {
        const double y = s.pos_y(j);
        const bool band = std::fabs(y - 0.5) < 0.25;
        for (int i = 0; i < s.sz.xc; ++i) {
            const uint64_t idx = (uint64_t)j * s.sz.xc + i;
            const double vx = (band ? 0.5 : -0.5) + amp * (2.0 * uniform01(seed, 2 * idx) - 1.0);
            const double vy = amp * (2.0 * uniform01(seed, 2 * idx + 1) - 1.0);
            set_prim_2d(s, sim.eos.gamma, j, i, band ? 2.0 : 1.0, vx, vy, 2.5);
            for (int k = 0; k < s.num_tracers; ++k)  // passive dye marking the band
                s.q(4 + k, 0, j, i) = (band ? 2.0 : 1.0) * (band ? 1.0 : 0.0);
        }
    }
}

// Sedov blast: rho = 1, v = 0, p = p_ambient, and the energy E spread
// uniformly over the N interior cells whose centre lies within
// r = radius_cells*dx of the origin, p = (gamma - 1) E / (N dx^d) there in d =
// 2 or 3 dimensions (ghost cells inside the radius get the same pressure).
void config_sedov(Simulation& sim, const Config& config) {
    if (sim.num_dim < 2) throw std::runtime_error("sedov only handles 2d and 3d problems");
    const double energy = get_or<double>(config, "problem.energy", 1.0);
    const double radius_cells = get_or<double>(config, "problem.radius_cells", 3.5);
    const double p_ambient = get_or<double>(config, "problem.p_ambient", 1e-5);
    State& s = sim.state;
    const double r = radius_cells * s.dx;
    const double r2 = r * r;
    const int g = s.sz.ng;
    if (sim.num_dim == 3) {  // a sphere: p = (gamma - 1) E / (N dx^3)
        auto inside3 = [&](int k, int j, int i) {
            const double x = s.pos_x(i), y = s.pos_y(j), z = s.pos_z(k);
            return (x * x + y * y + z * z) < r2;
        };
        long long n3 = 0;
        for (int k = g; k < s.sz.zc - g; ++k)
            for (int j = g; j < s.sz.yc - g; ++j)
                for (int i = g; i < s.sz.xc - g; ++i) n3 += inside3(k, j, i) ? 1 : 0;
        if (n3 == 0) throw std::runtime_error("sedov: no cell centre lies within problem.radius_cells");
        const double p3 = (sim.eos.gamma - 1.0) * energy / (n3 * s.dx * s.dx * s.dx);
        for (int k = 0; k < s.sz.zc; ++k)
            for (int j = 0; j < s.sz.yc; ++j)
                for (int i = 0; i < s.sz.xc; ++i)
                    set_rest_3d(s, sim.eos.gamma, k, j, i, 1.0, inside3(k, j, i) ? p3 : p_ambient);
        return;
    }
    auto inside = [&](int j, int i) {
        const double x = s.pos_x(i), y = s.pos_y(j);
        return (x * x + y * y) < r2;
    };
    long long n_in = 0;
    for (int j = g; j < s.sz.yc - g; ++j)
        for (int i = g; i < s.sz.xc - g; ++i) n_in += inside(j, i) ? 1 : 0;
    if (n_in == 0) throw std::runtime_error("sedov: no cell centre lies within problem.radius_cells");
    const double p_in = (sim.eos.gamma - 1.0) * energy / (n_in * s.dx * s.dx);
    for (int j = 0; j < s.sz.yc; ++j)
        for (int i = 0; i < s.sz.xc; ++i) set_prim_2d(s, sim.eos.gamma, j, i, 1.0, 0.0, 0.0, inside(j, i) ? p_in : p_ambient);
}

// The reference's gravity-driven shear layer (KelvinHelmholtz.cpp:11-70):
// rho = 2 where the coordinate along problem.axis is positive, else 1;
// p = 2.5 + rho g x_axis; a cosine bump in the velocity along the axis; and
// a gravity source. The reference writes -0.1 into sources.gravity.<axis>
// unconditionally (set_if_not_present assigns, MosscapConfig.hpp:91-103) and
// setup_gravity reads the other components with its defaults (x: -1), so with
// axis 1 the gravity is (-1, -0.1) unless sources.gravity.x is given; both
// are kept here.
void config_kelvin_helmholtz(Simulation& sim, const Config& config) {
    require_2d(sim, "kelvin_helmholtz");
    const int axis = get_or<int>(config, "problem.axis", 1);
    const double grav = -0.1;
    State& s = sim.state;
    const int nx = s.sz.xc - 2 * s.sz.ng, ny = s.sz.yc - 2 * s.sz.ng;
    if (axis == 0 && nx / ny != 3) throw std::runtime_error("Aspect ratio should be 3 (x / y) == 3 for axis = 0");
    if (axis != 0 && ny / nx != 3) throw std::runtime_error("Aspect ratio should be 3 (y / x) == 3 for axis = 1");
    const int other = axis == 0 ? 1 : 0;
    const int iv = axis == 0 ? 1 : 2;
    auto axis_length = [&](int a) { return (a == 0 ? nx : ny) * s.dx; };  // State::get_axis_length
    const double L_axis = axis_length(axis), L_other = axis_length(other);
    for (int j = 0; j < s.sz.yc; ++j)
        for (int i = 0; i < s.sz.xc; ++i) {
            const double pos[2] = {s.pos_x(i), s.pos_y(j)};
            const double pres_0 = 2.5;
            double w[4] = {0.0, 0.0, 0.0, 0.0};
            w[0] = pos[axis] > 0.0 ? 2.0 : 1.0;
            w[3] = pres_0 + w[0] * grav * pos[axis];
            w[iv] = 0.01 * (1.0 + std::cos(2.0 * M_PI / L_other * pos[other])) *
                    (1.0 + std::cos(2.0 * M_PI / L_axis * pos[axis])) * 0.25;
            set_prim_2d(s, sim.eos.gamma, j, i, w[0], w[1], w[2], w[3]);
        }
    double forced[3] = {NAN, NAN, NAN};
    forced[axis == 0 ? 0 : 1] = grav;
    setup_gravity(sim, config, forced);
}

const bool registered = [] {
    auto& f = get_problem_generator();
    f.register_method("shock_tube_2d", config_shock_tube_2d);
    f.register_method("shock_tube_1d", config_shock_tube_1d);
    f.register_method("kh_band", config_kh_band);
    f.register_method("sedov", config_sedov);
    f.register_method("kelvin_helmholtz", config_kelvin_helmholtz);
    return true;
}();

}  // namespace

}  // namespace mh::host
