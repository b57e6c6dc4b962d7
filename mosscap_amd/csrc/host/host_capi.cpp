// C entry points of the host layer (libmosscap_host.so): run a configuration
// through setup and the driver loop, and expose the configuration semantics
// to scripts. None of these throw; failures return non-zero with the message
// in `err`.
#include <cstdio>
#include <cstring>
#include <sstream>

#include "../../../include/mosscap_host.h"
#include "config.hpp"
#include "simulation.hpp"

using namespace mh::host;

namespace {

int fail(char* err, int64_t len, const std::string& msg) {
    if (err && len > 0) {
        std::strncpy(err, msg.c_str(), (size_t)len - 1);
        err[len - 1] = '\0';
    }
    return 1;
}

int put(char* out, int64_t len, const std::string& s) {
    if (!out || len <= (int64_t)s.size()) return 2;
    std::memcpy(out, s.c_str(), s.size() + 1);
    return 0;
}

std::string json_str(const std::string& s) {
    std::string o = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o + "\"";
}

std::string num(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.17g", v);
    return b;
}

std::string summary(const Simulation& sim) {
    const auto& s = sim.state;
    const auto& b = s.boundaries;
    std::ostringstream o;
    auto consts = [&](const double* a) {
        std::string r = "[";
        for (int i = 0; i < 5; ++i) r += (i ? "," : "") + num(a[i]);
        return r + "]";
    };
    o << "{\"num_dim\":" << sim.num_dim << ",\"xc\":" << s.sz.xc << ",\"yc\":" << s.sz.yc << ",\"zc\":" << s.sz.zc
      << ",\"ng\":" << s.sz.ng << ",\"dx\":" << num(s.dx) << ",\"loc\":[" << num(s.loc.x) << "," << num(s.loc.y)
      << "," << num(s.loc.z) << "],\"num_tracers\":" << s.num_tracers << ",\"bc\":[" << (int)b.xs << ","
      << (int)b.xe << "," << (int)b.ys << "," << (int)b.ye << "," << (int)b.zs << "," << (int)b.ze
      << "],\"bc_const\":[" << consts(b.xs_const) << "," << consts(b.xe_const) << "," << consts(b.ys_const) << ","
      << consts(b.ye_const) << "," << consts(b.zs_const) << "," << consts(b.ze_const) << "]"
      << ",\"reconstruction\":" << (int)sim.scheme.reconstruction << ",\"slope_limit\":" << (int)sim.scheme.slope_limit
      << ",\"riemann_solver\":" << (int)sim.scheme.riemann_solver << ",\"time_stepper\":"
      << (int)sim.scheme.time_stepper << ",\"max_cfl\":" << num(sim.max_cfl) << ",\"max_time\":" << num(sim.max_time)
      << ",\"max_steps\":" << sim.max_steps << ",\"gamma\":" << num(sim.eos.gamma)
      << ",\"problem\":" << json_str(sim.out_cfg.problem_name) << ",\"output_name\":" << json_str(sim.out_cfg.filename)
      << ",\"single_file\":" << (sim.out_cfg.single_file ? "true" : "false") << ",\"delta\":" << num(sim.out_cfg.delta)
      << ",\"primitive\":" << (sim.out_cfg.variables.primitive ? "true" : "false")
      << ",\"precision_form\":" << sim.device.precision_form << ",\"span_rows\":" << sim.device.span_rows
      << ",\"rho_floor\":" << num(sim.device.rho_floor) << ",\"p_floor\":" << num(sim.device.p_floor)
      << ",\"gravity\":" << (sim.gravity ? "[" + num(sim.grav[0]) + "," + num(sim.grav[1]) + "," + num(sim.grav[2]) + "]"
                                              : std::string("null"));
    if (sim.eos.tabulated) {  // the table as read: axes and the sums of both tables
        const LteTable& t = sim.eos.table;
        double ys = 0.0, ts = 0.0;
        for (double v : t.y_table) ys += v;
        for (double v : t.log_T_table) ts += v;
        o << ",\"lte_table\":{\"log_eint_rho_start\":" << num(t.log_eint_rho_start)
          << ",\"log_eint_rho_step\":" << num(t.log_eint_rho_step) << ",\"n_eint_rho\":" << t.n_eint_rho
          << ",\"log_rho_start\":" << num(t.log_rho_start) << ",\"log_rho_step\":" << num(t.log_rho_step)
          << ",\"n_rho\":" << t.n_rho << ",\"y_sum\":" << num(ys) << ",\"log_T_sum\":" << num(ts) << "}";
    }
    o << "}";
    return o.str();
}

}  // namespace

extern "C" {

int mhh_setup_summary(const char* yaml_text, char* out, int64_t out_len, char* err, int64_t err_len) {
    try {
        const Simulation sim = setup_sim_host(parse_yaml(yaml_text ? yaml_text : ""));
        return put(out, out_len, summary(sim));
    } catch (const std::exception& e) {
        return fail(err, err_len, e.what());
    }
}

int mhh_initial_state(const char* yaml_text, double* q, int64_t n, char* err, int64_t err_len) {
    try {
        const Simulation sim = setup_sim_host(parse_yaml(yaml_text ? yaml_text : ""));
        if (n != (int64_t)sim.state.Q.size())
            return fail(err, err_len, "state has " + std::to_string(sim.state.Q.size()) + " values");
        std::memcpy(q, sim.state.Q.data(), sim.state.Q.size() * sizeof(double));
        return 0;
    } catch (const std::exception& e) {
        return fail(err, err_len, e.what());
    }
}

int mhh_get(const char* yaml_text, const char* key, int32_t type, const char* default_text, char* out,
            int64_t out_len, char* err, int64_t err_len) {
    try {
        const Config c = parse_yaml(yaml_text ? yaml_text : "");
        const std::string d = default_text ? default_text : "";
        switch (type) {
            case 0: {
                int dv = 0;
                convert_scalar(d, dv);
                return put(out, out_len, std::to_string(get_or<int>(c, key, dv)));
            }
            case 1: {
                double dv = 0;
                convert_scalar(d, dv);
                return put(out, out_len, num(get_or<double>(c, key, dv)));
            }
            case 2: {
                bool dv = false;
                convert_scalar(d, dv);
                return put(out, out_len, get_or<bool>(c, key, dv) ? "true" : "false");
            }
            default:
                return put(out, out_len, get_or<std::string>(c, key, d));
        }
    } catch (const std::exception& e) {
        return fail(err, err_len, e.what());
    }
}

int mhh_write_initial_output(const char* yaml_text, const char* directory, char* err, int64_t err_len) {
    try {
        Simulation sim = setup_sim_host(parse_yaml(yaml_text ? yaml_text : ""));
        sim.out_cfg.directory = directory ? directory : ".";
        sim.write_output(sim);
        sim.time = 0.5;  // a second record, same state
        sim.write_output(sim);
        return 0;
    } catch (const std::exception& e) {
        return fail(err, err_len, e.what());
    }
}

int mhh_run_config(const char* yaml_text, const char* directory, int32_t verbose, mhh_run_result* res, char* err,
                   int64_t err_len) {
    try {
        Simulation sim = setup_sim_host(parse_yaml(yaml_text ? yaml_text : ""));
        sim.out_cfg.directory = directory ? directory : ".";
        attach_device(sim, true);
        const DriverResult r = run_driver(sim, verbose != 0);
        if (res) {
            res->steps = r.steps;
            res->time = r.time;
            res->outputs = r.outputs + 1;
            res->last_dt = sim.dt;
        }
        return 0;
    } catch (const std::exception& e) {
        return fail(err, err_len, e.what());
    }
}

}  // extern "C"
