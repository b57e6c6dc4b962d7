/*
 * mosscap_host.h - C entry points of the host layer (libmosscap_host.so).
 *
 * The host layer is the reference's setup and driver restated over the HIP
 * library: YAML configuration with get_or defaults (MosscapConfig.hpp:62-123,
 * SetupModules.hpp:9-262), the problem registry (ProblemGenerator.hpp:12-53),
 * the Simulation operator slots bound to mosscap_hip.h, netCDF output
 * (Output.cpp:88-278) and the main loop (main.cpp:50-73). The C++ interface
 * is mosscap_amd/csrc/host/simulation.hpp; these functions serve scripts and
 * tests. All take the configuration as YAML text, return 0 on success and
 * otherwise write a message into err (err_len bytes including the NUL).
 */
#ifndef MOSSCAP_HOST_H
#define MOSSCAP_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int64_t steps;
    double time;
    int32_t outputs;
    double last_dt;
} mhh_run_result;

/* Host setup only (no device): JSON summary of the resolved configuration. */
int mhh_setup_summary(const char* yaml_text, char* out, int64_t out_len, char* err, int64_t err_len);
/* The problem generator's initial state (n values, layout of Q, ghosts included). */
int mhh_initial_state(const char* yaml_text, double* q, int64_t n, char* err, int64_t err_len);
/* get_or<T>(config, key, default): type 0 int, 1 double, 2 bool, 3 string; result as text. */
int mhh_get(const char* yaml_text, const char* key, int32_t type, const char* default_text, char* out,
            int64_t out_len, char* err, int64_t err_len);
/* Writes the initial state as output (two records at t = 0 and 0.5) into directory; no device. */
int mhh_write_initial_output(const char* yaml_text, const char* directory, char* err, int64_t err_len);
/* setup_sim + the driver loop on the GPU, output files into directory. */
int mhh_run_config(const char* yaml_text, const char* directory, int32_t verbose, mhh_run_result* res, char* err,
                   int64_t err_len);

#ifdef __cplusplus
}
#endif

#endif
