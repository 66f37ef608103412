// chkn_io.h -- chicken job-directory formats: config.json, grid and flow files, times.data.
//
// Formats follow the reference reader/writers: read_config_file (config.cu:347-525),
// Block::readGrid / readFlow / writeFlow (block.cu:785-976), and the times.data handling of
// initialize_simulation / write_flow_data (simulate.cu:134-149, :158-179).
#pragma once

#include <string>
#include <vector>

#include "chkn_real.h"

namespace chkn {

struct JobBlock {
    int i = 0, j = 0, k = 0;
    int nic = 0, njc = 0, nkc = 0;
    bool active = true;
    int bc[6] = {0, 0, 0, 0, 0, 0};
    int bc_fs[6] = {0, 0, 0, 0, 0, 0};
    double bc_twall[6] = {0, 0, 0, 0, 0, 0};  // fixed-T walls; NaN when TWall is a function
    int bc_fun[6] = {0, 0, 0, 0, 0, 0};       // inflow_function sides: face.cu:51-59 index
    std::string label;
};

struct JobConfig {
    std::string title;
    int nib = 0, njb = 0, nkb = 0;
    std::vector<int> blk_ids;  // (i*njb + j)*nkb + k
    std::vector<JobBlock> blocks;
    std::vector<double> flow_states;  // 6 per state: p, T, YB, vx, vy, vz
    int x_order = 2, t_order = 2, flux_calc = 0, cfl_count = 10, max_step = 100, print_count = 20;
    double max_time = 1.0e-3, dt_init = 1.0e-6;
    std::vector<double> cfl_times, cfl_values, dt_plot_times, dt_plot_values;
    bool reacting = false, viscous = false;
    int source_terms = 0;
    int time_scheme = 0;  // "time_scheme": "rk3" (0) or "predictor_corrector" (1); not a reference key
    std::vector<double> bl_profile;  // p_e, then (y, T, velx) x 12, when a side uses it
    std::vector<double> gas_model;   // gamma, R, q, alpha, Ti, Cv
};

// The job's gas model (gamma, R, q, alpha, Ti, Cv) is the built one, to 1e-4 absolute as
// config.cu:369-386 checks it.
bool gas_model_matches(const double* model);

// Throw std::runtime_error with the reference's message text on failure.  Numbers are
// returned as double holding the value the reference reads into its number type: text is
// parsed as real (float in the FP32 flavour, whose binary readers refuse, block.cu:794-796,
// :887-889), configuration values are rounded to real where they are used.
JobConfig read_config_json(const std::string& path);
// Vertices (nic+1)(njc+1)(nkc+1) xyz triples, i fastest.
std::vector<double> read_grid(const std::string& path, bool binary, int nic, int njc, int nkc);
// 13*ncells IO variables, variable-major.
std::vector<double> read_flow(const std::string& path, bool binary, int ncells);
// Binary files hold the real type's bytes; text is %g (6 significant digits).
void write_flow(const std::string& path, bool binary, const real* iovars, int ncells);
// Start time and next plot index for time index tindx from times.data (t = 0 if absent).
void read_times(const std::string& job, int tindx, double& t, int& next_plot_indx);
void append_times(const std::string& job, int tindx, real t);
// Laminar boundary-layer profile file (config.cu:37-60): p_e, then y, T, velx of 12 samples.
std::vector<double> read_bl_profile(const std::string& path);
// TWall of every face of a fixed-T side given as a function: <job>/TWall-i-j-k-<face>.gz
// (block.cu:566-611), nfaces values in the file's order.
std::vector<double> read_twall(const std::string& job, const JobBlock& b, int side, int nfaces);
std::string block_file(const std::string& job, const char* kind, int tindx, int i, int j, int k,
                       bool binary);

}  // namespace chkn
