// ref_harness.cpp -- test-only driver around the reference's own chicken sources.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/build_ref.sh into oracle/_ref/ and used by the
// golden-vector generator (tests/golden/make_golden.py) and as the "reference" CPU baseline.
//
// The reference's pointwise and block-level numerics (number/vector3/gas/flow/face/cell/
// config/block/bcs .cu under src/chicken/) are #included unmodified from the reference
// tree and compiled as HIP host code: HIP defines the __host__/__device__ qualifiers those
// files use, so no CUDA stand-in is involved.  simulate.cu and main.cu are not compiled:
// they call the CUDA runtime API unconditionally.  This file therefore supplies the two
// globals bcs.cu expects (simulate.cu:42-46) and restates, in its own code, the driver
// sequence of simulate.cu:55-155 (initialise), :158-179 (write), :182-342 (CPU march)
// and :768-792 (finalise), calling the reference's Block/BC functions for all arithmetic.
//
// Built per gas model and number type as the reference's makefile builds them (GAS_MODEL,
// FLOAT_NUMBERS); the time and time step are the reference's number type, as SimState holds
// them (simulate.cu:32-39).
//
// Modes:
//   chkn_ref run  <job> <tindx> <nsteps|-1> <dtlog|-> [text|textin]
//                               march like chkn-run --binary (with "text": like chkn-run
//                               without --binary, gzipped text files; with "textin": read
//                               gzipped text, write binary flow files -- the FP32 build
//                               reads text only, block.cu:794-796)
//   chkn_ref flux <in> <out>    flux probe: records of 4x9 states + 9 frame doubles
//                               + (x_order, flux_calc) as doubles -> 6 fluxes each
//   chkn_ref geom <in> <out>    geometry probe: 8 hex vertices -> pos, vol, lengths;
//                               then 4 quad vertices (first four) -> quad properties
#include <vector>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <iomanip>
#include <chrono>

#include "config.cu"
#include "block.cu"

std::vector<Block> fluidBlocks;
std::vector<BConfig> blk_configs;

#include "bcs.cu"

namespace H {
number dt = 0.0, t = 0.0, t_plot = 0.0;
int step = 0, next_plot_indx = 1, steps_since_last_plot = 0;
}

static void init_sim(int tindx, bool binary_in)
{
    blk_configs = read_config_file(Config::job + "/config.json");
    char buf[256];
    for (int id = 0; id < Config::nFluidBlocks; ++id) {
        BConfig& cfg = blk_configs[id];
        int i = cfg.i, j = cfg.j, k = cfg.k;
        if (id != Config::blk_ids[i][j][k]) throw runtime_error("inconsistent blk_id");
        cfg.fill_in_dimensions(Config::nics[i], Config::njcs[j], Config::nkcs[k],
                               Config::threads_per_GPUblock);
        Block blk;
        blk.configure(cfg);
        snprintf(buf, sizeof buf, "/grid/grid-%04d-%04d-%04d", i, j, k);
        blk.readGrid(cfg, Config::job + buf + (binary_in ? ".bin" : ".gz"), binary_in);
        snprintf(buf, sizeof buf, "/flow/t%04d/flow-%04d-%04d-%04d", tindx, i, j, k);
        blk.readFlow(cfg, Config::job + buf + (binary_in ? ".bin" : ".gz"), binary_in);
        blk.computeGeometry(cfg);
        fluidBlocks.push_back(blk);
    }
    configure_exchange_info(fluidBlocks, blk_configs);
    H::t = 0.0;
    std::ifstream tf(Config::job + "/times.data");
    std::string line;
    while (std::getline(tf, line)) {
        if (line.empty() || line.find('#') != std::string::npos) continue;
        std::stringstream ss(line);
        int ti; number tm;
        ss >> ti >> tm;
        if (ti == tindx) { H::t = tm; H::next_plot_indx = ti + 1; }
    }
    H::t_plot = H::t + Config::dt_plot_schedule.get_value(H::t);
    H::steps_since_last_plot = 0;
}

static void write_flow(int tindx, number tme, bool binary)
{
    char buf[512];
    snprintf(buf, sizeof buf, "%s/flow/t%04d", Config::job.c_str(), tindx);
    std::string dir(buf);
    if (!filesystem::exists(dir)) filesystem::create_directories(dir);
    for (int id = 0; id < Config::nFluidBlocks; ++id) {
        BConfig& c = blk_configs[id];
        snprintf(buf, sizeof buf, "%s/flow-%04d-%04d-%04d", dir.c_str(), c.i, c.j, c.k);
        fluidBlocks[id].writeFlow(c, std::string(buf) + (binary ? ".bin" : ".gz"), binary);
    }
    std::ofstream tf(Config::job + "/times.data", std::ofstream::app);
    tf << tindx << " " << tme << std::endl;
}

static int stage(int which)
{
    apply_boundary_conditions_for_convective_fluxes();
    int nb = Config::nFluidBlocks;
    std::vector<int> bad(nb, 0);
#pragma omp parallel for schedule(dynamic)
    for (int ib = 0; ib < nb; ++ib) {
        BConfig& cfg = blk_configs[ib];
        Block& blk = fluidBlocks[ib];
        if (!cfg.active) continue;
        blk.calculate_convective_fluxes(Config::flux_calc, Config::x_order);
        if (Config::viscous) { // simulate.cu:235-238
            apply_viscous_boundary_conditions(blk, cfg);
            blk.add_viscous_fluxes();
        }
        if (which == 1) bad[ib] = blk.update_stage_1(cfg, H::dt);
        else if (which == 2) bad[ib] = blk.update_stage_2(cfg, H::dt);
        else bad[ib] = blk.update_stage_3(cfg, H::dt);
    }
    int total = 0;
    for (int b : bad) total += b;
    return total;
}

static int run(const char* job, int tindx, int nsteps, const char* dtlog, bool binary_in,
               bool binary)
{
    Config::job = job;
    init_sim(tindx, binary_in);
    H::dt = Config::dt_init;
    H::step = 0;
    int nb = Config::nFluidBlocks;
#pragma omp parallel for
    for (int ib = 0; ib < nb; ++ib)
        if (blk_configs[ib].active) {
            fluidBlocks[ib].encodeConserved(blk_configs[ib], 0);
            if (Config::viscous) fluidBlocks[ib].setup_LSQ_arrays(); // simulate.cu:201
        }
    FILE* fdt = (dtlog && std::string(dtlog) != "-") ? std::fopen(dtlog, "w") : nullptr;
    auto t_start = std::chrono::steady_clock::now();
    int done = 0;
    while (H::step < Config::max_step && H::t < Config::max_time && (nsteps < 0 || done < nsteps)) {
        if (H::step > 0 && (H::step % Config::cfl_count) == 0) {
            number smallest = numeric_limits<number>::max();
            std::vector<number> adt(nb, smallest);
            number cfl = Config::cfl_schedule.get_value(H::t);
#pragma omp parallel for
            for (int ib = 0; ib < nb; ++ib)
                if (blk_configs[ib].active) adt[ib] = fluidBlocks[ib].estimate_allowed_dt(blk_configs[ib], cfl);
            for (number a : adt) smallest = fmin(smallest, a);
            H::dt = fmin((number)1.5 * H::dt, smallest);
        }
        for (int s = 1; s <= 3; ++s) {
            int bad = stage(s);
            if (bad > 0) {
                std::cerr << "Stage " << s << " bad cell count: " << bad << std::endl;
                if (fdt) std::fclose(fdt);
                return 10 + s;
            }
        }
#pragma omp parallel for
        for (int ib = 0; ib < nb; ++ib)
            if (blk_configs[ib].active) fluidBlocks[ib].copy_conserved_data(blk_configs[ib], 1, 0);
        if (Config::reacting) { // simulate.cu:296-303
#pragma omp parallel for
            for (int ib = 0; ib < nb; ++ib)
                if (blk_configs[ib].active) fluidBlocks[ib].update_chemistry(blk_configs[ib], H::dt);
        }
        if (fdt) std::fprintf(fdt, "%d %.17g\n", H::step, (double)H::dt);
        H::t += H::dt;
        H::step += 1;
        H::steps_since_last_plot += 1;
        done += 1;
        if (H::t >= H::t_plot) {
            write_flow(H::next_plot_indx, H::t, binary);
            H::steps_since_last_plot = 0;
            H::next_plot_indx += 1;
            H::t_plot = H::t + Config::dt_plot_schedule.get_value(H::t);
        }
    }
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    if (H::steps_since_last_plot > 0) write_flow(H::next_plot_indx, H::t, binary);
    if (fdt) std::fclose(fdt);
    long long cells = 0;
    for (auto& c : blk_configs) if (c.active) cells += (long long)c.nic * c.njc * c.nkc;
    std::printf("steps=%d t=%.17g dt=%.17g cells=%lld march_seconds=%.6f\n", H::step, (double)H::t,
                (double)H::dt, cells, secs);
    return 0;
}

static std::vector<double> read_all(const char* path)
{
    std::ifstream f(path, std::ios::binary);
    std::vector<char> raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::vector<double> v(raw.size() / sizeof(double));
    std::memcpy(v.data(), raw.data(), v.size() * sizeof(double));
    return v;
}

static void write_all(const char* path, const std::vector<double>& v)
{
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(double));
}

static FlowState fs9(const double* w)
{
    FlowState s;
    s.gas.rho = w[0]; s.gas.e = w[1]; s.gas.p = w[2]; s.gas.T = w[3]; s.gas.YB = w[4]; s.gas.a = w[5];
    s.vel.set(w[6], w[7], w[8]);
    return s;
}

static int flux_probe(const char* in, const char* out)
{
    std::vector<double> v = read_all(in);
    const size_t rec = 36 + 9 + 2;
    size_t n = v.size() / rec;
    std::vector<double> res(6 * n);
    for (size_t r = 0; r < n; ++r) {
        const double* p = &v[r * rec];
        FVFace f;
        f.n.set(p[36], p[37], p[38]); f.t1.set(p[39], p[40], p[41]); f.t2.set(p[42], p[43], p[44]);
        FlowState a = fs9(p), b = fs9(p + 9), c = fs9(p + 18), d = fs9(p + 27);
        f.calculate_convective_flux(a, b, c, d, int(p[46]), int(p[45]));
        for (int q = 0; q < 6; ++q) res[6 * r + q] = f.F[q];
    }
    write_all(out, res);
    return 0;
}

static int geom_probe(const char* in, const char* out)
{
    std::vector<double> v = read_all(in);
    size_t n = v.size() / 24;
    std::vector<double> res(20 * n);
    for (size_t r = 0; r < n; ++r) {
        const double* p = &v[r * 24];
        Vector3 q[8];
        for (int i = 0; i < 8; ++i) q[i].set(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
        Vector3 c; number vol, iL, jL, kL;
        hex_cell_properties(q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], false, c, vol, iL, jL, kL);
        double* o = &res[20 * r];
        o[0] = c.x; o[1] = c.y; o[2] = c.z; o[3] = vol; o[4] = iL; o[5] = jL; o[6] = kL;
        Vector3 fc, n_, t1, t2; number area;
        quad_properties(q[0], q[1], q[2], q[3], fc, n_, t1, t2, area);
        o[7] = fc.x; o[8] = fc.y; o[9] = fc.z; o[10] = n_.x; o[11] = n_.y; o[12] = n_.z;
        o[13] = t1.x; o[14] = t1.y; o[15] = t1.z; o[16] = t2.x; o[17] = t2.y; o[18] = t2.z;
        o[19] = area;
    }
    write_all(out, res);
    return 0;
}

int main(int argc, char** argv)
{
    if (argc < 2) { std::cerr << "usage: chkn_ref run|flux|geom ..." << std::endl; return 2; }
    std::string mode = argv[1];
    try {
        if (mode == "run" && argc >= 6) {
            const std::string io = argc >= 7 ? argv[6] : "";
            return run(argv[2], std::atoi(argv[3]), std::atoi(argv[4]), argv[5],
                       io != "text" && io != "textin", io != "text");
        }
        if (mode == "flux" && argc >= 4) return flux_probe(argv[2], argv[3]);
        if (mode == "geom" && argc >= 4) return geom_probe(argv[2], argv[3]);
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 3;
    }
    std::cerr << "bad arguments" << std::endl;
    return 2;
}
