// chkn_io.cpp -- readers and writers for chicken job directories (see chkn_io.h).
#include "chkn_io.h"

#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "chkn_json.h"
#include "chkn_physics.h"

namespace chkn {

namespace {

std::string slurp(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("Could not open ifstream for " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// Whole gzip file as text.
std::string slurp_gz(const std::string& path)
{
    gzFile g = gzopen(path.c_str(), "rb");
    if (!g) throw std::runtime_error("Did not open gzipped file successfully: " + path);
    std::string out;
    char buf[1 << 16];
    int n;
    while ((n = gzread(g, buf, sizeof buf)) > 0) out.append(buf, (size_t)n);
    gzclose(g);
    if (n < 0) throw std::runtime_error("Error reading gzipped file: " + path);
    return out;
}

// face.cu:38-49: unknown tags fall back to wall_with_slip.
int bc_code(const std::string& name)
{
    if (name == "wall_with_slip") return 0;
    if (name == "wall_no_slip_adiabatic" || name == "wall_no_slip") return 1;
    if (name == "wall_no_slip_fixed_T") return 2;
    if (name == "exchange") return 3;
    if (name == "inflow") return 4;
    if (name == "outflow") return 5;
    if (name == "inflow_function") return 6;
    return 0;
}

// face.cu:61-68: unknown names fall back to none.
}  // namespace

bool gas_model_matches(const double* m)
{
    // number.cu:48-50 on the build's constants and the file's values read as number
    // (config.cu:369-377).
    auto close = [](real a, double b) { return std::fabs((double)a - (double)(real)b) < 1.0e-4; };
    return close(kGamma, m[0]) && close(kR, m[1]) && close(kQ, m[2]) && close(kAlpha, m[3]) &&
           close(kTi, m[4]) && close(kCv, m[5]);
}

// One number of a text file as the reference reads it into its number type (`>> number`,
// "%lf" / "%f"): strtod, or strtof in the FP32 flavour; widened to double (exactly).
double parse_real(const char* p, char** end)
{
    if (sizeof(real) == sizeof(float)) return (double)std::strtof(p, end);
    return std::strtod(p, end);
}

namespace {

int bc_function(const std::string& name)
{
    if (name == "supersonic_vortex") return 1;
    if (name == "laminar_boundary_layer") return 2;
    if (name == "manufactured_solution") return 3;
    return 0;
}

const char* kFaceNames[6] = {"iminus", "iplus", "jminus", "jplus", "kminus", "kplus"};
const char* kIoNames[13] = {"posx", "posy", "posz", "vol", "p", "T", "rho", "e", "YB", "a",
                            "velx", "vely", "velz"};

std::vector<double> num_list(const Json& j)
{
    std::vector<double> v;
    for (size_t q = 0; q < j.size(); ++q) v.push_back(j[q].as_num());
    return v;
}

}  // namespace

JobConfig read_config_json(const std::string& path)
{
    Json js = JsonParser(slurp(path)).parse();
    JobConfig c;
    c.title = js["title"].as_str();
    // Gas-model consistency check against the built model, config.cu:369-386.
    const Json& gm = js["gas_model"];
    const double model[6] = {gm["gamma"].as_num(), gm["R"].as_num(), gm["q"].as_num(),
                             gm["alpha"].as_num(), gm["Ti"].as_num(), gm["Cv"].as_num()};
    if (!gas_model_matches(model)) throw std::runtime_error("Incorrect gas model data");
    c.gas_model.assign(model, model + 6);
    const Json& ion = js["iovar_names"];
    for (int q = 0; q < 13 && q < (int)ion.size(); ++q)
        if (ion[q].as_str() != kIoNames[q])
            std::fprintf(stderr, "Mismatched iovar: %s%s\n", ion[q].as_str().c_str(), kIoNames[q]);
    const Json& fs = js["flow_states"];
    for (size_t q = 0; q < fs.size(); ++q) {
        const Json& s = fs[q];
        c.flow_states.push_back(s["gas"]["p"].as_num());
        c.flow_states.push_back(s["gas"]["T"].as_num());
        c.flow_states.push_back(s["gas"]["YB"].as_num());
        const Json& vel = s["vel"];
        for (int d = 0; d < 3; ++d) c.flow_states.push_back(vel[d].as_num());
    }
    int nblocks = js["n_fluid_blocks"].as_int();
    c.nib = js["nib"].as_int();
    c.njb = js["njb"].as_int();
    c.nkb = js["nkb"].as_int();
    std::vector<double> nics = num_list(js["nics"]), njcs = num_list(js["njcs"]), nkcs = num_list(js["nkcs"]);
    c.blk_ids.assign((size_t)c.nib * c.njb * c.nkb, -1);
    const Json& ids = js["blk_ids"];
    for (int i = 0; i < c.nib; ++i)
        for (int j = 0; j < c.njb; ++j)
            for (int k = 0; k < c.nkb; ++k) c.blk_ids[((size_t)i * c.njb + j) * c.nkb + k] = ids[i][j][k].as_int();
    const Json& fb = js["fluid_blocks"];
    for (size_t b = 0; b < fb.size(); ++b) {
        const Json& bj = fb[b];
        JobBlock blk;
        blk.i = bj["i"].as_int();
        blk.j = bj["j"].as_int();
        blk.k = bj["k"].as_int();
        blk.active = bj["active"].as_bool();
        if (bj.has("label") && bj["label"].kind == Json::String) blk.label = bj["label"].as_str();
        const Json& bcs = bj["bcs"];
        for (int f = 0; f < 6; ++f) {
            const Json& bc = bcs[kFaceNames[f]];
            blk.bc[f] = bc_code(bc["tag"].as_str());
            blk.bc_fs[f] = 0;
            if (blk.bc[f] == 4) blk.bc_fs[f] = bc["flow_state_index"].as_int();
            blk.bc_twall[f] = 0.0;
            if (blk.bc[f] == 2) {  // config.cu:463-467
                blk.bc_twall[f] = bc["value"].as_num();
                if (bc["form"].as_str() == "fun") blk.bc_twall[f] = std::nan("");
            }
            if (blk.bc[f] == 6) {  // config.cu:469-478; the profile is read from the current directory
                blk.bc_fun[f] = bc_function(bc["fun_name"].as_str());
                if (blk.bc_fun[f] == 2 && c.bl_profile.empty())
                    c.bl_profile = read_bl_profile("./InflowBoundaryLayerBC.data");
            }
        }
        if (blk.i < 0 || blk.i >= c.nib || blk.j < 0 || blk.j >= c.njb || blk.k < 0 || blk.k >= c.nkb)
            throw std::runtime_error("block position outside the block array");
        blk.nic = (int)nics.at(blk.i);
        blk.njc = (int)njcs.at(blk.j);
        blk.nkc = (int)nkcs.at(blk.k);
        c.blocks.push_back(blk);
    }
    if ((int)c.blocks.size() != nblocks)
        throw std::runtime_error("Did not read the correct number of block configurations: " +
                                 std::to_string(nblocks) + " " + std::to_string(c.blocks.size()));
    c.x_order = js["x_order"].as_int();
    c.t_order = js["t_order"].as_int();
    std::string fc = js["flux_calc"].as_str();
    c.flux_calc = (fc == "sbp_asf" || fc == "sbp-asf") ? 1 : 0;  // config.cu:304-309
    c.source_terms = (js["source_terms"].as_str() == "manufactured_solution") ? 1 : 0;
    c.dt_init = js["dt_init"].as_num();
    c.max_time = js["max_time"].as_num();
    c.max_step = js["max_step"].as_int();
    c.cfl_times = num_list(js["cfl_times"]);
    c.cfl_values = num_list(js["cfl_values"]);
    c.cfl_count = js["cfl_count"].as_int();
    c.print_count = js["print_count"].as_int();
    c.dt_plot_times = num_list(js["t_change"]);
    c.dt_plot_values = num_list(js["dt_plot"]);
    c.reacting = js["reacting"].as_bool();
    c.viscous = js["viscous"].as_bool();
    if (js.has("time_scheme")) {
        const std::string ts = js["time_scheme"].as_str();
        if (ts == "rk3") c.time_scheme = 0;
        else if (ts == "predictor_corrector") c.time_scheme = 1;
        else throw std::runtime_error("Unknown time_scheme: " + ts);
    }
    return c;
}

std::vector<double> read_grid(const std::string& path, bool binary, int nic, int njc, int nkc)
{
    int niv, njv, nkv;
    std::vector<double> v;
    if (binary) {  // block.cu:793-827: [dims, 0, 0, niv, njv, nkv] then xyz per vertex
        if (sizeof(real) != sizeof(double))  // block.cu:794-796
            throw std::runtime_error("Do not ask to FP32 code to read grid file with binary data.");
        std::string raw = slurp(path);
        if (raw.size() < 6 * sizeof(double)) throw std::runtime_error("Did not open binary grid file successfully: " + path);
        const double* d = reinterpret_cast<const double*>(raw.data());
        niv = (int)d[3]; njv = (int)d[4]; nkv = (int)d[5];
        if (nic != niv - 1 || njc != njv - 1 || nkc != nkv - 1)
            throw std::runtime_error("Unexpected grid size: niv=" + std::to_string(niv) + " njv=" +
                                     std::to_string(njv) + " nkv=" + std::to_string(nkv));
        size_t nv = (size_t)niv * njv * nkv;
        if (raw.size() < (6 + 3 * nv) * sizeof(double)) throw std::runtime_error("Grid file too short: " + path);
        v.assign(d + 6, d + 6 + 3 * nv);
    } else {  // block.cu:828-877: six header lines then "x y z" per line
        std::string text = slurp_gz(path);
        std::istringstream in(text);
        std::string line;
        for (int q = 0; q < 3; ++q) std::getline(in, line);
        std::getline(in, line); std::sscanf(line.c_str(), "niv: %d", &niv);
        std::getline(in, line); std::sscanf(line.c_str(), "njv: %d", &njv);
        std::getline(in, line); std::sscanf(line.c_str(), "nkv: %d", &nkv);
        if (nic != niv - 1 || njc != njv - 1 || nkc != nkv - 1)
            throw std::runtime_error("Unexpected grid size: niv=" + std::to_string(niv) + " njv=" +
                                     std::to_string(njv) + " nkv=" + std::to_string(nkv));
        size_t nv = (size_t)niv * njv * nkv;
        v.resize(3 * nv);
        for (size_t q = 0; q < nv; ++q) {
            std::getline(in, line);
            const char* p = line.c_str();
            char* end = nullptr;
            double xyz[3] = {0.0, 0.0, 0.0};  // block.cu:866-871 (sscanf %lf, or %f for FP32)
            for (int a = 0; a < 3; ++a) {
                const double val = parse_real(p, &end);
                if (end == p) break;
                xyz[a] = val;
                p = end;
            }
            v[3 * q] = xyz[0]; v[3 * q + 1] = xyz[1]; v[3 * q + 2] = xyz[2];
        }
    }
    return v;
}

std::vector<double> read_flow(const std::string& path, bool binary, int ncells)
{
    std::vector<double> v((size_t)13 * ncells);
    if (binary) {  // block.cu:886-907
        if (sizeof(real) != sizeof(double))  // block.cu:887-889
            throw std::runtime_error("Do not ask to FP32 code to read flow file with binary data.");
        std::string raw = slurp(path);
        if (raw.size() < v.size() * sizeof(double)) throw std::runtime_error("Did not open binary flow file successfully: " + path);
        std::memcpy(v.data(), raw.data(), v.size() * sizeof(double));
    } else {  // block.cu:908-928: one value per line
        std::string text = slurp_gz(path);
        const char* p = text.c_str();
        for (size_t q = 0; q < v.size(); ++q) {
            char* end = nullptr;
            v[q] = parse_real(p, &end);
            if (end == p) throw std::runtime_error("Flow file too short: " + path);
            p = end;
        }
    }
    return v;
}

void write_flow(const std::string& path, bool binary, const real* iovars, int ncells)
{
    size_t n = (size_t)13 * ncells;
    if (binary) {  // block.cu:938-955: the number type's bytes (4 in the FP32 flavour)
        std::ofstream f(path, std::ios::binary);
        if (!f) throw std::runtime_error("Did not open binary flow file successfully for writing: " + path);
        f.write(reinterpret_cast<const char*>(iovars), (std::streamsize)(n * sizeof(real)));
    } else {  // block.cu:956-974: ostream default formatting, i.e. %g with 6 digits
        gzFile g = gzopen(path.c_str(), "wb");
        if (!g) throw std::runtime_error("Did not open gzipped flow file successfully for writing: " + path);
        std::string buf;
        buf.reserve(1 << 20);
        char tmp[64];
        for (size_t q = 0; q < n; ++q) {
            int len = std::snprintf(tmp, sizeof tmp, "%g\n", (double)iovars[q]);
            buf.append(tmp, (size_t)len);
            if (buf.size() > (1 << 20)) { gzwrite(g, buf.data(), (unsigned)buf.size()); buf.clear(); }
        }
        if (!buf.empty()) gzwrite(g, buf.data(), (unsigned)buf.size());
        gzclose(g);
    }
}

void read_times(const std::string& job, int tindx, double& t, int& next_plot_indx)
{
    t = 0.0;  // simulate.cu:132-149 reads the time as number
    next_plot_indx = 1;
    std::ifstream f(job + "/times.data");
    if (!f.good()) return;
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line.find('#') != std::string::npos) continue;
        std::istringstream ss(line);
        int ti;
        real tm;
        if (ss >> ti >> tm) {
            if (ti == tindx) { t = (double)tm; next_plot_indx = ti + 1; }
        }
    }
}

void append_times(const std::string& job, int tindx, real t)
{
    std::ofstream f(job + "/times.data", std::ios::app);
    char buf[64];
    std::snprintf(buf, sizeof buf, "%d %g\n", tindx, (double)t);
    f << buf;
}

std::vector<double> read_bl_profile(const std::string& path)
{
    std::error_code ec;
    if (!std::filesystem::is_regular_file(path, ec))
        throw std::runtime_error("Boundary-layer file is not present in current directory.");
    std::istringstream in(slurp(path));
    real p_e = 0;  // config.cu:37-60 reads number values
    int nsample = 0;
    in >> p_e >> nsample;
    if (nsample != 12) throw std::runtime_error("Boundary-layer file has unexpected number of sample points.");
    std::vector<double> out(1, (double)p_e);
    for (int q = 0; q < 3 * nsample; ++q) {
        real v;
        if (!(in >> v)) throw std::runtime_error("Boundary-layer file is too short: " + path);
        out.push_back((double)v);
    }
    return out;
}

std::vector<double> read_twall(const std::string& job, const JobBlock& b, int side, int nfaces)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, "/TWall-%04d-%04d-%04d-%s.gz", b.i, b.j, b.k, kFaceNames[side]);
    const std::string path = job + buf;
    std::string text;
    try {
        text = slurp_gz(path);
    } catch (const std::runtime_error&) {
        throw std::runtime_error("Did not open gzipped TWall file successfully: " + path);
    }
    std::vector<double> v((size_t)nfaces);
    const char* p = text.c_str();
    for (int q = 0; q < nfaces; ++q) {
        char* end = nullptr;
        v[q] = parse_real(p, &end);
        if (end == p) throw std::runtime_error("TWall file too short: " + path);
        p = end;
    }
    return v;
}

std::string block_file(const std::string& job, const char* kind, int tindx, int i, int j, int k,
                       bool binary)
{
    char buf[256];
    if (std::strcmp(kind, "grid") == 0)
        std::snprintf(buf, sizeof buf, "/grid/grid-%04d-%04d-%04d", i, j, k);
    else
        std::snprintf(buf, sizeof buf, "/flow/t%04d/flow-%04d-%04d-%04d", tindx, i, j, k);
    return job + buf + (binary ? ".bin" : ".gz");
}

}  // namespace chkn
