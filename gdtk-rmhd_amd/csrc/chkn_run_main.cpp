// chkn_run_main.cpp -- chkn-run command line, the drop-in for the reference's main
// (src/chicken/main.cu:25-89): same flags, same console protocol, same exit behaviour
// (a failed run prints the error and still exits 0, main.cu:82-88).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/chkn.h"

static bool take(int argc, char** argv, int& i, const char* lng, const char* shrt, std::string& val)
{
    std::string a = argv[i];
    std::string l = std::string("--") + lng;
    if (a.rfind(l + "=", 0) == 0) { val = a.substr(l.size() + 1); return true; }
    if ((a == l || a == shrt) && i + 1 < argc) { val = argv[++i]; return true; }
    return false;
}

int main(int argc, char** argv)
{
    std::printf("Chicken compressible-flow CFD\n");
    std::printf("Revision-id: chkn-mi355x\n");
    std::printf("MI355X (HIP, gfx950) flavour of program\n");
    auto t0 = std::chrono::steady_clock::now();
    std::string job, v;
    int tindx = 0, verbosity = 0;
    bool binary = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "-h" || a == "--help") {
            std::fprintf(stderr, "chkn-run --job=<string> [--tindx=<int>] [--verbosity=<int>] [--binary]\n");
            return 0;
        } else if (take(argc, argv, i, "job", "-j", v)) {
            job = v;
        } else if (take(argc, argv, i, "tindx", "-t", v)) {
            tindx = std::atoi(v.c_str());
        } else if (take(argc, argv, i, "verbosity", "-v", v)) {
            verbosity = std::atoi(v.c_str());
        } else if (a == "--binary" || a == "-b") {
            binary = true;
        } else {
            std::fprintf(stderr, "Unknown argument: %s\n", a.c_str());
            return 1;
        }
    }
    if (verbosity > 0) std::printf("job name: %s\n", job.c_str());
    if (job.empty()) {
        std::fprintf(stderr, "You need to provide a job name.\n");
        return 1;
    }
    if (verbosity > 0) std::printf("tindx_start: %d\nbinary_data: %d\n", tindx, binary ? 1 : 0);
    chkn_set_verbosity(verbosity);
    chkn_sim* sim = nullptr;
    int rc = chkn_load_job(job.c_str(), tindx, binary ? 1 : 0, &sim);
    if (rc != CHKN_OK) {
        std::fprintf(stderr, "%s\n", chkn_last_error(nullptr));
    } else {
        std::printf("Cells in simulation: %lld\n", chkn_active_cells(sim));  // simulate.cu:89-90
        std::printf("Bytes allocated on GPU: %.3fMB\n", (double)chkn_device_bytes(sim) / 1.0e6);
        chkn_report rep;
        rc = chkn_march(sim, binary ? 1 : 0, &rep);
        if (rc != CHKN_OK) std::fprintf(stderr, "%s\n", chkn_last_error(sim));
        else chkn_finalize(sim, binary ? 1 : 0);
        chkn_destroy(sim);
    }
    double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("Done in %.3fs\n", s);
    return 0;
}
