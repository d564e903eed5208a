// main.cpp -- command-line driver, the counterpart of the reference's EU3D
// binary (src/Main.cpp): reads a ctrl file, runs the time loop on the GPU and
// writes the EU3D_STATE_V1_LE dump when EU3D_DUMP_STATE=1 (src/Euler.cpp:1540).
//   eu3d_run [ctrl.txt] [--base DIR] [--steps N] [--dump PATH] [--no-compact]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/eu3d_host.h"

int main(int argc, char **argv)
{
    std::string ctrl = "./input/ctrl.txt", base = ".", dump, outdir;
    int steps = 1 << 30, flags = EU3D_CHEM_COMPACT | EU3D_CHEM_KEEP_CMS;
    for (int a = 1; a < argc; a++) {
        std::string s = argv[a];
        if (s == "--steps" && a + 1 < argc)
            steps = std::atoi(argv[++a]);
        else if (s == "--base" && a + 1 < argc)
            base = argv[++a];
        else if (s == "--dump" && a + 1 < argc)
            dump = argv[++a];
        else if (s == "--output" && a + 1 < argc)
            outdir = argv[++a];
        else if (s == "--no-compact")
            flags &= ~EU3D_CHEM_COMPACT;
        else
            ctrl = s;
    }
    const char *env = std::getenv("EU3D_DUMP_STATE");
    if (dump.empty() && env && std::strcmp(env, "1") == 0)
        dump = "state_check_1_rank0.bin";
    char err[512] = {0};
    eu3d_solver *s = nullptr;
    int rc = eu3d_solver_open(ctrl.c_str(), base.c_str(), &s, err, sizeof err);
    if (rc) {
        std::fprintf(stderr, "eu3d: %s\n", err);
        return 2;
    }
    if (!outdir.empty() && (rc = eu3d_solver_set_output(s, outdir.c_str()))) {
        std::fprintf(stderr, "eu3d: %s\n", eu3d_solver_error(s));
        eu3d_solver_close(s);
        return 2;
    }
    auto t0 = std::chrono::steady_clock::now();
    rc = eu3d_solver_run(s, steps, flags);
    eu3d_synchronize(eu3d_solver_ctx(s));
    auto t1 = std::chrono::steady_clock::now();
    if (rc) {
        std::fprintf(stderr, "eu3d: %s\n", eu3d_solver_error(s));
        eu3d_solver_close(s);
        return 3;
    }
    eu3d_health_info h;
    int hr = eu3d_health(eu3d_solver_ctx(s), &h);
    int dims[10];
    double info[6];
    eu3d_solver_dims(s, dims);
    eu3d_solver_info(s, info);
    double sec = std::chrono::duration<double>(t1 - t0).count();
    double cells = (double)dims[0] * dims[1] * dims[2];
    std::printf("iterations=%d time=%.9g main_loop=%.6f s cell_steps_per_s=%.6g max_nchem=%d stiff=%lld\n",
                (int)info[5], info[3], sec, cells * info[5] / sec, h.max_nchem, h.stiff_cells);
    if (!dump.empty()) {
        rc = eu3d_solver_dump(s, dump.c_str(), 1);
        if (rc)
            std::fprintf(stderr, "eu3d: %s\n", eu3d_solver_error(s));
    }
    if (!outdir.empty()) { // the reference's final Output(count) (src/Euler.cpp:101)
        int rc2 = eu3d_solver_output(s, -1);
        if (rc2)
            std::fprintf(stderr, "eu3d: %s\n", eu3d_solver_error(s));
        rc = rc ? rc : rc2;
    }
    eu3d_solver_close(s);
    return hr ? 4 : (rc ? 3 : 0);
}
