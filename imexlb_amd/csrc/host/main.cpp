// Driver with the reference's structure (tgv/main.cpp:8-214; cavity and
// sphere drivers lid_driven_cavity/main.cpp:6-51, sphere/main.cpp:6-51):
// read input.in, build the LBM, Initialize, (TGV) 20 warm-up steps and a
// re-Initialize, then the timed Collision / pack / exchange / unpack /
// Streaming / Update loop with a wall-clock accumulator per phase, and the
// same "Avg time spent in ..." report.  One rank, one GPU.
//
//   imexlb [--case tgv|cavity|sphere] [--mode faithful|config]
//          [--engine fused|phased] [--dtype f64|f32|f16] [--storage pull|aa]
//          [--input input.in] [--steps N] [--honour-input] [--fast]
//          [--lid U] [--sphere cx,cy,cz,r] [--tune CACHE] [--device D]
//          [--output]   (Tecplot output<n>.plt every inter steps, as the
//                        cavity driver does: lid_driven_cavity/main.cpp:26,39-44)
//          [--restart FILE] (resume from a restart file; no warm-up)
//          [--checkpoint FILE] (write a restart file after the last step)
//          [--place N] (time N lattice allocations, keep the fastest)
//          [--window GIB] (slide the lattice through an address range of
//                          lattice + GIB and keep the fastest offset; the
//                          rest of the range is given back)
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "lbm_facade.hpp"

using namespace imexlb;
using clk = std::chrono::steady_clock;

static double now_s() { return std::chrono::duration<double>(clk::now().time_since_epoch()).count(); }

static int pick(const std::string &v, std::initializer_list<std::pair<const char *, int>> table, const char *what)
{
    for (auto &p : table)
        if (v == p.first) return p.second;
    std::fprintf(stderr, "unknown %s '%s'\n", what, v.c_str());
    std::exit(2);
}

int main(int argc, char **argv)
{
    Options opt;
    std::string input = "input.in";
    int steps = -1;
    bool honour_input = false, fast = false, have_sphere = false, output = false;
    std::string tune, restart, checkpoint;
    int place = 1;
    double window_gib = 0.0;
    for (int a = 1; a < argc; ++a) {
        std::string k = argv[a];
        auto val = [&]() -> std::string {
            if (a + 1 >= argc) {
                std::fprintf(stderr, "%s needs a value\n", k.c_str());
                std::exit(2);
            }
            return argv[++a];
        };
        if (k == "--case") opt.case_id = pick(val(), {{"tgv", 0}, {"cavity", 1}, {"ldc", 1}, {"sphere", 2}}, "case");
        else if (k == "--mode") opt.mode = pick(val(), {{"faithful", 0}, {"config", 1}}, "mode");
        else if (k == "--engine") opt.engine = pick(val(), {{"fused", 0}, {"phased", 1}}, "engine");
        else if (k == "--dtype") opt.dtype = pick(val(), {{"f64", 0}, {"f32", 1}, {"f16", 2}}, "dtype");
        else if (k == "--storage") opt.storage = pick(val(), {{"pull", 0}, {"aa", 1}}, "storage");
        else if (k == "--input") input = val();
        else if (k == "--steps") steps = std::atoi(val().c_str());
        else if (k == "--honour-input") honour_input = true;
        else if (k == "--fast") fast = true;
        else if (k == "--output") output = true;
        else if (k == "--lid") opt.lid_u = std::atof(val().c_str());
        else if (k == "--device") opt.device = std::atoi(val().c_str());
        else if (k == "--tune") tune = val();
        else if (k == "--restart") restart = val();
        else if (k == "--checkpoint") checkpoint = val();
        else if (k == "--place") place = std::atoi(val().c_str());
        else if (k == "--window") window_gib = std::atof(val().c_str());
        else if (k == "--sphere") {
            std::string s = val();
            if (std::sscanf(s.c_str(), "%lf,%lf,%lf,%lf", &opt.sphere[0], &opt.sphere[1], &opt.sphere[2],
                            &opt.sphere[3]) != 4) {
                std::fprintf(stderr, "--sphere cx,cy,cz,r\n");
                return 2;
            }
            have_sphere = true;
        } else if (k == "-h" || k == "--help") {
            std::printf("see the header of main.cpp for the options\n");
            return 0;
        } else {
            std::fprintf(stderr, "unknown option %s\n", k.c_str());
            return 2;
        }
    }
    // Every case runs on the fused engine (the reference's own cavity and
    // sphere in FP64 pull storage, with their first ghost layer as a ring)
    // unless --engine phased asks for the reference's ghost-3 layout, phase
    // by phase: the same bits, kept as the reference-layout cross-check.

    try {
        System s1(input, opt.case_id, !honour_input);
        s1.Monitor();
        if (steps >= 0) s1.Time = steps;
        if (opt.case_id == IMEXLB_CASE_SPHERE && opt.mode == IMEXLB_MODE_CONFIG && !have_sphere) {
            // the channel benchmark's sphere: diameter 40/256 of the height, at x = nx/4
            opt.sphere[0] = s1.sx / 4.0;
            opt.sphere[1] = s1.sy / 2.0;
            opt.sphere[2] = s1.sz / 2.0;
            opt.sphere[3] = 20.0 * s1.sy / 256.0;
        }
        CommHelper comm(1, 0, opt.case_id == IMEXLB_CASE_TGV, s1.sx, s1.sy, s1.sz);
        LBM l1(comm, s1.sx, s1.sy, s1.sz, s1.tau, s1.rho0, s1.u0, opt);
        if (opt.engine == IMEXLB_ENGINE_PHASED && (!tune.empty() || place > 1 || window_gib > 0)) {
            std::cerr << "note: --tune, --place and --window apply to the fused engine; the phased engine keeps "
                         "the reference's layout" << std::endl;
        } else {
            // placement first, then the launch shape on the kept lattice
            if (window_gib > 0) {
                int npos = 0;
                const int kept = l1.tune_placement_window(window_gib, &npos);
                std::cout << "Lattice placement window: kept offset " << kept << " of " << npos << std::endl;
            } else if (place > 1) {
                std::cout << "Lattice placement kept: " << l1.tune_placement(place) << " of " << place << std::endl;
            }
            if (!tune.empty()) l1.autotune(tune.c_str());
        }
        l1.Initialize();
        l1.setup_subdomain();
        std::cout << "Currently using HIP device: " << opt.device << std::endl;

        if (!restart.empty()) {
            l1.restart(restart);
            if (l1.comm.me == 0) std::cout << "Restarted from " << restart << std::endl;
        } else if (opt.case_id == IMEXLB_CASE_TGV) {
            if (l1.comm.me == 0) std::cout << "Warm-Up" << std::endl;
            for (int it = 1; it <= 20; it++) {  // tgv/main.cpp:51-62
                l1.Collision();
                l1.pack();
                l1.exchange();
                l1.unpack();
                l1.Streaming();
                l1.Update();
            }
            l1.sync();
            if (l1.comm.me == 0) std::cout << "Reset and Run Main-Loop" << std::endl;
            l1.Initialize();  // tgv/main.cpp:68
        }

        // the initial fields, as the cavity driver writes them (lid_driven_cavity/main.cpp:26)
        if (output) l1.MPIoutput(0);

        double t_col = 0, t_pack = 0, t_exch = 0, t_unpack = 0, t_stream = 0, t_update = 0;
        double start = 0, end = 0;
        if (fast) {
            l1.sync();
            start = now_s();
            // moments every `inter` steps; with inter = 0 only the last step
            // writes them, for the final diagnostics
            const float ms = l1.step_timed(s1.Time, s1.inter > 0 ? s1.inter : s1.Time);
            end = now_s();
            t_stream = ms * 1e-3;
        } else {
            l1.sync();
            start = now_s();
            for (int it = 1; it <= s1.Time; it++) {
                double t0 = now_s();
                l1.Collision();
                l1.sync();
                double t1 = now_s();
                l1.pack();
                l1.sync();
                double t2 = now_s();
                l1.exchange();
                l1.sync();
                double t3 = now_s();
                l1.unpack();
                l1.sync();
                double t4 = now_s();
                l1.Streaming();
                l1.sync();
                double t5 = now_s();
                l1.Update();
                l1.sync();
                double t6 = now_s();
                t_col += t1 - t0;
                t_pack += t2 - t1;
                t_exch += t3 - t2;
                t_unpack += t4 - t3;
                t_stream += t5 - t4;
                t_update += t6 - t5;
                if (s1.inter > 0 && it % s1.inter == 0 && l1.comm.me == 0) {
                    double ke = 0, mass = 0;
                    l1.diagnostics(ke, mass);
                    std::printf("time-step = %d  KE = %.12e  mass = %.12e\n", it, ke, mass);
                    if (output) l1.MPIoutput(it / s1.inter);
                }
            }
            end = now_s();
        }
        const double total = end - start;
        const double cells = (double)s1.sx * s1.sy * s1.sz;
        std::printf("Avg time spent in Collision:  %lf\n", t_col);
        std::printf("Avg time spent in Pack:  %lf\n", t_pack);
        std::printf("Avg time spent in Exchange:  %lf\n", t_exch);
        std::printf("Avg time spent in Unpack:  %lf\n", t_unpack);
        std::printf("Avg time spent in Stream:  %lf\n", t_stream);
        std::printf("Avg time spent in Update:  %lf\n", t_update);
        std::printf("Avg. Total Solver Time:  %lf\n", total);
        std::printf("Avg. Time per time-step:  %lf\n", s1.Time ? total / s1.Time : 0.0);
        if (total > 0) std::printf("Throughput:  %.3f MLUPS\n", cells * s1.Time / total / 1e6);
        if (s1.Time > 0) {
            double ke = 0, mass = 0;
            l1.diagnostics(ke, mass);
            std::printf("Final KE = %.15e  mass = %.15e\n", ke, mass);
        }
        if (!checkpoint.empty()) {
            l1.checkpoint(checkpoint);
            std::printf("Restart file %s written\n", checkpoint.c_str());
        }
    } catch (const Error &e) {
        std::fprintf(stderr, "imexlb: %s (status %d)\n", e.what(), e.status);
        return 1;
    }
    return 0;
}
