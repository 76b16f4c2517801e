// C++ host side of the drop-in: the reference's System, CommHelper and LBM
// types (tgv/src/System.hpp, tgv/src/lbm.hpp) with the same member and method
// names, backed by the C-ABI in include/imexlb.h.  Only plain C calls cross
// into libimexlb.so.
#pragma once

#include <stdexcept>
#include <string>

#include "imexlb.h"

namespace imexlb {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string &what) : std::runtime_error(what), status(status) {}
    int status;
};

// System (tgv/src/System.hpp:7-20): input.in scalars and derived tau.
class System {
public:
    System(const std::string &path, int case_id, bool keep_hardcoded);
    void Monitor() const;  // tgv/src/System.cpp:26-42 (banner)

    int sx, sy, sz;
    int Time, inter;
    double miu, u0, R, rho0, Ma, tau;
    double Re;  // an int in the reference (tgv/src/System.hpp:18)
    double cs2, cs;
};

// CommHelper (tgv/src/lbm.hpp:34-117) for this process's rank; one rank per
// GPU, and this build runs one rank.
struct CommHelper {
    CommHelper(int nranks = 1, int me = 0, bool periodic = true, int glx = 1, int gly = 1, int glz = 1);
    int rx, ry, rz, me, px, py, pz, nranks;
    int neighbours[26];
    int l_l[3], lo[3], hi[3];
};

struct Options {
    int case_id = IMEXLB_CASE_TGV;
    int mode = IMEXLB_MODE_FAITHFUL;
    int engine = IMEXLB_ENGINE_FUSED;
    int dtype = IMEXLB_F64;
    int storage = IMEXLB_PULL;
    double lid_u = 0.1;
    double sphere[4] = {0, 0, 0, 0};
    int device = 0;
};

// struct LBM (tgv/src/lbm.hpp:119-252): the six-call step surface.
struct LBM {
    LBM(CommHelper comm_, int sx, int sy, int sz, double &tau, double &rho0, double &u0, const Options &opt);
    ~LBM();
    LBM(const LBM &) = delete;
    LBM &operator=(const LBM &) = delete;

    void Initialize();
    void Collision();
    void setup_subdomain();
    void pack();
    void exchange();
    void unpack();
    void Streaming();
    void Update();
    void MPIoutput(int n);  // output<n>.plt, tgv/src/lbm.cpp:335-669

    // beyond the reference: fused fast path, data access, diagnostics
    void step(int nsteps, int moments_every);
    float step_timed(int nsteps, int moments_every);
    void sync();
    void moments(double *p, double *u, double *v, double *w);
    void diagnostics(double &ke, double &mass);
    void autotune(const char *cache);
    int tune_placement(int trials);  // index of the lattice kept
    // window placement search over lattice + slack_gib GiB; *npos receives
    // the number of offsets timed; returns the index kept
    int tune_placement_window(double slack_gib, int *npos);
    void checkpoint(const std::string &path);
    void restart(const std::string &path);
    imexlb_info info();

    CommHelper comm;
    int glx, gly, glz;
    double tau0, rho0, u0;

private:
    void check(int status);
    imexlb_handle *h_ = nullptr;
};

}  // namespace imexlb
