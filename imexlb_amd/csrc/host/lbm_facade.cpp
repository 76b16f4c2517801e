#include "lbm_facade.hpp"

#include <cstdio>
#include <cstring>
#include <iostream>
#include <vector>

namespace imexlb {

System::System(const std::string &path, int case_id, bool keep_hardcoded)
{
    imexlb_system s;
    int st = imexlb_system_read(path.c_str(), case_id, keep_hardcoded ? 1 : 0, &s);
    if (st != IMEXLB_OK) throw Error(st, "cannot read " + path + ": " + imexlb_strerror(st));
    sx = s.sx;
    sy = s.sy;
    sz = s.sz;
    Time = s.Time;
    inter = s.inter;
    miu = s.miu;
    u0 = s.u0;
    R = s.R;
    rho0 = s.rho0;
    Ma = s.Ma;
    tau = s.tau;
    Re = s.Re;
    cs2 = s.cs2;
    cs = s.cs;
}

void System::Monitor() const
{
    std::cout << "============================" << std::endl
              << "D3Q27 BGK (imexlb on MI355X)" << std::endl
              << "Re    =" << Re << std::endl
              << "Ma    =" << Ma << std::endl
              << "rho   =" << rho0 << std::endl
              << "miu   =" << miu << std::endl
              << "tau   =" << tau << std::endl
              << "Time  =" << Time << std::endl
              << "inter =" << inter << std::endl
              << "nx    =" << sx << std::endl
              << "ny    =" << sy << std::endl
              << "nz    =" << sz << std::endl
              << "============================" << std::endl;
}

CommHelper::CommHelper(int nranks_, int me_, bool periodic, int glx, int gly, int glz)
{
    imexlb_comm c;
    int st = imexlb_comm_layout(nranks_, me_, periodic ? 1 : 0, glx, gly, glz, &c);
    if (st != IMEXLB_OK) throw Error(st, "invalid rank layout");
    nranks = c.nranks;
    me = c.me;
    rx = c.rx;
    ry = c.ry;
    rz = c.rz;
    px = c.px;
    py = c.py;
    pz = c.pz;
    std::memcpy(neighbours, c.neighbours, sizeof neighbours);
    std::memcpy(l_l, c.l_l, sizeof l_l);
    std::memcpy(lo, c.lo, sizeof lo);
    std::memcpy(hi, c.hi, sizeof hi);
}

LBM::LBM(CommHelper comm_, int sx, int sy, int sz, double &tau, double &rho0_, double &u0_, const Options &opt)
    : comm(comm_), glx(sx), gly(sy), glz(sz), tau0(tau), rho0(rho0_), u0(u0_)
{
    imexlb_config c;
    imexlb_config_init(&c);
    c.case_id = opt.case_id;
    c.mode = opt.mode;
    c.engine = opt.engine;
    c.dtype = opt.dtype;
    c.storage = opt.storage;
    c.nx = comm.l_l[0];
    c.ny = comm.l_l[1];
    c.nz = comm.l_l[2];
    c.tau0 = tau;
    c.rho0 = rho0_;
    c.u0 = u0_;
    c.lid_u = opt.lid_u;
    c.sphere_cx = opt.sphere[0];
    c.sphere_cy = opt.sphere[1];
    c.sphere_cz = opt.sphere[2];
    c.sphere_r = opt.sphere[3];
    c.device = opt.device;
    int st = imexlb_create(&c, &h_);
    if (st != IMEXLB_OK) throw Error(st, std::string("imexlb_create: ") + imexlb_last_error(nullptr));
}

LBM::~LBM() { imexlb_destroy(h_); }

void LBM::check(int st)
{
    if (st != IMEXLB_OK) throw Error(st, imexlb_last_error(h_));
}

void LBM::Initialize() { check(imexlb_initialize(h_)); }
void LBM::Collision() { check(imexlb_collide(h_)); }
void LBM::setup_subdomain() { check(imexlb_setup_subdomain(h_)); }
void LBM::pack() { check(imexlb_pack(h_)); }
void LBM::exchange() { check(imexlb_exchange(h_)); }
void LBM::unpack() { check(imexlb_unpack(h_)); }
void LBM::Streaming() { check(imexlb_stream(h_)); }
void LBM::Update() { check(imexlb_update(h_)); }
void LBM::MPIoutput(int n)
{
    const size_t cells = (size_t)comm.l_l[0] * comm.l_l[1] * comm.l_l[2];
    std::vector<double> p(cells), u(cells), v(cells), w(cells);
    check(imexlb_get_moments(h_, p.data(), u.data(), v.data(), w.data()));
    const std::string path = "output" + std::to_string(n) + ".plt";
    check(imexlb_tecplot_write(path.c_str(), (double)n, comm.l_l[0], comm.l_l[1], comm.l_l[2], p.data(), u.data(),
                               v.data(), w.data()));
    if (comm.me == 0) std::printf("\nThe result %d is writen\n\n============================\n", n);
}

void LBM::step(int n, int every) { check(imexlb_step(h_, n, every)); }
float LBM::step_timed(int n, int every)
{
    float ms = 0;
    check(imexlb_step_timed(h_, n, every, &ms));
    return ms;
}
void LBM::sync() { check(imexlb_sync(h_)); }
void LBM::moments(double *p, double *u, double *v, double *w) { check(imexlb_get_moments(h_, p, u, v, w)); }
void LBM::diagnostics(double &ke, double &mass) { check(imexlb_diagnostics(h_, &ke, &mass)); }
void LBM::autotune(const char *cache) { check(imexlb_autotune(h_, cache, 10, nullptr)); }
int LBM::tune_placement(int trials)
{
    int32_t kept = 0;
    check(imexlb_tune_placement(h_, trials, 10, nullptr, &kept));
    return kept;
}
int LBM::tune_placement_window(double slack_gib, int *npos)
{
    // steps of 2 GiB for lattices of 16 GiB and more, 1 GiB from 4 GiB, 2 MiB
    // multiples down to 64 MiB below; at most 256 offsets
    const imexlb_info i = info();
    const double lat = (double)i.device_bytes;
    const int64_t mib2 = 2ll << 20;
    int64_t step = lat >= 16.0 * (1 << 30) ? 2ll << 30 : lat >= 4.0 * (1 << 30) ? 1ll << 30 : 64ll << 20;
    int64_t slack = (int64_t)(slack_gib * (double)(1ll << 30)) / mib2 * mib2;
    while (slack / step > 255) step *= 2;
    slack = slack / step * step;
    const int n = (int)(slack / step) + 1;
    std::vector<float> ms((size_t)n, 0.0f);
    int32_t kept = 0;
    check(imexlb_tune_placement_window(h_, slack, step, 10, ms.data(), n, &kept));
    if (npos) *npos = n;
    return kept;
}
void LBM::checkpoint(const std::string &path) { check(imexlb_checkpoint_write(h_, path.c_str())); }
void LBM::restart(const std::string &path) { check(imexlb_checkpoint_read(h_, path.c_str())); }
imexlb_info LBM::info()
{
    imexlb_info i;
    check(imexlb_get_info(h_, &i));
    return i;
}

}  // namespace imexlb
