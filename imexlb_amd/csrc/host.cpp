// Host-only pieces of the reference behind the C-ABI: the input.in parser
// (System) and the rank grid (CommHelper + the LBM constructor's extents).
// Neither touches the GPU.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/imexlb.h"

extern "C" {

// System::System (tgv/src/System.cpp:5-24; lid_driven_cavity/src/System.cpp:
// 5-20; sphere/src/System.cpp:5-23).  The reference reads Re into an int
// (tgv/src/System.hpp:18); it is a double here, which reads the same for the
// integer values the reference accepts.  A missing or short file is an error
// here (the reference leaves the fields uninitialised).
int imexlb_system_read(const char *path, int case_id, int keep_hardcoded, imexlb_system *o)
{
    if (!path || !o || case_id < 0 || case_id > 2) return IMEXLB_EINVAL;
    std::memset(o, 0, sizeof *o);
    std::ifstream in(path);
    if (!in) return IMEXLB_EIO;
    double Time = 0, inter = 0;
    if (!(in >> o->rho0 >> o->R >> o->Re)) return IMEXLB_EIO;
    if (!(in >> o->u0 >> Time >> inter)) return IMEXLB_EIO;
    if (!(in >> o->sx >> o->sy >> o->sz)) return IMEXLB_EIO;
    if (o->sx <= 0 || o->sy <= 0 || o->sz <= 0 || o->Re == 0.0) return IMEXLB_EIO;
    o->Time = (int32_t)Time;
    o->inter = (int32_t)inter;
    o->cs2 = 1.0 / 3.0;
    o->cs = std::sqrt(o->cs2);
    o->Ma = o->u0 / o->cs;
    if (case_id == IMEXLB_CASE_TGV) {
        if (keep_hardcoded) {
            o->Time = 100;  // tgv/src/System.cpp:17
            o->inter = 10;  // tgv/src/System.cpp:18
        }
        o->miu = o->rho0 * o->u0 * o->sx / 2 / 3.1415926 / o->Re;
        o->tau = o->u0 * o->sx / o->Re / o->cs2 / 2 / 3.1415926;
    } else {
        o->miu = o->rho0 * o->u0 * o->R / o->Re;
        o->tau = o->u0 * o->R / o->Re / o->cs2;
        if (case_id == IMEXLB_CASE_SPHERE && keep_hardcoded) {
            o->Time = (int32_t)(std::pow(2 * o->R, 2) / o->miu);  // sphere/src/System.cpp:21
            o->inter = 1;                                          // :22
        }
    }
    return IMEXLB_OK;
}

// CommHelper::CommHelper (tgv/src/lbm.hpp:52-116, periodic; non-periodic
// variant lid_driven_cavity/src/lbm.hpp:52-101) and the extents and global
// offsets of LBM::LBM (tgv/src/lbm.hpp:161-236; lid_driven_cavity/src/lbm.hpp:
// 145-194).  The offsets gather every rank's l_l as the reference's
// MPI_Allgather would, including its indexing of the gathered arrays.
int imexlb_comm_layout(int nranks, int me, int periodic, int glx, int gly, int glz, imexlb_comm *o)
{
    if (!o || nranks < 1 || me < 0 || me >= nranks || glx < 1 || gly < 1 || glz < 1) return IMEXLB_EINVAL;
    std::memset(o, 0, sizeof *o);
    int rx = (int)std::pow(1.0 * nranks, 1.0 / 3.0);
    if (rx < 1) rx = 1;
    while (nranks % rx != 0) rx++;
    int rz = (int)std::sqrt(1.0 * (nranks / rx));
    if (rz < 1) rz = 1;
    while ((nranks / rx) % rz != 0) rz++;
    const int ry = nranks / rx / rz;
    auto coords = [&](int r, int &px, int &py, int &pz) {
        px = r % rx;
        pz = (r / rx) % rz;
        py = (r / rx / rz);
    };
    int px, py, pz;
    coords(me, px, py, pz);
    o->nranks = nranks;
    o->me = me;
    o->rx = rx;
    o->ry = ry;
    o->rz = rz;
    o->px = px;
    o->py = py;
    o->pz = pz;

    int *nb = o->neighbours;
    if (periodic) {
        const int left = px == 0 ? me + rx - 1 : me - 1;
        const int right = px == rx - 1 ? me - rx + 1 : me + 1;
        const int down = pz == 0 ? me + rx * (rz - 1) : me - rx;
        const int up = pz == rz - 1 ? me - rx * (rz - 1) : me + rx;
        const int front = py == 0 ? me + rx * rz * (ry - 1) : me - rx * rz;
        const int back = py == ry - 1 ? me - rx * rz * (ry - 1) : me + rx * rz;
        const int leftup = (pz == rz - 1) ? left - rx * (rz - 1) : left + rx;
        const int rightup = (pz == rz - 1) ? right - rx * (rz - 1) : right + rx;
        const int leftdown = (pz == 0) ? left + rx * (rz - 1) : left - rx;
        const int rightdown = (pz == 0) ? right + rx * (rz - 1) : right - rx;
        const int frontup = (pz == rz - 1) ? front - rx * (rz - 1) : front + rx;
        const int frontright = (px == rx - 1) ? front - rx + 1 : front + 1;
        const int frontdown = (pz == 0) ? front + rx * (rz - 1) : front - rx;
        const int frontleft = (px == 0) ? front + rx - 1 : front - 1;
        const int backup = (pz == rz - 1) ? back - rx * (rz - 1) : back + rx;
        const int backright = (px == rx - 1) ? back - rx + 1 : back + 1;
        const int backdown = (pz == 0) ? back + rx * (rz - 1) : back - rx;
        const int backleft = (px == 0) ? back + rx - 1 : back - 1;
        const int backrightdown = (px == rx - 1) ? backdown - rx + 1 : backdown + 1;
        const int frontrightdown = (px == rx - 1) ? frontdown - rx + 1 : frontdown + 1;
        const int frontrightup = (px == rx - 1) ? frontup - rx + 1 : frontup + 1;
        const int backrightup = (px == rx - 1) ? backup - rx + 1 : backup + 1;
        const int frontleftup = (px == 0) ? frontup + rx - 1 : frontup - 1;
        const int backleftdown = (px == 0) ? backdown + rx - 1 : backdown - 1;
        const int frontleftdown = (px == 0) ? frontdown + rx - 1 : frontdown - 1;
        const int backleftup = (px == 0) ? backup + rx - 1 : backup - 1;
        const int v[26] = {left, right, down, up, front, back,
                           frontup, frontright, frontdown, frontleft, backup, backright, backdown, backleft,
                           leftup, rightup, leftdown, rightdown,
                           frontleftup, frontleftdown, frontrightup, frontrightdown,
                           backleftup, backleftdown, backrightup, backrightdown};
        std::memcpy(nb, v, sizeof v);
    } else {
        const int left = px == 0 ? -1 : me - 1;
        const int right = px == rx - 1 ? -1 : me + 1;
        const int down = pz == 0 ? -1 : me - rx;
        const int up = pz == rz - 1 ? -1 : me + rx;
        const int front = py == 0 ? -1 : me - rx * rz;
        const int back = py == ry - 1 ? -1 : me + rx * rz;
        const int leftup = (px == 0 || pz == rz - 1) ? -1 : me - 1 + rx;
        const int rightup = (px == rx - 1 || pz == rz - 1) ? -1 : me + 1 + rx;
        const int leftdown = (px == 0 || pz == 0) ? -1 : me - 1 - rx;
        const int rightdown = (px == rx - 1 || pz == 0) ? -1 : me + 1 - rx;
        const int frontup = (py == 0 || pz == rz - 1) ? -1 : me - rx * rz + rx;
        const int frontdown = (py == 0 || pz == 0) ? -1 : me - rx * rz - rx;
        const int frontleft = (py == 0 || px == 0) ? -1 : me - rx * rz - 1;
        const int frontright = (py == 0 || px == rx - 1) ? -1 : me - rx * rz + 1;
        const int backup = (py == ry - 1 || pz == rz - 1) ? -1 : me + rx * rz + rx;
        const int backdown = (py == ry - 1 || pz == 0) ? -1 : me + rx * rz - rx;
        const int backleft = (py == ry - 1 || px == 0) ? -1 : me + rx * rz - 1;
        const int backright = (py == ry - 1 || px == rx - 1) ? -1 : me + rx * rz + 1;
        const int frontleftdown = (py == 0 || px == 0 || pz == 0) ? -1 : me - rx * rz - rx - 1;
        const int frontrightdown = (py == 0 || px == rx - 1 || pz == 0) ? -1 : me - rx * rz - rx + 1;
        const int frontrightup = (py == 0 || px == rx - 1 || pz == rz - 1) ? -1 : me - rx * rz + rx + 1;
        const int frontleftup = (py == 0 || px == 0 || pz == rz - 1) ? -1 : me - rx * rz + rx - 1;
        const int backleftdown = (py == ry - 1 || px == 0 || pz == 0) ? -1 : me + rx * rz - rx - 1;
        const int backrightdown = (py == ry - 1 || px == rx - 1 || pz == 0) ? -1 : me + rx * rz - rx + 1;
        const int backrightup = (py == ry - 1 || px == rx - 1 || pz == rz - 1) ? -1 : me + rx * rz + rx + 1;
        const int backleftup = (py == ry - 1 || px == 0 || pz == rz - 1) ? -1 : me + rx * rz + rx - 1;
        const int v[26] = {left, right, down, up, front, back,
                           frontup, frontright, frontdown, frontleft, backup, backright, backdown, backleft,
                           leftup, rightup, leftdown, rightdown,
                           frontleftup, frontleftdown, frontrightup, frontrightdown,
                           backleftup, backleftdown, backrightup, backrightdown};
        std::memcpy(nb, v, sizeof v);
    }

    // local extents (tgv/src/lbm.hpp:164-166)
    auto ll = [&](int r, int d) {
        int qx, qy, qz;
        coords(r, qx, qy, qz);
        const int gl[3] = {glx, gly, glz}, rr[3] = {rx, ry, rz}, pp[3] = {qx, qy, qz};
        return (pp[d] - gl[d] % rr[d] >= 0) ? gl[d] / rr[d] : gl[d] / rr[d] + 1;
    };
    for (int d = 0; d < 3; ++d) o->l_l[d] = ll(me, d);
    // gathered extents, indexed as the reference indexes them
    std::vector<int> his[3];
    for (int d = 0; d < 3; ++d) {
        his[d].resize(nranks);
        for (int r = 0; r < nranks; ++r) his[d][r] = ll(r, d);
    }
    int hi[3] = {0, 0, 0};
    if (periodic) {
        // ax_his[i][j][k] = x_his[i + j rx + k rx ry] (tgv/src/lbm.hpp:196-221)
        for (int i = 0; i <= px; ++i) hi[0] += his[0][i];
        for (int j = 0; j <= py; ++j) hi[1] += his[1][j * rx];
        for (int k = 0; k <= pz; ++k) hi[2] += his[2][k * rx * ry];
    } else {
        // x_his[i], y_his[j], z_his[k] (lid_driven_cavity/src/lbm.hpp:168-193)
        for (int i = 0; i <= px; ++i) hi[0] += his[0][i];
        for (int j = 0; j <= py; ++j) hi[1] += his[1][j];
        for (int k = 0; k <= pz; ++k) hi[2] += his[2][k];
    }
    for (int d = 0; d < 3; ++d) {
        o->lo[d] = hi[d] - o->l_l[d];
        o->hi[d] = hi[d] - 1;
    }
    return IMEXLB_OK;
}

// Tecplot binary (#!TDV112) file of one ordered zone with the variables
// x y z u v w p, the layout LBM::MPIoutput writes (tgv/src/lbm.cpp:335-669,
// lid_driven_cavity/src/lbm.cpp:351-674): a 332-byte header -- title, seven
// variable names, zone "ZONE 001" with the solution time and I/J/K, all
// variables FP64 -- then the 14 min/max doubles and the seven arrays, x
// fastest.  Coordinates are x / (nx - 1) etc. on [0, 1]; min/max of u v w p
// are taken over the given arrays.
int imexlb_tecplot_write(const char *path, double solution_time, int nx, int ny, int nz, const double *p,
                         const double *u, const double *v, const double *w)
{
    if (!path || nx < 1 || ny < 1 || nz < 1 || !p || !u || !v || !w) return IMEXLB_EINVAL;
    const size_t n = (size_t)nx * ny * nz;
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return IMEXLB_EIO;
    bool ok = true;
    auto I = [&](int32_t v_) { ok = ok && std::fwrite(&v_, 4, 1, fp) == 1; };
    auto F = [&](float v_) { ok = ok && std::fwrite(&v_, 4, 1, fp) == 1; };
    auto D = [&](double v_) { ok = ok && std::fwrite(&v_, 8, 1, fp) == 1; };
    ok = std::fwrite("#!TDV112", 1, 8, fp) == 8;
    I(1);  // byte order
    I(0);  // file type: full
    I(0);  // empty title
    I(7);
    for (int c : {'x', 'y', 'z', 'u', 'v', 'w', 'p'}) {
        I(c);
        I(0);
    }
    F(299.0f);  // zone marker
    for (int c : {'Z', 'O', 'N', 'E', ' ', '0', '0', '1'}) I(c);
    I(0);
    I(-1);              // parent zone
    D(solution_time);   // solution time
    I(0);               // strand (static)
    I(-1);              // zone colour
    I(0);               // ordered zone
    I(0);               // no variable location list
    I(0);               // no raw face neighbours
    I(0);               // no user face connections
    I(nx);
    I(ny);
    I(nz);
    I(0);               // no auxiliary data
    F(357.0f);          // end of header
    F(299.0f);          // data section
    for (int k = 0; k < 7; ++k) I(2);  // FP64
    I(0);               // no passive variables
    I(0);               // no shared variables
    I(-1);              // no shared connectivity
    const double *fields[4] = {u, v, w, p};
    double mn[4], mx[4];
    for (int k = 0; k < 4; ++k) {
        mn[k] = *std::min_element(fields[k], fields[k] + n);
        mx[k] = *std::max_element(fields[k], fields[k] + n);
    }
    for (int k = 0; k < 3; ++k) {
        D(0.0);
        D(1.0);
    }
    for (int k = 0; k < 4; ++k) {
        D(mn[k]);
        D(mx[k]);
    }
    // coordinates, then the fields, x fastest
    std::vector<double> row((size_t)nx);
    const int ext[3] = {nx, ny, nz};
    for (int axis = 0; axis < 3 && ok; ++axis) {
        const double den = ext[axis] > 1 ? (double)(ext[axis] - 1) : 1.0;
        for (int z = 0; z < nz && ok; ++z)
            for (int y = 0; y < ny && ok; ++y) {
                for (int x = 0; x < nx; ++x) {
                    const int c = axis == 0 ? x : (axis == 1 ? y : z);
                    row[x] = (double)c / den;
                }
                ok = std::fwrite(row.data(), 8, (size_t)nx, fp) == (size_t)nx;
            }
    }
    for (int k = 0; k < 4 && ok; ++k) ok = std::fwrite(fields[k], 8, n, fp) == n;
    if (std::fclose(fp) != 0) ok = false;
    return ok ? IMEXLB_OK : IMEXLB_EIO;
}

}  // extern "C"
