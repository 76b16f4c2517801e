// D3Q27 lattice tables, shared by host and device code.
//
// Order, weights and bounce-back partners are the reference's
// (tgv/src/lbm.cpp:27-53 weights t, :55-81 bb, :84-190 velocities e); the same
// table appears in lid_driven_cavity/src/lbm.cpp and sphere/src/lbm.cpp.  The
// order has to match so that population i means the same thing on both sides
// of the C-ABI (imexlb_get_f / imexlb_set_f exchange populations by index).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define IMX_HD __host__ __device__
#else
#define IMX_HD
#endif

namespace imx {

constexpr int Q = 27;
constexpr int GHOST = 3;            // reference ghost depth (tgv/src/lbm.hpp:22)
constexpr double PI_REF = 3.1415926; // reference pi (tgv/src/lbm.cpp:5)

constexpr int EX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
constexpr int EY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1, -1, 1};
constexpr int EZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1, -1, 1, -1, 1};
constexpr int BB[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17, 20, 19, 22, 21, 24, 23, 26, 25};

// Weight class of direction i: 0 rest, 1 face, 2 edge, 3 corner.
constexpr int WCLASS[Q] = {0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3, 3};
constexpr double WVAL[4] = {8.0 / 27.0, 2.0 / 27.0, 1.0 / 54.0, 1.0 / 216.0};

IMX_HD constexpr double w(int i) { return WVAL[WCLASS[i]]; }

// FP16 storage keeps (f_i - w_i S) * 2^k with 2^k nearest to 1/w_i, so every
// direction's stored value is O(|e.u|) and far from the FP16 subnormals.
constexpr float H_SCALE[4] = {4.0f, 16.0f, 64.0f, 256.0f};
constexpr float H_INV_SCALE[4] = {0.25f, 0.0625f, 0.015625f, 0.00390625f};

// The d-th direction, in index order, of the nine with e_AXIS * SIGN > 0
// (d < 9), and a direction's velocity and opposite, for kernels that give each
// direction its own row of workgroups.
template <int AXIS, int SIGN>
IMX_HD inline int nth_dir(int d)
{
    int q = 0, cnt = 0;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        const int e = AXIS == 0 ? EX[k] : AXIS == 1 ? EY[k] : EZ[k];
        if (e * SIGN > 0) {
            if (cnt == d) q = k;
            ++cnt;
        }
    }
    return q;
}
// The nine directions with e_x = +1 and with e_x = -1, in index order
// (nth_dir<0, 1> and nth_dir<0, -1>), as tables: kernels that index a
// register array with them need the index known at compile time.
constexpr int XDIR_P[9] = {1, 7, 9, 11, 13, 19, 21, 23, 25};
constexpr int XDIR_M[9] = {2, 8, 10, 12, 14, 20, 22, 24, 26};
constexpr bool xdir_tables_ok()
{
    for (int r = 0; r < 9; ++r) {
        int p = 0, m = 0, cp = 0, cm = 0;
        for (int k = 0; k < Q; ++k) {
            if (EX[k] > 0) {
                if (cp == r) p = k;
                ++cp;
            }
            if (EX[k] < 0) {
                if (cm == r) m = k;
                ++cm;
            }
        }
        if (XDIR_P[r] != p || XDIR_M[r] != m) return false;
    }
    return true;
}
static_assert(xdir_tables_ok(), "XDIR_P / XDIR_M must list the e_x = +1 / -1 directions in index order");

// The rank of direction q among the nine with its e_x (e_x != 0): the d
// with nth_dir<0, EX[q]>(d) == q.
IMX_HD constexpr int x_rank(int q)
{
    int r = 0;
    for (int k = 0; k < q; ++k) r += EX[k] == EX[q] ? 1 : 0;
    return r;
}
IMX_HD inline void dir_of(int q, int (&e)[3], int &bq)
{
    e[0] = e[1] = e[2] = bq = 0;
#pragma unroll
    for (int k = 0; k < Q; ++k)
        if (k == q) {
            e[0] = EX[k];
            e[1] = EY[k];
            e[2] = EZ[k];
            bq = BB[k];
        }
}

}  // namespace imx
