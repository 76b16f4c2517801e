// Host-side unit checks of the address-mode choice (kernels.h) and the
// placement window's chunk size (window.h).  Built with hipcc's host pass
// (--cuda-host-only) and run on the CPU; prints "ok <name>" per check.
#include <cstdio>
#include <cstdint>

#include "../imexlb_amd/csrc/kernels.h"
#include "../imexlb_amd/csrc/window.h"

#define CHECK(c)                                                              \
    do {                                                                      \
        if (!(c)) {                                                           \
            std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
            return 1;                                                         \
        }                                                                     \
    } while (0)

int main()
{
    using imx::fused_addr_mode;
    const uint64_t g28 = uint64_t(1) << 28, g29 = uint64_t(1) << 29, g30 = uint64_t(1) << 30,
                   g31 = uint64_t(1) << 31;
    // grouped bases while 16 bytes of stride per cell span at most 4 GiB
    CHECK(fused_addr_mode(8, g28, 2) == 2);
    CHECK(fused_addr_mode(8, g28 + 1, 2) == 1);
    CHECK(fused_addr_mode(2, g28, 2) == 2 && fused_addr_mode(4, g28, 2) == 2);
    // a base per population while one population spans at most 4 GiB
    CHECK(fused_addr_mode(8, g29, 2) == 1 && fused_addr_mode(8, g29 + 1, 2) == 0);
    CHECK(fused_addr_mode(4, g30, 2) == 1 && fused_addr_mode(4, g30 + 1, 2) == 0);
    CHECK(fused_addr_mode(2, g31, 2) == 1 && fused_addr_mode(2, g31 + 1, 2) == 0);
    // 512^3 with the default pad takes the grouped mode in every storage type
    const uint64_t ps512 = (uint64_t(1) << 27) + 2304 / 8;
    CHECK(fused_addr_mode(8, ps512, 2) == 2 && fused_addr_mode(2, ps512, 2) == 2);
    // the cap
    CHECK(fused_addr_mode(8, 1000, 1) == 1 && fused_addr_mode(8, 1000, 0) == 0);
    CHECK(fused_addr_mode(8, 1000, -1) == 0);
    std::printf("ok fused_addr_mode\n");

    // the cells per thread a launch runs with: pairs on periodic lattices
    // with nx even and grouped bases only, two cells with grouped bases only
    using imx::fused_cells_per_thread;
    CHECK(fused_cells_per_thread(3, imx::BC_PERIODIC, 512, 2) == 3);
    CHECK(fused_cells_per_thread(3, imx::BC_PERIODIC, 2, 2) == 3);
    CHECK(fused_cells_per_thread(3, imx::BC_PERIODIC, 67, 2) == 1);
    CHECK(fused_cells_per_thread(3, imx::BC_CAVITY, 128, 2) == 1);
    CHECK(fused_cells_per_thread(3, imx::BC_FAITHFUL, 256, 2) == 1);
    CHECK(fused_cells_per_thread(3, imx::BC_PERIODIC, 512, 1) == 1);
    CHECK(fused_cells_per_thread(2, imx::BC_CHANNEL, 512, 2) == 2);
    CHECK(fused_cells_per_thread(2, imx::BC_PERIODIC, 512, 1) == 1);
    CHECK(fused_cells_per_thread(1, imx::BC_PERIODIC, 512, 2) == 1);
    std::printf("ok fused_cells_per_thread\n");

    using imx::window_chunk;
    const size_t mib2 = size_t(2) << 20, gib = size_t(1) << 30;
    // the step itself while the window holds at most 512 chunks
    CHECK(window_chunk(125 * gib, 2 * gib, mib2) == 2 * gib);
    CHECK(window_chunk(103 * gib, gib, mib2) == gib);
    CHECK(window_chunk(22 << 20, mib2, mib2) == mib2);
    // a multiple of the step beyond that, and a multiple of the granularity
    const size_t c = window_chunk(64 * gib, mib2, mib2);
    CHECK(c % mib2 == 0 && (64 * gib) / c <= 512 && (64 * gib) / (c - mib2) > 512);
    CHECK(window_chunk(3 * mib2, 3 << 20, mib2) == 4 << 20);
    std::printf("ok window_chunk\n");
    return 0;
}
