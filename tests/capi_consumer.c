/* A plain C99 program against include/imexlb.h, as a C or FFI caller of the
 * library would be built: it uses only the header's C types and calls the
 * entry points that need no GPU (configuration checks, input.in, the rank
 * grid, the Tecplot writer, status strings).  Built and run by
 * tests/test_capi_c.py.  Prints one "ok" line per check and exits 0 when
 * every check holds. */
#include <stdio.h>
#include <string.h>

#include "imexlb.h"

#define CHECK(cond)                                              \
    do {                                                         \
        if (!(cond)) {                                           \
            fprintf(stderr, "check failed line %d: %s\n", __LINE__, #cond); \
            return 1;                                            \
        }                                                        \
    } while (0)

int main(int argc, char **argv)
{
    const char *input = argc > 1 ? argv[1] : "input.in";
    const char *plt = argc > 2 ? argv[2] : "output3.plt";

    imexlb_config c;
    CHECK(imexlb_config_init(&c) == IMEXLB_OK);
    CHECK(c.case_id == IMEXLB_CASE_TGV && c.dtype == IMEXLB_F64);
    CHECK(c.addr_mode == -1 && c.nt_loads == -1);
    imexlb_handle *h = NULL;
    c.addr_mode = 3; /* no such address mode: refused before any device call */
    CHECK(imexlb_create(&c, &h) == IMEXLB_EINVAL && h == NULL);
    c.addr_mode = -1;
    c.cells_per_thread = 4; /* 1, 2 or 3 cells per thread (0: default) */
    CHECK(imexlb_create(&c, &h) == IMEXLB_EINVAL && h == NULL);
    c.cells_per_thread = 0;
    c.xcd_remap = 2; /* -1 default, 0 or 1 */
    CHECK(imexlb_create(&c, &h) == IMEXLB_EINVAL && h == NULL);
    c.xcd_remap = -1;
    c.nt_loads = 5; /* -1 default, 0 or 1 */
    CHECK(imexlb_create(&c, &h) == IMEXLB_EINVAL && h == NULL);
    c.nt_loads = -1;
    c.nx = 1; /* extents below 2 are refused before any device call */
    CHECK(imexlb_create(&c, &h) == IMEXLB_EINVAL && h == NULL);
    CHECK(strlen(imexlb_last_error(NULL)) > 0);
    printf("ok create rejects: %s\n", imexlb_last_error(NULL));
    CHECK(strcmp(imexlb_strerror(IMEXLB_OK), imexlb_strerror(IMEXLB_EINVAL)) != 0);

    imexlb_system s;
    CHECK(imexlb_system_read(input, IMEXLB_CASE_CAVITY, 1, &s) == IMEXLB_OK);
    CHECK(s.sx == 24 && s.sy == 20 && s.sz == 16 && s.Time == 20 && s.inter == 10);
    printf("ok system tau %.6f\n", s.tau);

    imexlb_comm m;
    CHECK(imexlb_comm_layout(8, 5, 1, 64, 64, 64, &m) == IMEXLB_OK);
    CHECK(m.nranks == 8 && m.me == 5 && m.l_l[0] * m.rx == 64);
    printf("ok comm rx %d ry %d rz %d\n", m.rx, m.ry, m.rz);

    double p[6], u[6], v[6], w[6];
    for (int k = 0; k < 6; ++k) {
        p[k] = k;
        u[k] = -k;
        v[k] = 0.5 * k;
        w[k] = 2.0;
    }
    CHECK(imexlb_tecplot_write(plt, 3.0, 3, 2, 1, p, u, v, w) == IMEXLB_OK);
    FILE *fp = fopen(plt, "rb");
    CHECK(fp != NULL);
    fseek(fp, 0, SEEK_END);
    CHECK(ftell(fp) == 332 + 7 * 6 * 8);
    fclose(fp);
    printf("ok tecplot\n");
    return 0;
}
