/* A plain C caller of the library: a host written against include/skirt7_amd.h alone, with no Python and
 * no C++ (the way SKIRT7's own code, or any other FFI, would bind it). It builds a Cartesian grid and a
 * binary tree, asks for DustGrid::path on a few rays and for whichcell, provokes one error, and prints
 * everything as lines that tests/test_gpu_c_abi.py compares with the Python binding's results. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "skirt7_amd.h"

static int report(const char* what, int st)
{
    if (st != 0) printf("error %s %d %s\n", what, st, sg_last_error());
    return st;
}

static void print_paths(const char* name, sg_grid* g, int n, const double* rk)
{
    enum { MAXSEG = 256 };
    int32_t counts[8];
    int32_t* m = (int32_t*)malloc(sizeof(int32_t) * n * MAXSEG);
    double* ds = (double*)malloc(sizeof(double) * n * MAXSEG);
    if (report("paths", sg_grid_paths(g, n, rk, MAXSEG, SG_FP64, counts, m, ds)) == 0)
        for (int i = 0; i < n; i++) {
            printf("path %s %d %d", name, i, counts[i]);
            for (int s = 0; s < counts[i]; s++) printf(" %d:%.17g", m[i * MAXSEG + s], ds[i * MAXSEG + s]);
            printf("\n");
        }
    free(m);
    free(ds);
}

int main(void)
{
    if (report("init", sg_init(0))) return 1;
    char name[128];
    if (sg_device_name(name, (int)sizeof name) == 0) printf("device %s\n", name);
    const double ext[6] = {-1.0, -1.0, -1.0, 1.0, 1.0, 1.0};
    double xv[5], yv[4], zv[6];
    for (int i = 0; i < 5; i++) xv[i] = -1.0 + 0.5 * i;
    for (int i = 0; i < 4; i++) yv[i] = -1.0 + 2.0 * i / 3.0;
    for (int i = 0; i < 6; i++) zv[i] = -1.0 + 0.4 * i;
    sg_grid* cart = NULL;
    if (report("cartesian", sg_grid_create_cartesian(ext, xv, 5, yv, 4, zv, 6, &cart))) return 1;
    int64_t nc = 0;
    sg_grid_num_cells(cart, &nc);
    printf("cells cartesian %lld\n", (long long)nc);
    const double rk[4 * 6] = {-0.9, -0.8, -0.7, 0.48, 0.6, 0.64,    /* inside, oblique */
                              -3.0, 0.1, 0.2, 1.0, 0.0, 0.0,        /* outside, entering along x */
                              0.3, -0.2, 0.9, -0.6, 0.0, -0.8,      /* inside, in the xz plane */
                              2.0, 2.0, 2.0, 1.0, 0.0, 0.0};        /* outside, moving away */
    print_paths("cartesian", cart, 4, rk);
    const double pts[3 * 3] = {0.1, 0.2, 0.3, -0.99, 0.99, -0.5, 5.0, 0.0, 0.0};
    int32_t cells[3];
    if (report("whichcell", sg_grid_whichcell(cart, 3, pts, cells)) == 0)
        printf("whichcell cartesian %d %d %d\n", cells[0], cells[1], cells[2]);
    /* a binary tree: the root and its first child subdivided (5 nodes) */
    const uint8_t flags[5] = {1, 1, 0, 0, 0};
    sg_grid* tree = NULL;
    if (report("tree", sg_grid_create_tree(2, ext, flags, 5, SG_NEIGHBOR, &tree))) return 1;
    sg_grid_num_cells(tree, &nc);
    printf("cells tree %lld\n", (long long)nc);
    print_paths("tree", tree, 4, rk);
    /* an invalid call: a Cartesian grid without bins */
    sg_grid* bad = NULL;
    int st = sg_grid_create_cartesian(ext, xv, 1, yv, 4, zv, 6, &bad);
    printf("invalid status %d message %s\n", st, sg_last_error());
    sg_grid_destroy(tree);
    sg_grid_destroy(cart);
    printf("done\n");
    return 0;
}
