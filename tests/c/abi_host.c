/* A C99 (GNU) consumer of include/cfdx.h: the host-side entry points a reference-side driver
 * calls before it touches a device (parameters, block sizing, C-grid geometry and metrics, tile
 * shapes, restart files, error reporting). Built and run by tests/test_c_header.py; prints one
 * line per check and exits non-zero on the first failure. */
#include "cfdx.h"

#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define CHECK(cond, ...)                          \
    do {                                          \
        if (!(cond)) {                            \
            fprintf(stderr, "FAIL: " __VA_ARGS__); \
            fprintf(stderr, "\n");                \
            return 1;                             \
        }                                         \
    } while (0)

int main(int argc, char **argv)
{
    const char *dir = argc > 1 ? argv[1] : ".";

    /* the reference's constants (Viscous.cl:30-33, :490-491, Weno.cl:59) */
    cfdx_params p;
    cfdx_params_default(&p);
    CHECK(p.mach == 1.5 && p.reyl == 500.0 && p.pr == 0.72 && p.free_t == 273.0 && p.c_mts == 0.03 && p.c_t == 10.0 &&
              p.eps == 1e-7 && p.precision == CFDX_FP64 && p.viscous == 1 && p.sgs == CFDX_SGS_MTS && p.c_s == 0.17,
          "params_default");
    CHECK(sizeof(cfdx_params) == 80 && offsetof(cfdx_params, c_s) == 72, "cfdx_params layout");
    printf("params ok\n");

    /* layouts of functions.h:49-112 */
    CHECK(sizeof(cfdx_flow) == 72 && sizeof(cfdx_vec5) == 40 && sizeof(cfdx_grid) == 56 && sizeof(cfdx_roe) == 288 &&
              sizeof(cfdx_metric) == 504 && sizeof(cfdx_deter) == 48,
          "struct sizes");
    printf("structs ok\n");

    /* a NACA 2412 C-grid block: descriptor, padded sizes, coordinates, metrics */
    const cfdx_cgrid g = {0.02, 0.4, 0.12, 20.0, 20.0, 1e-3, 0.25, 12, 0};
    cfdx_block_desc d;
    CHECK(cfdx_geom_cgrid_desc(48, 16, 4, CFDX_MIXED, &g, 2.0 * M_PI / 180.0, &d) == CFDX_OK, "cgrid_desc: %s",
          cfdx_last_error());
    int32_t dims[3], off[3];
    CHECK(cfdx_block_padded(&d, dims, off) == CFDX_OK, "padded");
    CHECK(dims[0] == 54 && dims[1] == 22 && dims[2] == 4 && off[0] == 3 && off[1] == 3 && off[2] == 0,
          "padded dims %d %d %d off %d %d %d", dims[0], dims[1], dims[2], off[0], off[1], off[2]);
    const size_t n = (size_t)dims[0] * dims[1] * dims[2];
    double *xyz = malloc(3 * n * sizeof(double)), *geo = malloc(22 * n * sizeof(double));
    CHECK(xyz && geo, "malloc");
    CHECK(cfdx_geom_cgrid_coords(&d, &g, xyz) == CFDX_OK, "cgrid_coords: %s", cfdx_last_error());
    const double L[3] = {0.0, 0.0, g.span};
    CHECK(cfdx_geom_block_metrics(&d, xyz, L, geo) == CFDX_OK, "block_metrics: %s", cfdx_last_error());
    size_t bad = 0;
    for (size_t q = 0; q < 22 * n; ++q) bad += !isfinite(geo[q]);
    for (size_t q = 0; q < n; ++q) bad += !(geo[q] > 0.0);   /* det of every stored node */
    CHECK(bad == 0, "geometry: %zu values not finite or non-positive dets", bad);
    printf("cgrid ok %d x %d x %d stored\n", dims[0], dims[1], dims[2]);

    /* device bytes of a block: mixed below FP64; a refused descriptor gives a negative code and a message */
    const int64_t b32 = cfdx_block_bytes(&d);
    d.precision = CFDX_FP64;
    const int64_t b64 = cfdx_block_bytes(&d);
    CHECK(b32 > 0 && b64 > b32 && b64 >= (int64_t)n * 520, "block_bytes %lld %lld", (long long)b32, (long long)b64);
    d.ni = 0;
    CHECK(cfdx_block_bytes(&d) < 0 && strlen(cfdx_last_error()) > 0, "block_bytes of ni = 0");
    printf("bytes ok (%s)\n", cfdx_last_error());

    /* tile shapes: every variant is 256 threads */
    for (int kind = 0; kind < 3; ++kind) {
        int32_t count = 0, shapes[32];
        CHECK(cfdx_tile_variants(kind, &count, NULL) == CFDX_OK && count > 0 && count <= 16, "tile count");
        CHECK(cfdx_tile_variants(kind, &count, shapes) == CFDX_OK, "tile shapes");
        for (int v = 0; v < count; ++v) CHECK(shapes[2 * v] * shapes[2 * v + 1] == 256, "tile %d/%d", kind, v);
    }
    CHECK(cfdx_tile_variants(3, NULL, NULL) < 0, "tile kind 3");
    printf("tiles ok\n");

    /* restart files: 17 digits round-trip doubles; the newer complete slot is found */
    cfdx_flow f[7], r[7];
    for (int q = 0; q < 7; ++q) {
        f[q].u = 0.1 * q + 1.0 / 3.0; f[q].v = -0.2 * q; f[q].w = 1e-9 * q; f[q].p = 0.7 + q;
        f[q].t = 1.0 + 1.0 / (q + 3); f[q].rho = 1.25 - 0.01 * q; f[q].mu = 1.1e-3; f[q].a = 0.0; f[q].e = 2.5 + q;
    }
    CHECK(cfdx_write_restart(dir, 1, 7, f, 40, 17) == CFDX_OK, "write_restart 1: %s", cfdx_last_error());
    CHECK(cfdx_write_restart(dir, 2, 7, f, 50, 17) == CFDX_OK, "write_restart 2: %s", cfdx_last_error());
    int32_t slot = 0;
    int64_t iter = 0;
    CHECK(cfdx_restart_latest(dir, 7, &slot, &iter) == CFDX_OK && slot == 2 && iter == 50, "restart_latest %d %lld",
          slot, (long long)iter);
    CHECK(cfdx_read_restart(dir, slot, 7, r, &iter) == CFDX_OK && iter == 50, "read_restart: %s", cfdx_last_error());
    for (int q = 0; q < 7; ++q)
        CHECK(r[q].u == f[q].u && r[q].v == f[q].v && r[q].w == f[q].w && r[q].p == f[q].p && r[q].t == f[q].t &&
                  r[q].rho == f[q].rho && r[q].mu == f[q].mu && r[q].e == f[q].e,
              "restart node %d", q);
    printf("restart ok\n");
    free(xyz);
    free(geo);
    printf("ALL OK\n");
    return 0;
}
