/*
 * abi_harness.c -- drives include/aeroflux.h in the order the Go binding
 * (go/aeroflux) issues its calls, for a caller that cannot run Go here:
 *
 *   Config.Normalize        af_config_normalize
 *   New                     af_new, af_arena_bytes, af_field_desc x 8
 *   Mesh.Data(U/V/W) writes af_field_read (the mirror refresh), then the
 *                           first Step's flush: af_field_write of U, V, W
 *   Step x steps            af_step, one call per Step
 *   Mesh.ReadData (8)       af_field_read
 *   MaxDivergence, KineticEnergy, MaxVelocity, CFL
 *   EnableSparseSolver      af_enable_sparse_solver (must fail, reference text)
 *   Close, Close            af_close + af_free, then nothing (idempotent)
 *
 * usage: abi_harness IN OUT NX NY NZ H DT NU STEPS
 *   IN:  U, V, W padded fields, little-endian doubles, back to back
 *   OUT: the eight fields after the steps in field order, then MaxDivergence,
 *        KineticEnergy, MaxVelocity, CFL
 * Exit status 0 on success; otherwise the failing call is printed.
 * With IN = "-" it only checks the calls that need no device (normalisation
 * and its error path) and exits 0.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "aeroflux.h"

static int check(int rc, const char *what)
{
    if (rc != AF_OK) fprintf(stderr, "%s: rc=%d: %s\n", what, rc, af_last_error());
    return rc;
}

static int host_only(void)
{
    af_config c;
    memset(&c, 0, sizeof c);
    c.NX = 0; c.NY = 8; c.NZ = 8;
    if (af_config_normalize(&c) != AF_ERR_CONFIG ||
        strncmp(af_last_error(), "aeroflux: ", 10) != 0) {
        fprintf(stderr, "normalize accepted a zero extent\n");
        return 1;
    }
    memset(&c, 0, sizeof c);
    c.NX = c.NY = c.NZ = 16;
    if (check(af_config_normalize(&c), "normalize")) return 1;
    if (c.GaussSeidelIters != 24 || c.Ghost != 1 || c.H != 1.0 / 16) {
        fprintf(stderr, "normalize defaults differ\n");
        return 1;
    }
    if (af_mac_bytes(16, 16, 16, 1) <= 0) return 1;
    printf("host-only checks passed\n");
    return 0;
}

int main(int argc, char **argv)
{
    if (argc >= 2 && strcmp(argv[1], "-") == 0) return host_only();
    if (argc != 10) {
        fprintf(stderr, "usage: %s IN OUT NX NY NZ H DT NU STEPS\n", argv[0]);
        return 2;
    }
    af_config c;
    memset(&c, 0, sizeof c);
    c.NX = atol(argv[3]); c.NY = atol(argv[4]); c.NZ = atol(argv[5]);
    c.H = atof(argv[6]); c.DT = atof(argv[7]); c.Viscosity = atof(argv[8]);
    const long steps = atol(argv[9]);
    if (check(af_config_normalize(&c), "normalize")) return 1;

    af_engine *e = NULL;
    if (check(af_new(&c, &e), "new")) return 1;
    if (af_arena_bytes(e) != af_mac_bytes(c.NX, c.NY, c.NZ, c.Ghost)) {
        fprintf(stderr, "arena bytes differ from MACBytes\n");
        return 1;
    }
    af_field f[AF_FIELD_COUNT];
    double *mirror[AF_FIELD_COUNT];
    for (int id = 0; id < AF_FIELD_COUNT; id++) {
        if (check(af_field_desc(e, id, &f[id]), "field_desc")) return 1;
        mirror[id] = (double *)malloc(sizeof(double) * (size_t)f[id].Count);
        if (!mirror[id]) return 1;
    }

    FILE *in = fopen(argv[1], "rb");
    if (!in) return 1;
    for (int id = AF_FIELD_U; id <= AF_FIELD_W; id++) {
        /* Mesh.Data: refresh the mirror, then the caller writes through it. */
        if (check(af_field_read(e, id, mirror[id], f[id].Count), "field_read")) return 1;
        if (fread(mirror[id], sizeof(double), (size_t)f[id].Count, in) != (size_t)f[id].Count) {
            fprintf(stderr, "short input\n");
            return 1;
        }
    }
    fclose(in);
    /* The first Step uploads the mirrors written through Data. */
    for (int id = AF_FIELD_U; id <= AF_FIELD_W; id++)
        if (check(af_field_write(e, id, mirror[id], f[id].Count), "field_write")) return 1;
    if (check(af_apply_velocity_bc(e), "apply_velocity_bc")) return 1;
    for (long s = 0; s < steps; s++)
        if (check(af_step(e), "step")) return 1;

    FILE *out = fopen(argv[2], "wb");
    if (!out) return 1;
    for (int id = 0; id < AF_FIELD_COUNT; id++) {
        af_field cur;
        if (check(af_field_desc(e, id, &cur), "field_desc")) return 1;
        if (check(af_field_read(e, id, mirror[id], cur.Count), "field_read")) return 1;
        fwrite(mirror[id], sizeof(double), (size_t)cur.Count, out);
    }
    double d[4];
    if (check(af_max_divergence(e, &d[0]), "max_divergence") ||
        check(af_kinetic_energy(e, &d[1]), "kinetic_energy") ||
        check(af_max_velocity(e, &d[2]), "max_velocity") || check(af_cfl(e, &d[3]), "cfl"))
        return 1;
    fwrite(d, sizeof(double), 4, out);
    fclose(out);

    if (af_enable_sparse_solver(e) != AF_ERR_UNSUPPORTED ||
        strcmp(af_last_error(),
               "aeroflux: sparse direct solver requires darwin with cgo (Accelerate); the "
               "Red-Black iterative solver is used instead") != 0) {
        fprintf(stderr, "EnableSparseSolver: unexpected result '%s'\n", af_last_error());
        return 1;
    }
    if (check(af_close(e), "close")) return 1;
    if (af_step(e) != AF_ERR_CLOSED) {
        fprintf(stderr, "step after close did not fail\n");
        return 1;
    }
    if (check(af_close(e), "second close")) return 1;
    af_free(e);
    for (int id = 0; id < AF_FIELD_COUNT; id++) free(mirror[id]);
    printf("abi harness: %ld steps, max|div| %.17g\n", steps, d[0]);
    return 0;
}
