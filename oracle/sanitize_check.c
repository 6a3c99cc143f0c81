/* sanitize_check.c -- drives the oracle (dmd_oracle.c) through every entry
 * point on small systems, for a build under AddressSanitizer and
 * UndefinedBehaviorSanitizer (oracle/Makefile, target `sanitize`).  TEST
 * INFRASTRUCTURE ONLY.  Any out-of-bounds access, leak, signed overflow or
 * misaligned load aborts the run with a non-zero status; the printed checks
 * (finite energies, pair sets of the two list builders equal) catch the
 * grosser mistakes.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dmd_oracle.h"

static int g_fail = 0;

static void expect(int ok, const char *what)
{
    if (!ok) {
        fprintf(stderr, "sanitize_check: FAILED: %s\n", what);
        g_fail = 1;
    }
}

static uint64_t g_rng = 0x2545F4914F6CDD1Dull;
static double uni(void)
{
    g_rng ^= g_rng << 13;
    g_rng ^= g_rng >> 7;
    g_rng ^= g_rng << 17;
    return (double)(g_rng >> 11) * (1.0 / 9007199254740992.0);
}

/* n atoms of a charged two-type LJ system in a cubic box L, chained by
 * bonds, angles, dihedrals and impropers, with exclusions and PME */
static void check_sim(int64_t n, double L, int coulomb, int thermo, int baro, int cons,
                      int physics)
{
    double *buf = calloc((size_t)n * 8, sizeof(double));
    double *x = buf, *y = buf + n, *z = buf + 2 * n, *vx = buf + 3 * n, *vy = buf + 4 * n,
           *vz = buf + 5 * n, *m = buf + 6 * n, *q = buf + 7 * n;
    int32_t *types = calloc((size_t)n, sizeof(int32_t));
    const int64_t nb = n - 1, na = n - 2, nd = n - 3;
    int32_t *ib = calloc((size_t)(4 * n), sizeof(int32_t));
    double *pb = calloc((size_t)(4 * n), sizeof(double));
    int32_t *per = calloc((size_t)n, sizeof(int32_t));
    /* a lattice so no two atoms overlap */
    int side = 1;
    while ((int64_t)side * side * side < n) ++side;
    for (int64_t i = 0; i < n; ++i) {
        x[i] = (i % side + 0.5 + 0.1 * (uni() - 0.5)) * L / side;
        y[i] = ((i / side) % side + 0.5 + 0.1 * (uni() - 0.5)) * L / side;
        z[i] = (i / ((int64_t)side * side) + 0.5 + 0.1 * (uni() - 0.5)) * L / side;
        vx[i] = uni() - 0.5;
        vy[i] = uni() - 0.5;
        vz[i] = uni() - 0.5;
        m[i] = 12.0 + 4.0 * (i % 3);
        q[i] = (i % 2) ? -0.3 : 0.3;
        types[i] = (int32_t)(i % 2);
    }
    for (int64_t k = 0; k < n; ++k) {
        ib[k] = (int32_t)k;
        ib[n + k] = (int32_t)((k + 1) % n);
        ib[2 * n + k] = (int32_t)((k + 2) % n);
        ib[3 * n + k] = (int32_t)((k + 3) % n);
        pb[k] = 100.0;
        pb[n + k] = L / side;
        pb[2 * n + k] = 1.9;
        pb[3 * n + k] = 0.1;
        per[k] = 1 + (int32_t)(k % 3);
    }
    const double sig[4] = {0.3, 0.31, 0.31, 0.32}, eps[4] = {0.4, 0.45, 0.45, 0.5};
    dmdx_system_desc d;
    memset(&d, 0, sizeof(d));
    d.n_atoms = n;
    d.box[0] = d.box[1] = d.box[2] = L;
    d.pos_x = x, d.pos_y = y, d.pos_z = z;
    d.vel_x = vx, d.vel_y = vy, d.vel_z = vz;
    d.masses = m;
    d.charges = q;
    d.atom_types = types;
    d.use_lj = 1;
    d.lj_n_types = 2;
    d.lj_sigma = sig;
    d.lj_epsilon = eps;
    d.lj_cutoff = 0.9;
    d.list_skin = 0.1;
    d.rebuild_interval = physics == DMDX_CORRECTED ? -1 : 10;
    d.step_since_rebuild = 10;
    d.n_bonds = nb, d.bond_i = ib, d.bond_j = ib + n, d.bond_k = pb, d.bond_r0 = pb + n;
    d.n_angles = na, d.angle_i = ib, d.angle_j = ib + n, d.angle_k = ib + 2 * n;
    d.angle_k_theta = pb, d.angle_theta0 = pb + 2 * n;
    d.n_dihedrals = nd, d.dih_i = ib, d.dih_j = ib + n, d.dih_k = ib + 2 * n;
    d.dih_l = ib + 3 * n, d.dih_periodicity = per, d.dih_k_phi = pb + 3 * n;
    d.dih_phi0 = pb + 3 * n;
    d.n_impropers = nd / 2, d.imp_i = ib, d.imp_j = ib + n, d.imp_k = ib + 2 * n;
    d.imp_l = ib + 3 * n, d.imp_k_phi = pb, d.imp_phi0 = pb + 3 * n;
    d.coulomb_type = coulomb;
    d.coulomb_cutoff = 0.9;
    d.ewald_coeff = 3.4;
    d.pme_grid[0] = 12, d.pme_grid[1] = 15, d.pme_grid[2] = 16;
    d.pme_order = 4;
    d.n_excl = nb, d.excl_i = ib, d.excl_j = ib + n;
    d.dt = 0.0005;
    d.thermostat = thermo;
    d.thermo_temperature = 300.0, d.thermo_tau = 0.1;
    d.andersen_frequency = 5.0;
    d.barostat = baro;
    d.baro_pressure = 1.0, d.baro_tau = 1.0, d.baro_compressibility = 4.5e-5;
    d.physics = physics;
    d.precision = DMDX_FP64;
    /* constraints between x-neighbours of a lattice row, at their initial
     * distance (SHAKE needs a reachable target) */
    int32_t *ci = calloc((size_t)n, sizeof(int32_t)), *cj = calloc((size_t)n, sizeof(int32_t));
    double *cd = calloc((size_t)n, sizeof(double));
    int64_t nc = 0;
    for (int64_t k = 0; k + 1 < n; k += 2) {
        if (k % side == side - 1) continue;
        const double dx = x[k + 1] - x[k], dy = y[k + 1] - y[k], dz = z[k + 1] - z[k];
        ci[nc] = (int32_t)k;
        cj[nc] = (int32_t)(k + 1);
        cd[nc++] = sqrt(dx * dx + dy * dy + dz * dz);
    }
    if (cons) {
        d.n_constraints = nc;
        d.con_i = ci;
        d.con_j = cj;
        d.con_distance = cd;
        d.con_tolerance = 1e-6;
        d.n_bonds = 0;
    }
    orc_sim *s = orc_sim_create(&d);
    expect(s != NULL, "orc_sim_create");
    if (s) {
        orc_sim_step(s, 25);
        double *f = calloc((size_t)n * 9, sizeof(double));
        orc_sim_get(s, f, f + n, f + 2 * n, f + 3 * n, f + 4 * n, f + 5 * n, f + 6 * n,
                    f + 7 * n, f + 8 * n);
        double pe = 0.0, sc[11];
        orc_sim_forces_exact(s, f + 6 * n, f + 7 * n, f + 8 * n, &pe);
        orc_sim_compute(s);
        orc_sim_set_positions(s, x, y, z);
        orc_sim_set_velocities(s, vx, vy, vz);
        orc_sim_step(s, 3);
        orc_sim_scalars(s, sc);
        if (!(isfinite(pe) && isfinite(sc[0]) && isfinite(sc[1])))
            fprintf(stderr, "n=%ld coulomb=%d thermo=%d baro=%d cons=%d physics=%d: pe %g %g ke %g\n",
                    (long)n, coulomb, thermo, baro, cons, physics, pe, sc[0], sc[1]);
        expect(isfinite(pe) && isfinite(sc[0]) && isfinite(sc[1]), "finite energies");
        free(f);
        orc_sim_destroy(s);
    }
    free(ci);
    free(cj);
    free(cd);
    free(buf);
    free(types);
    free(ib);
    free(pb);
    free(per);
}

static void check_lists_and_pme(void)
{
    const int64_t n = 300;
    const double L = 2.7;
    double cell[9] = {L, 0, 0, 0, L, 0, 0, 0, L};
    double *p = malloc(sizeof(double) * 3 * n), *q = malloc(sizeof(double) * n);
    for (int64_t i = 0; i < 3 * n; ++i) p[i] = (uni() * 1.6 - 0.3) * L;
    for (int64_t i = 0; i < n; ++i) q[i] = uni() - 0.5;
    const int64_t cap = n * n;
    int32_t *a = malloc(sizeof(int32_t) * 2 * cap), *b = malloc(sizeof(int32_t) * 2 * cap);
    for (int t = 0; t < 2; ++t) {
        orc_set_cell_type(t);
        int64_t k1 = orc_pairs_n2(n, p, p + n, p + 2 * n, cell, 0.8, a, a + cap, cap);
        int64_t k2 = orc_pairs_cells(n, p, p + n, p + 2 * n, cell, 0.8, b, b + cap, cap);
        expect(k1 == k2 && !memcmp(a, b, sizeof(int32_t) * k1) &&
                   !memcmp(a + cap, b + cap, sizeof(int32_t) * k1),
               "n2 and cell-list pair sets agree");
        double *f = calloc((size_t)3 * n, sizeof(double)), e = 0.0;
        const double s1 = 0.3, e1 = 0.5;
        orc_lj_pairs(k1, a, a + cap, p, p + n, p + 2 * n, cell, 0.8, NULL, 1, &s1, &e1, f,
                     f + n, f + 2 * n, &e);
        orc_coulomb_pairs(k1, a, a + cap, p, p + n, p + 2 * n, cell, 0.8, 3.3, q, f, f + n,
                          f + 2 * n, &e);
        orc_exclusion(k1 < 50 ? k1 : 50, a, a + cap, p, p + n, p + 2 * n, cell, 3.3, q, t, f,
                      f + n, f + 2 * n, &e);
        expect(isfinite(e), "finite pair energy");
        free(f);
    }
    orc_set_cell_type(0);
    for (int mode = 0; mode < 2; ++mode) {
        double *f = calloc((size_t)3 * n, sizeof(double)), e = 0.0;
        orc_pme(n, p, p + n, p + 2 * n, q, cell, 10, 12, 9, 3.1, mode, f, f + n, f + 2 * n, &e);
        expect(isfinite(e), "finite PME energy");
        memset(f, 0, sizeof(double) * 3 * n);
        orc_ewald(60, p, p + n, p + 2 * n, q, cell, 3.1, 1.0, 3, f, f + n, f + 2 * n, &e);
        free(f);
    }
    free(p);
    free(q);
    free(a);
    free(b);
}

int main(void)
{
    check_lists_and_pme();
    check_sim(216, 2.2, DMDX_COULOMB_PME, DMDX_THERMO_BERENDSEN, DMDX_BARO_NONE, 0,
              DMDX_CORRECTED);
    check_sim(216, 2.2, DMDX_COULOMB_PME, DMDX_THERMO_NOSE_HOOVER, DMDX_BARO_BERENDSEN, 0,
              DMDX_FAITHFUL);
    check_sim(125, 2.0, DMDX_COULOMB_CUTOFF, DMDX_THERMO_ANDERSEN, DMDX_BARO_ANDERSEN, 1,
              DMDX_CORRECTED);
    check_sim(64, 1.6, DMDX_COULOMB_NONE, DMDX_THERMO_NONE, DMDX_BARO_NONE, 1, DMDX_FAITHFUL);
    printf("{\"sanitize_check\": \"oracle\", \"ok\": %s}\n", g_fail ? "false" : "true");
    return g_fail;
}
