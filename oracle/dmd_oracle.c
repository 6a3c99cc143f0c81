/*
 * dmd_oracle.c -- FP64 CPU restatement of DMD's per-step force and
 * integration path.  TEST INFRASTRUCTURE ONLY: the checker of the HIP
 * library and the timed CPU baseline ("port"); nothing in the product path
 * links, loads or calls it.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off).  In FAITHFUL mode
 * every expression keeps the reference's operation order so that results
 * are bit-identical to the reference compiled with the same flags.
 *
 * Pinning: the faithful functions are checked bit-for-bit against the
 * reference's own translation units (oracle/_ref, built by
 * oracle/build_ref.sh) where those compile, and every function against the
 * known-answer values of the reference's unit tests, finite differences,
 * closed forms, brute force and a plain Ewald sum (tests/test_oracle*.py).
 */
#include "dmd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ core */

/* Cell type of every cell passed to the functions below (CellType,
 * src/core/cell.h:3-8: 0 orthorhombic, 1 triclinic, 2 hexagonal,
 * 3 rhombic dodecahedron).  Test-side switch, orthorhombic by default. */
static int g_cell_type = 0;

void orc_set_cell_type(int type) { g_cell_type = type; }

/* src/core/cell.cpp:26-38: the orthorhombic image; every other cell type
 * leaves the displacement unchanged (the reference's stub). */
void orc_min_image(const double cell[9], double *dx, double *dy, double *dz)
{
    if (g_cell_type != 0) return;
    double inv_a = 1.0 / cell[0];
    double inv_b = 1.0 / cell[4];
    double inv_c = 1.0 / cell[8];
    *dx = *dx - round(*dx * inv_a) * cell[0];
    *dy = *dy - round(*dy * inv_b) * cell[4];
    *dz = *dz - round(*dz * inv_c) * cell[8];
}

/* src/core/cell.cpp:19-24 */
double orc_cell_volume(const double m[9])
{
    double a1 = m[0], a2 = m[1], a3 = m[2];
    double b1 = m[3], b2 = m[4], b3 = m[5];
    double c1 = m[6], c2 = m[7], c3 = m[8];
    return a1 * (b2 * c3 - b3 * c2) - a2 * (b1 * c3 - b3 * c1) + a3 * (b1 * c2 - b2 * c1);
}

/* ------------------------------------------------------------ pair lists */

/* src/force/lennard_jones.cpp:18-41 (LennardJones::rebuild_pair_list) and
 * the identical loop of src/force/coulomb_direct.cpp:14-36. */
int64_t orc_pairs_n2(int64_t n, const double *x, const double *y, const double *z,
                     const double cell[9], double rc, int32_t *pi, int32_t *pj, int64_t cap)
{
    double rc2 = rc * rc;
    int64_t np = 0;
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t j = i + 1; j < n; ++j) {
            double dx = x[j] - x[i];
            double dy = y[j] - y[i];
            double dz = z[j] - z[i];
            orc_min_image(cell, &dx, &dy, &dz);
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < rc2) {
                if (np < cap) {
                    pi[np] = (int32_t)i;
                    pj[np] = (int32_t)j;
                }
                ++np;
            }
        }
    }
    return np;
}

static int cmp_pair(const void *a, const void *b)
{
    const int64_t *pa = (const int64_t *)a, *pb = (const int64_t *)b;
    return (*pa > *pb) - (*pa < *pb);
}

/* Same set as orc_pairs_n2, found through a cell grid (own algorithm; the
 * decision r2 < rc2 is the reference's expression).  Falls back to the
 * all-pairs loop when a box edge holds fewer than 3 cells. */
int64_t orc_pairs_cells(int64_t n, const double *x, const double *y, const double *z,
                        const double cell[9], double rc, int32_t *pi, int32_t *pj, int64_t cap)
{
    double L[3] = {cell[0], cell[4], cell[8]};
    int nc[3];
    for (int d = 0; d < 3; ++d) {
        nc[d] = (int)floor(L[d] / (rc * (1.0 + 1e-9)));
        if (nc[d] < 3) return orc_pairs_n2(n, x, y, z, cell, rc, pi, pj, cap);
    }
    int64_t ncell = (int64_t)nc[0] * nc[1] * nc[2];
    int32_t *head = (int32_t *)malloc(sizeof(int32_t) * ncell);
    int32_t *next = (int32_t *)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
    int32_t *cid = (int32_t *)malloc(sizeof(int32_t) * 3 * (n > 0 ? n : 1));
    for (int64_t c = 0; c < ncell; ++c) head[c] = -1;
    for (int64_t i = n - 1; i >= 0; --i) {
        const double p[3] = {x[i], y[i], z[i]};
        int c3[3];
        for (int d = 0; d < 3; ++d) {
            double f = p[d] / L[d];
            f -= floor(f);
            int c = (int)(f * nc[d]);
            if (c >= nc[d]) c = nc[d] - 1;
            if (c < 0) c = 0;
            c3[d] = c;
            cid[3 * i + d] = c;
        }
        int64_t c = ((int64_t)c3[2] * nc[1] + c3[1]) * nc[0] + c3[0];
        next[i] = head[c];
        head[c] = (int32_t)i;
    }
    double rc2 = rc * rc;
    int64_t keycap = 1024, nk = 0;
    int64_t *keys = (int64_t *)malloc(sizeof(int64_t) * keycap);
    for (int64_t i = 0; i < n; ++i) {
        for (int oz = -1; oz <= 1; ++oz)
            for (int oy = -1; oy <= 1; ++oy)
                for (int ox = -1; ox <= 1; ++ox) {
                    int cx = (cid[3 * i] + ox + nc[0]) % nc[0];
                    int cy = (cid[3 * i + 1] + oy + nc[1]) % nc[1];
                    int cz = (cid[3 * i + 2] + oz + nc[2]) % nc[2];
                    int64_t c = ((int64_t)cz * nc[1] + cy) * nc[0] + cx;
                    for (int32_t j = head[c]; j >= 0; j = next[j]) {
                        if (j <= i) continue;
                        double dx = x[j] - x[i];
                        double dy = y[j] - y[i];
                        double dz = z[j] - z[i];
                        orc_min_image(cell, &dx, &dy, &dz);
                        double r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 < rc2) {
                            if (nk == keycap) {
                                keycap *= 2;
                                keys = (int64_t *)realloc(keys, sizeof(int64_t) * keycap);
                            }
                            keys[nk++] = (i << 32) | (int64_t)j;
                        }
                    }
                }
    }
    qsort(keys, (size_t)nk, sizeof(int64_t), cmp_pair);
    for (int64_t p = 0; p < nk && p < cap; ++p) {
        pi[p] = (int32_t)(keys[p] >> 32);
        pj[p] = (int32_t)(keys[p] & 0xffffffff);
    }
    free(keys);
    free(head);
    free(next);
    free(cid);
    return nk;
}

int64_t orc_pairs_filter_excl(int64_t np, int32_t *pi, int32_t *pj,
                              int64_t nex, const int32_t *ei, const int32_t *ej)
{
    if (nex == 0) return np;
    int64_t *ex = (int64_t *)malloc(sizeof(int64_t) * nex);
    for (int64_t e = 0; e < nex; ++e) {
        int64_t a = ei[e] < ej[e] ? ei[e] : ej[e];
        int64_t b = ei[e] < ej[e] ? ej[e] : ei[e];
        ex[e] = (a << 32) | b;
    }
    qsort(ex, (size_t)nex, sizeof(int64_t), cmp_pair);
    int64_t w = 0;
    for (int64_t p = 0; p < np; ++p) {
        int64_t key = ((int64_t)pi[p] << 32) | (int64_t)pj[p];
        if (bsearch(&key, ex, (size_t)nex, sizeof(int64_t), cmp_pair)) continue;
        pi[w] = pi[p];
        pj[w] = pj[p];
        ++w;
    }
    free(ex);
    return w;
}

/* ------------------------------------------------------ force components */

/* src/force/lennard_jones.cpp:57-91 (pair loop of LennardJones::compute). */
void orc_lj_pairs(int64_t np, const int32_t *pi, const int32_t *pj,
                  const double *x, const double *y, const double *z, const double cell[9],
                  double rc, const int32_t *types, int32_t n_types,
                  const double *sigma, const double *epsilon,
                  double *fx, double *fy, double *fz, double *energy_out)
{
    double cutoff2 = rc * rc;
    double energy = 0.0;
    for (int64_t p = 0; p < np; ++p) {
        int i = pi[p], j = pj[p];
        double dx = x[j] - x[i];
        double dy = y[j] - y[i];
        double dz = z[j] - z[i];
        orc_min_image(cell, &dx, &dy, &dz);
        double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 >= cutoff2) continue;
        int ti = types ? types[i] : 0;
        int tj = types ? types[j] : 0;
        double eps = epsilon[ti * n_types + tj];
        double sig = sigma[ti * n_types + tj];
        double sig2 = sig * sig;
        double sig6 = sig2 * sig2 * sig2;
        double r6 = r2 * r2 * r2;
        double sr6 = sig6 / r6;
        double f_scalar = 24.0 * eps * (2.0 * sr6 * sr6 - sr6) / r2;
        fx[i] -= f_scalar * dx;
        fy[i] -= f_scalar * dy;
        fz[i] -= f_scalar * dz;
        fx[j] += f_scalar * dx;
        fy[j] += f_scalar * dy;
        fz[j] += f_scalar * dz;
        energy += 4.0 * eps * (sr6 * sr6 - sr6);
    }
    *energy_out += energy;
}

/* src/force/coulomb_direct.cpp:38-86 (CoulombDirect::compute). */
void orc_coulomb_pairs(int64_t np, const int32_t *pi, const int32_t *pj,
                       const double *x, const double *y, const double *z, const double cell[9],
                       double rc, double alpha, const double *q,
                       double *fx, double *fy, double *fz, double *energy_out)
{
    double k_e = ORC_KE;
    double cutoff2 = rc * rc;
    double energy = 0.0;
    for (int64_t p = 0; p < np; ++p) {
        int i = pi[p], j = pj[p];
        double dx = x[j] - x[i];
        double dy = y[j] - y[i];
        double dz = z[j] - z[i];
        orc_min_image(cell, &dx, &dy, &dz);
        double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 >= cutoff2) continue;
        double r = sqrt(r2);
        double qiqj = q[i] * q[j];
        double erfc_ar = erfc(alpha * r);
        double force_scalar = k_e * qiqj * (erfc_ar / (r2 * r) +
                              2.0 * alpha * exp(-alpha * alpha * r2) / (sqrt(ORC_PI) * r2));
        fx[i] -= force_scalar * dx;
        fy[i] -= force_scalar * dy;
        fz[i] -= force_scalar * dz;
        fx[j] += force_scalar * dx;
        fy[j] += force_scalar * dy;
        fz[j] += force_scalar * dz;
        energy += k_e * qiqj * erfc_ar / r;
    }
    *energy_out += energy;
}

/* FAITHFUL: src/force/coulomb_exclusion.cpp:16-56 (subtracts k q q / r but
 * adds the force of +k q q / r).  CORRECTED: removes the reciprocal-space
 * part k q q erf(alpha r)/r of excluded pairs with the consistent force. */
void orc_exclusion(int64_t ne, const int32_t *ei, const int32_t *ej,
                   const double *x, const double *y, const double *z, const double cell[9],
                   double alpha, const double *q, int mode,
                   double *fx, double *fy, double *fz, double *energy_out)
{
    double k_e = ORC_KE;
    double energy = 0.0;
    for (int64_t p = 0; p < ne; ++p) {
        int i = ei[p], j = ej[p];
        double dx = x[j] - x[i];
        double dy = y[j] - y[i];
        double dz = z[j] - z[i];
        orc_min_image(cell, &dx, &dy, &dz);
        double r2 = dx * dx + dy * dy + dz * dz;
        double r = sqrt(r2);
        double qiqj = q[i] * q[j];
        if (mode == DMDX_FAITHFUL) {
            double force_scalar = k_e * qiqj / (r2 * r);
            fx[i] -= force_scalar * dx;
            fy[i] -= force_scalar * dy;
            fz[i] -= force_scalar * dz;
            fx[j] += force_scalar * dx;
            fy[j] += force_scalar * dy;
            fz[j] += force_scalar * dz;
            energy += k_e * qiqj / r;
        } else {
            /* E_pair = -k q q erf(a r)/r ;  F_j = s d with
             * s = -(dE/dr)/r = k q q (2a/sqrt(pi) e^{-a^2 r^2}/r^2 - erf(a r)/r^3) */
            double erf_ar = erf(alpha * r);
            double s = k_e * qiqj * (2.0 * alpha / sqrt(ORC_PI) * exp(-alpha * alpha * r2) / r2 -
                                     erf_ar / (r2 * r));
            fx[i] -= s * dx;
            fy[i] -= s * dy;
            fz[i] -= s * dz;
            fx[j] += s * dx;
            fy[j] += s * dy;
            fz[j] += s * dz;
            energy += k_e * qiqj * erf_ar / r;
        }
    }
    *energy_out -= energy;
}

/* src/force/harmonic_bond.cpp:12-47 (correct as written: same in both modes) */
void orc_bonds(int64_t nb, const int32_t *bi, const int32_t *bj, const double *k, const double *r0,
               const double *x, const double *y, const double *z, const double cell[9],
               double *fx, double *fy, double *fz, double *energy_out)
{
    double energy = 0.0;
    for (int64_t n = 0; n < nb; ++n) {
        int ai = bi[n], aj = bj[n];
        double dx = x[aj] - x[ai];
        double dy = y[aj] - y[ai];
        double dz = z[aj] - z[ai];
        orc_min_image(cell, &dx, &dy, &dz);
        double r2 = dx * dx + dy * dy + dz * dz;
        double r = sqrt(r2);
        double dr = r - r0[n];
        double f_scalar = -k[n] * dr / r;
        double gx = f_scalar * dx, gy = f_scalar * dy, gz = f_scalar * dz;
        fx[ai] -= gx; fy[ai] -= gy; fz[ai] -= gz;
        fx[aj] += gx; fy[aj] += gy; fz[aj] += gz;
        energy += 0.5 * k[n] * dr * dr;
    }
    *energy_out += energy;
}

/* src/force/harmonic_angle.cpp:10-63.  FAITHFUL keeps the reference's
 * prefactor -k dtheta / sin(theta), which points the force up the gradient;
 * CORRECTED uses +k dtheta / sin(theta) (F = -grad E). */
void orc_angles(int64_t na, const int32_t *ai_, const int32_t *aj_, const int32_t *ak_,
                const double *kt, const double *t0,
                const double *x, const double *y, const double *z, const double cell[9], int mode,
                double *fx, double *fy, double *fz, double *energy_out)
{
    double energy = 0.0;
    for (int64_t n = 0; n < na; ++n) {
        int ai = ai_[n], aj = aj_[n], ak = ak_[n];
        double dx_ij = x[ai] - x[aj];
        double dy_ij = y[ai] - y[aj];
        double dz_ij = z[ai] - z[aj];
        orc_min_image(cell, &dx_ij, &dy_ij, &dz_ij);
        double dx_kj = x[ak] - x[aj];
        double dy_kj = y[ak] - y[aj];
        double dz_kj = z[ak] - z[aj];
        orc_min_image(cell, &dx_kj, &dy_kj, &dz_kj);
        double r_ij2 = dx_ij * dx_ij + dy_ij * dy_ij + dz_ij * dz_ij;
        double r_kj2 = dx_kj * dx_kj + dy_kj * dy_kj + dz_kj * dz_kj;
        double r_ij = sqrt(r_ij2);
        double r_kj = sqrt(r_kj2);
        double dot = dx_ij * dx_kj + dy_ij * dy_kj + dz_ij * dz_kj;
        double cos_theta = dot / (r_ij * r_kj);
        cos_theta = fmax(-1.0, fmin(1.0, cos_theta));
        double theta = acos(cos_theta);
        double dtheta = theta - t0[n];
        double f_scalar;
        if (mode == DMDX_FAITHFUL)
            f_scalar = -kt[n] * dtheta / sqrt(1.0 - cos_theta * cos_theta);
        else
            f_scalar = kt[n] * dtheta / sqrt(1.0 - cos_theta * cos_theta);
        double fx_ij = f_scalar * (dx_kj / (r_ij * r_kj) - cos_theta * dx_ij / r_ij2);
        double fy_ij = f_scalar * (dy_kj / (r_ij * r_kj) - cos_theta * dy_ij / r_ij2);
        double fz_ij = f_scalar * (dz_kj / (r_ij * r_kj) - cos_theta * dz_ij / r_ij2);
        double fx_kj = f_scalar * (dx_ij / (r_ij * r_kj) - cos_theta * dx_kj / r_kj2);
        double fy_kj = f_scalar * (dy_ij / (r_ij * r_kj) - cos_theta * dy_kj / r_kj2);
        double fz_kj = f_scalar * (dz_ij / (r_ij * r_kj) - cos_theta * dz_kj / r_kj2);
        fx[ai] += fx_ij; fy[ai] += fy_ij; fz[ai] += fz_ij;
        fx[aj] -= fx_ij + fx_kj; fy[aj] -= fy_ij + fy_kj; fz[aj] -= fz_ij + fz_kj;
        fx[ak] += fx_kj; fy[ak] += fy_kj; fz[ak] += fz_kj;
        energy += 0.5 * kt[n] * dtheta * dtheta;
    }
    *energy_out += energy;
}

/* Shared torsion geometry of src/force/periodic_dihedral.cpp:21-60 and
 * src/force/harmonic_improper.cpp:21-59.  Returns 0 when the reference skips
 * the term (|n1|^2, |n2|^2 or |bc|^2 below 1e-30). */
typedef struct {
    double ba[3], bc[3], dc[3], n1[3], n2[3];
    double n1_sq, n2_sq, bc_sq, phi;
} torsion_geom;

static int torsion(const double *x, const double *y, const double *z, const double cell[9],
                   int ai, int aj, int ak, int al, torsion_geom *g)
{
    double ba_x = x[ai] - x[aj], ba_y = y[ai] - y[aj], ba_z = z[ai] - z[aj];
    orc_min_image(cell, &ba_x, &ba_y, &ba_z);
    double bc_x = x[ak] - x[aj], bc_y = y[ak] - y[aj], bc_z = z[ak] - z[aj];
    orc_min_image(cell, &bc_x, &bc_y, &bc_z);
    double dc_x = x[ak] - x[al], dc_y = y[ak] - y[al], dc_z = z[ak] - z[al];
    orc_min_image(cell, &dc_x, &dc_y, &dc_z);
    double n1_x = ba_y * bc_z - ba_z * bc_y;
    double n1_y = ba_z * bc_x - ba_x * bc_z;
    double n1_z = ba_x * bc_y - ba_y * bc_x;
    double n2_x = bc_y * dc_z - bc_z * dc_y;
    double n2_y = bc_z * dc_x - bc_x * dc_z;
    double n2_z = bc_x * dc_y - bc_y * dc_x;
    double n1_sq = n1_x * n1_x + n1_y * n1_y + n1_z * n1_z;
    double n2_sq = n2_x * n2_x + n2_y * n2_y + n2_z * n2_z;
    double bc_sq = bc_x * bc_x + bc_y * bc_y + bc_z * bc_z;
    g->ba[0] = ba_x; g->ba[1] = ba_y; g->ba[2] = ba_z;
    g->bc[0] = bc_x; g->bc[1] = bc_y; g->bc[2] = bc_z;
    g->dc[0] = dc_x; g->dc[1] = dc_y; g->dc[2] = dc_z;
    g->n1[0] = n1_x; g->n1[1] = n1_y; g->n1[2] = n1_z;
    g->n2[0] = n2_x; g->n2[1] = n2_y; g->n2[2] = n2_z;
    g->n1_sq = n1_sq; g->n2_sq = n2_sq; g->bc_sq = bc_sq;
    if (n1_sq < 1e-30 || n2_sq < 1e-30 || bc_sq < 1e-30) return 0;
    double n1_norm = sqrt(n1_sq);
    double n2_norm = sqrt(n2_sq);
    double bc_norm = sqrt(bc_sq);
    double dot = n1_x * n2_x + n1_y * n2_y + n1_z * n2_z;
    double cos_phi = dot / (n1_norm * n2_norm);
    cos_phi = fmax(-1.0, fmin(1.0, cos_phi));
    double triple = bc_x * (n1_y * n2_z - n1_z * n2_y)
                  + bc_y * (n1_z * n2_x - n1_x * n2_z)
                  + bc_z * (n1_x * n2_y - n1_y * n2_x);
    double sin_phi = triple / (bc_norm * n1_norm * n2_norm);
    sin_phi = fmax(-1.0, fmin(1.0, sin_phi));
    g->phi = atan2(sin_phi, cos_phi);
    return 1;
}

/* FAITHFUL force of src/force/periodic_dihedral.cpp:86-134 (identical in
 * src/force/harmonic_improper.cpp:81-127): F_m += pref * g_m with in-plane
 * g vectors. */
static void torsion_force_faithful(const torsion_geom *t, double pref, int ai, int aj, int ak,
                                   int al, double *fx, double *fy, double *fz)
{
    double bc_x = t->bc[0], bc_y = t->bc[1], bc_z = t->bc[2];
    double ba_x = t->ba[0], ba_y = t->ba[1], ba_z = t->ba[2];
    double dc_x = t->dc[0], dc_y = t->dc[1], dc_z = t->dc[2];
    double n1_x = t->n1[0], n1_y = t->n1[1], n1_z = t->n1[2];
    double n2_x = t->n2[0], n2_y = t->n2[1], n2_z = t->n2[2];
    double n1_sq = t->n1_sq, n2_sq = t->n2_sq;
    double gi_x = (bc_y * n1_z - bc_z * n1_y) / n1_sq;
    double gi_y = (bc_z * n1_x - bc_x * n1_z) / n1_sq;
    double gi_z = (bc_x * n1_y - bc_y * n1_x) / n1_sq;
    double gl_x = -(bc_y * n2_z - bc_z * n2_y) / n2_sq;
    double gl_y = -(bc_z * n2_x - bc_x * n2_z) / n2_sq;
    double gl_z = -(bc_x * n2_y - bc_y * n2_x) / n2_sq;
    double n1xba_x = n1_y * ba_z - n1_z * ba_y;
    double n1xba_y = n1_z * ba_x - n1_x * ba_z;
    double n1xba_z = n1_x * ba_y - n1_y * ba_x;
    double dcxn2_x = dc_y * n2_z - dc_z * n2_y;
    double dcxn2_y = dc_z * n2_x - dc_x * n2_z;
    double dcxn2_z = dc_x * n2_y - dc_y * n2_x;
    double gj_x = n1xba_x / n1_sq + dcxn2_x / n2_sq - gi_x;
    double gj_y = n1xba_y / n1_sq + dcxn2_y / n2_sq - gi_y;
    double gj_z = n1xba_z / n1_sq + dcxn2_z / n2_sq - gi_z;
    double bxn2_x = (bc_y * n2_z - bc_z * n2_y) / n2_sq;
    double bxn2_y = (bc_z * n2_x - bc_x * n2_z) / n2_sq;
    double bxn2_z = (bc_x * n2_y - bc_y * n2_x) / n2_sq;
    double gk_x = bxn2_x - n1xba_x / n1_sq - dcxn2_x / n2_sq;
    double gk_y = bxn2_y - n1xba_y / n1_sq - dcxn2_y / n2_sq;
    double gk_z = bxn2_z - n1xba_z / n1_sq - dcxn2_z / n2_sq;
    fx[ai] += pref * gi_x; fy[ai] += pref * gi_y; fz[ai] += pref * gi_z;
    fx[aj] += pref * gj_x; fy[aj] += pref * gj_y; fz[aj] += pref * gj_z;
    fx[ak] += pref * gk_x; fy[ak] += pref * gk_y; fz[ak] += pref * gk_z;
    fx[al] += pref * gl_x; fy[al] += pref * gl_y; fz[al] += pref * gl_z;
}

/* CORRECTED torsion force, F = -(dE/dphi) grad(phi), with the gradient of
 * Bekker (1996) / Blondel & Karplus (1996) written for r_ij = ba,
 * r_kj = bc, r_kl = dc, m = n1, n = n2 (the same phi as the reference). */
static void torsion_force_corrected(const torsion_geom *t, double dEdphi, int ai, int aj, int ak,
                                    int al, double *fx, double *fy, double *fz)
{
    double nrkj = sqrt(t->bc_sq);
    double a = -dEdphi * nrkj / t->n1_sq;
    double b = dEdphi * nrkj / t->n2_sq;
    double f_i[3], f_l[3], f_j[3], f_k[3];
    for (int d = 0; d < 3; ++d) {
        f_i[d] = a * t->n1[d];
        f_l[d] = b * t->n2[d];
    }
    double p = (t->ba[0] * t->bc[0] + t->ba[1] * t->bc[1] + t->ba[2] * t->bc[2]) / t->bc_sq;
    double q = (t->dc[0] * t->bc[0] + t->dc[1] * t->bc[1] + t->dc[2] * t->bc[2]) / t->bc_sq;
    for (int d = 0; d < 3; ++d) {
        double s = p * f_i[d] - q * f_l[d];
        f_j[d] = f_i[d] - s;
        f_k[d] = f_l[d] + s;
    }
    fx[ai] += f_i[0]; fy[ai] += f_i[1]; fz[ai] += f_i[2];
    fx[aj] -= f_j[0]; fy[aj] -= f_j[1]; fz[aj] -= f_j[2];
    fx[ak] -= f_k[0]; fy[ak] -= f_k[1]; fz[ak] -= f_k[2];
    fx[al] += f_l[0]; fy[al] += f_l[1]; fz[al] += f_l[2];
}

/* src/force/periodic_dihedral.cpp:10-137 */
void orc_dihedrals(int64_t nd, const int32_t *di, const int32_t *dj, const int32_t *dk,
                   const int32_t *dl, const int32_t *per, const double *kp, const double *p0,
                   const double *x, const double *y, const double *z, const double cell[9],
                   int mode, double *fx, double *fy, double *fz, double *energy_out)
{
    double energy = 0.0;
    for (int64_t n = 0; n < nd; ++n) {
        torsion_geom t;
        if (!torsion(x, y, z, cell, di[n], dj[n], dk[n], dl[n], &t)) continue;
        int p = per[n];
        double k = kp[n];
        double phi0 = p0[n];
        double cos_np = cos(p * t.phi - phi0);
        energy += k * (1.0 + cos_np);
        double pref = k * p * sin(p * t.phi - phi0);
        if (fabs(pref) < 1e-30) continue;
        if (mode == DMDX_FAITHFUL)
            torsion_force_faithful(&t, pref, di[n], dj[n], dk[n], dl[n], fx, fy, fz);
        else
            torsion_force_corrected(&t, -pref, di[n], dj[n], dk[n], dl[n], fx, fy, fz);
    }
    *energy_out += energy;
}

/* src/force/harmonic_improper.cpp:10-130 */
void orc_impropers(int64_t nd, const int32_t *di, const int32_t *dj, const int32_t *dk,
                   const int32_t *dl, const double *kp, const double *p0,
                   const double *x, const double *y, const double *z, const double cell[9],
                   int mode, double *fx, double *fy, double *fz, double *energy_out)
{
    double energy = 0.0;
    for (int64_t n = 0; n < nd; ++n) {
        torsion_geom t;
        if (!torsion(x, y, z, cell, di[n], dj[n], dk[n], dl[n], &t)) continue;
        double dchi = t.phi - p0[n];
        double k = kp[n];
        energy += 0.5 * k * dchi * dchi;
        double pref = -k * dchi;
        if (fabs(pref) < 1e-30) continue;
        if (mode == DMDX_FAITHFUL)
            torsion_force_faithful(&t, pref, di[n], dj[n], dk[n], dl[n], fx, fy, fz);
        else
            torsion_force_corrected(&t, -pref, di[n], dj[n], dk[n], dl[n], fx, fy, fz);
    }
    *energy_out += energy;
}

/* ------------------------------------------------------------------ SPME */

/* src/force/coulomb_pme.cpp:50-58 */
double orc_theta4(double u)
{
    double a = fabs(u);
    if (a >= 2.0) return 0.0;
    if (a >= 1.0) {
        double t = 2.0 - a;
        return t * t * t / 6.0;
    }
    return (4.0 - 6.0 * a * a + 3.0 * a * a * a) / 6.0;
}

/* src/force/coulomb_pme.cpp:60-69 */
double orc_dtheta4(double u)
{
    double a = fabs(u);
    if (a >= 2.0) return 0.0;
    double s = (u >= 0.0) ? 1.0 : -1.0;
    if (a >= 1.0) {
        double t = 2.0 - a;
        return -s * t * t / 2.0;
    }
    return s * (9.0 * a * a - 12.0 * a) / 6.0;
}

/* src/force/coulomb_pme.cpp:71-75 */
double orc_pme_self_energy(double alpha, int64_t n, const double *q, double k_e)
{
    double sum_q2 = 0.0;
    for (int64_t i = 0; i < n; ++i) sum_q2 += q[i] * q[i];
    return -k_e * alpha / sqrt(ORC_PI) * sum_q2;
}

/* Recursive mixed-radix decimation-in-time DFT of one line (own code; it
 * stands in for FFTW's unnormalised transform at coulomb_pme.cpp:154,198). */
static void dft_rec(int n, const double *ir, const double *ii, int stride,
                    double *or_, double *oi, int sign, double *tr, double *ti)
{
    if (n == 1) {
        or_[0] = ir[0];
        oi[0] = ii[0];
        return;
    }
    int p = 2;
    while (n % p) ++p;
    int m = n / p;
    for (int r = 0; r < p; ++r)
        dft_rec(m, ir + r * stride, ii + r * stride, stride * p, or_ + r * m, oi + r * m, sign,
                tr, ti);
    for (int k = 0; k < m; ++k)
        for (int s = 0; s < p; ++s) {
            double accr = 0.0, acci = 0.0;
            int kk = k + s * m;
            for (int r = 0; r < p; ++r) {
                long e = ((long)r * kk) % n;
                double ang = sign * 2.0 * ORC_PI * (double)e / (double)n;
                double c = cos(ang), sn = sin(ang);
                double xr = or_[r * m + k], xi = oi[r * m + k];
                accr += xr * c - xi * sn;
                acci += xr * sn + xi * c;
            }
            tr[kk] = accr;
            ti[kk] = acci;
        }
    memcpy(or_, tr, sizeof(double) * n);
    memcpy(oi, ti, sizeof(double) * n);
}

static void dft_line(int n, double *re, double *im, long stride, int sign)
{
    double *buf = (double *)malloc(sizeof(double) * 6 * n);
    double *ir = buf, *ii = buf + n, *or_ = buf + 2 * n, *oi = buf + 3 * n;
    double *tr = buf + 4 * n, *ti = buf + 5 * n;
    for (int k = 0; k < n; ++k) {
        ir[k] = re[k * stride];
        ii[k] = im[k * stride];
    }
    dft_rec(n, ir, ii, 1, or_, oi, sign, tr, ti);
    for (int k = 0; k < n; ++k) {
        re[k * stride] = or_[k];
        im[k * stride] = oi[k];
    }
    free(buf);
}

void orc_fft3d(int nx, int ny, int nz, double *re, double *im, int sign)
{
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y) {
            long o = ((long)z * ny + y) * nx;
            dft_line(nx, re + o, im + o, 1, sign);
        }
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x) {
            long o = (long)z * ny * nx + x;
            dft_line(ny, re + o, im + o, nx, sign);
        }
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            long o = (long)y * nx + x;
            dft_line(nz, re + o, im + o, (long)nx * ny, sign);
        }
}

/* FAITHFUL B factor of src/force/coulomb_pme.cpp:159-163 (multiplies by
 * (sin t / t)^(2p)); CORRECTED uses the exact Euler exponential-spline
 * modulus of the centred order-4 spline (Essmann et al. 1995, eq. 4.4):
 * |b(m)|^2 = 1 / (2/3 + cos(2 pi m / K) / 3)^2. */
static double bmod(int m, int K, int mode)
{
    if (mode == DMDX_FAITHFUL) {
        if (m == 0) return 1.0;
        double t = ORC_PI * abs(m) / K;
        return pow(sin(t) / t, 2 * 4);
    }
    double d = 2.0 / 3.0 + cos(2.0 * ORC_PI * m / K) / 3.0;
    return 1.0 / (d * d);
}

/* src/force/coulomb_pme.cpp:79-136 (spread), :143-201 (reciprocal energy and
 * potential), :205-283 (force interpolation), :287-318 (compute).  The grid
 * is stored unpadded with x fastest, consistently for spread, FFT and
 * gather (the reference mixes FFTW's padded layout with unpadded indexing,
 * which cannot be reproduced; see the quirk register).  FAITHFUL keeps the
 * multiplied B factor and the 1/(nx ny nz) scaling of the force grid;
 * CORRECTED divides by |b|^2, leaves the grid unscaled and adds the
 * neutralising-background energy -pi k_e Q^2 / (2 V alpha^2). */
void orc_pme(int64_t n, const double *x, const double *y, const double *z, const double *q,
             const double cell[9], int nx, int ny, int nz, double alpha, int mode,
             double *fx, double *fy, double *fz, double *energy_out)
{
    double k_e = ORC_KE;
    double Lx = cell[0], Ly = cell[4], Lz = cell[8];
    long K3 = (long)nx * ny * nz;
    double *gr = (double *)calloc(K3, sizeof(double));
    double *gi = (double *)calloc(K3, sizeof(double));
    /* spread */
    for (int64_t i = 0; i < n; ++i) {
        double qi = q[i];
        if (qi == 0.0) continue;
        double frac_x = x[i] / Lx, frac_y = y[i] / Ly, frac_z = z[i] / Lz;
        frac_x -= floor(frac_x);
        frac_y -= floor(frac_y);
        frac_z -= floor(frac_z);
        double ux = frac_x * nx, uy = frac_y * ny, uz = frac_z * nz;
        int m0x = (int)floor(ux) - 1, m0y = (int)floor(uy) - 1, m0z = (int)floor(uz) - 1;
        double wx[4], wy[4], wz[4];
        for (int d = 0; d < 4; ++d) {
            wx[d] = orc_theta4(ux - (m0x + d));
            wy[d] = orc_theta4(uy - (m0y + d));
            wz[d] = orc_theta4(uz - (m0z + d));
        }
        for (int iz = 0; iz < 4; ++iz) {
            int gz = (m0z + iz) % nz;
            if (gz < 0) gz += nz;
            double w_z = wz[iz];
            for (int iy = 0; iy < 4; ++iy) {
                int gy = (m0y + iy) % ny;
                if (gy < 0) gy += ny;
                double w_yz = wy[iy] * w_z;
                for (int ix = 0; ix < 4; ++ix) {
                    int gx = (m0x + ix) % nx;
                    if (gx < 0) gx += nx;
                    gr[((long)gz * ny + gy) * nx + gx] += qi * wx[ix] * w_yz;
                }
            }
        }
    }
    /* reciprocal energy and potential */
    orc_fft3d(nx, ny, nz, gr, gi, -1);
    double V = orc_cell_volume(cell);
    double e_rec = 0.0;
    for (int kz = 0; kz < nz; ++kz) {
        int mz = (kz <= nz / 2) ? kz : kz - nz;
        for (int ky = 0; ky < ny; ++ky) {
            int my = (ky <= ny / 2) ? ky : ky - ny;
            for (int kx = 0; kx < nx; ++kx) {
                int mx = (kx <= nx / 2) ? kx : kx - nx;
                long idx = ((long)kz * ny + ky) * nx + kx;
                if (mx == 0 && my == 0 && mz == 0) {
                    gr[idx] = 0.0;
                    gi[idx] = 0.0;
                    continue;
                }
                double msq = ((double)mx * mx) / (Lx * Lx) + ((double)my * my) / (Ly * Ly) +
                             ((double)mz * mz) / (Lz * Lz);
                double k2 = 4.0 * ORC_PI * ORC_PI * msq;
                double B = bmod(mx, nx, mode) * bmod(my, ny, mode) * bmod(mz, nz, mode);
                double green = 4.0 * ORC_PI * k_e / V * exp(-k2 / (4.0 * alpha * alpha)) * B / k2;
                double re = gr[idx], im = gi[idx];
                e_rec += 0.5 * green * (re * re + im * im);
                gr[idx] = re * green;
                gi[idx] = im * green;
            }
        }
    }
    orc_fft3d(nx, ny, nz, gr, gi, +1);
    if (mode == DMDX_FAITHFUL) {
        double norm = 1.0 / ((double)nx * ny * nz);
        for (long i = 0; i < K3; ++i) gr[i] *= norm;
    }
    /* gather */
    double fx_factor = (double)nx / Lx, fy_factor = (double)ny / Ly, fz_factor = (double)nz / Lz;
    double qtot = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double qi = q[i];
        qtot += qi;
        if (qi == 0.0) continue;
        double frac_x = x[i] / Lx, frac_y = y[i] / Ly, frac_z = z[i] / Lz;
        frac_x -= floor(frac_x);
        frac_y -= floor(frac_y);
        frac_z -= floor(frac_z);
        double ux = frac_x * nx, uy = frac_y * ny, uz = frac_z * nz;
        int m0x = (int)floor(ux) - 1, m0y = (int)floor(uy) - 1, m0z = (int)floor(uz) - 1;
        double wx[4], wy[4], wz[4], dwx[4], dwy[4], dwz[4];
        for (int d = 0; d < 4; ++d) {
            wx[d] = orc_theta4(ux - (m0x + d));
            wy[d] = orc_theta4(uy - (m0y + d));
            wz[d] = orc_theta4(uz - (m0z + d));
            dwx[d] = orc_dtheta4(ux - (m0x + d));
            dwy[d] = orc_dtheta4(uy - (m0y + d));
            dwz[d] = orc_dtheta4(uz - (m0z + d));
        }
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int iz = 0; iz < 4; ++iz) {
            int gz = (m0z + iz) % nz;
            if (gz < 0) gz += nz;
            double w_z = wz[iz], dw_z = dwz[iz];
            for (int iy = 0; iy < 4; ++iy) {
                int gy = (m0y + iy) % ny;
                if (gy < 0) gy += ny;
                double w_y = wy[iy], dw_y = dwy[iy];
                long row = ((long)gz * ny + gy) * nx;
                for (int ix = 0; ix < 4; ++ix) {
                    int gx = (m0x + ix) % nx;
                    if (gx < 0) gx += nx;
                    double phi = gr[row + gx];
                    sx += phi * dwx[ix] * w_y * w_z;
                    sy += phi * wx[ix] * dw_y * w_z;
                    sz += phi * wx[ix] * w_y * dw_z;
                }
            }
        }
        fx[i] -= qi * fx_factor * sx;
        fy[i] -= qi * fy_factor * sy;
        fz[i] -= qi * fz_factor * sz;
    }
    double energy = e_rec + orc_pme_self_energy(alpha, n, q, k_e);
    if (mode != DMDX_FAITHFUL)
        energy += -ORC_PI * k_e * qtot * qtot / (2.0 * V * alpha * alpha);
    *energy_out += energy;
    free(gr);
    free(gi);
}

/* Plain Ewald sum, corrected physics (own code, for pinning orc_pme). */
void orc_ewald(int64_t n, const double *x, const double *y, const double *z, const double *q,
               const double cell[9], double alpha, double rc_real, int kmax,
               double *fx, double *fy, double *fz, double *energy_out)
{
    double k_e = ORC_KE;
    double L[3] = {cell[0], cell[4], cell[8]};
    double V = L[0] * L[1] * L[2];
    double e = 0.0;
    int nim[3];
    for (int d = 0; d < 3; ++d) nim[d] = (int)ceil(rc_real / L[d]);
    double rc2 = rc_real * rc_real;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = i; j < n; ++j)
            for (int a = -nim[0]; a <= nim[0]; ++a)
                for (int b = -nim[1]; b <= nim[1]; ++b)
                    for (int c = -nim[2]; c <= nim[2]; ++c) {
                        if (i == j && a == 0 && b == 0 && c == 0) continue;
                        double dx = x[j] - x[i] + a * L[0];
                        double dy = y[j] - y[i] + b * L[1];
                        double dz = z[j] - z[i] + c * L[2];
                        double r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 >= rc2) continue;
                        double r = sqrt(r2);
                        double qq = q[i] * q[j];
                        double scale = (i == j) ? 0.5 : 1.0;
                        e += scale * k_e * qq * erfc(alpha * r) / r;
                        if (i != j) {
                            double s = k_e * qq * (erfc(alpha * r) / (r2 * r) +
                                       2.0 * alpha / sqrt(ORC_PI) * exp(-alpha * alpha * r2) / r2);
                            fx[i] -= s * dx; fy[i] -= s * dy; fz[i] -= s * dz;
                            fx[j] += s * dx; fy[j] += s * dy; fz[j] += s * dz;
                        }
                    }
    double qtot = 0.0;
    for (int64_t i = 0; i < n; ++i) qtot += q[i];
    for (int mx = -kmax; mx <= kmax; ++mx)
        for (int my = -kmax; my <= kmax; ++my)
            for (int mz = -kmax; mz <= kmax; ++mz) {
                if (!mx && !my && !mz) continue;
                double m[3] = {mx / L[0], my / L[1], mz / L[2]};
                double m2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
                double f = exp(-ORC_PI * ORC_PI * m2 / (alpha * alpha)) / m2;
                double sr = 0.0, si = 0.0;
                for (int64_t i = 0; i < n; ++i) {
                    double th = 2.0 * ORC_PI * (m[0] * x[i] + m[1] * y[i] + m[2] * z[i]);
                    sr += q[i] * cos(th);
                    si += q[i] * sin(th);
                }
                e += k_e / (2.0 * ORC_PI * V) * f * (sr * sr + si * si);
                for (int64_t i = 0; i < n; ++i) {
                    double th = 2.0 * ORC_PI * (m[0] * x[i] + m[1] * y[i] + m[2] * z[i]);
                    double g = 2.0 * k_e * q[i] / V * f * (sin(th) * sr - cos(th) * si);
                    fx[i] += g * m[0];
                    fy[i] += g * m[1];
                    fz[i] += g * m[2];
                }
            }
    double sq2 = 0.0;
    for (int64_t i = 0; i < n; ++i) sq2 += q[i] * q[i];
    e += -k_e * alpha / sqrt(ORC_PI) * sq2;
    e += -ORC_PI * k_e * qtot * qtot / (2.0 * V * alpha * alpha);
    *energy_out += e;
}

/* ------------------------------------------------- integration, coupling */

/* src/integrate/integrator.cpp:4-11 */
void orc_half_kick(int64_t n, double *vx, double *vy, double *vz,
                   const double *fx, const double *fy, const double *fz,
                   const double *m, double dt)
{
    double half_dt = 0.5 * dt;
    for (int64_t i = 0; i < n; ++i) {
        double inv_mass = 1.0 / m[i];
        vx[i] += fx[i] * inv_mass * half_dt;
        vy[i] += fy[i] * inv_mass * half_dt;
        vz[i] += fz[i] * inv_mass * half_dt;
    }
}

/* src/integrate/integrator.cpp:14-20 */
void orc_advance(int64_t n, double *x, double *y, double *z,
                 const double *vx, const double *vy, const double *vz, double dt)
{
    for (int64_t i = 0; i < n; ++i) {
        x[i] += vx[i] * dt;
        y[i] += vy[i] * dt;
        z[i] += vz[i] * dt;
    }
}

/* src/thermostat/thermostat.cpp:6-15 */
double orc_kinetic_energy(int64_t n, const double *m, const double *vx, const double *vy,
                          const double *vz)
{
    double ekin = 0.0;
    for (int64_t i = 0; i < n; ++i)
        ekin += 0.5 * m[i] * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
    return ekin;
}

/* src/thermostat/thermostat.cpp:17-21 */
double orc_temperature(double ekin, int64_t n)
{
    if (n < 2) return 0.0;
    double dof = 3.0 * (double)n - 3.0;
    return 2.0 * ekin / (dof * ORC_KB);
}

/* src/thermostat/berendsen.cpp:13-28 */
void orc_berendsen(int64_t n, const double *m, double *vx, double *vy, double *vz,
                   double dt, double T0, double tau)
{
    if (tau <= 0.0) return;
    double ekin = orc_kinetic_energy(n, m, vx, vy, vz);
    if (ekin <= 0.0) return;
    double T_inst = orc_temperature(ekin, n);
    double lambda = sqrt(1.0 + dt / tau * (T0 / T_inst - 1.0));
    if (lambda <= 0.0) return;
    for (int64_t i = 0; i < n; ++i) {
        vx[i] *= lambda;
        vy[i] *= lambda;
        vz[i] *= lambda;
    }
}

/* src/thermostat/nose_hoover.cpp:6-11 (Q) and :18-32 (apply) */
void orc_nose_hoover(int64_t n, const double *m, double *vx, double *vy, double *vz,
                     double dt, double T0, double tau, double *zeta)
{
    double dof = 3.0 * (double)n - 3.0;
    if (dof < 1.0) dof = 1.0;
    double Q = dof * ORC_KB * T0 * tau * tau;
    double ekin = orc_kinetic_energy(n, m, vx, vy, vz);
    double T_inst = orc_temperature(ekin, n);
    *zeta += 0.5 * dt * ORC_KB * (T_inst - T0) / Q;
    double scale = exp(-*zeta * 0.5 * dt);
    for (int64_t i = 0; i < n; ++i) {
        vx[i] *= scale;
        vy[i] *= scale;
        vz[i] *= scale;
    }
}

/* src/barostat/barostat.cpp:6-14 */
double orc_virial(int64_t n, const double *x, const double *y, const double *z,
                  const double *fx, const double *fy, const double *fz)
{
    double w = 0.0;
    for (int64_t i = 0; i < n; ++i) w += x[i] * fx[i] + y[i] * fy[i] + z[i] * fz[i];
    return w;
}

/* src/barostat/barostat.cpp:16-19 */
double orc_pressure(double ekin, double w, double volume)
{
    if (volume <= 0.0) return 0.0;
    return (2.0 * ekin + w) / (3.0 * volume);
}

/* src/barostat/berendsen_barostat.cpp:14-66 (bar / internal-unit mix kept) */
void orc_berendsen_barostat(int64_t n, const double *m, const double *vx, const double *vy,
                            const double *vz, double *x, double *y, double *z,
                            const double *fx, const double *fy, const double *fz,
                            double cell[9], double dt, double P0, double tau, double beta)
{
    if (tau <= 0.0) return;
    double volume = orc_cell_volume(cell);
    if (volume <= 0.0) return;
    double ekin = orc_kinetic_energy(n, m, vx, vy, vz);
    double P_inst = orc_pressure(ekin, orc_virial(n, x, y, z, fx, fy, fz), volume);
    double beta_internal = beta / 0.0602214086;
    double mu = 1.0 - (dt / tau) * beta_internal * (P0 - P_inst);
    if (mu <= 0.0) mu = 0.0;
    double scale = cbrt(mu);
    for (int64_t i = 0; i < n; ++i) {
        x[i] *= scale;
        y[i] *= scale;
        z[i] *= scale;
    }
    cell[0] *= scale;
    cell[4] *= scale;
    cell[8] *= scale;
}

/* src/barostat/andersen_barostat.cpp:8-45 (W uses P0 where T0 belongs) */
void orc_andersen_barostat(int64_t n, const double *m, const double *vx, const double *vy,
                           const double *vz, double *x, double *y, double *z,
                           const double *fx, const double *fy, const double *fz,
                           double cell[9], double dt, double P0, double tau, double *eps_dot)
{
    double dof = 3.0 * (double)n - 3.0;
    if (dof < 1.0) dof = 1.0;
    double W = dof * ORC_KB * P0 * tau * tau;
    double volume = orc_cell_volume(cell);
    if (volume <= 0.0) return;
    double ekin = orc_kinetic_energy(n, m, vx, vy, vz);
    double P_inst = orc_pressure(ekin, orc_virial(n, x, y, z, fx, fy, fz), volume);
    double forcing = 3.0 * volume * (P_inst - P0) / W;
    *eps_dot = *eps_dot + 0.5 * dt * forcing;
    double scale = exp(*eps_dot * dt);
    for (int64_t i = 0; i < n; ++i) {
        x[i] *= scale;
        y[i] *= scale;
        z[i] *= scale;
    }
    cell[0] *= scale;
    cell[4] *= scale;
    cell[8] *= scale;
    *eps_dot = *eps_dot + 0.5 * dt * forcing;
}

/* ------------------------------------------------------------ simulation */

/* ------------------------------------------------------------ constraints */

/* src/integrate/constraints.cpp:9-49: Gauss-Seidel sweeps over all
 * constraints in order, at most 100, until a sweep's largest |r^2 - d^2| is
 * below the tolerance; positions only. */
void orc_shake(int64_t nc, const int32_t *ci, const int32_t *cj, const double *dist, double tol,
               const double cell[9], const double *m, double *x, double *y, double *z)
{
    for (int iter = 0; iter < 100; ++iter) {
        double max_error = 0.0;
        for (int64_t c = 0; c < nc; ++c) {
            int ai = ci[c], aj = cj[c];
            double dx = x[aj] - x[ai];
            double dy = y[aj] - y[ai];
            double dz = z[aj] - z[ai];
            orc_min_image(cell, &dx, &dy, &dz);
            double r2 = dx * dx + dy * dy + dz * dz;
            double d2 = dist[c] * dist[c];
            double error = r2 - d2;
            if (fabs(error) > max_error) max_error = fabs(error);
            if (fabs(error) > tol) {
                double imi = 1.0 / m[ai];
                double imj = 1.0 / m[aj];
                double ims = imi + imj;
                double lambda = error / (2.0 * d2 * ims);
                x[ai] += lambda * dx * imi;
                y[ai] += lambda * dy * imi;
                z[ai] += lambda * dz * imi;
                x[aj] -= lambda * dx * imj;
                y[aj] -= lambda * dy * imj;
                z[aj] -= lambda * dz * imj;
            }
        }
        if (max_error < tol) break;
    }
}

/* src/integrate/constraints.cpp:51-91: the same sweeps on |r . v_rel|,
 * velocities only. */
void orc_rattle(int64_t nc, const int32_t *ci, const int32_t *cj, double tol,
                const double cell[9], const double *m, const double *x, const double *y,
                const double *z, double *vx, double *vy, double *vz)
{
    for (int iter = 0; iter < 100; ++iter) {
        double max_error = 0.0;
        for (int64_t c = 0; c < nc; ++c) {
            int ai = ci[c], aj = cj[c];
            double dx = x[aj] - x[ai];
            double dy = y[aj] - y[ai];
            double dz = z[aj] - z[ai];
            orc_min_image(cell, &dx, &dy, &dz);
            double dvx = vx[aj] - vx[ai];
            double dvy = vy[aj] - vy[ai];
            double dvz = vz[aj] - vz[ai];
            double r_dot_v = dx * dvx + dy * dvy + dz * dvz;
            double r2 = dx * dx + dy * dy + dz * dz;
            double err = fabs(r_dot_v);
            if (err > max_error) max_error = err;
            if (err > tol) {
                double imi = 1.0 / m[ai];
                double imj = 1.0 / m[aj];
                double ims = imi + imj;
                double lambda = r_dot_v / (r2 * ims);
                vx[ai] += lambda * dx * imi;
                vy[ai] += lambda * dy * imi;
                vz[ai] += lambda * dz * imi;
                vx[aj] -= lambda * dx * imj;
                vy[aj] -= lambda * dy * imj;
                vz[aj] -= lambda * dz * imj;
            }
        }
        if (max_error < tol) break;
    }
}

/* ------------------------------------------------------------ Andersen */

/* mt19937_64 as the C++ standard defines it ([rand.eng.mers], parameters
 * from [rand.predef]), written from that definition: seeding
 * x_i = f (x_{i-1} ^ (x_{i-1} >> 62)) + i, the twist over the 312-word
 * state, then tempering. */
void orc_mt64_seed(orc_mt64 *g, uint64_t seed)
{
    g->mt[0] = seed;
    for (int i = 1; i < 312; ++i)
        g->mt[i] = 6364136223846793005ull * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->idx = 312;
}

uint64_t orc_mt64_next(orc_mt64 *g)
{
    if (g->idx >= 312) {
        for (int i = 0; i < 312; ++i) {
            uint64_t x = (g->mt[i] & 0xFFFFFFFF80000000ull) |
                         (g->mt[(i + 1) % 312] & 0x7FFFFFFFull);
            uint64_t xa = x >> 1;
            if (x & 1ull) xa ^= 0xB5026F5AA96619E9ull;
            g->mt[i] = g->mt[(i + 156) % 312] ^ xa;
        }
        g->idx = 0;
    }
    uint64_t y = g->mt[g->idx++];
    y ^= (y >> 29) & 0x5555555555555555ull;
    y ^= (y << 17) & 0x71D67FFFEDA60000ull;
    y ^= (y << 37) & 0xFFF7EEE000000000ull;
    y ^= y >> 43;
    return y;
}

/* std::generate_canonical<double, 53> over a 64-bit engine (libstdc++,
 * GCC 11): one output, converted and scaled by 2^-64, kept below 1. */
static double mt64_canonical(orc_mt64 *g)
{
    double r = (double)orc_mt64_next(g) / 18446744073709551616.0;
    if (r >= 1.0) r = nextafter(1.0, 0.0);
    return r;
}

/* std::normal_distribution<double>(0, 1) of libstdc++: Marsaglia's polar
 * method, returning y * mult and keeping x * mult for the next draw. */
typedef struct {
    double saved;
    int available;
} orc_normal;

static double orc_normal_draw(orc_normal *nd, orc_mt64 *g)
{
    double ret;
    if (nd->available) {
        nd->available = 0;
        ret = nd->saved;
    } else {
        double x, y, r2;
        do {
            x = 2.0 * mt64_canonical(g) - 1.0;
            y = 2.0 * mt64_canonical(g) - 1.0;
            r2 = x * x + y * y;
        } while (r2 > 1.0 || r2 == 0.0);
        double mult = sqrt(-2.0 * log(r2) / r2);
        nd->saved = x * mult;
        nd->available = 1;
        ret = y * mult;
    }
    return ret * 1.0 + 0.0;
}

/* src/thermostat/andersen.cpp:14-32, one apply() call with the generator
 * seeded by `seed` (42 + calls made before). */
void orc_andersen(int64_t n, const double *m, double *vx, double *vy, double *vz, double dt,
                  double T0, double frequency, uint32_t seed)
{
    double prob = frequency * dt;
    if (prob <= 0.0) return;
    orc_mt64 g;
    orc_mt64_seed(&g, (uint64_t)seed);
    orc_normal nd = {0.0, 0};
    for (int64_t i = 0; i < n; ++i) {
        if (mt64_canonical(&g) < prob) {
            double sigma = sqrt(ORC_KB * T0 / m[i]);
            vx[i] = orc_normal_draw(&nd, &g) * sigma;
            vy[i] = orc_normal_draw(&nd, &g) * sigma;
            vz[i] = orc_normal_draw(&nd, &g) * sigma;
        }
    }
}

struct orc_sim {
    int64_t n;
    double cell[9];
    double *x, *y, *z, *vx, *vy, *vz, *fx, *fy, *fz, *m, *q;
    int32_t *types;
    int use_lj, nt;
    double *sig, *eps;
    double rc_lj, skin;
    int interval, counter;
    int64_t nb, na, nd, ni, ne;
    int32_t *bi, *bj, *ai, *aj, *ak, *di, *dj, *dk, *dl, *dper, *ii, *ij, *ik, *il, *ei, *ej;
    double *bk, *br0, *akt, *at0, *dkp, *dp0, *ikp, *ip0;
    int coulomb;
    double rc_c, alpha;
    int K[3];
    double dt;
    int thermo;
    double T0, ttau;
    int baro;
    double P0, ptau, beta;
    int physics;
    int64_t nc;
    int32_t *ci, *cj;
    double *cd, ctol;
    double and_freq;
    int64_t and_calls;
    /* state */
    int32_t *pi, *pj;
    int64_t np, pcap;
    double pe;
    int64_t step;
    double time, zeta, eps_dot;
};

static void *dup_bytes(const void *src, size_t bytes)
{
    void *p = malloc(bytes > 0 ? bytes : 1);
    if (src && bytes) memcpy(p, src, bytes);
    else memset(p, 0, bytes > 0 ? bytes : 1);
    return p;
}

orc_sim *orc_sim_create(const dmdx_system_desc *d)
{
    orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
    int64_t n = d->n_atoms;
    size_t nd8 = sizeof(double) * n;
    s->n = n;
    s->cell[0] = d->box[0];
    s->cell[4] = d->box[1];
    s->cell[8] = d->box[2];
    s->x = dup_bytes(d->pos_x, nd8);
    s->y = dup_bytes(d->pos_y, nd8);
    s->z = dup_bytes(d->pos_z, nd8);
    s->vx = dup_bytes(d->vel_x, nd8);
    s->vy = dup_bytes(d->vel_y, nd8);
    s->vz = dup_bytes(d->vel_z, nd8);
    s->fx = dup_bytes(NULL, nd8);
    s->fy = dup_bytes(NULL, nd8);
    s->fz = dup_bytes(NULL, nd8);
    s->m = dup_bytes(d->masses, nd8);
    s->q = dup_bytes(d->charges, nd8);
    s->types = dup_bytes(d->atom_types, sizeof(int32_t) * n);
    s->use_lj = d->use_lj;
    s->nt = d->lj_n_types > 0 ? d->lj_n_types : 1;
    s->sig = dup_bytes(d->lj_sigma, sizeof(double) * s->nt * s->nt);
    s->eps = dup_bytes(d->lj_epsilon, sizeof(double) * s->nt * s->nt);
    s->rc_lj = d->lj_cutoff;
    s->skin = d->list_skin;
    s->interval = d->rebuild_interval > 0 ? d->rebuild_interval : 10;
    s->counter = d->step_since_rebuild;
#define DUPI(dst, src, cnt) s->dst = dup_bytes(d->src, sizeof(int32_t) * (cnt))
#define DUPD(dst, src, cnt) s->dst = dup_bytes(d->src, sizeof(double) * (cnt))
    s->nb = d->n_bonds;
    DUPI(bi, bond_i, s->nb); DUPI(bj, bond_j, s->nb); DUPD(bk, bond_k, s->nb); DUPD(br0, bond_r0, s->nb);
    s->na = d->n_angles;
    DUPI(ai, angle_i, s->na); DUPI(aj, angle_j, s->na); DUPI(ak, angle_k, s->na);
    DUPD(akt, angle_k_theta, s->na); DUPD(at0, angle_theta0, s->na);
    s->nd = d->n_dihedrals;
    DUPI(di, dih_i, s->nd); DUPI(dj, dih_j, s->nd); DUPI(dk, dih_k, s->nd); DUPI(dl, dih_l, s->nd);
    DUPI(dper, dih_periodicity, s->nd); DUPD(dkp, dih_k_phi, s->nd); DUPD(dp0, dih_phi0, s->nd);
    s->ni = d->n_impropers;
    DUPI(ii, imp_i, s->ni); DUPI(ij, imp_j, s->ni); DUPI(ik, imp_k, s->ni); DUPI(il, imp_l, s->ni);
    DUPD(ikp, imp_k_phi, s->ni); DUPD(ip0, imp_phi0, s->ni);
    s->ne = d->n_excl;
    DUPI(ei, excl_i, s->ne); DUPI(ej, excl_j, s->ne);
    s->nc = d->n_constraints > 0 ? d->n_constraints : 0;
    DUPI(ci, con_i, s->nc); DUPI(cj, con_j, s->nc); DUPD(cd, con_distance, s->nc);
    s->ctol = d->con_tolerance;
    s->and_freq = d->andersen_frequency;
    s->and_calls = d->andersen_calls;
#undef DUPI
#undef DUPD
    s->coulomb = d->coulomb_type;
    s->rc_c = d->coulomb_cutoff;
    s->alpha = d->ewald_coeff;
    s->K[0] = d->pme_grid[0];
    s->K[1] = d->pme_grid[1];
    s->K[2] = d->pme_grid[2];
    s->dt = d->dt;
    s->thermo = d->thermostat;
    s->T0 = d->thermo_temperature;
    s->ttau = d->thermo_tau;
    s->baro = d->barostat;
    s->P0 = d->baro_pressure;
    s->ptau = d->baro_tau;
    s->beta = d->baro_compressibility;
    s->physics = d->physics;
    s->pcap = 0;
    return s;
}

void orc_sim_destroy(orc_sim *s)
{
    if (!s) return;
    void *ptrs[] = {s->x, s->y, s->z, s->vx, s->vy, s->vz, s->fx, s->fy, s->fz, s->m, s->q,
                    s->types, s->sig, s->eps, s->bi, s->bj, s->bk, s->br0, s->ai, s->aj, s->ak,
                    s->akt, s->at0, s->di, s->dj, s->dk, s->dl, s->dper, s->dkp, s->dp0, s->ii,
                    s->ij, s->ik, s->il, s->ikp, s->ip0, s->ei, s->ej, s->pi, s->pj,
                    s->ci, s->cj, s->cd};
    for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); ++k) free(ptrs[k]);
    free(s);
}

/* list radius: FAITHFUL lists only LJ pairs at lj_cutoff (+skin); CORRECTED
 * shares one list for LJ and real-space Coulomb at the larger cutoff. */
static double list_radius(const orc_sim *s)
{
    double rc = s->use_lj ? s->rc_lj : 0.0;
    if (s->physics != DMDX_FAITHFUL && s->coulomb != DMDX_COULOMB_NONE && s->rc_c > rc)
        rc = s->rc_c;
    return rc + s->skin;
}

static int64_t build_list(const orc_sim *s, int32_t **pi, int32_t **pj, int64_t *cap, int cells)
{
    double rc = list_radius(s);
    int64_t np;
    for (;;) {
        if (cells)
            np = orc_pairs_cells(s->n, s->x, s->y, s->z, s->cell, rc, *pi, *pj, *cap);
        else
            np = orc_pairs_n2(s->n, s->x, s->y, s->z, s->cell, rc, *pi, *pj, *cap);
        if (np <= *cap) break;
        *cap = np + np / 4 + 16;
        *pi = (int32_t *)realloc(*pi, sizeof(int32_t) * (*cap));
        *pj = (int32_t *)realloc(*pj, sizeof(int32_t) * (*cap));
    }
    if (s->physics != DMDX_FAITHFUL)
        np = orc_pairs_filter_excl(np, *pi, *pj, s->ne, s->ei, s->ej);
    return np;
}

static int has_list(const orc_sim *s)
{
    return s->use_lj || (s->physics != DMDX_FAITHFUL && s->coulomb != DMDX_COULOMB_NONE);
}

/* ForceEngine::compute (src/force/force_engine.cpp:10-25) over the
 * components in build_force_engine order (src/sim/simulation_engine.cpp:
 * 147-203): LJ, bonds, angles, dihedrals, impropers, CoulombDirect, PME,
 * exclusion.  `np`/`pi`/`pj` is the pair list used for the pair terms. */
static void forces_from_list(orc_sim *s, int64_t np, const int32_t *pi, const int32_t *pj,
                             double *fx, double *fy, double *fz, double *pe)
{
    int mode = s->physics;
    if (s->use_lj)
        orc_lj_pairs(np, pi, pj, s->x, s->y, s->z, s->cell, s->rc_lj, s->types, s->nt, s->sig,
                     s->eps, fx, fy, fz, pe);
    orc_bonds(s->nb, s->bi, s->bj, s->bk, s->br0, s->x, s->y, s->z, s->cell, fx, fy, fz, pe);
    orc_angles(s->na, s->ai, s->aj, s->ak, s->akt, s->at0, s->x, s->y, s->z, s->cell, mode, fx,
               fy, fz, pe);
    orc_dihedrals(s->nd, s->di, s->dj, s->dk, s->dl, s->dper, s->dkp, s->dp0, s->x, s->y, s->z,
                  s->cell, mode, fx, fy, fz, pe);
    orc_impropers(s->ni, s->ii, s->ij, s->ik, s->il, s->ikp, s->ip0, s->x, s->y, s->z, s->cell,
                  mode, fx, fy, fz, pe);
    if (s->coulomb == DMDX_COULOMB_NONE) return;
    if (mode == DMDX_FAITHFUL) {
        /* CoulombDirect's list is never built by build_simulation (Q3): 0 */
        if (s->coulomb == DMDX_COULOMB_PME)
            orc_pme(s->n, s->x, s->y, s->z, s->q, s->cell, s->K[0], s->K[1], s->K[2], s->alpha,
                    mode, fx, fy, fz, pe);
        if (s->coulomb == DMDX_COULOMB_PME || s->ne > 0)
            orc_exclusion(s->ne, s->ei, s->ej, s->x, s->y, s->z, s->cell, s->alpha, s->q, mode,
                          fx, fy, fz, pe);
    } else {
        orc_coulomb_pairs(np, pi, pj, s->x, s->y, s->z, s->cell, s->rc_c, s->alpha, s->q, fx, fy,
                          fz, pe);
        if (s->coulomb == DMDX_COULOMB_PME) {
            orc_pme(s->n, s->x, s->y, s->z, s->q, s->cell, s->K[0], s->K[1], s->K[2], s->alpha,
                    mode, fx, fy, fz, pe);
            orc_exclusion(s->ne, s->ei, s->ej, s->x, s->y, s->z, s->cell, s->alpha, s->q, mode,
                          fx, fy, fz, pe);
        }
    }
}

void orc_sim_compute(orc_sim *s)
{
    memset(s->fx, 0, sizeof(double) * s->n);
    memset(s->fy, 0, sizeof(double) * s->n);
    memset(s->fz, 0, sizeof(double) * s->n);
    s->pe = 0.0;
    if (has_list(s)) {
        /* src/force/lennard_jones.cpp:53-55: rebuild when the counter has
         * reached the interval; the counter increments per compute (:91) */
        if (s->counter >= s->interval) {
            s->np = build_list(s, &s->pi, &s->pj, &s->pcap, 0);
            s->counter = 0;
        }
    }
    forces_from_list(s, s->np, s->pi, s->pj, s->fx, s->fy, s->fz, &s->pe);
    if (has_list(s)) ++s->counter;
}

void orc_sim_forces_exact(orc_sim *s, double *fx, double *fy, double *fz, double *pe)
{
    memset(fx, 0, sizeof(double) * s->n);
    memset(fy, 0, sizeof(double) * s->n);
    memset(fz, 0, sizeof(double) * s->n);
    *pe = 0.0;
    int32_t *pi = NULL, *pj = NULL;
    int64_t cap = 0, np = 0;
    if (has_list(s)) np = build_list(s, &pi, &pj, &cap, 1);
    forces_from_list(s, np, pi, pj, fx, fy, fz, pe);
    free(pi);
    free(pj);
}

/* SimulationEngine::step, src/sim/simulation_engine.cpp:53-86 */
void orc_sim_step(orc_sim *s, int64_t nsteps)
{
    for (int64_t it = 0; it < nsteps; ++it) {
        orc_sim_compute(s);
        orc_half_kick(s->n, s->vx, s->vy, s->vz, s->fx, s->fy, s->fz, s->m, s->dt);
        orc_advance(s->n, s->x, s->y, s->z, s->vx, s->vy, s->vz, s->dt);
        if (s->nc > 0)
            orc_shake(s->nc, s->ci, s->cj, s->cd, s->ctol, s->cell, s->m, s->x, s->y, s->z);
        orc_sim_compute(s);
        orc_half_kick(s->n, s->vx, s->vy, s->vz, s->fx, s->fy, s->fz, s->m, s->dt);
        if (s->nc > 0)
            orc_rattle(s->nc, s->ci, s->cj, s->ctol, s->cell, s->m, s->x, s->y, s->z, s->vx,
                       s->vy, s->vz);
        if (s->thermo == DMDX_THERMO_ANDERSEN) {
            if (!(s->and_freq * s->dt <= 0.0)) {  /* andersen.cpp:16 */
                orc_andersen(s->n, s->m, s->vx, s->vy, s->vz, s->dt, s->T0, s->and_freq,
                             (uint32_t)(42u + (uint32_t)s->and_calls));
                ++s->and_calls;
            }
        } else if (s->thermo == DMDX_THERMO_BERENDSEN)
            orc_berendsen(s->n, s->m, s->vx, s->vy, s->vz, s->dt, s->T0, s->ttau);
        else if (s->thermo == DMDX_THERMO_NOSE_HOOVER)
            orc_nose_hoover(s->n, s->m, s->vx, s->vy, s->vz, s->dt, s->T0, s->ttau, &s->zeta);
        if (s->baro == DMDX_BARO_BERENDSEN)
            orc_berendsen_barostat(s->n, s->m, s->vx, s->vy, s->vz, s->x, s->y, s->z, s->fx,
                                   s->fy, s->fz, s->cell, s->dt, s->P0, s->ptau, s->beta);
        else if (s->baro == DMDX_BARO_ANDERSEN)
            orc_andersen_barostat(s->n, s->m, s->vx, s->vy, s->vz, s->x, s->y, s->z, s->fx,
                                  s->fy, s->fz, s->cell, s->dt, s->P0, s->ptau, &s->eps_dot);
        ++s->step;
        s->time += s->dt;
    }
}

void orc_sim_get(const orc_sim *s, double *x, double *y, double *z,
                 double *vx, double *vy, double *vz, double *fx, double *fy, double *fz)
{
    size_t b = sizeof(double) * s->n;
    if (x) memcpy(x, s->x, b);
    if (y) memcpy(y, s->y, b);
    if (z) memcpy(z, s->z, b);
    if (vx) memcpy(vx, s->vx, b);
    if (vy) memcpy(vy, s->vy, b);
    if (vz) memcpy(vz, s->vz, b);
    if (fx) memcpy(fx, s->fx, b);
    if (fy) memcpy(fy, s->fy, b);
    if (fz) memcpy(fz, s->fz, b);
}

void orc_sim_set_positions(orc_sim *s, const double *x, const double *y, const double *z)
{
    size_t b = sizeof(double) * s->n;
    memcpy(s->x, x, b);
    memcpy(s->y, y, b);
    memcpy(s->z, z, b);
}

void orc_sim_set_velocities(orc_sim *s, const double *vx, const double *vy, const double *vz)
{
    size_t b = sizeof(double) * s->n;
    memcpy(s->vx, vx, b);
    memcpy(s->vy, vy, b);
    memcpy(s->vz, vz, b);
}

void orc_sim_scalars(const orc_sim *s, double out[11])
{
    out[0] = s->pe;
    out[1] = orc_kinetic_energy(s->n, s->m, s->vx, s->vy, s->vz);
    out[2] = (double)s->step;
    out[3] = s->time;
    out[4] = s->zeta;
    out[5] = s->eps_dot;
    out[6] = s->cell[0];
    out[7] = s->cell[4];
    out[8] = s->cell[8];
    out[9] = (double)s->counter;
    out[10] = (double)s->np;
}
