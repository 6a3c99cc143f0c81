/*
 * dmd_oracle.h -- FP64 CPU restatement of DMD's step path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * library (dmd_amd/libdmdx.so) and the timed CPU baseline; it is never the
 * thing measured on the GPU nor a fallback of the product path.
 *
 * Every function restates one reference function in the same operation
 * order (see the file:line cited at each definition in dmd_oracle.c) and is
 * compiled with -ffp-contract=off so that, in FAITHFUL mode, its results are
 * bit-identical to the reference translation units compiled the same way.
 * CORRECTED mode swaps in textbook formulas where DMD is physically wrong.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   min_image, cell_volume      DMD's test_cell.cpp values; half-away rounding
 *   pairs_n2 / pairs_cells      O(N^2) brute force = DMD's algorithm; the two
 *                               agree as sorted sets; DMD's own build (_ref)
 *   lj, coulomb pairs, exclusion,
 *   bonds, angles, dihedrals,
 *   impropers                   bit-exact against DMD's translation units
 *                               (oracle/_ref) in FAITHFUL mode; DMD's unit
 *                               test values; finite differences of the
 *                               energy (CORRECTED) and the documented
 *                               anti-gradient quirks (FAITHFUL)
 *   theta4, dtheta4, self energy
 *                               DMD's test_coulomb_pme.cpp values; closed form
 *   fft3d                       numpy.fft on random data
 *   pme (CORRECTED)             plain Ewald sum (order-4 convergence), NaCl
 *                               Madelung constant, finite differences
 *   pme (FAITHFUL)              parity partially pinned: forces are exactly
 *                               -grad(E)/(nx ny nz) (finite differences); the
 *                               multiplied B-spline modulus is pinned only as
 *                               "farther from Ewald than CORRECTED"; DMD's
 *                               padded-grid indexing cannot be reproduced and
 *                               is replaced by a consistent layout
 *   ewald                       Madelung constant, finite differences
 *   integrator, KE, T, thermostats, virial, pressure, barostats,
 *   sim_* (step sequence)       bit-exact against DMD's translation units
 *                               (_ref); closed forms; DMD's thermostat and
 *                               barostat unit-test values
 *   shake, rattle               DMD's test_constraints.cpp cases; bit-exact
 *                               against DMD's Constraints (_ref); converged
 *                               bond lengths and r.v
 *   mt64                        the C++ standard's known answer (10000th
 *                               output of the default-seeded mt19937_64)
 *   andersen                    bit-exact against DMD's AndersenThermostat
 *                               (_ref, GCC 11 libstdc++ distributions);
 *                               collision fraction and Maxwell variance
 */
#ifndef DMD_ORACLE_H
#define DMD_ORACLE_H

#include <stdint.h>
#include "../include/dmdx.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_KE 138.935456          /* k_e, src/force/coulomb_direct.cpp:42  */
#define ORC_KB 0.008314462618      /* src/core/constants.h:6                */
#define ORC_PI 3.14159265358979323846

/* ---- core */
void orc_set_cell_type(int type);
void orc_min_image(const double cell[9], double *dx, double *dy, double *dz);
double orc_cell_volume(const double cell[9]);

/* ---- pair lists: all i<j with minimum-image r^2 < rc^2, lexicographic.
 * Return the number of pairs; only the first `cap` are written. */
int64_t orc_pairs_n2(int64_t n, const double *x, const double *y, const double *z,
                     const double cell[9], double rc, int32_t *pi, int32_t *pj, int64_t cap);
int64_t orc_pairs_cells(int64_t n, const double *x, const double *y, const double *z,
                        const double cell[9], double rc, int32_t *pi, int32_t *pj, int64_t cap);
/* remove excluded pairs (excl given as any-order pairs); returns new count */
int64_t orc_pairs_filter_excl(int64_t np, int32_t *pi, int32_t *pj,
                              int64_t nex, const int32_t *ei, const int32_t *ej);

/* ---- force components (accumulate into f and *energy) */
void orc_lj_pairs(int64_t np, const int32_t *pi, const int32_t *pj,
                  const double *x, const double *y, const double *z, const double cell[9],
                  double rc, const int32_t *types, int32_t n_types,
                  const double *sigma, const double *epsilon,
                  double *fx, double *fy, double *fz, double *energy);
void orc_coulomb_pairs(int64_t np, const int32_t *pi, const int32_t *pj,
                       const double *x, const double *y, const double *z, const double cell[9],
                       double rc, double alpha, const double *q,
                       double *fx, double *fy, double *fz, double *energy);
void orc_exclusion(int64_t ne, const int32_t *ei, const int32_t *ej,
                   const double *x, const double *y, const double *z, const double cell[9],
                   double alpha, const double *q, int mode,
                   double *fx, double *fy, double *fz, double *energy);
void orc_bonds(int64_t nb, const int32_t *bi, const int32_t *bj, const double *k, const double *r0,
               const double *x, const double *y, const double *z, const double cell[9],
               double *fx, double *fy, double *fz, double *energy);
void orc_angles(int64_t na, const int32_t *ai, const int32_t *aj, const int32_t *ak,
                const double *kt, const double *t0,
                const double *x, const double *y, const double *z, const double cell[9], int mode,
                double *fx, double *fy, double *fz, double *energy);
void orc_dihedrals(int64_t nd, const int32_t *di, const int32_t *dj, const int32_t *dk,
                   const int32_t *dl, const int32_t *per, const double *kp, const double *p0,
                   const double *x, const double *y, const double *z, const double cell[9],
                   int mode, double *fx, double *fy, double *fz, double *energy);
void orc_impropers(int64_t nd, const int32_t *di, const int32_t *dj, const int32_t *dk,
                   const int32_t *dl, const double *kp, const double *p0,
                   const double *x, const double *y, const double *z, const double cell[9],
                   int mode, double *fx, double *fy, double *fz, double *energy);

/* ---- SPME */
double orc_theta4(double u);
double orc_dtheta4(double u);
double orc_pme_self_energy(double alpha, int64_t n, const double *q, double k_e);
/* in-place complex 3D DFT, x fastest; sign -1 forward, +1 backward, unnormalised */
void orc_fft3d(int nx, int ny, int nz, double *re, double *im, int sign);
void orc_pme(int64_t n, const double *x, const double *y, const double *z, const double *q,
             const double cell[9], int nx, int ny, int nz, double alpha, int mode,
             double *fx, double *fy, double *fz, double *energy);
/* plain Ewald sum (real space over images within rc_real, reciprocal over
 * |m_d| <= kmax), corrected physics; for pinning orc_pme */
void orc_ewald(int64_t n, const double *x, const double *y, const double *z, const double *q,
               const double cell[9], double alpha, double rc_real, int kmax,
               double *fx, double *fy, double *fz, double *energy);

/* ---- integration and coupling */
void orc_half_kick(int64_t n, double *vx, double *vy, double *vz,
                   const double *fx, const double *fy, const double *fz,
                   const double *m, double dt);
void orc_advance(int64_t n, double *x, double *y, double *z,
                 const double *vx, const double *vy, const double *vz, double dt);
/* ---- constraints (src/integrate/constraints.cpp) */
void orc_shake(int64_t nc, const int32_t *ci, const int32_t *cj, const double *dist, double tol,
               const double cell[9], const double *m, double *x, double *y, double *z);
void orc_rattle(int64_t nc, const int32_t *ci, const int32_t *cj, double tol,
                const double cell[9], const double *m, const double *x, const double *y,
                const double *z, double *vx, double *vy, double *vz);

/* ---- Andersen thermostat (src/thermostat/andersen.cpp) and its stream */
typedef struct orc_mt64 {
    uint64_t mt[312];
    int idx;
} orc_mt64;
void orc_mt64_seed(orc_mt64 *g, uint64_t seed);
uint64_t orc_mt64_next(orc_mt64 *g);
void orc_andersen(int64_t n, const double *m, double *vx, double *vy, double *vz, double dt,
                  double T0, double frequency, uint32_t seed);

double orc_kinetic_energy(int64_t n, const double *m, const double *vx, const double *vy,
                          const double *vz);
double orc_temperature(double ekin, int64_t n);
void orc_berendsen(int64_t n, const double *m, double *vx, double *vy, double *vz,
                   double dt, double T0, double tau);
void orc_nose_hoover(int64_t n, const double *m, double *vx, double *vy, double *vz,
                     double dt, double T0, double tau, double *zeta);
double orc_virial(int64_t n, const double *x, const double *y, const double *z,
                  const double *fx, const double *fy, const double *fz);
double orc_pressure(double ekin, double w, double volume);
void orc_berendsen_barostat(int64_t n, const double *m, const double *vx, const double *vy,
                            const double *vz, double *x, double *y, double *z,
                            const double *fx, const double *fy, const double *fz,
                            double cell[9], double dt, double P0, double tau, double beta);
void orc_andersen_barostat(int64_t n, const double *m, const double *vx, const double *vy,
                           const double *vz, double *x, double *y, double *z,
                           const double *fx, const double *fy, const double *fz,
                           double cell[9], double dt, double P0, double tau, double *eps_dot);

/* ---- simulation (SimulationEngine::step sequence) */
typedef struct orc_sim orc_sim;
orc_sim *orc_sim_create(const dmdx_system_desc *d);
void orc_sim_destroy(orc_sim *s);
void orc_sim_step(orc_sim *s, int64_t nsteps);
/* one ForceEngine::compute (zeroes F and PE first, as step() does) */
void orc_sim_compute(orc_sim *s);
/* forces/energy at the current positions from a freshly built exact list;
 * does not touch the list counter or any state */
void orc_sim_forces_exact(orc_sim *s, double *fx, double *fy, double *fz, double *pe);
void orc_sim_get(const orc_sim *s, double *x, double *y, double *z,
                 double *vx, double *vy, double *vz, double *fx, double *fy, double *fz);
void orc_sim_set_positions(orc_sim *s, const double *x, const double *y, const double *z);
void orc_sim_set_velocities(orc_sim *s, const double *vx, const double *vy, const double *vz);
/* scalars: [pe, ke, step, time, zeta, eps_dot, box_x, box_y, box_z, counter, n_pairs] */
void orc_sim_scalars(const orc_sim *s, double out[11]);

#ifdef __cplusplus
}
#endif
#endif
