/*
 * dmdx.h -- C-ABI of the MI355X (gfx950) step engine for DMD.
 *
 * The library replaces DMD's per-step force + integration path
 * (SimulationEngine::step, src/sim/simulation_engine.cpp:53-86, and the
 * ForceComponent implementations under src/force/) with hand-written HIP
 * kernels.  Every entry point takes plain pointers and sizes; nothing here
 * depends on C++ or on torch.  Every function returns an int status
 * (DMDX_OK == 0, negative on error); the message of the last error raised on
 * the calling thread is returned by dmdx_last_error().  Nothing is thrown
 * across the ABI.  A handle must be used by one host thread at a time.
 *
 * Two levels are exported:
 *   1. ForceComponent level (dmdx_component_*): the drop-in for
 *      ForceComponent::compute (src/force/force_component.h:8-22).  Positions
 *      and forces are HOST arrays, exactly like the std::span arguments of the
 *      reference; forces and energy ACCUMULATE (+=), nothing is zeroed
 *      (pinned by tests/unit/force/test_force_engine.cpp:46-57).
 *   2. Engine level (dmdx_engine_*): the drop-in for SimulationEngine
 *      (src/sim/simulation_engine.h:16-76).  State stays resident in HBM
 *      between calls; host copies happen only in get/set calls.
 */
#ifndef DMDX_H
#define DMDX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: the first five mirror ForceError (src/core/error.h:6-12) */
enum {
    DMDX_OK = 0,
    DMDX_E_NAN = -1,            /* ForceError::NaN: non-finite force/energy   */
    DMDX_E_BOND_EXPLOSION = -2, /* ForceError::BondExplosion                  */
    DMDX_E_MISMATCH = -3,       /* ForceError::Mismatch: array sizes disagree */
    DMDX_E_PARM = -4,           /* ForceError::ParmError: bad parameter       */
    DMDX_E_GPU = -5,            /* ForceError::GPUError: HIP runtime failure  */
    DMDX_E_OOM = -6,            /* device allocation failed / over budget     */
    DMDX_E_INVALID = -7,        /* invalid handle or argument                 */
    DMDX_E_BUSY = -8,           /* server: request cannot be admitted now;
                                   snapshots: every staging slot is in use    */
    DMDX_E_IO = -9              /* file I/O (H5MD trajectory) failed          */
};

/* ---- physics mode.
 * FAITHFUL reproduces DMD's arithmetic as written, including its known
 * quirks (empty real-space Coulomb list in engine builds, anti-restoring
 * angle force, in-plane dihedral/improper gradients, exclusion force sign,
 * multiplied B-spline modulus and 1/K^3-scaled PME forces).
 * CORRECTED uses the textbook formulas for the same energy functions. */
enum { DMDX_FAITHFUL = 0, DMDX_CORRECTED = 1 };

/* ---- precision of the device path.
 * FP64: every kernel in double precision (the "plain" path).
 * MIXED: pair/spline/FFT arithmetic in FP32, forces accumulated in FP64,
 *        integration, thermostat and energy/virial reductions in FP64. */
enum { DMDX_FP64 = 0, DMDX_MIXED = 1 };

/* ---- cell type (CellType, src/core/cell.h:3-8).  The minimum image wraps
 *      by the matrix diagonal for ORTHORHOMBIC and is the identity for every
 *      other type (src/core/cell.cpp:26-38); the volume is the determinant
 *      of the matrix for all types (:19-24). */
enum {
    DMDX_CELL_ORTHORHOMBIC = 0,
    DMDX_CELL_TRICLINIC = 1,
    DMDX_CELL_HEXAGONAL = 2,
    DMDX_CELL_RHOMBIC_DODECAHEDRON = 3
};

/* ---- electrostatics (SimulationConfig::coulomb_type,
 *      src/sim/simulation_config.h:52-56) */
enum { DMDX_COULOMB_NONE = 0, DMDX_COULOMB_CUTOFF = 1, DMDX_COULOMB_PME = 2 };

/* ---- coupling (ThermostatConfig / BarostatConfig,
 *      src/sim/simulation_config.h:13-25) */
enum {
    DMDX_THERMO_NONE = 0,
    DMDX_THERMO_BERENDSEN = 1,
    DMDX_THERMO_NOSE_HOOVER = 2,
    DMDX_THERMO_ANDERSEN = 3
};
enum { DMDX_BARO_NONE = 0, DMDX_BARO_BERENDSEN = 1, DMDX_BARO_ANDERSEN = 2 };

/*
 * One system, as a POD.  Mirrors the fields of SimulationConfig
 * (src/sim/simulation_config.h:27-84) that build_simulation copies into
 * SystemData and the force components (src/sim/simulation_engine.cpp:
 * 248-386).  All arrays are host pointers; the library copies them, matching
 * the copy semantics of the reference setters.  NULL optional arrays mean:
 * velocities 0, charges 0, atom types 0.
 */
typedef struct dmdx_system_desc {
    int64_t n_atoms;
    double box[3];                 /* orthorhombic edge lengths (nm)          */
    const double *pos_x, *pos_y, *pos_z;
    const double *vel_x, *vel_y, *vel_z;
    const double *masses;
    const double *charges;
    const int32_t *atom_types;

    /* Lennard-Jones (LennardJonesParams, src/force/lennard_jones.h:7-11) */
    int32_t use_lj;
    int32_t lj_n_types;
    const double *lj_sigma;        /* n_types*n_types, row-major              */
    const double *lj_epsilon;      /* n_types*n_types, row-major              */
    double lj_cutoff;
    double list_skin;              /* list built at cutoff+skin (0 = DMD)     */
    int32_t rebuild_interval;      /* > 0: computes between rebuilds; 0: DMD's
                                      10; -1: rebuild when an atom has moved
                                      more than list_skin/2 since the last
                                      build (Verlet criterion)               */
    int32_t step_since_rebuild;    /* initial counter (DMD: 10)               */

    /* bonded terms (HarmonicBondParams etc., src/force/harmonic_*.h) */
    int64_t n_bonds;
    const int32_t *bond_i, *bond_j;
    const double *bond_k, *bond_r0;
    int64_t n_angles;
    const int32_t *angle_i, *angle_j, *angle_k;
    const double *angle_k_theta, *angle_theta0;
    int64_t n_dihedrals;
    const int32_t *dih_i, *dih_j, *dih_k, *dih_l;
    const int32_t *dih_periodicity;
    const double *dih_k_phi, *dih_phi0;
    int64_t n_impropers;
    const int32_t *imp_i, *imp_j, *imp_k, *imp_l;
    const double *imp_k_phi, *imp_phi0;

    /* electrostatics */
    int32_t coulomb_type;          /* DMDX_COULOMB_*                          */
    double coulomb_cutoff;
    double ewald_coeff;            /* alpha, 1/nm                             */
    int32_t pme_grid[3];           /* nx, ny, nz                              */
    int32_t pme_order;             /* only 4 is implemented (as DMD)          */
    int64_t n_excl;                /* excluded pairs (SystemData::excl_i/j)   */
    const int32_t *excl_i, *excl_j;

    /* integration and coupling */
    double dt;                     /* ps                                      */
    int32_t thermostat;            /* DMDX_THERMO_*                           */
    double thermo_temperature, thermo_tau;
    int32_t barostat;              /* DMDX_BARO_*                             */
    double baro_pressure, baro_tau, baro_compressibility;

    int32_t physics;               /* DMDX_FAITHFUL | DMDX_CORRECTED          */
    int32_t precision;             /* DMDX_FP64 | DMDX_MIXED                  */

    /* SHAKE / RATTLE distance constraints (ConstraintsParams,
     * src/integrate/constraints.h:8-13): positions are corrected after the
     * drift and velocities after the closing kick
     * (src/sim/simulation_engine.cpp:67,78).  build_simulation creates them
     * for constraint_type "shake" (:293-301).  0 constraints: none. */
    int64_t n_constraints;
    const int32_t *con_i, *con_j;
    const double *con_distance;    /* target distances (nm)                   */
    double con_tolerance;          /* SimulationConfig::constraint_tolerance  */

    /* Andersen thermostat (DMDX_THERMO_ANDERSEN; AndersenParams,
     * src/thermostat/andersen.h:5-8): thermo_temperature is its T0. */
    double andersen_frequency;     /* collisions per ps                       */
    int64_t andersen_calls;        /* apply() calls already made: the call
                                      k draws from mt19937_64 seeded with
                                      42 + k (src/thermostat/andersen.h:18,
                                      andersen.cpp:19,31)                     */
} dmdx_system_desc;

/* Scalars of one system after the last step / compute. */
typedef struct dmdx_scalars {
    double potential_energy;       /* SystemData::potential_energy            */
    double energy_terms[8];        /* lj, coulomb real space, bond, angle,
                                      dihedral, improper, pme (incl. self),
                                      exclusion (subtracted)                  */
    double kinetic_energy;         /* of the velocities at the end of step()  */
    double zeta;                   /* NoseHooverThermostat::zeta()            */
    double piston_velocity;        /* AndersenBarostat::piston_velocity()     */
    double virial;                 /* virial() of the last barostat call      */
    double pressure;               /* pressure() of the last barostat call    */
    double box[3];                 /* Cell::matrix[0], [4], [8]               */
    int32_t step_since_rebuild;    /* LennardJones::step_since_rebuild()      */
    int32_t flags;                 /* 1: neighbour-list overflow, 2: non-finite */
    int64_t steps_done;
    int64_t list_builds;           /* neighbour-list builds since creation    */
} dmdx_scalars;

/* ---- errors and device ------------------------------------------------ */

/* Message of the last error raised on the calling thread ("" if none). */
const char *dmdx_last_error(void);

/* Number of visible GPUs and free/total HBM of the current one (any pointer
 * may be NULL). */
int dmdx_device_info(int32_t *device_count, int64_t *free_bytes, int64_t *total_bytes);

/* ---- engine: SimulationEngine (src/sim/simulation_engine.h:16-76) ------ */

typedef struct dmdx_engine dmdx_engine;

/* build_simulation (src/sim/simulation_engine.cpp:248-386) for n_systems
 * independent systems that step together in the same launches.  All systems
 * must share one precision.  State is copied to HBM and stays there. */
int dmdx_engine_create(const dmdx_system_desc *descs, int32_t n_systems, dmdx_engine **out);
int dmdx_engine_destroy(dmdx_engine *e);

/* reuse_forces: skip a first compute whose result equals the previous
 * compute's (default 1); use_graph: replay hipGraph-captured blocks of
 * graph_steps steps (default 0). */
int dmdx_engine_set_options(dmdx_engine *e, int32_t reuse_forces, int32_t use_graph,
                            int32_t graph_steps);

/* SimulationEngine::step() repeated nsteps times
 * (src/sim/simulation_engine.cpp:53-86).  Asynchronous: returns once the
 * work is queued; dmdx_engine_sync or any get call waits for it. */
int dmdx_engine_step(dmdx_engine *e, int64_t nsteps);

/* One ForceEngine::compute on the current positions with forces and energy
 * zeroed first (src/force/force_engine.cpp:10-25 as called by step()). */
int dmdx_engine_compute(dmdx_engine *e);

int dmdx_engine_sync(dmdx_engine *e);
int dmdx_engine_n_systems(dmdx_engine *e, int32_t *n);

/* what: 0 positions, 1 velocities, 2 forces (SystemData::pos_* / vel_* /
 * forces_*).  Host arrays of n_atoms doubles; NULL skips a component. */
int dmdx_engine_get_arrays(dmdx_engine *e, int32_t sys, int32_t what, double *x, double *y,
                           double *z);
int dmdx_engine_set_arrays(dmdx_engine *e, int32_t sys, int32_t what, const double *x,
                           const double *y, const double *z);
int dmdx_engine_get_scalars(dmdx_engine *e, int32_t sys, dmdx_scalars *out);

/* LennardJones::set_step_since_rebuild (src/force/lennard_jones.h:37), as
 * restored by SimulationEngine::load_checkpoint (simulation_engine.cpp:127-137). */
int dmdx_engine_set_step_since_rebuild(dmdx_engine *e, int32_t sys, int32_t counter);

/* Exclude (0) or include (1) a system in subsequent steps. */
int dmdx_engine_set_active(dmdx_engine *e, int32_t sys, int32_t active);

/* Move system src_sys of src, with its whole state (positions, velocities,
 * forces, list rows, cells, frames, counters, thermostat and barostat
 * variables, step count), into slot dst_sys of dst, replacing what was there.
 * Both slots must have been created from the same description (same atoms,
 * terms, cell grid, list rows and PME grid; DMDX_E_MISMATCH otherwise) in
 * engines of one precision.  The moved system's later trajectory is the one
 * it would have had in src.  No counterpart in the reference (one
 * SimulationEngine owns one system); the server merges cohorts with it. */
int dmdx_engine_adopt(dmdx_engine *dst, int32_t dst_sys, dmdx_engine *src, int32_t src_sys);

/* Current pair list of a system as the sorted set {(i, j) : i < j} (local
 * atom indices), i.e. the pairs_i_/pairs_j_ of LennardJones
 * (src/force/lennard_jones.h:43).  *count receives the number of pairs; at
 * most cap are written. */
int dmdx_engine_get_pairs(dmdx_engine *e, int32_t sys, int32_t *pi, int32_t *pj, int64_t cap,
                          int64_t *count);

/* Frames of a running engine without stopping it: the state behind
 * SimulationEngine::save_frame / save_checkpoint (simulation_engine.cpp:
 * 88-106, :117-143), taken where the host reads sys_ after a step.
 * _begin enqueues, after the work enqueued so far, a copy of system sys's
 * positions and velocities (velocities with any pending thermostat scale
 * applied; the engine's own state is not changed), its box edges and its
 * steps_done into a device staging slot, then an asynchronous copy of the
 * slot to pinned host memory on a copy stream of its own; it returns at
 * once with a ticket, and later steps overlap the copy.  _read waits for
 * that ticket's copy and writes positions and velocities as interleaved
 * (n, 3) rows (the layout of the H5MD position/velocity datasets,
 * h5md_writer.cpp:200-225; either may be NULL), the box edges and
 * steps_done.  At most DMDX_SNAPSHOT_SLOTS frames are in flight: a further
 * _begin returns DMDX_E_BUSY until one is read.  The staging (device and
 * pinned host, DMDX_SNAPSHOT_SLOTS x (64 + 48 n_max) bytes) is allocated on
 * the first snapshot and freed with the engine. */
#define DMDX_SNAPSHOT_SLOTS 4
int dmdx_engine_snapshot_begin(dmdx_engine *e, int32_t sys, int64_t *ticket);
int dmdx_engine_snapshot_read(dmdx_engine *e, int64_t ticket, double *pos_xyz, double *vel_xyz,
                              double box[3], int64_t *steps_done);

/* ---- components: ForceComponent (src/force/force_component.h:8-22) ----- */

typedef struct dmdx_component dmdx_component;

/* LennardJones(LennardJonesParams, cutoff) (src/force/lennard_jones.h:15);
 * sigma/epsilon are n_types*n_types row-major (lennard_jones.h:7-11).  The
 * list has no skin and is rebuilt every 10 computes, the counter starting
 * at 10 (lennard_jones.cpp:9-10, :53-55, :91). */
int dmdx_lj_create(int32_t n_types, const double *sigma, const double *epsilon, double cutoff,
                   dmdx_component **out);
/* LennardJones::set_atom_types (lennard_jones.h:28) */
int dmdx_lj_set_atom_types(dmdx_component *c, const int32_t *types, int64_t n);
/* LennardJones::step_since_rebuild / set_step_since_rebuild (lennard_jones.h:36-37) */
int dmdx_lj_get_step_since_rebuild(dmdx_component *c, int32_t *out);
int dmdx_lj_set_step_since_rebuild(dmdx_component *c, int32_t s);

/* CoulombDirect(CoulombDirectParams{cutoff, ewald_coefficient})
 * (src/force/coulomb_direct.h:6-19): erfc real-space pairs over a list that
 * only dmdx_component_rebuild_pair_list fills (as in the reference). */
int dmdx_coulomb_direct_create(double cutoff, double ewald_coefficient, dmdx_component **out);

/* CoulombExclusion (src/force/coulomb_exclusion.h:6-25) */
int dmdx_coulomb_exclusion_create(dmdx_component **out);
/* CoulombExclusion::set_exclusions (coulomb_exclusion.h:21) */
int dmdx_coulomb_exclusion_set_exclusions(dmdx_component *c, const int32_t *excl_i,
                                          const int32_t *excl_j, int64_t n);
/* No counterpart in the reference (its exclusion term is the bare 1/r of
 * FAITHFUL, coulomb_exclusion.cpp:29-54): the Ewald coefficient of the erf
 * form that DMDX_CORRECTED subtracts, which must be the PME's.  Default 3.5,
 * PMEParams' default (coulomb_pme.h:7-12); unused under DMDX_FAITHFUL. */
int dmdx_coulomb_exclusion_set_ewald_coefficient(dmdx_component *c, double alpha);

/* CoulombPME(PMEParams{cutoff, ewald_coefficient, nx, ny, nz, spline_order})
 * (src/force/coulomb_pme.h:7-16); spline_order 4 only, as the reference. */
int dmdx_pme_create(double cutoff, double ewald_coefficient, int32_t nx, int32_t ny, int32_t nz,
                    int32_t spline_order, dmdx_component **out);
/* CoulombPME::theta4 / dtheta4 / calc_self_energy (coulomb_pme.h:39-41) */
int dmdx_pme_theta4(double u, double *out);
int dmdx_pme_dtheta4(double u, double *out);
int dmdx_pme_self_energy(double alpha, const double *charges, int64_t n, double k_e, double *out);

/* HarmonicBond, HarmonicAngle, PeriodicDihedral, HarmonicImproper from their
 * parameter structs (src/force/harmonic_bond.h:7-11, harmonic_angle.h:7-11,
 * periodic_dihedral.h:7-12, harmonic_improper.h:7-11). */
int dmdx_harmonic_bond_create(int64_t n, const int32_t *i, const int32_t *j, const double *k,
                              const double *r0, dmdx_component **out);
int dmdx_harmonic_angle_create(int64_t n, const int32_t *i, const int32_t *j, const int32_t *k,
                               const double *k_theta, const double *theta0, dmdx_component **out);
int dmdx_periodic_dihedral_create(int64_t n, const int32_t *i, const int32_t *j, const int32_t *k,
                                  const int32_t *l, const int32_t *periodicity,
                                  const double *k_phi, const double *phi0, dmdx_component **out);
int dmdx_harmonic_improper_create(int64_t n, const int32_t *i, const int32_t *j, const int32_t *k,
                                  const int32_t *l, const double *k_phi, const double *phi0,
                                  dmdx_component **out);

/* set_charges of CoulombDirect / CoulombPME / CoulombExclusion
 * (coulomb_direct.h:26, coulomb_pme.h:26, coulomb_exclusion.h:20) */
int dmdx_component_set_charges(dmdx_component *c, const double *charges, int64_t n);

/* rebuild_pair_list of LennardJones / CoulombDirect (lennard_jones.h:29-34,
 * coulomb_direct.h:27-32); cell is Cell::matrix (row-major 3x3) of an
 * orthorhombic Cell. */
int dmdx_component_rebuild_pair_list(dmdx_component *c, const double *x, const double *y,
                                     const double *z, int64_t n, const double cell[9]);

/* ForceComponent::compute: forces and *energy_out are ADDED to. */
int dmdx_component_compute(dmdx_component *c, const double *x, const double *y, const double *z,
                           int64_t n, const double cell[9], double *fx, double *fy, double *fz,
                           double *energy_out);

/* The same two calls for a Cell of any CellType (cell_type = DMDX_CELL_*):
 * the Cell(matrix, type) constructor (src/core/cell.h:12) and the
 * std::span/Cell signatures of rebuild_pair_list / compute
 * (force_component.h:15-16).  The matrix diagonal must be positive.  Cells
 * other than orthorhombic run on the FP64 path only (DMDX_MIXED keeps
 * cell-local frames that assume a wrapped box) and return DMDX_E_PARM under
 * DMDX_MIXED. */
int dmdx_component_rebuild_pair_list_cell(dmdx_component *c, const double *x, const double *y,
                                          const double *z, int64_t n, const double cell[9],
                                          int32_t cell_type);
int dmdx_component_compute_cell(dmdx_component *c, const double *x, const double *y,
                                const double *z, int64_t n, const double cell[9],
                                int32_t cell_type, double *fx, double *fy, double *fz,
                                double *energy_out);

/* Allocation-free ForceComponent::compute.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  x, y, z,
 * fx, fy, fz and *energy_out are DEVICE arrays (energy_out: one double);
 * forces and energy accumulate as in dmdx_component_compute, with the same
 * result bit for bit.  Everything else the kernels need lives in one
 * file-scope device array of DMDX_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends; no hipMalloc, hipFree or other device allocation is
 * made.  A system that does not fit returns DMDX_E_OOM.  Calls on one GPU
 * share the array: calls that overlap on one GPU (several host threads)
 * race.  Stateless components only (bond, angle, dihedral, improper,
 * exclusion, PME); LennardJones and CoulombDirect keep a pair list between
 * calls and return DMDX_E_PARM.  Returns after the device work is done. */
#define DMDX_NOALLOC_SCRATCH_BYTES (64ull << 20)
int dmdx_component_compute_noalloc(dmdx_component *c, const double *x, const double *y,
                                   const double *z, int64_t n, const double cell[9], double *fx,
                                   double *fy, double *fz, double *energy_out);

/* ForceComponent::name(): "LennardJones", "CoulombDirect", ... */
int dmdx_component_name(dmdx_component *c, char *buf, int64_t cap);

/* DMDX_FAITHFUL (default) | DMDX_CORRECTED, DMDX_FP64 (default) | DMDX_MIXED */
int dmdx_component_set_mode(dmdx_component *c, int32_t physics, int32_t precision);
int dmdx_component_destroy(dmdx_component *c);

/* ---- measurement (no counterpart in the reference) -------------------- */

/* Bracket every kernel group with HIP events on the engine's stream (eager
 * launches only).  Groups: 0 pair forces, 1 list build, 2 bonded terms,
 * 3 PME, 4 force gather, 5 kick/drift.  dmdx_engine_kernel_time waits for
 * the stream and returns the accumulated time and launch count of a group. */
int dmdx_engine_set_profiling(dmdx_engine *e, int32_t on);
/* Bracket only the groups whose bit (1 << group) is set in mask (0: off);
 * fewer events per step leave fewer gaps between the untimed kernels. */
int dmdx_engine_set_profiling_groups(dmdx_engine *e, uint32_t mask);
int dmdx_engine_kernel_time(dmdx_engine *e, int32_t group, double *total_ms, int64_t *launches);

/* Launch shape of the mixed-precision kernels (no counterpart in the
 * reference): key 0 = lanes per atom of the pair kernel (2, 4, 8, 16 or 32),
 * key 1 = lanes per atom of the list build (8, 16 or 32), key 2 = list
 * entries per lane per pass of the pair kernel (2, 4 or 8), key 3 = list
 * build per cell (1, default: the atoms of a cell share their candidates;
 * systems that need a minimum image are still built per atom) or per atom
 * (0), key 4 = waves per cell of the per-cell build (1, 2, 4 or 8), key 5
 * = whether dmdx_engine_autotune may change list skins, cells and the FFT
 * tiling (1, default, or 0),
 * key 6 = LJ-only systems' pair kernel: 4 or 8 = row tiles staged in LDS
 * with that many list entries per lane block, 0 (default) = the per-atom
 * kernel of keys 0 and 2 (charged systems always use the per-atom kernel),
 * key 7 = the per-cell build's kernel: 8 or 16 = each cell's candidates
 * staged in LDS and scanned by that many lanes per atom, 0 (default) =
 * candidates shared in registers by the waves of key 4.
 * Key 8 (read only) = the FFT lines per workgroup tile the autotuner chose
 * for PME (4, 8, 16; 0 = the grid-size rule).  Key 9 (write only, before
 * the first step; the batch is made again) = cells per list radius of every
 * mixed system with a list: 1, 2, or 0 for the density rule (read per system
 * with dmdx_engine_get_cell_div).  Keys 10 and 11 = lanes per atom and
 * entries per lane pass of the pair kernel for the systems with real-space
 * Coulomb pairs (default 8 x 2: their rows are several times longer than
 * LJ rows); keys 0 and 2 then apply to the LJ-only systems.  Key 12
 * (before the first step; the batch is made again) = the mixed PME
 * transforms' x and y passes as one xy plane per workgroup: 0 never, 1
 * wherever the plane fits in 64 KiB of LDS, 2 (default) when the batch has
 * at least 256 such planes; key 13 (read only) = 1 if some system's
 * transforms run on planes.  Planes give the same grid bit for bit.
 * The list rows do not depend on keys 1, 3, 4 or 7.  Pair shapes
 * (lanes x entries): 2x8, 4x8, 8x8, 4x4, 8x4, 16x4, 32x4, 8x2, 16x2, 32x2;
 * another combination runs as the nearest of these (8 lanes where fewer
 * are not built for the width, at most 8 lanes with 8 entries).  Results do
 * not depend on the shape beyond FP32 summation order; a row-tile system's
 * results do not depend on keys 0 and 2 at all. */
int dmdx_engine_set_tuning(dmdx_engine *e, int32_t key, int32_t value);
/* The current value of a tuning key (as dmdx_engine_set_tuning numbers them). */
int dmdx_engine_get_tuning(dmdx_engine *e, int32_t key, int32_t *value);

/* Autotuner (no counterpart in the reference): times launch shapes of the
 * mixed pair kernel (lanes per atom x list entries per lane pass) and of the
 * list build (8/16/32 lanes per atom, or per cell with 2/4/8 waves per
 * cell) on a scratch copy of the engine's
 * systems stepped from their initial state (the engine's own state is
 * untouched), `steps` steps per candidate, and applies the fastest of each
 * (a pair shape replaces the engine's current one only if at least 2%
 * faster: near ties in a few steps are noise).
 * cache_path (may be NULL): text file of earlier results keyed by a
 * signature of the systems; a hit applies the cached choice without timing.
 * times_ms (DMDX_TUNE_TIMES doubles, may be NULL): per-step ms of the pair
 * kernels with the LJ-only systems' shape (keys 0, 2) at 2x8, 4x8, 8x8,
 * 4x4, 8x4, 16x4, 8x2, 16x2, then with LJ-only systems on row tiles of 4
 * and 8 entries per block (tuning key 6) beside the fastest of those
 * shapes, then per-step ms of the list build at 8, 16, 32 lanes per atom,
 * per cell at 2, 4, 8 waves per cell and per cell staged in LDS at 8, 16
 * lanes per atom, then of the pair kernels with the Coulomb systems' shape
 * (keys 10, 11) at the eight shapes, timed first, with the LJ-only shape
 * unchanged (-1 when not timed, e.g. for a variant no system uses);
 * dmdx_engine_get_tuning reads the chosen Coulomb shape, row tiles and
 * build (keys 10, 11, 6, 1, 3, 4, 7).  FP64 engines have nothing to tune.
 * First, on an engine that has not stepped or been changed since creation
 * (and unless tuning key 5 is 0): the list skin of systems that rebuild by
 * the displacement criterion (rebuild_interval -1), scaled by 1, 1.3, 1.6 or
 * 2 per group of systems with the same cutoffs and skin, then each group's
 * cells per list radius (1 or 2, against the density rule), timed on scratch
 * batches (20 steps from the start and 20 after 60 more); a change must win
 * by 2%.  Then, for a batch with PME, the FFT lines per workgroup tile (the
 * grid-size rule, 4, 8 or 16) timed on the PME kernels.  The engine's batch
 * is then made again with the choices (dmdx_engine_get_list_skin,
 * dmdx_engine_get_cell_div and tuning key 8 read them).
 * Mixed-precision results depend on the shape, the skin, the cells and the
 * FFT tiling only through FP32 summation order (well inside the
 * mixed-precision tolerance). */
#define DMDX_TUNE_TIMES 26
int dmdx_engine_autotune(dmdx_engine *e, int32_t steps, const char *cache_path,
                         int32_t *pair_lanes, int32_t *pair_width, int32_t *build_lanes,
                         double *times_ms);
/* The list skin system sys is built with (its description's, or the
 * autotuner's choice). */
int dmdx_engine_get_list_skin(dmdx_engine *e, int32_t sys, double *skin);
/* Cells per list radius along an edge of system `sys`'s list grid (1, or 2
 * for cells of half the list radius searched +-2): the density rule's, or
 * the autotuner's choice.  No counterpart in the reference. */
int dmdx_engine_get_cell_div(dmdx_engine *e, int32_t sys, int32_t *cdiv);

/* Device-to-device copy of nbytes, repeated; reports the achieved GB/s
 * (read + write).  Calibrates hardware byte counters. */
int dmdx_copy_bandwidth(int64_t nbytes, int32_t repeats, double *gbps);

/* ---- serving many simulations on one device ----------------------------
 * Replaces running many SimulationEngine::run loops side by side
 * (src/sim/simulation_engine.cpp:89-106, driven per process by
 * python/dmd/runner.py:44-58).  A request is a system plus a step count.
 * Waiting requests are admitted when their estimated device memory fits in
 * the budget and grouped into cohorts: one batch per cohort, each on its own
 * stream, several cohorts stepped concurrently; after each pump the
 * cohorts are merged into one batch (each system moves with its whole
 * state, so its trajectory does not change), which keeps the batches large
 * as requests come and go.  A request leaves its batch exactly at its step
 * target, or when it fails (neighbour-list overflow, non-finite energy) --
 * the other systems are not disturbed, and a request's result is the same
 * alone or in any mix.  One host thread drives
 * a server: submit at any time, then call dmdx_server_pump repeatedly. */
typedef struct dmdx_server dmdx_server;

typedef struct dmdx_server_options {
    int64_t hbm_budget_bytes;      /* device memory the server may hold;
                                      0 = 90% of the free memory at creation */
    int32_t max_cohorts;           /* cohorts stepped concurrently (0: 8)     */
    int32_t chunk_steps;           /* steps per cohort per pump (0: 100)      */
    int32_t graph_steps;           /* > 0: replay captured blocks of steps    */
    int32_t min_admit;             /* form a cohort once this many requests
                                      wait (0: 1) ...                          */
    double max_wait_ms;            /* ... or once the oldest waited this long */
    int32_t max_cohort_systems;    /* systems per cohort (0: no limit)        */
    int32_t keep_cohorts;          /* 0: after each pump, merge the cohorts of
                                      one precision into one batch (systems
                                      move with their whole state, results
                                      unchanged); 1: never merge             */
} dmdx_server_options;

enum {
    DMDX_REQ_QUEUED = 0,
    DMDX_REQ_RUNNING = 1,
    DMDX_REQ_DONE = 2,
    DMDX_REQ_FAILED = 3,
    DMDX_REQ_CANCELLED = 4
};

typedef struct dmdx_request_info {
    int32_t state;                 /* DMDX_REQ_*                              */
    int32_t flags;                 /* as dmdx_scalars::flags                  */
    int64_t n_atoms;
    int64_t steps_target, steps_done;
    double t_submit, t_admit, t_done;  /* seconds on the server clock (-1: not yet) */
    int64_t est_bytes;             /* admission estimate of device memory     */
    char error[128];               /* why a FAILED request failed             */
} dmdx_request_info;

typedef struct dmdx_server_stats {
    int64_t queued, running, retired, failed;
    int64_t cohorts, cohorts_made, rounds, merges;
    int64_t bytes_in_use, bytes_budget;
    double estimate_scale;         /* learned actual/estimate memory ratio    */
    double clock_s;                /* server clock now                        */
    int64_t atom_steps;            /* atoms x steps advanced so far           */
    int64_t cancelled;             /* requests cancelled (not in retired)     */
} dmdx_server_stats;

int dmdx_server_create(const dmdx_server_options *opt, dmdx_server **out);
int dmdx_server_destroy(dmdx_server *srv);
/* copies desc; returns the request id */
int dmdx_server_submit(dmdx_server *srv, const dmdx_system_desc *desc, int64_t n_steps,
                       int64_t *id);
/* one round: admit, advance every cohort by up to chunk_steps, wait, retire */
int dmdx_server_pump(dmdx_server *srv, int32_t *n_retired);
/* Stop a request.  A queued request leaves the queue.  A running request
 * takes no further step: its state at the end of the last pump (positions,
 * velocities, forces, scalars, steps_done) is kept as its result, as for a
 * finished request, and the next pump's merge or compaction of its batch
 * returns its device memory; the other systems of the batch are untouched.
 * DMDX_E_INVALID for a request that has already finished.  The reference
 * stops a run only by ending its loop (src/sim/simulation_engine.cpp:89-106). */
int dmdx_server_cancel(dmdx_server *srv, int64_t id);
int dmdx_server_request_info(dmdx_server *srv, int64_t id, dmdx_request_info *out);
/* final state of a finished request: what 0 positions, 1 velocities, 2 forces */
int dmdx_server_result(dmdx_server *srv, int64_t id, int32_t what, double *x, double *y,
                       double *z);
int dmdx_server_result_scalars(dmdx_server *srv, int64_t id, dmdx_scalars *out);
/* forget a finished request and its results */
int dmdx_server_release(dmdx_server *srv, int64_t id);
int dmdx_server_get_stats(dmdx_server *srv, dmdx_server_stats *out);
/* per-step latency samples: each pump contributes, per cohort, the round's
 * wall time divided by the steps it advanced (ms), weighted by the steps the
 * requests advanced; count = number of samples (copied up to cap) */
int dmdx_server_step_latency(dmdx_server *srv, double *ms, int64_t *weight, int64_t cap,
                             int64_t *count);

#ifdef __cplusplus
}
#endif

#endif /* DMDX_H */
