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
    DMDX_E_BUSY = -8            /* server: request cannot be admitted now     */
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

/* ---- electrostatics (SimulationConfig::coulomb_type,
 *      src/sim/simulation_config.h:52-56) */
enum { DMDX_COULOMB_NONE = 0, DMDX_COULOMB_CUTOFF = 1, DMDX_COULOMB_PME = 2 };

/* ---- coupling (ThermostatConfig / BarostatConfig,
 *      src/sim/simulation_config.h:13-25) */
enum { DMDX_THERMO_NONE = 0, DMDX_THERMO_BERENDSEN = 1, DMDX_THERMO_NOSE_HOOVER = 2 };
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
    int32_t rebuild_interval;      /* computes between rebuilds (DMD: 10)     */
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
} dmdx_system_desc;

#ifdef __cplusplus
}
#endif

#endif /* DMDX_H */
