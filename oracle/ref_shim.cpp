// ref_shim.cpp -- extern "C" entry points over DMD's own translation units.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/build_ref.sh together with
// the reference sources, where they lie, into oracle/_ref/libdmdref.so.  It is
// used to pin the oracle bit-for-bit (tests/test_oracle_vs_ref.py) and can be
// timed as the "reference" CPU baseline for PME-free workloads.  The shim
// holds no physics: each function builds the reference objects and calls
// their methods.  src/force/coulomb_pme.cpp (needs fftw3.h) and
// src/sim/simulation_engine.cpp (needs fftw3.h and hdf5.h through its
// includes) cannot be compiled in this image and are not used; ref_sim_*
// therefore drives ForceEngine/Integrator/thermostats in the order of
// SimulationEngine::step (src/sim/simulation_engine.cpp:53-86).
#include "core/cell.h"
#include "core/system_data.h"
#include "force/force_engine.h"
#include "force/lennard_jones.h"
#include "force/coulomb_direct.h"
#include "force/coulomb_exclusion.h"
#include "force/harmonic_bond.h"
#include "force/harmonic_angle.h"
#include "force/periodic_dihedral.h"
#include "force/harmonic_improper.h"
#include "integrate/integrator.h"
#include "integrate/constraints.h"
#include "thermostat/thermostat.h"
#include "thermostat/andersen.h"
#include "thermostat/berendsen.h"
#include "thermostat/nose_hoover.h"
#include "barostat/barostat.h"
#include "barostat/berendsen_barostat.h"
#include "barostat/andersen_barostat.h"
#include "sim/checkpoint.h"

#include <cstdint>
#include <memory>
#include <span>
#include <vector>

namespace {

SystemData make_sys(int64_t n, const double* x, const double* y, const double* z)
{
    SystemData s(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
        s.pos_x[i] = x[i];
        s.pos_y[i] = y[i];
        s.pos_z[i] = z[i];
    }
    return s;
}

void add_forces(const SystemData& s, double* fx, double* fy, double* fz)
{
    for (size_t i = 0; i < s.n_atoms; ++i) {
        fx[i] += s.forces_x[i];
        fy[i] += s.forces_y[i];
        fz[i] += s.forces_z[i];
    }
}

template <class T>
std::vector<T> vec(const T* p, int64_t n)
{
    return std::vector<T>(p, p + n);
}

// Cell type of the cells built below (set by ref_set_cell_type): DMD's own
// Cell(matrix, type) with the box on the diagonal.
CellType g_cell_type = CellType::orthorhombic;

Cell make_cell(const double box[3])
{
    if (g_cell_type == CellType::orthorhombic) return Cell(box[0], box[1], box[2]);
    const double m[9] = {box[0], 0.0, 0.0, 0.0, box[1], 0.0, 0.0, 0.0, box[2]};
    return Cell(m, g_cell_type);
}

}  // namespace

extern "C" {

void ref_min_image(const double cell[9], double* dx, double* dy, double* dz)
{
    Cell c = g_cell_type == CellType::orthorhombic ? Cell(cell[0], cell[4], cell[8])
                                                    : Cell(cell, g_cell_type);
    c.minimum_image(*dx, *dy, *dz, *dx, *dy, *dz);
}

// CellType of the cells the functions of this file build (0 orthorhombic,
// 1 triclinic, 2 hexagonal, 3 rhombic dodecahedron; src/core/cell.h:3-8)
void ref_set_cell_type(int type) { g_cell_type = static_cast<CellType>(type); }

// Cell::volume of a full 3x3 matrix (src/core/cell.cpp:19-24)
double ref_cell_volume(const double m[9], int type)
{
    return Cell(m, static_cast<CellType>(type)).volume();
}

// LennardJones: rebuild (as the first compute does) then compute `ncompute`
// times on the same positions (forces accumulate, as the reference does).
void ref_lj(int64_t n, const double* x, const double* y, const double* z, const double box[3],
            double cutoff, int32_t n_types, const double* sigma, const double* epsilon,
            const int32_t* types, double* fx, double* fy, double* fz, double* energy)
{
    LennardJonesParams p;
    p.n_types = static_cast<size_t>(n_types);
    p.sigma = vec(sigma, int64_t(n_types) * n_types);
    p.epsilon = vec(epsilon, int64_t(n_types) * n_types);
    LennardJones lj(std::move(p), cutoff);
    SystemData s = make_sys(n, x, y, z);
    if (types)
        for (int64_t i = 0; i < n; ++i) s.atom_types[i] = types[i];
    lj.set_atom_types(s.atom_types);
    Cell cell = make_cell(box);
    lj.compute(s.x(), s.y(), s.z(), cell, s.fx(), s.fy(), s.fz(), *energy);
    add_forces(s, fx, fy, fz);
}

void ref_coulomb_direct(int64_t n, const double* x, const double* y, const double* z,
                        const double box[3], double cutoff, double alpha, const double* q,
                        int rebuild, double* fx, double* fy, double* fz, double* energy)
{
    CoulombDirectParams p;
    p.cutoff = cutoff;
    p.ewald_coefficient = alpha;
    CoulombDirect cd(p);
    SystemData s = make_sys(n, x, y, z);
    cd.set_charges(std::span<const double>(q, size_t(n)));
    Cell cell = make_cell(box);
    if (rebuild) cd.rebuild_pair_list(s.x(), s.y(), s.z(), cell);
    cd.compute(s.x(), s.y(), s.z(), cell, s.fx(), s.fy(), s.fz(), *energy);
    add_forces(s, fx, fy, fz);
}

void ref_exclusion(int64_t n, const double* x, const double* y, const double* z,
                   const double box[3], const double* q, int64_t ne, const int32_t* ei,
                   const int32_t* ej, double* fx, double* fy, double* fz, double* energy)
{
    CoulombExclusion ce;
    SystemData s = make_sys(n, x, y, z);
    ce.set_charges(std::span<const double>(q, size_t(n)));
    std::vector<int> vi(ei, ei + ne), vj(ej, ej + ne);
    ce.set_exclusions(vi, vj);
    Cell cell = make_cell(box);
    ce.compute(s.x(), s.y(), s.z(), cell, s.fx(), s.fy(), s.fz(), *energy);
    add_forces(s, fx, fy, fz);
}

void ref_bonds(int64_t n, const double* x, const double* y, const double* z, const double box[3],
               int64_t nb, const int32_t* bi, const int32_t* bj, const double* k,
               const double* r0, double* fx, double* fy, double* fz, double* energy)
{
    HarmonicBondParams p;
    p.n = size_t(nb);
    p.i.assign(bi, bi + nb);
    p.j.assign(bj, bj + nb);
    p.k = vec(k, nb);
    p.r0 = vec(r0, nb);
    HarmonicBond b(std::move(p));
    SystemData s = make_sys(n, x, y, z);
    Cell cell = make_cell(box);
    b.compute(s.x(), s.y(), s.z(), cell, s.fx(), s.fy(), s.fz(), *energy);
    add_forces(s, fx, fy, fz);
}

void ref_angles(int64_t n, const double* x, const double* y, const double* z, const double box[3],
                int64_t na, const int32_t* ai, const int32_t* aj, const int32_t* ak,
                const double* kt, const double* t0, double* fx, double* fy, double* fz,
                double* energy)
{
    HarmonicAngleParams p;
    p.n = size_t(na);
    p.i.assign(ai, ai + na);
    p.j.assign(aj, aj + na);
    p.k.assign(ak, ak + na);
    p.k_theta = vec(kt, na);
    p.theta0 = vec(t0, na);
    HarmonicAngle a(std::move(p));
    SystemData s = make_sys(n, x, y, z);
    Cell cell = make_cell(box);
    a.compute(s.x(), s.y(), s.z(), cell, s.fx(), s.fy(), s.fz(), *energy);
    add_forces(s, fx, fy, fz);
}

void ref_dihedrals(int64_t n, const double* x, const double* y, const double* z,
                   const double box[3], int64_t nd, const int32_t* di, const int32_t* dj,
                   const int32_t* dk, const int32_t* dl, const int32_t* per, const double* kp,
                   const double* p0, double* fx, double* fy, double* fz, double* energy)
{
    PeriodicDihedralParams p;
    p.n = size_t(nd);
    p.i.assign(di, di + nd);
    p.j.assign(dj, dj + nd);
    p.k.assign(dk, dk + nd);
    p.l.assign(dl, dl + nd);
    p.periodicity.assign(per, per + nd);
    p.k_phi = vec(kp, nd);
    p.phi0 = vec(p0, nd);
    PeriodicDihedral d(std::move(p));
    SystemData s = make_sys(n, x, y, z);
    Cell cell = make_cell(box);
    d.compute(s.x(), s.y(), s.z(), cell, s.fx(), s.fy(), s.fz(), *energy);
    add_forces(s, fx, fy, fz);
}

void ref_impropers(int64_t n, const double* x, const double* y, const double* z,
                   const double box[3], int64_t nd, const int32_t* di, const int32_t* dj,
                   const int32_t* dk, const int32_t* dl, const double* kp, const double* p0,
                   double* fx, double* fy, double* fz, double* energy)
{
    ImproperParams p;
    p.n = size_t(nd);
    p.i.assign(di, di + nd);
    p.j.assign(dj, dj + nd);
    p.k.assign(dk, dk + nd);
    p.l.assign(dl, dl + nd);
    p.k_phi = vec(kp, nd);
    p.phi0 = vec(p0, nd);
    HarmonicImproper d(std::move(p));
    SystemData s = make_sys(n, x, y, z);
    Cell cell = make_cell(box);
    d.compute(s.x(), s.y(), s.z(), cell, s.fx(), s.fy(), s.fz(), *energy);
    add_forces(s, fx, fy, fz);
}

double ref_kinetic_energy(int64_t n, const double* m, const double* vx, const double* vy,
                          const double* vz)
{
    SystemData s(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
        s.masses[i] = m[i];
        s.vel_x[i] = vx[i];
        s.vel_y[i] = vy[i];
        s.vel_z[i] = vz[i];
    }
    return kinetic_energy(s);
}

double ref_temperature(double ekin, int64_t n) { return temperature(ekin, size_t(n)); }

double ref_virial(int64_t n, const double* x, const double* y, const double* z, const double* fx,
                  const double* fy, const double* fz)
{
    SystemData s = make_sys(n, x, y, z);
    for (int64_t i = 0; i < n; ++i) {
        s.forces_x[i] = fx[i];
        s.forces_y[i] = fy[i];
        s.forces_z[i] = fz[i];
    }
    return virial(s);
}

double ref_pressure(double ekin, double w, double volume) { return pressure(ekin, w, volume); }

// Integrator / thermostat / barostat on a SystemData filled from the arrays;
// updated arrays are written back.
static SystemData fill_state(int64_t n, const double* m, const double* x, const double* y,
                             const double* z, const double* vx, const double* vy,
                             const double* vz, const double* fx, const double* fy,
                             const double* fz)
{
    SystemData s = make_sys(n, x, y, z);
    for (int64_t i = 0; i < n; ++i) {
        s.masses[i] = m[i];
        s.vel_x[i] = vx[i];
        s.vel_y[i] = vy[i];
        s.vel_z[i] = vz[i];
        s.forces_x[i] = fx[i];
        s.forces_y[i] = fy[i];
        s.forces_z[i] = fz[i];
    }
    return s;
}

static void back(const SystemData& s, double* x, double* y, double* z, double* vx, double* vy,
                 double* vz)
{
    for (size_t i = 0; i < s.n_atoms; ++i) {
        x[i] = s.pos_x[i];
        y[i] = s.pos_y[i];
        z[i] = s.pos_z[i];
        vx[i] = s.vel_x[i];
        vy[i] = s.vel_y[i];
        vz[i] = s.vel_z[i];
    }
}

// kind: 0 half_kick, 1 advance, 2 Berendsen thermostat, 3 Nose-Hoover
// thermostat (state[0] = zeta in/out), 4 Berendsen barostat, 5 Andersen
// barostat (state[0] = eps_dot in/out).  box[0..2] in/out.
void ref_apply(int kind, int64_t n, const double* m, double* x, double* y, double* z, double* vx,
               double* vy, double* vz, const double* fx, const double* fy, const double* fz,
               double box[3], double dt, double target, double tau, double beta, double* state)
{
    SystemData s = fill_state(n, m, x, y, z, vx, vy, vz, fx, fy, fz);
    Cell cell = make_cell(box);
    Integrator integ;
    switch (kind) {
    case 0: integ.half_kick(s, dt); break;
    case 1: integ.advance(s, dt); break;
    case 2: {
        BerendsenParams p;
        p.temperature = target;
        p.tau = tau;
        BerendsenThermostat t(p);
        t.apply(s, dt);
        break;
    }
    case 3: {
        NoseHooverParams p;
        p.temperature = target;
        p.tau = tau;
        NoseHooverThermostat t(p, size_t(n));
        t.set_zeta(state[0]);
        t.apply(s, dt);
        state[0] = t.zeta();
        break;
    }
    case 4: {
        BerendsenBarostatParams p;
        p.pressure = target;
        p.tau = tau;
        p.compressibility = beta;
        BerendsenBarostat b(p);
        b.apply(s, cell, dt);
        break;
    }
    case 5: {
        AndersenBarostatParams p;
        p.pressure = target;
        p.tau = tau;
        AndersenBarostat b(p, size_t(n));
        b.set_piston_velocity(state[0]);
        b.apply(s, cell, dt);
        state[0] = b.piston_velocity();
        break;
    }
    default: break;
    }
    back(s, x, y, z, vx, vy, vz);
    box[0] = cell.matrix[0];
    box[1] = cell.matrix[4];
    box[2] = cell.matrix[8];
}

// SHAKE (kind 0, positions) or RATTLE (kind 1, velocities) of DMD's
// Constraints on one state; box is the orthorhombic cell.
void ref_constraints(int kind, int64_t nc, const int32_t* ci, const int32_t* cj,
                     const double* dist, double tol, int64_t n, const double* m, double* x,
                     double* y, double* z, double* vx, double* vy, double* vz,
                     const double box[3])
{
    std::vector<double> zero(size_t(n), 0.0);
    SystemData s = fill_state(n, m, x, y, z, vx, vy, vz, zero.data(), zero.data(), zero.data());
    Cell cell = make_cell(box);
    ConstraintsParams p;
    p.i.assign(ci, ci + nc);
    p.j.assign(cj, cj + nc);
    p.distance_target.assign(dist, dist + nc);
    p.tolerance = tol;
    Constraints c(std::move(p));
    if (kind == 0) c.apply(s, cell);
    else c.apply_rattle(s, cell);
    back(s, x, y, z, vx, vy, vz);
}

// DMD's AndersenThermostat: `calls` apply() calls on a scratch copy (each
// advances the member seed), then one on the state.
void ref_andersen(int64_t n, const double* m, double* vx, double* vy, double* vz, double dt,
                  double T0, double frequency, int64_t calls)
{
    std::vector<double> zero(size_t(n), 0.0);
    SystemData s = fill_state(n, m, zero.data(), zero.data(), zero.data(), vx, vy, vz,
                              zero.data(), zero.data(), zero.data());
    AndersenParams p;
    p.temperature = T0;
    p.frequency = frequency;
    AndersenThermostat t(p);
    for (int64_t k = 0; k < calls; ++k) {
        SystemData scratch = s;
        t.apply(scratch, dt);
    }
    t.apply(s, dt);
    for (int64_t i = 0; i < n; ++i) {
        vx[i] = s.vel_x[i];
        vy[i] = s.vel_y[i];
        vz[i] = s.vel_z[i];
    }
}

// ---- step-level driver: ForceEngine + Integrator + thermostat/barostat in
// the order of SimulationEngine::step.  Components are added in
// build_force_engine order (LJ, bonds, angles, dihedrals, impropers,
// CoulombDirect, exclusion); PME is not available.
struct ref_sim {
    ForceEngine fe;
    Integrator integ;
    std::unique_ptr<Thermostat> thermo;
    std::unique_ptr<Barostat> baro;
    Cell cell;
    SystemData sys;
    double dt = 0.0;
    LennardJones* lj = nullptr;
    std::unique_ptr<Constraints> con;
};

ref_sim* ref_sim_create(int64_t n, const double* x, const double* y, const double* z,
                        const double* vx, const double* vy, const double* vz, const double* m,
                        const double* q, const int32_t* types, const double box[3], int use_lj,
                        int32_t n_types, const double* sigma, const double* epsilon,
                        double lj_cutoff, int64_t nb, const int32_t* bi, const int32_t* bj,
                        const double* bk, const double* br0, int64_t na, const int32_t* ai,
                        const int32_t* aj, const int32_t* ak, const double* akt,
                        const double* at0, int64_t nd, const int32_t* di, const int32_t* dj,
                        const int32_t* dk, const int32_t* dl, const int32_t* dper,
                        const double* dkp, const double* dp0, int64_t ne, const int32_t* ei,
                        const int32_t* ej, int coulomb_cutoff_mode, double coulomb_cutoff,
                        double alpha, double dt, int thermostat, double T0, double ttau,
                        int barostat, double P0, double ptau, double beta)
{
    auto* r = new ref_sim();
    r->cell = Cell(box[0], box[1], box[2]);
    r->sys = SystemData(size_t(n));
    r->dt = dt;
    for (int64_t i = 0; i < n; ++i) {
        r->sys.pos_x[i] = x[i];
        r->sys.pos_y[i] = y[i];
        r->sys.pos_z[i] = z[i];
        r->sys.vel_x[i] = vx ? vx[i] : 0.0;
        r->sys.vel_y[i] = vy ? vy[i] : 0.0;
        r->sys.vel_z[i] = vz ? vz[i] : 0.0;
        r->sys.masses[i] = m[i];
        r->sys.charges[i] = q ? q[i] : 0.0;
        r->sys.atom_types[i] = types ? types[i] : 0;
    }
    if (use_lj) {
        LennardJonesParams p;
        p.n_types = size_t(n_types);
        p.sigma = vec(sigma, int64_t(n_types) * n_types);
        p.epsilon = vec(epsilon, int64_t(n_types) * n_types);
        auto lj = std::make_unique<LennardJones>(std::move(p), lj_cutoff);
        lj->set_atom_types(r->sys.atom_types);
        r->lj = lj.get();
        r->fe.add_component(std::move(lj));
    }
    if (nb > 0) {
        HarmonicBondParams p;
        p.n = size_t(nb);
        p.i.assign(bi, bi + nb);
        p.j.assign(bj, bj + nb);
        p.k = vec(bk, nb);
        p.r0 = vec(br0, nb);
        r->fe.add_component(std::make_unique<HarmonicBond>(std::move(p)));
    }
    if (na > 0) {
        HarmonicAngleParams p;
        p.n = size_t(na);
        p.i.assign(ai, ai + na);
        p.j.assign(aj, aj + na);
        p.k.assign(ak, ak + na);
        p.k_theta = vec(akt, na);
        p.theta0 = vec(at0, na);
        r->fe.add_component(std::make_unique<HarmonicAngle>(std::move(p)));
    }
    if (nd > 0) {
        PeriodicDihedralParams p;
        p.n = size_t(nd);
        p.i.assign(di, di + nd);
        p.j.assign(dj, dj + nd);
        p.k.assign(dk, dk + nd);
        p.l.assign(dl, dl + nd);
        p.periodicity.assign(dper, dper + nd);
        p.k_phi = vec(dkp, nd);
        p.phi0 = vec(dp0, nd);
        r->fe.add_component(std::make_unique<PeriodicDihedral>(std::move(p)));
    }
    if (coulomb_cutoff_mode) {
        CoulombDirectParams p;
        p.cutoff = coulomb_cutoff;
        p.ewald_coefficient = alpha;
        auto cd = std::make_unique<CoulombDirect>(p);
        cd->set_charges(r->sys.charges);
        r->fe.add_component(std::move(cd));
        if (ne > 0) {
            auto ce = std::make_unique<CoulombExclusion>();
            std::vector<int> vi(ei, ei + ne), vj(ej, ej + ne);
            ce->set_exclusions(vi, vj);
            ce->set_charges(r->sys.charges);
            r->fe.add_component(std::move(ce));
        }
    }
    if (thermostat == 1) {
        BerendsenParams p;
        p.temperature = T0;
        p.tau = ttau;
        r->thermo = std::make_unique<BerendsenThermostat>(p);
    } else if (thermostat == 2) {
        NoseHooverParams p;
        p.temperature = T0;
        p.tau = ttau;
        r->thermo = std::make_unique<NoseHooverThermostat>(p, size_t(n));
    }
    if (barostat == 1) {
        BerendsenBarostatParams p;
        p.pressure = P0;
        p.tau = ptau;
        p.compressibility = beta;
        r->baro = std::make_unique<BerendsenBarostat>(p);
    } else if (barostat == 2) {
        AndersenBarostatParams p;
        p.pressure = P0;
        p.tau = ptau;
        r->baro = std::make_unique<AndersenBarostat>(p, size_t(n));
    }
    return r;
}

// constraint_type "shake" (src/sim/simulation_engine.cpp:293-301)
void ref_sim_set_constraints(ref_sim* r, int64_t nc, const int32_t* ci, const int32_t* cj,
                             const double* dist, double tol)
{
    ConstraintsParams p;
    p.i.assign(ci, ci + nc);
    p.j.assign(cj, cj + nc);
    p.distance_target.assign(dist, dist + nc);
    p.tolerance = tol;
    r->con = std::make_unique<Constraints>(std::move(p));
}

// thermostat type "andersen" (src/sim/simulation_engine.cpp:220-225)
void ref_sim_set_andersen(ref_sim* r, double T0, double frequency)
{
    AndersenParams p;
    p.temperature = T0;
    p.frequency = frequency;
    r->thermo = std::make_unique<AndersenThermostat>(p);
}

void ref_sim_step(ref_sim* r, int64_t nsteps)
{
    auto& s = r->sys;
    for (int64_t it = 0; it < nsteps; ++it) {
        s.forces_x.assign(s.n_atoms, 0.0);
        s.forces_y.assign(s.n_atoms, 0.0);
        s.forces_z.assign(s.n_atoms, 0.0);
        s.potential_energy = 0.0;
        r->fe.compute(s, r->cell);
        r->integ.half_kick(s, r->dt);
        r->integ.advance(s, r->dt);
        if (r->con) r->con->apply(s, r->cell);
        s.forces_x.assign(s.n_atoms, 0.0);
        s.forces_y.assign(s.n_atoms, 0.0);
        s.forces_z.assign(s.n_atoms, 0.0);
        s.potential_energy = 0.0;
        r->fe.compute(s, r->cell);
        r->integ.half_kick(s, r->dt);
        if (r->con) r->con->apply_rattle(s, r->cell);
        if (r->thermo) r->thermo->apply(s, r->dt);
        if (r->baro) r->baro->apply(s, r->cell, r->dt);
    }
}

void ref_sim_get(const ref_sim* r, double* x, double* y, double* z, double* vx, double* vy,
                 double* vz, double* fx, double* fy, double* fz, double* scalars)
{
    const auto& s = r->sys;
    for (size_t i = 0; i < s.n_atoms; ++i) {
        x[i] = s.pos_x[i];
        y[i] = s.pos_y[i];
        z[i] = s.pos_z[i];
        vx[i] = s.vel_x[i];
        vy[i] = s.vel_y[i];
        vz[i] = s.vel_z[i];
        fx[i] = s.forces_x[i];
        fy[i] = s.forces_y[i];
        fz[i] = s.forces_z[i];
    }
    scalars[0] = s.potential_energy;
    scalars[1] = kinetic_energy(s);
    scalars[2] = r->cell.matrix[0];
    scalars[3] = r->lj ? double(r->lj->step_since_rebuild()) : -1.0;
}

void ref_sim_destroy(ref_sim* r) { delete r; }

// checkpoint::save of a SystemData filled from the arrays: v holds the 11
// FP64 vectors in SystemData order (pos xyz, vel xyz, forces xyz, masses,
// charges), n each
void ref_checkpoint_save(const char* path, int64_t n, int64_t step, double time, double pe,
                         const double* v, const int32_t* types, int32_t lj_counter)
{
    SystemData s(static_cast<size_t>(n));
    std::vector<double>* dst[11] = {&s.pos_x,    &s.pos_y,    &s.pos_z,    &s.vel_x,
                                    &s.vel_y,    &s.vel_z,    &s.forces_x, &s.forces_y,
                                    &s.forces_z, &s.masses,   &s.charges};
    for (int k = 0; k < 11; ++k) dst[k]->assign(v + k * n, v + (k + 1) * n);
    s.atom_types.assign(types, types + n);
    s.step = static_cast<size_t>(step);
    s.time = time;
    s.potential_energy = pe;
    checkpoint::save(path, s, lj_counter);
}

}  // extern "C"
