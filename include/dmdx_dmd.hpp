// dmdx_dmd.hpp -- C++ adapters that put the MI355X step engine behind DMD's
// own interfaces.  Header-only; compile inside DMD's tree (its src/ on the
// include path) and link libdmdx.so and libdmdx_h5md.so (trajectory frames
// of DeviceSimulationEngine::run; HDF5, which DMD already links).
//
//   * dmdx_dmd::GpuLennardJones        : LennardJones (src/force/lennard_jones.h)
//   * dmdx_dmd::GpuCoulombDirect, GpuCoulombExclusion, GpuCoulombPME,
//     GpuHarmonicBond, GpuHarmonicAngle, GpuPeriodicDihedral,
//     GpuHarmonicImproper               : ForceComponent (force_component.h:8-22)
//   * dmdx_dmd::build_force_engine_gpu  : drop-in for build_force_engine
//                                         (src/sim/simulation_engine.cpp:147-203)
//   * dmdx_dmd::DeviceSimulationEngine  : SimulationEngine's step()/run()/
//                                         checkpoint surface with state in HBM
//
// Status codes from the C-ABI become ForceException (src/core/error.h:14-21)
// with the matching ForceError.
#pragma once

#include <algorithm>
#include <cstring>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

#include "core/cell.h"
#include "core/error.h"
#include "core/system_data.h"
#include "force/force_component.h"
#include "force/force_engine.h"
#include "force/lennard_jones.h"
#include "sim/checkpoint.h"
#include "sim/simulation_config.h"
#include "dmdx.h"
#include "dmdx_h5md.h"

namespace dmdx_dmd {

inline void check(int rc, const std::string& component)
{
    if (rc == DMDX_OK) return;
    ForceError e = ForceError::GPUError;
    switch (rc) {
    case DMDX_E_NAN: e = ForceError::NaN; break;
    case DMDX_E_BOND_EXPLOSION: e = ForceError::BondExplosion; break;
    case DMDX_E_MISMATCH: e = ForceError::Mismatch; break;
    case DMDX_E_PARM: e = ForceError::ParmError; break;
    default: break;
    }
    throw ForceException(e, component.c_str(), std::string(dmdx_last_error()));
}

struct ComponentHandle {
    dmdx_component* h = nullptr;
    ComponentHandle() = default;
    ComponentHandle(const ComponentHandle&) = delete;
    ComponentHandle& operator=(const ComponentHandle&) = delete;
    ~ComponentHandle() { dmdx_component_destroy(h); }
};

// CellType (src/core/cell.h:3-8) has the values of DMDX_CELL_*: the device
// path takes the type with the matrix (identity minimum image unless
// orthorhombic, src/core/cell.cpp:26-38; volume from the determinant).
inline int32_t cell_type_of(const Cell& cell) { return (int32_t)cell.type; }
static_assert((int)CellType::triclinic == DMDX_CELL_TRICLINIC &&
                  (int)CellType::rhombic_dodecahedron == DMDX_CELL_RHOMBIC_DODECAHEDRON,
              "CellType and DMDX_CELL_* differ");

inline void compute_on(dmdx_component* h, const std::string& name, std::span<const double> x,
                       std::span<const double> y, std::span<const double> z, const Cell& cell,
                       std::span<double> fx, std::span<double> fy, std::span<double> fz,
                       double& energy_out)
{
    if (y.size() != x.size() || z.size() != x.size() || fx.size() != x.size() ||
        fy.size() != x.size() || fz.size() != x.size())
        throw ForceException(ForceError::Mismatch, name.c_str(), "span sizes differ");
    check(dmdx_component_compute_cell(h, x.data(), y.data(), z.data(), (int64_t)x.size(),
                                      cell.matrix, cell_type_of(cell), fx.data(), fy.data(),
                                      fz.data(), &energy_out),
          name);
}

// LennardJones on the GPU.  Stays a LennardJones, so name() == "LennardJones"
// and the dynamic_cast<LennardJones*> of SimulationEngine::save_checkpoint /
// load_checkpoint (src/sim/simulation_engine.cpp:108-137) still finds it; the
// rebuild counter of the base object is the state seen between calls.
class GpuLennardJones : public LennardJones {
public:
    GpuLennardJones(LennardJonesParams params, double cutoff, std::span<const int> types)
        : LennardJones(params, cutoff)
    {
        check(dmdx_lj_create((int32_t)params.n_types, params.sigma.data(), params.epsilon.data(),
                             cutoff, &h_.h),
              "LennardJones");
        if (!types.empty()) {
            std::vector<int32_t> t(types.begin(), types.end());
            check(dmdx_lj_set_atom_types(h_.h, t.data(), (int64_t)t.size()), "LennardJones");
        }
    }
    std::string name() const override { return "LennardJones"; }
    void compute(std::span<const double> x, std::span<const double> y, std::span<const double> z,
                 const Cell& cell, std::span<double> fx, std::span<double> fy,
                 std::span<double> fz, double& energy_out) override
    {
        check(dmdx_lj_set_step_since_rebuild(h_.h, LennardJones::step_since_rebuild()),
              "LennardJones");
        compute_on(h_.h, "LennardJones", x, y, z, cell, fx, fy, fz, energy_out);
        int32_t c = 0;
        check(dmdx_lj_get_step_since_rebuild(h_.h, &c), "LennardJones");
        LennardJones::set_step_since_rebuild(c);
    }
    void set_atom_types_gpu(std::span<const int> types)
    {
        std::vector<int32_t> t(types.begin(), types.end());
        check(dmdx_lj_set_atom_types(h_.h, t.data(), (int64_t)t.size()), "LennardJones");
    }
    void rebuild_pair_list_gpu(std::span<const double> x, std::span<const double> y,
                               std::span<const double> z, const Cell& cell)
    {
        check(dmdx_component_rebuild_pair_list_cell(h_.h, x.data(), y.data(), z.data(),
                                                    (int64_t)x.size(), cell.matrix,
                                                    cell_type_of(cell)),
              "LennardJones");
        LennardJones::set_step_since_rebuild(0);
    }

private:
    ComponentHandle h_;
};

// Any other component: owns a handle made by `make`.
class GpuComponent : public ForceComponent {
public:
    std::string name() const override { return name_; }
    void compute(std::span<const double> x, std::span<const double> y, std::span<const double> z,
                 const Cell& cell, std::span<double> fx, std::span<double> fy,
                 std::span<double> fz, double& energy_out) override
    {
        compute_on(h_.h, name_, x, y, z, cell, fx, fy, fz, energy_out);
    }
    dmdx_component* handle() { return h_.h; }

protected:
    explicit GpuComponent(std::string name) : name_(std::move(name)) {}
    ComponentHandle h_;
    std::string name_;
};

class GpuCoulombDirect : public GpuComponent {
public:
    GpuCoulombDirect(double cutoff, double alpha, std::span<const double> charges)
        : GpuComponent("CoulombDirect")
    {
        check(dmdx_coulomb_direct_create(cutoff, alpha, &h_.h), name_);
        check(dmdx_component_set_charges(h_.h, charges.data(), (int64_t)charges.size()), name_);
    }
    void rebuild_pair_list(std::span<const double> x, std::span<const double> y,
                           std::span<const double> z, const Cell& cell)
    {
        check(dmdx_component_rebuild_pair_list_cell(h_.h, x.data(), y.data(), z.data(),
                                                    (int64_t)x.size(), cell.matrix,
                                                    cell_type_of(cell)),
              name_);
    }
};

class GpuCoulombExclusion : public GpuComponent {
public:
    // alpha: Ewald coefficient of the DMDX_CORRECTED erf form (unused by
    // DMD's own bare 1/r term)
    GpuCoulombExclusion(std::span<const double> charges, std::span<const int> ei,
                        std::span<const int> ej, double alpha = 3.5)
        : GpuComponent("CoulombExclusion")
    {
        check(dmdx_coulomb_exclusion_create(&h_.h), name_);
        check(dmdx_coulomb_exclusion_set_ewald_coefficient(h_.h, alpha), name_);
        check(dmdx_component_set_charges(h_.h, charges.data(), (int64_t)charges.size()), name_);
        std::vector<int32_t> a(ei.begin(), ei.end()), b(ej.begin(), ej.end());
        check(dmdx_coulomb_exclusion_set_exclusions(h_.h, a.data(), b.data(), (int64_t)a.size()),
              name_);
    }
};

class GpuCoulombPME : public GpuComponent {
public:
    GpuCoulombPME(double cutoff, double alpha, int nx, int ny, int nz, int order,
                  std::span<const double> charges)
        : GpuComponent("CoulombPME")
    {
        check(dmdx_pme_create(cutoff, alpha, nx, ny, nz, order, &h_.h), name_);
        check(dmdx_component_set_charges(h_.h, charges.data(), (int64_t)charges.size()), name_);
    }
};

class GpuHarmonicBond : public GpuComponent {
public:
    explicit GpuHarmonicBond(const HarmonicBondParams& p) : GpuComponent("HarmonicBond")
    {
        std::vector<int32_t> i(p.i.begin(), p.i.end()), j(p.j.begin(), p.j.end());
        check(dmdx_harmonic_bond_create((int64_t)p.n, i.data(), j.data(), p.k.data(), p.r0.data(),
                                        &h_.h),
              name_);
    }
};

class GpuHarmonicAngle : public GpuComponent {
public:
    explicit GpuHarmonicAngle(const HarmonicAngleParams& p) : GpuComponent("HarmonicAngle")
    {
        std::vector<int32_t> i(p.i.begin(), p.i.end()), j(p.j.begin(), p.j.end()),
            k(p.k.begin(), p.k.end());
        check(dmdx_harmonic_angle_create((int64_t)p.n, i.data(), j.data(), k.data(),
                                         p.k_theta.data(), p.theta0.data(), &h_.h),
              name_);
    }
};

class GpuPeriodicDihedral : public GpuComponent {
public:
    explicit GpuPeriodicDihedral(const PeriodicDihedralParams& p)
        : GpuComponent("PeriodicDihedral")
    {
        std::vector<int32_t> i(p.i.begin(), p.i.end()), j(p.j.begin(), p.j.end()),
            k(p.k.begin(), p.k.end()), l(p.l.begin(), p.l.end()),
            per(p.periodicity.begin(), p.periodicity.end());
        check(dmdx_periodic_dihedral_create((int64_t)p.n, i.data(), j.data(), k.data(), l.data(),
                                            per.data(), p.k_phi.data(), p.phi0.data(), &h_.h),
              name_);
    }
};

class GpuHarmonicImproper : public GpuComponent {
public:
    explicit GpuHarmonicImproper(const ImproperParams& p) : GpuComponent("HarmonicImproper")
    {
        std::vector<int32_t> i(p.i.begin(), p.i.end()), j(p.j.begin(), p.j.end()),
            k(p.k.begin(), p.k.end()), l(p.l.begin(), p.l.end());
        check(dmdx_harmonic_improper_create((int64_t)p.n, i.data(), j.data(), k.data(), l.data(),
                                            p.k_phi.data(), p.phi0.data(), &h_.h),
              name_);
    }
};

// build_force_engine (src/sim/simulation_engine.cpp:147-203) with the GPU
// components, in the same order and with the same grid rule; charges, types
// and exclusions come from cfg, which build_simulation also copies.
inline std::unique_ptr<ForceEngine> build_force_engine_gpu(const SimulationConfig& cfg)
{
    auto fe = std::make_unique<ForceEngine>();
    std::vector<int> types = cfg.atom_types;
    if (types.empty()) types.assign(cfg.pos_x.size(), 0);
    std::vector<double> q = cfg.charges;
    if (q.empty()) q.assign(cfg.pos_x.size(), 0.0);
    if (cfg.use_lj) fe->add_component(std::make_unique<GpuLennardJones>(cfg.lj, cfg.lj_cutoff, types));
    for (auto& b : cfg.bonds) fe->add_component(std::make_unique<GpuHarmonicBond>(b));
    for (auto& a : cfg.angles) fe->add_component(std::make_unique<GpuHarmonicAngle>(a));
    for (auto& d : cfg.dihedrals) fe->add_component(std::make_unique<GpuPeriodicDihedral>(d));
    for (auto& im : cfg.impropers) fe->add_component(std::make_unique<GpuHarmonicImproper>(im));
    if (cfg.coulomb_type == "direct" || cfg.coulomb_type == "cutoff" || cfg.coulomb_type == "pme")
        fe->add_component(std::make_unique<GpuCoulombDirect>(cfg.coulomb_cutoff, cfg.ewald_coeff, q));
    if (cfg.coulomb_type == "pme") {
        int g = std::max(8, static_cast<int>(cfg.box_size / cfg.pme_grid_spacing + 0.5));
        fe->add_component(std::make_unique<GpuCoulombPME>(cfg.coulomb_cutoff, cfg.ewald_coeff, g, g,
                                                          g, cfg.pme_order, q));
        fe->add_component(
            std::make_unique<GpuCoulombExclusion>(q, cfg.excl_i, cfg.excl_j, cfg.ewald_coeff));
    } else if ((cfg.coulomb_type == "direct" || cfg.coulomb_type == "cutoff") &&
               !cfg.excl_i.empty()) {
        fe->add_component(
            std::make_unique<GpuCoulombExclusion>(q, cfg.excl_i, cfg.excl_j, cfg.ewald_coeff));
    }
    return fe;
}

// SimulationEngine's step()/run()/checkpoint surface over dmdx_engine: the
// whole step runs on the device; system() copies the state back into a
// SystemData.  Checkpoints use DMD's own checkpoint::save/load (binary format
// of src/sim/checkpoint.cpp, including the LJ counter), so a device run and a
// CPU run can restart from each other's files.
class DeviceSimulationEngine {
public:
    // `sys` is the state build_simulation produced (positions, velocities
    // from gen_vel, masses, charges, types, exclusions); precision
    // DMDX_FP64 or DMDX_MIXED, physics DMDX_FAITHFUL or DMDX_CORRECTED.
    DeviceSimulationEngine(const SimulationConfig& cfg, const SystemData& sys, int precision,
                           int physics)
        : n_(sys.n_atoms), n_steps_(cfg.n_steps), traj_interval_(cfg.nstxout),
          ckpt_interval_(cfg.checkpoint_interval), traj_path_(cfg.trajectory_path),
          ckpt_path_(cfg.checkpoint_path), statics_(sys)
    {
        clock_.dt = cfg.dt;
        dmdx_system_desc d;
        std::memset(&d, 0, sizeof(d));
        d.n_atoms = (int64_t)sys.n_atoms;
        d.box[0] = d.box[1] = d.box[2] = cfg.box_size;
        d.pos_x = sys.pos_x.data();
        d.pos_y = sys.pos_y.data();
        d.pos_z = sys.pos_z.data();
        d.vel_x = sys.vel_x.data();
        d.vel_y = sys.vel_y.data();
        d.vel_z = sys.vel_z.data();
        d.masses = sys.masses.data();
        d.charges = sys.charges.data();
        types_.assign(sys.atom_types.begin(), sys.atom_types.end());
        d.atom_types = types_.data();
        d.use_lj = cfg.use_lj;
        d.lj_n_types = (int32_t)cfg.lj.n_types;
        d.lj_sigma = cfg.lj.sigma.data();
        d.lj_epsilon = cfg.lj.epsilon.data();
        d.lj_cutoff = cfg.lj_cutoff;
        d.rebuild_interval = 10;
        d.step_since_rebuild = 10;
        for (auto& b : cfg.bonds) append(bi_, bj_, b.i, b.j), cat(bk_, b.k), cat(br_, b.r0);
        d.n_bonds = (int64_t)bi_.size();
        d.bond_i = bi_.data();
        d.bond_j = bj_.data();
        d.bond_k = bk_.data();
        d.bond_r0 = br_.data();
        for (auto& a : cfg.angles) {
            append(ai_, aj_, a.i, a.j);
            cati(ak_, a.k);
            cat(akt_, a.k_theta);
            cat(at0_, a.theta0);
        }
        d.n_angles = (int64_t)ai_.size();
        d.angle_i = ai_.data();
        d.angle_j = aj_.data();
        d.angle_k = ak_.data();
        d.angle_k_theta = akt_.data();
        d.angle_theta0 = at0_.data();
        for (auto& h : cfg.dihedrals) {
            append(di_, dj_, h.i, h.j);
            append(dk_, dl_, h.k, h.l);
            cati(dp_, h.periodicity);
            cat(dkp_, h.k_phi);
            cat(dp0_, h.phi0);
        }
        d.n_dihedrals = (int64_t)di_.size();
        d.dih_i = di_.data();
        d.dih_j = dj_.data();
        d.dih_k = dk_.data();
        d.dih_l = dl_.data();
        d.dih_periodicity = dp_.data();
        d.dih_k_phi = dkp_.data();
        d.dih_phi0 = dp0_.data();
        for (auto& h : cfg.impropers) {
            append(ii_, ij_, h.i, h.j);
            append(ik_, il_, h.k, h.l);
            cat(ikp_, h.k_phi);
            cat(ip0_, h.phi0);
        }
        d.n_impropers = (int64_t)ii_.size();
        d.imp_i = ii_.data();
        d.imp_j = ij_.data();
        d.imp_k = ik_.data();
        d.imp_l = il_.data();
        d.imp_k_phi = ikp_.data();
        d.imp_phi0 = ip0_.data();
        d.coulomb_type = cfg.coulomb_type == "pme" ? DMDX_COULOMB_PME
                         : (cfg.coulomb_type == "direct" || cfg.coulomb_type == "cutoff")
                             ? DMDX_COULOMB_CUTOFF
                             : DMDX_COULOMB_NONE;
        d.coulomb_cutoff = cfg.coulomb_cutoff;
        d.ewald_coeff = cfg.ewald_coeff;
        int g = std::max(8, static_cast<int>(cfg.box_size / cfg.pme_grid_spacing + 0.5));
        d.pme_grid[0] = d.pme_grid[1] = d.pme_grid[2] = g;
        d.pme_order = cfg.pme_order;
        ei_.assign(cfg.excl_i.begin(), cfg.excl_i.end());
        ej_.assign(cfg.excl_j.begin(), cfg.excl_j.end());
        d.n_excl = (int64_t)ei_.size();
        d.excl_i = ei_.data();
        d.excl_j = ej_.data();
        d.dt = cfg.dt;
        const auto& t = cfg.thermostat;
        d.thermostat = t.type == "berendsen"     ? DMDX_THERMO_BERENDSEN
                       : t.type == "nose-hoover" ? DMDX_THERMO_NOSE_HOOVER
                       : t.type == "andersen"    ? DMDX_THERMO_ANDERSEN
                                                 : DMDX_THERMO_NONE;
        if (t.type != "none" && d.thermostat == DMDX_THERMO_NONE)
            throw std::runtime_error("Unknown thermostat type: " + t.type);
        d.thermo_temperature = t.temperature;
        d.thermo_tau = t.tau;
        d.andersen_frequency = t.frequency;
        // constraint_type "shake" with pairs (src/sim/simulation_engine.cpp:293-301)
        if (cfg.constraint_type == "shake" && !cfg.constraint_i.empty()) {
            ci_.assign(cfg.constraint_i.begin(), cfg.constraint_i.end());
            cj_.assign(cfg.constraint_j.begin(), cfg.constraint_j.end());
            d.n_constraints = (int64_t)ci_.size();
            d.con_i = ci_.data();
            d.con_j = cj_.data();
            d.con_distance = cfg.constraint_dist.data();
            d.con_tolerance = cfg.constraint_tolerance;
        }
        const auto& b = cfg.barostat;
        d.barostat = b.type == "berendsen" ? DMDX_BARO_BERENDSEN
                     : b.type == "andersen" ? DMDX_BARO_ANDERSEN
                                            : DMDX_BARO_NONE;
        if (b.type != "none" && d.barostat == DMDX_BARO_NONE)
            throw std::runtime_error("Unknown barostat type: " + b.type);
        d.baro_pressure = b.pressure;
        d.baro_tau = b.tau;
        d.baro_compressibility = b.compressibility;
        d.physics = physics;
        d.precision = precision;
        check(dmdx_engine_create(&d, 1, &e_), "SimulationEngine");
    }
    ~DeviceSimulationEngine()
    {
        if (traj_) dmdx_h5md_close(traj_);
        dmdx_engine_destroy(e_);
    }
    DeviceSimulationEngine(const DeviceSimulationEngine&) = delete;
    DeviceSimulationEngine& operator=(const DeviceSimulationEngine&) = delete;

    void step() { steps(1); }
    void steps(int64_t k)
    {
        check(dmdx_engine_step(e_, k), "SimulationEngine");
        for (int64_t i = 0; i < k; ++i) {  // SimulationEngine::step's clock (:54, :82-85)
            clock_.sys_time = clock_.time;
            clock_.step += 1;
            clock_.time += clock_.dt;
        }
    }
    // SimulationEngine::run (simulation_engine.cpp:88-106): n_steps steps, an
    // H5MD frame every nstxout steps (taken asynchronously and written while
    // the device runs on), a checkpoint every checkpoint_interval steps
    void run()
    {
        if (traj_interval_ > 0 && !traj_)
            h5(dmdx_h5md_open(traj_path_.c_str(), (int64_t)n_, 1, &traj_));
        int64_t left = n_steps_;
        while (left > 0) {
            int64_t k = left;
            if (ckpt_interval_ > 0)
                k = std::min<int64_t>(k, ckpt_interval_ - clock_.step % ckpt_interval_);
            h5(dmdx_h5md_run(e_, 0, k, traj_interval_ > 0 ? traj_interval_ : 0, &clock_, traj_));
            left -= k;
            if (ckpt_interval_ > 0 && clock_.step % ckpt_interval_ == 0)
                save_checkpoint(ckpt_path_, statics_);
        }
    }
    int current_step() const { return (int)clock_.step; }
    double current_time() const { return clock_.time; }
    // frames written so far by run()
    int64_t trajectory_frames() const
    {
        int64_t f = 0;
        if (traj_) dmdx_h5md_frame_count(traj_, &f);
        return f;
    }

    SystemData system() const
    {
        SystemData s(n_);
        check(dmdx_engine_get_arrays(e_, 0, 0, s.pos_x.data(), s.pos_y.data(), s.pos_z.data()),
              "SimulationEngine");
        check(dmdx_engine_get_arrays(e_, 0, 1, s.vel_x.data(), s.vel_y.data(), s.vel_z.data()),
              "SimulationEngine");
        check(dmdx_engine_get_arrays(e_, 0, 2, s.forces_x.data(), s.forces_y.data(),
                                     s.forces_z.data()),
              "SimulationEngine");
        dmdx_scalars sc;
        check(dmdx_engine_get_scalars(e_, 0, &sc), "SimulationEngine");
        s.potential_energy = sc.potential_energy;
        s.step = (size_t)clock_.step;
        s.time = clock_.sys_time;  // set at the start of the last step (:55)
        return s;
    }

    // checkpoint::save / load of src/sim/checkpoint.cpp; masses, charges and
    // types are those of the SystemData the engine was built from
    void save_checkpoint(const std::string& path, const SystemData& statics) const
    {
        SystemData s = system();
        s.masses = statics.masses;
        s.charges = statics.charges;
        s.atom_types = statics.atom_types;
        dmdx_scalars sc;
        check(dmdx_engine_get_scalars(e_, 0, &sc), "SimulationEngine");
        checkpoint::save(path, s, sc.step_since_rebuild);
    }
    void load_checkpoint(const std::string& path)
    {
        SystemData s(n_);
        int lj_counter = 0;
        checkpoint::load(path, s, lj_counter);
        check(dmdx_engine_set_arrays(e_, 0, 0, s.pos_x.data(), s.pos_y.data(), s.pos_z.data()),
              "SimulationEngine");
        check(dmdx_engine_set_arrays(e_, 0, 1, s.vel_x.data(), s.vel_y.data(), s.vel_z.data()),
              "SimulationEngine");
        check(dmdx_engine_set_step_since_rebuild(e_, 0, lj_counter), "SimulationEngine");
        clock_.step = (int64_t)s.step;
        clock_.time = s.time;
        clock_.sys_time = s.time;
    }

private:
    static void append(std::vector<int32_t>& a, std::vector<int32_t>& b, const std::vector<int>& x,
                       const std::vector<int>& y)
    {
        a.insert(a.end(), x.begin(), x.end());
        b.insert(b.end(), y.begin(), y.end());
    }
    static void cat(std::vector<double>& a, const std::vector<double>& x)
    {
        a.insert(a.end(), x.begin(), x.end());
    }
    static void cati(std::vector<int32_t>& a, const std::vector<int>& x)
    {
        a.insert(a.end(), x.begin(), x.end());
    }

    static void h5(int rc)
    {
        if (rc != DMDX_OK) throw std::runtime_error(std::string("trajectory: ") +
                                                    dmdx_h5md_last_error());
    }

    dmdx_engine* e_ = nullptr;
    size_t n_;
    int n_steps_;
    int traj_interval_, ckpt_interval_;
    std::string traj_path_, ckpt_path_;
    SystemData statics_;  // masses, charges and types for checkpoints
    dmdx_clock clock_{0, 0.0, 0.0, 0.0};
    dmdx_h5md* traj_ = nullptr;
    std::vector<int32_t> types_, bi_, bj_, ai_, aj_, ak_, di_, dj_, dk_, dl_, dp_, ii_, ij_, ik_,
        il_, ei_, ej_, ci_, cj_;
    std::vector<double> bk_, br_, akt_, at0_, dkp_, dp0_, ikp_, ip0_;
};

}  // namespace dmdx_dmd
