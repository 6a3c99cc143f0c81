// adapter_check.cpp -- runs the adapters of include/dmdx_dmd.hpp inside DMD's
// own classes, compiled against DMD's headers and translation units
// (oracle/build_ref.sh builds it into oracle/_ref/adapter_check).
//
//   adapter_check --no-gpu   interface checks only: the GPU LennardJones is a
//                            LennardJones to dynamic_cast, names match, span
//                            size mismatches and a missing device surface as
//                            ForceException with the right ForceError.
//   adapter_check            the same, then three runs of one chain system
//                            (LJ, bonds, angles, dihedrals, impropers,
//                            charges, exclusions, cutoff Coulomb):
//                              A: DMD's ForceEngine with DMD's components
//                                 stepped by DMD's Integrator in the order of
//                                 SimulationEngine::step
//                                 (src/sim/simulation_engine.cpp:54-86);
//                              B: the same loop with build_force_engine_gpu;
//                              B': A and B again under a triclinic Cell
//                                  (identity minimum image);
//                              C: DeviceSimulationEngine (FP64, faithful),
//                                 plus a checkpoint save/load round trip in
//                                 DMD's checkpoint format, and run() with
//                                 H5MD frames and checkpoints on cadence.
//                            Prints one JSON line; exit status 0 when B and C
//                            match A to 1e-10 relative and the restart is
//                            bit-exact.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "core/cell.h"
#include "core/system_data.h"
#include "force/coulomb_direct.h"
#include "force/coulomb_exclusion.h"
#include "force/force_engine.h"
#include "force/harmonic_angle.h"
#include "force/harmonic_bond.h"
#include "force/harmonic_improper.h"
#include "force/lennard_jones.h"
#include "force/periodic_dihedral.h"
#include "integrate/integrator.h"
#include "sim/checkpoint.h"
#include "sim/simulation_config.h"

#include "dmdx_dmd.hpp"

#include <hdf5.h>

namespace {

int g_fail = 0;

void expect(bool ok, const char* what)
{
    if (!ok) {
        std::fprintf(stderr, "adapter_check: FAILED: %s\n", what);
        g_fail = 1;
    }
}

// 10x10x10 lattice sites (0.3 nm) in a 3.0 nm box; 8 chains of 25 atoms
// walk the lattice in serpentine order, positions jittered by +-0.02 nm.
SimulationConfig chain_config()
{
    SimulationConfig c;
    c.dt = 0.001;
    c.box_size = 3.0;
    const int n_chain = 8, len = 25, n = n_chain * len;
    uint64_t s = 0x9E3779B97F4A7C15ull;
    auto uni = [&]() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (double)(s >> 11) * (1.0 / 9007199254740992.0);
    };
    std::vector<int> order;
    for (int z = 0; z < 10; ++z)
        for (int y = 0; y < 10; ++y)
            for (int x = 0; x < 10; ++x) {
                int yy = (z % 2) ? 9 - y : y;
                int xx = (yy % 2) ? 9 - x : x;
                order.push_back((z * 10 + yy) * 10 + xx);
            }
    for (int a = 0; a < n; ++a) {
        int site = order[a];
        c.pos_x.push_back(0.15 + 0.3 * (site % 10) + 0.04 * (uni() - 0.5));
        c.pos_y.push_back(0.15 + 0.3 * ((site / 10) % 10) + 0.04 * (uni() - 0.5));
        c.pos_z.push_back(0.15 + 0.3 * (site / 100) + 0.04 * (uni() - 0.5));
        c.vel_x.push_back(0.4 * (uni() - 0.5));
        c.vel_y.push_back(0.4 * (uni() - 0.5));
        c.vel_z.push_back(0.4 * (uni() - 0.5));
        c.masses.push_back(12.0 + 4.0 * (a % 3));
        c.charges.push_back((a % 2) ? -0.25 : 0.25);
        c.atom_types.push_back(a % 2);
    }
    c.use_lj = true;
    c.lj_cutoff = 1.0;
    c.lj.n_types = 2;
    c.lj.sigma = {0.30, 0.31, 0.31, 0.32};
    c.lj.epsilon = {0.40, 0.45, 0.45, 0.50};
    HarmonicBondParams b;
    HarmonicAngleParams an;
    PeriodicDihedralParams d;
    ImproperParams im;
    for (int ch = 0; ch < n_chain; ++ch) {
        int o = ch * len;
        for (int k = 0; k + 1 < len; ++k) {
            b.i.push_back(o + k), b.j.push_back(o + k + 1), b.k.push_back(2.0e4),
                b.r0.push_back(0.3);
            c.excl_i.push_back(o + k), c.excl_j.push_back(o + k + 1);
        }
        for (int k = 0; k + 2 < len; ++k)
            an.i.push_back(o + k), an.j.push_back(o + k + 1), an.k.push_back(o + k + 2),
                an.k_theta.push_back(300.0), an.theta0.push_back(1.9);
        for (int k = 0; k + 3 < len; ++k)
            d.i.push_back(o + k), d.j.push_back(o + k + 1), d.k.push_back(o + k + 2),
                d.l.push_back(o + k + 3), d.periodicity.push_back(1 + k % 3),
                d.k_phi.push_back(3.0), d.phi0.push_back(0.2 * (k % 5));
        for (int k = 0; k + 3 < len; k += 4)
            im.i.push_back(o + k), im.j.push_back(o + k + 1), im.k.push_back(o + k + 2),
                im.l.push_back(o + k + 3), im.k_phi.push_back(40.0), im.phi0.push_back(0.1);
    }
    b.n = b.i.size();
    an.n = an.i.size();
    d.n = d.i.size();
    im.n = im.i.size();
    c.bonds.push_back(b);
    c.angles.push_back(an);
    c.dihedrals.push_back(d);
    c.impropers.push_back(im);
    c.coulomb_type = "cutoff";
    c.coulomb_cutoff = 1.0;
    c.ewald_coeff = 3.12;
    return c;
}

// The components and setters that build_force_engine and build_simulation
// (src/sim/simulation_engine.cpp:147-203, 248-336) create for this config.
std::unique_ptr<ForceEngine> cpu_force_engine(const SimulationConfig& c)
{
    auto fe = std::make_unique<ForceEngine>();
    auto lj = std::make_unique<LennardJones>(c.lj, c.lj_cutoff);
    lj->set_atom_types(c.atom_types);
    fe->add_component(std::move(lj));
    for (auto& b : c.bonds) fe->add_component(std::make_unique<HarmonicBond>(b));
    for (auto& a : c.angles) fe->add_component(std::make_unique<HarmonicAngle>(a));
    for (auto& d : c.dihedrals) fe->add_component(std::make_unique<PeriodicDihedral>(d));
    for (auto& im : c.impropers) fe->add_component(std::make_unique<HarmonicImproper>(im));
    CoulombDirectParams p;
    p.cutoff = c.coulomb_cutoff;
    p.ewald_coefficient = c.ewald_coeff;
    auto cd = std::make_unique<CoulombDirect>(p);
    cd->set_charges(c.charges);
    fe->add_component(std::move(cd));
    auto ex = std::make_unique<CoulombExclusion>();
    ex->set_exclusions(c.excl_i, c.excl_j);
    ex->set_charges(c.charges);
    fe->add_component(std::move(ex));
    return fe;
}

SystemData initial_state(const SimulationConfig& c)
{
    SystemData s(c.pos_x.size());
    s.pos_x = c.pos_x, s.pos_y = c.pos_y, s.pos_z = c.pos_z;
    s.vel_x = c.vel_x, s.vel_y = c.vel_y, s.vel_z = c.vel_z;
    s.masses = c.masses;
    s.charges = c.charges;
    s.atom_types = c.atom_types;
    return s;
}

void zero_forces(SystemData& s)
{
    s.forces_x.assign(s.n_atoms, 0.0);
    s.forces_y.assign(s.n_atoms, 0.0);
    s.forces_z.assign(s.n_atoms, 0.0);
    s.potential_energy = 0.0;
}

// One velocity-Verlet step in the order of SimulationEngine::step with no
// constraints, thermostat or barostat.
void dmd_step(ForceEngine& fe, Integrator& in, SystemData& s, const Cell& cell, double dt)
{
    zero_forces(s);
    fe.compute(s, cell);
    in.half_kick(s, dt);
    in.advance(s, dt);
    zero_forces(s);
    fe.compute(s, cell);
    in.half_kick(s, dt);
}

double rel_diff(const std::vector<double>& a, const std::vector<double>& b)
{
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < a.size(); ++i) {
        num = std::fmax(num, std::fabs(a[i] - b[i]));
        den = std::fmax(den, std::fabs(a[i]));
    }
    return den > 0.0 ? num / den : num;
}

// a whole dataset of a file written by the device engine's run()
template <class T>
std::vector<T> read_h5(const std::string& path, const char* name, hid_t type)
{
    std::vector<T> out;
    hid_t f = H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return out;
    hid_t ds = H5Dopen2(f, name, H5P_DEFAULT);
    if (ds >= 0) {
        hid_t sp = H5Dget_space(ds);
        out.resize((size_t)H5Sget_simple_extent_npoints(sp));
        if (!out.empty() && H5Dread(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) < 0)
            out.clear();
        H5Sclose(sp);
        H5Dclose(ds);
    }
    H5Fclose(f);
    return out;
}

double rel3(const SystemData& a, const SystemData& b, int what)
{
    if (what == 0)
        return std::fmax(rel_diff(a.pos_x, b.pos_x),
                         std::fmax(rel_diff(a.pos_y, b.pos_y), rel_diff(a.pos_z, b.pos_z)));
    if (what == 1)
        return std::fmax(rel_diff(a.vel_x, b.vel_x),
                         std::fmax(rel_diff(a.vel_y, b.vel_y), rel_diff(a.vel_z, b.vel_z)));
    return std::fmax(rel_diff(a.forces_x, b.forces_x),
                     std::fmax(rel_diff(a.forces_y, b.forces_y),
                               rel_diff(a.forces_z, b.forces_z)));
}

void interface_checks(bool have_gpu)
{
    LennardJonesParams p;
    p.n_types = 1;
    p.sigma = {0.3};
    p.epsilon = {0.5};
    std::vector<int> types(4, 0);
    std::unique_ptr<ForceComponent> c =
        std::make_unique<dmdx_dmd::GpuLennardJones>(p, 1.0, types);
    expect(c->name() == "LennardJones", "GpuLennardJones::name()");
    auto* lj = dynamic_cast<LennardJones*>(c.get());
    expect(lj != nullptr, "dynamic_cast<LennardJones*> finds the GPU component");
    expect(lj && lj->step_since_rebuild() == 10, "LJ counter starts at the rebuild interval");
    std::vector<double> x = {0.1, 0.5, 0.9, 1.3}, y(4, 0.2), z(4, 0.3), f3(3), f4(4, 0.0);
    Cell cell(3.0, 3.0, 3.0);
    double e = 0.0;
    bool mismatch = false;
    try {
        c->compute(x, y, z, cell, f3, f4, f4, e);
    } catch (const ForceException& ex) {
        mismatch = ex.err == ForceError::Mismatch && ex.component == "LennardJones";
    }
    expect(mismatch, "span size mismatch raises ForceException(Mismatch)");
    bool refused = false;
    try {
        const double flat[9] = {3.0, 0.0, 0.0, 0.5, 0.0, 0.0, 0.0, 0.0, 3.0};
        std::vector<double> fx(4, 0.0), fy(4, 0.0), fz(4, 0.0);
        c->compute(x, y, z, Cell(flat, CellType::triclinic), fx, fy, fz, e);
    } catch (const ForceException& ex) {
        refused = ex.err == ForceError::ParmError;
    }
    expect(refused, "a Cell with a zero diagonal entry raises ForceException(ParmError)");
    if (!have_gpu) {
        bool gpu_err = false;
        std::vector<double> fx(4, 0.0), fy(4, 0.0), fz(4, 0.0);
        try {
            c->compute(x, y, z, cell, fx, fy, fz, e);
        } catch (const ForceException& ex) {
            gpu_err = ex.err == ForceError::GPUError;
        }
        expect(gpu_err, "no device raises ForceException(GPUError)");
    }
    ForceComponent* cd = new dmdx_dmd::GpuCoulombExclusion(std::vector<double>(4, 0.1),
                                                           std::vector<int>{0},
                                                           std::vector<int>{1});
    expect(cd->name() == "CoulombExclusion", "GpuCoulombExclusion::name()");
    delete cd;
}

}  // namespace

int main(int argc, char** argv)
{
    bool no_gpu = argc > 1 && std::strcmp(argv[1], "--no-gpu") == 0;
    interface_checks(!no_gpu);
    if (no_gpu) {
        std::printf("{\"adapter_check\": \"interface\", \"ok\": %s}\n", g_fail ? "false" : "true");
        return g_fail;
    }

    const SimulationConfig cfg = chain_config();
    const Cell cell(cfg.box_size, cfg.box_size, cfg.box_size);
    const int n_steps = 30;
    Integrator in;

    auto fe_cpu = cpu_force_engine(cfg);
    SystemData a = initial_state(cfg);
    auto fe_gpu = dmdx_dmd::build_force_engine_gpu(cfg);
    SystemData b = initial_state(cfg);
    double worst_b = 0.0;
    std::vector<SystemData> a_at;  // DMD's state after steps 10, 20, 30
    for (int k = 0; k < n_steps; ++k) {
        dmd_step(*fe_cpu, in, a, cell, cfg.dt);
        dmd_step(*fe_gpu, in, b, cell, cfg.dt);
        worst_b = std::fmax(worst_b, rel3(a, b, 2));
        if ((k + 1) % 10 == 0) a_at.push_back(a);
    }
    double pos_b = rel3(a, b, 0), vel_b = rel3(a, b, 1);
    double pe_b = std::fabs(a.potential_energy - b.potential_energy) /
                  std::fmax(1e-300, std::fabs(a.potential_energy));
    auto* lj_a = dynamic_cast<LennardJones*>(fe_cpu->components()[0].get());
    auto* lj_b = dynamic_cast<LennardJones*>(fe_gpu->components()[0].get());
    expect(lj_a && lj_b && lj_a->step_since_rebuild() == lj_b->step_since_rebuild(),
           "LJ rebuild counters agree");

    // the same loop under a triclinic Cell: DMD's minimum image is then the
    // identity (src/core/cell.cpp:26-38), on DMD's components and on the GPU
    const double tri_m[9] = {3.0, 0.0, 0.0, 0.4, 3.0, 0.0, 0.3, 0.2, 3.0};
    const Cell tri(tri_m, CellType::triclinic);
    auto fe_cpu_t = cpu_force_engine(cfg);
    auto fe_gpu_t = dmdx_dmd::build_force_engine_gpu(cfg);
    SystemData at = initial_state(cfg), bt = initial_state(cfg);
    double worst_t = 0.0;
    for (int k = 0; k < n_steps; ++k) {
        dmd_step(*fe_cpu_t, in, at, tri, cfg.dt);
        dmd_step(*fe_gpu_t, in, bt, tri, cfg.dt);
        worst_t = std::fmax(worst_t, rel3(at, bt, 2));
    }
    const double pos_t = rel3(at, bt, 0);
    const double pe_t = std::fabs(at.potential_energy - bt.potential_energy) /
                        std::fmax(1e-300, std::fabs(at.potential_energy));
    // the identity image changes the physics: the periodic run differs
    expect(rel3(a, at, 2) > 1e-6, "the triclinic run differs from the periodic one");

    dmdx_dmd::DeviceSimulationEngine dev(cfg, initial_state(cfg), DMDX_FP64, DMDX_FAITHFUL);
    dev.steps(n_steps);
    SystemData c = dev.system();
    double pos_c = rel3(a, c, 0), vel_c = rel3(a, c, 1), frc_c = rel3(a, c, 2);

    // restart: save at n_steps, run 10 more; a fresh engine loads the file
    // and runs the same 10 steps
    const char* tmp = std::getenv("TMPDIR");
    std::string path = std::string(tmp ? tmp : "/tmp") + "/dmdx_adapter_check.ckpt";
    SystemData statics = initial_state(cfg);
    dev.save_checkpoint(path, statics);
    dev.steps(10);
    SystemData c1 = dev.system();
    dmdx_dmd::DeviceSimulationEngine dev2(cfg, initial_state(cfg), DMDX_FP64, DMDX_FAITHFUL);
    dev2.load_checkpoint(path);
    expect(dev2.current_step() == n_steps, "checkpoint restores the step");
    dev2.steps(10);
    SystemData c2 = dev2.system();
    double restart = std::fmax(rel3(c1, c2, 0), rel3(c1, c2, 1));
    SystemData chk(statics.n_atoms);
    int lj_counter = -1;
    checkpoint::load(path, chk, lj_counter);
    std::remove(path.c_str());
    expect(chk.step == (size_t)n_steps, "checkpoint::load reads the device engine's file");

    // SimulationEngine::run on the device: an H5MD frame every 10 steps, a
    // checkpoint every 15 (the file keeps the last, at step 30)
    double traj_rel = 1.0;
    {
        SimulationConfig rc = cfg;
        rc.n_steps = n_steps;
        rc.nstxout = 10;
        rc.checkpoint_interval = 15;
        rc.trajectory_path = std::string(tmp ? tmp : "/tmp") + "/dmdx_adapter_check.h5";
        rc.checkpoint_path = std::string(tmp ? tmp : "/tmp") + "/dmdx_adapter_check_run.ckpt";
        SystemData last(statics.n_atoms);
        {
            dmdx_dmd::DeviceSimulationEngine dr(rc, initial_state(cfg), DMDX_FP64,
                                                DMDX_FAITHFUL);
            dr.run();
            expect(dr.current_step() == n_steps && dr.trajectory_frames() == 3,
                   "run() takes n_steps steps and writes a frame every nstxout");
            last = dr.system();
        }  // closes the trajectory
        const size_t n = statics.n_atoms;
        auto pos = read_h5<double>(rc.trajectory_path, "/particles/atoms/position/value",
                                   H5T_NATIVE_DOUBLE);
        auto steps = read_h5<int>(rc.trajectory_path, "/particles/atoms/position/step",
                                  H5T_NATIVE_INT);
        auto times = read_h5<double>(rc.trajectory_path, "/particles/atoms/position/time",
                                     H5T_NATIVE_DOUBLE);
        std::remove(rc.trajectory_path.c_str());
        expect(pos.size() == 3 * n * 3 && steps == std::vector<int>({10, 20, 30}),
               "the trajectory holds frames 10, 20, 30");
        // SystemData::time of a frame: the engine's running time at the start
        // of the frame's last step (simulation_engine.cpp:55, :85)
        double t = 0.0, t_start = 0.0;
        std::vector<double> want_t;
        for (int k = 1; k <= n_steps; ++k) {
            t_start = t;
            t += cfg.dt;
            if (k % 10 == 0) want_t.push_back(t_start);
        }
        expect(times == want_t, "frame times are SystemData::time");
        if (pos.size() == 3 * n * 3) {
            traj_rel = 0.0;
            for (int f = 0; f < 3; ++f) {
                SystemData fr(n);
                for (size_t i = 0; i < n; ++i) {
                    fr.pos_x[i] = pos[(f * n + i) * 3 + 0];
                    fr.pos_y[i] = pos[(f * n + i) * 3 + 1];
                    fr.pos_z[i] = pos[(f * n + i) * 3 + 2];
                }
                traj_rel = std::fmax(traj_rel, rel3(a_at[f], fr, 0));
            }
        }
        SystemData ck(n);
        int ck_counter = -1;
        checkpoint::load(rc.checkpoint_path, ck, ck_counter);
        std::remove(rc.checkpoint_path.c_str());
        expect(ck.step == (size_t)n_steps && ck.time == want_t.back() &&
                   ck.pos_x == last.pos_x && ck.vel_z == last.vel_z,
               "run() checkpoints the state at checkpoint_interval");
    }

    const double tol = 1e-10;
    expect(worst_b <= tol && pos_b <= tol && vel_b <= tol && pe_b <= tol,
           "GPU components inside DMD's loop match DMD's components");
    expect(worst_t <= tol && pos_t <= tol && pe_t <= tol,
           "GPU components under a triclinic Cell match DMD's components");
    expect(pos_c <= tol && vel_c <= tol && frc_c <= tol,
           "DeviceSimulationEngine matches DMD's loop");
    expect(restart == 0.0, "restart from the checkpoint is bit-exact");
    expect(traj_rel <= tol, "run()'s frames match DMD's loop at the frame steps");
    std::printf(
        "{\"adapter_check\": \"full\", \"n_atoms\": %zu, \"steps\": %d, "
        "\"components_force_rel\": %.3e, \"components_pos_rel\": %.3e, "
        "\"components_vel_rel\": %.3e, \"components_pe_rel\": %.3e, "
        "\"device_pos_rel\": %.3e, \"device_vel_rel\": %.3e, \"device_force_rel\": %.3e, "
        "\"triclinic_force_rel\": %.3e, \"triclinic_pos_rel\": %.3e, "
        "\"triclinic_pe_rel\": %.3e, "
        "\"restart_rel\": %.3e, \"trajectory_pos_rel\": %.3e, \"lj_counter\": %d, "
        "\"ok\": %s}\n",
        a.n_atoms, n_steps, worst_b, pos_b, vel_b, pe_b, pos_c, vel_c, frc_c, worst_t, pos_t,
        pe_t, restart, traj_rel, lj_counter, g_fail ? "false" : "true");
    return g_fail;
}
