// components.cpp -- ForceComponent-level C-ABI.
//
// Each handle stands for one of DMD's force components
// (src/force/force_component.h:8-22) and keeps the component's state on the
// device: the LJ pair list and its rebuild counter, the CoulombDirect list,
// charges, exclusions, bonded parameters.  dmdx_component_compute has the
// reference's accumulate semantics -- forces and energy are added to what
// the caller's arrays already hold (src/force/force_engine.cpp:13-23,
// tests/unit/force/test_force_engine.cpp:46-57).  Positions and forces are
// host arrays, like the std::span arguments of the reference, so every call
// copies them to and from HBM; the device-resident path is dmdx_engine_*.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "engine.h"

using dmdx::Batch;
using dmdx::BatchOptions;

struct dmdx_component {
    int type = dmdx::COMP_NONE;
    std::string name;
    int physics = DMDX_FAITHFUL;
    int precision = DMDX_FP64;
    // parameters (host copies, as the reference setters copy)
    int32_t n_types = 0;
    std::vector<double> sigma, epsilon;
    double cutoff = 1.2, alpha = 3.5;
    int32_t grid[3] = {32, 32, 32};
    int32_t order = 4;
    std::vector<int32_t> types, ei, ej;
    std::vector<double> charges;
    std::vector<int32_t> ti, tj, tk, tl, per;
    std::vector<double> k1, k2;
    int64_t n_terms = 0;
    int32_t counter = 10;  // LennardJones::step_since_rebuild_ (lennard_jones.cpp:9-10)
    int32_t max_nbr = 0;   // list row width; 0 = estimated from the density
    // device side
    std::unique_ptr<Batch> b;
    int64_t n = -1;
    double box[3] = {0, 0, 0};
};

namespace {

// Cell (src/core/cell.h:10-23): any row-major matrix with a positive
// diagonal; its type selects the minimum image (identity unless
// orthorhombic, src/core/cell.cpp:26-38) and its determinant is the volume.
int check_cell(const double cell[9], int type)
{
    if (!cell) return dmdx::fail(DMDX_E_INVALID, "null cell");
    if (type < DMDX_CELL_ORTHORHOMBIC || type > DMDX_CELL_RHOMBIC_DODECAHEDRON)
        return dmdx::fail(DMDX_E_PARM, "unknown cell type");
    if (!(cell[0] > 0.0 && cell[4] > 0.0 && cell[8] > 0.0))
        return dmdx::fail(DMDX_E_PARM, "cell edges (the matrix diagonal) must be positive");
    return DMDX_OK;
}

// One-system batch of the component for n atoms at the given (host)
// positions and box, with the launch options `opt` (component fields set
// here).
int make_batch(dmdx_component* c, int64_t n, const double* x, const double* y, const double* z,
               const double cell[9], BatchOptions opt, Batch& b)
{
    if (c->type == dmdx::COMP_LJ && !c->types.empty() && (int64_t)c->types.size() != n)
        return dmdx::fail(DMDX_E_MISMATCH, "atom types were set for a different atom count");
    if (!c->charges.empty() && (int64_t)c->charges.size() != n)
        return dmdx::fail(DMDX_E_MISMATCH, "charges were set for a different atom count");
    std::vector<double> masses(n, 1.0);
    std::vector<int32_t> types = c->types;
    if (types.empty()) types.assign(n, 0);
    std::vector<double> q = c->charges;
    if (q.empty()) q.assign(n, 0.0);
    dmdx_system_desc D;
    std::memset(&D, 0, sizeof(D));
    D.n_atoms = n;
    D.box[0] = cell[0];
    D.box[1] = cell[4];
    D.box[2] = cell[8];
    D.pos_x = x;
    D.pos_y = y;
    D.pos_z = z;
    D.masses = masses.data();
    D.charges = q.data();
    D.atom_types = types.data();
    D.physics = c->physics;
    D.precision = c->precision;
    D.dt = 0.001;
    D.rebuild_interval = 10;
    D.step_since_rebuild = c->counter;
    D.lj_cutoff = c->cutoff;
    D.coulomb_cutoff = c->cutoff;
    D.ewald_coeff = c->alpha;
    opt.component = c->type;
    opt.reuse_forces = false;
    opt.max_nbr = c->max_nbr;
    switch (c->type) {
    case dmdx::COMP_LJ:
        D.use_lj = 1;
        D.lj_n_types = c->n_types;
        D.lj_sigma = c->sigma.data();
        D.lj_epsilon = c->epsilon.data();
        break;
    case dmdx::COMP_COULOMB_DIRECT:
        D.coulomb_type = DMDX_COULOMB_CUTOFF;
        break;
    case dmdx::COMP_EXCLUSION:
        D.n_excl = (int64_t)c->ei.size();
        D.excl_i = c->ei.data();
        D.excl_j = c->ej.data();
        break;
    case dmdx::COMP_PME:
        D.coulomb_type = DMDX_COULOMB_PME;
        D.pme_grid[0] = c->grid[0];
        D.pme_grid[1] = c->grid[1];
        D.pme_grid[2] = c->grid[2];
        D.pme_order = c->order;
        break;
    case dmdx::COMP_BOND:
        D.n_bonds = c->n_terms;
        D.bond_i = c->ti.data();
        D.bond_j = c->tj.data();
        D.bond_k = c->k1.data();
        D.bond_r0 = c->k2.data();
        break;
    case dmdx::COMP_ANGLE:
        D.n_angles = c->n_terms;
        D.angle_i = c->ti.data();
        D.angle_j = c->tj.data();
        D.angle_k = c->tk.data();
        D.angle_k_theta = c->k1.data();
        D.angle_theta0 = c->k2.data();
        break;
    case dmdx::COMP_DIHEDRAL:
        D.n_dihedrals = c->n_terms;
        D.dih_i = c->ti.data();
        D.dih_j = c->tj.data();
        D.dih_k = c->tk.data();
        D.dih_l = c->tl.data();
        D.dih_periodicity = c->per.data();
        D.dih_k_phi = c->k1.data();
        D.dih_phi0 = c->k2.data();
        break;
    case dmdx::COMP_IMPROPER:
        D.n_impropers = c->n_terms;
        D.imp_i = c->ti.data();
        D.imp_j = c->tj.data();
        D.imp_k = c->tk.data();
        D.imp_l = c->tl.data();
        D.imp_k_phi = c->k1.data();
        D.imp_phi0 = c->k2.data();
        break;
    default:
        return dmdx::fail(DMDX_E_INVALID, "unknown component");
    }
    return b.create(&D, 1, opt);
}

// (Re)create the component's own batch for n atoms at the given positions/box.
int ensure_batch(dmdx_component* c, int64_t n, const double* x, const double* y, const double* z,
                 const double cell[9])
{
    if (c->b && c->n == n) return DMDX_OK;
    if (c->b) c->counter = c->b->host[0].counter;
    auto b = std::make_unique<Batch>();
    int rc = make_batch(c, n, x, y, z, cell, BatchOptions(), *b);
    if (rc != DMDX_OK) return rc;
    c->b = std::move(b);
    c->n = n;
    c->box[0] = cell[0];
    c->box[1] = cell[4];
    c->box[2] = cell[8];
    return DMDX_OK;
}

int upload_state(dmdx_component* c, const double* x, const double* y, const double* z,
                 const double cell[9], int type)
{
    int rc = c->b->set_cell(0, cell, type);
    if (rc) return rc;
    return c->b->set_arrays(0, 0, x, y, z);
}

// DMD's pair lists are vectors that grow with the pairs found
// (lennard_jones.cpp:18-40); the device rows have a fixed width, estimated
// from the mean density.  After a build that overflowed (atoms far denser
// than the box average), the batch is made again with rows twice as wide and
// the caller repeats the call; `counter` is the list counter to restore.
int widen_list(dmdx_component* c, int64_t n, int32_t counter)
{
    const int have = c->b ? c->b->hsys[0].max_nbr : 0;
    if (have >= 2048 || have >= n - 1)
        return dmdx::fail(DMDX_E_MISMATCH, "neighbour list overflow (more than 2048 neighbours "
                                           "of one atom)");
    c->max_nbr = std::min(std::max(2 * have, 64), 2048);
    c->b.reset();
    c->n = -1;
    c->counter = counter;
    return DMDX_OK;
}

bool overflowed(dmdx_component* c, int* rc)
{
    *rc = c->b->fetch_state();
    return *rc == DMDX_OK && (c->b->hstate[0].flags & dmdx::FLAG_NBR_OVERFLOW);
}

int out(dmdx_component* c, dmdx_component** dst)
{
    *dst = c;
    return DMDX_OK;
}

template <class T>
std::vector<T> copy_of(const T* p, int64_t n)
{
    return p ? std::vector<T>(p, p + n) : std::vector<T>();
}

}  // namespace

extern "C" {

int dmdx_lj_create(int32_t n_types, const double* sigma, const double* epsilon, double cutoff,
                   dmdx_component** outp)
{
    if (!outp || !sigma || !epsilon || n_types <= 0)
        return dmdx::fail(DMDX_E_INVALID, "LennardJones needs n_types > 0 and sigma/epsilon tables");
    auto* c = new dmdx_component();
    c->type = dmdx::COMP_LJ;
    c->name = "LennardJones";
    c->n_types = n_types;
    c->sigma = copy_of(sigma, (int64_t)n_types * n_types);
    c->epsilon = copy_of(epsilon, (int64_t)n_types * n_types);
    c->cutoff = cutoff;
    return out(c, outp);
}

int dmdx_coulomb_direct_create(double cutoff, double ewald_coefficient, dmdx_component** outp)
{
    if (!outp) return dmdx::fail(DMDX_E_INVALID, "null out");
    auto* c = new dmdx_component();
    c->type = dmdx::COMP_COULOMB_DIRECT;
    c->name = "CoulombDirect";
    c->cutoff = cutoff;
    c->alpha = ewald_coefficient;
    return out(c, outp);
}

int dmdx_coulomb_exclusion_create(dmdx_component** outp)
{
    if (!outp) return dmdx::fail(DMDX_E_INVALID, "null out");
    auto* c = new dmdx_component();
    c->type = dmdx::COMP_EXCLUSION;
    c->name = "CoulombExclusion";
    return out(c, outp);
}

int dmdx_pme_create(double cutoff, double ewald_coefficient, int32_t nx, int32_t ny, int32_t nz,
                    int32_t spline_order, dmdx_component** outp)
{
    if (!outp) return dmdx::fail(DMDX_E_INVALID, "null out");
    auto* c = new dmdx_component();
    c->type = dmdx::COMP_PME;
    c->name = "CoulombPME";
    c->cutoff = cutoff;
    c->alpha = ewald_coefficient;
    c->grid[0] = nx;
    c->grid[1] = ny;
    c->grid[2] = nz;
    c->order = spline_order;
    return out(c, outp);
}

int dmdx_harmonic_bond_create(int64_t n, const int32_t* i, const int32_t* j, const double* k,
                              const double* r0, dmdx_component** outp)
{
    if (!outp || n < 0 || (n > 0 && (!i || !j || !k || !r0)))
        return dmdx::fail(DMDX_E_INVALID, "bad HarmonicBond parameters");
    auto* c = new dmdx_component();
    c->type = dmdx::COMP_BOND;
    c->name = "HarmonicBond";
    c->n_terms = n;
    c->ti = copy_of(i, n);
    c->tj = copy_of(j, n);
    c->k1 = copy_of(k, n);
    c->k2 = copy_of(r0, n);
    return out(c, outp);
}

int dmdx_harmonic_angle_create(int64_t n, const int32_t* i, const int32_t* j, const int32_t* k,
                               const double* k_theta, const double* theta0, dmdx_component** outp)
{
    if (!outp || n < 0 || (n > 0 && (!i || !j || !k || !k_theta || !theta0)))
        return dmdx::fail(DMDX_E_INVALID, "bad HarmonicAngle parameters");
    auto* c = new dmdx_component();
    c->type = dmdx::COMP_ANGLE;
    c->name = "HarmonicAngle";
    c->n_terms = n;
    c->ti = copy_of(i, n);
    c->tj = copy_of(j, n);
    c->tk = copy_of(k, n);
    c->k1 = copy_of(k_theta, n);
    c->k2 = copy_of(theta0, n);
    return out(c, outp);
}

int dmdx_periodic_dihedral_create(int64_t n, const int32_t* i, const int32_t* j, const int32_t* k,
                                  const int32_t* l, const int32_t* periodicity,
                                  const double* k_phi, const double* phi0, dmdx_component** outp)
{
    if (!outp || n < 0 || (n > 0 && (!i || !j || !k || !l || !periodicity || !k_phi || !phi0)))
        return dmdx::fail(DMDX_E_INVALID, "bad PeriodicDihedral parameters");
    auto* c = new dmdx_component();
    c->type = dmdx::COMP_DIHEDRAL;
    c->name = "PeriodicDihedral";
    c->n_terms = n;
    c->ti = copy_of(i, n);
    c->tj = copy_of(j, n);
    c->tk = copy_of(k, n);
    c->tl = copy_of(l, n);
    c->per = copy_of(periodicity, n);
    c->k1 = copy_of(k_phi, n);
    c->k2 = copy_of(phi0, n);
    return out(c, outp);
}

int dmdx_harmonic_improper_create(int64_t n, const int32_t* i, const int32_t* j, const int32_t* k,
                                  const int32_t* l, const double* k_phi, const double* phi0,
                                  dmdx_component** outp)
{
    if (!outp || n < 0 || (n > 0 && (!i || !j || !k || !l || !k_phi || !phi0)))
        return dmdx::fail(DMDX_E_INVALID, "bad HarmonicImproper parameters");
    auto* c = new dmdx_component();
    c->type = dmdx::COMP_IMPROPER;
    c->name = "HarmonicImproper";
    c->n_terms = n;
    c->ti = copy_of(i, n);
    c->tj = copy_of(j, n);
    c->tk = copy_of(k, n);
    c->tl = copy_of(l, n);
    c->k1 = copy_of(k_phi, n);
    c->k2 = copy_of(phi0, n);
    return out(c, outp);
}

int dmdx_component_destroy(dmdx_component* c)
{
    delete c;
    return DMDX_OK;
}

int dmdx_component_name(dmdx_component* c, char* buf, int64_t cap)
{
    if (!c || !buf || cap <= 0) return dmdx::fail(DMDX_E_INVALID, "bad arguments");
    std::strncpy(buf, c->name.c_str(), (size_t)cap - 1);
    buf[cap - 1] = '\0';
    return DMDX_OK;
}

int dmdx_component_set_mode(dmdx_component* c, int32_t physics, int32_t precision)
{
    if (!c) return dmdx::fail(DMDX_E_INVALID, "null component");
    if ((physics != DMDX_FAITHFUL && physics != DMDX_CORRECTED) ||
        (precision != DMDX_FP64 && precision != DMDX_MIXED))
        return dmdx::fail(DMDX_E_PARM, "unknown physics or precision");
    if (c->b) c->counter = c->b->host[0].counter;
    c->physics = physics;
    c->precision = precision;
    c->b.reset();  // rebuilt with the new mode on the next call
    c->n = -1;
    return DMDX_OK;
}

int dmdx_lj_set_atom_types(dmdx_component* c, const int32_t* types, int64_t n)
{
    if (!c || c->type != dmdx::COMP_LJ) return dmdx::fail(DMDX_E_INVALID, "not a LennardJones");
    if (n < 0 || (n > 0 && !types)) return dmdx::fail(DMDX_E_INVALID, "bad types");
    for (int64_t a = 0; a < n; ++a)
        if (types[a] < 0 || types[a] >= c->n_types)
            return dmdx::fail(DMDX_E_PARM, "atom type out of range of the LJ tables");
    c->types = copy_of(types, n);
    if (c->b && c->n == n) {  // keep the list: update in place
        int rc = c->b->copy(c->b->d.type, c->types.data(), sizeof(int32_t) * n,
                            hipMemcpyHostToDevice);
        if (rc) return rc;
    }
    return DMDX_OK;
}

int dmdx_component_set_charges(dmdx_component* c, const double* q, int64_t n)
{
    if (!c || (c->type != dmdx::COMP_COULOMB_DIRECT && c->type != dmdx::COMP_EXCLUSION &&
               c->type != dmdx::COMP_PME))
        return dmdx::fail(DMDX_E_INVALID, "component has no charges");
    if (n < 0 || (n > 0 && !q)) return dmdx::fail(DMDX_E_INVALID, "bad charges");
    c->charges = copy_of(q, n);
    if (c->b && c->n == n) {
        int rc = c->b->copy(c->b->d.q, c->charges.data(), sizeof(double) * n,
                            hipMemcpyHostToDevice);
        if (rc) return rc;
        // CoulombPME::set_charges recomputes the self energy
        // (src/force/coulomb_pme.cpp:42-45)
        double sum_q2 = 0.0, qs = 0.0;
        for (double v : c->charges) {
            sum_q2 += v * v;
            qs += v;
        }
        dmdx::SysDev& S = c->b->hsys[0];
        rc = c->b->copy(&S, c->b->d.sys, sizeof(S), hipMemcpyDeviceToHost);
        if (rc) return rc;
        S.self_energy = -dmdx::KE * c->alpha / std::sqrt(dmdx::PI) * sum_q2;
        S.qsum = qs;
        rc = c->b->upload_sys(0);
        if (rc) return rc;
        rc = c->b->sync();
        if (rc) return rc;
    }
    return DMDX_OK;
}

int dmdx_coulomb_exclusion_set_exclusions(dmdx_component* c, const int32_t* ei, const int32_t* ej,
                                          int64_t n)
{
    if (!c || c->type != dmdx::COMP_EXCLUSION)
        return dmdx::fail(DMDX_E_INVALID, "not a CoulombExclusion");
    if (n < 0 || (n > 0 && (!ei || !ej))) return dmdx::fail(DMDX_E_INVALID, "bad exclusions");
    c->ei = copy_of(ei, n);
    c->ej = copy_of(ej, n);
    c->b.reset();  // the term arrays are rebuilt on the next call
    c->n = -1;
    return DMDX_OK;
}

int dmdx_coulomb_exclusion_set_ewald_coefficient(dmdx_component* c, double alpha)
{
    if (!c || c->type != dmdx::COMP_EXCLUSION)
        return dmdx::fail(DMDX_E_INVALID, "not a CoulombExclusion");
    if (!(alpha > 0.0)) return dmdx::fail(DMDX_E_PARM, "the Ewald coefficient must be positive");
    c->alpha = alpha;
    c->b.reset();  // alpha is part of the system description
    c->n = -1;
    return DMDX_OK;
}

int dmdx_component_rebuild_pair_list(dmdx_component* c, const double* x, const double* y,
                                     const double* z, int64_t n, const double cell[9])
{
    return dmdx_component_rebuild_pair_list_cell(c, x, y, z, n, cell, DMDX_CELL_ORTHORHOMBIC);
}

int dmdx_component_rebuild_pair_list_cell(dmdx_component* c, const double* x, const double* y,
                                          const double* z, int64_t n, const double cell[9],
                                          int32_t cell_type)
{
    if (!c || (c->type != dmdx::COMP_LJ && c->type != dmdx::COMP_COULOMB_DIRECT))
        return dmdx::fail(DMDX_E_INVALID, "component has no pair list");
    if (!x || !y || !z || n <= 0) return dmdx::fail(DMDX_E_INVALID, "bad positions");
    int rc = check_cell(cell, cell_type);
    if (rc) return rc;
    for (;;) {
        rc = ensure_batch(c, n, x, y, z, cell);
        if (rc) return rc;
        rc = upload_state(c, x, y, z, cell, cell_type);
        if (rc) return rc;
        rc = c->b->rebuild_only();
        if (rc) return rc;
        if (!overflowed(c, &rc)) return rc;
        rc = widen_list(c, n, 0);
        if (rc) return rc;
    }
}

int dmdx_lj_get_step_since_rebuild(dmdx_component* c, int32_t* outv)
{
    if (!c || c->type != dmdx::COMP_LJ || !outv)
        return dmdx::fail(DMDX_E_INVALID, "not a LennardJones");
    *outv = c->b ? c->b->host[0].counter : c->counter;
    return DMDX_OK;
}

int dmdx_lj_set_step_since_rebuild(dmdx_component* c, int32_t s)
{
    if (!c || c->type != dmdx::COMP_LJ) return dmdx::fail(DMDX_E_INVALID, "not a LennardJones");
    c->counter = s;
    if (c->b) return c->b->set_counter(0, s);
    return DMDX_OK;
}

int dmdx_component_compute(dmdx_component* c, const double* x, const double* y, const double* z,
                           int64_t n, const double cell[9], double* fx, double* fy, double* fz,
                           double* energy_out)
{
    return dmdx_component_compute_cell(c, x, y, z, n, cell, DMDX_CELL_ORTHORHOMBIC, fx, fy, fz,
                                       energy_out);
}

int dmdx_component_compute_cell(dmdx_component* c, const double* x, const double* y,
                                const double* z, int64_t n, const double cell[9],
                                int32_t cell_type, double* fx, double* fy, double* fz,
                                double* energy_out)
{
    if (!c) return dmdx::fail(DMDX_E_INVALID, "null component");
    if (!x || !y || !z || !fx || !fy || !fz || !energy_out || n <= 0)
        return dmdx::fail(DMDX_E_INVALID, "null array or empty system");
    int rc = check_cell(cell, cell_type);
    if (rc) return rc;
    for (;;) {
        rc = ensure_batch(c, n, x, y, z, cell);
        if (rc) return rc;
        const int32_t counter = c->b->host[0].counter;
        rc = upload_state(c, x, y, z, cell, cell_type);
        if (rc) return rc;
        rc = c->b->set_arrays(0, 2, fx, fy, fz);
        if (rc) return rc;
        rc = c->b->compute(1);
        if (rc) return rc;
        if (!overflowed(c, &rc)) break;
        // the list was built in this call (a kept list never overflows):
        // repeat it with wider rows, the caller's forces untouched
        rc = widen_list(c, n, counter);
        if (rc) return rc;
    }
    if (rc) return rc;
    rc = c->b->get_arrays(0, 2, fx, fy, fz);
    if (rc) return rc;
    const dmdx::SysState& T = c->b->hstate[0];
    *energy_out += T.pe;
    if (!std::isfinite(T.pe)) return dmdx::fail(DMDX_E_NAN, c->name + ": non-finite energy");
    return DMDX_OK;
}

// dmdx_component_compute on device arrays with no device allocation: the
// component's one-system batch is laid out in the device's scratch array
// for the length of the call (same kernels, same order, so the same bits as
// dmdx_component_compute).  Components whose compute depends on a pair list
// kept between calls (LennardJones, CoulombDirect) are refused.
int dmdx_component_compute_noalloc(dmdx_component* c, const double* x, const double* y,
                                   const double* z, int64_t n, const double cell[9], double* fx,
                                   double* fy, double* fz, double* energy_out)
{
    if (!c) return dmdx::fail(DMDX_E_INVALID, "null component");
    if (!x || !y || !z || !fx || !fy || !fz || !energy_out || n <= 0)
        return dmdx::fail(DMDX_E_INVALID, "null array or empty system");
    if (c->type == dmdx::COMP_LJ || c->type == dmdx::COMP_COULOMB_DIRECT)
        return dmdx::fail(DMDX_E_PARM, c->name + " keeps its pair list between calls: it has no "
                                                 "allocation-free form");
    int rc = check_cell(cell, DMDX_CELL_ORTHORHOMBIC);
    if (rc) return rc;
    dmdx::NoallocDevice* nd = nullptr;
    rc = dmdx::noalloc_device(&nd);
    if (rc) return rc;
    BatchOptions opt;
    opt.arena = nd->scratch;
    opt.arena_bytes = nd->bytes;
    opt.arena_stream = nd->stream;
    opt.arena_side = nd->side;
    for (int k = 0; k < 3; ++k) opt.arena_events[k] = nd->events[k];
    // the batch is filled from host arrays; positions arrive below, on the device
    std::vector<double> zero((size_t)n, 0.0);
    Batch b;
    rc = make_batch(c, n, zero.data(), zero.data(), zero.data(), cell, opt, b);
    if (rc == DMDX_OK) rc = b.set_cell(0, cell, DMDX_CELL_ORTHORHOMBIC);
    if (rc == DMDX_OK) rc = b.set_arrays(0, 0, x, y, z, true);
    if (rc == DMDX_OK) rc = b.set_arrays(0, 2, fx, fy, fz, true);
    if (rc == DMDX_OK) rc = b.compute(1);
    if (rc == DMDX_OK) rc = b.get_arrays(0, 2, fx, fy, fz, true);
    if (rc == DMDX_OK) rc = b.fetch_state();
    if (rc != DMDX_OK) {
        (void)hipStreamSynchronize(nd->stream);
        (void)hipStreamSynchronize(nd->side);
        return rc;
    }
    const dmdx::SysState& T = b.hstate[0];
    if (!std::isfinite(T.pe)) return dmdx::fail(DMDX_E_NAN, c->name + ": non-finite energy");
    DMDX_HIP(dmdx::launch_add_energy(energy_out, b.d.st, nd->stream));
    DMDX_HIP(hipStreamSynchronize(nd->stream));
    return DMDX_OK;
}

// CoulombPME static helpers (src/force/coulomb_pme.h:39-41), host side
int dmdx_pme_theta4(double u, double* outv)
{
    if (!outv) return dmdx::fail(DMDX_E_INVALID, "null out");
    double a = std::fabs(u);
    if (a >= 2.0) *outv = 0.0;
    else if (a >= 1.0) {
        double t = 2.0 - a;
        *outv = t * t * t / 6.0;
    } else
        *outv = (4.0 - 6.0 * a * a + 3.0 * a * a * a) / 6.0;
    return DMDX_OK;
}

int dmdx_pme_dtheta4(double u, double* outv)
{
    if (!outv) return dmdx::fail(DMDX_E_INVALID, "null out");
    double a = std::fabs(u);
    if (a >= 2.0) {
        *outv = 0.0;
        return DMDX_OK;
    }
    double s = (u >= 0.0) ? 1.0 : -1.0;
    if (a >= 1.0) {
        double t = 2.0 - a;
        *outv = -s * t * t / 2.0;
    } else
        *outv = s * (9.0 * a * a - 12.0 * a) / 6.0;
    return DMDX_OK;
}

int dmdx_pme_self_energy(double alpha, const double* charges, int64_t n, double k_e, double* outv)
{
    if (!outv || n < 0 || (n > 0 && !charges)) return dmdx::fail(DMDX_E_INVALID, "bad arguments");
    double sum_q2 = 0.0;
    for (int64_t i = 0; i < n; ++i) sum_q2 += charges[i] * charges[i];
    *outv = -k_e * alpha / std::sqrt(dmdx::PI) * sum_q2;
    return DMDX_OK;
}

}  // extern "C"
