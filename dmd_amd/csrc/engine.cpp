// engine.cpp -- host side of the step engine: packing systems into device
// buffers, the step sequence and the engine-level C-ABI.
//
// The step replaces SimulationEngine::step (src/sim/simulation_engine.cpp:
// 53-86): compute, kick + drift, compute, kick, thermostat, barostat.  One
// difference in what is launched, not in what is computed: the first compute
// of a step sees the same positions, list and parameters as the second
// compute of the previous step unless the list is rebuilt or a barostat moved
// the atoms, and then yields the same forces and energy, so it is skipped
// (the list counter still advances exactly as in the reference).
#include "engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

namespace dmdx {

static thread_local std::string g_last_error;

void set_error(int status, const std::string& msg)
{
    (void)status;
    g_last_error = msg;
}

int fail(int status, const std::string& msg)
{
    set_error(status, msg);
    return status;
}

const char* last_error() { return g_last_error.c_str(); }

static int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

// Region of n elements: a new allocation, or in arena mode the next region of
// the arena, cleared by a fill kernel (arena contents are left over from the
// previous user).
template <class T>
static int dev_region(T** dst, size_t n, DevMem& owned, bool zero)
{
    const size_t bytes = (n * sizeof(T) + 3) / 4 * 4;
    if (owned.arena) {
        void* p = owned.take(bytes);
        if (!p)
            return fail(DMDX_E_OOM, "the system needs more device scratch than the " +
                                        std::to_string(owned.cap) + "-byte array holds");
        *dst = (T*)p;
        if (zero) DMDX_HIP(launch_zero_bytes(p, bytes, owned.zero_stream));
        return DMDX_OK;
    }
    DMDX_HIP(hipMalloc((void**)dst, bytes));
    owned.add(*dst, bytes);
    if (zero) DMDX_HIP(hipMemset(*dst, 0, bytes));
    return DMDX_OK;
}

template <class T>
static int dev_upload(T** dst, const std::vector<T>& src, DevMem& owned)
{
    const int rc = dev_region(dst, std::max<size_t>(src.size(), 1), owned, src.empty());
    if (rc != DMDX_OK) return rc;
    if (!src.empty()) DMDX_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return DMDX_OK;
}

template <class T>
static int dev_alloc_zero(T** dst, size_t n, DevMem& owned)
{
    return dev_region(dst, std::max<size_t>(n, 1), owned, true);
}

#define TRY(expr)                 \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != DMDX_OK) return rc_; \
    } while (0)

// A mixed-path list that has (or, with a barostat, may come to have) fewer
// than 2 cdiv + 1 cells along an edge, as cells_in_use counts them on the
// device: such a system's list is built per atom with a minimum image.
static bool mixed_list_wraps(const SysDev& S)
{
    if (!S.has_list || S.precision == DMDX_FP64) return false;
    bool wraps = S.barostat != DMDX_BARO_NONE || S.cdiv > 2;
    for (int k = 0; k < 3; ++k) {
        int nu = (int)std::floor(S.L[k] * S.cdiv / (S.rlist * (1.0 + 1e-9)));
        nu = std::max(1, std::min(nu, S.nc[k]));
        wraps = wraps || nu < 2 * S.cdiv + 1;
    }
    return wraps;
}

// radices for a Stockham plan: 4s first, then 2, 3, 5, 7, then other primes
static bool plan_radices(int K, FftPlan& P)
{
    P.K = K;
    P.nrad = 0;
    int n = K;
    const int pref[] = {4, 2, 3, 5, 7, 11, 13};
    for (int r : pref)
        while (n % r == 0) {
            if (P.nrad >= 16) return false;
            P.radix[P.nrad++] = r;
            n /= r;
        }
    return n == 1;
}

Batch::~Batch()
{
    snap.reset();  // waits for frames in flight, which read the state arrays
    for (auto& u : ev_used) {
        (void)hipEventDestroy(u.second.first);
        (void)hipEventDestroy(u.second.second);
    }
    for (auto ev : ev_pool) (void)hipEventDestroy(ev);
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (void* p : owned.ptrs) (void)hipFree(p);
    if (borrowed) return;  // arena mode: the streams and events are the caller's
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (ev_list) (void)hipEventDestroy(ev_list);
    if (side) (void)hipStreamDestroy(side);
    if (stream) (void)hipStreamDestroy(stream);
}

int Batch::create(const dmdx_system_desc* descs, int nsys, const BatchOptions& opt)
{
    if (nsys <= 0) return fail(DMDX_E_INVALID, "no systems");
    options = opt;
    if (opt.arena) {
        if (opt.use_graph || opt.profile || !opt.arena_stream || !opt.arena_side ||
            !opt.arena_events[0] || !opt.arena_events[1] || !opt.arena_events[2])
            return fail(DMDX_E_INVALID, "an arena batch needs streams and events and runs eagerly");
        borrowed = true;
        stream = opt.arena_stream;
        side = opt.arena_side;
        ev_fork = opt.arena_events[0];
        ev_join = opt.arena_events[1];
        ev_list = opt.arena_events[2];
        owned.arena = (char*)opt.arena;
        owned.cap = opt.arena_bytes;
        owned.zero_stream = stream;
    } else {
        DMDX_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        DMDX_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        DMDX_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        DMDX_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
        DMDX_HIP(hipEventCreateWithFlags(&ev_list, hipEventDisableTiming));
    }
    const int prec = descs[0].precision;
    for (int s = 0; s < nsys; ++s)
        if (descs[s].precision != prec)
            return fail(DMDX_E_PARM, "all systems of a batch must share one precision");
    d.precision = prec;
    d.nsys = nsys;
    hsys.resize(nsys);
    hstate.assign(nsys, SysState{});
    host.resize(nsys);

    std::vector<int> slot_sys;
    std::vector<double> x, y, z, vx, vy, vz, invm, m, q;
    std::vector<int> type;
    std::vector<double> lj_sig, lj_eps;
    std::vector<float> c12, c6;
    std::vector<int> cell_sys;
    std::vector<int> row_sys;
    std::vector<int> b_sys, bi, bj, a_sys, ai, aj, ak, d_sys, di, dj, dk, dl, dper, i_sys, ii, ij,
        ik, il, e_sys, ei, ej;
    std::vector<double> bk, br0, akt, at0, dkp, dp0, ikp, ip0;
    std::vector<std::vector<int>> ex_of;  // per slot, corrected-mode list exclusions
    int64_t slot = 0, nbr_total = 0, grid_total = 0, tw_total = 0, bm_total = 0;
    int ncells = 0, lj_total = 0, pw_total = 0;
    std::vector<int> grid_sys;

    // ---- pass 1: atoms, parameters, offsets
    for (int s = 0; s < nsys; ++s) {
        const dmdx_system_desc& D = descs[s];
        SysDev& S = hsys[s];
        std::memset(&S, 0, sizeof(S));
        HostSys& H = host[s];
        if (D.n_atoms <= 0) return fail(DMDX_E_INVALID, "system with no atoms");
        if (!D.pos_x || !D.pos_y || !D.pos_z || !D.masses)
            return fail(DMDX_E_INVALID, "positions and masses are required");
        for (int k = 0; k < 3; ++k)
            if (!(D.box[k] > 0.0)) return fail(DMDX_E_PARM, "box edges must be positive");
        const int n = (int)D.n_atoms;
        const int n_pad = (int)round_up(n, WAVE);
        S.atom_off = slot;
        S.n = n;
        S.n_pad = n_pad;
        S.physics = D.physics;
        S.precision = D.precision;
        for (int k = 0; k < 3; ++k) {
            S.L[k] = D.box[k];
            S.invL[k] = 1.0 / D.box[k];
            S.iminv[k] = S.invL[k];
        }
        S.vdet = 0.0;
        const bool faithful = D.physics == DMDX_FAITHFUL;
        // pair terms: LJ always from the list; real-space erfc only in
        // CORRECTED (build_simulation never fills CoulombDirect's list)
        S.use_lj = D.use_lj ? 1 : 0;
        S.coul_pairs = (!faithful && D.coulomb_type != DMDX_COULOMB_NONE) ? 1 : 0;
        if (opt.component == COMP_COULOMB_DIRECT) S.coul_pairs = 1;
        else if (opt.component != COMP_NONE) S.coul_pairs = 0;  // e.g. PME: reciprocal part only
        S.ntypes = S.use_lj ? std::max(1, (int)D.lj_n_types) : 1;
        S.lj_off = lj_total;
        for (int t = 0; t < S.ntypes * S.ntypes; ++t) {
            double sg = S.use_lj ? D.lj_sigma[t] : 1.0;
            double ep = S.use_lj ? D.lj_epsilon[t] : 0.0;
            lj_sig.push_back(sg);
            lj_eps.push_back(ep);
            double s6 = sg * sg * sg * sg * sg * sg;
            c12.push_back((float)(4.0 * ep * s6 * s6));
            c6.push_back((float)(4.0 * ep * s6));
        }
        lj_total += S.ntypes * S.ntypes;
        S.rc_lj2 = D.lj_cutoff * D.lj_cutoff;
        S.rc_c2 = D.coulomb_cutoff * D.coulomb_cutoff;
        S.alpha = D.ewald_coeff;
        double rl = 0.0;
        if (S.use_lj) rl = D.lj_cutoff;
        if (S.coul_pairs) rl = std::max(rl, D.coulomb_cutoff);
        S.has_list = (S.use_lj || S.coul_pairs) ? 1 : 0;
        S.rlist = rl + D.list_skin;
        S.rlist2 = S.rlist * S.rlist;
        // rebuild_interval: > 0 cadence, 0 DMD's 10, -1 Verlet criterion,
        // <= -2 explicit rebuilds only
        if (D.rebuild_interval > 0) S.interval = D.rebuild_interval;
        else if (D.rebuild_interval == 0) S.interval = 10;
        else if (D.rebuild_interval == -1) S.interval = 0;
        else S.interval = -1;
        if (opt.component == COMP_COULOMB_DIRECT) S.interval = -1;
        S.half_skin2 = 0.25 * D.list_skin * D.list_skin;
        if (S.has_list) {
            double vol = D.box[0] * D.box[1] * D.box[2];
            double dens = n / vol;
            double expect = 4.0 / 3.0 * PI * S.rlist * S.rlist * S.rlist * dens;
            int M = (int)std::ceil(1.6 * expect + 48.0);
            M = std::min(M, n - 1);
            M = std::max(M, 1);
            if (opt.max_nbr > 0) M = std::min(std::max(opt.max_nbr, 1), std::max(n - 1, 1));
            M = (M + 7) / 8 * 8;  // rows hold whole 8-entry vector loads
            if (M > 2048)
                return fail(DMDX_E_PARM, "neighbour list would exceed 2048 entries per atom; "
                                         "reduce the cutoff");
            S.max_nbr = M;
            S.nbr_off = nbr_total;
            nbr_total += (int64_t)M * n_pad;
            // mixed path, dense systems (many atoms per list-radius cell, e.g.
            // water): cells of half the list radius, searched +-2, cut the
            // candidates by ~40%; sparse systems keep one cell per radius
            const double per_cell = n / (D.box[0] * D.box[1] * D.box[2]) * S.rlist * S.rlist * S.rlist;
            const int forced = (size_t)s < opt.cdiv.size() ? opt.cdiv[s] : 0;
            S.cdiv = D.precision == DMDX_FP64 ? 1
                     : (forced == 1 || forced == 2) ? forced
                     : (per_cell >= 40.0 ? 2 : 1);
            for (int k = 0; k < 3; ++k) {
                int c = (int)std::floor(D.box[k] * S.cdiv / (S.rlist * (1.0 + 1e-9)));
                // a barostat can grow the box: leave room for more cells
                if (D.barostat != DMDX_BARO_NONE) c = (int)std::floor(1.2 * c) + 1;
                S.nc[k] = std::max(1, c);
            }
            for (int k = 0; k < 3; ++k)
                if (D.box[k] < 2.0 * (S.use_lj ? D.lj_cutoff : D.coulomb_cutoff))
                    return fail(DMDX_E_PARM, "cutoff must be below half the box (minimum image)");
        } else {
            S.cdiv = 1;
            S.nc[0] = S.nc[1] = S.nc[2] = 1;
        }
        if (D.precision != DMDX_FP64 && (int64_t)n_pad > (int64_t)NBR_SLOT_MASK)
            return fail(DMDX_E_PARM, "a mixed-precision system holds at most 2^23 atoms");
        S.cell_off = ncells;
        int nc3 = S.nc[0] * S.nc[1] * S.nc[2];
        S.bin_wg = (S.has_list && n <= BIN_WG_ATOMS && nc3 <= BIN_WG_CELLS) ? 1 : 0;
        ncells += nc3;
        for (int c = 0; c < nc3; ++c) cell_sys.push_back(s);
        S.row_off = (int32_t)row_sys.size();
        for (int r = 0; r < S.nc[1] * S.nc[2]; ++r) row_sys.push_back(s);

        double qsum = 0.0, sum_q2 = 0.0;
        for (int i = 0; i < n_pad; ++i) {
            slot_sys.push_back(s);
            bool real = i < n;
            x.push_back(real ? D.pos_x[i] : 0.0);
            y.push_back(real ? D.pos_y[i] : 0.0);
            z.push_back(real ? D.pos_z[i] : 0.0);
            vx.push_back(real && D.vel_x ? D.vel_x[i] : 0.0);
            vy.push_back(real && D.vel_y ? D.vel_y[i] : 0.0);
            vz.push_back(real && D.vel_z ? D.vel_z[i] : 0.0);
            double mm = real ? D.masses[i] : 1.0;
            m.push_back(mm);
            invm.push_back(1.0 / mm);
            double qq = (real && D.charges) ? D.charges[i] : 0.0;
            q.push_back(qq);
            if (real) {
                qsum += qq;
                sum_q2 += qq * qq;
            }
            int tt = (real && D.atom_types) ? D.atom_types[i] : 0;
            if (S.use_lj && (tt < 0 || tt >= S.ntypes))
                return fail(DMDX_E_PARM, "atom type out of range of the LJ tables");
            type.push_back(tt);
        }
        S.qsum = qsum;
        // PME (src/force/coulomb_pme.cpp:71-75 for the self energy)
        S.pme = (D.coulomb_type == DMDX_COULOMB_PME || opt.component == COMP_PME) ? 1 : 0;
        if (opt.component != COMP_NONE && opt.component != COMP_PME) S.pme = 0;
        if (S.pme) {
            if (D.pme_order != 4 && D.pme_order != 0)
                return fail(DMDX_E_PARM, "only order-4 B-splines are implemented (as DMD)");
            for (int k = 0; k < 3; ++k) {
                S.K[k] = D.pme_grid[k];
                if (S.K[k] < 4 || S.K[k] > 256) return fail(DMDX_E_PARM, "PME grid must be 4..256");
                FftPlan P;
                if (!plan_radices(S.K[k], P))
                    return fail(DMDX_E_PARM, "PME grid size must factor into primes <= 13");
            }
            S.self_energy = -KE * D.ewald_coeff / std::sqrt(PI) * sum_q2;
            S.grid_off = grid_total;
            int64_t k3 = (int64_t)S.K[0] * S.K[1] * S.K[2];
            grid_total += k3;
            for (int64_t e = 0; e < k3; ++e) grid_sys.push_back(s);
            for (int k = 0; k < 3; ++k) {
                S.tw_off[k] = tw_total;
                tw_total += S.K[k];
                S.bm_off[k] = bm_total;
                bm_total += S.K[k];
            }
        }
        // integration (src/thermostat/nose_hoover.cpp:6-11,
        // src/barostat/andersen_barostat.cpp:8-14)
        S.dt = D.dt;
        if (D.thermostat < DMDX_THERMO_NONE || D.thermostat > DMDX_THERMO_ANDERSEN)
            return fail(DMDX_E_PARM, "unknown thermostat");
        if (D.barostat < DMDX_BARO_NONE || D.barostat > DMDX_BARO_ANDERSEN)
            return fail(DMDX_E_PARM, "unknown barostat");
        S.thermostat = D.thermostat;
        S.and_prob = D.andersen_frequency * D.dt;  // src/thermostat/andersen.cpp:15
        S.T0 = D.thermo_temperature;
        S.ttau = D.thermo_tau;
        double dof = 3.0 * (double)n - 3.0;
        if (dof < 1.0) dof = 1.0;
        S.Qnh = dof * KB * S.T0 * S.ttau * S.ttau;
        S.barostat = D.barostat;
        S.P0 = D.baro_pressure;
        S.ptau = D.baro_tau;
        S.beta = D.baro_compressibility;
        S.Wab = dof * KB * S.P0 * S.ptau * S.ptau;
        S.aw_off = (int)(slot / WAVE);
        S.aw_n = n_pad / WAVE;
        slot += n_pad;

        SysState& T = hstate[s];
        std::memset(&T, 0, sizeof(T));
        T.counter = D.step_since_rebuild;
        T.and_calls = D.andersen_calls;
        T.lambda = 1.0;
        T.active = 1;
        T.valid = 0;

        H.n = n;
        H.has_barostat = D.barostat != DMDX_BARO_NONE;
    }
    d.nslots = slot;
    d.ncells = ncells;

    // ---- pass 2: bonded terms (global slot indices), exclusions, contributions
    ex_of.assign(slot, {});
    auto pad_terms = [](int64_t cnt) { return round_up(cnt, WAVE); };
    auto check_atom = [](int a, int n) { return a >= 0 && a < n; };
    std::vector<std::vector<int64_t>> pre(slot), post(slot);  // per slot, term-force indices
    // term-force offsets need the padded totals first
    int64_t NB = 0, NA = 0, ND = 0, NI = 0, NE = 0;
    for (int s = 0; s < nsys; ++s) {
        const dmdx_system_desc& D = descs[s];
        bool excl_term = exclusion_term_active(D, opt);
        NB += D.n_bonds > 0 ? pad_terms(D.n_bonds) : 0;
        NA += D.n_angles > 0 ? pad_terms(D.n_angles) : 0;
        ND += D.n_dihedrals > 0 ? pad_terms(D.n_dihedrals) : 0;
        NI += D.n_impropers > 0 ? pad_terms(D.n_impropers) : 0;
        NE += (excl_term && D.n_excl > 0) ? pad_terms(D.n_excl) : 0;
    }
    d.tf_bond = 0;
    d.tf_angle = 2 * NB;
    d.tf_dih = d.tf_angle + 3 * NA;
    d.tf_imp = d.tf_dih + 4 * ND;
    d.tf_ex = d.tf_imp + 4 * NI;
    int64_t tf_total = d.tf_ex + 2 * NE;
    int64_t ob = 0, oa = 0, od = 0, oi = 0, oe = 0;
    for (int s = 0; s < nsys; ++s) {
        const dmdx_system_desc& D = descs[s];
        SysDev& S = hsys[s];
        const int n = S.n;
        const int64_t base = S.atom_off;
        bool comp_ok_bond = opt.component == COMP_NONE || opt.component == COMP_BOND;
        bool comp_ok_angle = opt.component == COMP_NONE || opt.component == COMP_ANGLE;
        bool comp_ok_dih = opt.component == COMP_NONE || opt.component == COMP_DIHEDRAL;
        bool comp_ok_imp = opt.component == COMP_NONE || opt.component == COMP_IMPROPER;
        // bonds
        S.bond_off = ob;
        S.n_bond = comp_ok_bond ? (int)D.n_bonds : 0;
        if (S.n_bond > 0) {
            int64_t np = pad_terms(S.n_bond);
            for (int64_t u = 0; u < np; ++u) {
                bool real = u < S.n_bond;
                int a = real ? D.bond_i[u] : 0, b = real ? D.bond_j[u] : 0;
                if (real && (!check_atom(a, n) || !check_atom(b, n)))
                    return fail(DMDX_E_PARM, "bond atom index out of range");
                b_sys.push_back(s);
                bi.push_back((int)(base + a));
                bj.push_back((int)(base + b));
                bk.push_back(real ? D.bond_k[u] : 0.0);
                br0.push_back(real ? D.bond_r0[u] : 1.0);
                if (real) {
                    pre[base + a].push_back(d.tf_bond + 2 * (ob + u));
                    pre[base + b].push_back(d.tf_bond + 2 * (ob + u) + 1);
                }
            }
            ob += np;
        }
        // angles
        S.angle_off = oa;
        S.n_angle = comp_ok_angle ? (int)D.n_angles : 0;
        if (S.n_angle > 0) {
            int64_t np = pad_terms(S.n_angle);
            for (int64_t u = 0; u < np; ++u) {
                bool real = u < S.n_angle;
                int a = real ? D.angle_i[u] : 0, b = real ? D.angle_j[u] : 1 % n,
                    c = real ? D.angle_k[u] : 2 % n;
                if (real && (!check_atom(a, n) || !check_atom(b, n) || !check_atom(c, n)))
                    return fail(DMDX_E_PARM, "angle atom index out of range");
                a_sys.push_back(s);
                ai.push_back((int)(base + a));
                aj.push_back((int)(base + b));
                ak.push_back((int)(base + c));
                akt.push_back(real ? D.angle_k_theta[u] : 0.0);
                at0.push_back(real ? D.angle_theta0[u] : 0.0);
                if (real) {
                    pre[base + a].push_back(d.tf_angle + 3 * (oa + u));
                    pre[base + b].push_back(d.tf_angle + 3 * (oa + u) + 1);
                    pre[base + c].push_back(d.tf_angle + 3 * (oa + u) + 2);
                }
            }
            oa += np;
        }
        // dihedrals
        S.dih_off = od;
        S.n_dih = comp_ok_dih ? (int)D.n_dihedrals : 0;
        if (S.n_dih > 0) {
            int64_t np = pad_terms(S.n_dih);
            for (int64_t u = 0; u < np; ++u) {
                bool real = u < S.n_dih;
                int a = real ? D.dih_i[u] : 0, b = real ? D.dih_j[u] : 0, c = real ? D.dih_k[u] : 0,
                    e = real ? D.dih_l[u] : 0;
                if (real && (!check_atom(a, n) || !check_atom(b, n) || !check_atom(c, n) ||
                             !check_atom(e, n)))
                    return fail(DMDX_E_PARM, "dihedral atom index out of range");
                d_sys.push_back(s);
                di.push_back((int)(base + a));
                dj.push_back((int)(base + b));
                dk.push_back((int)(base + c));
                dl.push_back((int)(base + e));
                dper.push_back(real ? D.dih_periodicity[u] : 1);
                dkp.push_back(real ? D.dih_k_phi[u] : 0.0);
                dp0.push_back(real ? D.dih_phi0[u] : 0.0);
                if (real)
                    for (int r = 0; r < 4; ++r)
                        pre[base + (r == 0 ? a : r == 1 ? b : r == 2 ? c : e)].push_back(
                            d.tf_dih + 4 * (od + u) + r);
            }
            od += np;
        }
        // impropers
        S.imp_off = oi;
        S.n_imp = comp_ok_imp ? (int)D.n_impropers : 0;
        if (S.n_imp > 0) {
            int64_t np = pad_terms(S.n_imp);
            for (int64_t u = 0; u < np; ++u) {
                bool real = u < S.n_imp;
                int a = real ? D.imp_i[u] : 0, b = real ? D.imp_j[u] : 0, c = real ? D.imp_k[u] : 0,
                    e = real ? D.imp_l[u] : 0;
                if (real && (!check_atom(a, n) || !check_atom(b, n) || !check_atom(c, n) ||
                             !check_atom(e, n)))
                    return fail(DMDX_E_PARM, "improper atom index out of range");
                i_sys.push_back(s);
                ii.push_back((int)(base + a));
                ij.push_back((int)(base + b));
                ik.push_back((int)(base + c));
                il.push_back((int)(base + e));
                ikp.push_back(real ? D.imp_k_phi[u] : 0.0);
                ip0.push_back(real ? D.imp_phi0[u] : 0.0);
                if (real)
                    for (int r = 0; r < 4; ++r)
                        pre[base + (r == 0 ? a : r == 1 ? b : r == 2 ? c : e)].push_back(
                            d.tf_imp + 4 * (oi + u) + r);
            }
            oi += np;
        }
        // exclusions: list filter (CORRECTED) and the exclusion term
        for (int64_t e = 0; e < D.n_excl; ++e) {
            int a = D.excl_i[e], b = D.excl_j[e];
            if (!check_atom(a, n) || !check_atom(b, n))
                return fail(DMDX_E_PARM, "exclusion atom index out of range");
        }
        if (D.physics != DMDX_FAITHFUL && S.has_list && opt.component == COMP_NONE) {
            for (int64_t e = 0; e < D.n_excl; ++e) {
                int a = D.excl_i[e], b = D.excl_j[e];
                ex_of[base + a].push_back((int)(base + b));
                ex_of[base + b].push_back((int)(base + a));
            }
        }
        S.ex_off = oe;
        S.n_ex = exclusion_term_active(D, opt) ? (int)D.n_excl : 0;
        S.ex_mode = S.n_ex > 0;
        if (S.n_ex > 0) {
            int64_t np = pad_terms(S.n_ex);
            for (int64_t u = 0; u < np; ++u) {
                bool real = u < S.n_ex;
                int a = real ? D.excl_i[u] : 0, b = real ? D.excl_j[u] : 1 % n;
                e_sys.push_back(s);
                ei.push_back((int)(base + a));
                ej.push_back((int)(base + b));
                if (real) {
                    post[base + a].push_back(d.tf_ex + 2 * (oe + u));
                    post[base + b].push_back(d.tf_ex + 2 * (oe + u) + 1);
                }
            }
            oe += np;
        }
        S.bw_off = (int)(S.bond_off / WAVE);
        S.bw_n = (int)(S.n_bond > 0 ? pad_terms(S.n_bond) / WAVE : 0);
        S.gw_off = (int)(S.angle_off / WAVE);
        S.gw_n = (int)(S.n_angle > 0 ? pad_terms(S.n_angle) / WAVE : 0);
        S.dw_off = (int)(S.dih_off / WAVE);
        S.dw_n = (int)(S.n_dih > 0 ? pad_terms(S.n_dih) / WAVE : 0);
        S.iw_off = (int)(S.imp_off / WAVE);
        S.iw_n = (int)(S.n_imp > 0 ? pad_terms(S.n_imp) / WAVE : 0);
        S.xw_off = (int)(S.ex_off / WAVE);
        S.xw_n = (int)(S.n_ex > 0 ? pad_terms(S.n_ex) / WAVE : 0);
    }
    d.nb = ob;
    d.na = oa;
    d.nd = od;
    d.ni = oi;
    d.ne = oe;
    std::vector<int> c_start(slot + 1, 0), c_mid(slot, 0), contrib;
    std::vector<int> ex_start(slot + 1, 0), ex_list;
    for (int64_t g = 0; g < slot; ++g) {
        c_start[g] = (int)contrib.size();
        for (int64_t v : pre[g]) contrib.push_back((int)v);
        c_mid[g] = (int)contrib.size();
        for (int64_t v : post[g]) contrib.push_back((int)v);
        ex_start[g] = (int)ex_list.size();
        for (int v : ex_of[g]) ex_list.push_back(v);
    }
    c_start[slot] = (int)contrib.size();
    ex_start[slot] = (int)ex_list.size();

    // ---- PME tables, plans and tiles
    std::vector<double> twd, bmod;
    std::vector<FftPlan> plans(3 * nsys);
    std::vector<FftTile> tiles[3];
    std::vector<PlaneTile> ptiles;
    const size_t cx = prec == DMDX_FP64 ? 16 : 8;
    // mixed path: an xy plane (two buffers and the x and y twiddles) within
    // 64 KiB of LDS can be transformed along x and y by one workgroup.  That
    // saves two launches and a grid round trip per direction, but gives a
    // system only K_z workgroups: by default (opt.fft_planes 2) planes are
    // used when the batch has at least one per CU (a single 48^3 system is
    // faster on line tiles, a served cohort of several water boxes on
    // planes).
    auto plane_bytes = [&](const SysDev& S) {
        return (2 * (size_t)S.K[0] * S.K[1] + S.K[0] + S.K[1]) * cx;
    };
    int64_t plane_count = 0;
    for (int s = 0; s < nsys; ++s)
        if (hsys[s].pme && prec != DMDX_FP64 && plane_bytes(hsys[s]) <= 64 * 1024)
            plane_count += hsys[s].K[2];
    const bool use_planes =
        opt.fft_planes == 1 ? plane_count > 0 : (opt.fft_planes == 2 && plane_count >= 256);
    for (int s = 0; s < nsys; ++s) {
        SysDev& S = hsys[s];
        if (!S.pme) continue;
        for (int k = 0; k < 3; ++k) {
            plan_radices(S.K[k], plans[3 * s + k]);
            for (int j = 0; j < S.K[k]; ++j) {
                double ang = -2.0 * PI * (double)j / (double)S.K[k];
                twd.push_back(std::cos(ang));
                twd.push_back(std::sin(ang));
            }
            for (int j = 0; j < S.K[k]; ++j) {
                int mm = (j <= S.K[k] / 2) ? j : j - S.K[k];
                double B;
                if (S.physics == DMDX_FAITHFUL) {
                    // src/force/coulomb_pme.cpp:159-163
                    if (mm == 0) B = 1.0;
                    else {
                        double t = PI * std::abs(mm) / S.K[k];
                        B = std::pow(std::sin(t) / t, 2 * 4);
                    }
                } else {
                    double dd = 2.0 / 3.0 + std::cos(2.0 * PI * mm / S.K[k]) / 3.0;
                    B = 1.0 / (dd * dd);
                }
                bmod.push_back(B);
            }
        }
        const size_t plane = plane_bytes(S);
        const bool planes = use_planes && prec != DMDX_FP64 && plane <= 64 * 1024;
        if (planes) {
            for (int z = 0; z < S.K[2]; ++z) ptiles.push_back(PlaneTile{s, z});
            d.plane_lds = std::max(d.plane_lds, plane);
        }
        // tiles: lines along dim k are indexed by the other two coordinates
        S.pw_off = pw_total;
        for (int k = 0; k < 3; ++k) {
            if (planes && k < 2) continue;
            int K = S.K[k];
            int nlines = (k == 0) ? S.K[1] * S.K[2] : (k == 1 ? S.K[0] * S.K[2] : S.K[0] * S.K[1]);
            // lines per workgroup: enough workgroups to spread a small grid
            // over the chip (about nlines/256, a power of two in 4..16: 16
            // lines are whole 128-byte runs of the strided y and z passes in
            // FP32 and keep several workgroups per CU on large grids), within
            // 48 KiB of LDS for the two ping-pong buffers and the twiddles
            int B = 4;
            while (B < 16 && 2 * B * 256 <= nlines) B *= 2;
            if (opt.fft_lines > 0) B = opt.fft_lines;  // the autotuner's choice
            B = (int)std::max<size_t>(
                1, std::min<size_t>((size_t)B, (48 * 1024 - (size_t)K * cx) / (2 * K * cx)));
            int local = 0;
            for (int l0 = 0; l0 < nlines; l0 += B) {
                FftTile t;
                t.sys = s;
                t.line0 = l0;
                t.nlines = std::min(B, nlines - l0);
                t.tile_local = local++;
                tiles[k].push_back(t);
                d.lds_bytes[k] = std::max(d.lds_bytes[k], (size_t)((2 * t.nlines + 1) * K * cx));
            }
            if (k == 2) {
                S.pw_n = local;
                pw_total += local;
            }
        }
    }
    for (int k = 0; k < 3; ++k) d.ntiles[k] = (int)tiles[k].size();
    d.nptiles = (int)ptiles.size();
    // mixed path: the influence function of a fixed box, tabulated once
    // (same expression as the on-the-fly FP64 form in k_fft_pass)
    std::vector<float> ginf;
    if (prec != DMDX_FP64) {
        ginf.assign((size_t)std::max<int64_t>(grid_total, 1), 0.0f);
        for (int s = 0; s < nsys; ++s) {
            SysDev& S = hsys[s];
            if (!S.pme || S.barostat != DMDX_BARO_NONE) continue;
            S.ginf_static = 1;
            const int Kx = S.K[0], Ky = S.K[1], Kz = S.K[2];
            const double Lx = S.L[0], Ly = S.L[1], Lz = S.L[2];
            const double V = sys_volume(S);
            const double* bmx = bmod.data() + S.bm_off[0];
            const double* bmy = bmod.data() + S.bm_off[1];
            const double* bmz = bmod.data() + S.bm_off[2];
            for (int kz = 0; kz < Kz; ++kz)
                for (int ky = 0; ky < Ky; ++ky)
                    for (int kx = 0; kx < Kx; ++kx) {
                        int mx = (kx <= Kx / 2) ? kx : kx - Kx;
                        int my = (ky <= Ky / 2) ? ky : ky - Ky;
                        int mz = (kz <= Kz / 2) ? kz : kz - Kz;
                        double green = 0.0;
                        if (mx != 0 || my != 0 || mz != 0) {
                            double msq = ((double)mx * mx) / (Lx * Lx) +
                                         ((double)my * my) / (Ly * Ly) +
                                         ((double)mz * mz) / (Lz * Lz);
                            double k2 = 4.0 * PI * PI * msq;
                            double B = bmx[kx] * bmy[ky] * bmz[kz];
                            green = 4.0 * PI * KE / V * std::exp(-k2 / (4.0 * S.alpha * S.alpha)) *
                                    B / k2;
                        }
                        ginf[S.grid_off + ((int64_t)kz * Ky + ky) * Kx + kx] = (float)green;
                    }
        }
    }

    // ---- SHAKE / RATTLE constraints (src/integrate/constraints.cpp): group
    // each system's constraints by connected component of the atoms they
    // link, keeping the caller's order inside a component; components are
    // ordered by their first constraint.  The reference's sweeps touch
    // different components' atoms independently, so one thread per component
    // reproduces them exactly.
    std::vector<int> k_sys, k_start, k_i, k_j;
    std::vector<double> k_d;
    for (int s = 0; s < nsys; ++s) {
        const dmdx_system_desc& D = descs[s];
        SysDev& S = hsys[s];
        if (D.n_constraints <= 0 || opt.component != COMP_NONE) continue;
        if (!D.con_i || !D.con_j || !D.con_distance)
            return fail(DMDX_E_INVALID, "constraint arrays are required");
        if (D.n_constraints > (int64_t)1 << 30) return fail(DMDX_E_PARM, "too many constraints");
        const int n = S.n;
        const int64_t nc = D.n_constraints;
        std::vector<int> parent(n);
        for (int a = 0; a < n; ++a) parent[a] = a;
        auto root = [&](int a) {
            while (parent[a] != a) a = parent[a] = parent[parent[a]];
            return a;
        };
        for (int64_t c = 0; c < nc; ++c) {
            const int a = D.con_i[c], b = D.con_j[c];
            if (a < 0 || a >= n || b < 0 || b >= n || a == b)
                return fail(DMDX_E_PARM, "constraint atom index out of range or repeated");
            if (!(D.con_distance[c] > 0.0))
                return fail(DMDX_E_PARM, "constraint distances must be positive");
            const int ra = root(a), rb = root(b);
            if (ra != rb) parent[std::max(ra, rb)] = std::min(ra, rb);
        }
        std::vector<int> comp_of(n, -1);
        std::vector<std::vector<int64_t>> members;
        for (int64_t c = 0; c < nc; ++c) {
            const int r = root(D.con_i[c]);
            if (comp_of[r] < 0) {
                comp_of[r] = (int)members.size();
                members.emplace_back();
            }
            members[comp_of[r]].push_back(c);
        }
        for (const auto& mem : members) {
            k_sys.push_back(s);
            k_start.push_back((int)k_i.size());
            for (int64_t c : mem) {
                k_i.push_back((int)(S.atom_off + D.con_i[c]));
                k_j.push_back((int)(S.atom_off + D.con_j[c]));
                k_d.push_back(D.con_distance[c]);
            }
        }
        S.n_con = (int32_t)nc;
        S.con_tol = D.con_tolerance;
    }
    d.ncomp = (int)k_sys.size();
    k_start.push_back((int)k_i.size());
    if (d.ncomp > 0) {
        TRY(dev_upload(&d.k_sys, k_sys, owned));
        TRY(dev_upload(&d.k_start, k_start, owned));
        TRY(dev_upload(&d.k_i, k_i, owned));
        TRY(dev_upload(&d.k_j, k_j, owned));
        TRY(dev_upload(&d.k_d, k_d, owned));
    }

    // ---- what exists
    for (int s = 0; s < nsys; ++s) {
        const SysDev& S = hsys[s];
        d.any_list |= S.has_list != 0;
        d.any_pme |= S.pme != 0;
        d.any_coul_pairs |= S.coul_pairs != 0;
        d.any_plain_pairs |= S.coul_pairs == 0;
        d.any_plain_nolist |= S.coul_pairs == 0 && !S.has_list;
        if (S.has_list) {
            if (S.bin_wg) ++d.n_bin_wg;
            else d.any_bin_multi = true;
            d.any_wrap_mixed = d.any_wrap_mixed || mixed_list_wraps(S);
        }
        d.any_excl |= S.n_ex > 0;
        d.any_baro |= S.barostat != DMDX_BARO_NONE;
        d.any_andersen |= S.thermostat == DMDX_THERMO_ANDERSEN;
        host[s].has_list = S.has_list != 0;
        host[s].interval = S.interval;
        host[s].counter = hstate[s].counter;
        d.any_auto |= S.has_list && S.interval == 0;
    }
    d.any_gather = d.nb || d.na || d.nd || d.ni || d.ne || d.any_pme;

    // ---- device allocations
    TRY(dev_upload(&d.sys, hsys, owned));
    TRY(dev_upload(&d.st, hstate, owned));
    TRY(dev_upload(&d.slot_sys, slot_sys, owned));
    TRY(dev_upload(&d.x, x, owned));
    TRY(dev_upload(&d.y, y, owned));
    TRY(dev_upload(&d.z, z, owned));
    TRY(dev_upload(&d.vx, vx, owned));
    TRY(dev_upload(&d.vy, vy, owned));
    TRY(dev_upload(&d.vz, vz, owned));
    TRY(dev_alloc_zero(&d.fx, slot, owned));
    TRY(dev_alloc_zero(&d.fy, slot, owned));
    TRY(dev_alloc_zero(&d.fz, slot, owned));
    if (d.precision != DMDX_FP64 && d.any_gather) {
        TRY(dev_alloc_zero(&d.fbx, slot, owned));
        TRY(dev_alloc_zero(&d.fby, slot, owned));
        TRY(dev_alloc_zero(&d.fbz, slot, owned));
    }
    TRY(dev_upload(&d.invm, invm, owned));
    TRY(dev_upload(&d.m, m, owned));
    TRY(dev_upload(&d.q, q, owned));
    TRY(dev_upload(&d.type, type, owned));
    TRY(dev_alloc_zero(&d.acell, slot, owned));
    TRY(dev_alloc_zero(&d.ccount, ncells, owned));
    TRY(dev_alloc_zero(&d.cstart, ncells, owned));
    TRY(dev_alloc_zero(&d.cfill, ncells, owned));
    TRY(dev_alloc_zero(&d.cell_atoms, slot, owned));
    TRY(dev_upload(&d.cell_sys, cell_sys, owned));
    d.nrows = (int)row_sys.size();
    TRY(dev_upload(&d.row_sys, row_sys, owned));
    TRY(dev_alloc_zero(&d.rb_sys, (size_t)nsys, owned));
    TRY(dev_alloc_zero(&d.rb_pre, (size_t)nsys + 1, owned));
    TRY(dev_alloc_zero(&d.rb_cpre, (size_t)nsys + 1, owned));
    TRY(dev_alloc_zero(&d.rb_n, 1, owned));
    TRY(dev_alloc_zero(&d.rbm_sys, (size_t)nsys, owned));
    TRY(dev_alloc_zero(&d.rbm_pre, (size_t)nsys + 1, owned));
    TRY(dev_alloc_zero(&d.rbm_cpre, (size_t)nsys + 1, owned));
    TRY(dev_alloc_zero(&d.rbm_n, 1, owned));
    TRY(dev_alloc_zero(&d.rbw_sys, (size_t)nsys, owned));
    TRY(dev_alloc_zero(&d.rbw_n, 1, owned));
    {
        std::vector<int> ident(slot);
        for (int64_t g = 0; g < slot; ++g) ident[g] = (int)g;
        TRY(dev_upload(&d.order, ident, owned));
    }
    TRY(dev_alloc_zero(&d.nn, slot, owned));
    TRY(dev_alloc_zero(&d.nbr, nbr_total, owned));
    TRY(dev_upload(&d.ex_start, ex_start, owned));
    TRY(dev_upload(&d.ex_list, ex_list, owned));
    TRY(dev_upload(&d.lj_sig, lj_sig, owned));
    TRY(dev_upload(&d.lj_eps, lj_eps, owned));
    TRY(dev_upload(&d.lj_c12, c12, owned));
    TRY(dev_upload(&d.lj_c6, c6, owned));
    TRY(dev_upload(&d.b_sys, b_sys, owned));
    TRY(dev_upload(&d.bi, bi, owned));
    TRY(dev_upload(&d.bj, bj, owned));
    TRY(dev_upload(&d.bk, bk, owned));
    TRY(dev_upload(&d.br0, br0, owned));
    TRY(dev_upload(&d.a_sys, a_sys, owned));
    TRY(dev_upload(&d.ai, ai, owned));
    TRY(dev_upload(&d.aj, aj, owned));
    TRY(dev_upload(&d.ak, ak, owned));
    TRY(dev_upload(&d.akt, akt, owned));
    TRY(dev_upload(&d.at0, at0, owned));
    TRY(dev_upload(&d.d_sys, d_sys, owned));
    TRY(dev_upload(&d.di, di, owned));
    TRY(dev_upload(&d.dj, dj, owned));
    TRY(dev_upload(&d.dk, dk, owned));
    TRY(dev_upload(&d.dl, dl, owned));
    TRY(dev_upload(&d.dper, dper, owned));
    TRY(dev_upload(&d.dkp, dkp, owned));
    TRY(dev_upload(&d.dp0, dp0, owned));
    TRY(dev_upload(&d.i_sys, i_sys, owned));
    TRY(dev_upload(&d.ii, ii, owned));
    TRY(dev_upload(&d.ij, ij, owned));
    TRY(dev_upload(&d.ik, ik, owned));
    TRY(dev_upload(&d.il, il, owned));
    TRY(dev_upload(&d.ikp, ikp, owned));
    TRY(dev_upload(&d.ip0, ip0, owned));
    TRY(dev_upload(&d.e_sys, e_sys, owned));
    TRY(dev_upload(&d.ei, ei, owned));
    TRY(dev_upload(&d.ej, ej, owned));
    TRY(dev_alloc_zero(&d.tf, tf_total, owned));
    TRY(dev_upload(&d.c_start, c_start, owned));
    TRY(dev_upload(&d.c_mid, c_mid, owned));
    TRY(dev_upload(&d.contrib, contrib, owned));
    d.grid_total = grid_total;
    TRY(dev_alloc_zero(&d.qgrid, grid_total, owned));
    if (prec == DMDX_FP64) {
        TRY(dev_alloc_zero((double**)&d.grid, 2 * grid_total, owned));
        TRY(dev_upload((double**)&d.tw, twd, owned));
    } else {
        std::vector<float> twf(twd.begin(), twd.end());
        TRY(dev_alloc_zero((float**)&d.grid, 2 * grid_total, owned));
        TRY(dev_upload((float**)&d.tw, twf, owned));
        TRY(dev_upload(&d.ginf, ginf, owned));
    }
    TRY(dev_upload(&d.grid_sys, grid_sys, owned));
    {
        // owner-computes spreading where atoms stay within skin/2 of their
        // build cells (mixed path, Verlet criterion, fixed box); 8^3 blocks
        std::vector<SpreadTile> st_tiles;
        std::vector<GridTile> g_tiles;
        for (int s = 0; s < nsys; ++s) {
            const SysDev& S = hsys[s];
            if (!S.pme) continue;
            bool gather = prec != DMDX_FP64 && S.has_list && S.interval == 0 &&
                          S.barostat == DMDX_BARO_NONE && opt.component == COMP_NONE;
            if (gather) {  // a block scans at most 255 cells (k_pme_spread_gather)
                int64_t cells = 1;
                for (int k = 0; k < 3; ++k) {
                    int nu = (int)std::floor(S.L[k] * S.cdiv / (S.rlist * (1.0 + 1e-9)));
                    nu = std::max(1, std::min(nu, S.nc[k]));
                    const double w = S.L[k] / nu;
                    const double span = 11.0 * S.L[k] / S.K[k] + 2.0 * std::sqrt(S.half_skin2);
                    cells *= std::min<int64_t>(nu, (int64_t)std::floor(span / w) + 2);
                }
                gather = cells <= 255;
            }
            if (gather) {
                for (int gz = 0; gz < S.K[2]; gz += 8)
                    for (int gy = 0; gy < S.K[1]; gy += 8)
                        for (int gx = 0; gx < S.K[0]; gx += 8) {
                            GridTile G;
                            G.sys = s;
                            G.g0[0] = gx;
                            G.g0[1] = gy;
                            G.g0[2] = gz;
                            g_tiles.push_back(G);
                        }
                continue;
            }
            for (int64_t t0 = 0; t0 < S.n; t0 += 256) {
                SpreadTile T;
                T.sys = s;
                T.t0 = S.atom_off + t0;
                T.cnt = (int32_t)std::min<int64_t>(256, S.n - t0);
                st_tiles.push_back(T);
            }
        }
        d.nstiles = (int)st_tiles.size();
        TRY(dev_upload(&d.stiles, st_tiles, owned));
        d.ngtiles = (int)g_tiles.size();
        TRY(dev_upload(&d.gtiles, g_tiles, owned));
    }
    TRY(dev_upload(&d.bmod, bmod, owned));
    TRY(dev_upload(&d.plans, plans, owned));
    for (int k = 0; k < 3; ++k) TRY(dev_upload(&d.tiles[k], tiles[k], owned));
    TRY(dev_upload(&d.ptiles, ptiles, owned));
    const int64_t nw = slot / WAVE;
    TRY(dev_alloc_zero(&d.p_lj, slot, owned));
    TRY(dev_alloc_zero(&d.p_coul, slot, owned));
    TRY(dev_alloc_zero(&d.pos4s, slot, owned));
    TRY(dev_alloc_zero(&d.xref, slot, owned));
    TRY(dev_alloc_zero(&d.p_disp, nw, owned));
    if (prec != DMDX_FP64) {
        TRY(dev_alloc_zero(&d.p_rows, slot, owned));
        TRY(dev_alloc_zero(&d.typ_s, slot, owned));
        TRY(dev_alloc_zero(&d.pos4f, slot, owned));
        TRY(dev_alloc_zero(&d.org, slot, owned));
        TRY(dev_alloc_zero(&d.lpos, slot, owned));
        TRY(dev_alloc_zero(&d.slot_of, slot, owned));
    }
    TRY(dev_alloc_zero(&d.p_ke, nw, owned));
    TRY(dev_alloc_zero(&d.p_w, nw, owned));
    TRY(dev_alloc_zero(&d.p_bond, std::max<int64_t>(ob / WAVE, 1), owned));
    TRY(dev_alloc_zero(&d.p_angle, std::max<int64_t>(oa / WAVE, 1), owned));
    TRY(dev_alloc_zero(&d.p_dih, std::max<int64_t>(od / WAVE, 1), owned));
    TRY(dev_alloc_zero(&d.p_imp, std::max<int64_t>(oi / WAVE, 1), owned));
    TRY(dev_alloc_zero(&d.p_excl, std::max<int64_t>(oe / WAVE, 1), owned));
    TRY(dev_alloc_zero(&d.p_pme, std::max(pw_total, 1), owned));
    DMDX_HIP(hipDeviceSynchronize());
    return DMDX_OK;
}

bool Batch::exclusion_term_active(const dmdx_system_desc& D, const BatchOptions& opt)
{
    if (opt.component == COMP_EXCLUSION) return true;
    if (opt.component != COMP_NONE) return false;
    // build_force_engine (src/sim/simulation_engine.cpp:168-200): PME always
    // adds CoulombExclusion; "cutoff" adds it when exclusions exist.
    // CORRECTED: only PME needs the reciprocal-space correction (excluded
    // pairs are removed from the real-space list instead).
    if (D.physics == DMDX_FAITHFUL)
        return D.coulomb_type == DMDX_COULOMB_PME ||
               (D.coulomb_type == DMDX_COULOMB_CUTOFF && D.n_excl > 0);
    return D.coulomb_type == DMDX_COULOMB_PME;
}

#define LAUNCH(expr)                                                          \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) return fail(DMDX_E_GPU, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// host mirror of the per-system list counters (the device keeps its own
// authoritative copy); decides whether any list kernel has to be launched
bool Batch::mirror_compute(int phase, int force_rebuild, bool* any_force)
{
    bool any_rebuild = false;
    *any_force = false;
    for (size_t s = 0; s < host.size(); ++s) {
        HostSys& H = host[s];
        if (!H.active) continue;
        bool rebuild = false, maybe = false;
        if (H.has_list) {
            if (force_rebuild) rebuild = true;
            else if (H.interval > 0) rebuild = H.counter >= H.interval;
            else if (H.interval == 0) maybe = !H.built || !(phase == 1 && H.valid);
        }
        bool reuse = phase == 1 && options.reuse_forces && H.valid && !rebuild && !maybe;
        if (H.has_list && H.interval > 0) H.counter = (rebuild ? 0 : H.counter) + 1;
        // the device builds a list at the first compute of every system
        if (H.has_list) H.built = true;
        H.valid = true;
        any_rebuild |= rebuild || maybe;
        *any_force |= !reuse;
    }
    return any_rebuild;
}

static int g_acc = 0;
static hipError_t launch_pair_acc(const Dev& d, hipStream_t s) { return launch_pair(d, s, g_acc); }

// launch fn on stream `on` (default: the batch stream), bracketed by a pair
// of HIP events on that stream when profiling is on (events are read back by
// kernel_time)
hipError_t Batch::timed(int group, hipError_t (*fn)(const Dev&, hipStream_t), hipStream_t on)
{
    if (!on) on = stream;
    if (!options.profile || !(options.profile_mask >> group & 1u)) return fn(d, on);
    hipEvent_t a, b;
    for (hipEvent_t* ev : {&a, &b}) {
        if (!ev_pool.empty()) {
            *ev = ev_pool.back();
            ev_pool.pop_back();
        } else {
            hipError_t e = hipEventCreate(ev);
            if (e != hipSuccess) return e;
        }
    }
    hipError_t e = hipEventRecord(a, on);
    if (e != hipSuccess) return e;
    e = fn(d, on);
    if (e != hipSuccess) return e;
    e = hipEventRecord(b, on);
    ev_used.push_back({group, {a, b}});
    return e;
}

int Batch::kernel_time(int group, double* total_ms, int64_t* launches)
{
    if (group < 0 || group >= KG_COUNT) return fail(DMDX_E_INVALID, "unknown kernel group");
    TRY(sync());
    for (auto& u : ev_used) {
        float ms = 0.0f;
        DMDX_HIP(hipEventElapsedTime(&ms, u.second.first, u.second.second));
        kg_ms[u.first] += ms;
        kg_n[u.first] += 1;
        ev_pool.push_back(u.second.first);
        ev_pool.push_back(u.second.second);
    }
    ev_used.clear();
    *total_ms = kg_ms[group];
    *launches = kg_n[group];
    return DMDX_OK;
}

int Batch::enqueue_compute(int phase, int accumulate, int force_rebuild, bool uniform,
                           bool* defer_pre)
{
    bool any_force = true, any_rebuild = true;
    if (!uniform) any_rebuild = mirror_compute(phase, force_rebuild, &any_force);
    else {
        bool dummy;
        mirror_compute(phase, force_rebuild, &dummy);
    }
    // a compute that only reuses forces (no system rebuilds or computes)
    // leaves its per-system decisions to the caller's next kernel
    if (defer_pre && !any_force) {
        *defer_pre = true;
        return DMDX_OK;
    }
    // after the drift, the mixed path has its Verlet check and frames from
    // k_kick_drift already, unless SHAKE moved atoms since
    const bool after_drift = phase == 2 && d.precision != DMDX_FP64 && d.ncomp == 0;
    if (d.any_auto && any_rebuild && !after_drift) LAUNCH(launch_displacement(d, stream));
    LAUNCH(launch_pre_compute(d, stream, phase, force_rebuild, options.reuse_forces));
    if (!any_force) return DMDX_OK;
    // The bonded terms and PME read the positions and write their own term
    // forces, grid and energy partials; the list build and the pair kernel
    // write the lists, cells, frames and per-atom forces.  So the bonded
    // terms run on the side stream while the list build and the pair kernel
    // run here (a rebuild often leaves most of the device idle), PME follows
    // them there once the list build has left the cell order its charge
    // spreading reads, and the gather, which reads everything, waits for the
    // side stream.  On the mixed path a step's closing compute gathers the
    // term forces on the side stream too, into their own arrays, and the
    // closing kick adds them to the pair forces.
    const bool bonded = d.nb || d.na || d.nd || d.ni || d.ne;
    const bool fork = bonded || d.any_pme;
    const bool defer = phase == 2 && d.fbx != nullptr;
    if (fork) {
        DMDX_HIP(hipEventRecord(ev_fork, stream));
        DMDX_HIP(hipStreamWaitEvent(side, ev_fork, 0));
        if (bonded) LAUNCH(timed(KG_BONDED, launch_bonded, side));
    }
    if (d.any_list && any_rebuild) LAUNCH(timed(KG_LIST, launch_build_lists));
    if (d.any_pme) {
        DMDX_HIP(hipEventRecord(ev_list, stream));
        DMDX_HIP(hipStreamWaitEvent(side, ev_list, 0));
        LAUNCH(timed(KG_PME, launch_pme, side));
    }
    if (defer) LAUNCH(timed(KG_GATHER, launch_gather_terms, side));
    if (fork) DMDX_HIP(hipEventRecord(ev_join, side));
    g_acc = accumulate;
    // after the drift, k_kick_drift (kept lists) and k_cell_sort (rebuilt
    // lists) have written every frame position already
    if (!after_drift) LAUNCH(launch_pack_phase(d, stream, phase == 2 ? 1 : 0));
    LAUNCH(timed(KG_PAIR, launch_pair_acc));
    if (fork) DMDX_HIP(hipStreamWaitEvent(stream, ev_join, 0));
    if (d.any_gather && !defer) LAUNCH(timed(KG_GATHER, launch_gather));
    return DMDX_OK;
}

static hipError_t launch_kick_drift_plain(const Dev& d, hipStream_t s)
{
    return launch_kick_drift(d, s, 0);
}
static hipError_t launch_kick_drift_pre(const Dev& d, hipStream_t s)
{
    return launch_kick_drift(d, s, 1);
}
static hipError_t launch_kick_drift_pre_reuse(const Dev& d, hipStream_t s)
{
    return launch_kick_drift(d, s, 2);
}

int Batch::enqueue_step(bool uniform)
{
    // eager steps whose first compute only reuses forces take its decisions
    // in the drift kernel (one launch fewer per step)
    bool deferred = false;
    TRY(enqueue_compute(1, 0, 0, uniform, uniform ? nullptr : &deferred));
    if (!deferred) LAUNCH(timed(KG_INTEGRATE, launch_kick_drift_plain));
    else if (options.reuse_forces) LAUNCH(timed(KG_INTEGRATE, launch_kick_drift_pre_reuse));
    else LAUNCH(timed(KG_INTEGRATE, launch_kick_drift_pre));
    if (d.ncomp) LAUNCH(timed(KG_INTEGRATE, launch_shake));
    TRY(enqueue_compute(2, 0, 0, uniform));
    LAUNCH(timed(KG_INTEGRATE, launch_kick_ke));
    if (d.ncomp) LAUNCH(timed(KG_INTEGRATE, launch_rattle));
    if (d.any_andersen) LAUNCH(timed(KG_INTEGRATE, launch_andersen));
    if (d.ncomp || d.any_andersen) LAUNCH(launch_ke_refresh(d, stream));
    LAUNCH(launch_finalize(d, stream, 1));
    if (d.any_baro) {
        LAUNCH(launch_barostat(d, stream));
        for (auto& H : host)
            if (H.has_barostat) H.valid = false;
    }
    return DMDX_OK;
}

int Batch::step(int64_t nsteps)
{
    pristine = false;
    if (nsteps < 0) return fail(DMDX_E_INVALID, "negative step count");
    if (!options.use_graph) {
        for (int64_t i = 0; i < nsteps; ++i) TRY(enqueue_step(false));
        return DMDX_OK;
    }
    // graph mode: every compute launches every kernel (per-system decisions
    // are taken on the device), so one captured block of graph_steps steps
    // is valid for any counter state
    const int G = std::max(1, options.graph_steps);
    int64_t done = 0;
    if (nsteps >= G && !graph_exec) {
        DMDX_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        std::vector<HostSys> saved = host;
        int rc = DMDX_OK;
        for (int i = 0; i < G && rc == DMDX_OK; ++i) rc = enqueue_step(true);
        hipGraph_t g = nullptr;
        hipError_t ec = hipStreamEndCapture(stream, &g);
        host = saved;  // the capture only recorded; nothing ran
        if (rc != DMDX_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        DMDX_HIP(ec);
        graph = g;
        DMDX_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
    }
    while (nsteps - done >= G && graph_exec) {
        DMDX_HIP(hipGraphLaunch(graph_exec, stream));
        for (int i = 0; i < G; ++i) {
            bool dummy;
            mirror_compute(1, 0, &dummy);
            mirror_compute(2, 0, &dummy);
            for (auto& H : host)
                if (H.has_barostat) H.valid = false;
        }
        done += G;
    }
    for (; done < nsteps; ++done) TRY(enqueue_step(false));
    return DMDX_OK;
}

void Batch::invalidate_graph()
{
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
}

int Batch::compute(int accumulate)
{
    pristine = false;
    TRY(enqueue_compute(0, accumulate, 0, false));
    LAUNCH(launch_finalize(d, stream, 0));
    return DMDX_OK;
}

int Batch::rebuild_only()
{
    pristine = false;
    LAUNCH(launch_pre_rebuild_only(d, stream));
    if (d.any_list) LAUNCH(launch_build_lists(d, stream));
    for (auto& H : host)
        if (H.has_list) {
            H.counter = 0;
            H.built = true;
        }
    return DMDX_OK;
}

int Batch::copy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind)
{
    DMDX_HIP(hipMemcpyAsync(dst, src, bytes, kind, stream));
    DMDX_HIP(hipStreamSynchronize(stream));
    return DMDX_OK;
}

int Batch::sync()
{
    DMDX_HIP(hipStreamSynchronize(stream));
    return DMDX_OK;
}

int Batch::fetch_state()
{
    TRY(sync());
    TRY(copy(hstate.data(), d.st, sizeof(SysState) * hstate.size(), hipMemcpyDeviceToHost));
    return DMDX_OK;
}

int Batch::check_sys(int s) const
{
    if (s < 0 || s >= d.nsys) return fail(DMDX_E_INVALID, "system index out of range");
    return DMDX_OK;
}

int Batch::get_arrays(int s, int what, double* a, double* b, double* c, bool on_device)
{
    TRY(check_sys(s));
    if (what == 1) LAUNCH(launch_materialize(d, stream));
    TRY(sync());
    const SysDev& S = hsys[s];
    double* src[3];
    if (what == 0) { src[0] = d.x; src[1] = d.y; src[2] = d.z; }
    else if (what == 1) { src[0] = d.vx; src[1] = d.vy; src[2] = d.vz; }
    else if (what == 2) { src[0] = d.fx; src[1] = d.fy; src[2] = d.fz; }
    else return fail(DMDX_E_INVALID, "what must be 0 (positions), 1 (velocities) or 2 (forces)");
    double* dst[3] = {a, b, c};
    for (int k = 0; k < 3; ++k)
        if (dst[k])
            TRY(copy(dst[k], src[k] + S.atom_off, sizeof(double) * S.n,
                     on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost));
    return DMDX_OK;
}

int Batch::set_arrays(int s, int what, const double* a, const double* b, const double* c,
                      bool on_device)
{
    TRY(check_sys(s));
    pristine = false;
    TRY(sync());
    const SysDev& S = hsys[s];
    double* dst[3];
    if (what == 0) { dst[0] = d.x; dst[1] = d.y; dst[2] = d.z; }
    else if (what == 1) {
        LAUNCH(launch_materialize(d, stream));
        TRY(sync());
        dst[0] = d.vx; dst[1] = d.vy; dst[2] = d.vz;
    } else if (what == 2) { dst[0] = d.fx; dst[1] = d.fy; dst[2] = d.fz; }
    else return fail(DMDX_E_INVALID, "what must be 0, 1 or 2");
    const double* src[3] = {a, b, c};
    for (int k = 0; k < 3; ++k)
        if (src[k])
            TRY(copy(dst[k] + S.atom_off, src[k], sizeof(double) * S.n,
                     on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
    if (what == 0) {
        // positions changed: forces of the last compute no longer apply
        TRY(fetch_state());
        hstate[s].valid = 0;
        TRY(copy(&d.st[s].valid, &hstate[s].valid, sizeof(int32_t), hipMemcpyHostToDevice));
        host[s].valid = false;
    }
    return DMDX_OK;
}

int Batch::set_cell(int s, const double m[9], int type)
{
    TRY(check_sys(s));
    if (type < DMDX_CELL_ORTHORHOMBIC || type > DMDX_CELL_RHOMBIC_DODECAHEDRON)
        return fail(DMDX_E_PARM, "unknown cell type");
    const double L[3] = {m[0], m[4], m[8]};
    for (int k = 0; k < 3; ++k)
        if (!(L[k] > 0.0)) return fail(DMDX_E_PARM, "cell edges (the matrix diagonal) must be positive");
    // Cell::volume (src/core/cell.cpp:19-24), in its operation order
    const double vdet = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                        m[2] * (m[3] * m[7] - m[4] * m[6]);
    if (!(vdet > 0.0)) return fail(DMDX_E_PARM, "the cell matrix must have a positive determinant");
    const bool open = type != DMDX_CELL_ORTHORHOMBIC;
    SysDev& S = hsys[s];
    if (open && S.precision != DMDX_FP64)
        return fail(DMDX_E_PARM, "non-orthorhombic cells (identity minimum image) run on the FP64 "
                                 "path only: the mixed path's cell-local frames are periodic");
    const bool same_box = S.L[0] == L[0] && S.L[1] == L[1] && S.L[2] == L[2];
    if (same_box && S.vdet == vdet && (S.iminv[0] == 0.0) == open) return DMDX_OK;
    TRY(sync());
    // the device copy may have been changed by a barostat: read it back first
    TRY(copy(&S, d.sys + s, sizeof(SysDev), hipMemcpyDeviceToHost));
    const bool box_changed = !(S.L[0] == L[0] && S.L[1] == L[1] && S.L[2] == L[2]);
    const double v_before = sys_volume(S);
    for (int k = 0; k < 3; ++k) {
        S.L[k] = L[k];
        S.invL[k] = 1.0 / L[k];
        // the reference's image is the identity for non-orthorhombic types
        S.iminv[k] = open ? 0.0 : S.invL[k];
    }
    S.vdet = vdet;
    // the mixed path's influence function was tabulated for the box and
    // volume the system was created with: another cell computes it per call
    if (box_changed || vdet != v_before) S.ginf_static = 0;
    if (!d.any_wrap_mixed && mixed_list_wraps(S)) {
        d.any_wrap_mixed = true;  // the per-atom build joins the launches
        invalidate_graph();
    }
    pristine = false;
    TRY(copy(d.sys + s, &S, sizeof(SysDev), hipMemcpyHostToDevice));
    host[s].valid = false;
    TRY(fetch_state());
    hstate[s].valid = 0;
    TRY(copy(&d.st[s].valid, &hstate[s].valid, sizeof(int32_t), hipMemcpyHostToDevice));
    return DMDX_OK;
}

int Batch::upload_sys(int s)
{
    TRY(copy(d.sys + s, &hsys[s], sizeof(SysDev), hipMemcpyHostToDevice));
    return DMDX_OK;
}

int Batch::set_counter(int s, int counter)
{
    pristine = false;
    TRY(check_sys(s));
    TRY(fetch_state());
    hstate[s].counter = counter;
    host[s].counter = counter;
    TRY(copy(&d.st[s].counter, &counter, sizeof(int32_t), hipMemcpyHostToDevice));
    return DMDX_OK;
}

int Batch::set_active(int s, int active)
{
    pristine = false;
    TRY(check_sys(s));
    TRY(fetch_state());
    hstate[s].active = active ? 1 : 0;
    host[s].active = active != 0;
    // re-admitting a system clears its failure flags (otherwise it would
    // leave again before its next step)
    if (active) hstate[s].flags = 0;
    TRY(copy(&d.st[s], &hstate[s], sizeof(SysState), hipMemcpyHostToDevice));
    return DMDX_OK;
}

int Batch::adopt(Batch& src, int ss, int ds)
{
    TRY(check_sys(ds));
    pristine = false;
    TRY(src.check_sys(ss));
    if (src.d.precision != d.precision) return fail(DMDX_E_MISMATCH, "precisions differ");
    TRY(src.fetch_state());
    TRY(sync());
    SysDev A, B;
    TRY(src.copy(&A, src.d.sys + ss, sizeof(SysDev), hipMemcpyDeviceToHost));
    TRY(copy(&B, d.sys + ds, sizeof(SysDev), hipMemcpyDeviceToHost));
    // the copies below move whole per-system regions (list rows, cells,
    // frames): both systems must have the same layout of every one of them
    bool same = A.n == B.n && A.n_pad == B.n_pad && A.max_nbr == B.max_nbr &&
                A.has_list == B.has_list && A.cdiv == B.cdiv && A.bin_wg == B.bin_wg &&
                A.pme == B.pme && A.precision == B.precision && A.n_bond == B.n_bond &&
                A.n_angle == B.n_angle && A.n_dih == B.n_dih && A.n_imp == B.n_imp &&
                A.n_ex == B.n_ex && A.n_con == B.n_con;
    for (int k = 0; k < 3; ++k) same = same && A.nc[k] == B.nc[k] && A.K[k] == B.K[k];
    if (!same)
        return fail(DMDX_E_MISMATCH, "systems were not created from the same description");
    for (int k = 0; k < 3; ++k) {  // a barostat may have changed the box
        B.L[k] = A.L[k];
        B.invL[k] = A.invL[k];
        B.iminv[k] = A.iminv[k];
    }
    B.vdet = A.vdet;
    hsys[ds].L[0] = B.L[0], hsys[ds].L[1] = B.L[1], hsys[ds].L[2] = B.L[2];
    hsys[ds].invL[0] = B.invL[0], hsys[ds].invL[1] = B.invL[1], hsys[ds].invL[2] = B.invL[2];
    if (!d.any_wrap_mixed && mixed_list_wraps(B)) {
        d.any_wrap_mixed = true;  // the per-atom build joins the launches
        invalidate_graph();
    }
    TRY(copy(d.sys + ds, &B, sizeof(SysDev), hipMemcpyHostToDevice));
    hstate[ds] = src.hstate[ss];
    TRY(copy(d.st + ds, &hstate[ds], sizeof(SysState), hipMemcpyHostToDevice));
    host[ds].counter = src.host[ss].counter;
    host[ds].valid = src.host[ss].valid;
    host[ds].built = src.host[ss].built;
    host[ds].active = src.host[ss].active;
    const int64_t n = A.n, np = A.n_pad, ga = A.atom_off, gb = B.atom_off;
    auto d2d = [&](void* dst, const void* from, size_t bytes) -> int {
        DMDX_HIP(hipMemcpyAsync(dst, from, bytes, hipMemcpyDeviceToDevice, stream));
        return DMDX_OK;
    };
    double* const sa[9] = {src.d.x, src.d.y, src.d.z, src.d.vx, src.d.vy, src.d.vz,
                           src.d.fx, src.d.fy, src.d.fz};
    double* const sb[9] = {d.x, d.y, d.z, d.vx, d.vy, d.vz, d.fx, d.fy, d.fz};
    for (int k = 0; k < 9; ++k) TRY(d2d(sb[k] + gb, sa[k] + ga, sizeof(double) * n));
    TRY(d2d(d.xref + gb, src.d.xref + ga, sizeof(double4) * np));
    if (A.has_list) {
        TRY(d2d(d.order + gb, src.d.order + ga, sizeof(int) * np));
        TRY(d2d(d.nn + gb, src.d.nn + ga, sizeof(int) * np));
        const int64_t cnt = (int64_t)A.max_nbr * np;
        TRY(d2d(d.nbr + B.nbr_off, src.d.nbr + A.nbr_off, sizeof(int) * cnt));
        if (d.org) TRY(d2d(d.org + gb, src.d.org + ga, sizeof(double4) * np));
        // the drift writes the frames through the cell-order map (relative
        // to the system); the types in cell order change only at a build
        if (d.slot_of) TRY(d2d(d.slot_of + gb, src.d.slot_of + ga, sizeof(int) * np));
        if (d.typ_s) TRY(d2d(d.typ_s + gb, src.d.typ_s + ga, sizeof(int) * np));
        // cells of the last build (read between builds by the owner-computes
        // charge spreading); starts are relative to the system's first slot
        const int64_t nc3 = (int64_t)A.nc[0] * A.nc[1] * A.nc[2];
        TRY(d2d(d.ccount + B.cell_off, src.d.ccount + A.cell_off, sizeof(int) * nc3));
        TRY(d2d(d.cstart + B.cell_off, src.d.cstart + A.cell_off, sizeof(int) * nc3));
        // FP64 entries are atom slots; mixed entries are local to the system
        LAUNCH(launch_remap_rows(stream, d.nbr + B.nbr_off, d.precision == DMDX_FP64 ? cnt : 0,
                                 d.order + gb, np, gb - ga));
    }
    TRY(sync());
    return DMDX_OK;
}

int Batch::set_stop_at(const std::vector<int64_t>& stop)
{
    pristine = false;
    if ((int)stop.size() != d.nsys) return fail(DMDX_E_MISMATCH, "one step target per system");
    TRY(fetch_state());
    for (int s = 0; s < d.nsys; ++s) hstate[s].stop_at = stop[s];
    TRY(copy(d.st, hstate.data(), sizeof(SysState) * hstate.size(), hipMemcpyHostToDevice));
    return DMDX_OK;
}

int Batch::get_pairs(int s, int32_t* pi, int32_t* pj, int64_t cap, int64_t* count)
{
    TRY(check_sys(s));
    TRY(sync());
    const SysDev& S = hsys[s];
    std::vector<int> order(S.n_pad), nn(S.n_pad);
    TRY(copy(order.data(), d.order + S.atom_off, sizeof(int) * S.n_pad, hipMemcpyDeviceToHost));
    TRY(copy(nn.data(), d.nn + S.atom_off, sizeof(int) * S.n_pad, hipMemcpyDeviceToHost));
    std::vector<int> blk;
    if (S.has_list) {
        blk.resize((size_t)S.max_nbr * S.n_pad);
        TRY(copy(blk.data(), d.nbr + S.nbr_off, sizeof(int) * blk.size(), hipMemcpyDeviceToHost));
    }
    // FP64 rows hold atom slots, mixed rows hold cell-order slots
    const bool sorted_ids = S.precision != DMDX_FP64;
    std::vector<int64_t> keys;
    for (int t = 0; t < S.n; ++t) {
        int i = order[t] - (int)S.atom_off;
        for (int k = 0; k < nn[t]; ++k) {
            int e = blk[(size_t)t * S.max_nbr + k];
            if (sorted_ids) e = (int)((uint32_t)e & NBR_SLOT_MASK);  // local cell-order slot
            int j = (sorted_ids ? order[e] : e) - (int)S.atom_off;
            if (j > i) keys.push_back(((int64_t)i << 32) | (int64_t)j);
        }
    }
    std::sort(keys.begin(), keys.end());
    *count = (int64_t)keys.size();
    for (int64_t p = 0; p < (int64_t)keys.size() && p < cap; ++p) {
        pi[p] = (int32_t)(keys[p] >> 32);
        pj[p] = (int32_t)(keys[p] & 0xffffffff);
    }
    return DMDX_OK;
}

}  // namespace dmdx

// ===================================================================== C-ABI
using dmdx::Batch;


extern "C" {

const char* dmdx_last_error(void) { return dmdx::last_error(); }

int dmdx_device_info(int32_t* device_count, int64_t* free_bytes, int64_t* total_bytes)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return dmdx::fail(DMDX_E_GPU, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
    if (device_count) *device_count = n;
    if (n > 0 && (free_bytes || total_bytes)) {
        size_t f = 0, t = 0;
        e = hipMemGetInfo(&f, &t);
        if (e != hipSuccess) return dmdx::fail(DMDX_E_GPU, std::string("hipMemGetInfo: ") + hipGetErrorString(e));
        if (free_bytes) *free_bytes = (int64_t)f;
        if (total_bytes) *total_bytes = (int64_t)t;
    }
    return DMDX_OK;
}

int dmdx_engine_create(const dmdx_system_desc* descs, int32_t n_systems, dmdx_engine** out)
{
    if (!descs || !out) return dmdx::fail(DMDX_E_INVALID, "null argument");
    auto* e = new dmdx_engine();
    dmdx::BatchOptions opt;
    int rc = e->b.create(descs, n_systems, opt);
    if (rc != DMDX_OK) {
        delete e;
        return rc;
    }
    for (int s = 0; s < n_systems; ++s) e->descs.push_back(std::make_unique<dmdx::DescCopy>(descs[s]));
    *out = e;
    return DMDX_OK;
}

int dmdx_engine_destroy(dmdx_engine* e)
{
    delete e;
    return DMDX_OK;
}

int dmdx_engine_set_options(dmdx_engine* e, int32_t reuse_forces, int32_t use_graph,
                            int32_t graph_steps)
{
    if (!e) return dmdx::fail(DMDX_E_INVALID, "null engine");
    e->b.invalidate_graph();
    e->b.options.reuse_forces = reuse_forces != 0;
    e->b.options.use_graph = use_graph != 0;
    e->b.options.graph_steps = graph_steps > 0 ? graph_steps : 10;
    return DMDX_OK;
}

int dmdx_engine_step(dmdx_engine* e, int64_t nsteps)
{
    if (!e) return dmdx::fail(DMDX_E_INVALID, "null engine");
    return e->b.step(nsteps);
}

int dmdx_engine_compute(dmdx_engine* e)
{
    if (!e) return dmdx::fail(DMDX_E_INVALID, "null engine");
    return e->b.compute(0);
}

int dmdx_engine_sync(dmdx_engine* e)
{
    if (!e) return dmdx::fail(DMDX_E_INVALID, "null engine");
    return e->b.sync();
}

int dmdx_engine_get_arrays(dmdx_engine* e, int32_t sys, int32_t what, double* x, double* y, double* z)
{
    if (!e) return dmdx::fail(DMDX_E_INVALID, "null engine");
    return e->b.get_arrays(sys, what, x, y, z);
}

int dmdx_engine_set_arrays(dmdx_engine* e, int32_t sys, int32_t what, const double* x,
                           const double* y, const double* z)
{
    if (!e) return dmdx::fail(DMDX_E_INVALID, "null engine");
    return e->b.set_arrays(sys, what, x, y, z);
}

int dmdx_engine_get_scalars(dmdx_engine* e, int32_t sys, dmdx_scalars* out)
{
    if (!e || !out) return dmdx::fail(DMDX_E_INVALID, "null argument");
    int rc = e->b.check_sys(sys);
    if (rc) return rc;
    rc = e->b.fetch_state();
    if (rc) return rc;
    dmdx::SysDev S;
    rc = e->b.copy(&S, e->b.d.sys + sys, sizeof(S), hipMemcpyDeviceToHost);
    if (rc) return rc;
    const dmdx::SysState& T = e->b.hstate[sys];
    out->potential_energy = T.pe;
    for (int k = 0; k < 8; ++k) out->energy_terms[k] = T.pe_terms[k];
    out->kinetic_energy = T.ke_out;
    out->zeta = T.zeta;
    out->piston_velocity = T.eps_dot;
    out->virial = T.virial;
    out->pressure = T.pressure;
    for (int k = 0; k < 3; ++k) out->box[k] = S.L[k];
    out->step_since_rebuild = T.counter;
    out->flags = T.flags;
    out->steps_done = T.steps_done;
    out->list_builds = T.rebuilds;
    return DMDX_OK;
}

int dmdx_engine_set_step_since_rebuild(dmdx_engine* e, int32_t sys, int32_t counter)
{
    if (!e) return dmdx::fail(DMDX_E_INVALID, "null engine");
    return e->b.set_counter(sys, counter);
}

int dmdx_engine_set_active(dmdx_engine* e, int32_t sys, int32_t active)
{
    if (!e) return dmdx::fail(DMDX_E_INVALID, "null engine");
    return e->b.set_active(sys, active);
}

int dmdx_engine_adopt(dmdx_engine* dst, int32_t dst_sys, dmdx_engine* src, int32_t src_sys)
{
    if (!dst || !src) return dmdx::fail(DMDX_E_INVALID, "null engine");
    if (dst == src) return dmdx::fail(DMDX_E_INVALID, "source and destination are one engine");
    return dst->b.adopt(src->b, src_sys, dst_sys);
}

int dmdx_engine_get_pairs(dmdx_engine* e, int32_t sys, int32_t* pi, int32_t* pj, int64_t cap,
                          int64_t* count)
{
    if (!e || !count) return dmdx::fail(DMDX_E_INVALID, "null argument");
    return e->b.get_pairs(sys, pi, pj, cap, count);
}

int dmdx_engine_set_profiling(dmdx_engine* e, int32_t on)
{
    if (!e) return dmdx::fail(DMDX_E_INVALID, "null engine");
    if (on && e->b.options.use_graph)
        return dmdx::fail(DMDX_E_INVALID, "kernel timing needs eager launches (use_graph = 0)");
    e->b.options.profile = on != 0;
    e->b.options.profile_mask = ~0u;
    return DMDX_OK;
}

int dmdx_engine_set_profiling_groups(dmdx_engine* e, uint32_t mask)
{
    if (!e) return dmdx::fail(DMDX_E_INVALID, "null engine");
    if (mask && e->b.options.use_graph)
        return dmdx::fail(DMDX_E_INVALID, "kernel timing needs eager launches (use_graph = 0)");
    e->b.options.profile = mask != 0;
    e->b.options.profile_mask = mask;
    return DMDX_OK;
}

int dmdx_engine_kernel_time(dmdx_engine* e, int32_t group, double* total_ms, int64_t* launches)
{
    if (!e || !total_ms || !launches) return dmdx::fail(DMDX_E_INVALID, "null argument");
    return e->b.kernel_time(group, total_ms, launches);
}

int dmdx_engine_set_tuning(dmdx_engine* e, int32_t key, int32_t value)
{
    if (!e) return dmdx::fail(DMDX_E_INVALID, "null engine");
    if (key == 2) {
        if (value != 2 && value != 4 && value != 8)
            return dmdx::fail(DMDX_E_PARM, "entries per lane must be 2, 4 or 8");
        e->b.d.pair_width = value;
    } else if (key == 0) {
        if (value != 2 && value != 4 && value != 8 && value != 16 && value != 32)
            return dmdx::fail(DMDX_E_PARM, "pair lanes per atom must be 2, 4, 8, 16 or 32");
        e->b.d.pair_tpa = value;
    } else if (key == 1) {
        if (value != 8 && value != 16 && value != 32)
            return dmdx::fail(DMDX_E_PARM, "build lanes per atom must be 8, 16 or 32");
        e->b.d.build_tpa = value;
    } else if (key == 3) {
        if (value != 0 && value != 1) return dmdx::fail(DMDX_E_PARM, "per-cell build is 0 or 1");
        e->b.d.build_cells = value != 0;
    } else if (key == 4) {
        if (value != 1 && value != 2 && value != 4 && value != 8)
            return dmdx::fail(DMDX_E_PARM, "waves per cell must be 1, 2, 4 or 8");
        e->b.d.build_tasks = value;
    } else if (key == 5) {
        if (value != 0 && value != 1) return dmdx::fail(DMDX_E_PARM, "skin tuning is 0 or 1");
        e->tune_skin = value != 0;
    } else if (key == 6) {
        if (value != 0 && value != 4 && value != 8)
            return dmdx::fail(DMDX_E_PARM, "row-tile pair kernel is 0 (off), 4 or 8 entries per block");
        e->b.d.pair_rows = value;
    } else if (key == 9) {
        // cells per list radius of every list system, for a batch that has
        // not left its initial state (the batch is made again)
        if (value < 0 || value > 2)
            return dmdx::fail(DMDX_E_PARM, "cells per list radius is 0 (density rule), 1 or 2");
        if (!e->b.pristine)
            return dmdx::fail(DMDX_E_INVALID, "the cell grid is set before the first step");
        dmdx::BatchOptions opt = e->b.options;
        opt.cdiv.assign(e->descs.size(), value);
        return dmdx_engine_remake(e, opt);
    } else if (key == 12) {
        // xy planes of the mixed PME transforms, for a batch that has not
        // left its initial state (the batch is made again)
        if (value < 0 || value > 2)
            return dmdx::fail(DMDX_E_PARM, "FFT planes are 0 (never), 1 (where they fit) or 2 (rule)");
        if (!e->b.pristine)
            return dmdx::fail(DMDX_E_INVALID, "the FFT planes are set before the first step");
        dmdx::BatchOptions opt = e->b.options;
        opt.fft_planes = value;
        return dmdx_engine_remake(e, opt);
    } else if (key == 7) {
        if (value != 0 && value != 8 && value != 16)
            return dmdx::fail(DMDX_E_PARM, "staged per-cell build is 0 (off), 8 or 16 lanes per atom");
        e->b.d.build_stage = value;
    } else if (key == 10) {
        if (value != 2 && value != 4 && value != 8 && value != 16 && value != 32)
            return dmdx::fail(DMDX_E_PARM, "Coulomb pair lanes per atom must be 2, 4, 8, 16 or 32");
        e->b.d.coul_tpa = value;
    } else if (key == 11) {
        if (value != 2 && value != 4 && value != 8)
            return dmdx::fail(DMDX_E_PARM, "Coulomb entries per lane must be 2, 4 or 8");
        e->b.d.coul_width = value;
    } else {
        return dmdx::fail(DMDX_E_INVALID, "unknown tuning key");
    }
    e->b.invalidate_graph();
    return DMDX_OK;
}

int dmdx_engine_get_tuning(dmdx_engine* e, int32_t key, int32_t* value)
{
    if (!e || !value) return dmdx::fail(DMDX_E_INVALID, "null argument");
    const dmdx::Dev& d = e->b.d;
    switch (key) {
    case 0: *value = d.pair_tpa; break;
    case 1: *value = d.build_tpa; break;
    case 2: *value = d.pair_width; break;
    case 3: *value = d.build_cells ? 1 : 0; break;
    case 4: *value = d.build_tasks; break;
    case 5: *value = e->tune_skin ? 1 : 0; break;
    case 6: *value = d.pair_rows; break;
    case 7: *value = d.build_stage; break;
    case 8: *value = e->b.options.fft_lines; break;
    case 10: *value = d.coul_tpa; break;
    case 11: *value = d.coul_width; break;
    case 12: *value = e->b.options.fft_planes; break;
    case 13: *value = d.nptiles > 0 ? 1 : 0; break;
    default: return dmdx::fail(DMDX_E_INVALID, "unknown tuning key");
    }
    return DMDX_OK;
}

int dmdx_engine_get_cell_div(dmdx_engine* e, int32_t sys, int32_t* cdiv)
{
    if (!e || !cdiv) return dmdx::fail(DMDX_E_INVALID, "null argument");
    if (sys < 0 || sys >= (int32_t)e->descs.size()) return dmdx::fail(DMDX_E_INVALID, "bad system");
    *cdiv = e->b.hsys[sys].cdiv;
    return DMDX_OK;
}

int dmdx_engine_get_list_skin(dmdx_engine* e, int32_t sys, double* skin)
{
    if (!e || !skin) return dmdx::fail(DMDX_E_INVALID, "null argument");
    if (sys < 0 || sys >= (int32_t)e->descs.size()) return dmdx::fail(DMDX_E_INVALID, "bad system");
    *skin = e->descs[sys]->d.list_skin;
    return DMDX_OK;
}

int dmdx_engine_n_systems(dmdx_engine* e, int32_t* n)
{
    if (!e || !n) return dmdx::fail(DMDX_E_INVALID, "null argument");
    *n = e->b.d.nsys;
    return DMDX_OK;
}

}  // extern "C"
