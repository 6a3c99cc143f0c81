// server.cpp -- many independent simulations served by one device.
//
// Requests (a system plus a step count) are submitted at any time.  The
// server groups waiting requests into cohorts: each cohort is one Batch (all
// its systems advanced by the same launches, on the cohort's own stream), and
// several cohorts run concurrently.  After each round the cohorts of one
// precision are merged into one batch, each system moving with its whole
// state (Batch::adopt), so new arrivals join the big batch after one round
// and finished requests stop occupying it.  A request is admitted when its estimated
// device memory fits in what the budget has left; it leaves its cohort
// exactly at its step target (the device freezes it at the start of the
// first step after steps_done reaches the target) or when it fails
// (neighbour-list overflow, non-finite energy), without disturbing the other
// systems of the cohort.  A cohort's memory is returned when its last
// request retires.  A system's arithmetic never depends on the other
// systems of its batch, so a request's result is the same whether it runs
// alone or in any mix.
//
// The reference runs one SimulationEngine per process
// (src/sim/simulation_engine.cpp:89-106, python/dmd/runner.py:44-58); this
// is the device-side replacement for running many of those side by side.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "engine.h"

using dmdx::Batch;
using dmdx::BatchOptions;

namespace {

using Clock = std::chrono::steady_clock;

// Device bytes a system will hold in a batch: per-slot state and scratch,
// the neighbour list at the width the batch will choose, PME grids and
// bonded term buffers.  An estimate; the server scales it by the largest
// actual/estimate ratio it has seen.
int64_t estimate_bytes(const dmdx_system_desc& D)
{
    const double PI = 3.14159265358979323846;
    const int64_t n = D.n_atoms, n_pad = (n + 63) / 64 * 64;
    int64_t b = n_pad * 512;  // FP64 state, packed copies, cell data, partials
    const bool coul = D.physics != DMDX_FAITHFUL && D.coulomb_type != DMDX_COULOMB_NONE;
    double rl = 0.0;
    if (D.use_lj) rl = D.lj_cutoff;
    if (coul) rl = std::max(rl, D.coulomb_cutoff);
    if (rl > 0.0) {
        rl += D.list_skin;
        const double dens = n / (D.box[0] * D.box[1] * D.box[2]);
        const double expect = 4.0 / 3.0 * PI * rl * rl * rl * dens;
        int64_t M = (int64_t)std::ceil(1.6 * expect + 48.0);
        M = std::max<int64_t>(1, std::min<int64_t>(M, n - 1));
        M = (M + 7) / 8 * 8;
        b += M * n_pad * 4;
        int64_t cells = 1;
        for (int k = 0; k < 3; ++k)
            cells *= std::max<int64_t>(1, (int64_t)std::floor(D.box[k] / rl) + 1);
        b += cells * 16;
    }
    if (D.coulomb_type == DMDX_COULOMB_PME) {
        const int64_t k3 = (int64_t)D.pme_grid[0] * D.pme_grid[1] * D.pme_grid[2];
        b += k3 * 2 * 16 + 4 * k3;
    }
    const int64_t roles = 2 * D.n_bonds + 3 * D.n_angles + 4 * D.n_dihedrals + 4 * D.n_impropers +
                          2 * D.n_excl;
    b += roles * 40 + (D.n_bonds + D.n_angles + D.n_dihedrals + D.n_impropers + D.n_excl) * 48;
    if (D.n_constraints > 0) b += D.n_constraints * 24;  // slots, distance, component index
    return b + 4096;
}

struct Request {
    int64_t id = 0;
    std::unique_ptr<dmdx::DescCopy> desc;
    int64_t n_steps = 0;
    int32_t state = DMDX_REQ_QUEUED;
    int32_t flags = 0;
    int64_t steps_done = 0;
    double t_submit = 0.0, t_admit = -1.0, t_done = -1.0;
    int64_t est_bytes = 0;
    int cohort = -1;  // index into cohorts while running
    int sys = -1;     // system index in the cohort's batch
    std::string error;
    // results, kept after retirement until released
    std::vector<double> res[9];
    dmdx_scalars sc;
};

struct Cohort {
    std::unique_ptr<Batch> b;
    std::vector<int64_t> ids;  // request of each system
    int64_t bytes = 0;
    int live = 0;
    int64_t chunk = 0;  // steps enqueued in the current round
};

}  // namespace

struct dmdx_server {
    dmdx_server_options opt;
    std::map<int64_t, Request> reqs;
    std::deque<int64_t> pending;
    std::vector<std::unique_ptr<Cohort>> cohorts;
    int64_t next_id = 1;
    int64_t budget = 0, in_use = 0;
    double est_scale = 1.25;
    Clock::time_point t0;
    int64_t rounds = 0, retired = 0, failed = 0, cohorts_made = 0, merges = 0, cancelled = 0;
    int64_t atom_steps = 0;  // atoms x steps advanced over all requests
    std::vector<double> lat_ms;
    std::vector<int64_t> lat_w;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

int retire(dmdx_server* srv, Cohort& c, Request& r, int32_t state)
{
    const bool failed = state == DMDX_REQ_FAILED;
    Batch& b = *c.b;
    const int s = r.sys;
    const int64_t n = r.desc->d.n_atoms;
    for (auto& v : r.res) v.assign(n, 0.0);
    int rc = b.get_arrays(s, 0, r.res[0].data(), r.res[1].data(), r.res[2].data());
    if (!rc) rc = b.get_arrays(s, 1, r.res[3].data(), r.res[4].data(), r.res[5].data());
    if (!rc) rc = b.get_arrays(s, 2, r.res[6].data(), r.res[7].data(), r.res[8].data());
    if (rc) return rc;
    const dmdx::SysState& T = b.hstate[s];
    dmdx::SysDev S;
    rc = b.copy(&S, b.d.sys + s, sizeof(S), hipMemcpyDeviceToHost);  // box after a barostat
    if (rc) return rc;
    std::memset(&r.sc, 0, sizeof(r.sc));
    r.sc.potential_energy = T.pe;
    for (int k = 0; k < 8; ++k) r.sc.energy_terms[k] = T.pe_terms[k];
    r.sc.kinetic_energy = T.ke_out;
    r.sc.zeta = T.zeta;
    r.sc.piston_velocity = T.eps_dot;
    r.sc.virial = T.virial;
    r.sc.pressure = T.pressure;
    for (int k = 0; k < 3; ++k) r.sc.box[k] = S.L[k];
    r.sc.step_since_rebuild = T.counter;
    r.sc.flags = T.flags;
    r.sc.steps_done = T.steps_done;
    r.sc.list_builds = T.rebuilds;
    r.state = state;
    r.steps_done = T.steps_done;
    if (failed)
        r.error = (T.flags & dmdx::FLAG_NBR_OVERFLOW) ? "neighbour list overflow"
                                                      : "non-finite energy";
    r.t_done = srv->now();
    r.desc.reset();
    r.cohort = -1;
    c.live -= 1;
    if (state == DMDX_REQ_CANCELLED) {
        srv->cancelled += 1;
    } else {
        srv->retired += 1;
        if (failed) srv->failed += 1;
    }
    return DMDX_OK;
}

// Device memory of a cohort's systems that no longer run, by estimate.
int64_t dead_bytes(const dmdx_server* srv, const Cohort& c)
{
    int64_t dead = 0;
    for (int64_t id : c.ids) {
        auto it = srv->reqs.find(id);
        if (it == srv->reqs.end() || it->second.state != DMDX_REQ_RUNNING)
            dead += it == srv->reqs.end() ? 0 : it->second.est_bytes;
    }
    return dead;
}

constexpr int COHORT_OVER_BUDGET = 1;

// Build one cohort from `ids` (all sharing one precision).  On a failure
// that is not an out-of-memory, the requests are tried one by one so that a
// bad request fails alone.  Returns COHORT_OVER_BUDGET (and learns a larger
// estimate scale) when the batch turned out bigger than the budget left.
int make_cohort(dmdx_server* srv, const std::vector<int64_t>& ids, std::vector<int64_t>* rejected)
{
    std::vector<dmdx_system_desc> descs;
    for (int64_t id : ids) descs.push_back(srv->reqs[id].desc->d);
    BatchOptions bo;
    bo.reuse_forces = true;
    bo.use_graph = srv->opt.graph_steps > 0;
    bo.graph_steps = std::max(1, (int)srv->opt.graph_steps);
    auto b = std::make_unique<Batch>();
    int rc = b->create(descs.data(), (int)descs.size(), bo);
    if (rc == DMDX_E_OOM || rc == DMDX_E_GPU) return rc;
    if (rc != DMDX_OK) {
        if (ids.size() == 1) {
            rejected->push_back(ids[0]);
            return rc;
        }
        std::vector<int64_t> good;
        for (int64_t id : ids) {
            Batch probe;
            const dmdx_system_desc& d1 = srv->reqs[id].desc->d;
            int r1 = probe.create(&d1, 1, bo);
            if (r1 == DMDX_OK) good.push_back(id);
            else if (r1 == DMDX_E_OOM || r1 == DMDX_E_GPU) return r1;
            else {
                srv->reqs[id].error = dmdx::last_error();
                rejected->push_back(id);
            }
        }
        if (good.empty()) return DMDX_OK;
        return make_cohort(srv, good, rejected);
    }
    const int64_t actual = (int64_t)b->device_bytes();
    int64_t est = 0;
    for (int64_t id : ids) est += srv->reqs[id].est_bytes;
    if (est > 0) srv->est_scale = std::max(srv->est_scale, 1.05 * (double)actual / (double)est);
    // the estimate was low: give the memory back before anything runs
    if (srv->in_use + actual > srv->budget) return COHORT_OVER_BUDGET;
    std::vector<int64_t> stop;
    for (int64_t id : ids) stop.push_back(srv->reqs[id].n_steps);
    rc = b->set_stop_at(stop);
    if (rc) return rc;
    auto c = std::make_unique<Cohort>();
    c->bytes = actual;
    c->b = std::move(b);
    c->ids = ids;
    c->live = (int)ids.size();
    const double t = srv->now();
    const int ci = (int)srv->cohorts.size();
    for (size_t k = 0; k < ids.size(); ++k) {
        Request& r = srv->reqs[ids[k]];
        r.state = DMDX_REQ_RUNNING;
        r.t_admit = t;
        r.cohort = ci;
        r.sys = (int)k;
    }
    srv->in_use += c->bytes;
    srv->cohorts.push_back(std::move(c));
    srv->cohorts_made += 1;
    return DMDX_OK;
}

// Merge every cohort of each precision into one batch: a new batch is built
// from the live requests' descriptions and each system continues there with
// its whole state (Batch::adopt).  A lone cohort is rebuilt the same way
// (compacted) once systems that no longer run -- retired or cancelled --
// hold a quarter of its memory.  Skipped when the new batch would not fit
// next to the old ones.
int merge(dmdx_server* srv)
{
    for (int prec : {DMDX_FP64, DMDX_MIXED}) {
        std::vector<int> cs;
        for (size_t ci = 0; ci < srv->cohorts.size(); ++ci)
            if (srv->cohorts[ci]->b->d.precision == prec) cs.push_back((int)ci);
        if (cs.empty()) continue;
        if (cs.size() == 1) {
            const Cohort& c = *srv->cohorts[cs[0]];
            int64_t est = 0;
            for (int64_t id : c.ids) {
                auto it = srv->reqs.find(id);
                if (it != srv->reqs.end()) est += it->second.est_bytes;
            }
            if (c.live == 0 || 4 * dead_bytes(srv, c) < est) continue;
        }
        std::vector<int64_t> ids;
        int64_t need = 0;
        for (int ci : cs)
            for (int64_t id : srv->cohorts[ci]->ids) {
                const Request& r = srv->reqs[id];
                if (r.state != DMDX_REQ_RUNNING) continue;
                ids.push_back(id);
                need += (int64_t)(r.est_bytes * srv->est_scale);
            }
        if (ids.empty() || srv->in_use + need > srv->budget) continue;
        std::vector<dmdx_system_desc> descs;
        for (int64_t id : ids) descs.push_back(srv->reqs[id].desc->d);
        BatchOptions bo;
        bo.reuse_forces = true;
        bo.use_graph = srv->opt.graph_steps > 0;
        bo.graph_steps = std::max(1, (int)srv->opt.graph_steps);
        auto b = std::make_unique<Batch>();
        int rc = b->create(descs.data(), (int)descs.size(), bo);
        if (rc == DMDX_E_OOM) continue;
        if (rc) return rc;
        const int64_t actual = (int64_t)b->device_bytes();
        if (srv->in_use + actual > srv->budget) continue;
        // launch shapes of the first cohort (all cohorts use the defaults)
        b->d.pair_tpa = srv->cohorts[cs[0]]->b->d.pair_tpa;
        b->d.pair_width = srv->cohorts[cs[0]]->b->d.pair_width;
        b->d.coul_tpa = srv->cohorts[cs[0]]->b->d.coul_tpa;
        b->d.coul_width = srv->cohorts[cs[0]]->b->d.coul_width;
        b->d.pair_rows = srv->cohorts[cs[0]]->b->d.pair_rows;
        b->d.build_tpa = srv->cohorts[cs[0]]->b->d.build_tpa;
        for (size_t k = 0; k < ids.size(); ++k) {
            Request& r = srv->reqs[ids[k]];
            rc = b->adopt(*srv->cohorts[r.cohort]->b, r.sys, (int)k);
            if (rc) return rc;
        }
        auto c = std::make_unique<Cohort>();
        c->bytes = actual;
        c->b = std::move(b);
        c->ids = ids;
        c->live = (int)ids.size();
        // retire the old cohorts (their live systems now run in c)
        std::vector<std::unique_ptr<Cohort>> keep;
        for (size_t ci = 0; ci < srv->cohorts.size(); ++ci) {
            if (std::find(cs.begin(), cs.end(), (int)ci) != cs.end()) srv->in_use -= srv->cohorts[ci]->bytes;
            else keep.push_back(std::move(srv->cohorts[ci]));
        }
        keep.push_back(std::move(c));
        srv->in_use += actual;
        srv->cohorts = std::move(keep);
        for (size_t ci = 0; ci < srv->cohorts.size(); ++ci) {
            const auto& ids2 = srv->cohorts[ci]->ids;
            for (size_t k = 0; k < ids2.size(); ++k) {
                Request& r = srv->reqs[ids2[k]];
                if (r.state != DMDX_REQ_RUNNING) continue;
                r.cohort = (int)ci;
                r.sys = (int)k;
            }
        }
        srv->merges += 1;
    }
    return DMDX_OK;
}

int admit(dmdx_server* srv)
{
    while (!srv->pending.empty() && (int)srv->cohorts.size() < srv->opt.max_cohorts) {
        const double t = srv->now();
        const Request& head = srv->reqs[srv->pending.front()];
        const bool enough = (int64_t)srv->pending.size() >= srv->opt.min_admit;
        const bool waited = (t - head.t_submit) * 1e3 >= srv->opt.max_wait_ms;
        if (!enough && !waited && !srv->cohorts.empty()) break;
        // FIFO scan: requests of the head's precision whose estimate fits
        const int prec = head.desc->d.precision;
        int64_t room = srv->budget - srv->in_use;
        std::vector<int64_t> pick;
        std::deque<int64_t> rest;
        for (int64_t id : srv->pending) {
            const Request& r = srv->reqs[id];
            const int64_t need = (int64_t)(r.est_bytes * srv->est_scale);
            const bool cap_ok = srv->opt.max_cohort_systems <= 0 ||
                                (int)pick.size() < srv->opt.max_cohort_systems;
            if (r.desc->d.precision == prec && need <= room && cap_ok) {
                pick.push_back(id);
                room -= need;
            } else {
                rest.push_back(id);
            }
        }
        if (pick.empty()) {
            // the head can never fit: nothing else holds memory
            if (srv->cohorts.empty()) {
                Request& r = srv->reqs[srv->pending.front()];
                r.state = DMDX_REQ_FAILED;
                r.error = "request does not fit in the device-memory budget";
                r.t_done = t;
                r.desc.reset();
                srv->pending.pop_front();
                srv->retired += 1;
                srv->failed += 1;
                continue;
            }
            break;
        }
        std::vector<int64_t> rejected;
        int rc = make_cohort(srv, pick, &rejected);
        if (rc == COHORT_OVER_BUDGET) continue;  // retry with the learned scale
        if (rc == DMDX_E_OOM) {
            // the device had less than the budget says: learn and wait for
            // memory to come back
            srv->est_scale *= 1.5;
            if (srv->cohorts.empty()) continue;
            return DMDX_OK;
        }
        if (rc != DMDX_OK && rc != DMDX_E_PARM && rc != DMDX_E_INVALID && rc != DMDX_E_MISMATCH)
            return rc;
        for (int64_t id : rejected) {
            Request& r = srv->reqs[id];
            r.state = DMDX_REQ_FAILED;
            if (r.error.empty()) r.error = dmdx::last_error();
            r.t_done = t;
            r.desc.reset();
            srv->retired += 1;
            srv->failed += 1;
        }
        srv->pending = rest;
    }
    return DMDX_OK;
}

}  // namespace

extern "C" {

int dmdx_server_create(const dmdx_server_options* opt, dmdx_server** out)
{
    if (!out) return dmdx::fail(DMDX_E_INVALID, "null argument");
    *out = nullptr;
    auto* srv = new dmdx_server();
    std::memset(&srv->opt, 0, sizeof(srv->opt));
    if (opt) srv->opt = *opt;
    if (srv->opt.max_cohorts <= 0) srv->opt.max_cohorts = 8;
    if (srv->opt.chunk_steps <= 0) srv->opt.chunk_steps = 100;
    if (srv->opt.min_admit <= 0) srv->opt.min_admit = 1;
    if (srv->opt.max_wait_ms < 0.0) srv->opt.max_wait_ms = 0.0;
    srv->budget = srv->opt.hbm_budget_bytes;
    if (srv->budget <= 0) {
        size_t f = 0, t = 0;
        hipError_t e = hipMemGetInfo(&f, &t);
        if (e != hipSuccess) {
            delete srv;
            return dmdx::fail(DMDX_E_GPU, std::string("hipMemGetInfo: ") + hipGetErrorString(e));
        }
        srv->budget = (int64_t)(0.9 * (double)f);
    }
    srv->t0 = Clock::now();
    *out = srv;
    return DMDX_OK;
}

int dmdx_server_destroy(dmdx_server* srv)
{
    delete srv;
    return DMDX_OK;
}

int dmdx_server_submit(dmdx_server* srv, const dmdx_system_desc* desc, int64_t n_steps, int64_t* id)
{
    if (!srv || !desc || !id) return dmdx::fail(DMDX_E_INVALID, "null argument");
    if (n_steps <= 0) return dmdx::fail(DMDX_E_INVALID, "a request needs at least one step");
    if (desc->n_atoms <= 0 || !desc->pos_x || !desc->pos_y || !desc->pos_z || !desc->masses)
        return dmdx::fail(DMDX_E_INVALID, "a request needs atoms, positions and masses");
    if (desc->precision != DMDX_FP64 && desc->precision != DMDX_MIXED)
        return dmdx::fail(DMDX_E_INVALID, "unknown precision");
    Request r;
    r.id = srv->next_id++;
    r.desc = std::make_unique<dmdx::DescCopy>(*desc);
    r.n_steps = n_steps;
    r.t_submit = srv->now();
    r.est_bytes = estimate_bytes(*desc);
    const int64_t key = r.id;
    srv->reqs.emplace(key, std::move(r));
    srv->pending.push_back(key);
    *id = key;
    return DMDX_OK;
}

int dmdx_server_pump(dmdx_server* srv, int32_t* n_retired)
{
    if (!srv) return dmdx::fail(DMDX_E_INVALID, "null server");
    const int64_t before = srv->retired;
    int rc = admit(srv);
    if (rc) return rc;
    if (srv->cohorts.empty()) {
        if (n_retired) *n_retired = (int32_t)(srv->retired - before);
        return DMDX_OK;
    }
    // advance every cohort by up to chunk_steps, never past its longest
    // remaining request
    const auto a = Clock::now();
    for (auto& cp : srv->cohorts) {
        Cohort& c = *cp;
        int64_t remain = 0;
        for (int64_t id : c.ids) {
            const Request& r = srv->reqs[id];
            if (r.state == DMDX_REQ_RUNNING) remain = std::max(remain, r.n_steps - r.steps_done);
        }
        c.chunk = std::min<int64_t>(srv->opt.chunk_steps, remain);
        if (c.chunk > 0) {
            rc = c.b->step(c.chunk);
            if (rc) return rc;
        }
    }
    for (auto& cp : srv->cohorts) {
        rc = cp->b->sync();
        if (rc) return rc;
    }
    const double round_ms = std::chrono::duration<double, std::milli>(Clock::now() - a).count();
    srv->rounds += 1;
    // progress, step latency samples, retirement
    for (auto& cp : srv->cohorts) {
        Cohort& c = *cp;
        rc = c.b->fetch_state();
        if (rc) return rc;
        for (size_t k = 0; k < c.ids.size(); ++k) {
            Request& r = srv->reqs[c.ids[k]];
            if (r.state != DMDX_REQ_RUNNING) continue;
            const dmdx::SysState& T = c.b->hstate[k];
            const int64_t adv = T.steps_done - r.steps_done;
            if (adv > 0 && c.chunk > 0) {
                srv->lat_ms.push_back(round_ms / (double)c.chunk);
                srv->lat_w.push_back(adv);
                srv->atom_steps += adv * r.desc->d.n_atoms;
            }
            r.steps_done = T.steps_done;
            r.flags = T.flags;
            if (T.flags != 0) rc = retire(srv, c, r, DMDX_REQ_FAILED);
            else if (r.steps_done >= r.n_steps) rc = retire(srv, c, r, DMDX_REQ_DONE);
            if (rc) return rc;
        }
    }
    // cohorts whose requests have all retired give their memory back
    std::vector<std::unique_ptr<Cohort>> keep;
    for (auto& cp : srv->cohorts) {
        if (cp->live > 0) keep.push_back(std::move(cp));
        else srv->in_use -= cp->bytes;
    }
    srv->cohorts = std::move(keep);
    for (size_t ci = 0; ci < srv->cohorts.size(); ++ci)
        for (int64_t id : srv->cohorts[ci]->ids) {
            Request& r = srv->reqs[id];
            if (r.state == DMDX_REQ_RUNNING) r.cohort = (int)ci;
        }
    if (srv->opt.keep_cohorts == 0) {
        rc = merge(srv);
        if (rc) return rc;
    }
    if (n_retired) *n_retired = (int32_t)(srv->retired - before);
    return DMDX_OK;
}

int dmdx_server_cancel(dmdx_server* srv, int64_t id)
{
    if (!srv) return dmdx::fail(DMDX_E_INVALID, "null server");
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end()) return dmdx::fail(DMDX_E_INVALID, "unknown request");
    Request& r = it->second;
    if (r.state == DMDX_REQ_QUEUED) {
        auto p = std::find(srv->pending.begin(), srv->pending.end(), id);
        if (p != srv->pending.end()) srv->pending.erase(p);
        r.state = DMDX_REQ_CANCELLED;
        r.t_done = srv->now();
        r.desc.reset();
        srv->cancelled += 1;
        return DMDX_OK;
    }
    if (r.state != DMDX_REQ_RUNNING) return dmdx::fail(DMDX_E_INVALID, "request has finished");
    // the batch is idle between pumps: switch the system off (it takes no
    // further step), keep its state as the result; the next merge or
    // compaction leaves it out, which returns its device memory
    Cohort& c = *srv->cohorts[r.cohort];
    int rc = c.b->set_active(r.sys, 0);
    if (rc) return rc;
    rc = c.b->fetch_state();
    if (rc) return rc;
    return retire(srv, c, r, DMDX_REQ_CANCELLED);
}

int dmdx_server_request_info(dmdx_server* srv, int64_t id, dmdx_request_info* out)
{
    if (!srv || !out) return dmdx::fail(DMDX_E_INVALID, "null argument");
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end()) return dmdx::fail(DMDX_E_INVALID, "unknown request");
    const Request& r = it->second;
    std::memset(out, 0, sizeof(*out));
    out->state = r.state;
    out->flags = r.flags;
    out->n_atoms = r.desc ? r.desc->d.n_atoms : (int64_t)r.res[0].size();
    out->steps_target = r.n_steps;
    out->steps_done = r.steps_done;
    out->t_submit = r.t_submit;
    out->t_admit = r.t_admit;
    out->t_done = r.t_done;
    out->est_bytes = (int64_t)(r.est_bytes * srv->est_scale);
    std::strncpy(out->error, r.error.c_str(), sizeof(out->error) - 1);
    return DMDX_OK;
}

int dmdx_server_result(dmdx_server* srv, int64_t id, int32_t what, double* x, double* y, double* z)
{
    if (!srv) return dmdx::fail(DMDX_E_INVALID, "null server");
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end()) return dmdx::fail(DMDX_E_INVALID, "unknown request");
    const Request& r = it->second;
    if (r.state == DMDX_REQ_QUEUED || r.state == DMDX_REQ_RUNNING)
        return dmdx::fail(DMDX_E_BUSY, "request has not finished");
    if (what < 0 || what > 2) return dmdx::fail(DMDX_E_INVALID, "what must be 0, 1 or 2");
    if (r.res[0].empty()) return dmdx::fail(DMDX_E_INVALID, "request produced no state");
    double* dst[3] = {x, y, z};
    for (int k = 0; k < 3; ++k)
        if (dst[k]) std::memcpy(dst[k], r.res[3 * what + k].data(), sizeof(double) * r.res[0].size());
    return DMDX_OK;
}

int dmdx_server_result_scalars(dmdx_server* srv, int64_t id, dmdx_scalars* out)
{
    if (!srv || !out) return dmdx::fail(DMDX_E_INVALID, "null argument");
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end()) return dmdx::fail(DMDX_E_INVALID, "unknown request");
    const Request& r = it->second;
    if (r.res[0].empty()) return dmdx::fail(DMDX_E_BUSY, "request has not finished");
    *out = r.sc;
    return DMDX_OK;
}

int dmdx_server_release(dmdx_server* srv, int64_t id)
{
    if (!srv) return dmdx::fail(DMDX_E_INVALID, "null server");
    auto it = srv->reqs.find(id);
    if (it == srv->reqs.end()) return dmdx::fail(DMDX_E_INVALID, "unknown request");
    if (it->second.state == DMDX_REQ_QUEUED || it->second.state == DMDX_REQ_RUNNING)
        return dmdx::fail(DMDX_E_BUSY, "request has not finished");
    srv->reqs.erase(it);
    return DMDX_OK;
}

int dmdx_server_get_stats(dmdx_server* srv, dmdx_server_stats* out)
{
    if (!srv || !out) return dmdx::fail(DMDX_E_INVALID, "null argument");
    std::memset(out, 0, sizeof(*out));
    out->queued = (int64_t)srv->pending.size();
    int64_t running = 0;
    for (auto& c : srv->cohorts) running += c->live;
    out->running = running;
    out->retired = srv->retired;
    out->failed = srv->failed;
    out->cancelled = srv->cancelled;
    out->cohorts = (int64_t)srv->cohorts.size();
    out->cohorts_made = srv->cohorts_made;
    out->merges = srv->merges;
    out->rounds = srv->rounds;
    out->bytes_in_use = srv->in_use;
    out->bytes_budget = srv->budget;
    out->estimate_scale = srv->est_scale;
    out->clock_s = srv->now();
    out->atom_steps = srv->atom_steps;
    return DMDX_OK;
}

int dmdx_server_step_latency(dmdx_server* srv, double* ms, int64_t* weight, int64_t cap,
                             int64_t* count)
{
    if (!srv || !count) return dmdx::fail(DMDX_E_INVALID, "null argument");
    const int64_t n = (int64_t)srv->lat_ms.size();
    *count = n;
    if (ms && weight) {
        const int64_t m = std::min(cap, n);
        for (int64_t k = 0; k < m; ++k) {
            ms[k] = srv->lat_ms[k];
            weight[k] = srv->lat_w[k];
        }
    }
    return DMDX_OK;
}

}  // extern "C"
