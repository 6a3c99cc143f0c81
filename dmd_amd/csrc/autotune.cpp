// autotune.cpp -- launch-shape autotuner for the mixed-precision kernels.
//
// The per-atom pair kernel of the mixed path gives each atom a group of 2
// to 32 lanes, each taking 2, 4 or 8 list entries per pass (more entries per
// lane keep more gathers in flight per wave, fewer need fewer registers so
// more waves fit; fewer lanes per atom waste fewer lanes on the ragged end of
// a row); LJ-only systems can instead run on row tiles staged in LDS with 4
// or 8 entries per lane block, timed after the best per-atom shape (which
// charged systems keep); and the list build gives each atom 8, 16 or 32 lanes or shares each
// cell's candidates among 2, 4 or 8 waves or stages them in LDS for groups
// of 8 or 16 lanes per atom.  Which is fastest depends on the systems (how
// many neighbours an atom has, how many atoms a batch holds, how many rebuild
// at once), so the choice is measured: a scratch copy of the engine's batch
// is stepped with each candidate and the times of the two kernel groups are
// read from HIP events -- per step for the pair kernel and for the list
// build, whose candidates take turns step by step so that each sees the
// same mix of full and partial rebuilds.  The engine's own state is not
// touched.  Results are cached in a text file keyed by a signature of the
// systems, so a workload is timed once.  There is no counterpart in the
// reference (its loops are serial).
//
// Before the shapes, on an engine that has not stepped yet: the list skin of
// systems that rebuild by the displacement criterion (rebuild_interval -1).
// A wider skin means longer rows (the pair kernel's work grows as
// (rc + skin)^3) and fewer rebuilds (an atom may move skin / 2 between
// them); hot, dense systems such as the melts of a replica mix rebuild every
// few steps at their default skin.  Systems of the same cutoffs and skin form
// a group; each group's skin is scaled by 1, 1.3, 1.6 or 2, one group at a
// time (largest first), timing a scratch batch over 20 steps from the start
// and 20 steps after 60 more; the fastest is taken if it wins by 2%.  Then, the same way,
// each group's cell grid: cells of the list radius or of half of it (the
// density rule picks half-width cells from 40 atoms per list-radius cube;
// finer cells cut the candidates of the list build and change the order in
// which the pair kernel gathers, coarser ones are fewer to scan).  The
// chosen skins and cells replace the engine's (the batch is made again from
// its descriptions, which it has not left).  The forces do not depend on
// either: only the FP32 summation order of the mixed path changes.
//
// Batches with PME then time the FFT's lines per workgroup tile (the
// grid-size rule, or 4, 8, 16) on the PME group of a scratch batch, and the
// engine's batch is made again with a tiling that wins by 2%.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <sstream>
#include <string>
#include <tuple>

#include "engine.h"

namespace {

uint64_t fnv(uint64_t h, const void* p, size_t n)
{
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t k = 0; k < n; ++k) {
        h ^= c[k];
        h *= 1099511628211ull;
    }
    return h;
}

template <class T>
uint64_t mix(uint64_t h, T v)
{
    return fnv(h, &v, sizeof(v));
}

// what the launch shapes depend on: sizes, densities, cutoffs and terms,
// and the cell grid and FFT tiling the batch was made with (so the shapes
// are timed again, not read back, after the autotuner has changed those)
uint64_t signature(const dmdx_engine* e)
{
    uint64_t h = 1469598103934665603ull;
    h = mix(h, (int64_t)e->descs.size());
    for (const auto& dc : e->descs) {
        const dmdx_system_desc& D = dc->d;
        h = mix(h, D.n_atoms);
        for (int k = 0; k < 3; ++k) h = mix(h, (int64_t)std::llround(D.box[k] * 1e3));
        h = mix(h, D.use_lj);
        h = mix(h, (int64_t)std::llround(D.lj_cutoff * 1e4));
        h = mix(h, (int64_t)std::llround(D.list_skin * 1e4));
        h = mix(h, D.coulomb_type);
        h = mix(h, (int64_t)std::llround(D.coulomb_cutoff * 1e4));
        h = mix(h, D.n_excl);
        h = mix(h, D.precision);
        h = mix(h, D.physics);
    }
    for (const auto& S : e->b.hsys) h = mix(h, S.cdiv);
    h = mix(h, e->b.options.fft_lines);
    return h;
}

bool pair_candidate(int p, int w);

// A cached line: key, pair lanes, pair width, per-atom build lanes, then
// optionally the per-cell build switch and its waves per cell (a line
// without them stands for the per-atom build), the row-tile setting, the
// staged per-cell build's lanes per atom (0 when absent) and the Coulomb
// variant's pair lanes and width (8 x 2 when absent).
bool cache_lookup(const char* path, uint64_t key, int* pair, int* width, int* build, int* cells,
                  int* tasks, int* rows, int* stage, int* coul_pair, int* coul_width)
{
    if (!path || !*path) return false;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream ss(line);
        std::string k;
        int p = 0, w = 0, b = 0;
        if (!(ss >> k >> p >> w >> b)) continue;
        if (std::strtoull(k.c_str(), nullptr, 16) == key) {
            *pair = p;
            *width = w;
            *build = b;
            int c = 0, t = 4, r = 0, g = 0, cp = 8, cw = 2;
            if (!(ss >> c >> t) || (c != 0 && c != 1)) c = 0, t = 4;
            else if (!(ss >> r) || (r != 0 && r != 4 && r != 8)) r = 0;
            else if (!(ss >> g) || (g != 0 && g != 8 && g != 16)) g = 0;
            else if (!(ss >> cp >> cw) || !pair_candidate(cp, cw)) cp = 8, cw = 2;
            *cells = c;
            *tasks = t;
            *rows = r;
            *stage = g;
            *coul_pair = cp;
            *coul_width = cw;
            return true;
        }
    }
    return false;
}

// pair kernel candidates: (lanes per atom, entries per lane per pass) of the
// per-atom kernel (for each of its two variants), then row tiles with 4 or
// 8 entries per block; list build candidates: per atom with 8, 16 or 32
// lanes, then per cell with 2, 4 or 8 waves per cell, then per cell staged
// in LDS with 8 or 16 lanes per atom
constexpr int NPAIR = 8, NROWS = 2, NBUILD = 8, NBUILD_ATOM = 3;
constexpr int PAIR_LANES[NPAIR] = {2, 4, 8, 4, 8, 16, 8, 16};
constexpr int PAIR_WIDTH[NPAIR] = {8, 8, 8, 4, 4, 4, 2, 2};
constexpr int ROWS_BLOCK[NROWS] = {4, 8};
constexpr int BUILD_LANES[NBUILD] = {8, 16, 32, 0, 0, 0, 0, 0};
constexpr int BUILD_TASKS[NBUILD] = {0, 0, 0, 2, 4, 8, 4, 4};
constexpr int BUILD_STAGE[NBUILD] = {0, 0, 0, 0, 0, 0, 8, 16};
constexpr int T_ROWS = NPAIR, T_BUILD = NPAIR + NROWS;  // times_ms layout
constexpr int T_COUL = NPAIR + NROWS + NBUILD;
static_assert(T_COUL + NPAIR == DMDX_TUNE_TIMES, "times_ms layout");

// the launch shapes of one batch given to another
void copy_shapes(dmdx::Dev& to, const dmdx::Dev& from)
{
    to.pair_tpa = from.pair_tpa;
    to.pair_width = from.pair_width;
    to.coul_tpa = from.coul_tpa;
    to.coul_width = from.coul_width;
    to.pair_rows = from.pair_rows;
    to.build_tpa = from.build_tpa;
    to.build_cells = from.build_cells;
    to.build_tasks = from.build_tasks;
    to.build_stage = from.build_stage;
}

bool pair_candidate(int p, int w)
{
    for (int c = 0; c < NPAIR; ++c)
        if (PAIR_LANES[c] == p && PAIR_WIDTH[c] == w) return true;
    return false;
}

void cache_store(const char* path, uint64_t key, int pair, int width, int build, int cells,
                 int tasks, int rows, int stage, int coul_pair, int coul_width, const double* t)
{
    if (!path || !*path) return;
    std::ostringstream o;
    char buf[96];
    std::snprintf(buf, sizeof(buf), "%016llx %d %d %d %d %d %d %d %d %d  # pair ms",
                  (unsigned long long)key, pair, width, build, cells, tasks, rows, stage,
                  coul_pair, coul_width);
    o << buf;
    for (int c = 0; c < NPAIR; ++c) {
        std::snprintf(buf, sizeof(buf), " %dx%d %.4f", PAIR_LANES[c], PAIR_WIDTH[c], t[c]);
        o << buf;
    }
    for (int c = 0; c < NROWS; ++c) {
        std::snprintf(buf, sizeof(buf), " rows%d %.4f", ROWS_BLOCK[c], t[T_ROWS + c]);
        o << buf;
    }
    o << " coulomb";
    for (int c = 0; c < NPAIR; ++c) {
        std::snprintf(buf, sizeof(buf), " %dx%d %.4f", PAIR_LANES[c], PAIR_WIDTH[c], t[T_COUL + c]);
        o << buf;
    }
    o << " build ms";
    for (int c = 0; c < NBUILD; ++c) {
        if (BUILD_LANES[c])
            std::snprintf(buf, sizeof(buf), " atom%d %.4f", BUILD_LANES[c], t[T_BUILD + c]);
        else if (BUILD_STAGE[c])
            std::snprintf(buf, sizeof(buf), " stage%d %.4f", BUILD_STAGE[c], t[T_BUILD + c]);
        else
            std::snprintf(buf, sizeof(buf), " cell%d %.4f", BUILD_TASKS[c], t[T_BUILD + c]);
        o << buf;
    }
    std::ofstream f(path, std::ios::app);
    f << o.str() << "\n";
}

constexpr int NSKIN = 4;
constexpr double SKIN_MULT[NSKIN] = {1.0, 1.3, 1.6, 2.0};
constexpr int SKIN_WINDOW = 20;  // steps per timed window
constexpr int SKIN_GAP = 60;     // steps between the two windows

struct SkinGroup {
    std::vector<int> sys;
    int64_t atoms = 0;
};

// Verlet-criterion systems with a skin and a pair list, grouped by cutoffs
// and skin, largest group first
std::vector<SkinGroup> skin_groups(const dmdx_engine* e)
{
    std::map<std::tuple<int64_t, int64_t, int64_t>, int> index;
    std::vector<SkinGroup> g;
    for (size_t s = 0; s < e->descs.size(); ++s) {
        const dmdx_system_desc& D = e->descs[s]->d;
        const bool pairs = D.use_lj || (D.coulomb_type != DMDX_COULOMB_NONE &&
                                        D.physics == DMDX_CORRECTED);
        if (D.rebuild_interval != -1 || !(D.list_skin > 0.0) || !pairs) continue;
        const auto key = std::make_tuple((int64_t)std::llround(D.lj_cutoff * 1e6),
                                         (int64_t)std::llround(D.coulomb_cutoff * 1e6),
                                         (int64_t)std::llround(D.list_skin * 1e6));
        auto it = index.find(key);
        if (it == index.end()) {
            it = index.emplace(key, (int)g.size()).first;
            g.emplace_back();
        }
        g[it->second].sys.push_back((int)s);
        g[it->second].atoms += D.n_atoms;
    }
    std::stable_sort(g.begin(), g.end(),
                     [](const SkinGroup& a, const SkinGroup& b) { return a.atoms > b.atoms; });
    return g;
}

std::vector<dmdx_system_desc> descs_with(const dmdx_engine* e, const std::vector<SkinGroup>& g,
                                         const std::vector<int>& choice)
{
    std::vector<dmdx_system_desc> ds;
    for (const auto& dc : e->descs) ds.push_back(dc->d);
    for (size_t k = 0; k < g.size(); ++k)
        for (int s : g[k].sys) ds[s].list_skin *= SKIN_MULT[choice[k]];
    return ds;
}

// ms per step of a scratch batch with these skins and the engine's launch
// shapes, the list built per cell (cells) or per atom: a window from the
// start and one later (the hot first steps rebuild most)
int time_skins_build(dmdx_engine* e, const std::vector<dmdx_system_desc>& ds,
                     const std::vector<int>& cdiv, bool cells, double* ms)
{
    dmdx::BatchOptions bo;
    bo.reuse_forces = e->b.options.reuse_forces;
    bo.cdiv = cdiv;
    bo.fft_lines = e->b.options.fft_lines;
    dmdx::Batch b;
    int rc = b.create(ds.data(), (int)ds.size(), bo);
    if (rc) return rc;
    copy_shapes(b.d, e->b.d);
    b.d.build_cells = cells;
    using clk = std::chrono::steady_clock;
    rc = b.step(1);
    if (!rc) rc = b.sync();
    const auto t0 = clk::now();
    if (!rc) rc = b.step(SKIN_WINDOW);
    if (!rc) rc = b.sync();
    const auto t1 = clk::now();
    if (!rc) rc = b.step(SKIN_GAP);
    if (!rc) rc = b.sync();
    const auto t2 = clk::now();
    if (!rc) rc = b.step(SKIN_WINDOW);
    if (!rc) rc = b.sync();
    const auto t3 = clk::now();
    if (rc) return rc;
    const double s = std::chrono::duration<double>((t1 - t0) + (t3 - t2)).count();
    *ms = s * 1e3 / (2 * SKIN_WINDOW);
    return DMDX_OK;
}

// per-system cells per list radius: each group's choice (0: the density
// rule), the engine's own setting for systems outside the groups
std::vector<int> cdiv_with(const dmdx_engine* e, const std::vector<SkinGroup>& g,
                           const std::vector<int>& cd)
{
    std::vector<int> v(e->descs.size(), 0);
    for (size_t s = 0; s < v.size() && s < e->b.options.cdiv.size(); ++s)
        v[s] = e->b.options.cdiv[s];
    for (size_t k = 0; k < g.size(); ++k)
        for (int s : g[k].sys) v[s] = cd[k];
    return v;
}

// the faster of the two list builds with these skins and cells (which build
// wins moves with the row length); *cells receives it
int time_skins(dmdx_engine* e, const std::vector<SkinGroup>& g, const std::vector<int>& choice,
               const std::vector<int>& cd, double* ms, bool* cells)
{
    const std::vector<dmdx_system_desc> ds = descs_with(e, g, choice);
    const std::vector<int> cdv = cdiv_with(e, g, cd);
    double tc = 0.0, ta = 0.0;
    int rc = time_skins_build(e, ds, cdv, true, &tc);
    if (!rc) rc = time_skins_build(e, ds, cdv, false, &ta);
    if (rc) return rc;
    *cells = tc <= ta;
    *ms = std::min(tc, ta);
    return DMDX_OK;
}

// "KEY skin s0 s1 ... [cdiv c0 c1 ...]": per group, the skin multiplier's
// index and the cells per list radius (0: the density rule; absent: 0)
bool skin_cache_lookup(const char* path, uint64_t key, size_t ngroups, std::vector<int>* choice,
                       std::vector<int>* cd)
{
    if (!path || !*path) return false;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream ss(line);
        std::string k, tag;
        if (!(ss >> k >> tag) || tag != "skin") continue;
        if (std::strtoull(k.c_str(), nullptr, 16) != key) continue;
        std::vector<int> c, d;
        std::string tok;
        bool in_cdiv = false;
        while (ss >> tok) {
            if (tok == "cdiv") {
                in_cdiv = true;
                continue;
            }
            const int v = std::atoi(tok.c_str());
            if (in_cdiv) {
                if (v < 0 || v > 2) return false;
                d.push_back(v);
            } else {
                if (v < 0 || v >= NSKIN) return false;
                c.push_back(v);
            }
        }
        if (c.size() != ngroups || (!d.empty() && d.size() != ngroups)) return false;
        if (d.empty()) d.assign(ngroups, 0);
        *choice = c;
        *cd = d;
        return true;
    }
    return false;
}

void skin_cache_store(const char* path, uint64_t key, const std::vector<int>& choice,
                      const std::vector<int>& cd)
{
    if (!path || !*path) return;
    std::ofstream f(path, std::ios::app);
    char buf[32];
    std::snprintf(buf, sizeof(buf), "%016llx skin", (unsigned long long)key);
    f << buf;
    for (int c : choice) f << " " << c;
    f << " cdiv";
    for (int c : cd) f << " " << c;
    f << "\n";
}

// Make the engine's batch again from these descriptions and options (only
// for a batch that has not left the state it was created with); on failure
// the engine keeps a batch of its own descriptions and options.
int remake(dmdx_engine* e, const std::vector<dmdx_system_desc>& ds, const dmdx::BatchOptions& opt)
{
    const dmdx::BatchOptions old_opt = e->b.options;
    const dmdx::Dev keep = e->b.d;  // launch shapes
    e->b.~Batch();
    new (&e->b) dmdx::Batch();
    int rc = e->b.create(ds.data(), (int)ds.size(), opt);
    if (rc) {
        // back to the descriptions the engine was created with
        std::vector<dmdx_system_desc> orig;
        for (const auto& dc : e->descs) orig.push_back(dc->d);
        e->b.~Batch();
        new (&e->b) dmdx::Batch();
        const int rc2 = e->b.create(orig.data(), (int)orig.size(), old_opt);
        return rc2 ? rc2 : rc;
    }
    for (size_t s = 0; s < ds.size(); ++s) e->descs[s]->d.list_skin = ds[s].list_skin;
    copy_shapes(e->b.d, keep);
    return DMDX_OK;
}

// *changed: some group's skin or cells changed (the batch was made again)
int tune_skins(dmdx_engine* e, const char* cache_path, bool* changed)
{
    *changed = false;
    const std::vector<SkinGroup> g = skin_groups(e);
    if (g.empty()) return DMDX_OK;
    const uint64_t key = signature(e);
    std::vector<int> best(g.size(), 0), cd(g.size(), 0);
    if (!skin_cache_lookup(cache_path, key, g.size(), &best, &cd)) {
        double tbest = 0.0;
        bool cells = true;
        int rc = time_skins(e, g, best, cd, &tbest, &cells);
        if (rc) return rc;
        // Per group, every candidate is timed and the fastest taken if it
        // beats the group's current setting by 2%.  (Taking a wider skin only
        // when it beat the last one taken by 2% stopped at 1.3x on the
        // replica mix's melts in some runs although 1.6x was faster in every
        // window: the steps just after the start and at steady state.)
        for (size_t k = 0; k < g.size(); ++k) {
            int pick = 0;
            double tpick = tbest;
            for (int c = 1; c < NSKIN; ++c) {
                std::vector<int> trial = best;
                trial[k] = c;
                double t = 0.0;
                rc = time_skins(e, g, trial, cd, &t, &cells);
                if (rc == DMDX_E_PARM) break;  // rows beyond their maximum: no wider skin
                if (rc) return rc;
                if (t < tpick) {
                    pick = c;
                    tpick = t;
                }
            }
            if (pick != 0 && tpick < 0.98 * tbest) {
                best[k] = pick;
                tbest = tpick;
            }
        }
        // cells of the list radius (1) or half of it (2), with the chosen skins
        for (size_t k = 0; k < g.size(); ++k) {
            int pick = 0;
            double tpick = tbest;
            for (int c = 1; c <= 2; ++c) {
                std::vector<int> trial = cd;
                trial[k] = c;
                double t = 0.0;
                rc = time_skins(e, g, best, trial, &t, &cells);
                if (rc) return rc;
                if (t < tpick) {
                    pick = c;
                    tpick = t;
                }
            }
            if (pick != 0 && tpick < 0.98 * tbest) {
                cd[k] = pick;
                tbest = tpick;
            }
        }
        skin_cache_store(cache_path, key, best, cd);
    }
    bool any = false;
    for (size_t k = 0; k < g.size(); ++k) any = any || best[k] != 0 || cd[k] != 0;
    if (!any) return DMDX_OK;
    *changed = true;
    dmdx::BatchOptions opt = e->b.options;
    opt.cdiv = cdiv_with(e, g, cd);
    return remake(e, descs_with(e, g, best), opt);
}

// FFT lines per workgroup tile for batches with PME (see the top of this
// file); *changed: the batch was made again with another tiling
constexpr int NFFT = 4;
constexpr int FFT_LINES[NFFT] = {0, 4, 8, 16};  // 0: the grid-size rule

int tune_fft(dmdx_engine* e, int32_t steps, const char* cache_path, bool* changed)
{
    *changed = false;
    if (!e->b.d.any_pme) return DMDX_OK;
    const uint64_t key = signature(e);
    int best = -1;
    if (cache_path && *cache_path) {
        std::ifstream f(cache_path);
        std::string line;
        while (std::getline(f, line)) {
            std::istringstream ss(line);
            std::string k, tag;
            int v = -1;
            if (!(ss >> k >> tag >> v) || tag != "fft") continue;
            if (std::strtoull(k.c_str(), nullptr, 16) == key &&
                (v == 0 || v == 4 || v == 8 || v == 16))
                best = v;
        }
    }
    if (best < 0) {
        std::vector<dmdx_system_desc> ds;
        for (const auto& dc : e->descs) ds.push_back(dc->d);
        const int n = steps > 0 ? steps : 5;
        double t[NFFT];
        for (int c = 0; c < NFFT; ++c) {
            dmdx::BatchOptions bo;
            bo.reuse_forces = e->b.options.reuse_forces;
            bo.cdiv = e->b.options.cdiv;
            bo.fft_lines = FFT_LINES[c];
            bo.profile = true;
            bo.profile_mask = 1u << dmdx::KG_PME;
            dmdx::Batch b;
            int rc = b.create(ds.data(), (int)ds.size(), bo);
            if (rc) return rc;
            copy_shapes(b.d, e->b.d);
            double ms0, ms1;
            int64_t n0, n1;
            rc = b.step(1);
            if (!rc) rc = b.kernel_time(dmdx::KG_PME, &ms0, &n0);
            if (!rc) rc = b.step(n);
            if (!rc) rc = b.kernel_time(dmdx::KG_PME, &ms1, &n1);
            if (rc) return rc;
            t[c] = (ms1 - ms0) / n;
        }
        int bc = 0;
        for (int c = 1; c < NFFT; ++c)
            if (t[c] < 0.98 * t[bc]) bc = c;
        best = FFT_LINES[bc];
        if (cache_path && *cache_path) {
            std::ofstream f(cache_path, std::ios::app);
            char buf[160];
            std::snprintf(buf, sizeof(buf),
                          "%016llx fft %d  # pme ms rule %.4f lines4 %.4f lines8 %.4f lines16 %.4f\n",
                          (unsigned long long)key, best, t[0], t[1], t[2], t[3]);
            f << buf;
        }
    }
    if (best == e->b.options.fft_lines) return DMDX_OK;
    *changed = true;
    std::vector<dmdx_system_desc> ds;
    for (const auto& dc : e->descs) ds.push_back(dc->d);
    dmdx::BatchOptions opt = e->b.options;
    opt.fft_lines = best;
    return remake(e, ds, opt);
}

// Launch shapes (see the top of this file) for the engine's systems as they
// are; t receives the per-step times of the candidates (-1: not timed).
int tune_shapes(dmdx_engine* e, int32_t steps, const char* cache_path, double* t)
{
    const int* pl = PAIR_LANES;
    const int* pw = PAIR_WIDTH;
    const uint64_t key = signature(e);
    int cp = 0, cw = 0, cb = 0, cc = 0, ct = 4, cr = 0, cg = 0, qp = 8, qw = 2;
    if (cache_lookup(cache_path, key, &cp, &cw, &cb, &cc, &ct, &cr, &cg, &qp, &qw) &&
        pair_candidate(cp, cw) && (cb == 8 || cb == 16 || cb == 32) &&
        (ct == 1 || ct == 2 || ct == 4 || ct == 8)) {
        e->b.d.pair_tpa = cp;
        e->b.d.pair_width = cw;
        e->b.d.coul_tpa = qp;
        e->b.d.coul_width = qw;
        e->b.d.pair_rows = cr;
        e->b.d.build_tpa = cb;
        e->b.d.build_cells = cc != 0;
        e->b.d.build_tasks = ct;
        e->b.d.build_stage = cg;
        e->b.invalidate_graph();
        return DMDX_OK;
    }
    // scratch batch: the same systems from their initial state, stepped as
    // the engine would step them (lists rebuilt when the systems' own rule
    // says so, i.e. mostly partial rebuilds)
    std::vector<dmdx_system_desc> ds;
    for (const auto& dc : e->descs) ds.push_back(dc->d);
    dmdx::BatchOptions bo;
    bo.profile = true;
    bo.reuse_forces = e->b.options.reuse_forces;
    dmdx::Batch scratch;
    int rc = scratch.create(ds.data(), (int)ds.size(), bo);
    if (rc) return rc;
    copy_shapes(scratch.d, e->b.d);
    const int n = steps > 0 ? steps : 5;
    double ms0, ms1;
    int64_t n0, n1;
    auto time_pair = [&](double* out) -> int {
        int r = scratch.step(1);  // first launches of this shape
        if (!r) r = scratch.kernel_time(dmdx::KG_PAIR, &ms0, &n0);
        if (!r) r = scratch.step(n);
        if (!r) r = scratch.kernel_time(dmdx::KG_PAIR, &ms1, &n1);
        if (!r) *out = (ms1 - ms0) / (double)n;
        return r;
    };
    // pair kernel: per-step time of each per-atom shape over n steps, first
    // of the Coulomb variant (the LJ-only systems on the engine's shape),
    // then of the LJ-only variant beside the fastest Coulomb shape, then of
    // the row tiles beside the fastest LJ-only shape; a variant no system
    // uses keeps the engine's shape
    scratch.d.pair_rows = 0;
    // the fastest candidate, unless the engine's current shape (cur, -1 if
    // not a candidate) is within 2% of it: a near tie in a few timed steps
    // is noise (4 x 4 has beaten 4 x 8 by 1% in the trial and lost by 6%
    // over the replica mix's run)
    auto best_of = [&](const double* tt, int cur) {
        int b = 0;
        for (int c = 1; c < NPAIR; ++c)
            if (tt[c] < tt[b]) b = c;
        if (cur >= 0 && tt[b] >= 0.98 * tt[cur]) b = cur;
        return b;
    };
    auto index_of = [&](int lanes, int width) {
        for (int c = 0; c < NPAIR; ++c)
            if (pl[c] == lanes && pw[c] == width) return c;
        return -1;
    };
    const int cur_lj = index_of(e->b.d.pair_tpa, e->b.d.pair_width);
    const int cur_coul = index_of(e->b.d.coul_tpa, e->b.d.coul_width);
    if (scratch.d.any_coul_pairs) {
        for (int c = 0; c < NPAIR; ++c) {
            scratch.d.coul_tpa = pl[c];
            scratch.d.coul_width = pw[c];
            rc = time_pair(&t[T_COUL + c]);
            if (rc) return rc;
        }
        const int bq = best_of(t + T_COUL, cur_coul);
        scratch.d.coul_tpa = pl[bq];
        scratch.d.coul_width = pw[bq];
    }
    if (scratch.d.any_plain_pairs) {
        for (int c = 0; c < NPAIR; ++c) {
            scratch.d.pair_tpa = pl[c];
            scratch.d.pair_width = pw[c];
            rc = time_pair(&t[c]);
            if (rc) return rc;
        }
        const int bp = best_of(t, cur_lj);
        scratch.d.pair_tpa = pl[bp];
        scratch.d.pair_width = pw[bp];
        int br = -1;  // per-atom kernel for every system
        for (int c = 0; c < NROWS; ++c) {
            scratch.d.pair_rows = ROWS_BLOCK[c];
            rc = time_pair(&t[T_ROWS + c]);
            if (rc) return rc;
            if (t[T_ROWS + c] < (br < 0 ? t[bp] : t[T_ROWS + br])) br = c;
        }
        scratch.d.pair_rows = br < 0 ? 0 : ROWS_BLOCK[br];
    }
    // list build: the candidates take turns step by step, so each sees the
    // same mix of full and partial rebuilds; time per step of the list group
    double acc[NBUILD] = {};
    rc = scratch.kernel_time(dmdx::KG_LIST, &ms0, &n0);
    if (rc) return rc;
    for (int k = 0; k < NBUILD * n; ++k) {
        const int c = k % NBUILD;
        scratch.d.build_cells = BUILD_LANES[c] == 0;
        if (BUILD_LANES[c]) scratch.d.build_tpa = BUILD_LANES[c];
        else scratch.d.build_tasks = BUILD_TASKS[c];
        scratch.d.build_stage = BUILD_STAGE[c];
        rc = scratch.step(1);
        if (!rc) rc = scratch.kernel_time(dmdx::KG_LIST, &ms1, &n1);
        if (rc) return rc;
        acc[k % NBUILD] += ms1 - ms0;
        ms0 = ms1;
    }
    for (int c = 0; c < NBUILD; ++c) t[T_BUILD + c] = acc[c] / n;
    // the fastest pair kernel; the fastest list build overall, and the
    // fastest per-atom lane count (the per-atom build serves the systems
    // that need a minimum image even when the per-cell build is chosen)
    int bb = 0, ba = 0;
    for (int c = 1; c < NBUILD; ++c)
        if (t[T_BUILD + c] < t[T_BUILD + bb]) bb = c;
    for (int c = 1; c < NBUILD_ATOM; ++c)
        if (t[T_BUILD + c] < t[T_BUILD + ba]) ba = c;
    e->b.d.pair_tpa = scratch.d.pair_tpa;
    e->b.d.pair_width = scratch.d.pair_width;
    e->b.d.coul_tpa = scratch.d.coul_tpa;
    e->b.d.coul_width = scratch.d.coul_width;
    e->b.d.pair_rows = scratch.d.pair_rows;
    e->b.d.build_tpa = BUILD_LANES[ba];
    e->b.d.build_cells = BUILD_LANES[bb] == 0;
    if (e->b.d.build_cells) e->b.d.build_tasks = BUILD_TASKS[bb];
    e->b.d.build_stage = BUILD_STAGE[bb];
    e->b.invalidate_graph();
    cache_store(cache_path, key, e->b.d.pair_tpa, e->b.d.pair_width, BUILD_LANES[ba],
                e->b.d.build_cells ? 1 : 0, e->b.d.build_tasks, e->b.d.pair_rows,
                e->b.d.build_stage, e->b.d.coul_tpa, e->b.d.coul_width, t);
    return DMDX_OK;
}

}  // namespace

int dmdx_engine_remake(dmdx_engine* e, const dmdx::BatchOptions& opt)
{
    std::vector<dmdx_system_desc> ds;
    for (const auto& dc : e->descs) ds.push_back(dc->d);
    return remake(e, ds, opt);
}

extern "C" int dmdx_engine_autotune(dmdx_engine* e, int32_t steps, const char* cache_path,
                                    int32_t* pair_lanes, int32_t* pair_width,
                                    int32_t* build_lanes, double* times_ms)
{
    if (!e) return dmdx::fail(DMDX_E_INVALID, "null engine");
    double t[DMDX_TUNE_TIMES];
    for (double& v : t) v = -1.0;
    auto report = [&]() {
        if (pair_lanes) *pair_lanes = e->b.d.pair_tpa;
        if (pair_width) *pair_width = e->b.d.pair_width;
        if (build_lanes) *build_lanes = e->b.d.build_tpa;
        if (times_ms)
            for (int k = 0; k < DMDX_TUNE_TIMES; ++k) times_ms[k] = t[k];
    };
    if (e->b.d.precision == DMDX_FP64) {  // one lane per atom, nothing to tune
        report();
        return DMDX_OK;
    }
    if (e->tune_skin && e->b.pristine && (!skin_groups(e).empty() || e->b.d.any_pme)) {
        // shapes for the batch as it is (the skin, cell and FFT trials run
        // with them), then the skins and cells, then the FFT tiling, then the
        // shapes again if the batch was made again
        int rc = tune_shapes(e, steps, cache_path, t);
        bool changed = false, fft_changed = false;
        if (!rc) rc = tune_skins(e, cache_path, &changed);
        if (!rc) rc = tune_fft(e, steps, cache_path, &fft_changed);
        if (rc) return rc;
        if (!changed && !fft_changed) {
            report();
            return DMDX_OK;
        }
        for (double& v : t) v = -1.0;
    }
    const int rc = tune_shapes(e, steps, cache_path, t);
    if (rc) return rc;
    report();
    return DMDX_OK;
}
