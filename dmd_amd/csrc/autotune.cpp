// autotune.cpp -- launch-shape autotuner for the mixed-precision kernels.
//
// The pair kernel of the mixed path gives each atom a group of 2 to 32
// lanes, each taking 2, 4 or 8 list entries per pass (more entries per lane
// keep more gathers in flight per wave, fewer need fewer registers so more
// waves fit; fewer lanes per atom waste fewer lanes on the ragged end of a
// row), and the list build gives each atom 8, 16 or 32 lanes or shares each
// cell's candidates among 2, 4 or 8 waves.  Which is fastest depends on the systems (how
// many neighbours an atom has, how many atoms a batch holds, how many rebuild
// at once), so the choice is measured: a scratch copy of the engine's batch
// is stepped with each candidate and the times of the two kernel groups are
// read from HIP events -- per step for the pair kernel and for the list
// build, whose candidates take turns step by step so that each sees the
// same mix of full and partial rebuilds.  The engine's own state is not
// touched.  Results are cached in a text file keyed by a signature of the
// systems, so a workload is timed once.  There is no counterpart in the
// reference (its loops are serial).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

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

// what the launch shapes depend on: sizes, densities, cutoffs and terms
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
    return h;
}

// A cached line: key, pair lanes, pair width, per-atom build lanes, then
// optionally the per-cell build switch and its waves per cell (a line
// without them stands for the per-atom build).
bool cache_lookup(const char* path, uint64_t key, int* pair, int* width, int* build, int* cells,
                  int* tasks)
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
            int c = 0, t = 4;
            if (!(ss >> c >> t) || (c != 0 && c != 1)) c = 0, t = 4;
            *cells = c;
            *tasks = t;
            return true;
        }
    }
    return false;
}

// pair kernel candidates: (lanes per atom, entries per lane per pass);
// list build candidates: per atom with 8, 16 or 32 lanes, then per cell
// with 2, 4 or 8 waves per cell
constexpr int NPAIR = 8, NBUILD = 6, NBUILD_ATOM = 3;
constexpr int PAIR_LANES[NPAIR] = {2, 4, 8, 4, 8, 16, 8, 16};
constexpr int PAIR_WIDTH[NPAIR] = {8, 8, 8, 4, 4, 4, 2, 2};
constexpr int BUILD_LANES[NBUILD] = {8, 16, 32, 0, 0, 0};
constexpr int BUILD_TASKS[NBUILD] = {0, 0, 0, 2, 4, 8};
static_assert(NPAIR + NBUILD == DMDX_TUNE_TIMES, "times_ms layout");

bool pair_candidate(int p, int w)
{
    for (int c = 0; c < NPAIR; ++c)
        if (PAIR_LANES[c] == p && PAIR_WIDTH[c] == w) return true;
    return false;
}

void cache_store(const char* path, uint64_t key, int pair, int width, int build, int cells,
                 int tasks, const double* t)
{
    if (!path || !*path) return;
    std::ostringstream o;
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%016llx %d %d %d %d %d  # pair ms", (unsigned long long)key,
                  pair, width, build, cells, tasks);
    o << buf;
    for (int c = 0; c < NPAIR; ++c) {
        std::snprintf(buf, sizeof(buf), " %dx%d %.4f", PAIR_LANES[c], PAIR_WIDTH[c], t[c]);
        o << buf;
    }
    o << " build ms";
    for (int c = 0; c < NBUILD; ++c) {
        if (BUILD_LANES[c])
            std::snprintf(buf, sizeof(buf), " atom%d %.4f", BUILD_LANES[c], t[NPAIR + c]);
        else
            std::snprintf(buf, sizeof(buf), " cell%d %.4f", BUILD_TASKS[c], t[NPAIR + c]);
        o << buf;
    }
    std::ofstream f(path, std::ios::app);
    f << o.str() << "\n";
}

}  // namespace

extern "C" int dmdx_engine_autotune(dmdx_engine* e, int32_t steps, const char* cache_path,
                                    int32_t* pair_lanes, int32_t* pair_width,
                                    int32_t* build_lanes, double* times_ms)
{
    if (!e) return dmdx::fail(DMDX_E_INVALID, "null engine");
    const int* pl = PAIR_LANES;
    const int* pw = PAIR_WIDTH;
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
    const uint64_t key = signature(e);
    int cp = 0, cw = 0, cb = 0, cc = 0, ct = 4;
    if (cache_lookup(cache_path, key, &cp, &cw, &cb, &cc, &ct) && pair_candidate(cp, cw) &&
        (cb == 8 || cb == 16 || cb == 32) && (ct == 1 || ct == 2 || ct == 4 || ct == 8)) {
        e->b.d.pair_tpa = cp;
        e->b.d.pair_width = cw;
        e->b.d.build_tpa = cb;
        e->b.d.build_cells = cc != 0;
        e->b.d.build_tasks = ct;
        e->b.invalidate_graph();
        report();
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
    scratch.d.pair_width = e->b.d.pair_width;
    const int n = steps > 0 ? steps : 5;
    double ms0, ms1;
    int64_t n0, n1;
    // pair kernel: per-step time of each shape over n steps
    for (int c = 0; c < NPAIR; ++c) {
        scratch.d.pair_tpa = pl[c];
        scratch.d.pair_width = pw[c];
        rc = scratch.step(1);  // first launches of this shape
        if (!rc) rc = scratch.kernel_time(dmdx::KG_PAIR, &ms0, &n0);
        if (!rc) rc = scratch.step(n);
        if (!rc) rc = scratch.kernel_time(dmdx::KG_PAIR, &ms1, &n1);
        if (rc) return rc;
        t[c] = (ms1 - ms0) / (double)n;
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
        rc = scratch.step(1);
        if (!rc) rc = scratch.kernel_time(dmdx::KG_LIST, &ms1, &n1);
        if (rc) return rc;
        acc[k % NBUILD] += ms1 - ms0;
        ms0 = ms1;
    }
    for (int c = 0; c < NBUILD; ++c) t[NPAIR + c] = acc[c] / n;
    // the fastest pair shape; the fastest list build overall, and the
    // fastest per-atom lane count (the per-atom build serves the systems
    // that need a minimum image even when the per-cell build is chosen)
    int bp = 0, bb = 0, ba = 0;
    for (int c = 1; c < NPAIR; ++c)
        if (t[c] < t[bp]) bp = c;
    for (int c = 1; c < NBUILD; ++c)
        if (t[NPAIR + c] < t[NPAIR + bb]) bb = c;
    for (int c = 1; c < NBUILD_ATOM; ++c)
        if (t[NPAIR + c] < t[NPAIR + ba]) ba = c;
    e->b.d.pair_tpa = pl[bp];
    e->b.d.pair_width = pw[bp];
    e->b.d.build_tpa = BUILD_LANES[ba];
    e->b.d.build_cells = BUILD_LANES[bb] == 0;
    if (e->b.d.build_cells) e->b.d.build_tasks = BUILD_TASKS[bb];
    e->b.invalidate_graph();
    cache_store(cache_path, key, pl[bp], pw[bp], BUILD_LANES[ba], e->b.d.build_cells ? 1 : 0,
                e->b.d.build_tasks, t);
    report();
    return DMDX_OK;
}
