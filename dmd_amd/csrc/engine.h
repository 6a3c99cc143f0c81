// engine.h -- host-side batch object of the step engine.
#pragma once

#include <memory>
#include <vector>

#include "dmdx_internal.h"

namespace dmdx {

const char* last_error();

// Which single ForceComponent a batch stands for (ForceComponent-level API);
// COMP_NONE = the whole ForceEngine of a SimulationEngine.
enum Component {
    COMP_NONE = 0,
    COMP_LJ,
    COMP_COULOMB_DIRECT,
    COMP_EXCLUSION,
    COMP_PME,
    COMP_BOND,
    COMP_ANGLE,
    COMP_DIHEDRAL,
    COMP_IMPROPER
};

struct BatchOptions {
    int component = COMP_NONE;
    bool reuse_forces = true;  // skip a first compute whose result is known
    bool use_graph = false;    // replay captured blocks of steps
    int graph_steps = 10;
    int max_nbr = 0;           // 0: estimate from the density
    // mixed path, per system: cells per list radius along an edge (1 or 2;
    // 0 or absent: the density rule, 2 from 40 atoms per list-radius cube)
    std::vector<int> cdiv;
    // PME: FFT lines per workgroup tile (4, 8 or 16; 0: the grid-size rule),
    // within the LDS limit
    int fft_lines = 0;
    // mixed PME: x and y passes as one xy plane per workgroup (planes within
    // 64 KiB of LDS): 0 never, 1 wherever a plane fits, 2 when the batch has
    // at least 256 such planes (one per CU)
    int fft_planes = 2;
    bool profile = false;      // HIP events around kernel groups
    uint32_t profile_mask = ~0u;  // which groups (bit = group) when profiling
    // allocation-free mode: device arrays are carved from `arena` (capacity
    // arena_bytes) and the streams and events are the caller's, so creating
    // and destroying the batch allocates and frees no device memory
    void* arena = nullptr;
    size_t arena_bytes = 0;
    hipStream_t arena_stream = nullptr, arena_side = nullptr;
    hipEvent_t arena_events[3] = {nullptr, nullptr, nullptr};  // fork, join, list
};

// kernel groups timed when profiling is on
enum KernelGroup { KG_PAIR = 0, KG_LIST, KG_BONDED, KG_PME, KG_GATHER, KG_INTEGRATE, KG_COUNT };

// staging of asynchronous frames (snapshot.hip)
struct SnapshotRing;
struct SnapshotRingDeleter {
    void operator()(SnapshotRing* r) const;
};

// host mirror of what the host needs to decide launches
struct HostSys {
    int n = 0;
    bool has_list = false;
    int interval = 10;
    int counter = 10;
    bool valid = false;
    bool built = false;
    bool active = true;
    bool has_barostat = false;
};

// deep copy of a system description (the library copies its inputs)
struct DescCopy {
    dmdx_system_desc d;
    std::vector<std::vector<double>> dv;
    std::vector<std::vector<int32_t>> iv;

    template <class T>
    const T* keep(std::vector<std::vector<T>>& store, const T* p, int64_t n)
    {
        if (!p || n <= 0) return nullptr;
        store.emplace_back(p, p + n);
        return store.back().data();
    }

    explicit DescCopy(const dmdx_system_desc& s) : d(s)
    {
        const int64_t n = s.n_atoms;
        dv.reserve(32);
        iv.reserve(32);
        d.pos_x = keep(dv, s.pos_x, n);
        d.pos_y = keep(dv, s.pos_y, n);
        d.pos_z = keep(dv, s.pos_z, n);
        d.vel_x = keep(dv, s.vel_x, n);
        d.vel_y = keep(dv, s.vel_y, n);
        d.vel_z = keep(dv, s.vel_z, n);
        d.masses = keep(dv, s.masses, n);
        d.charges = keep(dv, s.charges, n);
        d.atom_types = keep(iv, s.atom_types, n);
        const int64_t t2 = (int64_t)s.lj_n_types * s.lj_n_types;
        d.lj_sigma = keep(dv, s.lj_sigma, t2);
        d.lj_epsilon = keep(dv, s.lj_epsilon, t2);
        d.bond_i = keep(iv, s.bond_i, s.n_bonds);
        d.bond_j = keep(iv, s.bond_j, s.n_bonds);
        d.bond_k = keep(dv, s.bond_k, s.n_bonds);
        d.bond_r0 = keep(dv, s.bond_r0, s.n_bonds);
        d.angle_i = keep(iv, s.angle_i, s.n_angles);
        d.angle_j = keep(iv, s.angle_j, s.n_angles);
        d.angle_k = keep(iv, s.angle_k, s.n_angles);
        d.angle_k_theta = keep(dv, s.angle_k_theta, s.n_angles);
        d.angle_theta0 = keep(dv, s.angle_theta0, s.n_angles);
        d.dih_i = keep(iv, s.dih_i, s.n_dihedrals);
        d.dih_j = keep(iv, s.dih_j, s.n_dihedrals);
        d.dih_k = keep(iv, s.dih_k, s.n_dihedrals);
        d.dih_l = keep(iv, s.dih_l, s.n_dihedrals);
        d.dih_periodicity = keep(iv, s.dih_periodicity, s.n_dihedrals);
        d.dih_k_phi = keep(dv, s.dih_k_phi, s.n_dihedrals);
        d.dih_phi0 = keep(dv, s.dih_phi0, s.n_dihedrals);
        d.imp_i = keep(iv, s.imp_i, s.n_impropers);
        d.imp_j = keep(iv, s.imp_j, s.n_impropers);
        d.imp_k = keep(iv, s.imp_k, s.n_impropers);
        d.imp_l = keep(iv, s.imp_l, s.n_impropers);
        d.imp_k_phi = keep(dv, s.imp_k_phi, s.n_impropers);
        d.imp_phi0 = keep(dv, s.imp_phi0, s.n_impropers);
        d.excl_i = keep(iv, s.excl_i, s.n_excl);
        d.excl_j = keep(iv, s.excl_j, s.n_excl);
        d.con_i = keep(iv, s.con_i, s.n_constraints);
        d.con_j = keep(iv, s.con_j, s.n_constraints);
        d.con_distance = keep(dv, s.con_distance, s.n_constraints);
    }
    DescCopy(const DescCopy&) = delete;
    DescCopy& operator=(const DescCopy&) = delete;
};

// device allocations of a batch, with their total size.  In arena mode every
// array is a 256-byte-aligned region carved from [arena, arena + cap) in
// creation order; nothing is allocated and nothing is freed.
struct DevMem {
    std::vector<void*> ptrs;
    size_t bytes = 0;
    char* arena = nullptr;
    size_t cap = 0;
    hipStream_t zero_stream = nullptr;  // where arena regions are cleared
    void add(void* p, size_t b)
    {
        ptrs.push_back(p);
        bytes += b;
    }
    // next region of `b` bytes, or nullptr when the arena is full
    void* take(size_t b)
    {
        const size_t at = (bytes + 255) / 256 * 256;
        if (at > cap || b > cap - at) return nullptr;
        bytes = at + b;
        return arena + at;
    }
};

class Batch {
public:
    ~Batch();
    int create(const dmdx_system_desc* descs, int nsys, const BatchOptions& opt);
    int step(int64_t nsteps);
    int compute(int accumulate);  // one standalone ForceEngine::compute + energy
    int rebuild_only();           // rebuild_pair_list
    int sync();
    int copy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind);
    int fetch_state();
    int check_sys(int s) const;
    // host arrays, or device arrays when on_device is set
    int get_arrays(int s, int what, double* a, double* b, double* c, bool on_device = false);
    int set_arrays(int s, int what, const double* a, const double* b, const double* c,
                   bool on_device = false);
    // DMD's Cell: box edges from the diagonal of m, the minimum image of
    // `type` (DMDX_CELL_*: identity unless orthorhombic), PME volume from the
    // determinant
    int set_cell(int s, const double m[9], int type);
    int set_counter(int s, int counter);
    int set_active(int s, int active);
    // step targets: system s leaves the batch (becomes inactive) at the start
    // of the first step after steps_done reaches stop[s] (0 = no target)
    int set_stop_at(const std::vector<int64_t>& stop);
    // continue system `src_sys` of `src` as system `dst` of this batch: both
    // were created from the same description; copies the dynamic state
    // (positions, velocities, forces, box, scalars, thermostat/barostat
    // state, list and list counters, cell-local frames), so the system's
    // trajectory is the same as if it had stayed in `src`
    int adopt(Batch& src, int src_sys, int dst);
    int get_pairs(int s, int32_t* pi, int32_t* pj, int64_t cap, int64_t* count);
    // asynchronous frame of system s at the current end of the stream
    // (dmdx_engine_snapshot_begin / _read)
    int snapshot_begin(int s, int64_t* ticket);
    int snapshot_read(int64_t ticket, double* pos, double* vel, double box[3],
                      int64_t* steps_done);
    int upload_sys(int s);
    void invalidate_graph();

    Dev d;
    BatchOptions options;
    // no call has changed the state the batch was created with (no step,
    // compute, list build or set/adopt of any system): re-creating the batch
    // from its descriptions gives the same batch
    bool pristine = true;
    std::vector<SysDev> hsys;
    std::vector<SysState> hstate;
    std::vector<HostSys> host;
    hipStream_t stream = nullptr;
    // second stream for the terms that do not read the pair list (bonded,
    // PME): forked after the per-compute decisions, joined before the gather
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_list = nullptr, ev_join = nullptr;
    size_t device_bytes() const { return owned.bytes; }

private:
    static bool exclusion_term_active(const dmdx_system_desc& D, const BatchOptions& opt);
    bool mirror_compute(int phase, int force_rebuild, bool* any_force);
    int enqueue_compute(int phase, int accumulate, int force_rebuild, bool uniform,
                        bool* defer_pre = nullptr);
    int enqueue_step(bool uniform);

    hipError_t timed(int group, hipError_t (*fn)(const Dev&, hipStream_t),
                     hipStream_t on = nullptr);
    DevMem owned;
    std::unique_ptr<SnapshotRing, SnapshotRingDeleter> snap;
    bool borrowed = false;  // streams and events belong to the caller (arena mode)
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_used;

public:
    int kernel_time(int group, double* total_ms, int64_t* launches);
    double kg_ms[KG_COUNT] = {0};
    int64_t kg_n[KG_COUNT] = {0};

private:
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
};

}  // namespace dmdx

// handle behind dmdx_engine_* (include/dmdx.h)
struct dmdx_engine {
    dmdx::Batch b;
    // host copies of the descriptions (the autotuner builds scratch batches)
    std::vector<std::unique_ptr<dmdx::DescCopy>> descs;
    bool tune_skin = true;  // the autotuner may change skins, cells and FFT tiling
};

// Make an engine's batch again from its descriptions with these options
// (autotune.cpp; for an engine that has not stepped since creation)
int dmdx_engine_remake(dmdx_engine* e, const dmdx::BatchOptions& opt);
