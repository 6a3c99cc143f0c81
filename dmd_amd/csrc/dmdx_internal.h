// dmdx_internal.h -- shared definitions of the gfx950 step engine.
//
// Layout in HBM (one "batch" = one or more independent systems):
//   * atoms of every system are packed in one SoA index space ("slots"),
//     each system's range padded to a multiple of the wavefront (64), so a
//     wavefront never straddles two systems and per-system reductions are
//     per-wave partial sums summed in a fixed order (deterministic, and the
//     same whether a system runs alone or inside a batch);
//   * master state is FP64: x, y, z (unwrapped, as DMD never wraps),
//     vx, vy, vz, fx, fy, fz, 1/m, m, q;
//   * bonded terms of all systems are packed per kind, each system's range
//     padded to 64; every term writes its per-atom contributions to a term
//     force buffer and a per-atom gather adds them in DMD's component order
//     (no force atomics; FP64 accumulation order equals the reference's);
//   * neighbour lists are per system, column-major (neighbour k of all atoms
//     of a system contiguous), so one-thread-per-atom loops read coalesced.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/dmdx.h"

namespace dmdx {

constexpr int WAVE = 64;
constexpr double KB = 0.008314462618;  // src/core/constants.h:6
constexpr double KE = 138.935456;      // src/force/coulomb_direct.cpp:42
constexpr double PI = 3.14159265358979323846;

// ------------------------------------------------------------- errors
void set_error(int status, const std::string& msg);
int fail(int status, const std::string& msg);  // set_error + return status

#define DMDX_HIP(expr)                                                                 \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ::dmdx::fail(e_ == hipErrorOutOfMemory ? DMDX_E_OOM : DMDX_E_GPU,     \
                                std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

// ---------------------------------------------------- per-system constants
// Everything a kernel needs to know about one system.  Lives in device
// memory; box fields are updated on the device by the barostat.
struct SysDev {
    int64_t atom_off;   // first slot
    int32_t n;          // atoms
    int32_t n_pad;      // slots (multiple of 64)
    int32_t physics;    // DMDX_FAITHFUL / DMDX_CORRECTED
    int32_t precision;  // DMDX_FP64 / DMDX_MIXED
    double L[3];        // box
    double invL[3];     // 1.0 / L (FP64 division, as the reference)
    // factor of the minimum image d - round(d * iminv) * L: 1 / L for an
    // orthorhombic cell, 0 for DMD's other cell types, whose image is the
    // identity (src/core/cell.cpp:26-38)
    double iminv[3];
    double vdet;        // > 0: volume of the cell matrix (Cell::volume, the
                        // determinant), used by PME; 0: L[0] * L[1] * L[2]
    // pair terms
    int32_t use_lj;
    int32_t coul_pairs;  // real-space erfc pairs computed from the list
    int32_t ntypes;
    int32_t lj_off;      // offset of the n_types^2 sigma/epsilon tables
    double rc_lj2, rc_c2, alpha;
    // list
    int32_t has_list;
    int32_t interval;    // > 0: computes between rebuilds (DMD cadence);
                         //   0: rebuild when an atom moved > skin/2 (Verlet
                         //      criterion); < 0: explicit rebuilds only
    double half_skin2;   // (skin/2)^2 for the displacement criterion
    double rlist, rlist2;
    int64_t nbr_off;     // start of this system's M x n_pad list block
    int32_t max_nbr;
    int32_t cell_off;    // start of this system's cells
    int32_t nc[3];       // cells per edge at allocation (upper bound)
    int32_t excl_off;    // per-atom exclusion CSR base (slot-indexed)
    int32_t row_off;     // first of this system's nc[1] * nc[2] x-rows of cells
                         // (mixed row-tile pair kernel units)
    int32_t pad_row_;
    // bonded terms: [off, off + n_pad) slots of the packed term arrays
    int64_t bond_off, angle_off, dih_off, imp_off, ex_off;
    int32_t n_bond, n_angle, n_dih, n_imp, n_ex;
    int32_t ex_mode;     // exclusion term active (faithful / corrected form)
    // PME
    int32_t pme;
    int32_t ginf_static;  // mixed path: influence function tabulated (no barostat)
    int32_t cdiv;         // cells per list radius along an edge (FP64 1; mixed 1 or 2)
    int32_t bin_wg;       // binned by one workgroup (k_bin_system)
    int32_t K[3];
    int64_t grid_off;    // complex elements
    int64_t tw_off[3];   // twiddle table offsets (complex elements)
    int64_t bm_off[3];   // B-modulus table offsets
    double self_energy;
    double qsum;         // total charge (neutralising background, CORRECTED)
    // integration
    double dt;
    int32_t thermostat, barostat;
    double T0, ttau, Qnh;
    double P0, ptau, beta, Wab;
    double and_prob;     // Andersen collision probability per call (frequency * dt)
    // SHAKE / RATTLE
    int32_t n_con;       // constraints of this system
    int32_t pad_con_;
    double con_tol;      // ConstraintsParams::tolerance
    // partial-sum ranges (in waves) for the deterministic reductions
    int32_t aw_off, aw_n;      // atom waves
    int32_t bw_off, bw_n;      // bond waves
    int32_t gw_off, gw_n;      // angle waves
    int32_t dw_off, dw_n;      // dihedral waves
    int32_t iw_off, iw_n;      // improper waves
    int32_t xw_off, xw_n;      // exclusion waves
    int32_t pw_off, pw_n;      // PME workgroup partials
};

// Per-system mutable scalars.
struct SysState {
    double pe;           // potential energy of the last compute
    double pe_terms[8];  // lj, coul, bond, angle, dih, imp, pme, excl
    double ke;           // kinetic energy after the closing kick (thermostat input)
    double ke_out;       // kinetic energy of the velocities at the end of the step
    double lambda;       // pending velocity scale (thermostat), applied lazily
    double zeta;         // Nose-Hoover friction
    double eps_dot;      // Andersen piston velocity
    double virial;
    double pressure;
    double xscale;       // barostat position scale of this step
    int32_t counter;     // computes since the last list rebuild
    int32_t rebuild;     // this compute rebuilds the list
    int32_t reuse;       // this compute reuses the previous forces
    int32_t valid;       // forces/energies correspond to the current positions
    int32_t flags;       // bit0 neighbour overflow, bit1 non-finite energy
    int32_t active;      // system participates in steps
    int32_t built;       // a list has been built since creation
    int64_t rebuilds;    // list builds since creation
    int64_t steps_done;
    int64_t stop_at;     // > 0: the system leaves the batch once steps_done reaches it
    int32_t nub[3];      // cells per edge of the last list build
    int32_t pad_;
    int64_t and_calls;   // Andersen apply() calls made (generator seed 42 + calls)
};

enum { FLAG_NBR_OVERFLOW = 1, FLAG_NONFINITE = 2 };

// Mixed-path list entries: bits 0-22 the neighbour's cell-order slot
// relative to its system's first slot, bits 23-31 the neighbour cell's
// offset from the atom's cell along x, y, z (three 3-bit two's complement
// fields, -2..2; the mixed path's cells are at least rlist / cdiv wide with
// cdiv 1 or 2, so neighbours lie within +-cdiv cells).  With cell-local
// coordinates l and cell width w the displacement is d = l_j - l_i + offset
// * w (the minimum image when the system has >= 2 cdiv + 1 cells per edge).
// The pair kernels look the 9-bit code up in 512-entry LDS tables.  (A
// dense 7-bit code (ox + 2) + 5 (oy + 2) + 25 (oz + 2) in a 128-entry
// table made the per-atom pair kernel 8% slower on the replica mix: its
// table reads from LDS took longer.)
// k_bin_system: systems of at most BIN_WG_ATOMS atoms whose allocated cells
// fit in LDS are binned by one workgroup each
constexpr int BIN_WG_ATOMS = 8192;
constexpr int BIN_WG_CELLS = 4096;

constexpr int NBR_SLOT_BITS = 23;
constexpr uint32_t NBR_SLOT_MASK = (1u << NBR_SLOT_BITS) - 1u;
constexpr int NBR_CODES = 512;  // table size of the offset codes
__host__ __device__ __forceinline__ uint32_t nbr_code(int ox, int oy, int oz)
{
    return ((uint32_t)ox & 7u) | (((uint32_t)oy & 7u) << 3) | (((uint32_t)oz & 7u) << 6);
}
// Mixed path: rows are padded to a multiple of 8 entries with this code (an
// x offset of 3 cells, which no neighbour cell has) over the atom's own
// slot; the pair kernels map it to a position far beyond any cutoff, so a
// padding entry adds exactly zero.  (Padding over slot 0 instead made every
// padding gather of a system read one address, and the replica-mix step 2%
// slower.)
constexpr uint32_t NBR_PAD_CODE = 3u;
__host__ __device__ __forceinline__ uint32_t nbr_pad_entry(int own_slot)
{
    return (NBR_PAD_CODE << NBR_SLOT_BITS) | (uint32_t)own_slot;
}

// Charges are spread onto the PME grid as 64-bit fixed-point integers with
// 2^44 units per e: integer atomic adds are associative, so the grid (and
// everything after it) does not depend on the order in which atoms land,
// and a system gives the same result alone or in any batch.  Resolution
// 5.7e-14 e per contribution; |grid value| must stay below 2^19 e.
constexpr double QGRID_SCALE = 17592186044416.0;  // 2^44

// term force buffer: one double3 per (term, role)
struct TermF {
    double x, y, z;
};

// one tile = up to B lines of one system's grid along one dimension
struct FftTile {
    int32_t sys;
    int32_t line0;
    int32_t nlines;
    int32_t tile_local;  // index among this system's tiles of this pass
};

// one tile = up to 256 consecutive cell-order slots of one PME system
struct SpreadTile {
    int32_t sys;
    int32_t cnt;
    int64_t t0;
};

// one owner-computes spreading block of a system's grid
struct GridTile {
    int32_t sys;
    int32_t g0[3];
};

// one xy plane (fixed z) of a system's grid, transformed along x and y in
// LDS by one workgroup (k_fft_plane)
struct PlaneTile {
    int32_t sys;
    int32_t z;
};

struct FftPlan {  // per system and dimension
    int32_t K;
    int32_t nrad;
    int32_t radix[16];
};

// All device pointers of a batch (plain struct, passed to the launchers).
struct Dev {
    int nsys = 0;
    int64_t nslots = 0;  // multiple of 64
    int ncells = 0;
    int precision = DMDX_FP64;
    SysDev* sys = nullptr;
    SysState* st = nullptr;
    int* slot_sys = nullptr;
    double *x = nullptr, *y = nullptr, *z = nullptr;
    double *vx = nullptr, *vy = nullptr, *vz = nullptr;
    double *fx = nullptr, *fy = nullptr, *fz = nullptr;
    // mixed path with gathered terms: per-slot sum of the bonded, PME and
    // exclusion forces of a step's closing compute, gathered on the side
    // stream and added to the pair forces by the closing kick
    double *fbx = nullptr, *fby = nullptr, *fbz = nullptr;
    double *invm = nullptr, *m = nullptr, *q = nullptr;
    int* type = nullptr;
    // lists
    int *acell = nullptr, *ccount = nullptr, *cstart = nullptr, *cfill = nullptr;
    int *cell_atoms = nullptr, *cell_sys = nullptr, *order = nullptr, *nn = nullptr;
    int* nbr = nullptr;  // row-major: row of atom slot t at nbr_off + (t - atom_off) * M
    int nrows = 0;         // x-rows of cells over all systems (allocated)
    int* row_sys = nullptr;  // system of each x-row
    int *ex_start = nullptr, *ex_list = nullptr;
    double4* pos4s = nullptr;  // (x, y, z, q) in cell order (list build; mixed pair kernel)
    double4* xref = nullptr;   // positions at the last list build (displacement criterion)
    float4* pos4f = nullptr;   // mixed list build: wrapped FP32 positions in cell order
    double4* org = nullptr;    // mixed path: origin of each slot's cell-local frame (FP64)
    float4* lpos = nullptr;    // mixed pair kernel: cell-local FP32 positions and charge
    int* slot_of = nullptr;    // mixed path: cell-order slot of each atom slot, relative to its system
    double* p_disp = nullptr;  // per-wave max squared displacement
    bool any_auto = false;     // some system uses the displacement criterion
    int* typ_s = nullptr;      // mixed path: LJ type in cell order
    int pair_tpa = 4;          // mixed pair kernel: lanes per atom (2, 4, 8, 16, 32)
    int build_tpa = 32;        // mixed list build: lanes per atom (8, 16, 32)
    bool build_cells = true;   // mixed list build: per cell where no minimum image is needed
    int build_tasks = 4;       // mixed per-cell build: waves per cell (1, 2, 4, 8)
    // mixed per-cell build with each cell's candidates staged in LDS
    // (k_nbr_build_staged) with 8 or 16 lanes per atom; 0 (default) =
    // k_nbr_build_cells
    int build_stage = 0;
    int pair_width = 8;        // mixed pair kernel: entries per lane per pass (2, 4, 8)
    // the same for the systems with real-space Coulomb pairs (their rows are
    // several times longer and their entries cost several times more)
    int coul_tpa = 8;
    int coul_width = 2;
    // mixed LJ pair forces: 0 = per-atom kernel (k_pair_mixed), 4 or 8 =
    // row-tile kernel (k_pair_rows) with that many list entries per block
    int pair_rows = 0;
    // LJ tables
    double *lj_sig = nullptr, *lj_eps = nullptr;
    float *lj_c12 = nullptr, *lj_c6 = nullptr;
    // bonded (padded term counts)
    int64_t nb = 0, na = 0, nd = 0, ni = 0, ne = 0;
    int *b_sys = nullptr, *bi = nullptr, *bj = nullptr;
    double *bk = nullptr, *br0 = nullptr;
    int *a_sys = nullptr, *ai = nullptr, *aj = nullptr, *ak = nullptr;
    double *akt = nullptr, *at0 = nullptr;
    int *d_sys = nullptr, *di = nullptr, *dj = nullptr, *dk = nullptr, *dl = nullptr;
    int* dper = nullptr;
    double *dkp = nullptr, *dp0 = nullptr;
    int *i_sys = nullptr, *ii = nullptr, *ij = nullptr, *ik = nullptr, *il = nullptr;
    double *ikp = nullptr, *ip0 = nullptr;
    int *e_sys = nullptr, *ei = nullptr, *ej = nullptr;
    TermF* tf = nullptr;
    int64_t tf_bond = 0, tf_angle = 0, tf_dih = 0, tf_imp = 0, tf_ex = 0;
    int *c_start = nullptr, *c_mid = nullptr, *contrib = nullptr;
    // PME
    int64_t grid_total = 0;
    void* grid = nullptr;  // complex, double or float per precision
    unsigned long long* qgrid = nullptr;  // charge spreading, fixed point (QGRID_SCALE)
    int* grid_sys = nullptr;
    void* tw = nullptr;  // twiddles, same real type
    float* ginf = nullptr;  // mixed path: tabulated influence function (fixed boxes)
    // SHAKE / RATTLE: constraints grouped by connected component (atoms
    // linked by constraints), each component's constraints in the caller's
    // order; one thread iterates a component
    int ncomp = 0;
    int* k_sys = nullptr;      // ncomp: system of the component
    int* k_start = nullptr;    // ncomp + 1: first constraint of the component
    int *k_i = nullptr, *k_j = nullptr;  // global slots
    double* k_d = nullptr;     // target distances
    bool any_andersen = false;
    // rebuild worklists (k_rb_list): systems rebuilding their lists this
    // step -- all of them (rb_*), those binned by the multi-kernel path
    // (rbm_*) and those binned one workgroup per system (rbw_*)
    int* rb_sys = nullptr;       // nsys
    int64_t* rb_pre = nullptr;   // nsys + 1: running slot totals
    int64_t* rb_cpre = nullptr;  // nsys + 1: running allocated-cell totals
    int* rb_n = nullptr;         // 1
    int* rbm_sys = nullptr;
    int64_t* rbm_pre = nullptr;
    int64_t* rbm_cpre = nullptr;  // running cell totals
    int* rbm_n = nullptr;
    int* rbw_sys = nullptr;
    int* rbw_n = nullptr;
    int n_bin_wg = 0;            // systems binned by k_bin_system
    bool any_bin_multi = false;  // systems binned by the multi-kernel path
    // mixed path: some system has (or, with a barostat, may come to have)
    // fewer than 2 cdiv + 1 cells along an edge; its list is built per atom
    // with an FP32 minimum image (k_nbr_build_mixed), every other one per
    // cell (k_nbr_build_cells)
    bool any_wrap_mixed = false;
    SpreadTile* stiles = nullptr;  // charge-spreading tiles (atomic route)
    int nstiles = 0;
    GridTile* gtiles = nullptr;    // owner-computes spreading blocks
    int ngtiles = 0;
    double* bmod = nullptr;
    FftPlan* plans = nullptr;
    FftTile* tiles[3] = {nullptr, nullptr, nullptr};
    int ntiles[3] = {0, 0, 0};
    size_t lds_bytes[3] = {0, 0, 0};
    // mixed path: systems whose xy planes fit in LDS run their x and y
    // passes as one plane per workgroup (they have no x or y tiles)
    PlaneTile* ptiles = nullptr;
    int nptiles = 0;
    size_t plane_lds = 0;
    // partial sums (one per wave / PME tile)
    // pair energies: one partial per pair-kernel wave (per-atom kernel) or
    // per atom slot (row-tile kernel, p_rows)
    double *p_lj = nullptr, *p_coul = nullptr, *p_bond = nullptr, *p_angle = nullptr;
    double* p_rows = nullptr;
    double *p_dih = nullptr, *p_imp = nullptr, *p_pme = nullptr, *p_excl = nullptr;
    double *p_ke = nullptr, *p_w = nullptr;
    // what exists in the batch
    bool any_list = false, any_pair = false, any_pme = false, any_excl = false;
    bool any_coul_pairs = false;  // some system computes real-space erfc pairs
    bool any_plain_pairs = false;  // some system does not (LJ-only pair variant)
    bool any_plain_nolist = false;  // some system has no pair terms at all
    bool any_gather = false, any_baro = false;
};

// host launchers (one per kernel group); return hipGetLastError()
hipError_t launch_build_lists(const Dev& d, hipStream_t s);
hipError_t launch_pair(const Dev& d, hipStream_t s, int accumulate);
hipError_t launch_pair_rows(const Dev& d, hipStream_t s, int accumulate);  // LJ-only systems
// effective mixed pair-kernel launch shape of the LJ-only or the Coulomb variant
void pair_shape(const Dev& d, int& tpa, int& w, bool coul);
hipError_t launch_pack(const Dev& d, hipStream_t s);
hipError_t launch_bonded(const Dev& d, hipStream_t s);
hipError_t launch_pme(const Dev& d, hipStream_t s);
hipError_t launch_pre_compute(const Dev& d, hipStream_t s, int phase, int force_rebuild,
                              int reuse_enabled);
hipError_t launch_displacement(const Dev& d, hipStream_t s);
hipError_t launch_pre_rebuild_only(const Dev& d, hipStream_t s);
hipError_t launch_gather(const Dev& d, hipStream_t s);
hipError_t launch_gather_terms(const Dev& d, hipStream_t s);  // into fb (mixed)
// pre: 0, or the step's phase-1 decisions taken in the drift (1; 2 with force
// reuse enabled) when that compute only reuses forces
hipError_t launch_kick_drift(const Dev& d, hipStream_t s, int pre = 0);
hipError_t launch_kick_ke(const Dev& d, hipStream_t s);
hipError_t launch_finalize(const Dev& d, hipStream_t s, int with_kinetic);
hipError_t launch_barostat(const Dev& d, hipStream_t s);
hipError_t launch_materialize(const Dev& d, hipStream_t s);
hipError_t launch_pack_phase(const Dev& d, hipStream_t s, int after_drift);
hipError_t launch_shake(const Dev& d, hipStream_t s);
hipError_t launch_rattle(const Dev& d, hipStream_t s);
hipError_t launch_andersen(const Dev& d, hipStream_t s);
hipError_t launch_ke_refresh(const Dev& d, hipStream_t s);
hipError_t launch_remap_rows(hipStream_t s, int* nbr, int64_t count, int* order, int64_t n_order,
                             int64_t delta);

// allocation-free compute (noalloc.hip): zero `bytes` (a multiple of 4) at p
// with a fill kernel; add the potential energy of system 0 to *energy
hipError_t launch_zero_bytes(void* p, size_t bytes, hipStream_t s);
hipError_t launch_add_energy(double* energy, const SysState* st, hipStream_t s);

// The device-resident scratch array of the allocation-free entry point on
// the current device, with the streams and events its batches borrow.  Set
// up once per device on first use (the array's address is looked up with
// hipGetSymbolAddress while that device is current).
struct NoallocDevice {
    void* scratch = nullptr;  // device address of the array on this device
    size_t bytes = 0;
    hipStream_t stream = nullptr, side = nullptr;
    hipEvent_t events[3] = {nullptr, nullptr, nullptr};
};
int noalloc_device(NoallocDevice** out);

inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Fixed-size grids: a launch covers `units` work units with at most
// MAX_GRID workgroups (a multiple of 8) that loop over them, so launches whose
// units mostly have nothing to do (e.g. systems that are not rebuilding their
// lists) cost a few microseconds instead of the dispatch of one workgroup per
// unit.  See DMDX_FOR_UNITS.
constexpr int64_t MAX_GRID = 2048;
inline unsigned grid_for(int64_t units)
{
    int64_t g = (units + 7) / 8 * 8;
    if (g > MAX_GRID) g = MAX_GRID;
    if (g < 8) g = 8;
    return (unsigned)g;
}

// ------------------------------------------------------- device helpers
// Orthorhombic minimum image with DMD's expression order
// (src/core/cell.cpp:26-38): d - round(d * (1/L)) * L, round half away.
// `inv` must be 1.0 / L computed in FP64 (correctly rounded, so it equals
// the value the reference recomputes per call).
__device__ __forceinline__ double min_image_dmd(double d, double L, double inv)
{
    return d - round(d * inv) * L;
}

// Volume of a system's cell, Cell::volume (src/core/cell.cpp:19-24): the
// determinant of its matrix where one was given, else the determinant of the
// diagonal matrix of the edges, which that expansion evaluates as
// L0 * (L1 * L2) exactly (every other term adds a signed zero).
__host__ __device__ __forceinline__ double sys_volume(const SysDev& S)
{
    return S.vdet > 0.0 ? S.vdet : S.L[0] * (S.L[1] * S.L[2]);
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup p runs on
// XCD p % 8).  Give each XCD a contiguous range of logical blocks, so the
// atoms (and systems) one XCD touches stay in its own L2.
__device__ __forceinline__ unsigned xcd_block()
{
    const unsigned p = blockIdx.x, G = gridDim.x;
    const unsigned x = p & 7u, q = p >> 3;
    const unsigned n = G >> 3, r = G & 7u;
    return (x < r) ? x * (n + 1) + q : r * (n + 1) + (x - r) * n + q;
}

// Plain grid-stride loop over work units [0, U): consecutive units go to
// different workgroups (and XCDs) -- for launches whose work is concentrated
// in a few systems, where spreading beats L2 locality.
#define DMDX_FOR_UNITS_SPREAD(u, U) \
    for (int64_t u = blockIdx.x; u < (int64_t)(U); u += gridDim.x)

// Loop over work units [0, U): each XCD (workgroups p with p % 8 == x) owns
// the contiguous share [U*x/8, U*(x+1)/8) and its workgroups stride through
// it.  gridDim.x must be a multiple of 8 (grid_for).  All threads of a
// workgroup see the same unit sequence.
#define DMDX_FOR_UNITS(u, U)                                                          \
    for (int64_t u##_x = blockIdx.x & 7u, u##_n = gridDim.x >> 3,                    \
                 u##_end = ((int64_t)(U) * (u##_x + 1)) >> 3,                          \
                 u = (((int64_t)(U) * u##_x) >> 3) + (blockIdx.x >> 3);                \
         u < u##_end; u += u##_n)

template <class T>
__device__ __forceinline__ T wave_sum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// sum over the TPA lanes of a lane group (TPA a power of two <= 64)
template <int TPA, class T>
__device__ __forceinline__ T group_sum(T v)
{
    for (int o = TPA / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace dmdx
