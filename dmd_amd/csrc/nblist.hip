// nblist.hip -- cell-grid Verlet list build on gfx950.
//
// Replaces LennardJones::rebuild_pair_list (src/force/lennard_jones.cpp:
// 18-41), an O(N^2) all-pairs loop, with an O(N) cell grid.  The pair
// decision is the reference's own FP64 expression (minimum image with
// round-half-away, r2 = dx*dx + dy*dy + dz*dz left to right, r2 < rc^2), so the
// pair set equals the reference's {i<j : r2 < rc^2} exactly; the list is
// stored "full" (both (i,j) and (j,i)) because the force kernel gives each
// atom its own thread and needs no atomics.
//
// Kernels (all systems of a batch in one launch; systems whose st.rebuild is
// 0 return at once):
//   k_rb_list      worklists of the rebuilding systems (slots, cells)
//   k_bin_system   one workgroup bins one system (counts, scan, fill in LDS)
//   k_cells_reset  the multi-kernel binning of larger systems: zero per-cell
//   k_cell_count   counters; cell of each atom from its wrapped position,
//   k_cell_scan    count; per-system exclusive scan of the counts; scatter
//   k_cell_fill    atoms into cell order
//   k_cell_sort    sort each cell by slot (deterministic order), both paths,
//                  and pack its atoms in cell order
//   k_nbr_build    per atom (cell order): scan the neighbour cells
//   k_nbr_sort     per atom: ascending neighbour order (FP64 path: makes the
//                  per-atom force sum follow the reference's summation order)
//   k_pack         mixed path: cell-local FP32 positions of a compute that
//                  follows no drift, or SHAKE (after a drift, k_kick_drift
//                  and k_cell_sort have written them)
#include "dmdx_internal.h"

#include <type_traits>

namespace dmdx {

// cells in use along one edge: as many as fit with edge >= rlist / cdiv, at
// most the allocated count.  Small counts are handled by the offset sets of
// the scans.
__device__ __forceinline__ int cells_in_use(const SysDev& S, int d)
{
    int nu = (int)floor(S.L[d] * S.cdiv / (S.rlist * (1.0 + 1e-9)));
    if (nu > S.nc[d]) nu = S.nc[d];
    if (nu < 1) nu = 1;
    return nu;
}

// cell of a coordinate along one edge: the fractional coordinate x * (1/L)
// (a multiplication: the cells only bound where the list build looks, and
// the list decisions are exact distance tests, so no division is needed)
__device__ __forceinline__ int cell_coord(double x, double invL, int nu)
{
    double f = x * invL;
    f -= floor(f);
    if (!(f >= 0.0 && f <= 1.0)) f = 0.0;  // non-finite coordinate (a failed system)
    int c = (int)(f * nu);
    return c < 0 ? 0 : (c >= nu ? nu - 1 : c);
}

// ---- rebuild worklist: the systems whose lists are rebuilt in this launch,
// with running totals of their slots and cells, so the list kernels visit
// only those systems' slots and cells (a partial rebuild touches a few
// percent of a batch)
struct RbList {
    const int* sys;        // rebuilding systems, in system order
    const int64_t* pre;    // pre[k]: slots (n_pad) of rebuilding systems before k
    const int64_t* cpre;   // cpre[k]: allocated cells of rebuilding systems before k
    const int* n;          // number of rebuilding systems
};

__device__ __forceinline__ int rb_find(const int64_t* __restrict__ pre, int n, int64_t v)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// one RB_THREADS-thread workgroup: scans over the systems, RB_THREADS at a
// time (per wave with lane shuffles, then over the wave totals).  Three lists: every rebuilding
// system (slots and cells), those binned by the multi-kernel path (slots and
// cells), and those binned by k_bin_system (one workgroup each).
struct RbOut {
    int* sys;
    int64_t* pre;
    int64_t* cpre;
    int* n;
};

constexpr int RB_THREADS = 256;

__global__ void __launch_bounds__(RB_THREADS) k_rb_list(const SysDev* __restrict__ sys,
                                                        const SysState* __restrict__ st,
                                                        int nsys, RbOut all, RbOut multi,
                                                        RbOut wg)
{
    constexpr int NW = RB_THREADS / 64, NQ = 7;
    // quantities: 0-2 list lengths (all, multi, wg), 3 slots of all, 4 slots
    // of multi, 5 cells of multi, 6 cells of all
    __shared__ int64_t wtot[NQ][NW];
    __shared__ int64_t base[NQ];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t < NQ) base[t] = 0;
    __syncthreads();
    for (int s0 = 0; s0 < nsys; s0 += RB_THREADS) {
        const int s = s0 + t;
        const bool rb = s < nsys && st[s].rebuild;
        const bool w = rb && sys[s].bin_wg;
        const int64_t np = rb ? sys[s].n_pad : 0;
        const int64_t nc = rb ? (int64_t)sys[s].nc[0] * sys[s].nc[1] * sys[s].nc[2] : 0;
        const bool m = rb && !w;
        const int64_t v[NQ] = {rb ? 1 : 0, m ? 1 : 0, w ? 1 : 0, np, m ? np : 0, m ? nc : 0, nc};
        int64_t inc[NQ];
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            inc[k] = v[k];
            for (int o = 1; o < 64; o <<= 1) {
                const int64_t y = __shfl_up(inc[k], o, 64);
                if (lane >= o) inc[k] += y;
            }
        }
        if (lane == 63)
#pragma unroll
            for (int k = 0; k < NQ; ++k) wtot[k][wave] = inc[k];
        __syncthreads();
        int64_t ex[NQ];
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
            int64_t b = base[k];
            for (int q = 0; q < wave; ++q) b += wtot[k][q];
            ex[k] = b + inc[k] - v[k];
        }
        if (v[0]) {
            all.sys[ex[0]] = s;
            all.pre[ex[0]] = ex[3];
            all.cpre[ex[0]] = ex[6];
        }
        if (v[1]) {
            multi.sys[ex[1]] = s;
            multi.pre[ex[1]] = ex[4];
            multi.cpre[ex[1]] = ex[5];
        }
        if (v[2]) wg.sys[ex[2]] = s;
        __syncthreads();
        if (t < NQ)
            for (int q = 0; q < NW; ++q) base[t] += wtot[t][q];
        __syncthreads();
    }
    if (t == 0) {
        all.n[0] = (int)base[0];
        all.pre[base[0]] = base[3];
        all.cpre[base[0]] = base[6];
        multi.n[0] = (int)base[1];
        multi.pre[base[1]] = base[4];
        multi.cpre[base[1]] = base[5];
        wg.n[0] = (int)base[2];
    }
}

// global slot of compact slot v, or -1
__device__ __forceinline__ int64_t rb_slot(const RbList& R, const SysDev* __restrict__ sys, int64_t v)
{
    const int n = *R.n;
    if (v >= R.pre[n]) return -1;
    const int k = rb_find(R.pre, n, v);
    return sys[R.sys[k]].atom_off + (v - R.pre[k]);
}

// global cell of compact cell v, or -1
__device__ __forceinline__ int rb_cell(const RbList& R, const SysDev* __restrict__ sys, int64_t v)
{
    const int n = *R.n;
    if (v >= R.cpre[n]) return -1;
    const int k = rb_find(R.cpre, n, v);
    return sys[R.sys[k]].cell_off + (int)(v - R.cpre[k]);
}

// loop of a fixed grid over the compact units [0, U), U read on the device
#define DMDX_FOR_RB_UNITS(u, U) for (int64_t u = blockIdx.x; u < (int64_t)(U); u += gridDim.x)

// A worklist's length and running totals (slots, or cells when `cells`)
// staged in LDS by one coalesced read per workgroup when the list is short,
// so the binary search of every lookup reads LDS instead of a chain of
// dependent global loads.  Every thread of the workgroup calls it, before
// the unit loop (it synchronises the workgroup).
constexpr int RB_LDS = 512;
struct RbStage {
    int64_t tot[RB_LDS + 1];
    int n;
};

__device__ __forceinline__ RbList rb_stage(RbList R, bool cells, RbStage& L)
{
    const int n = *R.n;
    if (n <= RB_LDS) {  // uniform over the workgroup
        const int64_t* src = cells ? R.cpre : R.pre;
        for (int i = threadIdx.x; i <= n; i += blockDim.x) L.tot[i] = src[i];
        if (threadIdx.x == 0) L.n = n;
        __syncthreads();
        if (cells) R.cpre = L.tot;
        else R.pre = L.tot;
        R.n = &L.n;
    }
    return R;
}

__device__ __forceinline__ void k_cells_reset_unit(int c, const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                              const int* __restrict__ cell_sys, int ncells_total,
                              int* __restrict__ ccount, int* __restrict__ cfill)
{
    if (c < 0 || c >= ncells_total) return;
    int s = cell_sys[c];
    if (!st[s].rebuild) return;
    ccount[c] = 0;
    cfill[c] = 0;
}

__global__ void k_cells_reset(const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                              const int* __restrict__ cell_sys, int ncells_total,
                              int* __restrict__ ccount, int* __restrict__ cfill, RbList R)
{
    DMDX_FOR_RB_UNITS(u_, (R.cpre[*R.n] + 255) / 256)
        k_cells_reset_unit(rb_cell(R, sys, u_ * 256 + threadIdx.x), sys, st, cell_sys, ncells_total,
                           ccount, cfill);
}

__device__ __forceinline__ void k_cell_count_unit(int64_t g, const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                             const int* __restrict__ slot_sys, int64_t nslots,
                             const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, int* __restrict__ acell,
                             int* __restrict__ ccount)
{
    if (g < 0 || g >= nslots) return;
    int s = slot_sys[g];
    if (s < 0 || !st[s].rebuild) return;
    const SysDev& S = sys[s];
    if (g - S.atom_off >= S.n) return;
    int nx = cells_in_use(S, 0), ny = cells_in_use(S, 1), nz = cells_in_use(S, 2);
    int cx = cell_coord(x[g], S.invL[0], nx);
    int cy = cell_coord(y[g], S.invL[1], ny);
    int cz = cell_coord(z[g], S.invL[2], nz);
    int c = (cz * ny + cy) * nx + cx;
    acell[g] = c;
    atomicAdd(&ccount[S.cell_off + c], 1);
}

__global__ void k_cell_count(const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                             const int* __restrict__ slot_sys, int64_t nslots,
                             const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, int* __restrict__ acell,
                             int* __restrict__ ccount, RbList R)
{
    DMDX_FOR_RB_UNITS(u_, (R.pre[*R.n] + 255) / 256)
        k_cell_count_unit(rb_slot(R, sys, u_ * 256 + threadIdx.x), sys, st, slot_sys, nslots, x, y,
                          z, acell, ccount);
}

// one workgroup (256 threads) per system
__global__ void k_cell_scan(const SysDev* __restrict__ sys, SysState* __restrict__ st,
                            const int* __restrict__ ccount, int* __restrict__ cstart)
{
    int s = blockIdx.x;
    if (!st[s].rebuild) return;
    const SysDev& S = sys[s];
    if (S.bin_wg) return;  // binned by k_bin_system
    int ncell = cells_in_use(S, 0) * cells_in_use(S, 1) * cells_in_use(S, 2);
    if (threadIdx.x < 3) st[s].nub[threadIdx.x] = cells_in_use(S, threadIdx.x);
    __shared__ int part[256];
    int t = threadIdx.x;
    int chunk = (ncell + 255) / 256;
    int b = t * chunk, e = min(ncell, b + chunk);
    int sum = 0;
    for (int c = b; c < e; ++c) sum += ccount[S.cell_off + c];
    part[t] = sum;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        int v = (t >= o) ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - sum;  // exclusive prefix of this chunk
    for (int c = b; c < e; ++c) {
        cstart[S.cell_off + c] = run;
        run += ccount[S.cell_off + c];
    }
}

__device__ __forceinline__ void k_cell_fill_unit(int64_t g, const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                            const int* __restrict__ slot_sys, int64_t nslots,
                            const int* __restrict__ acell, const int* __restrict__ cstart,
                            int* __restrict__ cfill, int* __restrict__ cell_atoms)
{
    if (g < 0 || g >= nslots) return;
    int s = slot_sys[g];
    if (s < 0 || !st[s].rebuild) return;
    const SysDev& S = sys[s];
    if (g - S.atom_off >= S.n) return;
    int c = acell[g];
    int pos = atomicAdd(&cfill[S.cell_off + c], 1);
    cell_atoms[S.atom_off + cstart[S.cell_off + c] + pos] = (int)g;
}

__global__ void k_cell_fill(const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                            const int* __restrict__ slot_sys, int64_t nslots,
                            const int* __restrict__ acell, const int* __restrict__ cstart,
                            int* __restrict__ cfill, int* __restrict__ cell_atoms, RbList R)
{
    DMDX_FOR_RB_UNITS(u_, (R.pre[*R.n] + 255) / 256)
        k_cell_fill_unit(rb_slot(R, sys, u_ * 256 + threadIdx.x), sys, st, slot_sys, nslots, acell,
                         cstart, cfill, cell_atoms);
}

// one 64-thread workgroup per cell: bitonic sort of the cell's slots in LDS
// (deterministic order within a cell; cells above the LDS capacity fall back
// to an insertion sort by one thread)
// After the cell sort, per atom (global slot i at cell-order slot t of cell
// lc of system S): positions in cell order (contiguous per cell, read by the
// list build) and the reference positions of the displacement criterion;
// mixed path: the cell-local frame, its origin, and the pair kernel's frame
// position and type.
struct PackOut {
    double4* pos4s;
    double4* xref;
    float4* pos4f;  // null: FP64 path
    double4* org;
    float4* lpos;
    int* typ_s;
};

__device__ __forceinline__ void pack_atom(const SysDev& S, int64_t t, int i, int lc,
                                          const double* __restrict__ x,
                                          const double* __restrict__ y,
                                          const double* __restrict__ z,
                                          const double* __restrict__ q,
                                          const int* __restrict__ type, const PackOut& o)
{
    o.pos4s[t] = make_double4(x[i], y[i], z[i], q[i]);
    o.xref[i] = make_double4(x[i], y[i], z[i], 0.0);
    if (o.pos4f) {
        // mixed path: cell-local frame centred on the atom's cell,
        // l = wrapped position - cell centre (|l| <= w/2 at the build, so FP32
        // keeps ~3e-8 nm), origin = x - l; pos4f holds l for the list build
        double wx = x[i] - floor(x[i] * S.invL[0]) * S.L[0];
        double wy = y[i] - floor(y[i] * S.invL[1]) * S.L[1];
        double wz = z[i] - floor(z[i] * S.invL[2]) * S.L[2];
        const int nx = cells_in_use(S, 0), ny = cells_in_use(S, 1), nz = cells_in_use(S, 2);
        const int cx = lc % nx, cy = (lc / nx) % ny, cz = lc / (nx * ny);
        const double lx = wx - (cx + 0.5) * (S.L[0] / nx);
        const double ly = wy - (cy + 0.5) * (S.L[1] / ny);
        const double lz = wz - (cz + 0.5) * (S.L[2] / nz);
        o.pos4f[t] = make_float4((float)lx, (float)ly, (float)lz, 0.0f);
        const double4 g = make_double4(x[i] - lx, y[i] - ly, z[i] - lz, q[i]);
        o.org[t] = g;
        // the pair kernel's frame position and type, as k_pack writes them
        o.lpos[t] = make_float4((float)(x[i] - g.x), (float)(y[i] - g.y), (float)(z[i] - g.z),
                                (float)g.w);
        o.typ_s[t] = type[i];
    }
}

// Sort each cell of the rebuilding systems (both binning paths) by slot,
// the deterministic order inside a cell, and pack its atoms (pack_atom): one
// wave per cell, four cells per workgroup, spread over the whole device.  Up
// to 64 atoms: each lane holds one, ranked against the others read back lane
// by lane (distinct slots, so the ranks are a permutation); up to
// CELL_RANK_MAX: eight per lane; larger cells: one lane's insertion sort.
// Every lane has read all m values before any writes (one wave, and the
// ranks need every value).
constexpr int CELL_RANK_MAX = 512;

__global__ void __launch_bounds__(256) k_cell_sort(const SysDev* __restrict__ sys,
                                                   const int* __restrict__ ccount,
                                                   const int* __restrict__ cstart,
                                                   int* __restrict__ cell_atoms,
                                                   const double* __restrict__ x,
                                                   const double* __restrict__ y,
                                                   const double* __restrict__ z,
                                                   const double* __restrict__ q,
                                                   const int* __restrict__ type, PackOut po,
                                                   RbList R)
{
    constexpr int RPL = CELL_RANK_MAX / 64;
    const int lane = threadIdx.x & 63;
    __shared__ RbStage rs;
    R = rb_stage(R, true, rs);
    const int nrb = *R.n;
    const int64_t U = R.cpre[nrb];
    for (int64_t u = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); u < U;
         u += (int64_t)gridDim.x * 4) {
        const int k = rb_find(R.cpre, nrb, u);
        const SysDev& S = sys[R.sys[k]];
        const int lc = (int)(u - R.cpre[k]);
        if (lc >= cells_in_use(S, 0) * cells_in_use(S, 1) * cells_in_use(S, 2)) continue;
        const int c = S.cell_off + lc;
        const int m = ccount[c];
        if (m == 0) continue;
        const int64_t t0 = S.atom_off + cstart[c];  // cell-order slot of the cell's first atom
        int* a = cell_atoms + t0;
        if (m <= 64) {
            const int v = lane < m ? a[lane] : 0x7fffffff;
            int rank = 0;
            // j is wave-uniform: a lane read into a scalar register
            for (int j = 0; j < m; ++j) rank += __builtin_amdgcn_readlane(v, j) < v ? 1 : 0;
            if (lane < m) {
                a[rank] = v;
                pack_atom(S, t0 + rank, v, lc, x, y, z, q, type, po);
            }
        } else if (m <= CELL_RANK_MAX) {
            int v[RPL], rank[RPL];
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                v[r] = lane + 64 * r < m ? a[lane + 64 * r] : 0x7fffffff;
                rank[r] = 0;
            }
#pragma unroll
            for (int b = 0; b < RPL; ++b) {
                const int lim = min(64, m - 64 * b);
                for (int j = 0; j < lim; ++j) {
                    const int aj = __builtin_amdgcn_readlane(v[b], j);
#pragma unroll
                    for (int r = 0; r < RPL; ++r) rank[r] += aj < v[r] ? 1 : 0;
                }
            }
#pragma unroll
            for (int r = 0; r < RPL; ++r)
                if (lane + 64 * r < m) {
                    a[rank[r]] = v[r];
                    pack_atom(S, t0 + rank[r], v[r], lc, x, y, z, q, type, po);
                }
        } else {
            if (lane == 0) {
                for (int i = 1; i < m; ++i) {
                    const int v = a[i];
                    int j = i - 1;
                    while (j >= 0 && a[j] > v) {
                        a[j + 1] = a[j];
                        --j;
                    }
                    a[j + 1] = v;
                }
            }
            __threadfence_block();  // lane 0's stores, visible to the wave
            __builtin_amdgcn_wave_barrier();
            for (int k = lane; k < m; k += 64) pack_atom(S, t0 + k, a[k], lc, x, y, z, q, type, po);
        }
    }
}

// ---- one workgroup bins one system (systems of at most BIN_WG_ATOMS atoms
// whose allocated cells fit in LDS): counts, exclusive scan, fill and the
// deterministic order inside each cell all happen in LDS; global memory sees
// one coalesced read of the positions and the cell arrays written.  Same
// cells and the same cell order as the multi-kernel path above, in one
// launch for all rebuilding systems of that kind.
constexpr int BIN_THREADS = 1024;
constexpr int BIN_PER_THREAD = BIN_WG_ATOMS / BIN_THREADS;

__global__ void __launch_bounds__(BIN_THREADS) k_bin_system(
    const SysDev* __restrict__ sys, SysState* __restrict__ st, const int* __restrict__ wg_sys,
    const int* __restrict__ wg_n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, int* __restrict__ acell, int* __restrict__ ccount,
    int* __restrict__ cstart, int* __restrict__ cell_atoms)
{
    __shared__ int fill[BIN_WG_CELLS];   // counts, then fill pointers
    __shared__ int slots[BIN_WG_ATOMS];  // the system's atoms in cell order
    __shared__ int wsum[BIN_THREADS / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nw = *wg_n;
    for (int k = blockIdx.x; k < nw; k += gridDim.x) {
        const int s = wg_sys[k];
        const SysDev& S = sys[s];
        const int nx = cells_in_use(S, 0), ny = cells_in_use(S, 1), nz = cells_in_use(S, 2);
        const int ncell = nx * ny * nz;
        const int nc3 = S.nc[0] * S.nc[1] * S.nc[2];
        const int64_t a0 = S.atom_off;
        const int n = S.n;
        if (tid < 3) st[s].nub[tid] = tid == 0 ? nx : (tid == 1 ? ny : nz);
        for (int c = tid; c < ncell; c += BIN_THREADS) fill[c] = 0;
        __syncthreads();
        // counts; each thread keeps its atoms' cells in registers
        int cl[BIN_PER_THREAD];
#pragma unroll
        for (int r = 0; r < BIN_PER_THREAD; ++r) {
            const int g = tid + r * BIN_THREADS;
            cl[r] = 0;
            if (g < n) {
                const int64_t a = a0 + g;
                const int c = (cell_coord(z[a], S.invL[2], nz) * ny +
                               cell_coord(y[a], S.invL[1], ny)) * nx +
                              cell_coord(x[a], S.invL[0], nx);
                cl[r] = c;
                acell[a] = c;
                atomicAdd(&fill[c], 1);
            }
        }
        __syncthreads();
        // exclusive scan over the cells: each thread a contiguous chunk, then
        // a scan of the chunk sums over the workgroup
        const int chunk = (ncell + BIN_THREADS - 1) / BIN_THREADS;
        const int cb = tid * chunk, ce = min(ncell, cb + chunk);
        int sum = 0;
        for (int c = cb; c < ce; ++c) sum += fill[c];
        int incl = sum;  // inclusive scan within the wave
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(incl, o, 64);
            if (lane >= o) incl += v;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        int wbase = 0;
        for (int w = 0; w < wave; ++w) wbase += wsum[w];
        int run = wbase + incl - sum;
        for (int c = cb; c < ce; ++c) {
            const int cnt = fill[c];
            ccount[S.cell_off + c] = cnt;
            cstart[S.cell_off + c] = run;
            fill[c] = run;
            run += cnt;
        }
        for (int c = ncell + tid; c < nc3; c += BIN_THREADS) {  // allocated, not in use
            ccount[S.cell_off + c] = 0;
            cstart[S.cell_off + c] = n;
        }
        __syncthreads();
        // fill (arbitrary order inside a cell)
#pragma unroll
        for (int r = 0; r < BIN_PER_THREAD; ++r) {
            const int g = tid + r * BIN_THREADS;
            if (g < n) slots[atomicAdd(&fill[cl[r]], 1)] = (int)(a0 + g);
        }
        __syncthreads();
        // cells in order, atoms of a cell in arbitrary order: k_cell_sort
        // orders and packs each cell, spread over the whole device
        for (int lt = tid; lt < n; lt += BIN_THREADS) cell_atoms[a0 + lt] = slots[lt];
        __syncthreads();  // LDS is reused by the next system
    }
}

// per atom: squared displacement since the last build, per-wave maximum
__device__ __forceinline__ void k_displacement_unit(int64_t unit_, const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                               const int* __restrict__ slot_sys, int64_t nslots,
                               const double* __restrict__ x, const double* __restrict__ y,
                               const double* __restrict__ z, const double4* __restrict__ xref,
                               double* __restrict__ p_disp)
{
    int64_t g = unit_ * blockDim.x + threadIdx.x;
    if (g >= nslots) return;
    int s = slot_sys[g];
    if (s < 0) return;
    const SysDev& S = sys[s];
    double d2 = 0.0;
    if (S.interval == 0 && st[s].active && st[s].built && g - S.atom_off < S.n) {
        double4 r = xref[g];
        double dx = x[g] - r.x, dy = y[g] - r.y, dz = z[g] - r.z;
        d2 = dx * dx + dy * dy + dz * dz;
    }
    for (int o = 32; o > 0; o >>= 1) d2 = fmax(d2, __shfl_xor(d2, o, 64));
    if ((threadIdx.x & 63) == 0) p_disp[g >> 6] = d2;
}

__global__ void k_displacement(const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                               const int* __restrict__ slot_sys, int64_t nslots,
                               const double* __restrict__ x, const double* __restrict__ y,
                               const double* __restrict__ z, const double4* __restrict__ xref,
                               double* __restrict__ p_disp, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_) k_displacement_unit(u_, sys, st, slot_sys, nslots, x, y, z, xref, p_disp);
}

// neighbour-cell offsets along one edge: {-1,0,1} with >= 3 cells, {0,1}
// with 2, {0} with 1 -- every cell of a periodic edge visited exactly once
__device__ __forceinline__ int edge_offsets(int nu, int* off)
{
    if (nu >= 3) { off[0] = -1; off[1] = 0; off[2] = 1; return 3; }
    if (nu == 2) { off[0] = 0; off[1] = 1; return 2; }
    off[0] = 0;
    return 1;
}

__device__ __forceinline__ bool is_excluded(const int* __restrict__ ex_start,
                                            const int* __restrict__ ex_list, int64_t i, int j)
{
    int b = ex_start[i], e = ex_start[i + 1];
    for (int k = b; k < e; ++k)
        if (ex_list[k] == j) return true;
    return false;
}

// A group of TPA lanes handles the atom of slot t (cell order): the lanes
// test TPA candidates of a neighbour cell at a time, compact the hits with a
// wavefront ballot and append them to the atom's row of the list, so a row
// is written with contiguous stores.  The list is row-major: the neighbours of
// slot t occupy nbr[nbr_off + (t - atom_off) * M + k].
//   FP64 path (this kernel, one lane per atom): entries are atom slots
//               (summation order is restored by k_nbr_sort);
//   mixed path (k_nbr_build_mixed): entries are cell-order slots, i.e.
//               indices into the packed cell-ordered positions (k_pack),
//               which keeps the gathers of the pair kernel spatially local,
//               with the neighbour cell's offset in the top bits
//               (NBR_SLOT_BITS): the pair kernel rebuilds the displacement
//               from FP32 cell-local coordinates.
template <int TPA>
__device__ __forceinline__ void k_nbr_build_unit(int64_t unit_, 
    const SysDev* __restrict__ sys, SysState* __restrict__ st, const int* __restrict__ slot_sys,
    int64_t nslots, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const int* __restrict__ acell, const int* __restrict__ ccount,
    const int* __restrict__ cstart, const int* __restrict__ cell_atoms,
    const int* __restrict__ ex_start, const int* __restrict__ ex_list, int* __restrict__ order,
    int* __restrict__ nn, int* __restrict__ nbr, const double4* __restrict__ pos4s,
    const float4* __restrict__ pos4f)
{
    constexpr int GPB = 256 / TPA;
    const int lane = threadIdx.x % TPA;
    const int64_t t = unit_ * GPB + threadIdx.x / TPA;
    if (t >= nslots) return;
    const int s = slot_sys[t];
    if (s < 0 || !st[s].rebuild) return;
    const SysDev& S = sys[s];
    const int lt = (int)(t - S.atom_off);
    if (lt >= S.n) {
        if (lane == 0) {
            nn[t] = 0;
            order[t] = -1;
        }
        return;
    }
    const int i = cell_atoms[t];
    if (lane == 0) order[t] = i;
    const int nx = cells_in_use(S, 0), ny = cells_in_use(S, 1), nz = cells_in_use(S, 2);
    const int ci = acell[i];
    const int cx = ci % nx, cy = (ci / nx) % ny, cz = ci / (nx * ny);
    int ox[3], oy[3], oz[3];
    const int nox = edge_offsets(nx, ox), noy = edge_offsets(ny, oy), noz = edge_offsets(nz, oz);
    const double4 pi4 = pos4s[t];
    const double xi = pi4.x, yi = pi4.y, zi = pi4.z;
    const double Lx = S.L[0], Ly = S.L[1], Lz = S.L[2];
    const double ix = S.iminv[0], iy = S.iminv[1], iz = S.iminv[2];
    const double rl2 = S.rlist2;
    const bool check_ex = ex_start[i + 1] > ex_start[i];
    const bool sorted_ids = S.precision != DMDX_FP64;
    int* row = nbr + S.nbr_off + (int64_t)lt * S.max_nbr;
    const int M = S.max_nbr;
    const int gshift = (threadIdx.x & 63) / TPA * TPA;  // first lane of my group in the wave
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int k = 0;
    for (int a = 0; a < noz; ++a) {
        int zz = cz + oz[a];
        zz = zz < 0 ? zz + nz : (zz >= nz ? zz - nz : zz);
        for (int b = 0; b < noy; ++b) {
            int yy = cy + oy[b];
            yy = yy < 0 ? yy + ny : (yy >= ny ? yy - ny : yy);
            for (int c = 0; c < nox; ++c) {
                int xx = cx + ox[c];
                xx = xx < 0 ? xx + nx : (xx >= nx ? xx - nx : xx);
                const int cell = S.cell_off + (zz * ny + yy) * nx + xx;
                const uint32_t shift = nbr_code(ox[c], oy[b], oz[a]) << NBR_SLOT_BITS;
                const int64_t base = S.atom_off + cstart[cell];
                const int cnt = ccount[cell];
                for (int m0 = 0; m0 < cnt; m0 += TPA) {
                    const int m = m0 + lane;
                    bool hit = false;
                    int j = -1;
                    if (m < cnt && base + m != t) {
                        {
                            const double4 pj = pos4s[base + m];
                            double dx = pj.x - xi;
                            double dy = pj.y - yi;
                            double dz = pj.z - zi;
                            dx = min_image_dmd(dx, Lx, ix);
                            dy = min_image_dmd(dy, Ly, iy);
                            dz = min_image_dmd(dz, Lz, iz);
                            double r2 = dx * dx + dy * dy + dz * dz;
                            hit = r2 < rl2;
                        }
                        if (hit && (check_ex || !sorted_ids)) j = cell_atoms[base + m];
                        if (hit && check_ex) hit = !is_excluded(ex_start, ex_list, i, j);
                    }
                    if (TPA == 1) {
                        if (hit) {
                            if (k < M) row[k] = sorted_ids ? (int)((uint32_t)(base + m) | shift) : j;
                            ++k;
                        }
                    } else {
                        const uint64_t bal = __ballot(hit);
                        const uint64_t mine = (bal >> gshift) & (TPA == 64 ? ~0ull : ((1ull << TPA) - 1));
                        const int pos = k + __popcll(mine & below);
                        if (hit && pos < M) row[pos] = sorted_ids ? (int)((uint32_t)(base + m) | shift) : j;
                        k += __popcll(mine);
                    }
                }
            }
        }
    }
    if (k > M) {
        if (lane == 0) atomicOr(&st[s].flags, FLAG_NBR_OVERFLOW);
        k = M;
    }
    if (lane == 0) nn[t] = k;
}

template <int TPA>
__global__ void __launch_bounds__(256) k_nbr_build(
    const SysDev* __restrict__ sys, SysState* __restrict__ st, const int* __restrict__ slot_sys,
    int64_t nslots, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const int* __restrict__ acell, const int* __restrict__ ccount,
    const int* __restrict__ cstart, const int* __restrict__ cell_atoms,
    const int* __restrict__ ex_start, const int* __restrict__ ex_list, int* __restrict__ order,
    int* __restrict__ nn, int* __restrict__ nbr, const double4* __restrict__ pos4s,
    const float4* __restrict__ pos4f, int64_t nunits_)
{
    DMDX_FOR_UNITS_SPREAD(u_, nunits_) k_nbr_build_unit<TPA>(u_, sys, st, slot_sys, nslots, x, y, z, acell, ccount, cstart, cell_atoms, ex_start, ex_list, order, nn, nbr, pos4s, pos4f);
}

// Mixed path.  Cells of at least rlist / cdiv (SysDev::cdiv: 2 for dense
// systems, else 1), neighbour cells within +-cdiv, atoms of a cell in slot
// order, ballot compaction into contiguous rows, on the cell-local FP32
// frames: a candidate's displacement is l_j - l_i + offset * w, with the FP32
// minimum image only when a system has fewer than 2 cdiv + 1 cells along an
// edge.  Candidates within 1e-5 nm of the list
// radius are decided by the reference's own FP64 expression on pos4s, so
// the pair set is exactly the FP64 one.  Entries carry the cell offset
// (NBR_SLOT_BITS) for the pair kernel.
// neighbour cells of one atom on the mixed path: 5^3 with half-width cells,
// 3^3 otherwise, at most every cell of a small system (< 5 cells per edge)
constexpr int NB_CELLS = 125;

// A mixed-path system with at least 2 cdiv + 1 cells along every edge needs
// no minimum image in its list build: its list is built per cell.
__device__ __forceinline__ bool mixed_builds_per_cell(const SysDev& S)
{
    const int full = 2 * S.cdiv + 1;
    return full * full * full <= NB_CELLS && cells_in_use(S, 0) >= full &&
           cells_in_use(S, 1) >= full && cells_in_use(S, 2) >= full;
}

template <int TPA>
__device__ __forceinline__ void k_nbr_build_mixed_unit(int64_t t,
    const SysDev* __restrict__ sys, SysState* __restrict__ st, const int* __restrict__ slot_sys,
    int64_t nslots, const int* __restrict__ acell, const int* __restrict__ ccount,
    const int* __restrict__ cstart, const int* __restrict__ cell_atoms,
    const int* __restrict__ ex_start, const int* __restrict__ ex_list, int* __restrict__ order,
    int* __restrict__ slot_of, int* __restrict__ nn, int* __restrict__ nbr,
    const double4* __restrict__ pos4s, const float4* __restrict__ pos4f,
    int (*tab_base)[NB_CELLS], int (*tab_pre)[NB_CELLS + 1], float4 (*tab_sh)[NB_CELLS],
    bool cells_on)
{
    const int lane = threadIdx.x % TPA;
    if (t < 0 || t >= nslots) return;  // whole waves (compact ranges are multiples of 64)
    // the wave's atoms are consecutive slots of one system (systems own
    // whole 64-slot blocks): system constants are scalar
    const int s = __builtin_amdgcn_readfirstlane(slot_sys[t]);
    if (s < 0 || !st[s].rebuild) return;
    const SysDev& S = sys[s];
    if (cells_on && mixed_builds_per_cell(S)) return;  // k_nbr_build_cells owns this system
    const int lt = (int)(t - S.atom_off);
    if (lt >= S.n) {
        if (lane == 0) {
            nn[t] = 0;
            order[t] = -1;
        }
        return;
    }
    const int i = cell_atoms[t];
    if (lane == 0) {
        order[t] = i;
        slot_of[i] = lt;
    }
    const int nx = cells_in_use(S, 0), ny = cells_in_use(S, 1), nz = cells_in_use(S, 2);
    const int ci = acell[i];
    const int cx = ci % nx, cy = (ci / nx) % ny, cz = ci / (nx * ny);
    const float4 li = pos4f[t];
    const float wx = (float)(S.L[0] / nx), wy = (float)(S.L[1] / ny), wz = (float)(S.L[2] / nz);
    // neighbour cells: +-cdiv with >= 2 cdiv + 1 cells (cells are at least
    // rlist / cdiv wide); with fewer cells, every cell once, and an FP32
    // minimum image per candidate
    const int reach = S.cdiv, full = 2 * reach + 1;
    const bool wrap = nx < full || ny < full || nz < full;
    const float lo2 = (float)((S.rlist - 1e-5) * (S.rlist - 1e-5));
    const float hi2 = (float)((S.rlist + 1e-5) * (S.rlist + 1e-5));
    // the atom's excluded partners, held in registers when there are few
    // (chains: 1-2 and 1-3 partners; water: the other two atoms)
    const int exb = ex_start[i], exn = ex_start[i + 1] - exb;
    const bool check_ex = exn > 0;
    constexpr int EXR = 6;
    int exr[EXR];
#pragma unroll
    for (int q = 0; q < EXR; ++q) exr[q] = q < exn ? ex_list[exb + q] : -1;
    int* row = nbr + S.nbr_off + (int64_t)lt * S.max_nbr;
    const int M = S.max_nbr;
    const int gshift = (threadIdx.x & 63) / TPA * TPA;  // first lane of my group in the wave
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    auto lo_of = [&](int nu) { return nu >= full ? -reach : -((nu - 1) / 2); };
    auto hi_of = [&](int nu) { return nu >= full ? reach : nu - 1 - (nu - 1) / 2; };
    const int lox = lo_of(nx), hix = hi_of(nx);
    const int loy = lo_of(ny), hiy = hi_of(ny);
    const int loz = lo_of(nz), hiz = hi_of(nz);
    // The atom's candidates, as one stream: the neighbour cells in scan order
    // (z, then y, then x offset, each ascending), each cell's atoms in cell
    // order -- the order of the candidate numbers k written below.  The
    // group's TPA lanes first tabulate the cells (cell-order start, running
    // count, offsets) in LDS, then take candidates k, k + TPA, ...: every
    // pass but the last is full, however the atoms split over cells.
    const int ex_n = hix - lox + 1, ey_n = hiy - loy + 1;
    const int ncb = ex_n * ey_n * (hiz - loz + 1);  // <= NB_CELLS
    const int grp = (threadIdx.x / TPA);
    int* tb = tab_base[grp];
    int* tp = tab_pre[grp];
    // per neighbour cell: its offset times the cell widths, minus l_i (x, y,
    // z), and the offset fields of its list entries (w, as int bits)
    float4* tsh = tab_sh[grp];
    {
        // lane l takes cells l, l + TPA, ...; all their loads are issued
        // before any count is used
        constexpr int Q = (NB_CELLS + TPA - 1) / TPA;
        int cnt_q[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int jc = q * TPA + lane;
            cnt_q[q] = 0;
            if (jc < ncb) {
                const int ox = lox + jc % ex_n, oy = loy + (jc / ex_n) % ey_n,
                          oz = loz + jc / (ex_n * ey_n);
                int xx = cx + ox, yy = cy + oy, zz = cz + oz;
                xx = xx < 0 ? xx + nx : (xx >= nx ? xx - nx : xx);
                yy = yy < 0 ? yy + ny : (yy >= ny ? yy - ny : yy);
                zz = zz < 0 ? zz + nz : (zz >= nz ? zz - nz : zz);
                const int c = S.cell_off + (zz * ny + yy) * nx + xx;
                tb[jc] = cstart[c];
                cnt_q[q] = ccount[c];
                const uint32_t code = nbr_code(ox, oy, oz) << NBR_SLOT_BITS;
                tsh[jc] = make_float4((float)ox * wx - li.x, (float)oy * wy - li.y,
                                      (float)oz * wz - li.z, __uint_as_float(code));
            }
        }
        // running counts in scan order: a group scan per round of TPA cells
        int run = 0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            if (q * TPA >= ncb) break;
            int incl = cnt_q[q];
            for (int o = 1; o < TPA; o <<= 1) {
                const int v = __shfl_up(incl, o, TPA);
                if (lane >= o) incl += v;
            }
            const int jc = q * TPA + lane;
            if (jc < ncb) tp[jc] = run + incl - cnt_q[q];
            run += __shfl(incl, TPA - 1, TPA);
        }
        if (lane == 0) tp[ncb] = run;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const int ncand = tp[ncb];
    // the wave-uniform minimum-image switch selects a loop variant outside
    // the candidate loop
    auto scan = [&](auto wrap_c) -> int {
        constexpr bool WRAP = decltype(wrap_c)::value;
        int k = 0;
        int jc = 0;  // the lane's current cell: tp[jc] <= candidate < tp[jc + 1]
        auto locate = [&](int kc) {
            while (jc + 1 < ncb && tp[jc + 1] <= kc) ++jc;
        };
        locate(lane);
        int slot_n = lane < ncand ? tb[jc] + (lane - tp[jc]) : 0;
        float4 nxt = lane < ncand ? pos4f[S.atom_off + slot_n] : make_float4(0.f, 0.f, 0.f, 0.f);
        for (int k0 = 0; k0 < ncand; k0 += TPA) {
            const int kc = k0 + lane;
            const float4 lj = nxt;
            const int slot = slot_n;
            const float4 sh = tsh[kc < ncand ? jc : 0];
            const uint32_t offs = kc < ncand ? __float_as_uint(sh.w) : 0u;
            // the next pass's candidate is located and its position loaded
            // before this one is tested
            if (kc + TPA < ncand) {
                locate(kc + TPA);
                slot_n = tb[jc] + (kc + TPA - tp[jc]);
                nxt = pos4f[S.atom_off + slot_n];
            }
            bool hit = false;
            if (kc < ncand && S.atom_off + slot != t) {
                float dx = lj.x + sh.x, dy = lj.y + sh.y, dz = lj.z + sh.z;
                if constexpr (WRAP) {
                    dx -= rintf(dx * (float)S.iminv[0]) * (float)S.L[0];
                    dy -= rintf(dy * (float)S.iminv[1]) * (float)S.L[1];
                    dz -= rintf(dz * (float)S.iminv[2]) * (float)S.L[2];
                }
                const float r2f = dx * dx + dy * dy + dz * dz;
                hit = r2f < lo2;
                if (!hit && r2f < hi2) {
                    // the reference's expression (src/core/cell.cpp:26-38,
                    // src/force/lennard_jones.cpp:29-36)
                    const double4 pi4 = pos4s[t];
                    const double4 pj = pos4s[S.atom_off + slot];
                    double ddx = min_image_dmd(pj.x - pi4.x, S.L[0], S.iminv[0]);
                    double ddy = min_image_dmd(pj.y - pi4.y, S.L[1], S.iminv[1]);
                    double ddz = min_image_dmd(pj.z - pi4.z, S.L[2], S.iminv[2]);
                    hit = ddx * ddx + ddy * ddy + ddz * ddz < S.rlist2;
                }
                if (hit && check_ex) {
                    const int j = cell_atoms[S.atom_off + slot];
                    if (exn <= EXR) {
                        bool ex = false;
#pragma unroll
                        for (int q = 0; q < EXR; ++q) ex |= exr[q] == j;
                        hit = !ex;
                    } else {
                        hit = !is_excluded(ex_start, ex_list, i, j);
                    }
                }
            }
            const uint64_t bal = __ballot(hit);
            const uint64_t mine = (bal >> gshift) & (TPA == 64 ? ~0ull : ((1ull << TPA) - 1));
            const int pos = k + __popcll(mine & below);
            if (hit && pos < M) row[pos] = (int)((uint32_t)slot | offs);
            k += __popcll(mine);
        }
        return k;
    };
    int k = wrap ? scan(std::true_type{}) : scan(std::false_type{});
    if (k > M) {
        if (lane == 0) atomicOr(&st[s].flags, FLAG_NBR_OVERFLOW);
        k = M;
    }
    if (lane == 0) nn[t] = k;
    // pad the row to a multiple of 8 entries (M is one), so the pair kernel's
    // 8-entry vector loads past the row end read entries that add zero
    for (int p = k + lane; p < ((k + 7) & ~7); p += TPA) row[p] = (int)nbr_pad_entry(lt);
}

template <int TPA>
__global__ void __launch_bounds__(256) k_nbr_build_mixed(
    const SysDev* __restrict__ sys, SysState* __restrict__ st, const int* __restrict__ slot_sys,
    int64_t nslots, const int* __restrict__ acell, const int* __restrict__ ccount,
    const int* __restrict__ cstart, const int* __restrict__ cell_atoms,
    const int* __restrict__ ex_start, const int* __restrict__ ex_list, int* __restrict__ order,
    int* __restrict__ slot_of, int* __restrict__ nn, int* __restrict__ nbr,
    const double4* __restrict__ pos4s, const float4* __restrict__ pos4f, RbList R,
    bool cells_on)
{
    constexpr int GPB = 256 / TPA;  // atoms per workgroup pass
    __shared__ int tab_base[GPB][NB_CELLS], tab_pre[GPB][NB_CELLS + 1];
    __shared__ float4 tab_sh[GPB][NB_CELLS];
    __shared__ RbStage rs;
    R = rb_stage(R, false, rs);
    DMDX_FOR_RB_UNITS(u_, (R.pre[*R.n] + GPB - 1) / GPB)
        k_nbr_build_mixed_unit<TPA>(rb_slot(R, sys, u_ * GPB + threadIdx.x / TPA), sys, st,
                                    slot_sys, nslots, acell, ccount, cstart, cell_atoms,
                                    ex_start, ex_list, order, slot_of, nn, nbr, pos4s, pos4f,
                                    tab_base, tab_pre, tab_sh, cells_on);
}

// Mixed path, systems built per cell (mixed_builds_per_cell): `tasks`
// waves per cell of a rebuilding system, task t taking the cell's atoms t,
// t + tasks, ...  Every atom of a cell has the same candidate stream -- the
// neighbour cells in scan order (z, then y, then x offset, each ascending),
// each cell's atoms in cell order, as in k_nbr_build_mixed -- so a wave
// loads each candidate once for all its atoms: 64 at a time, CB_CHUNKS
// chunks per batch held in registers, shifted into the frame of the wave's
// cell (l_j + offset * w).  Each of the wave's atoms (broadcast from the
// lane that preloaded it) is then tested against the batch and its hits are
// appended to its row in stream order with a ballot.  Several waves per
// cell keep each wave's serial chain short and the device full.  The decisions are
// those of k_nbr_build_mixed (FP32 test outside a 1e-5 nm band around the
// list radius, the reference's FP64 expression inside it), so the rows are
// the same, entry for entry.
constexpr int CB_WAVES = 4;   // cells per workgroup pass
constexpr int CB_CHUNKS = 4;  // 64-candidate chunks per register batch
constexpr int CB_EXR = 6;     // excluded partners held in registers

__global__ void __launch_bounds__(64 * CB_WAVES) k_nbr_build_cells(
    const SysDev* __restrict__ sys, SysState* __restrict__ st, const int* __restrict__ ccount,
    const int* __restrict__ cstart, const int* __restrict__ cell_atoms,
    const int* __restrict__ ex_start, const int* __restrict__ ex_list, int* __restrict__ order,
    int* __restrict__ slot_of, int* __restrict__ nn, int* __restrict__ nbr,
    const double4* __restrict__ pos4s, const float4* __restrict__ pos4f, RbList R, int tasks)
{
    __shared__ int tb_all[CB_WAVES][NB_CELLS];        // cell-order start of each neighbour cell
    __shared__ int tp_all[CB_WAVES][NB_CELLS + 1];    // candidates before it in the stream
    __shared__ float4 tsh_all[CB_WAVES][NB_CELLS];    // offset * w, and the entry's offset bits
    __shared__ RbStage rs;
    R = rb_stage(R, true, rs);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int* tb = tb_all[wv];
    int* tp = tp_all[wv];
    float4* tsh = tsh_all[wv];
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int nrb = *R.n;
    const int64_t U = R.cpre[nrb];
    for (int64_t w = (int64_t)blockIdx.x * CB_WAVES + wv; w < U * tasks;
         w += (int64_t)gridDim.x * CB_WAVES) {
        const int64_t u = w / tasks;
        const int task = (int)(w - u * tasks);
        const int kq = rb_find(R.cpre, nrb, u);
        const int s = R.sys[kq];
        const SysDev& S = sys[s];
        if (S.precision == DMDX_FP64 || !mixed_builds_per_cell(S)) continue;
        const int lc = (int)(u - R.cpre[kq]);
        if (lc == 0 && task == 0)  // the system's padding slots hold no atom
            for (int p = S.n + lane; p < S.n_pad; p += 64) {
                nn[S.atom_off + p] = 0;
                order[S.atom_off + p] = -1;
            }
        const int nx = cells_in_use(S, 0), ny = cells_in_use(S, 1), nz = cells_in_use(S, 2);
        if (lc >= nx * ny * nz) continue;
        const int c = S.cell_off + lc;
        const int m = ccount[c];
        const int mt = (m - task + tasks - 1) / tasks;  // this task's atoms
        if (mt <= 0) continue;
        const int cx = lc % nx, cy = (lc / nx) % ny, cz = lc / (nx * ny);
        const float wx = (float)(S.L[0] / nx), wy = (float)(S.L[1] / ny),
                    wz = (float)(S.L[2] / nz);
        const int reach = S.cdiv, full = 2 * reach + 1;
        const int ncb = full * full * full;  // 27 or 125
        // neighbour-cell table, two rounds of 64 cells, then the exclusive
        // running count of candidates in scan order
        int cnt[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int jc = q * 64 + lane;
            cnt[q] = 0;
            if (jc < ncb) {
                const int ox = jc % full - reach, oy = (jc / full) % full - reach,
                          oz = jc / (full * full) - reach;
                int xx = cx + ox, yy = cy + oy, zz = cz + oz;
                xx = xx < 0 ? xx + nx : (xx >= nx ? xx - nx : xx);
                yy = yy < 0 ? yy + ny : (yy >= ny ? yy - ny : yy);
                zz = zz < 0 ? zz + nz : (zz >= nz ? zz - nz : zz);
                const int cc = S.cell_off + (zz * ny + yy) * nx + xx;
                tb[jc] = cstart[cc];
                cnt[q] = ccount[cc];
                const uint32_t code = nbr_code(ox, oy, oz) << NBR_SLOT_BITS;
                tsh[jc] = make_float4((float)ox * wx, (float)oy * wy, (float)oz * wz,
                                      __uint_as_float(code));
            }
        }
        int run = 0;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            int incl = cnt[q];
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(incl, o, 64);
                if (lane >= o) incl += v;
            }
            if (q * 64 + lane < ncb) tp[q * 64 + lane] = run + incl - cnt[q];
            run += __shfl(incl, 63, 64);
        }
        if (lane == 0) tp[ncb] = run;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const int ncand = run;
        const int M = S.max_nbr;
        const float lo2 = (float)((S.rlist - 1e-5) * (S.rlist - 1e-5));
        const float hi2 = (float)((S.rlist + 1e-5) * (S.rlist + 1e-5));
        const int64_t t0 = S.atom_off + cstart[c];  // the cell's first cell-order slot
        for (int ig = 0; ig < mt; ig += 64) {
            const int gm = min(64, mt - ig);
            // the group's atoms (the task's atoms ig, ig + 1, ...: cell atoms
            // task + tasks * (ig + lane)), one per lane: frame position, atom
            // slot, excluded partners; broadcast to the wave atom by atom below
            const bool mine = lane < gm;
            const int64_t tl = t0 + task + (int64_t)tasks * (ig + (mine ? lane : 0));
            const float4 lmy = pos4f[tl];
            const int imy = cell_atoms[tl];
            const int exb = ex_start[imy];
            const int exn = mine ? ex_start[imy + 1] - exb : 0;
            int exr[CB_EXR];
#pragma unroll
            for (int q = 0; q < CB_EXR; ++q) exr[q] = q < exn ? ex_list[exb + q] : -1;
            int kreg = 0;  // row length of the lane's atom so far
            int jc = 0;    // the lane's current neighbour cell (candidates only increase)
            for (int b0 = 0; b0 < ncand; b0 += 64 * CB_CHUNKS) {
                float px[CB_CHUNKS], py[CB_CHUNKS], pz[CB_CHUNKS];
                uint32_t ent[CB_CHUNKS];
                int jid[CB_CHUNKS];  // the candidate's atom slot
                const int nch = min(CB_CHUNKS, (ncand - b0 + 63) / 64);  // wave-uniform
#pragma unroll
                for (int r = 0; r < CB_CHUNKS; ++r) {
                    const int kc = b0 + r * 64 + lane;
                    // past the end of the stream: a candidate outside every cutoff
                    px[r] = py[r] = pz[r] = 1e18f;
                    ent[r] = 0u;
                    jid[r] = -1;
                    if (kc < ncand) {
                        while (jc + 1 < ncb && tp[jc + 1] <= kc) ++jc;
                        const int slot = tb[jc] + (kc - tp[jc]);
                        const float4 p = pos4f[S.atom_off + slot];
                        const float4 sh = tsh[jc];
                        px[r] = p.x + sh.x;
                        py[r] = p.y + sh.y;
                        pz[r] = p.z + sh.z;
                        ent[r] = (uint32_t)slot | __float_as_uint(sh.w);
                        jid[r] = cell_atoms[S.atom_off + slot];
                    }
                }
                for (int ii = 0; ii < gm; ++ii) {
                    const float lx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(lmy.x), ii));
                    const float ly = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(lmy.y), ii));
                    const float lz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(lmy.z), ii));
                    const int i = __builtin_amdgcn_readlane(imy, ii);
                    const int nex = __builtin_amdgcn_readlane(exn, ii);
                    int ex[CB_EXR];
#pragma unroll
                    for (int q = 0; q < CB_EXR; ++q) ex[q] = __builtin_amdgcn_readlane(exr[q], ii);
                    const int lti = (int)(t0 + task + (int64_t)tasks * (ig + ii) - S.atom_off);
                    int* row = nbr + S.nbr_off + (int64_t)lti * M;
                    int k = __builtin_amdgcn_readlane(kreg, ii);
#pragma unroll
                    for (int r = 0; r < CB_CHUNKS; ++r) {
                        if (r >= nch) break;
                        const uint32_t e = ent[r];
                        const int slot = (int)(e & NBR_SLOT_MASK);
                        const float dx = px[r] - lx, dy = py[r] - ly, dz = pz[r] - lz;
                        const float r2f = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                        const bool other = slot != lti;
                        bool hit = other && r2f < lo2;
                        const bool band = other && r2f >= lo2 && r2f < hi2;
                        if (__ballot(band) != 0ull && band) {
                            // within 1e-5 nm of the list radius: the reference's
                            // expression decides (src/core/cell.cpp:26-38,
                            // src/force/lennard_jones.cpp:29-36)
                            const double4 pi4 = pos4s[S.atom_off + lti];
                            const double4 pj = pos4s[S.atom_off + slot];
                            const double ddx = min_image_dmd(pj.x - pi4.x, S.L[0], S.iminv[0]);
                            const double ddy = min_image_dmd(pj.y - pi4.y, S.L[1], S.iminv[1]);
                            const double ddz = min_image_dmd(pj.z - pi4.z, S.L[2], S.iminv[2]);
                            hit = ddx * ddx + ddy * ddy + ddz * ddz < S.rlist2;
                        }
                        uint64_t bal = __ballot(hit);
                        if (nex > 0 && bal != 0ull) {  // wave-uniform
                            bool x = false;
                            if (nex <= CB_EXR) {
                                const int j = jid[r];
#pragma unroll
                                for (int q = 0; q < CB_EXR; ++q) x |= ex[q] == j;
                            } else if (hit) {
                                x = is_excluded(ex_start, ex_list, i, jid[r]);
                            }
                            hit = hit && !x;
                            bal = __ballot(hit);
                        }
                        const int pos = k + __popcll(bal & below);
                        if (hit && pos < M) row[pos] = (int)e;
                        k += __popcll(bal);
                    }
                    kreg = lane == ii ? k : kreg;
                }
            }
            if (mine) {
                const int64_t ti = tl;
                const int lti = (int)(ti - S.atom_off);
                order[ti] = imy;
                slot_of[imy] = lti;
                int kk = kreg;
                if (kk > M) {
                    atomicOr(&st[s].flags, FLAG_NBR_OVERFLOW);
                    kk = M;
                }
                nn[ti] = kk;
                // pad the row to a multiple of 8 entries (as k_nbr_build_mixed)
                int* row = nbr + S.nbr_off + (int64_t)lti * M;
                for (int p = kk; p < ((kk + 7) & ~7); ++p) row[p] = (int)nbr_pad_entry(lti);
            }
        }
    }
}

// Mixed path, systems built per cell (mixed_builds_per_cell), candidates
// staged in LDS: one workgroup per cell of a rebuilding system.  The cell's
// candidate stream -- the neighbour cells in scan order (z, then y, then x
// offset, each ascending), each cell's atoms in cell order, as in
// k_nbr_build_mixed -- is copied into LDS once, SB_CAP candidates at a time,
// each shifted into the frame of the workgroup's cell (l_j + offset * w)
// with its list entry (slot | offset code) and atom id beside it.  Groups of
// TPA lanes then scan it for the cell's atoms straight from LDS: no table
// walk and no global load per candidate.  Decisions and order are those of
// k_nbr_build_mixed (FP32 test outside a 1e-5 nm band around the list
// radius, the reference's FP64 expression inside it), so the rows are the
// same, entry for entry.
constexpr int SB_THREADS = 256;
constexpr int SB_CAP = 1024;    // staged candidates per pass over the stream
constexpr int SB_ATOMS = 256;   // row lengths kept in LDS (atoms of one cell)
constexpr int SB_EXR = 6;       // excluded partners held in registers

template <int TPA>
__global__ void __launch_bounds__(SB_THREADS) k_nbr_build_staged(
    const SysDev* __restrict__ sys, SysState* __restrict__ st, const int* __restrict__ ccount,
    const int* __restrict__ cstart, const int* __restrict__ cell_atoms,
    const int* __restrict__ ex_start, const int* __restrict__ ex_list, int* __restrict__ order,
    int* __restrict__ slot_of, int* __restrict__ nn, int* __restrict__ nbr,
    const double4* __restrict__ pos4s, const float4* __restrict__ pos4f, RbList R)
{
    __shared__ float4 stage[SB_CAP];   // shifted position, list entry (w bits)
    __shared__ int sid[SB_CAP];        // atom id (exclusions)
    __shared__ int tb[NB_CELLS];       // cell-order start of each neighbour cell
    __shared__ int tp[NB_CELLS + 1];   // candidates before it in the stream
    __shared__ float4 tsh[NB_CELLS];   // offset * w, and the entry's code bits
    __shared__ int kk[SB_ATOMS];       // row length so far of each atom of the cell
    __shared__ int wtot[2];
    __shared__ RbStage rs;
    R = rb_stage(R, true, rs);
    constexpr int G = SB_THREADS / TPA;  // atoms per round
    const int tid = threadIdx.x, lane = tid % TPA, grp = tid / TPA;
    const int gshift = (tid & 63) / TPA * TPA;  // first lane of my group in the wave
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const int nrb = *R.n;
    const int64_t U = R.cpre[nrb];
    for (int64_t u = blockIdx.x; u < U; u += gridDim.x) {
        const int kq = rb_find(R.cpre, nrb, u);  // uniform over the workgroup
        const int s = R.sys[kq];
        const SysDev& S = sys[s];
        if (S.precision == DMDX_FP64 || !mixed_builds_per_cell(S)) continue;
        const int lc = (int)(u - R.cpre[kq]);
        if (lc == 0)  // the system's padding slots hold no atom
            for (int p = S.n + tid; p < S.n_pad; p += SB_THREADS) {
                nn[S.atom_off + p] = 0;
                order[S.atom_off + p] = -1;
            }
        const int nx = cells_in_use(S, 0), ny = cells_in_use(S, 1), nz = cells_in_use(S, 2);
        if (lc >= nx * ny * nz) continue;
        const int c = S.cell_off + lc;
        const int m = ccount[c];
        if (m == 0) continue;
        const int cx = lc % nx, cy = (lc / nx) % ny, cz = lc / (nx * ny);
        const float wx = (float)(S.L[0] / nx), wy = (float)(S.L[1] / ny),
                    wz = (float)(S.L[2] / nz);
        const int reach = S.cdiv, full = 2 * reach + 1;
        const int ncb = full * full * full;  // 27 or 125
        // neighbour-cell table (threads 0-127), running candidate counts
        // over the two waves
        int cnt = 0;
        if (tid < 128) {
            if (tid < ncb) {
                const int ox = tid % full - reach, oy = (tid / full) % full - reach,
                          oz = tid / (full * full) - reach;
                int xx = cx + ox, yy = cy + oy, zz = cz + oz;
                xx = xx < 0 ? xx + nx : (xx >= nx ? xx - nx : xx);
                yy = yy < 0 ? yy + ny : (yy >= ny ? yy - ny : yy);
                zz = zz < 0 ? zz + nz : (zz >= nz ? zz - nz : zz);
                const int cc = S.cell_off + (zz * ny + yy) * nx + xx;
                tb[tid] = cstart[cc];
                cnt = ccount[cc];
                const uint32_t code = nbr_code(ox, oy, oz) << NBR_SLOT_BITS;
                tsh[tid] = make_float4((float)ox * wx, (float)oy * wy, (float)oz * wz,
                                       __uint_as_float(code));
            }
            int incl = cnt;
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(incl, o, 64);
                if ((tid & 63) >= o) incl += v;
            }
            if ((tid & 63) == 63) wtot[tid >> 6] = incl;
            cnt = incl - cnt;  // exclusive, within the wave
        }
        __syncthreads();
        if (tid < ncb) tp[tid] = cnt + (tid >= 64 ? wtot[0] : 0);
        if (tid == 0) tp[ncb] = wtot[0] + (ncb > 64 ? wtot[1] : 0);
        for (int a = tid; a < m; a += SB_THREADS)
            if (a < SB_ATOMS) kk[a] = 0;
        __syncthreads();
        const int ncand = tp[ncb];
        const int64_t a0 = S.atom_off;
        const int64_t t0 = a0 + cstart[c];  // the cell's first cell-order slot
        const int M = S.max_nbr;
        const float lo2 = (float)((S.rlist - 1e-5) * (S.rlist - 1e-5));
        const float hi2 = (float)((S.rlist + 1e-5) * (S.rlist + 1e-5));
        // the system's atoms have list exclusions (CSR over atom slots)
        const bool any_ex = ex_start[a0 + S.n_pad] != ex_start[a0];
        for (int k0 = 0; k0 < ncand; k0 += SB_CAP) {
            const int k1 = min(ncand, k0 + SB_CAP);
            {
                // stage candidates [k0, k1), four loads in flight per thread
                constexpr int UNR = 4;
                int jc = 0;  // neighbour cell of candidate k (k only grows)
                for (int kb = k0 + tid; kb < k1; kb += UNR * SB_THREADS) {
                    float4 p[UNR];
                    int slot[UNR], id[UNR], cell[UNR];
#pragma unroll
                    for (int q = 0; q < UNR; ++q) {
                        const int k = kb + q * SB_THREADS;
                        if (k < k1) {
                            while (jc + 1 < ncb && tp[jc + 1] <= k) ++jc;
                            cell[q] = jc;
                            slot[q] = tb[jc] + (k - tp[jc]);
                            p[q] = pos4f[a0 + slot[q]];
                            id[q] = any_ex ? cell_atoms[a0 + slot[q]] : -1;
                        }
                    }
#pragma unroll
                    for (int q = 0; q < UNR; ++q) {
                        const int k = kb + q * SB_THREADS;
                        if (k < k1) {
                            const float4 sh = tsh[cell[q]];
                            stage[k - k0] = make_float4(
                                p[q].x + sh.x, p[q].y + sh.y, p[q].z + sh.z,
                                __uint_as_float((uint32_t)slot[q] | __float_as_uint(sh.w)));
                            sid[k - k0] = id[q];
                        }
                    }
                }
            }
            __syncthreads();
            const int nk = k1 - k0;
            for (int ai = grp; ai < m; ai += G) {
                const int64_t t = t0 + ai;
                const int lt = (int)(t - a0);
                const float4 li = pos4f[t];
                const int i = cell_atoms[t];
                const int exb = ex_start[i], exn = ex_start[i + 1] - exb;
                int exr[SB_EXR];
#pragma unroll
                for (int q = 0; q < SB_EXR; ++q) exr[q] = q < exn ? ex_list[exb + q] : -1;
                int* row = nbr + S.nbr_off + (int64_t)lt * M;
                int k = ai < SB_ATOMS ? kk[ai] : 0;
                if (ai >= SB_ATOMS && k0 > 0) {
                    // an atom past the LDS table (a cell of more than SB_ATOMS
                    // atoms): its row so far is its own count of entries
                    k = nn[t];
                }
                for (int kc0 = 0; kc0 < nk; kc0 += TPA) {
                    const int kc = kc0 + lane;
                    bool hit = false;
                    uint32_t e = 0u;
                    if (kc < nk) {
                        const float4 pj = stage[kc];
                        e = __float_as_uint(pj.w);
                        const int slot = (int)(e & NBR_SLOT_MASK);
                        const float dx = pj.x - li.x, dy = pj.y - li.y, dz = pj.z - li.z;
                        const float r2f = dx * dx + dy * dy + dz * dz;
                        if (slot != lt) {
                            hit = r2f < lo2;
                            if (!hit && r2f < hi2) {
                                // the reference's expression (src/core/cell.cpp:26-38,
                                // src/force/lennard_jones.cpp:29-36)
                                const double4 pi4 = pos4s[t];
                                const double4 pq = pos4s[a0 + slot];
                                const double ddx = min_image_dmd(pq.x - pi4.x, S.L[0], S.iminv[0]);
                                const double ddy = min_image_dmd(pq.y - pi4.y, S.L[1], S.iminv[1]);
                                const double ddz = min_image_dmd(pq.z - pi4.z, S.L[2], S.iminv[2]);
                                hit = ddx * ddx + ddy * ddy + ddz * ddz < S.rlist2;
                            }
                            if (hit && exn > 0) {
                                const int j = sid[kc];
                                if (exn <= SB_EXR) {
                                    bool x = false;
#pragma unroll
                                    for (int q = 0; q < SB_EXR; ++q) x |= exr[q] == j;
                                    hit = !x;
                                } else {
                                    hit = !is_excluded(ex_start, ex_list, i, j);
                                }
                            }
                        }
                    }
                    const uint64_t bal = __ballot(hit);
                    const uint64_t mine = (bal >> gshift) & (TPA == 64 ? ~0ull : ((1ull << TPA) - 1));
                    const int pos = k + __popcll(mine & below);
                    if (hit && pos < M) row[pos] = (int)e;
                    k += __popcll(mine);
                }
                if (lane == 0) {
                    if (ai < SB_ATOMS) kk[ai] = k;
                    else nn[t] = k;  // carried to the next pass over the stream
                }
            }
            __syncthreads();
        }
        // rows done: length, overflow flag, padding, atom maps
        for (int ai = tid; ai < m; ai += SB_THREADS) {
            const int64_t t = t0 + ai;
            const int lt = (int)(t - a0);
            const int i = cell_atoms[t];
            order[t] = i;
            slot_of[i] = lt;
            int k = ai < SB_ATOMS ? kk[ai] : nn[t];
            if (k > M) {
                atomicOr(&st[s].flags, FLAG_NBR_OVERFLOW);
                k = M;
            }
            nn[t] = k;
            int* row = nbr + S.nbr_off + (int64_t)lt * M;
            for (int p = k; p < ((k + 7) & ~7); ++p) row[p] = (int)nbr_pad_entry(lt);
        }
        __syncthreads();  // the next cell rewrites the shared tables
    }
}

// one 64-thread workgroup per atom slot: bitonic sort of its row in LDS
constexpr int SORT_CAP = 2048;

__device__ __forceinline__ void k_nbr_sort_unit(int64_t unit_, const SysDev* __restrict__ sys,
                                                 const SysState* __restrict__ st,
                                                 const int* __restrict__ slot_sys,
                                                 const int* __restrict__ nn,
                                                 int* __restrict__ nbr)
{
    int64_t t = unit_;
    int s = slot_sys[t];
    if (s < 0 || !st[s].rebuild) return;
    const SysDev& S = sys[s];
    if (S.precision != DMDX_FP64) return;
    int lt = (int)(t - S.atom_off);
    if (lt >= S.n) return;
    int n = nn[t];
    if (n <= 1) return;
    __shared__ int buf[SORT_CAP];
    int* row = nbr + S.nbr_off + (int64_t)lt * S.max_nbr;
    int p = 1;
    while (p < n) p <<= 1;
    if (p > SORT_CAP) return;  // capacity is checked on the host (max_nbr <= SORT_CAP)
    for (int k = threadIdx.x; k < p; k += 64) buf[k] = (k < n) ? row[k] : 0x7fffffff;
    __syncthreads();
    for (int size = 2; size <= p; size <<= 1) {
        for (int stride2 = size >> 1; stride2 > 0; stride2 >>= 1) {
            for (int k = threadIdx.x; k < p; k += 64) {
                int l = k ^ stride2;
                if (l > k) {
                    bool up = (k & size) == 0;
                    int a = buf[k], b = buf[l];
                    if ((a > b) == up) {
                        buf[k] = b;
                        buf[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int k = threadIdx.x; k < n; k += 64) row[k] = buf[k];
}

__global__ void __launch_bounds__(64) k_nbr_sort(const SysDev* __restrict__ sys,
                                                 const SysState* __restrict__ st,
                                                 const int* __restrict__ slot_sys,
                                                 const int* __restrict__ nn,
                                                 int* __restrict__ nbr, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_)
    {
        k_nbr_sort_unit(u_, sys, st, slot_sys, nn, nbr);
        __syncthreads();  // LDS is reused by the next unit
    }
}

// mixed path, every compute: cell-local FP32 positions and charges in cell
// order (x - origin of the slot's frame, fixed at the last list build)
__device__ __forceinline__ void k_pack_unit(int64_t unit_, const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                       const int* __restrict__ slot_sys, int64_t nslots,
                       const int* __restrict__ order, const double* __restrict__ x,
                       const double* __restrict__ y, const double* __restrict__ z,
                       const int* __restrict__ type, const double4* __restrict__ org,
                       float4* __restrict__ lpos, int* __restrict__ typ_s, int after_drift)
{
    int64_t t = unit_ * blockDim.x + threadIdx.x;
    if (t >= nslots) return;
    int s = slot_sys[t];
    if (s < 0 || st[s].reuse || !st[s].active) return;
    const SysDev& S = sys[s];
    if (t - S.atom_off >= S.n) return;
    // after the drift, k_kick_drift has written the frames already, except
    // where the cell order just changed or SHAKE moved atoms since
    if (after_drift && !st[s].rebuild && S.n_con == 0) return;
    int i = order[t];
    const double4 o = org[t];
    lpos[t] = make_float4((float)(x[i] - o.x), (float)(y[i] - o.y), (float)(z[i] - o.z),
                          (float)o.w);
    typ_s[t] = type[i];
}

__global__ void k_pack(const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                       const int* __restrict__ slot_sys, int64_t nslots,
                       const int* __restrict__ order, const double* __restrict__ x,
                       const double* __restrict__ y, const double* __restrict__ z,
                       const int* __restrict__ type, const double4* __restrict__ org,
                       float4* __restrict__ lpos, int* __restrict__ typ_s, int after_drift,
                       int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_) k_pack_unit(u_, sys, st, slot_sys, nslots, order, x, y, z, type, org, lpos, typ_s, after_drift);
}

}  // namespace dmdx

namespace dmdx {

// moving a system between batches: FP64 list entries and the cell-order map
// refer to atom slots, which shift by `delta`
__global__ void k_remap_rows(int* __restrict__ nbr, int64_t count, int* __restrict__ order,
                             int64_t n_order, int64_t delta)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride)
        nbr[k] = (int)((uint32_t)nbr[k] + (uint32_t)delta);
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n_order; k += stride)
        if (order[k] >= 0) order[k] = (int)(order[k] + delta);
}

}  // namespace dmdx

namespace dmdx {

hipError_t launch_remap_rows(hipStream_t s, int* nbr, int64_t count, int* order,
                             int64_t n_order, int64_t delta)
{
    const int64_t n = std::max(count, n_order);
    const unsigned g = (unsigned)std::min<int64_t>(1024, std::max<int64_t>(1, (n + 255) / 256));
    k_remap_rows<<<g, 256, 0, s>>>(nbr, count, order, n_order, delta);
    return hipGetLastError();
}

hipError_t launch_build_lists(const Dev& d, hipStream_t s)
{
    const int BS = 256;
    // Worklists of rebuilding systems first; the binning and list kernels
    // then visit only their cells and slots.  Grids are sized for a full
    // rebuild and stride over the compact ranges, whose lengths they read on
    // the device.
    const RbOut all{d.rb_sys, d.rb_pre, d.rb_cpre, d.rb_n};
    const RbOut multi{d.rbm_sys, d.rbm_pre, d.rbm_cpre, d.rbm_n};
    const RbOut wg{d.rbw_sys, nullptr, nullptr, d.rbw_n};
    k_rb_list<<<1, RB_THREADS, 0, s>>>(d.sys, d.st, d.nsys, all, multi, wg);
    if (d.n_bin_wg > 0)
        k_bin_system<<<std::min(d.n_bin_wg, 1024), BIN_THREADS, 0, s>>>(
            d.sys, d.st, d.rbw_sys, d.rbw_n, d.x, d.y, d.z, d.acell, d.ccount, d.cstart,
            d.cell_atoms);
    if (d.any_bin_multi) {
        const RbList R{d.rbm_sys, d.rbm_pre, d.rbm_cpre, d.rbm_n};
        if (d.ncells > 0)
            k_cells_reset<<<grid_for(blocks_for(d.ncells, BS)), BS, 0, s>>>(
                d.sys, d.st, d.cell_sys, d.ncells, d.ccount, d.cfill, R);
        k_cell_count<<<grid_for(blocks_for(d.nslots, BS)), BS, 0, s>>>(
            d.sys, d.st, d.slot_sys, d.nslots, d.x, d.y, d.z, d.acell, d.ccount, R);
        k_cell_scan<<<d.nsys, 256, 0, s>>>(d.sys, d.st, d.ccount, d.cstart);
        k_cell_fill<<<grid_for(blocks_for(d.nslots, BS)), BS, 0, s>>>(
            d.sys, d.st, d.slot_sys, d.nslots, d.acell, d.cstart, d.cfill, d.cell_atoms, R);
    }
    const RbList R{d.rb_sys, d.rb_pre, d.rb_cpre, d.rb_n};
    const PackOut po{d.pos4s, d.xref, d.pos4f, d.org, d.lpos, d.typ_s};
    if (d.ncells > 0)
        k_cell_sort<<<grid_for(blocks_for(d.ncells, 4)), 256, 0, s>>>(
            d.sys, d.ccount, d.cstart, d.cell_atoms, d.x, d.y, d.z, d.q, d.type, po, R);
    if (d.precision == DMDX_FP64)
        k_nbr_build<1><<<grid_for(blocks_for(d.nslots, BS)), BS, 0, s>>>(
            d.sys, d.st, d.slot_sys, d.nslots, d.x, d.y, d.z, d.acell, d.ccount, d.cstart,
            d.cell_atoms, d.ex_start, d.ex_list, d.order, d.nn, d.nbr, d.pos4s, nullptr, (int64_t)(blocks_for(d.nslots, BS)));
    else {
        if (d.build_cells && d.ncells > 0 && d.build_stage != 0) {
            const unsigned g = (unsigned)std::min<int64_t>(2048, std::max(1, d.ncells));
#define DMDX_STAGED_ARGS                                                                  \
    d.sys, d.st, d.ccount, d.cstart, d.cell_atoms, d.ex_start, d.ex_list, d.order, d.slot_of, \
        d.nn, d.nbr, d.pos4s, d.pos4f, R
            if (d.build_stage == 16)
                k_nbr_build_staged<16><<<g, SB_THREADS, 0, s>>>(DMDX_STAGED_ARGS);
            else
                k_nbr_build_staged<8><<<g, SB_THREADS, 0, s>>>(DMDX_STAGED_ARGS);
#undef DMDX_STAGED_ARGS
        } else if (d.build_cells && d.ncells > 0)
            k_nbr_build_cells<<<grid_for(blocks_for((int64_t)d.ncells * d.build_tasks, CB_WAVES)),
                                64 * CB_WAVES, 0, s>>>(
                d.sys, d.st, d.ccount, d.cstart, d.cell_atoms, d.ex_start, d.ex_list, d.order,
                d.slot_of, d.nn, d.nbr, d.pos4s, d.pos4f, R, d.build_tasks);
        if (d.any_wrap_mixed || !d.build_cells) {
            const int tpa = d.build_tpa;
            const int64_t units = blocks_for(d.nslots * tpa, BS);
#define DMDX_BUILD_ARGS                                                                       \
    d.sys, d.st, d.slot_sys, d.nslots, d.acell, d.ccount, d.cstart, d.cell_atoms, d.ex_start, \
        d.ex_list, d.order, d.slot_of, d.nn, d.nbr, d.pos4s, d.pos4f, R, d.build_cells
            if (tpa == 8) k_nbr_build_mixed<8><<<grid_for(units), BS, 0, s>>>(DMDX_BUILD_ARGS);
            else if (tpa == 32) k_nbr_build_mixed<32><<<grid_for(units), BS, 0, s>>>(DMDX_BUILD_ARGS);
            else k_nbr_build_mixed<16><<<grid_for(units), BS, 0, s>>>(DMDX_BUILD_ARGS);
#undef DMDX_BUILD_ARGS
        }
    }
    if (d.precision == DMDX_FP64)
        k_nbr_sort<<<grid_for((unsigned)d.nslots), 64, 0, s>>>(d.sys, d.st, d.slot_sys, d.nn, d.nbr, (int64_t)((unsigned)d.nslots));
    return hipGetLastError();
}

hipError_t launch_displacement(const Dev& d, hipStream_t s)
{
    if (!d.any_auto) return hipSuccess;
    k_displacement<<<grid_for(blocks_for(d.nslots, 256)), 256, 0, s>>>(d.sys, d.st, d.slot_sys, d.nslots,
                                                             d.x, d.y, d.z, d.xref, d.p_disp, (int64_t)(blocks_for(d.nslots, 256)));
    return hipGetLastError();
}

hipError_t launch_pack_phase(const Dev& d, hipStream_t s, int after_drift)
{
    if (d.precision == DMDX_FP64) return hipSuccess;
    k_pack<<<grid_for(blocks_for(d.nslots, 256)), 256, 0, s>>>(
        d.sys, d.st, d.slot_sys, d.nslots, d.order, d.x, d.y, d.z, d.type, d.org, d.lpos, d.typ_s,
        after_drift, (int64_t)(blocks_for(d.nslots, 256)));
    return hipGetLastError();
}

hipError_t launch_pack(const Dev& d, hipStream_t s) { return launch_pack_phase(d, s, 0); }

}  // namespace dmdx
