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
//   k_cells_reset  zero per-cell counters
//   k_cell_count   cell of each atom from its wrapped position, count
//   k_cell_scan    per-system exclusive scan of the counts
//   k_cell_fill    scatter atoms into cell order
//   k_cell_sort    sort each cell by slot (deterministic order)
//   k_nbr_build    per atom (cell order): scan the neighbour cells
//   k_nbr_sort     per atom: ascending neighbour order (FP64 path: makes the
//                  per-atom force sum follow the reference's summation order)
#include "dmdx_internal.h"

namespace dmdx {

// cells in use along one edge: as many as fit with edge >= rlist, at most the
// allocated count.  Fewer than 3 is handled by the offset sets below.
__device__ __forceinline__ int cells_in_use(const SysDev& S, int d)
{
    int nu = (int)floor(S.L[d] / (S.rlist * (1.0 + 1e-9)));
    if (nu > S.nc[d]) nu = S.nc[d];
    if (nu < 1) nu = 1;
    return nu;
}

__device__ __forceinline__ int cell_coord(double x, double L, int nu)
{
    double f = x / L;
    f -= floor(f);
    int c = (int)(f * nu);
    return c < 0 ? 0 : (c >= nu ? nu - 1 : c);
}

__global__ void k_cells_reset(const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                              const int* __restrict__ cell_sys, int ncells_total,
                              int* __restrict__ ccount, int* __restrict__ cfill)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells_total) return;
    int s = cell_sys[c];
    if (!st[s].rebuild) return;
    ccount[c] = 0;
    cfill[c] = 0;
}

__global__ void k_cell_count(const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                             const int* __restrict__ slot_sys, int64_t nslots,
                             const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, int* __restrict__ acell,
                             int* __restrict__ ccount)
{
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nslots) return;
    int s = slot_sys[g];
    if (s < 0 || !st[s].rebuild) return;
    const SysDev& S = sys[s];
    if (g - S.atom_off >= S.n) return;
    int nx = cells_in_use(S, 0), ny = cells_in_use(S, 1), nz = cells_in_use(S, 2);
    int cx = cell_coord(x[g], S.L[0], nx);
    int cy = cell_coord(y[g], S.L[1], ny);
    int cz = cell_coord(z[g], S.L[2], nz);
    int c = (cz * ny + cy) * nx + cx;
    acell[g] = c;
    atomicAdd(&ccount[S.cell_off + c], 1);
}

// one workgroup (256 threads) per system
__global__ void k_cell_scan(const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                            const int* __restrict__ ccount, int* __restrict__ cstart)
{
    int s = blockIdx.x;
    if (!st[s].rebuild) return;
    const SysDev& S = sys[s];
    int ncell = cells_in_use(S, 0) * cells_in_use(S, 1) * cells_in_use(S, 2);
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

__global__ void k_cell_fill(const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                            const int* __restrict__ slot_sys, int64_t nslots,
                            const int* __restrict__ acell, const int* __restrict__ cstart,
                            int* __restrict__ cfill, int* __restrict__ cell_atoms)
{
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nslots) return;
    int s = slot_sys[g];
    if (s < 0 || !st[s].rebuild) return;
    const SysDev& S = sys[s];
    if (g - S.atom_off >= S.n) return;
    int c = acell[g];
    int pos = atomicAdd(&cfill[S.cell_off + c], 1);
    cell_atoms[S.atom_off + cstart[S.cell_off + c] + pos] = (int)g;
}

__global__ void k_cell_sort(const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                            const int* __restrict__ cell_sys, int ncells_total,
                            const int* __restrict__ ccount, const int* __restrict__ cstart,
                            int* __restrict__ cell_atoms)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells_total) return;
    int s = cell_sys[c];
    if (!st[s].rebuild) return;
    const SysDev& S = sys[s];
    int lc = c - S.cell_off;
    int ncell = cells_in_use(S, 0) * cells_in_use(S, 1) * cells_in_use(S, 2);
    if (lc >= ncell) return;
    int* a = cell_atoms + S.atom_off + cstart[c];
    int n = ccount[c];
    for (int i = 1; i < n; ++i) {
        int v = a[i];
        int j = i - 1;
        while (j >= 0 && a[j] > v) {
            a[j + 1] = a[j];
            --j;
        }
        a[j + 1] = v;
    }
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

// thread t (slot order) handles the atom cell_atoms[t]; its list lives in
// column (t - atom_off) of the system's M x n_pad block.
__global__ void k_nbr_build(const SysDev* __restrict__ sys, SysState* __restrict__ st,
                            const int* __restrict__ slot_sys, int64_t nslots,
                            const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ z, const int* __restrict__ acell,
                            const int* __restrict__ ccount, const int* __restrict__ cstart,
                            const int* __restrict__ cell_atoms, const int* __restrict__ ex_start,
                            const int* __restrict__ ex_list, int* __restrict__ order,
                            int* __restrict__ nn, int* __restrict__ nbr)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nslots) return;
    int s = slot_sys[t];
    if (s < 0 || !st[s].rebuild) return;
    const SysDev& S = sys[s];
    int lt = (int)(t - S.atom_off);
    if (lt >= S.n) {
        nn[t] = 0;
        order[t] = -1;
        return;
    }
    int i = cell_atoms[t];
    order[t] = i;
    int nx = cells_in_use(S, 0), ny = cells_in_use(S, 1), nz = cells_in_use(S, 2);
    int ci = acell[i];
    int cx = ci % nx, cy = (ci / nx) % ny, cz = ci / (nx * ny);
    int ox[3], oy[3], oz[3];
    int nox = edge_offsets(nx, ox), noy = edge_offsets(ny, oy), noz = edge_offsets(nz, oz);
    const double xi = x[i], yi = y[i], zi = z[i];
    const double Lx = S.L[0], Ly = S.L[1], Lz = S.L[2];
    const double ix = S.invL[0], iy = S.invL[1], iz = S.invL[2];
    const double rl2 = S.rlist2;
    const bool check_ex = ex_start[i + 1] > ex_start[i];
    int* col = nbr + S.nbr_off + lt;
    const int64_t stride = S.n_pad;
    int k = 0;
    const int M = S.max_nbr;
    for (int a = 0; a < noz; ++a) {
        int zz = cz + oz[a];
        zz = zz < 0 ? zz + nz : (zz >= nz ? zz - nz : zz);
        for (int b = 0; b < noy; ++b) {
            int yy = cy + oy[b];
            yy = yy < 0 ? yy + ny : (yy >= ny ? yy - ny : yy);
            for (int c = 0; c < nox; ++c) {
                int xx = cx + ox[c];
                xx = xx < 0 ? xx + nx : (xx >= nx ? xx - nx : xx);
                int cell = S.cell_off + (zz * ny + yy) * nx + xx;
                const int* ca = cell_atoms + S.atom_off + cstart[cell];
                int cnt = ccount[cell];
                for (int m = 0; m < cnt; ++m) {
                    int j = ca[m];
                    if (j == i) continue;
                    double dx = x[j] - xi;
                    double dy = y[j] - yi;
                    double dz = z[j] - zi;
                    dx = min_image_dmd(dx, Lx, ix);
                    dy = min_image_dmd(dy, Ly, iy);
                    dz = min_image_dmd(dz, Lz, iz);
                    double r2 = dx * dx + dy * dy + dz * dz;
                    if (!(r2 < rl2)) continue;
                    if (check_ex && is_excluded(ex_start, ex_list, i, j)) continue;
                    if (k < M) col[(int64_t)k * stride] = j;
                    ++k;
                }
            }
        }
    }
    if (k > M) {
        atomicOr(&st[s].flags, FLAG_NBR_OVERFLOW);
        k = M;
    }
    nn[t] = k;
}

// one 64-thread workgroup per atom slot: bitonic sort of its column in LDS
constexpr int SORT_CAP = 2048;

__global__ void __launch_bounds__(64) k_nbr_sort(const SysDev* __restrict__ sys,
                                                 const SysState* __restrict__ st,
                                                 const int* __restrict__ slot_sys,
                                                 const int* __restrict__ nn,
                                                 int* __restrict__ nbr)
{
    int64_t t = blockIdx.x;
    int s = slot_sys[t];
    if (s < 0 || !st[s].rebuild) return;
    const SysDev& S = sys[s];
    if (S.precision != DMDX_FP64) return;
    int lt = (int)(t - S.atom_off);
    if (lt >= S.n) return;
    int n = nn[t];
    if (n <= 1) return;
    __shared__ int buf[SORT_CAP];
    int* col = nbr + S.nbr_off + lt;
    const int64_t stride = S.n_pad;
    int p = 1;
    while (p < n) p <<= 1;
    if (p > SORT_CAP) return;  // capacity is checked on the host (max_nbr <= SORT_CAP)
    for (int k = threadIdx.x; k < p; k += 64) buf[k] = (k < n) ? col[(int64_t)k * stride] : 0x7fffffff;
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
    for (int k = threadIdx.x; k < n; k += 64) col[(int64_t)k * stride] = buf[k];
}

}  // namespace dmdx

namespace dmdx {

hipError_t launch_build_lists(const Dev& d, hipStream_t s)
{
    const int BS = 256;
    if (d.ncells > 0)
        k_cells_reset<<<blocks_for(d.ncells, BS), BS, 0, s>>>(d.sys, d.st, d.cell_sys, d.ncells,
                                                              d.ccount, d.cfill);
    k_cell_count<<<blocks_for(d.nslots, BS), BS, 0, s>>>(d.sys, d.st, d.slot_sys, d.nslots, d.x,
                                                         d.y, d.z, d.acell, d.ccount);
    k_cell_scan<<<d.nsys, 256, 0, s>>>(d.sys, d.st, d.ccount, d.cstart);
    k_cell_fill<<<blocks_for(d.nslots, BS), BS, 0, s>>>(d.sys, d.st, d.slot_sys, d.nslots,
                                                        d.acell, d.cstart, d.cfill,
                                                        d.cell_atoms);
    if (d.ncells > 0)
        k_cell_sort<<<blocks_for(d.ncells, BS), BS, 0, s>>>(d.sys, d.st, d.cell_sys, d.ncells,
                                                            d.ccount, d.cstart, d.cell_atoms);
    k_nbr_build<<<blocks_for(d.nslots, BS), BS, 0, s>>>(
        d.sys, d.st, d.slot_sys, d.nslots, d.x, d.y, d.z, d.acell, d.ccount, d.cstart,
        d.cell_atoms, d.ex_start, d.ex_list, d.order, d.nn, d.nbr);
    if (d.precision == DMDX_FP64)
        k_nbr_sort<<<(unsigned)d.nslots, 64, 0, s>>>(d.sys, d.st, d.slot_sys, d.nn, d.nbr);
    return hipGetLastError();
}

}  // namespace dmdx
