// pme.hip -- smooth particle-mesh Ewald, reciprocal part, on gfx950.
//
// Replaces CoulombPME::compute (src/force/coulomb_pme.cpp:287-318) and its
// FFTW calls (:154, :198).  Per compute:
//   k_pme_zero      clear the fixed-point charge grid (only when some system
//                   spreads with atomics: owner-computes blocks store every
//                   point of their grid)
//   k_pme_spread_tiles / k_pme_spread_gather  order-4 B-spline charge
//                   spreading (:79-136) in 64-bit fixed point: integer
//                   atomics staged in LDS per tile of cell-ordered atoms, or
//                   owner-computes grid blocks that gather the atoms of the
//                   nearby list cells (no global atomics); both give the
//                   same grid
//   k_fft_pass x,y  forward 1D FFT passes along x then y; the x pass reads
//                   the fixed-point grid and converts it as it loads its lines
//                   (k_pme_convert is the stand-alone conversion); on the
//                   mixed path, systems whose xy plane fits in LDS run both
//                   as k_fft_plane, one workgroup per plane
//   k_fft_pass z+   forward z pass, influence function G(m) with the energy
//                   0.5 G |Q(m)|^2 (:143-201), inverse z pass -- the whole
//                   convolution of each z line happens in LDS
//   k_fft_pass y,x  inverse passes (k_fft_plane for the plane systems)
//   (gather)        in k_gather (integrate.hip), per atom
// The grid is complex (real input, imaginary part 0) with x fastest.  The FFT
// is a mixed-radix Stockham autosort transform (Govindaraju et al. 2008,
// "High performance discrete Fourier transforms on graphics processors"):
// each workgroup loads a tile of B lines into LDS, runs all radix stages
// there (ping-pong buffers) and writes back; y and z tiles are B adjacent x
// columns so global loads/stores are coalesced.  Templated on the real type:
// double for the FP64 path, float for the mixed path.
#include "dmdx_internal.h"

namespace dmdx {

template <class R>
struct Cx {
    R re, im;
};

template <class R>
__device__ __forceinline__ Cx<R> cmul(Cx<R> a, Cx<R> b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// centred cubic B-spline and derivative, src/force/coulomb_pme.cpp:50-69
template <class R>
__device__ __forceinline__ R theta4(R u)
{
    R a = fabs(u);
    if (a >= R(2)) return R(0);
    if (a >= R(1)) {
        R t = R(2) - a;
        return t * t * t / R(6);
    }
    return (R(4) - R(6) * a * a + R(3) * a * a * a) / R(6);
}

template <class R>
__device__ __forceinline__ R dtheta4(R u)
{
    R a = fabs(u);
    if (a >= R(2)) return R(0);
    R s = (u >= R(0)) ? R(1) : R(-1);
    if (a >= R(1)) {
        R t = R(2) - a;
        return -s * t * t / R(2);
    }
    return s * (R(9) * a * a - R(12) * a) / R(6);
}

__device__ __forceinline__ void k_pme_zero_unit(int64_t unit_, const SysState* __restrict__ st,
                                                const int* __restrict__ gsys, int64_t ntot,
                                                unsigned long long* __restrict__ qgrid)
{
    int64_t e = unit_ * blockDim.x + threadIdx.x;
    if (e >= ntot) return;
    int s = gsys[e];
    if (s < 0 || st[s].reuse || !st[s].active) return;
    qgrid[e] = 0ull;
}

__global__ void k_pme_zero(const SysState* __restrict__ st, const int* __restrict__ gsys,
                           int64_t ntot, unsigned long long* __restrict__ qgrid, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_) k_pme_zero_unit(u_, st, gsys, ntot, qgrid);
}

// fixed-point grid -> complex grid of the FFT (imaginary part 0)
template <class R>
__device__ __forceinline__ void k_pme_convert_unit(int64_t unit_, const SysState* __restrict__ st,
                                                   const int* __restrict__ gsys, int64_t ntot,
                                                   const unsigned long long* __restrict__ qgrid,
                                                   Cx<R>* __restrict__ grid)
{
    int64_t e = unit_ * blockDim.x + threadIdx.x;
    if (e >= ntot) return;
    int s = gsys[e];
    if (s < 0 || st[s].reuse || !st[s].active) return;
    const double v = (double)(long long)qgrid[e] * (1.0 / QGRID_SCALE);
    grid[e] = {R(v), R(0)};
}

template <class R>
__global__ void k_pme_convert(const SysState* __restrict__ st, const int* __restrict__ gsys,
                              int64_t ntot, const unsigned long long* __restrict__ qgrid,
                              Cx<R>* __restrict__ grid, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_) k_pme_convert_unit<R>(u_, st, gsys, ntot, qgrid, grid);
}

// two's-complement fixed point: unsigned wrap-around addition is signed addition
__device__ __forceinline__ void atomic_add_fixed(unsigned long long* p, double v)
{
    atomicAdd(p, (unsigned long long)__double2ll_rn(v * QGRID_SCALE));
}

// per atom: fractional coordinate, first grid index and offset, in FP64 (the
// reference's expressions), then the weights in R
template <class R>
__device__ __forceinline__ void k_pme_spread_unit(int64_t unit_, const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                             const int* __restrict__ slot_sys, int64_t nslots,
                             const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, const double* __restrict__ q,
                             unsigned long long* __restrict__ qgrid)
{
    int64_t g = unit_ * blockDim.x + threadIdx.x;
    if (g >= nslots) return;
    int s = slot_sys[g];
    if (s < 0 || st[s].reuse || !st[s].active) return;
    const SysDev& S = sys[s];
    if (!S.pme || g - S.atom_off >= S.n) return;
    double qi = q[g];
    if (qi == 0.0) return;
    const int nx = S.K[0], ny = S.K[1], nz = S.K[2];
    double frac_x = x[g] / S.L[0], frac_y = y[g] / S.L[1], frac_z = z[g] / S.L[2];
    frac_x -= floor(frac_x);
    frac_y -= floor(frac_y);
    frac_z -= floor(frac_z);
    // non-finite coordinates (a failed system) must still index the grid
    frac_x = (frac_x >= 0.0 && frac_x <= 1.0) ? frac_x : 0.0;
    frac_y = (frac_y >= 0.0 && frac_y <= 1.0) ? frac_y : 0.0;
    frac_z = (frac_z >= 0.0 && frac_z <= 1.0) ? frac_z : 0.0;
    double ux = frac_x * nx, uy = frac_y * ny, uz = frac_z * nz;
    int m0x = (int)floor(ux) - 1, m0y = (int)floor(uy) - 1, m0z = (int)floor(uz) - 1;
    R wx[4], wy[4], wz[4];
    for (int d = 0; d < 4; ++d) {
        wx[d] = theta4<R>((R)(ux - (m0x + d)));
        wy[d] = theta4<R>((R)(uy - (m0y + d)));
        wz[d] = theta4<R>((R)(uz - (m0z + d)));
    }
    unsigned long long* G = qgrid + S.grid_off;
    const R qr = (R)qi;
    for (int iz = 0; iz < 4; ++iz) {
        int gz = (m0z + iz) % nz;
        if (gz < 0) gz += nz;
        R w_z = wz[iz];
        for (int iy = 0; iy < 4; ++iy) {
            int gy = (m0y + iy) % ny;
            if (gy < 0) gy += ny;
            R w_yz = wy[iy] * w_z;
            int64_t row = ((int64_t)gz * ny + gy) * nx;
            for (int ix = 0; ix < 4; ++ix) {
                int gx = (m0x + ix) % nx;
                if (gx < 0) gx += nx;
                atomic_add_fixed(&G[row + gx], (double)(qr * wx[ix] * w_yz));
            }
        }
    }
}

template <class R>
__global__ void k_pme_spread(const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                             const int* __restrict__ slot_sys, int64_t nslots,
                             const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, const double* __restrict__ q,
                             unsigned long long* __restrict__ qgrid, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_) k_pme_spread_unit<R>(u_, sys, st, slot_sys, nslots, x, y, z, q, qgrid);
}

// LDS-staged spreading.  A tile is up to 256 consecutive cell-order slots of
// one system (spatially compact: atoms sorted by cell).  The workgroup finds
// the bounding box of its atoms' 4x4x4 stencils; when the box fits in LDS
// (SPREAD_LDS_POINTS int64 values) the 64 contributions per atom are added
// there with LDS atomics and the box is flushed to the grid with one global
// atomic per touched point, else every contribution goes to the grid
// directly.  The sums are integers, so both routes give the same grid as
// k_pme_spread.
constexpr int SPREAD_LDS_POINTS = 4096;

template <class R>
__device__ __forceinline__ void spread_tile_unit(int64_t unit_, const SpreadTile* __restrict__ tiles,
                                                 const SysDev* __restrict__ sys,
                                                 const SysState* __restrict__ st,
                                                 const int* __restrict__ order,
                                                 const double* __restrict__ x,
                                                 const double* __restrict__ y,
                                                 const double* __restrict__ z,
                                                 const double* __restrict__ q,
                                                 unsigned long long* __restrict__ qgrid)
{
    __shared__ unsigned long long acc[SPREAD_LDS_POINTS];
    __shared__ int box[6];
    const SpreadTile T = tiles[unit_];
    const int s = T.sys;
    if (st[s].reuse || !st[s].active) return;  // uniform per workgroup
    const SysDev& S = sys[s];
    const int nx = S.K[0], ny = S.K[1], nz = S.K[2];
    const int64_t t = T.t0 + threadIdx.x;
    bool mine = threadIdx.x < T.cnt;
    int64_t g = 0;
    double qi = 0.0;
    if (mine) {
        g = S.has_list ? (int64_t)order[t] : t;  // cell order when a list exists
        qi = q[g];
        mine = qi != 0.0;
    }
    int m0x = 0, m0y = 0, m0z = 0;
    R wx[4], wy[4], wz[4];
    if (mine) {
        double frac_x = x[g] / S.L[0], frac_y = y[g] / S.L[1], frac_z = z[g] / S.L[2];
        frac_x -= floor(frac_x);
        frac_y -= floor(frac_y);
        frac_z -= floor(frac_z);
        frac_x = (frac_x >= 0.0 && frac_x <= 1.0) ? frac_x : 0.0;
        frac_y = (frac_y >= 0.0 && frac_y <= 1.0) ? frac_y : 0.0;
        frac_z = (frac_z >= 0.0 && frac_z <= 1.0) ? frac_z : 0.0;
        const double ux = frac_x * nx, uy = frac_y * ny, uz = frac_z * nz;
        m0x = (int)floor(ux) - 1;
        m0y = (int)floor(uy) - 1;
        m0z = (int)floor(uz) - 1;
        for (int d = 0; d < 4; ++d) {
            wx[d] = theta4<R>((R)(ux - (m0x + d)));
            wy[d] = theta4<R>((R)(uy - (m0y + d)));
            wz[d] = theta4<R>((R)(uz - (m0z + d)));
        }
    }
    if (threadIdx.x < 6) box[threadIdx.x] = (threadIdx.x & 1) ? -0x40000000 : 0x40000000;
    __syncthreads();
    if (mine) {
        atomicMin(&box[0], m0x);
        atomicMax(&box[1], m0x);
        atomicMin(&box[2], m0y);
        atomicMax(&box[3], m0y);
        atomicMin(&box[4], m0z);
        atomicMax(&box[5], m0z);
    }
    __syncthreads();
    const int bx0 = box[0], by0 = box[2], bz0 = box[4];
    const int bnx = box[1] - bx0 + 4, bny = box[3] - by0 + 4, bnz = box[5] - bz0 + 4;
    const bool any = box[1] >= box[0];
    const bool staged = any && (int64_t)bnx * bny * bnz <= SPREAD_LDS_POINTS;
    unsigned long long* G = qgrid + S.grid_off;
    const R qr = (R)qi;
    if (staged) {
        const int npts = bnx * bny * bnz;
        for (int k = threadIdx.x; k < npts; k += blockDim.x) acc[k] = 0ull;
        __syncthreads();
        if (mine) {
            for (int iz = 0; iz < 4; ++iz) {
                const R w_z = wz[iz];
                for (int iy = 0; iy < 4; ++iy) {
                    const R w_yz = wy[iy] * w_z;
                    const int rowl = ((m0z + iz - bz0) * bny + (m0y + iy - by0)) * bnx + (m0x - bx0);
                    for (int ix = 0; ix < 4; ++ix)
                        atomicAdd(&acc[rowl + ix], (unsigned long long)__double2ll_rn(
                                                       (double)(qr * wx[ix] * w_yz) * QGRID_SCALE));
                }
            }
        }
        __syncthreads();
        for (int k = threadIdx.x; k < npts; k += blockDim.x) {
            const unsigned long long v = acc[k];
            if (v == 0ull) continue;
            const int lx = k % bnx, ly = (k / bnx) % bny, lz = k / (bnx * bny);
            int gx = (bx0 + lx) % nx, gy = (by0 + ly) % ny, gz = (bz0 + lz) % nz;
            gx += gx < 0 ? nx : 0;
            gy += gy < 0 ? ny : 0;
            gz += gz < 0 ? nz : 0;
            atomicAdd(&G[((int64_t)gz * ny + gy) * nx + gx], v);
        }
    } else if (mine) {
        for (int iz = 0; iz < 4; ++iz) {
            int gz = (m0z + iz) % nz;
            if (gz < 0) gz += nz;
            const R w_z = wz[iz];
            for (int iy = 0; iy < 4; ++iy) {
                int gy = (m0y + iy) % ny;
                if (gy < 0) gy += ny;
                const R w_yz = wy[iy] * w_z;
                const int64_t row = ((int64_t)gz * ny + gy) * nx;
                for (int ix = 0; ix < 4; ++ix) {
                    int gx = (m0x + ix) % nx;
                    if (gx < 0) gx += nx;
                    atomic_add_fixed(&G[row + gx], (double)(qr * wx[ix] * w_yz));
                }
            }
        }
    }
}

// Owner-computes spreading (no global atomics).  A workgroup owns one
// GT x GT x GT block of a system's grid: it scans the atoms of the list cells
// that can reach the block (cells of the last list build, widened by the
// displacement the Verlet criterion allows, skin/2), adds each atom's
// contributions that fall inside the block into LDS (the same fixed-point
// values as k_pme_spread_tiles, so the grid is identical) and stores the
// block.  Used for systems whose atoms stay within skin/2 of their build
// cells: mixed path, Verlet criterion, fixed box.
constexpr int GT = 8;
constexpr int GCELLS = 255;  // cells a block may scan (host checks)

template <class R>
__device__ __forceinline__ void spread_gather_unit(int64_t unit_, const GridTile* __restrict__ tiles,
                                                   const SysDev* __restrict__ sys,
                                                   const SysState* __restrict__ st,
                                                   const int* __restrict__ ccount,
                                                   const int* __restrict__ cstart,
                                                   const int* __restrict__ order,
                                                   const double* __restrict__ x,
                                                   const double* __restrict__ y,
                                                   const double* __restrict__ z,
                                                   const double* __restrict__ q,
                                                   unsigned long long* __restrict__ qgrid)
{
    __shared__ unsigned long long acc[GT * GT * GT];
    const GridTile Tl = tiles[unit_];
    const int s = Tl.sys;
    const SysState& T = st[s];
    if (T.reuse || !T.active) return;  // uniform per workgroup
    const SysDev& S = sys[s];
    const int K0 = S.K[0], K1 = S.K[1], K2 = S.K[2];
    const int g0[3] = {Tl.g0[0], Tl.g0[1], Tl.g0[2]};
    for (int k = threadIdx.x; k < GT * GT * GT; k += blockDim.x) acc[k] = 0ull;
    // cell range per axis: atoms whose stencil reaches [g0, g0 + GT) have
    // u in [g0 - 2, g0 + GT + 1), i.e. x in that range times L/K, and are at
    // most skin/2 from where the cells were built
    const double drift = sqrt(S.half_skin2);
    int clo[3], ncl[3], nu[3];
    for (int d = 0; d < 3; ++d) {
        const int Kd = S.K[d];
        nu[d] = T.nub[d];
        const double h = S.L[d] / Kd, w = S.L[d] / nu[d];
        const double xa = (g0[d] - 2) * h - drift, xb = (g0[d] + GT + 1) * h + drift;
        int ca = (int)floor(xa / w), cb = (int)floor(xb / w);
        if (cb - ca + 1 >= nu[d]) {
            ca = 0;
            cb = nu[d] - 1;
        }
        clo[d] = ca;
        ncl[d] = cb - ca + 1;
    }
    // the cells' atom ranges, then all candidates spread over the workgroup
    // (cells hold ~10 atoms: a wave per cell would leave most lanes idle)
    __shared__ int cbase[GCELLS], cpre[GCELLS + 1];
    const int ncells = ncl[0] * ncl[1] * ncl[2];  // <= GCELLS (checked on the host)
    int mycnt = 0;
    if ((int)threadIdx.x < ncells) {
        const int c = threadIdx.x;
        int cx = clo[0] + c % ncl[0], cy = clo[1] + (c / ncl[0]) % ncl[1],
            cz = clo[2] + c / (ncl[0] * ncl[1]);
        cx += cx < 0 ? nu[0] : (cx >= nu[0] ? -nu[0] : 0);
        cy += cy < 0 ? nu[1] : (cy >= nu[1] ? -nu[1] : 0);
        cz += cz < 0 ? nu[2] : (cz >= nu[2] ? -nu[2] : 0);
        const int cell = S.cell_off + (cz * nu[1] + cy) * nu[0] + cx;
        cbase[c] = cstart[cell];
        mycnt = ccount[cell];
    }
    // exclusive scan of the counts over the workgroup (fixed tree)
    cpre[threadIdx.x < GCELLS ? threadIdx.x : 0] = 0;
    __syncthreads();
    if ((int)threadIdx.x < ncells) cpre[threadIdx.x + 1] = mycnt;
    __syncthreads();
    for (int o = 1; o < GCELLS; o <<= 1) {
        int v = 0;
        if ((int)threadIdx.x <= ncells && (int)threadIdx.x >= o) v = cpre[threadIdx.x - o];
        __syncthreads();
        if ((int)threadIdx.x <= ncells) cpre[threadIdx.x] += v;
        __syncthreads();
    }
    const int total = cpre[ncells];
    for (int k = threadIdx.x; k < total; k += blockDim.x) {
        // cell of candidate k: last c with cpre[c] <= k
        int lo = 0, hi = ncells - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (cpre[mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        const int64_t g = order[S.atom_off + cbase[lo] + (k - cpre[lo])];
        const double qi = q[g];
        if (qi == 0.0) continue;
        // x first: most candidates miss the block along some axis
        double frac_x = x[g] / S.L[0];
        frac_x -= floor(frac_x);
        frac_x = (frac_x >= 0.0 && frac_x <= 1.0) ? frac_x : 0.0;
        const double ux = frac_x * K0;
        const int m0x = (int)floor(ux) - 1;
        int lx[4];
        bool any = false;
        for (int d = 0; d < 4; ++d) {
            int p = m0x + d;
            p += p < 0 ? K0 : (p >= K0 ? -K0 : 0);
            lx[d] = (p >= g0[0] && p < g0[0] + GT) ? p - g0[0] : -1;
            any |= lx[d] >= 0;
        }
        if (!any) continue;
        double frac_y = y[g] / S.L[1];
        frac_y -= floor(frac_y);
        frac_y = (frac_y >= 0.0 && frac_y <= 1.0) ? frac_y : 0.0;
        const double uy = frac_y * K1;
        const int m0y = (int)floor(uy) - 1;
        int ly[4];
        any = false;
        for (int d = 0; d < 4; ++d) {
            int p = m0y + d;
            p += p < 0 ? K1 : (p >= K1 ? -K1 : 0);
            ly[d] = (p >= g0[1] && p < g0[1] + GT) ? p - g0[1] : -1;
            any |= ly[d] >= 0;
        }
        if (!any) continue;
        double frac_z = z[g] / S.L[2];
        frac_z -= floor(frac_z);
        frac_z = (frac_z >= 0.0 && frac_z <= 1.0) ? frac_z : 0.0;
        const double uz = frac_z * K2;
        const int m0z = (int)floor(uz) - 1;
        int lz[4];
        any = false;
        for (int d = 0; d < 4; ++d) {
            int p = m0z + d;
            p += p < 0 ? K2 : (p >= K2 ? -K2 : 0);
            lz[d] = (p >= g0[2] && p < g0[2] + GT) ? p - g0[2] : -1;
            any |= lz[d] >= 0;
        }
        if (!any) continue;
        R wx[4], wy[4], wz[4];
        for (int d = 0; d < 4; ++d) {
            wx[d] = theta4<R>((R)(ux - (m0x + d)));
            wy[d] = theta4<R>((R)(uy - (m0y + d)));
            wz[d] = theta4<R>((R)(uz - (m0z + d)));
        }
        const R qr = (R)qi;
        for (int iz = 0; iz < 4; ++iz) {
            if (lz[iz] < 0) continue;
            const R w_z = wz[iz];
            for (int iy = 0; iy < 4; ++iy) {
                if (ly[iy] < 0) continue;
                const R w_yz = wy[iy] * w_z;
                for (int ix = 0; ix < 4; ++ix) {
                    if (lx[ix] < 0) continue;
                    atomicAdd(&acc[(lz[iz] * GT + ly[iy]) * GT + lx[ix]],
                              (unsigned long long)__double2ll_rn(
                                  (double)(qr * wx[ix] * w_yz) * QGRID_SCALE));
                }
            }
        }
    }
    __syncthreads();
    unsigned long long* G = qgrid + S.grid_off;
    for (int k = threadIdx.x; k < GT * GT * GT; k += blockDim.x) {
        const int lx = k % GT, ly = (k / GT) % GT, lz = k / (GT * GT);
        const int px = g0[0] + lx, py = g0[1] + ly, pz = g0[2] + lz;
        if (px < K0 && py < K1 && pz < K2) G[((int64_t)pz * K1 + py) * K0 + px] = acc[k];
    }
}

template <class R>
__global__ void __launch_bounds__(256) k_pme_spread_gather(
    const GridTile* __restrict__ tiles, const SysDev* __restrict__ sys,
    const SysState* __restrict__ st, const int* __restrict__ ccount,
    const int* __restrict__ cstart, const int* __restrict__ order,
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ q, unsigned long long* __restrict__ qgrid, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_)
    {
        spread_gather_unit<R>(u_, tiles, sys, st, ccount, cstart, order, x, y, z, q, qgrid);
        __syncthreads();  // LDS is reused by the next unit
    }
}

template <class R>
__global__ void __launch_bounds__(256) k_pme_spread_tiles(
    const SpreadTile* __restrict__ tiles, const SysDev* __restrict__ sys,
    const SysState* __restrict__ st, const int* __restrict__ order,
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ q, unsigned long long* __restrict__ qgrid, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_)
    {
        spread_tile_unit<R>(u_, tiles, sys, st, order, x, y, z, q, qgrid);
        __syncthreads();  // LDS is reused by the next unit
    }
}

// RAD-point DFT of v[0..RAD) in place; w^m = tw[m * (K / RAD)] (m < RAD).
// RAD is a compile-time constant, so v stays in registers.
template <class R, int RAD>
__device__ __forceinline__ void small_dft(Cx<R>* v, const Cx<R>* __restrict__ tw, int K, int sign)
{
    if constexpr (RAD == 2) {
        Cx<R> a = v[0], b = v[1];
        v[0] = {a.re + b.re, a.im + b.im};
        v[1] = {a.re - b.re, a.im - b.im};
    } else if constexpr (RAD == 4) {
        Cx<R> a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3];
        Cx<R> s02 = {a0.re + a2.re, a0.im + a2.im}, d02 = {a0.re - a2.re, a0.im - a2.im};
        Cx<R> s13 = {a1.re + a3.re, a1.im + a3.im}, d13 = {a1.re - a3.re, a1.im - a3.im};
        // multiply d13 by w4 = sign * i
        Cx<R> jd13 = sign < 0 ? Cx<R>{d13.im, -d13.re} : Cx<R>{-d13.im, d13.re};
        v[0] = {s02.re + s13.re, s02.im + s13.im};
        v[1] = {d02.re + jd13.re, d02.im + jd13.im};
        v[2] = {s02.re - s13.re, s02.im - s13.im};
        v[3] = {d02.re - jd13.re, d02.im - jd13.im};
    } else {
        const int step = K / RAD;
        Cx<R> w[RAD];
#pragma unroll
        for (int m = 0; m < RAD; ++m) {
            w[m] = tw[m * step];
            if (sign > 0) w[m].im = -w[m].im;
        }
        Cx<R> out[RAD];
#pragma unroll
        for (int k = 0; k < RAD; ++k) {
            Cx<R> acc = v[0];
#pragma unroll
            for (int r = 1; r < RAD; ++r) {
                Cx<R> p = cmul(v[r], w[(r * k) % RAD]);
                acc.re += p.re;
                acc.im += p.im;
            }
            out[k] = acc;
        }
#pragma unroll
        for (int k = 0; k < RAD; ++k) v[k] = out[k];
    }
}

// one radix-RAD Stockham stage over nl lines of length K in LDS (a -> b)
template <class R, int RAD>
__device__ __forceinline__ void stockham_stage(const Cx<R>* a, Cx<R>* b, int nl, int K, int Ns,
                                               const Cx<R>* __restrict__ tw, int sign)
{
    const int nb = K / RAD;
    const int sub = K / (Ns * RAD);
    for (int w = threadIdx.x; w < nl * nb; w += blockDim.x) {
        const int line = w / nb, j = w - line * nb;
        const Cx<R>* src = a + line * K;
        Cx<R>* dst = b + line * K;
        const int k = j % Ns;
        Cx<R> v[RAD];
        v[0] = src[j];
#pragma unroll
        for (int r = 1; r < RAD; ++r) {
            Cx<R> tv = tw[k * r * sub];  // k r sub < K
            if (sign > 0) tv.im = -tv.im;
            v[r] = cmul(src[j + r * nb], tv);
        }
        small_dft<R, RAD>(v, tw, K, sign);
        const int idxD = (j - k) * RAD + k;  // (j / Ns) * Ns * RAD + k
#pragma unroll
        for (int r = 0; r < RAD; ++r) dst[idxD + r * Ns] = v[r];
    }
}

// all Stockham stages of nl lines of length K held in LDS (a -> b ping-pong);
// returns the buffer that holds the result
template <class R>
__device__ __forceinline__ Cx<R>* stockham_lds(Cx<R>* a, Cx<R>* b, int nl, const FftPlan& P,
                               const Cx<R>* __restrict__ tw, int sign)
{
    const int K = P.K;
    int Ns = 1;
    for (int st = 0; st < P.nrad; ++st) {
        const int Rr = P.radix[st];
        switch (Rr) {
        case 2: stockham_stage<R, 2>(a, b, nl, K, Ns, tw, sign); break;
        case 3: stockham_stage<R, 3>(a, b, nl, K, Ns, tw, sign); break;
        case 4: stockham_stage<R, 4>(a, b, nl, K, Ns, tw, sign); break;
        case 5: stockham_stage<R, 5>(a, b, nl, K, Ns, tw, sign); break;
        case 7: stockham_stage<R, 7>(a, b, nl, K, Ns, tw, sign); break;
        case 11: stockham_stage<R, 11>(a, b, nl, K, Ns, tw, sign); break;
        default: stockham_stage<R, 13>(a, b, nl, K, Ns, tw, sign); break;
        }
        __syncthreads();
        Cx<R>* t = a;
        a = b;
        b = t;
        Ns *= Rr;
    }
    return a;
}

// dim: 0 = x, 1 = y, 2 = z.  conv = 1 on the forward z pass: after the
// forward transform multiply by G(m), accumulate the energy and run the
// inverse z transform in place.
template <class R>
__device__ __forceinline__ void k_fft_pass_unit(int64_t unit_, const FftTile* __restrict__ tiles, const SysDev* __restrict__ sys,
                           const SysState* __restrict__ st, const FftPlan* __restrict__ plans,
                           Cx<R>* __restrict__ grid, const Cx<R>* __restrict__ twiddles,
                           const double* __restrict__ bmod, const float* __restrict__ ginf,
                           int dim, int sign, int conv, double* __restrict__ epart,
                           const unsigned long long* __restrict__ qin)
{
    extern __shared__ unsigned char smem_raw[];
    const FftTile T = tiles[unit_];
    const int s = T.sys;
    if (st[s].reuse || !st[s].active) return;
    const SysDev& S = sys[s];
    const FftPlan& P = plans[3 * s + dim];
    const int K = P.K;
    const int nl = T.nlines;
    const int Kx = S.K[0], Ky = S.K[1];
    Cx<R>* bufA = reinterpret_cast<Cx<R>*>(smem_raw);
    Cx<R>* bufB = bufA + nl * K;
    // the line's twiddle table (K entries) beside the two buffers: the
    // stages read their twiddles from LDS instead of global memory
    Cx<R>* twl = bufB + nl * K;
    Cx<R>* G = grid + S.grid_off;
    const Cx<R>* twg = twiddles + S.tw_off[dim];
    for (int e = threadIdx.x; e < K; e += blockDim.x) twl[e] = twg[e];
    const Cx<R>* tw = twl;
    // element (line b, index k) -> grid offset
    auto addr = [&](int b, int k) -> int64_t {
        int l = T.line0 + b;
        if (dim == 0) return (int64_t)l * Kx + k;
        if (dim == 1) {
            int xx = l % Kx, zz = l / Kx;
            return ((int64_t)zz * Ky + k) * Kx + xx;
        }
        int xx = l % Kx, yy = l / Kx;
        return ((int64_t)k * Ky + yy) * Kx + xx;
    };
    const int tot = nl * K;
    if (dim == 0 && qin) {
        // forward x pass: the fixed-point charge grid, converted as
        // k_pme_convert converts it
        const unsigned long long* Q = qin + S.grid_off + (int64_t)T.line0 * Kx;
        for (int e = threadIdx.x; e < tot; e += blockDim.x)
            bufA[e] = {R((double)(long long)Q[e] * (1.0 / QGRID_SCALE)), R(0)};
    } else if (dim == 0) {
        for (int e = threadIdx.x; e < tot; e += blockDim.x) bufA[e] = G[(int64_t)T.line0 * Kx + e];
    } else {
        for (int e = threadIdx.x; e < tot; e += blockDim.x) {
            int b = e % nl, k = e / nl;
            bufA[b * K + k] = G[addr(b, k)];
        }
    }
    __syncthreads();
    Cx<R>* res = stockham_lds<R>(bufA, bufB, nl, P, tw, sign);
    if (conv) {
        // influence function on the full spectrum: lines are z columns at
        // fixed (x, y); m_d = k <= K_d/2 ? k : k - K_d.  Reference
        // expressions of src/force/coulomb_pme.cpp:165-196.
        __shared__ double red[256];
        const double Lx = S.L[0], Ly = S.L[1], Lz = S.L[2];
        const double V = sys_volume(S);
        const double alpha = S.alpha;
        const double* bmx = bmod + S.bm_off[0];
        const double* bmy = bmod + S.bm_off[1];
        const double* bmz = bmod + S.bm_off[2];
        const int Kz = S.K[2];
        double esum = 0.0;
        for (int e = threadIdx.x; e < tot; e += blockDim.x) {
            int b = e / K, kz = e - b * K;
            int l = T.line0 + b;
            int kx = l % Kx, ky = l / Kx;
            int mx = (kx <= Kx / 2) ? kx : kx - Kx;
            int my = (ky <= Ky / 2) ? ky : ky - Ky;
            int mz = (kz <= Kz / 2) ? kz : kz - Kz;
            Cx<R> v = res[e];
            if (mx == 0 && my == 0 && mz == 0) {
                res[e] = {R(0), R(0)};
                continue;
            }
            double green;
            if (sizeof(R) == sizeof(float) && S.ginf_static) {
                green = (double)ginf[S.grid_off + ((int64_t)kz * Ky + ky) * Kx + kx];
            } else {
                double msq = ((double)mx * mx) / (Lx * Lx) + ((double)my * my) / (Ly * Ly) +
                             ((double)mz * mz) / (Lz * Lz);
                double k2 = 4.0 * PI * PI * msq;
                double B = bmx[kx] * bmy[ky] * bmz[kz];
                green = 4.0 * PI * KE / V * exp(-k2 / (4.0 * alpha * alpha)) * B / k2;
            }
            double re = (double)v.re, im = (double)v.im;
            esum += 0.5 * green * (re * re + im * im);
            R gr = (R)green;
            res[e] = {v.re * gr, v.im * gr};
        }
        red[threadIdx.x] = esum;
        __syncthreads();
        for (int o = blockDim.x / 2; o > 0; o >>= 1) {
            if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) epart[S.pw_off + T.tile_local] = red[0];
        Cx<R>* other = (res == bufA) ? bufB : bufA;
        res = stockham_lds<R>(res, other, nl, P, tw, -sign);
    }
    if (dim == 0) {
        for (int e = threadIdx.x; e < tot; e += blockDim.x) G[(int64_t)T.line0 * Kx + e] = res[e];
    } else {
        for (int e = threadIdx.x; e < tot; e += blockDim.x) {
            int b = e % nl, k = e / nl;
            G[addr(b, k)] = res[b * K + k];
        }
    }
}

template <class R>
__global__ void __launch_bounds__(256) k_fft_pass(const FftTile* __restrict__ tiles, const SysDev* __restrict__ sys,
                           const SysState* __restrict__ st, const FftPlan* __restrict__ plans,
                           Cx<R>* __restrict__ grid, const Cx<R>* __restrict__ twiddles,
                           const double* __restrict__ bmod, const float* __restrict__ ginf,
                           int dim, int sign, int conv, double* __restrict__ epart,
                           const unsigned long long* __restrict__ qin, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_)
    {
        k_fft_pass_unit<R>(u_, tiles, sys, st, plans, grid, twiddles, bmod, ginf, dim, sign, conv,
                           epart, qin);
        __syncthreads();  // LDS is reused by the next unit
    }
}

// One xy plane (fixed z) of a system's grid per workgroup, in LDS: the
// forward x then y passes (sign -1; the fixed-point grid is converted as it
// loads, as in the x pass of k_fft_pass) or the inverse y then x passes
// (sign +1), with the plane transposed in LDS between the two.  Every line
// goes through the same Stockham stages and twiddles as in k_fft_pass, so
// the plane holds what the two separate passes would leave.
template <class R>
__device__ __forceinline__ void k_fft_plane_unit(int64_t unit_, const PlaneTile* __restrict__ tiles,
                                                 const SysDev* __restrict__ sys,
                                                 const SysState* __restrict__ st,
                                                 const FftPlan* __restrict__ plans,
                                                 Cx<R>* __restrict__ grid,
                                                 const Cx<R>* __restrict__ twiddles, int sign,
                                                 const unsigned long long* __restrict__ qin)
{
    extern __shared__ unsigned char smem_raw[];
    const PlaneTile T = tiles[unit_];
    const int s = T.sys;
    if (st[s].reuse || !st[s].active) return;
    const SysDev& S = sys[s];
    const int Kx = S.K[0], Ky = S.K[1], n = Kx * Ky;
    const FftPlan& PX = plans[3 * s];
    const FftPlan& PY = plans[3 * s + 1];
    Cx<R>* a = reinterpret_cast<Cx<R>*>(smem_raw);
    Cx<R>* b = a + n;
    Cx<R>* twx = b + n;
    Cx<R>* twy = twx + Kx;
    for (int e = threadIdx.x; e < Kx; e += blockDim.x) twx[e] = twiddles[S.tw_off[0] + e];
    for (int e = threadIdx.x; e < Ky; e += blockDim.x) twy[e] = twiddles[S.tw_off[1] + e];
    const int64_t off = S.grid_off + (int64_t)T.z * n;  // the plane: y rows of Kx
    Cx<R>* G = grid + off;
    if (qin) {
        for (int e = threadIdx.x; e < n; e += blockDim.x)
            a[e] = {R((double)(long long)qin[off + e] * (1.0 / QGRID_SCALE)), R(0)};
    } else {
        for (int e = threadIdx.x; e < n; e += blockDim.x) a[e] = G[e];
    }
    __syncthreads();
    if (sign < 0) {
        // x lines (row y at y Kx), then y lines (column x at x Ky)
        Cx<R>* r = stockham_lds<R>(a, b, Ky, PX, twx, sign);
        Cx<R>* o = (r == a) ? b : a;
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            const int y = e / Kx, x = e - y * Kx;
            o[x * Ky + y] = r[e];
        }
        __syncthreads();
        Cx<R>* r2 = stockham_lds<R>(o, r, Kx, PY, twy, sign);
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            const int y = e / Kx, x = e - y * Kx;
            G[e] = r2[x * Ky + y];
        }
    } else {
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            const int y = e / Kx, x = e - y * Kx;
            b[x * Ky + y] = a[e];
        }
        __syncthreads();
        Cx<R>* r = stockham_lds<R>(b, a, Kx, PY, twy, sign);
        Cx<R>* o = (r == a) ? b : a;
        for (int e = threadIdx.x; e < n; e += blockDim.x) {
            const int y = e / Kx, x = e - y * Kx;
            o[e] = r[x * Ky + y];
        }
        __syncthreads();
        Cx<R>* r2 = stockham_lds<R>(o, r, Ky, PX, twx, sign);
        for (int e = threadIdx.x; e < n; e += blockDim.x) G[e] = r2[e];
    }
}

template <class R>
__global__ void __launch_bounds__(256) k_fft_plane(const PlaneTile* __restrict__ tiles,
                                                   const SysDev* __restrict__ sys,
                                                   const SysState* __restrict__ st,
                                                   const FftPlan* __restrict__ plans,
                                                   Cx<R>* __restrict__ grid,
                                                   const Cx<R>* __restrict__ twiddles, int sign,
                                                   const unsigned long long* __restrict__ qin,
                                                   int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_)
    {
        k_fft_plane_unit<R>(u_, tiles, sys, st, plans, grid, twiddles, sign, qin);
        __syncthreads();  // LDS is reused by the next unit
    }
}

}  // namespace dmdx

namespace dmdx {

template <class R>
static hipError_t launch_pme_t(const Dev& d, hipStream_t s)
{
    const int BS = 256;
    Cx<R>* grid = reinterpret_cast<Cx<R>*>(d.grid);
    const Cx<R>* tw = reinterpret_cast<const Cx<R>*>(d.tw);
    const int64_t gu = blocks_for(d.grid_total, BS);
    if (d.nstiles) k_pme_zero<<<grid_for(gu), BS, 0, s>>>(d.st, d.grid_sys, d.grid_total, d.qgrid, gu);
    if (d.nstiles)
        k_pme_spread_tiles<R><<<grid_for(d.nstiles), BS, 0, s>>>(d.stiles, d.sys, d.st, d.order,
                                                                   d.x, d.y, d.z, d.q, d.qgrid,
                                                                   (int64_t)d.nstiles);
    if (d.ngtiles)
        k_pme_spread_gather<R><<<grid_for(d.ngtiles), BS, 0, s>>>(
            d.gtiles, d.sys, d.st, d.ccount, d.cstart, d.order, d.x, d.y, d.z, d.q, d.qgrid,
            (int64_t)d.ngtiles);
    // the forward x pass and the forward planes convert the fixed-point grid
    // as they load it
    if (!d.ntiles[0] && !d.nptiles)
        k_pme_convert<R><<<grid_for(gu), BS, 0, s>>>(d.st, d.grid_sys, d.grid_total, d.qgrid, grid, gu);
    if (d.nptiles)
        k_fft_plane<R><<<grid_for(d.nptiles), 256, d.plane_lds, s>>>(
            d.ptiles, d.sys, d.st, d.plans, grid, tw, -1, d.qgrid, (int64_t)d.nptiles);
    // forward x, forward y, forward z + convolution + inverse z, inverse y, inverse x
    const int dims[5] = {0, 1, 2, 1, 0};
    const int signs[5] = {-1, -1, -1, +1, +1};
    for (int p = 0; p < 5; ++p) {
        int dim = dims[p];
        if (!d.ntiles[dim]) continue;
        k_fft_pass<R><<<grid_for(d.ntiles[dim]), 256, d.lds_bytes[dim], s>>>(
            d.tiles[dim], d.sys, d.st, d.plans, grid, tw, d.bmod, d.ginf, dim, signs[p], p == 2 ? 1 : 0,
            d.p_pme, p == 0 ? d.qgrid : nullptr, (int64_t)(d.ntiles[dim]));
    }
    if (d.nptiles)
        k_fft_plane<R><<<grid_for(d.nptiles), 256, d.plane_lds, s>>>(
            d.ptiles, d.sys, d.st, d.plans, grid, tw, +1, nullptr, (int64_t)d.nptiles);
    return hipGetLastError();
}

hipError_t launch_pme(const Dev& d, hipStream_t s)
{
    if (!d.any_pme) return hipSuccess;
    return d.precision == DMDX_FP64 ? launch_pme_t<double>(d, s) : launch_pme_t<float>(d, s);
}

}  // namespace dmdx
