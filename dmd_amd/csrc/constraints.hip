// constraints.hip -- SHAKE (positions) and RATTLE (velocities) distance
// constraints, FP64, in the reference's Gauss-Seidel order.
//
//   k_shake        after the drift      (src/integrate/constraints.cpp:9-49,
//                                        called at simulation_engine.cpp:67)
//   k_rattle       after the closing kick (constraints.cpp:51-91, called at
//                                        simulation_engine.cpp:78)
//   k_ke_refresh   kinetic-energy partials of the velocities left by RATTLE
//                  or the Andersen collisions, for the thermostat and the
//                  end-of-step kinetic energy
//
// The reference sweeps all constraints in order, correcting each one that
// is off by more than the tolerance with the positions the earlier ones in
// the same sweep left, and repeats (at most 100 sweeps) until a sweep sees
// every |error| below the tolerance.  Constraints of different connected
// components (atoms linked by constraints) touch disjoint atoms, so their
// updates commute, and a component whose constraints are all within the
// tolerance is left unchanged by every later sweep.  One thread therefore
// runs the sweeps of one component, in the caller's order, until that
// component converges or reaches 100 sweeps: the same positions and
// velocities, bit for bit, as the reference's global loop.

#include "dmdx_internal.h"

namespace dmdx {

namespace {

constexpr int MAX_SWEEPS = 100;  // constraints.cpp:10, :52

__device__ __forceinline__ void con_vector(const SysDev& S, const double* __restrict__ x,
                                           const double* __restrict__ y,
                                           const double* __restrict__ z, int ai, int aj,
                                           double& dx, double& dy, double& dz)
{
    dx = x[aj] - x[ai];
    dy = y[aj] - y[ai];
    dz = z[aj] - z[ai];
    dx = min_image_dmd(dx, S.L[0], S.iminv[0]);
    dy = min_image_dmd(dy, S.L[1], S.iminv[1]);
    dz = min_image_dmd(dz, S.L[2], S.iminv[2]);
}

}  // namespace

__global__ void __launch_bounds__(64) k_shake(const SysDev* __restrict__ sys,
                                              const SysState* __restrict__ st, int ncomp,
                                              const int* __restrict__ k_sys,
                                              const int* __restrict__ k_start,
                                              const int* __restrict__ k_i,
                                              const int* __restrict__ k_j,
                                              const double* __restrict__ k_d,
                                              const double* __restrict__ m,
                                              double* __restrict__ x, double* __restrict__ y,
                                              double* __restrict__ z)
{
    const int c = blockIdx.x * 64 + threadIdx.x;
    if (c >= ncomp) return;
    const int s = k_sys[c];
    if (!st[s].active) return;
    const SysDev& S = sys[s];
    const double tol = S.con_tol;
    const int k0 = k_start[c], k1 = k_start[c + 1];
    for (int it = 0; it < MAX_SWEEPS; ++it) {
        double max_error = 0.0;
        for (int k = k0; k < k1; ++k) {
            const int ai = k_i[k], aj = k_j[k];
            double dx, dy, dz;
            con_vector(S, x, y, z, ai, aj, dx, dy, dz);
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double d2 = k_d[k] * k_d[k];
            const double error = r2 - d2;
            if (fabs(error) > max_error) max_error = fabs(error);
            if (fabs(error) > tol) {
                const double imi = 1.0 / m[ai];
                const double imj = 1.0 / m[aj];
                const double ims = imi + imj;
                const double lambda = error / (2.0 * d2 * ims);
                x[ai] += lambda * dx * imi;
                y[ai] += lambda * dy * imi;
                z[ai] += lambda * dz * imi;
                x[aj] -= lambda * dx * imj;
                y[aj] -= lambda * dy * imj;
                z[aj] -= lambda * dz * imj;
            }
        }
        if (max_error < tol) break;
    }
}

__global__ void __launch_bounds__(64) k_rattle(const SysDev* __restrict__ sys,
                                               const SysState* __restrict__ st, int ncomp,
                                               const int* __restrict__ k_sys,
                                               const int* __restrict__ k_start,
                                               const int* __restrict__ k_i,
                                               const int* __restrict__ k_j,
                                               const double* __restrict__ m,
                                               const double* __restrict__ x,
                                               const double* __restrict__ y,
                                               const double* __restrict__ z,
                                               double* __restrict__ vx, double* __restrict__ vy,
                                               double* __restrict__ vz)
{
    const int c = blockIdx.x * 64 + threadIdx.x;
    if (c >= ncomp) return;
    const int s = k_sys[c];
    if (!st[s].active) return;
    const SysDev& S = sys[s];
    const double tol = S.con_tol;
    const int k0 = k_start[c], k1 = k_start[c + 1];
    for (int it = 0; it < MAX_SWEEPS; ++it) {
        double max_error = 0.0;
        for (int k = k0; k < k1; ++k) {
            const int ai = k_i[k], aj = k_j[k];
            double dx, dy, dz;
            con_vector(S, x, y, z, ai, aj, dx, dy, dz);
            const double dvx = vx[aj] - vx[ai];
            const double dvy = vy[aj] - vy[ai];
            const double dvz = vz[aj] - vz[ai];
            const double r_dot_v = dx * dvx + dy * dvy + dz * dvz;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double err = fabs(r_dot_v);
            if (err > max_error) max_error = err;
            if (err > tol) {
                const double imi = 1.0 / m[ai];
                const double imj = 1.0 / m[aj];
                const double ims = imi + imj;
                const double lambda = r_dot_v / (r2 * ims);
                vx[ai] += lambda * dx * imi;
                vy[ai] += lambda * dy * imi;
                vz[ai] += lambda * dz * imi;
                vx[aj] -= lambda * dx * imj;
                vy[aj] -= lambda * dy * imj;
                vz[aj] -= lambda * dz * imj;
            }
        }
        if (max_error < tol) break;
    }
}

// Kinetic-energy partials (one per 64-slot wave, the layout k_kick_ke
// writes) of the systems whose velocities changed after the closing kick.
// A wave never straddles two systems (n_pad is a multiple of 64), so the
// early exits are uniform over a wave.
__global__ void k_ke_refresh(const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                             const int* __restrict__ slot_sys, int64_t nslots,
                             const double* __restrict__ m, const double* __restrict__ vx,
                             const double* __restrict__ vy, const double* __restrict__ vz,
                             double* __restrict__ ke_part, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_)
    {
        const int64_t g = u_ * blockDim.x + threadIdx.x;
        if (g >= nslots) continue;
        const int s = slot_sys[g];
        if (s < 0 || !st[s].active) continue;
        const SysDev& S = sys[s];
        if (S.n_con == 0 && S.thermostat != DMDX_THERMO_ANDERSEN) continue;
        double ke = 0.0;
        if (g - S.atom_off < S.n) {
            const double a = vx[g], b = vy[g], c = vz[g];
            ke = 0.5 * m[g] * (a * a + b * b + c * c);
        }
        ke = wave_sum(ke);
        if (lane_id() == 0) ke_part[g >> 6] = ke;
    }
}

hipError_t launch_shake(const Dev& d, hipStream_t s)
{
    if (d.ncomp == 0) return hipSuccess;
    k_shake<<<blocks_for(d.ncomp, 64), 64, 0, s>>>(d.sys, d.st, d.ncomp, d.k_sys, d.k_start,
                                                   d.k_i, d.k_j, d.k_d, d.m, d.x, d.y, d.z);
    return hipGetLastError();
}

hipError_t launch_rattle(const Dev& d, hipStream_t s)
{
    if (d.ncomp == 0) return hipSuccess;
    k_rattle<<<blocks_for(d.ncomp, 64), 64, 0, s>>>(d.sys, d.st, d.ncomp, d.k_sys, d.k_start,
                                                    d.k_i, d.k_j, d.m, d.x, d.y, d.z, d.vx,
                                                    d.vy, d.vz);
    return hipGetLastError();
}

hipError_t launch_ke_refresh(const Dev& d, hipStream_t s)
{
    if (d.ncomp == 0 && !d.any_andersen) return hipSuccess;
    const int BS = 256;
    k_ke_refresh<<<grid_for(blocks_for(d.nslots, BS)), BS, 0, s>>>(
        d.sys, d.st, d.slot_sys, d.nslots, d.m, d.vx, d.vy, d.vz, d.p_ke,
        (int64_t)blocks_for(d.nslots, BS));
    return hipGetLastError();
}

}  // namespace dmdx
