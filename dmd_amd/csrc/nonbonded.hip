// nonbonded.hip -- Lennard-Jones and real-space Ewald (erfc) pair kernels.
//
// One thread per atom over its full neighbour list; the atom's force is a
// private accumulator, so there are no force atomics and the result does
// not depend on scheduling.
//
// k_pair_fp64: the plain path.  Per pair it evaluates the reference's
//   expressions in their order, with contraction off:
//     LJ  src/force/lennard_jones.cpp:62-88
//     erfc src/force/coulomb_direct.cpp:60-83
//   With d = x_j - x_i for every neighbour and F_i -= f*d, the contribution
//   of each pair to F_i has the same bits as in the reference loop (the
//   reference's (j,i) pair adds f*(x_i - x_j) = -(f*d) exactly).  Neighbours
//   are in ascending index order, which is the order in which the
//   reference's lexicographic pair loop reaches atom i, so an LJ-only system
//   gets bit-identical forces.
//
// k_pair_mixed: the mixed path.  The minimum-image separation is formed in
//   FP64 (FP32 positions of a 5-10 nm box would carry ~5e-7 nm of error into
//   every separation), then everything else -- 1/r^2, the LJ powers, erfc,
//   exp -- runs in FP32 and each pair force is added to an FP64 accumulator.
#include "dmdx_internal.h"

namespace dmdx {

// accumulate: 1 = start from the forces already in fx/fy/fz (ForceComponent
// semantics), 0 = start from zero (first term of a SimulationEngine compute)
__device__ __forceinline__ void k_pair_fp64_unit(int64_t unit_, const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                            const int* __restrict__ slot_sys, int64_t nslots,
                            const int* __restrict__ order, const int* __restrict__ nn,
                            const int* __restrict__ nbr, const double* __restrict__ x,
                            const double* __restrict__ y, const double* __restrict__ z,
                            const double* __restrict__ q, const int* __restrict__ type,
                            const double* __restrict__ lj_sig, const double* __restrict__ lj_eps,
                            double* __restrict__ fx, double* __restrict__ fy,
                            double* __restrict__ fz, double* __restrict__ e_lj_part,
                            double* __restrict__ e_c_part, int accumulate)
{
    int64_t t = unit_ * blockDim.x + threadIdx.x;
    if (t >= nslots) return;  // nslots is a multiple of 64: whole waves exit
    int s = slot_sys[t];
    if (s < 0) return;  // unused tail slots: whole waves
    if (st[s].reuse || !st[s].active) return;  // uniform per wave
    const SysDev& S = sys[s];
    int lt = (int)(t - S.atom_off);
    double e_lj = 0.0, e_c = 0.0;
    if (lt < S.n) {
        int i = order[t];
        const double xi = x[i], yi = y[i], zi = z[i];
        double fxi = 0.0, fyi = 0.0, fzi = 0.0;
        if (accumulate) {
            fxi = fx[i];
            fyi = fy[i];
            fzi = fz[i];
        }
        const double Lx = S.L[0], Ly = S.L[1], Lz = S.L[2];
        const double ix = S.invL[0], iy = S.invL[1], iz = S.invL[2];
        const int nt = S.ntypes;
        const double* sg = lj_sig + S.lj_off;
        const double* ep = lj_eps + S.lj_off;
        const int ti = S.use_lj ? type[i] : 0;
        const double qi = S.coul_pairs ? q[i] : 0.0;
        const double alpha = S.alpha;
        const double cutoff2 = S.rc_lj2, ccut2 = S.rc_c2;
        const int* row = nbr + S.nbr_off + (int64_t)lt * S.max_nbr;
        const int cnt = nn[t];
        for (int k = 0; k < cnt; ++k) {
            int j = row[k];
            double dx = x[j] - xi;
            double dy = y[j] - yi;
            double dz = z[j] - zi;
            dx = min_image_dmd(dx, Lx, ix);
            dy = min_image_dmd(dy, Ly, iy);
            dz = min_image_dmd(dz, Lz, iz);
            double r2 = dx * dx + dy * dy + dz * dz;
            if (S.use_lj && !(r2 >= cutoff2)) {
                int tj = type[j];
                double eps = ep[ti * nt + tj];
                double sig = sg[ti * nt + tj];
                double sig2 = sig * sig;
                double sig6 = sig2 * sig2 * sig2;
                double r6 = r2 * r2 * r2;
                double sr6 = sig6 / r6;
                double f_scalar = 24.0 * eps * (2.0 * sr6 * sr6 - sr6) / r2;
                fxi -= f_scalar * dx;
                fyi -= f_scalar * dy;
                fzi -= f_scalar * dz;
                if (j > i) e_lj += 4.0 * eps * (sr6 * sr6 - sr6);
            }
            if (S.coul_pairs && !(r2 >= ccut2)) {
                double r = sqrt(r2);
                double qiqj = qi * q[j];
                double erfc_ar = erfc(alpha * r);
                double force_scalar = KE * qiqj * (erfc_ar / (r2 * r) +
                                      2.0 * alpha * exp(-alpha * alpha * r2) / (sqrt(PI) * r2));
                fxi -= force_scalar * dx;
                fyi -= force_scalar * dy;
                fzi -= force_scalar * dz;
                if (j > i) e_c += KE * qiqj * erfc_ar / r;
            }
        }
        fx[i] = fxi;
        fy[i] = fyi;
        fz[i] = fzi;
    }
    e_lj_part[t] = e_lj;
    e_c_part[t] = e_c;
}

__global__ void k_pair_fp64(const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                            const int* __restrict__ slot_sys, int64_t nslots,
                            const int* __restrict__ order, const int* __restrict__ nn,
                            const int* __restrict__ nbr, const double* __restrict__ x,
                            const double* __restrict__ y, const double* __restrict__ z,
                            const double* __restrict__ q, const int* __restrict__ type,
                            const double* __restrict__ lj_sig, const double* __restrict__ lj_eps,
                            double* __restrict__ fx, double* __restrict__ fy,
                            double* __restrict__ fz, double* __restrict__ e_lj_part,
                            double* __restrict__ e_c_part, int accumulate, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_) k_pair_fp64_unit(u_, sys, st, slot_sys, nslots, order, nn, nbr, x, y, z, q, type, lj_sig, lj_eps, fx, fy, fz, e_lj_part, e_c_part, accumulate);
}

// Mixed precision.  A group of TPA lanes handles one atom (cell-order slot
// t): lane l takes neighbours l, l+TPA, ... of the atom's row, reads them from
// the cell-ordered packed positions (spatially local), and the group sums its
// partial forces and energies with shuffles.  lj_c12/lj_c6 hold 4 eps sig^12
// and 4 eps sig^6 in FP32.
template <int TPA>
__device__ __forceinline__ void k_pair_mixed_unit(int64_t unit_, 
    const SysDev* __restrict__ sys, const SysState* __restrict__ st,
    const int* __restrict__ slot_sys, int64_t nslots, const int* __restrict__ order,
    const int* __restrict__ nn, const int* __restrict__ nbr, const double4* __restrict__ pos4s,
    const int* __restrict__ typ_s, const float* __restrict__ lj_c12,
    const float* __restrict__ lj_c6, double* __restrict__ fx, double* __restrict__ fy,
    double* __restrict__ fz, double* __restrict__ e_lj_part, double* __restrict__ e_c_part,
    int accumulate)
{
    // FP32 arithmetic of this kernel may fuse into FMAs (the library is built
    // with contraction off for the FP64 parity kernels)
#pragma clang fp contract(fast)
    constexpr int GPB = 256 / TPA;
    const int lane = threadIdx.x % TPA;
    const int64_t t = unit_ * GPB + threadIdx.x / TPA;
    if (t >= nslots) return;
    const int s = slot_sys[t];
    if (s < 0) return;
    if (st[s].reuse || !st[s].active) return;
    const SysDev& S = sys[s];
    const int lt = (int)(t - S.atom_off);
    if (lt >= S.n) {
        if (lane == 0) {
            e_lj_part[t] = 0.0;
            e_c_part[t] = 0.0;
        }
        return;
    }
    const double4 pi = pos4s[t];
    const double Lx = S.L[0], Ly = S.L[1], Lz = S.L[2];
    const double ix = S.invL[0], iy = S.invL[1], iz = S.invL[2];
    const int nt = S.ntypes;
    const float* c12t = lj_c12 + S.lj_off;
    const float* c6t = lj_c6 + S.lj_off;
    const bool lj = S.use_lj, cp = S.coul_pairs;
    const int ti = lj ? typ_s[t] : 0;
    const float qi = cp ? (float)(KE * pi.w) : 0.0f;
    const float alpha = (float)S.alpha;
    const float two_a_rpi = (float)(2.0 * S.alpha / sqrt(PI));
    const float a2 = (float)(S.alpha * S.alpha);
    const float cut2 = (float)S.rc_lj2, ccut2 = (float)S.rc_c2;
    const int* row = nbr + S.nbr_off + (int64_t)lt * S.max_nbr;
    const int cnt = nn[t];
    double fxi = 0.0, fyi = 0.0, fzi = 0.0, ej = 0.0, ec = 0.0;
    // rows are 16-byte aligned with M a multiple of 4: each lane takes 4
    // consecutive entries per pass and issues their 4 gathers before any
    // arithmetic (memory-level parallelism for the dependent loads)
    for (int k0 = lane * 4; k0 < cnt; k0 += TPA * 4) {
        const int4 jj = *reinterpret_cast<const int4*>(row + k0);
        const int jv[4] = {jj.x, jj.y, jj.z, jj.w};
        double4 pj[4];
        int tj[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = (k0 + q < cnt) ? jv[q] : (int)t;
            pj[q] = pos4s[j];
            tj[q] = lj ? typ_s[j] : 0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (k0 + q >= cnt) continue;
            double ddx = pj[q].x - pi.x, ddy = pj[q].y - pi.y, ddz = pj[q].z - pi.z;
            ddx -= rint(ddx * ix) * Lx;
            ddy -= rint(ddy * iy) * Ly;
            ddz -= rint(ddz * iz) * Lz;
            const float dx = (float)ddx, dy = (float)ddy, dz = (float)ddz;
            const float r2 = dx * dx + dy * dy + dz * dz;
            const float rinv2 = __builtin_amdgcn_rcpf(r2);  // 1 ulp hardware reciprocal
            float fs = 0.0f, el = 0.0f, ecl = 0.0f;
            if (lj && r2 < cut2) {
                const float c12 = c12t[ti * nt + tj[q]], c6 = c6t[ti * nt + tj[q]];
                const float r6i = rinv2 * rinv2 * rinv2;
                const float a12 = c12 * r6i * r6i, a6 = c6 * r6i;
                fs += (12.0f * a12 - 6.0f * a6) * rinv2;
                el = a12 - a6;
            }
            if (cp && r2 < ccut2) {
                const float rinv = rsqrtf(r2);
                const float r = r2 * rinv;
                const float qq = qi * (float)pj[q].w;
                const float erfc_ar = erfcf(alpha * r);
                const float ex = __expf(-a2 * r2);
                fs += qq * (erfc_ar * rinv + two_a_rpi * ex) * rinv2;
                ecl = qq * erfc_ar * rinv;
            }
            const double fsd = (double)fs;
            fxi -= fsd * ddx;
            fyi -= fsd * ddy;
            fzi -= fsd * ddz;
            if (jv[q] > t) {  // each unordered pair once
                ej += (double)el;
                ec += (double)ecl;
            }
        }
    }
    fxi = group_sum<TPA>(fxi);
    fyi = group_sum<TPA>(fyi);
    fzi = group_sum<TPA>(fzi);
    ej = group_sum<TPA>(ej);
    ec = group_sum<TPA>(ec);
    if (lane == 0) {
        const int i = order[t];
        if (accumulate) {
            fxi += fx[i];
            fyi += fy[i];
            fzi += fz[i];
        }
        fx[i] = fxi;
        fy[i] = fyi;
        fz[i] = fzi;
        e_lj_part[t] = ej;
        e_c_part[t] = ec;
    }
}

template <int TPA>
__global__ void __launch_bounds__(256) k_pair_mixed(
    const SysDev* __restrict__ sys, const SysState* __restrict__ st,
    const int* __restrict__ slot_sys, int64_t nslots, const int* __restrict__ order,
    const int* __restrict__ nn, const int* __restrict__ nbr, const double4* __restrict__ pos4s,
    const int* __restrict__ typ_s, const float* __restrict__ lj_c12,
    const float* __restrict__ lj_c6, double* __restrict__ fx, double* __restrict__ fy,
    double* __restrict__ fz, double* __restrict__ e_lj_part, double* __restrict__ e_c_part,
    int accumulate, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_) k_pair_mixed_unit<TPA>(u_, sys, st, slot_sys, nslots, order, nn, nbr, pos4s, typ_s, lj_c12, lj_c6, fx, fy, fz, e_lj_part, e_c_part, accumulate);
}

}  // namespace dmdx

namespace dmdx {

hipError_t launch_pair(const Dev& d, hipStream_t s, int accumulate)
{
    const int BS = 256;
    if (d.precision == DMDX_FP64)
        k_pair_fp64<<<grid_for(blocks_for(d.nslots, BS)), BS, 0, s>>>(
            d.sys, d.st, d.slot_sys, d.nslots, d.order, d.nn, d.nbr, d.x, d.y, d.z, d.q, d.type,
            d.lj_sig, d.lj_eps, d.fx, d.fy, d.fz, d.p_lj, d.p_coul, accumulate, (int64_t)(blocks_for(d.nslots, BS)));
    else if (d.pair_tpa == 8)
        k_pair_mixed<8><<<grid_for(blocks_for(d.nslots * 8, BS)), BS, 0, s>>>(
            d.sys, d.st, d.slot_sys, d.nslots, d.order, d.nn, d.nbr, d.pos4s, d.typ_s, d.lj_c12,
            d.lj_c6, d.fx, d.fy, d.fz, d.p_lj, d.p_coul, accumulate, (int64_t)(blocks_for(d.nslots * 8, BS)));
    else if (d.pair_tpa == 32)
        k_pair_mixed<32><<<grid_for(blocks_for(d.nslots * 32, BS)), BS, 0, s>>>(
            d.sys, d.st, d.slot_sys, d.nslots, d.order, d.nn, d.nbr, d.pos4s, d.typ_s, d.lj_c12,
            d.lj_c6, d.fx, d.fy, d.fz, d.p_lj, d.p_coul, accumulate, (int64_t)(blocks_for(d.nslots * 32, BS)));
    else
        k_pair_mixed<16><<<grid_for(blocks_for(d.nslots * 16, BS)), BS, 0, s>>>(
            d.sys, d.st, d.slot_sys, d.nslots, d.order, d.nn, d.nbr, d.pos4s, d.typ_s, d.lj_c12,
            d.lj_c6, d.fx, d.fy, d.fz, d.p_lj, d.p_coul, accumulate, (int64_t)(blocks_for(d.nslots * 16, BS)));
    return hipGetLastError();
}

}  // namespace dmdx
