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
// k_pair_mixed: the mixed path.  Positions are FP32 offsets from the origin
//   of the atom's list-build cell (lpos = x - org, formed in FP64 and
//   rounded once), so they stay below a cell width and carry ~1e-8 nm of
//   error instead of the ~5e-7 nm of absolute FP32 coordinates in a 5-10 nm
//   box.  A list entry holds the neighbour's slot and the offset of its cell
//   from the atom's cell; the separation is l_j - l_i + offset * width, with
//   offset * width read from a per-wave table of the 125 offsets (the minimum
//   image when the system has >= 2 cdiv + 1 cells per edge; otherwise the
//   separation is wrapped by the box in FP32).  Everything else -- 1/r^2, the
//   LJ powers, erfc, exp -- runs in FP32 and each pair force is added to an
//   FP64 accumulator.
//
//   Charged systems also add the part of offset * width that FP32 drops:
//   the table holds offset * fl32(w) exactly (|offset| <= 2), which misses
//   offset * (w - fl32(w)), ~3e-8 nm per cell crossed.  That error is the
//   same for every pair of a cell offset and stays fixed between list
//   builds, so it is not noise: measured as the work the force error does
//   along a mixed NVE run of 12,288-atom water, it injected 1e-3
//   kJ/mol/ns/atom, and 3e-4 with the correction (fma(shift, rho, d),
//   rho = (w - fl32(w)) / fl32(w)); the wrap by the box gets the same
//   treatment.  LJ-only systems, where the same term was 1e-4, skip the
//   three FMAs.
#include "dmdx_internal.h"

#include <type_traits>

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
        const double ix = S.iminv[0], iy = S.iminv[1], iz = S.iminv[2];
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
// t): lane l takes neighbours l, l + TPA, l + 2 TPA, ... of the atom's row, reads
// their cell-local FP32 positions (spatially local gathers of 16 bytes) and
// rebuilds each displacement as l_j - l_i + offset * w from the offset
// stored in the list entry -- no FP64 and no minimum-image rounding per pair
// when the system has at least 2 cdiv + 1 cells per edge.  The W entries of a pass
// are summed in FP32 and the pass sums accumulate in FP64; the group sums
// its partial forces and energies with shuffles.
// lj_c12/lj_c6 hold 4 eps sig^12 and 4 eps sig^6 in FP32.
template <int TPA, int W, bool COUL>
__device__ __forceinline__ void k_pair_mixed_unit(int64_t unit_,
    const SysDev* __restrict__ sys, const SysState* __restrict__ st,
    const int* __restrict__ slot_sys, int64_t nslots, const int* __restrict__ order,
    const int* __restrict__ nn, const int* __restrict__ nbr, const float4* __restrict__ lpos,
    const int* __restrict__ typ_s, const float* __restrict__ lj_c12,
    const float* __restrict__ lj_c6, double* __restrict__ fx, double* __restrict__ fy,
    double* __restrict__ fz, double* __restrict__ e_lj_part, double* __restrict__ e_c_part,
    int accumulate, int rows_on, float4* __restrict__ tab, int& tab_sys)
{
    // FP32 arithmetic of this kernel may fuse into FMAs (the library is built
    // with contraction off for the FP64 parity kernels)
#pragma clang fp contract(fast)
    constexpr int GPB = 256 / TPA;
    const int lane = threadIdx.x % TPA;
    const int64_t t = unit_ * GPB + threadIdx.x / TPA;
    if (t >= nslots) return;  // whole waves: nslots is a multiple of 64
    // A wave's atoms are 64 / TPA consecutive slots inside one 64-slot block,
    // and a system owns whole blocks: the system is wave-uniform, so its
    // constants live in scalar registers and the branches on them are
    // scalar branches.
    const int s = __builtin_amdgcn_readfirstlane(slot_sys[t]);
    if (s < 0) return;
    const SysState& T = st[s];
    if (T.reuse || !T.active) return;
    const SysDev& S = sys[s];
    // Each system is computed by the one variant its own terms select
    // (COUL: real-space Coulomb pairs), whatever else shares the batch, so
    // its forces are bit-identical alone and in any mix; with row tiles on,
    // k_pair_rows takes every LJ-only system that has a list.
    if ((S.coul_pairs != 0) != COUL) return;
    if (!COUL && rows_on && S.has_list) return;
    const int lt = (int)(t - S.atom_off);
    // padding slots stay in the wave (no entries) for the wave's energy sum
    const bool real = lt < S.n;
    const float4 pi = real ? lpos[t] : make_float4(0.f, 0.f, 0.f, 0.f);
    // cell widths of the frames (cells of the last build; a barostat scales
    // the box, the positions and the frame origins together)
    const float wx = (float)(S.L[0] / T.nub[0]);
    const float wy = (float)(S.L[1] / T.nub[1]);
    const float wz = (float)(S.L[2] / T.nub[2]);
    // the wave's table of cell offsets times cell widths, indexed by the
    // 9-bit offset field of an entry; rebuilt when the wave moves to another
    // system (its units are mostly consecutive blocks of one system)
    if (s != tab_sys) {
        for (int c = threadIdx.x & 63; c < 125; c += 64) {
            const int ox = c % 5 - 2, oy = (c / 5) % 5 - 2, oz = c / 25 - 2;
            tab[nbr_code(ox, oy, oz)] =
                make_float4((float)ox * wx, (float)oy * wy, (float)oz * wz, 0.0f);
        }
        // row padding (the atom itself, 1e15 nm away): r2 = 1e30, outside
        // every cutoff, and fs * dx = 0 exactly
        if ((threadIdx.x & 63) == 0) tab[NBR_PAD_CODE] = make_float4(1e15f, 0.0f, 0.0f, 0.0f);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        tab_sys = s;
    }
    const int full = 2 * S.cdiv + 1;  // cells per edge for offsets to be minimum images
    const bool wrap = T.nub[0] < full || T.nub[1] < full || T.nub[2] < full;
    const int64_t a0 = S.atom_off;
    const float Lx = (float)S.L[0], Ly = (float)S.L[1], Lz = (float)S.L[2];
    const float ix = (float)S.iminv[0], iy = (float)S.iminv[1], iz = (float)S.iminv[2];
    // the FP32 shift table's and box's missing low parts (charged systems)
    const float rhox = (float)((S.L[0] / T.nub[0] - (double)wx) / (double)wx);
    const float rhoy = (float)((S.L[1] / T.nub[1] - (double)wy) / (double)wy);
    const float rhoz = (float)((S.L[2] / T.nub[2] - (double)wz) / (double)wz);
    const float Lxlo = (float)(S.L[0] - (double)Lx), Lylo = (float)(S.L[1] - (double)Ly),
                Lzlo = (float)(S.L[2] - (double)Lz);
    const int nt = S.ntypes;
    const bool lj = S.use_lj, cp = COUL;
    const bool typed = lj && nt > 1;
    const float* c12t = lj_c12 + S.lj_off;
    const float* c6t = lj_c6 + S.lj_off;
    const int ti = typed && real ? typ_s[t] : 0;
    const float c12_0 = c12t[0], c6_0 = c6t[0];
    const float c12x12 = 12.0f * c12_0, c6x6 = 6.0f * c6_0;
    const float qi = cp ? (float)(KE * (double)pi.w) : 0.0f;
    const float alpha = (float)S.alpha;
    const float two_a_rpi = (float)(2.0 * S.alpha / sqrt(PI));
    const float a2 = (float)(S.alpha * S.alpha);
    const float cut2 = (float)S.rc_lj2, ccut2 = (float)S.rc_c2;
    const int* row = nbr + S.nbr_off + (int64_t)lt * S.max_nbr;
    const float4* __restrict__ lps = lpos + a0;  // this system's slots
    const int* __restrict__ tys = typ_s + a0;
    const int cnt = real ? nn[t] : 0;
    double fxi = 0.0, fyi = 0.0, fzi = 0.0, ej = 0.0, ec = 0.0;
    // Each lane takes W (2, 4 or 8) entries per pass and issues their W gathers
    // before any arithmetic (memory-level parallelism for the dependent
    // loads; W=2 needs fewer registers, so more waves fit).  The wave-uniform
    // switches (minimum-image wrap, typed LJ tables) select a loop variant
    // outside the loop, so its body carries no branch on them.
    auto passes = [&](auto wrap_c, auto typed_c) {
        constexpr bool WRAP = decltype(wrap_c)::value;
        constexpr bool TYPED = decltype(typed_c)::value;
        constexpr int STRIDE = TPA * W;
        // Pass p covers entries [p STRIDE, (p + 1) STRIDE) of the row.
        // Coulomb variant: lane l takes entries l, l + TPA, ..., so each
        // gather instruction reads TPA consecutive entries of a row --
        // consecutive cell-order slots, a cache line or two -- and the
        // entries of pass p + 1 are loaded while pass p is gathered and
        // computed; entry reads are clamped into the row (index < M) and
        // entries at or past cnt never address memory.  LJ variant (less
        // arithmetic per entry, so occupancy matters more): lane l takes the
        // W consecutive entries W l .. W l + W - 1 with one or two vector
        // loads (rows are 16-byte aligned, M a multiple of 8).
        constexpr bool STRIDED = COUL;
        // Blocked lanes without the minimum image need no per-entry test of
        // the row end: a lane's W-entry pass (W = 2, 4, 8) starts at a
        // multiple of W below cnt, and the list build pads rows to a
        // multiple of 8 with entries that add zero.  (Strided lanes reach
        // past that padding; the minimum image would fold the padding's
        // shift back into the box.)
        constexpr bool CHECK_END = STRIDED || WRAP;
        auto entry_index = [&](int k0, int q) {
            return STRIDED ? k0 + lane + q * TPA : k0 + lane * W + q;
        };
        const int mlast = S.max_nbr - 1;
        auto load_entries = [&](int k0, uint32_t(&ev)[W]) {
            if constexpr (STRIDED) {
#pragma unroll
                for (int q = 0; q < W; ++q) {
                    const int k = k0 + lane + q * TPA;
                    ev[q] = (uint32_t)row[k < mlast ? k : mlast];
                }
            } else if constexpr (W == 8) {
                const int4 ja = *reinterpret_cast<const int4*>(row + k0 + lane * W);
                const int4 jb = *reinterpret_cast<const int4*>(row + k0 + lane * W + 4);
                ev[0] = (uint32_t)ja.x;
                ev[1] = (uint32_t)ja.y;
                ev[2] = (uint32_t)ja.z;
                ev[3] = (uint32_t)ja.w;
                ev[4] = (uint32_t)jb.x;
                ev[5] = (uint32_t)jb.y;
                ev[6] = (uint32_t)jb.z;
                ev[7] = (uint32_t)jb.w;
            } else if constexpr (W == 4) {
                const int4 jj = *reinterpret_cast<const int4*>(row + k0 + lane * W);
                ev[0] = (uint32_t)jj.x;
                ev[1] = (uint32_t)jj.y;
                ev[2] = (uint32_t)jj.z;
                ev[3] = (uint32_t)jj.w;
            } else {
                const int2 jj = *reinterpret_cast<const int2*>(row + k0 + lane * W);
                ev[0] = (uint32_t)jj.x;
                ev[1] = (uint32_t)jj.y;
            }
        };
        // blocked lanes start at lane W, so their loop runs while that is
        // below cnt, and they load a pass's entries only when it starts
        // below cnt (their vector loads stay inside the row)
        const int kbase = STRIDED ? 0 : lane * W;
        uint32_t ev[W];
        if (STRIDED || kbase < cnt) load_entries(0, ev);
        for (int kp = kbase; kp < cnt; kp += STRIDE) {
            const int k0 = kp - kbase;
            float4 pj[W];
            int tj[W];
#pragma unroll
            for (int q = 0; q < W; ++q) {
                const uint32_t j = (!CHECK_END || entry_index(k0, q) < cnt)
                                       ? (ev[q] & NBR_SLOT_MASK) : (uint32_t)lt;
                pj[q] = lps[j];
                tj[q] = TYPED ? tys[j] : 0;
            }
            // the next pass's entries load while this pass gathers
            uint32_t en[W];
            if (STRIDED || kp + STRIDE < cnt) load_entries(k0 + STRIDE, en);
            // the W entries of a pass are summed in FP32 (starting from the
            // first entry's terms), then added in FP64
            float px = 0.0f, py = 0.0f, pz = 0.0f, pel = 0.0f, pec = 0.0f;
            // Branch-free entries: an entry past the row end (its gather read
            // the atom itself, r2 = 0) or outside a cutoff computes on its own
            // r2 -- infinities and NaNs included -- and contributes exactly
            // zero through the selects on its force and energy terms.
#pragma unroll
            for (int q = 0; q < W; ++q) {
                const bool valid = !CHECK_END || entry_index(k0, q) < cnt;
                const uint32_t e = ev[q];
                // the neighbour cell's offset times the cell widths, from the
                // wave's table (the entry's top nine bits index it)
                const float4 sh = tab[valid ? e >> NBR_SLOT_BITS : 0u];
                float dx = (pj[q].x - pi.x) + sh.x;
                float dy = (pj[q].y - pi.y) + sh.y;
                float dz = (pj[q].z - pi.z) + sh.z;
                if constexpr (COUL) {
                    dx = fmaf(sh.x, rhox, dx);
                    dy = fmaf(sh.y, rhoy, dy);
                    dz = fmaf(sh.z, rhoz, dz);
                }
                if constexpr (WRAP) {
                    const float kx = rintf(dx * ix), ky = rintf(dy * iy), kz = rintf(dz * iz);
                    dx -= kx * Lx;
                    dy -= ky * Ly;
                    dz -= kz * Lz;
                    if constexpr (COUL) {
                        dx = fmaf(-kx, Lxlo, dx);
                        dy = fmaf(-ky, Lylo, dy);
                        dz = fmaf(-kz, Lzlo, dz);
                    }
                }
                const float r2 = dx * dx + dy * dy + dz * dz;
                const bool in_lj = lj && valid && r2 < cut2;
                const float rinv2 = __builtin_amdgcn_rcpf(r2);  // 1 ulp hardware reciprocal
                const float r6i = rinv2 * rinv2 * rinv2;
                // f/r = r^-8 (12 c12 r^-6 - 6 c6), e = r^-6 (c12 r^-6 - c6),
                // with c12 = 4 eps sig^12, c6 = 4 eps sig^6
                float fs, el;
                if constexpr (TYPED) {
                    const float c12 = c12t[ti * nt + tj[q]], c6 = c6t[ti * nt + tj[q]];
                    fs = in_lj ? fmaf(12.0f * c12, r6i, -6.0f * c6) * r6i * rinv2 : 0.0f;
                    el = in_lj ? fmaf(c12, r6i, -c6) * r6i : 0.0f;
                } else {
                    fs = in_lj ? fmaf(c12x12, r6i, -c6x6) * r6i * rinv2 : 0.0f;
                    el = in_lj ? fmaf(c12_0, r6i, -c6_0) * r6i : 0.0f;
                }
                float ecl = 0.0f;
                if constexpr (COUL) {
                    const bool in_c = cp && valid && r2 < ccut2;
                    const float rinv = rsqrtf(r2);
                    const float r = r2 * rinv;
                    const float qq = qi * pj[q].w;
                    const float ex = __expf(-a2 * r2);
                    // erfc(a r) by Abramowitz & Stegun 7.1.26 (|error| <= 1.5e-7),
                    // sharing exp(-(a r)^2) with the force term
                    const float tt = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * (alpha * r));
                    const float erfc_ar =
                        tt * (0.254829592f + tt * (-0.284496736f + tt * (1.421413741f +
                              tt * (-1.453152027f + tt * 1.061405429f)))) * ex;
                    fs += in_c ? qq * (erfc_ar * rinv + two_a_rpi * ex) * (rinv * rinv) : 0.0f;
                    ecl = in_c ? qq * erfc_ar * rinv : 0.0f;
                }
                if (q == 0) {
                    px = fs * dx;
                    py = fs * dy;
                    pz = fs * dz;
                    pel = el;
                    pec = ecl;
                } else {
                    px += fs * dx;
                    py += fs * dy;
                    pz += fs * dz;
                    pel += el;
                    pec += ecl;
                }
            }
            fxi -= (double)px;
            fyi -= (double)py;
            fzi -= (double)pz;
            ej += (double)pel;
            if (COUL) ec += (double)pec;
#pragma unroll
            for (int q = 0; q < W; ++q) ev[q] = en[q];
        }
    };
    using T_ = std::true_type;
    using F_ = std::false_type;
    if (wrap) {
        if (typed) passes(T_{}, T_{});
        else passes(T_{}, F_{});
    } else {
        if (typed) passes(F_{}, T_{});
        else passes(F_{}, F_{});
    }
    fxi = group_sum<TPA>(fxi);
    fyi = group_sum<TPA>(fyi);
    fzi = group_sum<TPA>(fzi);
    ej = group_sum<TPA>(ej);
    ec = group_sum<TPA>(ec);
    // one energy partial per wave (its 64 / TPA atom slots), a fixed tree
    const double ejw = wave_sum(lane == 0 ? ej : 0.0);
    const double ecw = COUL ? wave_sum(lane == 0 ? ec : 0.0) : 0.0;
    if (lane == 0 && real) {
        const int i = order[t];
        if (accumulate) {
            fxi += fx[i];
            fyi += fy[i];
            fzi += fz[i];
        }
        fx[i] = fxi;
        fy[i] = fyi;
        fz[i] = fzi;
    }
    // the partial of the wave's atoms lt .. lt + 64 / TPA - 1, at
    // atom_off + lt / (64 / TPA): inside the system's own slot range, whatever
    // shape the other systems' variant has
    if ((threadIdx.x & 63) == 0) {  // rows hold both orders of a pair: halved
        e_lj_part[a0 + lt / (64 / TPA)] = 0.5 * ejw;
        e_c_part[a0 + lt / (64 / TPA)] = 0.5 * ecw;
    }
}

template <int TPA, int W, bool COUL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(COUL ? 4 : 5))) k_pair_mixed(
    const SysDev* __restrict__ sys, const SysState* __restrict__ st,
    const int* __restrict__ slot_sys, int64_t nslots, const int* __restrict__ order,
    const int* __restrict__ nn, const int* __restrict__ nbr, const float4* __restrict__ lpos,
    const int* __restrict__ typ_s, const float* __restrict__ lj_c12,
    const float* __restrict__ lj_c6, double* __restrict__ fx, double* __restrict__ fy,
    double* __restrict__ fz, double* __restrict__ e_lj_part, double* __restrict__ e_c_part,
    int accumulate, int rows_on, int64_t nunits_)
{
    __shared__ float4 shift_tab[4][NBR_CODES];  // per wave
    int tab_sys = -1;
    DMDX_FOR_UNITS(u_, nunits_) k_pair_mixed_unit<TPA, W, COUL>(u_, sys, st, slot_sys, nslots, order, nn, nbr, lpos, typ_s, lj_c12, lj_c6, fx, fy, fz, e_lj_part, e_c_part, accumulate, rows_on, shift_tab[threadIdx.x >> 6], tab_sys);
}

}  // namespace dmdx

namespace dmdx {

// Launch shape of the mixed pair kernel from the tuning values: lanes per
// atom x entries per lane pass, for the LJ-only variant or (coul) the
// variant with real-space Coulomb pairs.  8 entries per pass come with 2, 4
// or 8 lanes; 2 or 4 entries with 4 (4 entries only), 8, 16 or 32 lanes.
// Other combinations take the nearest of these.
void pair_shape(const Dev& d, int& tpa, int& w, bool coul)
{
    w = coul ? d.coul_width : d.pair_width;
    tpa = coul ? d.coul_tpa : d.pair_tpa;
    if (w == 8) {
        if (tpa > 8) tpa = 8;
    } else {
        if (w != 2) w = 4;
        if (tpa < 8 && !(tpa == 4 && w == 4)) tpa = 8;
    }
}

template <bool COUL>
static hipError_t launch_pair_mixed(const Dev& d, hipStream_t s, int accumulate)
{
    const int BS = 256;
    int tpa, w;
    pair_shape(d, tpa, w, COUL);
    const int64_t units = blocks_for(d.nslots * tpa, BS);
#define DMDX_PAIR_ARGS                                                                       \
    d.sys, d.st, d.slot_sys, d.nslots, d.order, d.nn, d.nbr, d.lpos, d.typ_s, d.lj_c12, d.lj_c6, \
        d.fx, d.fy, d.fz, d.p_lj, d.p_coul, accumulate, d.pair_rows != 0 ? 1 : 0, units
    const unsigned g = grid_for(units);
#define DMDX_PAIR_SHAPE(T_, W_) \
    case T_ * 100 + W_: k_pair_mixed<T_, W_, COUL><<<g, BS, 0, s>>>(DMDX_PAIR_ARGS); break
    switch (tpa * 100 + w) {
        DMDX_PAIR_SHAPE(2, 8);
        DMDX_PAIR_SHAPE(4, 8);
        DMDX_PAIR_SHAPE(8, 8);
        DMDX_PAIR_SHAPE(4, 4);
        DMDX_PAIR_SHAPE(8, 2);
        DMDX_PAIR_SHAPE(8, 4);
        DMDX_PAIR_SHAPE(16, 2);
        DMDX_PAIR_SHAPE(16, 4);
        DMDX_PAIR_SHAPE(32, 2);
        DMDX_PAIR_SHAPE(32, 4);
    default: return hipErrorInvalidValue;
    }
#undef DMDX_PAIR_SHAPE
#undef DMDX_PAIR_ARGS
    return hipGetLastError();
}

hipError_t launch_pair(const Dev& d, hipStream_t s, int accumulate)
{
    const int BS = 256;
    if (d.precision == DMDX_FP64)
        k_pair_fp64<<<grid_for(blocks_for(d.nslots, BS)), BS, 0, s>>>(
            d.sys, d.st, d.slot_sys, d.nslots, d.order, d.nn, d.nbr, d.x, d.y, d.z, d.q, d.type,
            d.lj_sig, d.lj_eps, d.fx, d.fy, d.fz, d.p_lj, d.p_coul, accumulate, (int64_t)(blocks_for(d.nslots, BS)));
    else {
        // one launch per variant present in the batch; each system is taken
        // by exactly one of them
        if (d.any_coul_pairs) {
            const hipError_t e = launch_pair_mixed<true>(d, s, accumulate);
            if (e != hipSuccess) return e;
        }
        if (d.any_plain_pairs) {
            // LJ-only systems with a list on row tiles; those without a list
            // (no pair terms: zero forces) on the per-atom kernel
            if (d.pair_rows != 0) {
                const hipError_t e = launch_pair_rows(d, s, accumulate);
                if (e != hipSuccess || !d.any_plain_nolist) return e;
            }
            return launch_pair_mixed<false>(d, s, accumulate);
        }
    }
    return hipGetLastError();
}

}  // namespace dmdx
