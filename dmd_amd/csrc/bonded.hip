// bonded.hip -- harmonic bond, harmonic angle, periodic dihedral, harmonic
// improper and Coulomb-exclusion terms, FP64, one thread per term.
//
// A term does not touch the force arrays: it writes the vector each of its
// atoms receives into a term-force buffer (zero when the reference skips the
// term) and its energy into a per-wave partial.  k_gather then adds, per
// atom and in the reference's component order (bonds, angles, dihedrals,
// impropers, [PME], exclusions) and term order, every contribution that atom
// receives -- the same additions in the same order as the reference's
// sequential loops, with no atomics.
//
// Expressions follow, in order and with contraction off:
//   bond      src/force/harmonic_bond.cpp:23-45
//   angle     src/force/harmonic_angle.cpp:21-61 (CORRECTED flips the sign)
//   dihedral  src/force/periodic_dihedral.cpp:21-135
//   improper  src/force/harmonic_improper.cpp:21-128
//   exclusion src/force/coulomb_exclusion.cpp:29-54 (CORRECTED: erf form)
// CORRECTED torsion forces use the gradient of Bekker (1996) / Blondel &
// Karplus (1996) for the same angle phi.
#include "dmdx_internal.h"

namespace dmdx {

struct BondedArgs {
    const SysDev* sys;
    const SysState* st;
    const double *x, *y, *z, *q;
    TermF* tf;  // term force buffer
};

__device__ __forceinline__ void mi(const SysDev& S, double& dx, double& dy, double& dz)
{
    dx = min_image_dmd(dx, S.L[0], S.iminv[0]);
    dy = min_image_dmd(dy, S.L[1], S.iminv[1]);
    dz = min_image_dmd(dz, S.L[2], S.iminv[2]);
}

__device__ __forceinline__ void put(TermF* tf, int64_t k, double a, double b, double c)
{
    tf[k].x = a;
    tf[k].y = b;
    tf[k].z = c;
}

__device__ __forceinline__ void finish_energy(double e, double* part, int64_t u)
{
    e = wave_sum(e);
    if (lane_id() == 0) part[u >> 6] = e;
}

// ---------------------------------------------------------------- bonds
__device__ __forceinline__ void k_bonds_unit(int64_t unit_, BondedArgs a, const int* __restrict__ term_sys, int64_t nterm,
                        const int* __restrict__ bi, const int* __restrict__ bj,
                        const double* __restrict__ bk, const double* __restrict__ br0,
                        int64_t tf_off, double* __restrict__ part)
{
    int64_t u = unit_ * blockDim.x + threadIdx.x;
    if (u >= nterm) return;
    int s = term_sys[u];
    if (s < 0) return;
    if (a.st[s].reuse || !a.st[s].active) return;
    const SysDev& S = a.sys[s];
    double energy = 0.0;
    if (u - S.bond_off < S.n_bond) {
        int ai = bi[u], aj = bj[u];
        double dx = a.x[aj] - a.x[ai];
        double dy = a.y[aj] - a.y[ai];
        double dz = a.z[aj] - a.z[ai];
        mi(S, dx, dy, dz);
        double r2 = dx * dx + dy * dy + dz * dz;
        double r = sqrt(r2);
        double dr = r - br0[u];
        double f_scalar = -bk[u] * dr / r;
        double fx = f_scalar * dx, fy = f_scalar * dy, fz = f_scalar * dz;
        put(a.tf, tf_off + 2 * u, -fx, -fy, -fz);
        put(a.tf, tf_off + 2 * u + 1, fx, fy, fz);
        energy = 0.5 * bk[u] * dr * dr;
    }
    finish_energy(energy, part, u);
}

__global__ void k_bonds(BondedArgs a, const int* __restrict__ term_sys, int64_t nterm,
                        const int* __restrict__ bi, const int* __restrict__ bj,
                        const double* __restrict__ bk, const double* __restrict__ br0,
                        int64_t tf_off, double* __restrict__ part, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_) k_bonds_unit(u_, a, term_sys, nterm, bi, bj, bk, br0, tf_off, part);
}

// --------------------------------------------------------------- angles
__device__ __forceinline__ void k_angles_unit(int64_t unit_, BondedArgs a, const int* __restrict__ term_sys, int64_t nterm,
                         const int* __restrict__ ai_, const int* __restrict__ aj_,
                         const int* __restrict__ ak_, const double* __restrict__ kth,
                         const double* __restrict__ th0, int64_t tf_off,
                         double* __restrict__ part)
{
    int64_t u = unit_ * blockDim.x + threadIdx.x;
    if (u >= nterm) return;
    int s = term_sys[u];
    if (s < 0) return;
    if (a.st[s].reuse || !a.st[s].active) return;
    const SysDev& S = a.sys[s];
    double energy = 0.0;
    if (u - S.angle_off < S.n_angle) {
        int ai = ai_[u], aj = aj_[u], ak = ak_[u];
        double dx_ij = a.x[ai] - a.x[aj];
        double dy_ij = a.y[ai] - a.y[aj];
        double dz_ij = a.z[ai] - a.z[aj];
        mi(S, dx_ij, dy_ij, dz_ij);
        double dx_kj = a.x[ak] - a.x[aj];
        double dy_kj = a.y[ak] - a.y[aj];
        double dz_kj = a.z[ak] - a.z[aj];
        mi(S, dx_kj, dy_kj, dz_kj);
        double r_ij2 = dx_ij * dx_ij + dy_ij * dy_ij + dz_ij * dz_ij;
        double r_kj2 = dx_kj * dx_kj + dy_kj * dy_kj + dz_kj * dz_kj;
        double r_ij = sqrt(r_ij2);
        double r_kj = sqrt(r_kj2);
        double dot = dx_ij * dx_kj + dy_ij * dy_kj + dz_ij * dz_kj;
        double cos_theta = dot / (r_ij * r_kj);
        cos_theta = fmax(-1.0, fmin(1.0, cos_theta));
        double theta = acos(cos_theta);
        double dtheta = theta - th0[u];
        double kt = kth[u];
        double f_scalar = (S.physics == DMDX_FAITHFUL)
                              ? -kt * dtheta / sqrt(1.0 - cos_theta * cos_theta)
                              : kt * dtheta / sqrt(1.0 - cos_theta * cos_theta);
        double fx_ij = f_scalar * (dx_kj / (r_ij * r_kj) - cos_theta * dx_ij / r_ij2);
        double fy_ij = f_scalar * (dy_kj / (r_ij * r_kj) - cos_theta * dy_ij / r_ij2);
        double fz_ij = f_scalar * (dz_kj / (r_ij * r_kj) - cos_theta * dz_ij / r_ij2);
        double fx_kj = f_scalar * (dx_ij / (r_ij * r_kj) - cos_theta * dx_kj / r_kj2);
        double fy_kj = f_scalar * (dy_ij / (r_ij * r_kj) - cos_theta * dy_kj / r_kj2);
        double fz_kj = f_scalar * (dz_ij / (r_ij * r_kj) - cos_theta * dz_kj / r_kj2);
        put(a.tf, tf_off + 3 * u, fx_ij, fy_ij, fz_ij);
        put(a.tf, tf_off + 3 * u + 1, -(fx_ij + fx_kj), -(fy_ij + fy_kj), -(fz_ij + fz_kj));
        put(a.tf, tf_off + 3 * u + 2, fx_kj, fy_kj, fz_kj);
        energy = 0.5 * kt * dtheta * dtheta;
    }
    finish_energy(energy, part, u);
}

__global__ void k_angles(BondedArgs a, const int* __restrict__ term_sys, int64_t nterm,
                         const int* __restrict__ ai_, const int* __restrict__ aj_,
                         const int* __restrict__ ak_, const double* __restrict__ kth,
                         const double* __restrict__ th0, int64_t tf_off,
                         double* __restrict__ part, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_) k_angles_unit(u_, a, term_sys, nterm, ai_, aj_, ak_, kth, th0, tf_off, part);
}

// ------------------------------------------------------------- torsions
struct Torsion {
    double ba[3], bc[3], dc[3], n1[3], n2[3];
    double n1_sq, n2_sq, bc_sq, phi;
};

__device__ bool torsion_geom(const BondedArgs& a, const SysDev& S, int ai, int aj, int ak,
                             int al, Torsion& t)
{
    double ba_x = a.x[ai] - a.x[aj], ba_y = a.y[ai] - a.y[aj], ba_z = a.z[ai] - a.z[aj];
    mi(S, ba_x, ba_y, ba_z);
    double bc_x = a.x[ak] - a.x[aj], bc_y = a.y[ak] - a.y[aj], bc_z = a.z[ak] - a.z[aj];
    mi(S, bc_x, bc_y, bc_z);
    double dc_x = a.x[ak] - a.x[al], dc_y = a.y[ak] - a.y[al], dc_z = a.z[ak] - a.z[al];
    mi(S, dc_x, dc_y, dc_z);
    double n1_x = ba_y * bc_z - ba_z * bc_y;
    double n1_y = ba_z * bc_x - ba_x * bc_z;
    double n1_z = ba_x * bc_y - ba_y * bc_x;
    double n2_x = bc_y * dc_z - bc_z * dc_y;
    double n2_y = bc_z * dc_x - bc_x * dc_z;
    double n2_z = bc_x * dc_y - bc_y * dc_x;
    double n1_sq = n1_x * n1_x + n1_y * n1_y + n1_z * n1_z;
    double n2_sq = n2_x * n2_x + n2_y * n2_y + n2_z * n2_z;
    double bc_sq = bc_x * bc_x + bc_y * bc_y + bc_z * bc_z;
    t.ba[0] = ba_x; t.ba[1] = ba_y; t.ba[2] = ba_z;
    t.bc[0] = bc_x; t.bc[1] = bc_y; t.bc[2] = bc_z;
    t.dc[0] = dc_x; t.dc[1] = dc_y; t.dc[2] = dc_z;
    t.n1[0] = n1_x; t.n1[1] = n1_y; t.n1[2] = n1_z;
    t.n2[0] = n2_x; t.n2[1] = n2_y; t.n2[2] = n2_z;
    t.n1_sq = n1_sq; t.n2_sq = n2_sq; t.bc_sq = bc_sq;
    if (n1_sq < 1e-30 || n2_sq < 1e-30 || bc_sq < 1e-30) return false;
    double n1_norm = sqrt(n1_sq), n2_norm = sqrt(n2_sq), bc_norm = sqrt(bc_sq);
    double dot = n1_x * n2_x + n1_y * n2_y + n1_z * n2_z;
    double cos_phi = dot / (n1_norm * n2_norm);
    cos_phi = fmax(-1.0, fmin(1.0, cos_phi));
    double triple = bc_x * (n1_y * n2_z - n1_z * n2_y)
                  + bc_y * (n1_z * n2_x - n1_x * n2_z)
                  + bc_z * (n1_x * n2_y - n1_y * n2_x);
    double sin_phi = triple / (bc_norm * n1_norm * n2_norm);
    sin_phi = fmax(-1.0, fmin(1.0, sin_phi));
    t.phi = atan2(sin_phi, cos_phi);
    return true;
}

// FAITHFUL: F_m += pref * g_m with the reference's g vectors
// (src/force/periodic_dihedral.cpp:86-134)
__device__ void torsion_put_faithful(const Torsion& t, double pref, TermF* tf, int64_t base)
{
    const double bc_x = t.bc[0], bc_y = t.bc[1], bc_z = t.bc[2];
    const double ba_x = t.ba[0], ba_y = t.ba[1], ba_z = t.ba[2];
    const double dc_x = t.dc[0], dc_y = t.dc[1], dc_z = t.dc[2];
    const double n1_x = t.n1[0], n1_y = t.n1[1], n1_z = t.n1[2];
    const double n2_x = t.n2[0], n2_y = t.n2[1], n2_z = t.n2[2];
    const double n1_sq = t.n1_sq, n2_sq = t.n2_sq;
    double gi_x = (bc_y * n1_z - bc_z * n1_y) / n1_sq;
    double gi_y = (bc_z * n1_x - bc_x * n1_z) / n1_sq;
    double gi_z = (bc_x * n1_y - bc_y * n1_x) / n1_sq;
    double gl_x = -(bc_y * n2_z - bc_z * n2_y) / n2_sq;
    double gl_y = -(bc_z * n2_x - bc_x * n2_z) / n2_sq;
    double gl_z = -(bc_x * n2_y - bc_y * n2_x) / n2_sq;
    double n1xba_x = n1_y * ba_z - n1_z * ba_y;
    double n1xba_y = n1_z * ba_x - n1_x * ba_z;
    double n1xba_z = n1_x * ba_y - n1_y * ba_x;
    double dcxn2_x = dc_y * n2_z - dc_z * n2_y;
    double dcxn2_y = dc_z * n2_x - dc_x * n2_z;
    double dcxn2_z = dc_x * n2_y - dc_y * n2_x;
    double gj_x = n1xba_x / n1_sq + dcxn2_x / n2_sq - gi_x;
    double gj_y = n1xba_y / n1_sq + dcxn2_y / n2_sq - gi_y;
    double gj_z = n1xba_z / n1_sq + dcxn2_z / n2_sq - gi_z;
    double bxn2_x = (bc_y * n2_z - bc_z * n2_y) / n2_sq;
    double bxn2_y = (bc_z * n2_x - bc_x * n2_z) / n2_sq;
    double bxn2_z = (bc_x * n2_y - bc_y * n2_x) / n2_sq;
    double gk_x = bxn2_x - n1xba_x / n1_sq - dcxn2_x / n2_sq;
    double gk_y = bxn2_y - n1xba_y / n1_sq - dcxn2_y / n2_sq;
    double gk_z = bxn2_z - n1xba_z / n1_sq - dcxn2_z / n2_sq;
    put(tf, base + 0, pref * gi_x, pref * gi_y, pref * gi_z);
    put(tf, base + 1, pref * gj_x, pref * gj_y, pref * gj_z);
    put(tf, base + 2, pref * gk_x, pref * gk_y, pref * gk_z);
    put(tf, base + 3, pref * gl_x, pref * gl_y, pref * gl_z);
}

// CORRECTED: F = -(dE/dphi) grad(phi)
__device__ void torsion_put_corrected(const Torsion& t, double dEdphi, TermF* tf, int64_t base)
{
    double nrkj = sqrt(t.bc_sq);
    double a = -dEdphi * nrkj / t.n1_sq;
    double b = dEdphi * nrkj / t.n2_sq;
    double f_i[3], f_l[3], f_j[3], f_k[3];
    for (int d = 0; d < 3; ++d) {
        f_i[d] = a * t.n1[d];
        f_l[d] = b * t.n2[d];
    }
    double p = (t.ba[0] * t.bc[0] + t.ba[1] * t.bc[1] + t.ba[2] * t.bc[2]) / t.bc_sq;
    double qq = (t.dc[0] * t.bc[0] + t.dc[1] * t.bc[1] + t.dc[2] * t.bc[2]) / t.bc_sq;
    for (int d = 0; d < 3; ++d) {
        double sv = p * f_i[d] - qq * f_l[d];
        f_j[d] = f_i[d] - sv;
        f_k[d] = f_l[d] + sv;
    }
    put(tf, base + 0, f_i[0], f_i[1], f_i[2]);
    put(tf, base + 1, -f_j[0], -f_j[1], -f_j[2]);
    put(tf, base + 2, -f_k[0], -f_k[1], -f_k[2]);
    put(tf, base + 3, f_l[0], f_l[1], f_l[2]);
}

__device__ __forceinline__ void torsion_zero(TermF* tf, int64_t base)
{
    for (int r = 0; r < 4; ++r) put(tf, base + r, 0.0, 0.0, 0.0);
}

__device__ __forceinline__ void k_dihedrals_unit(int64_t unit_, BondedArgs a, const int* __restrict__ term_sys, int64_t nterm,
                            const int* __restrict__ di, const int* __restrict__ dj,
                            const int* __restrict__ dk, const int* __restrict__ dl,
                            const int* __restrict__ per, const double* __restrict__ kp,
                            const double* __restrict__ p0, int64_t tf_off,
                            double* __restrict__ part)
{
    int64_t u = unit_ * blockDim.x + threadIdx.x;
    if (u >= nterm) return;
    int s = term_sys[u];
    if (s < 0) return;
    if (a.st[s].reuse || !a.st[s].active) return;
    const SysDev& S = a.sys[s];
    double energy = 0.0;
    if (u - S.dih_off < S.n_dih) {
        Torsion t;
        int64_t base = tf_off + 4 * u;
        if (!torsion_geom(a, S, di[u], dj[u], dk[u], dl[u], t)) {
            torsion_zero(a.tf, base);
        } else {
            int p = per[u];
            double k = kp[u];
            double phi0 = p0[u];
            double cos_np = cos(p * t.phi - phi0);
            energy = k * (1.0 + cos_np);
            double pref = k * p * sin(p * t.phi - phi0);
            if (fabs(pref) < 1e-30)
                torsion_zero(a.tf, base);
            else if (S.physics == DMDX_FAITHFUL)
                torsion_put_faithful(t, pref, a.tf, base);
            else
                torsion_put_corrected(t, -pref, a.tf, base);
        }
    }
    finish_energy(energy, part, u);
}

__global__ void k_dihedrals(BondedArgs a, const int* __restrict__ term_sys, int64_t nterm,
                            const int* __restrict__ di, const int* __restrict__ dj,
                            const int* __restrict__ dk, const int* __restrict__ dl,
                            const int* __restrict__ per, const double* __restrict__ kp,
                            const double* __restrict__ p0, int64_t tf_off,
                            double* __restrict__ part, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_) k_dihedrals_unit(u_, a, term_sys, nterm, di, dj, dk, dl, per, kp, p0, tf_off, part);
}

__device__ __forceinline__ void k_impropers_unit(int64_t unit_, BondedArgs a, const int* __restrict__ term_sys, int64_t nterm,
                            const int* __restrict__ di, const int* __restrict__ dj,
                            const int* __restrict__ dk, const int* __restrict__ dl,
                            const double* __restrict__ kp, const double* __restrict__ p0,
                            int64_t tf_off, double* __restrict__ part)
{
    int64_t u = unit_ * blockDim.x + threadIdx.x;
    if (u >= nterm) return;
    int s = term_sys[u];
    if (s < 0) return;
    if (a.st[s].reuse || !a.st[s].active) return;
    const SysDev& S = a.sys[s];
    double energy = 0.0;
    if (u - S.imp_off < S.n_imp) {
        Torsion t;
        int64_t base = tf_off + 4 * u;
        if (!torsion_geom(a, S, di[u], dj[u], dk[u], dl[u], t)) {
            torsion_zero(a.tf, base);
        } else {
            double dchi = t.phi - p0[u];
            double k = kp[u];
            energy = 0.5 * k * dchi * dchi;
            double pref = -k * dchi;
            if (fabs(pref) < 1e-30)
                torsion_zero(a.tf, base);
            else if (S.physics == DMDX_FAITHFUL)
                torsion_put_faithful(t, pref, a.tf, base);
            else
                torsion_put_corrected(t, -pref, a.tf, base);
        }
    }
    finish_energy(energy, part, u);
}

__global__ void k_impropers(BondedArgs a, const int* __restrict__ term_sys, int64_t nterm,
                            const int* __restrict__ di, const int* __restrict__ dj,
                            const int* __restrict__ dk, const int* __restrict__ dl,
                            const double* __restrict__ kp, const double* __restrict__ p0,
                            int64_t tf_off, double* __restrict__ part, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_) k_impropers_unit(u_, a, term_sys, nterm, di, dj, dk, dl, kp, p0, tf_off, part);
}

// ------------------------------------------------------------ exclusions
// The reference subtracts the whole term sum from the energy
// (energy_out -= energy, src/force/coulomb_exclusion.cpp:55): the partials
// hold +sum and the finalize subtracts.
__device__ __forceinline__ void k_exclusions_unit(int64_t unit_, BondedArgs a, const int* __restrict__ term_sys, int64_t nterm,
                             const int* __restrict__ ei, const int* __restrict__ ej,
                             int64_t tf_off, double* __restrict__ part)
{
    int64_t u = unit_ * blockDim.x + threadIdx.x;
    if (u >= nterm) return;
    int s = term_sys[u];
    if (s < 0) return;
    if (a.st[s].reuse || !a.st[s].active) return;
    const SysDev& S = a.sys[s];
    double energy = 0.0;
    if (u - S.ex_off < S.n_ex) {
        int i = ei[u], j = ej[u];
        double dx = a.x[j] - a.x[i];
        double dy = a.y[j] - a.y[i];
        double dz = a.z[j] - a.z[i];
        mi(S, dx, dy, dz);
        double r2 = dx * dx + dy * dy + dz * dz;
        double r = sqrt(r2);
        double qiqj = a.q[i] * a.q[j];
        double fs;
        if (S.physics == DMDX_FAITHFUL) {
            fs = KE * qiqj / (r2 * r);
            energy = KE * qiqj / r;
        } else {
            double alpha = S.alpha;
            double erf_ar = erf(alpha * r);
            fs = KE * qiqj * (2.0 * alpha / sqrt(PI) * exp(-alpha * alpha * r2) / r2 -
                              erf_ar / (r2 * r));
            energy = KE * qiqj * erf_ar / r;
        }
        double gx = fs * dx, gy = fs * dy, gz = fs * dz;
        put(a.tf, tf_off + 2 * u, -gx, -gy, -gz);
        put(a.tf, tf_off + 2 * u + 1, gx, gy, gz);
    }
    finish_energy(energy, part, u);
}

__global__ void k_exclusions(BondedArgs a, const int* __restrict__ term_sys, int64_t nterm,
                             const int* __restrict__ ei, const int* __restrict__ ej,
                             int64_t tf_off, double* __restrict__ part, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_) k_exclusions_unit(u_, a, term_sys, nterm, ei, ej, tf_off, part);
}

}  // namespace dmdx

namespace dmdx {

hipError_t launch_bonded(const Dev& d, hipStream_t s)
{
    const int BS = 256;
    BondedArgs a{d.sys, d.st, d.x, d.y, d.z, d.q, d.tf};
    if (d.nb)
        k_bonds<<<grid_for(blocks_for(d.nb, BS)), BS, 0, s>>>(a, d.b_sys, d.nb, d.bi, d.bj, d.bk, d.br0,
                                                    d.tf_bond, d.p_bond, (int64_t)(blocks_for(d.nb, BS)));
    if (d.na)
        k_angles<<<grid_for(blocks_for(d.na, BS)), BS, 0, s>>>(a, d.a_sys, d.na, d.ai, d.aj, d.ak, d.akt,
                                                     d.at0, d.tf_angle, d.p_angle, (int64_t)(blocks_for(d.na, BS)));
    if (d.nd)
        k_dihedrals<<<grid_for(blocks_for(d.nd, BS)), BS, 0, s>>>(a, d.d_sys, d.nd, d.di, d.dj, d.dk, d.dl,
                                                        d.dper, d.dkp, d.dp0, d.tf_dih, d.p_dih, (int64_t)(blocks_for(d.nd, BS)));
    if (d.ni)
        k_impropers<<<grid_for(blocks_for(d.ni, BS)), BS, 0, s>>>(a, d.i_sys, d.ni, d.ii, d.ij, d.ik, d.il,
                                                        d.ikp, d.ip0, d.tf_imp, d.p_imp, (int64_t)(blocks_for(d.ni, BS)));
    if (d.ne)
        k_exclusions<<<grid_for(blocks_for(d.ne, BS)), BS, 0, s>>>(a, d.e_sys, d.ne, d.ei, d.ej, d.tf_ex,
                                                         d.p_excl, (int64_t)(blocks_for(d.ne, BS)));
    return hipGetLastError();
}

}  // namespace dmdx
