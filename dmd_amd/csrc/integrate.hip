// integrate.hip -- per-atom force gather, velocity-Verlet kick/drift,
// deterministic per-system reductions, thermostats and barostats (FP64).
//
//   k_pre_compute   per system: list-rebuild decision and counter
//                   (src/force/lennard_jones.cpp:53-55, :91), force reuse
//   k_gather        per atom: bonded contributions, PME interpolation
//                   (src/force/coulomb_pme.cpp:205-283), exclusion
//                   contributions -- in the reference's component order
//   k_kick_drift    v = lambda v + F/m dt/2 ; x += v dt
//                   (src/integrate/integrator.cpp:4-20; lambda is the
//                   thermostat scale of the previous step, applied here)
//   k_kick_ke       v += F/m dt/2 and per-wave kinetic-energy partials
//                   (src/thermostat/thermostat.cpp:6-15)
//   k_finalize      per system: energies, temperature, Berendsen
//                   (src/thermostat/berendsen.cpp:13-28) or Nose-Hoover
//                   (src/thermostat/nose_hoover.cpp:18-32) scale
//   k_baro_*        virial, pressure and Berendsen / Andersen scaling
//                   (src/barostat/*.cpp), only launched with a barostat
#include "dmdx_internal.h"

namespace dmdx {

template <class R>
struct Cx2 {
    R re, im;
};

__device__ __forceinline__ bool pre_retires(const SysState& T);
__device__ void pre_decide(const SysDev& S, SysState& T, int phase, int force_rebuild,
                           int reuse_enabled, double dmax);

// phase: 0 standalone compute, 1 first compute of a step, 2 second compute;
// force_rebuild: 1 = rebuild_pair_list requested by the host.
// Rebuild rule (SysDev::interval): > 0 DMD's cadence (counter reaches the
// interval, src/force/lennard_jones.cpp:53-55); 0 the Verlet criterion: some
// atom moved more than skin/2 since the last build (checked whenever the
// positions may have changed since the last compute).
__global__ void __launch_bounds__(64) k_pre_compute(const SysDev* __restrict__ sys,
                                                    SysState* __restrict__ st, int nsys,
                                                    int phase, int force_rebuild,
                                                    int reuse_enabled,
                                                    const double* __restrict__ p_disp)
{
    // one wave per system; lane 0 decides
    int s = blockIdx.x;
    if (s >= nsys) return;
    SysState& T = st[s];
    const SysDev& S0 = sys[s];
    double dmax = 0.0;
    if (p_disp && S0.has_list && S0.interval == 0 && T.active && T.built && !(phase == 1 && T.valid))
        for (int w = lane_id(); w < S0.aw_n; w += 64) dmax = fmax(dmax, p_disp[S0.aw_off + w]);
    for (int o = 32; o > 0; o >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, o, 64));
    if (lane_id() != 0) return;
    pre_decide(sys[s], T, phase, force_rebuild, reuse_enabled, dmax);
}

__device__ __forceinline__ bool pre_retires(const SysState& T)
{
    return T.flags != 0 || (T.stop_at > 0 && T.steps_done >= T.stop_at);
}

// One system's decisions before a compute (k_pre_compute's lane 0, or the
// system's first wave of k_kick_drift when that compute only reuses forces).
__device__ void pre_decide(const SysDev& S, SysState& T, int phase, int force_rebuild,
                           int reuse_enabled, double dmax)
{
    // a system leaves the batch before the first compute of a step once it
    // has reached its step target or has failed (neighbour overflow,
    // non-finite energy): its state stays that of the end of the last step,
    // and the other systems of the batch are unaffected
    if (phase == 1 && T.active && pre_retires(T)) T.active = 0;
    if (!T.active) {
        T.rebuild = 0;
        T.reuse = 1;
        return;
    }
    int rebuild = 0;
    if (S.has_list) {
        if (force_rebuild) {
            rebuild = 1;
        } else if (S.interval > 0) {
            rebuild = T.counter >= S.interval;
        } else if (S.interval == 0) {
            if (!T.built) {
                rebuild = 1;
            } else if (!(phase == 1 && T.valid)) {
                rebuild = dmax > S.half_skin2;
            }
        }
    }
    T.rebuild = rebuild;
    if (rebuild) {
        T.built = 1;
        T.rebuilds += 1;
    }
    T.reuse = (phase == 1 && reuse_enabled && T.valid && !rebuild) ? 1 : 0;
    if (S.has_list && S.interval >= 0) T.counter = (rebuild ? 0 : T.counter) + 1;
    T.valid = 1;
}

// explicit rebuild_pair_list (src/force/lennard_jones.cpp:18-41): builds the
// list and leaves the counter at 0, without computing forces
__global__ void k_pre_rebuild_only(const SysDev* __restrict__ sys, SysState* __restrict__ st,
                                   int nsys)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    st[s].rebuild = sys[s].has_list ? 1 : 0;
    st[s].reuse = 1;  // no force kernels
    if (sys[s].has_list) {
        st[s].counter = 0;
        st[s].built = 1;
    }
}

// F += tf[contrib[c]] for c = c0, c0 + 1, ..., c1 - 1, in that order: eight
// indices, then their eight term forces, are loaded before any is added
__device__ __forceinline__ void add_terms(const int* __restrict__ contrib,
                                          const TermF* __restrict__ tf, int c0, int c1,
                                          double& Fx, double& Fy, double& Fz)
{
    for (int c = c0; c < c1; c += 8) {
        const int n = c1 - c;
        int idx[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) idx[u] = u < n ? contrib[c + u] : 0;
        TermF v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (u < n) v[u] = tf[idx[u]];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (u < n) {
                Fx += v[u].x;
                Fy += v[u].y;
                Fz += v[u].z;
            }
    }
}

// the system has bonded, exclusion or PME forces to gather
__device__ __forceinline__ bool system_gathers(const SysDev& S)
{
    return (S.bw_n | S.gw_n | S.dw_n | S.iw_n | S.xw_n | S.pme) != 0;
}

// Per-atom gather.  cstart[g] .. cmid[g]: bonded contributions (bonds,
// angles, dihedrals, impropers in term order); cmid[g] .. cstart[g+1]:
// exclusion contributions.  PME sits between them, as in build_force_engine
// (src/sim/simulation_engine.cpp:147-203).  TERMS = false: added to the
// pair forces in fx, fy, fz.  TERMS = true (mixed path, on the side stream
// while the pair kernel runs): their sum from zero is written to fx, fy, fz
// (the term-force arrays) for every atom of a system that has such terms;
// the closing kick adds it to the pair forces.
template <class R, bool TERMS>
__device__ __forceinline__ void k_gather_unit(int64_t unit_, const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                         const int* __restrict__ slot_sys, int64_t nslots,
                         const int* __restrict__ cstart, const int* __restrict__ cmid,
                         const int* __restrict__ contrib, const TermF* __restrict__ tf,
                         const double* __restrict__ x, const double* __restrict__ y,
                         const double* __restrict__ z, const double* __restrict__ q,
                         const Cx2<R>* __restrict__ grid, double* __restrict__ fx,
                         double* __restrict__ fy, double* __restrict__ fz)
{
    int64_t g = unit_ * blockDim.x + threadIdx.x;
    if (g >= nslots) return;
    int s = slot_sys[g];
    if (s < 0 || st[s].reuse || !st[s].active) return;
    const SysDev& S = sys[s];
    if (g - S.atom_off >= S.n) return;
    if (TERMS && !system_gathers(S)) return;
    int c0 = cstart[g], c1 = cmid[g], c2 = cstart[g + 1];
    double qi = S.pme ? q[g] : 0.0;
    if (!TERMS && c0 == c2 && qi == 0.0) return;
    double Fx = TERMS ? 0.0 : fx[g], Fy = TERMS ? 0.0 : fy[g], Fz = TERMS ? 0.0 : fz[g];
    add_terms(contrib, tf, c0, c1, Fx, Fy, Fz);
    if (qi != 0.0) {
        const int nx = S.K[0], ny = S.K[1], nz = S.K[2];
        const double Lx = S.L[0], Ly = S.L[1], Lz = S.L[2];
        double frac_x = x[g] / Lx, frac_y = y[g] / Ly, frac_z = z[g] / Lz;
        frac_x -= floor(frac_x);
        frac_y -= floor(frac_y);
        frac_z -= floor(frac_z);
        // non-finite coordinates (a failed system) must still index the grid
        frac_x = (frac_x >= 0.0 && frac_x <= 1.0) ? frac_x : 0.0;
        frac_y = (frac_y >= 0.0 && frac_y <= 1.0) ? frac_y : 0.0;
        frac_z = (frac_z >= 0.0 && frac_z <= 1.0) ? frac_z : 0.0;
        double ux = frac_x * nx, uy = frac_y * ny, uz = frac_z * nz;
        int m0x = (int)floor(ux) - 1, m0y = (int)floor(uy) - 1, m0z = (int)floor(uz) - 1;
        R wx[4], wy[4], wz[4], dwx[4], dwy[4], dwz[4];
        for (int d = 0; d < 4; ++d) {
            R tx = (R)(ux - (m0x + d)), ty = (R)(uy - (m0y + d)), tz = (R)(uz - (m0z + d));
            R a;
            // theta4 / dtheta4, src/force/coulomb_pme.cpp:50-69
            a = fabs(tx);
            wx[d] = a >= R(2) ? R(0) : (a >= R(1) ? (R(2) - a) * (R(2) - a) * (R(2) - a) / R(6)
                                                 : (R(4) - R(6) * a * a + R(3) * a * a * a) / R(6));
            dwx[d] = a >= R(2) ? R(0)
                               : (a >= R(1) ? -(tx >= R(0) ? R(1) : R(-1)) * (R(2) - a) * (R(2) - a) / R(2)
                                            : (tx >= R(0) ? R(1) : R(-1)) * (R(9) * a * a - R(12) * a) / R(6));
            a = fabs(ty);
            wy[d] = a >= R(2) ? R(0) : (a >= R(1) ? (R(2) - a) * (R(2) - a) * (R(2) - a) / R(6)
                                                 : (R(4) - R(6) * a * a + R(3) * a * a * a) / R(6));
            dwy[d] = a >= R(2) ? R(0)
                               : (a >= R(1) ? -(ty >= R(0) ? R(1) : R(-1)) * (R(2) - a) * (R(2) - a) / R(2)
                                            : (ty >= R(0) ? R(1) : R(-1)) * (R(9) * a * a - R(12) * a) / R(6));
            a = fabs(tz);
            wz[d] = a >= R(2) ? R(0) : (a >= R(1) ? (R(2) - a) * (R(2) - a) * (R(2) - a) / R(6)
                                                 : (R(4) - R(6) * a * a + R(3) * a * a * a) / R(6));
            dwz[d] = a >= R(2) ? R(0)
                               : (a >= R(1) ? -(tz >= R(0) ? R(1) : R(-1)) * (R(2) - a) * (R(2) - a) / R(2)
                                            : (tz >= R(0) ? R(1) : R(-1)) * (R(9) * a * a - R(12) * a) / R(6));
        }
        // FAITHFUL keeps the 1/(nx ny nz) scaling of the potential grid
        // (src/force/coulomb_pme.cpp:303-306)
        const bool norm = S.physics == DMDX_FAITHFUL;
        const R nrm = (R)(1.0 / ((double)nx * ny * nz));
        const Cx2<R>* G = grid + S.grid_off;
        R sx = R(0), sy = R(0), sz = R(0);
        for (int iz = 0; iz < 4; ++iz) {
            int gz = (m0z + iz) % nz;
            if (gz < 0) gz += nz;
            R w_z = wz[iz], dw_z = dwz[iz];
            for (int iy = 0; iy < 4; ++iy) {
                int gy = (m0y + iy) % ny;
                if (gy < 0) gy += ny;
                R w_y = wy[iy], dw_y = dwy[iy];
                int64_t row = ((int64_t)gz * ny + gy) * nx;
                for (int ix = 0; ix < 4; ++ix) {
                    int gx = (m0x + ix) % nx;
                    if (gx < 0) gx += nx;
                    R phi = G[row + gx].re;
                    if (norm) phi *= nrm;
                    sx += phi * dwx[ix] * w_y * w_z;
                    sy += phi * wx[ix] * dw_y * w_z;
                    sz += phi * wx[ix] * w_y * dw_z;
                }
            }
        }
        double fxf = (double)nx / Lx, fyf = (double)ny / Ly, fzf = (double)nz / Lz;
        Fx -= qi * fxf * (double)sx;
        Fy -= qi * fyf * (double)sy;
        Fz -= qi * fzf * (double)sz;
    }
    add_terms(contrib, tf, c1, c2, Fx, Fy, Fz);
    fx[g] = Fx;
    fy[g] = Fy;
    fz[g] = Fz;
}

template <class R, bool TERMS>
__global__ void k_gather(const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                         const int* __restrict__ slot_sys, int64_t nslots,
                         const int* __restrict__ cstart, const int* __restrict__ cmid,
                         const int* __restrict__ contrib, const TermF* __restrict__ tf,
                         const double* __restrict__ x, const double* __restrict__ y,
                         const double* __restrict__ z, const double* __restrict__ q,
                         const Cx2<R>* __restrict__ grid, double* __restrict__ fx,
                         double* __restrict__ fy, double* __restrict__ fz, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_) k_gather_unit<R, TERMS>(u_, sys, st, slot_sys, nslots, cstart, cmid, contrib, tf, x, y, z, q, grid, fx, fy, fz);
}


// Mixed path, besides the kick and drift: the cell-local FP32 frame of each
// moved atom (lpos[t] = x - org[t], t its cell-order slot), so the compute
// that follows needs no separate packing pass, and the squared displacement
// since the last list build (per-wave maximum, the Verlet criterion) -- both
// only where SHAKE does not move the atoms again before that compute.
struct DriftOut {
    const double4* org;
    float4* lpos;
    const int* slot_of;
    const double4* xref;
    double* p_disp;
};

// pre (1, or 2 with force reuse enabled): the step's first compute reuses
// every system's forces (the host knows it from its mirror of the
// decisions), so k_pre_compute's phase-1 decisions are taken here instead of
// in a launch of their own: by lane 0 of the system's first wave, while
// every wave reads the retirement condition itself (a system that retires
// does not move).  The other waves read only fields those decisions leave
// alone or set to the value the wave derives itself (active).
__device__ __forceinline__ void k_kick_drift_unit(int64_t unit_, const SysDev* __restrict__ sys, SysState* __restrict__ st,
                             const int* __restrict__ slot_sys, int64_t nslots,
                             const double* __restrict__ invm, const double* __restrict__ fx,
                             const double* __restrict__ fy, const double* __restrict__ fz,
                             double* __restrict__ vx, double* __restrict__ vy,
                             double* __restrict__ vz, double* __restrict__ x,
                             double* __restrict__ y, double* __restrict__ z, DriftOut o, int pre)
{
    const int64_t g = unit_ * blockDim.x + threadIdx.x;
    if (g >= nslots) return;  // whole waves: nslots is a multiple of 64
    const int s = __builtin_amdgcn_readfirstlane(slot_sys[g]);  // one system per wave
    if (s < 0) return;
    const SysDev& S = sys[s];
    if (pre) {
        SysState& T = st[s];
        const bool moves = T.active && !pre_retires(T);
        if (g == S.atom_off) pre_decide(S, T, 1, 0, pre == 2 ? 1 : 0, 0.0);
        if (!moves) return;
    } else if (!st[s].active) {
        return;
    }
    const bool real = g - S.atom_off < S.n;
    double d2 = 0.0;
    if (real) {
        const double lam = st[s].lambda;
        const double dt = S.dt;
        const double half_dt = 0.5 * dt;
        const double im = invm[g];
        double a = lam * vx[g], b = lam * vy[g], c = lam * vz[g];
        a += fx[g] * im * half_dt;
        b += fy[g] * im * half_dt;
        c += fz[g] * im * half_dt;
        vx[g] = a;
        vy[g] = b;
        vz[g] = c;
        const double nx = x[g] + a * dt, ny = y[g] + b * dt, nz = z[g] + c * dt;
        x[g] = nx;
        y[g] = ny;
        z[g] = nz;
        if (o.lpos && S.n_con == 0 && st[s].built) {
            const int64_t t = S.atom_off + o.slot_of[g];
            const double4 f = o.org[t];
            o.lpos[t] = make_float4((float)(nx - f.x), (float)(ny - f.y), (float)(nz - f.z),
                                    (float)f.w);
            if (S.interval == 0) {
                const double4 r = o.xref[g];
                const double ex = nx - r.x, ey = ny - r.y, ez = nz - r.z;
                d2 = ex * ex + ey * ey + ez * ez;
            }
        }
    }
    if (o.lpos && S.n_con == 0) {
        for (int k = 32; k > 0; k >>= 1) d2 = fmax(d2, __shfl_xor(d2, k, 64));
        if ((threadIdx.x & 63) == 0) o.p_disp[g >> 6] = d2;
    }
}

__global__ void k_kick_drift(const SysDev* __restrict__ sys, SysState* __restrict__ st,
                             const int* __restrict__ slot_sys, int64_t nslots,
                             const double* __restrict__ invm, const double* __restrict__ fx,
                             const double* __restrict__ fy, const double* __restrict__ fz,
                             double* __restrict__ vx, double* __restrict__ vy,
                             double* __restrict__ vz, double* __restrict__ x,
                             double* __restrict__ y, double* __restrict__ z, DriftOut o, int pre,
                             int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_) k_kick_drift_unit(u_, sys, st, slot_sys, nslots, invm, fx, fy, fz, vx, vy, vz, x, y, z, o, pre);
}

// Closing kick.  Mixed path with term forces gathered beside the pair
// kernel (fbx non-null): a system with such terms first completes its forces,
// F = pair force + term forces, and stores them for the next step's kick.
__device__ __forceinline__ void k_kick_ke_unit(int64_t unit_, const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                          const int* __restrict__ slot_sys, int64_t nslots,
                          const double* __restrict__ invm, const double* __restrict__ m,
                          double* __restrict__ fx, double* __restrict__ fy,
                          double* __restrict__ fz, const double* __restrict__ fbx,
                          const double* __restrict__ fby, const double* __restrict__ fbz,
                          double* __restrict__ vx, double* __restrict__ vy,
                          double* __restrict__ vz, double* __restrict__ ke_part)
{
    int64_t g = unit_ * blockDim.x + threadIdx.x;
    if (g >= nslots) return;
    int s = slot_sys[g];
    if (s < 0 || !st[s].active) return;
    const SysDev& S = sys[s];
    double ke = 0.0;
    if (g - S.atom_off < S.n) {
        const double half_dt = 0.5 * S.dt;
        const double im = invm[g];
        double Fx = fx[g], Fy = fy[g], Fz = fz[g];
        if (fbx && system_gathers(S)) {
            Fx += fbx[g];
            Fy += fby[g];
            Fz += fbz[g];
            fx[g] = Fx;
            fy[g] = Fy;
            fz[g] = Fz;
        }
        double a = vx[g] + Fx * im * half_dt;
        double b = vy[g] + Fy * im * half_dt;
        double c = vz[g] + Fz * im * half_dt;
        vx[g] = a;
        vy[g] = b;
        vz[g] = c;
        ke = 0.5 * m[g] * (a * a + b * b + c * c);
    }
    ke = wave_sum(ke);
    if (lane_id() == 0) ke_part[g >> 6] = ke;
}

__global__ void k_kick_ke(const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                          const int* __restrict__ slot_sys, int64_t nslots,
                          const double* __restrict__ invm, const double* __restrict__ m,
                          double* __restrict__ fx, double* __restrict__ fy,
                          double* __restrict__ fz, const double* __restrict__ fbx,
                          const double* __restrict__ fby, const double* __restrict__ fbz,
                          double* __restrict__ vx, double* __restrict__ vy,
                          double* __restrict__ vz, double* __restrict__ ke_part, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_) k_kick_ke_unit(u_, sys, st, slot_sys, nslots, invm, m, fx, fy, fz, fbx, fby, fbz, vx, vy, vz, ke_part);
}

__device__ __forceinline__ double strided_sum_from(const double* __restrict__ part, int64_t off,
                                                   int64_t n, double a, int64_t k);

// strided partial sum of part[off .. off+n) for a 256-thread workgroup:
// thread t sums t, t+256, ... in order
__device__ __forceinline__ double strided_sum(const double* __restrict__ part, int64_t off,
                                              int64_t n)
{
    // eight loads in flight before the in-order adds (the same sum, bit for
    // bit, as the plain loop)
    return strided_sum_from(part, off, n, 0.0, threadIdx.x);
}

// the same from element k on, starting from a
__device__ __forceinline__ double strided_sum_from(const double* __restrict__ part, int64_t off,
                                                   int64_t n, double a, int64_t k)
{
    for (; k + 7 * 256 < n; k += 8 * 256) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = part[off + k + u * 256];
#pragma unroll
        for (int u = 0; u < 8; ++u) a += v[u];
    }
    for (; k < n; k += 256) a += part[off + k];
    return a;
}

// sum of part[off .. off+n) by a 256-thread workgroup: the strided sums,
// then a fixed tree over the workgroup -- deterministic; every thread gets
// the result
__device__ double block_range_sum(const double* part, int64_t off, int64_t n, double* red)
{
    double a = wave_sum(strided_sum(part, off, n));
    __syncthreads();
    if (lane_id() == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    double r = (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
    return r;
}

struct Parts {
    const double *lj, *coul, *bond, *angle, *dih, *imp, *pme, *excl, *ke;
    // atom slots per pair-energy partial (FP64: 1; mixed: one per wave) of
    // the LJ-only systems and of the systems with Coulomb pairs (the two
    // variants of the mixed pair kernel have their own launch shapes); a
    // system's partials start at its first slot
    int pair_div, coul_div;
    // row-tile pair kernel on: LJ-only systems with a list have their LJ
    // energy as one partial per atom slot here (their Coulomb partials stay
    // zero in coul)
    const double* rows;
};

__device__ __forceinline__ double temperature_of(double ekin, int n)
{
    if (n < 2) return 0.0;
    double dof = 3.0 * (double)n - 3.0;
    return 2.0 * ekin / (dof * KB);
}

// a + part[off + k] + part[off + k + BS] + ... (elements below n), in that
// order, loading eight at a time (guarded) before adding them
template <int BS>
__device__ __forceinline__ double strided_sum_batched(const double* __restrict__ part,
                                                      int64_t off, int64_t n, double a, int64_t k)
{
    for (; k < n; k += 8 * BS) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = k + u * BS < n ? part[off + k + u * BS] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (k + u * BS < n) a += v[u];
    }
    return a;
}

// one FIN_THREADS-thread workgroup per system, after the closing kick
constexpr int FIN_THREADS = 512;
constexpr int FIN_WAVES = FIN_THREADS / 64;

__global__ void __launch_bounds__(FIN_THREADS) k_finalize(const SysDev* __restrict__ sys,
                                                          SysState* __restrict__ st, Parts P,
                                                          int with_kinetic)
{
    __shared__ double red[9][FIN_WAVES];
    int s = blockIdx.x;
    SysState& T = st[s];
    if (!T.active) return;
    const SysDev& S = sys[s];
    // the nine sums (8 energy terms and the kinetic energy) in one pass: each
    // thread's strided sums, then a fixed tree over each wave and a fixed
    // pairwise tree over the FIN_WAVES waves -- deterministic, one barrier
    // pair energies are per atom slot, the others per wave or workgroup
    const bool rows = P.rows && S.has_list && !S.coul_pairs;  // uniform per workgroup
    const double* arr[9] = {rows ? P.rows : P.lj, P.coul, P.bond, P.angle, P.dih, P.imp,
                            P.pme, P.excl, P.ke};
    const int div = S.coul_pairs ? P.coul_div : P.pair_div;
    const int64_t po = S.atom_off, pn = S.n_pad / div;
    const int64_t off[9] = {rows ? S.atom_off : po, po, S.bw_off, S.gw_off, S.dw_off,
                            S.iw_off, S.pw_off, S.xw_off, S.aw_off};
    const int64_t cnt[9] = {rows ? S.n_pad : pn, rows ? 0 : pn, S.bw_n, S.gw_n, S.dw_n,
                            S.iw_n, S.pw_n, S.xw_n, with_kinetic ? S.aw_n : 0};
    // the first element of each of the nine sums is loaded before any is
    // added (nine loads in flight), then the strided remainders
    double a[9];
    const int64_t t0 = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 9; ++k) a[k] = (t0 < cnt[k]) ? arr[k][off[k] + t0] : 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k)
        a[k] = strided_sum_batched<FIN_THREADS>(arr[k], off[k], cnt[k], 0.0 + a[k],
                                                t0 + FIN_THREADS);
#pragma unroll
    for (int k = 0; k < 9; ++k) a[k] = wave_sum(a[k]);
    if (lane_id() == 0)
#pragma unroll
        for (int k = 0; k < 9; ++k) red[k][threadIdx.x >> 6] = a[k];
    __syncthreads();
    if (threadIdx.x != 0) return;
    double r[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        double w[FIN_WAVES];
#pragma unroll
        for (int i = 0; i < FIN_WAVES; ++i) w[i] = red[k][i];
#pragma unroll
        for (int h = FIN_WAVES / 2; h > 0; h /= 2)
#pragma unroll
            for (int i = 0; i < h; ++i) w[i] = w[2 * i] + w[2 * i + 1];
        r[k] = w[0];
    }
    double e[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) e[k] = r[k];
    double ke = r[8];
    if (S.pme) {
        e[6] += S.self_energy;
        if (S.physics != DMDX_FAITHFUL && S.qsum != 0.0) {
            double V = sys_volume(S);
            e[6] += -PI * KE * S.qsum * S.qsum / (2.0 * V * S.alpha * S.alpha);
        }
    }
    // the pair terms, bonded terms and PME add, the exclusion term subtracts
    double pe = 0.0;
    for (int k = 0; k < 7; ++k) pe += e[k];
    pe -= e[7];
    for (int k = 0; k < 8; ++k) T.pe_terms[k] = e[k];
    T.pe = pe;
    if (!with_kinetic) {
        if (!isfinite(pe)) T.flags |= FLAG_NONFINITE;
        return;
    }
    T.ke = ke;
    double lam = 1.0;
    const int n = S.n;
    if (S.thermostat == DMDX_THERMO_BERENDSEN) {
        if (S.ttau > 0.0 && ke > 0.0) {
            double T_inst = temperature_of(ke, n);
            double l = sqrt(1.0 + S.dt / S.ttau * (S.T0 / T_inst - 1.0));
            if (!(l <= 0.0)) lam = l;
        }
    } else if (S.thermostat == DMDX_THERMO_NOSE_HOOVER) {
        double T_inst = temperature_of(ke, n);
        T.zeta += 0.5 * S.dt * KB * (T_inst - S.T0) / S.Qnh;
        lam = exp(-T.zeta * 0.5 * S.dt);
    }
    T.lambda = lam;
    T.ke_out = ke * lam * lam;
    T.steps_done += 1;
    if (!isfinite(pe) || !isfinite(ke) || !isfinite(lam)) T.flags |= FLAG_NONFINITE;
}

// ---- barostat path: materialise lambda, kinetic energy and virial
__device__ __forceinline__ void k_baro_ke_virial_unit(int64_t unit_, const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                                 const int* __restrict__ slot_sys, int64_t nslots,
                                 const double* __restrict__ m, double* __restrict__ vx,
                                 double* __restrict__ vy, double* __restrict__ vz,
                                 const double* __restrict__ x, const double* __restrict__ y,
                                 const double* __restrict__ z, const double* __restrict__ fx,
                                 const double* __restrict__ fy, const double* __restrict__ fz,
                                 double* __restrict__ ke_part, double* __restrict__ w_part)
{
    int64_t g = unit_ * blockDim.x + threadIdx.x;
    if (g >= nslots) return;
    int s = slot_sys[g];
    if (s < 0 || !st[s].active) return;
    const SysDev& S = sys[s];
    if (S.barostat == DMDX_BARO_NONE) return;
    double ke = 0.0, w = 0.0;
    if (g - S.atom_off < S.n) {
        const double lam = st[s].lambda;
        double a = vx[g] * lam, b = vy[g] * lam, c = vz[g] * lam;
        vx[g] = a;
        vy[g] = b;
        vz[g] = c;
        ke = 0.5 * m[g] * (a * a + b * b + c * c);
        w = x[g] * fx[g] + y[g] * fy[g] + z[g] * fz[g];
    }
    ke = wave_sum(ke);
    w = wave_sum(w);
    if (lane_id() == 0) {
        ke_part[g >> 6] = ke;
        w_part[g >> 6] = w;
    }
}

__global__ void k_baro_ke_virial(const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                                 const int* __restrict__ slot_sys, int64_t nslots,
                                 const double* __restrict__ m, double* __restrict__ vx,
                                 double* __restrict__ vy, double* __restrict__ vz,
                                 const double* __restrict__ x, const double* __restrict__ y,
                                 const double* __restrict__ z, const double* __restrict__ fx,
                                 const double* __restrict__ fy, const double* __restrict__ fz,
                                 double* __restrict__ ke_part, double* __restrict__ w_part, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_) k_baro_ke_virial_unit(u_, sys, st, slot_sys, nslots, m, vx, vy, vz, x, y, z, fx, fy, fz, ke_part, w_part);
}

__global__ void __launch_bounds__(256) k_baro_finalize(SysDev* __restrict__ sys,
                                                       SysState* __restrict__ st,
                                                       const double* __restrict__ ke_part,
                                                       const double* __restrict__ w_part)
{
    __shared__ double red[4];
    int s = blockIdx.x;
    SysState& T = st[s];
    SysDev& S = sys[s];
    if (!T.active || S.barostat == DMDX_BARO_NONE) return;
    double ke = block_range_sum(ke_part, S.aw_off, S.aw_n, red);
    double w = block_range_sum(w_part, S.aw_off, S.aw_n, red);
    if (threadIdx.x != 0) return;
    T.lambda = 1.0;
    T.ke_out = ke;
    double volume = sys_volume(S);
    double scale = 1.0;
    T.xscale = 1.0;
    if (volume <= 0.0) return;
    double P_inst = (2.0 * ke + w) / (3.0 * volume);
    T.virial = w;
    T.pressure = P_inst;
    if (S.barostat == DMDX_BARO_BERENDSEN) {
        // src/barostat/berendsen_barostat.cpp:14-66
        if (S.ptau <= 0.0) return;
        double beta_internal = S.beta / 0.0602214086;
        double mu = 1.0 - (S.dt / S.ptau) * beta_internal * (S.P0 - P_inst);
        if (mu <= 0.0) mu = 0.0;
        scale = cbrt(mu);
    } else {
        // src/barostat/andersen_barostat.cpp:20-45
        double forcing = 3.0 * volume * (P_inst - S.P0) / S.Wab;
        T.eps_dot = T.eps_dot + 0.5 * S.dt * forcing;
        scale = exp(T.eps_dot * S.dt);
        T.eps_dot = T.eps_dot + 0.5 * S.dt * forcing;
    }
    T.xscale = scale;
    for (int d = 0; d < 3; ++d) {
        S.L[d] *= scale;
        S.invL[d] = 1.0 / S.L[d];
        S.iminv[d] = S.invL[d];  // engine systems have orthorhombic cells
    }
    S.vdet = 0.0;  // the volume follows the scaled edges
    T.valid = 0;
}

__device__ __forceinline__ void k_scale_positions_unit(int64_t unit_, const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                                  const int* __restrict__ slot_sys, int64_t nslots,
                                  double* __restrict__ x, double* __restrict__ y,
                                  double* __restrict__ z, double4* __restrict__ org)
{
    int64_t g = unit_ * blockDim.x + threadIdx.x;
    if (g >= nslots) return;
    int s = slot_sys[g];
    if (s < 0 || !st[s].active) return;
    const SysDev& S = sys[s];
    if (S.barostat == DMDX_BARO_NONE || g - S.atom_off >= S.n) return;
    double sc = st[s].xscale;
    x[g] *= sc;
    y[g] *= sc;
    z[g] *= sc;
    if (org) {  // mixed path: cell-local frames scale with the box
        double4 o = org[g];
        org[g] = make_double4(o.x * sc, o.y * sc, o.z * sc, o.w);
    }
}

__global__ void k_scale_positions(const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                                  const int* __restrict__ slot_sys, int64_t nslots,
                                  double* __restrict__ x, double* __restrict__ y,
                                  double* __restrict__ z, double4* __restrict__ org, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_) k_scale_positions_unit(u_, sys, st, slot_sys, nslots, x, y, z, org);
}

// apply the pending thermostat scale to the stored velocities (state reads)
__device__ __forceinline__ void k_materialize_unit(int64_t unit_, const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                              const int* __restrict__ slot_sys, int64_t nslots,
                              double* __restrict__ vx, double* __restrict__ vy,
                              double* __restrict__ vz)
{
    int64_t g = unit_ * blockDim.x + threadIdx.x;
    if (g >= nslots) return;
    int s = slot_sys[g];
    if (s < 0) return;
    const SysDev& S = sys[s];
    if (g - S.atom_off >= S.n) return;
    double lam = st[s].lambda;
    vx[g] = vx[g] * lam;
    vy[g] = vy[g] * lam;
    vz[g] = vz[g] * lam;
}

__global__ void k_materialize(const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                              const int* __restrict__ slot_sys, int64_t nslots,
                              double* __restrict__ vx, double* __restrict__ vy,
                              double* __restrict__ vz, int64_t nunits_)
{
    DMDX_FOR_UNITS(u_, nunits_) k_materialize_unit(u_, sys, st, slot_sys, nslots, vx, vy, vz);
}

__global__ void k_lambda_reset(SysState* __restrict__ st, int nsys)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsys) st[s].lambda = 1.0;
}

}  // namespace dmdx

namespace dmdx {

hipError_t launch_pre_compute(const Dev& d, hipStream_t s, int phase, int force_rebuild,
                              int reuse_enabled)
{
    k_pre_compute<<<d.nsys, 64, 0, s>>>(d.sys, d.st, d.nsys, phase, force_rebuild,
                                        reuse_enabled, d.any_auto ? d.p_disp : nullptr);
    return hipGetLastError();
}

hipError_t launch_pre_rebuild_only(const Dev& d, hipStream_t s)
{
    k_pre_rebuild_only<<<blocks_for(d.nsys, 64), 64, 0, s>>>(d.sys, d.st, d.nsys);
    return hipGetLastError();
}

hipError_t launch_gather(const Dev& d, hipStream_t s)
{
    if (!d.any_gather) return hipSuccess;
    const int BS = 256;
    if (d.precision == DMDX_FP64)
        k_gather<double, false><<<grid_for(blocks_for(d.nslots, BS)), BS, 0, s>>>(
            d.sys, d.st, d.slot_sys, d.nslots, d.c_start, d.c_mid, d.contrib, d.tf, d.x, d.y, d.z,
            d.q, reinterpret_cast<const Cx2<double>*>(d.grid), d.fx, d.fy, d.fz, (int64_t)(blocks_for(d.nslots, BS)));
    else
        k_gather<float, false><<<grid_for(blocks_for(d.nslots, BS)), BS, 0, s>>>(
            d.sys, d.st, d.slot_sys, d.nslots, d.c_start, d.c_mid, d.contrib, d.tf, d.x, d.y, d.z,
            d.q, reinterpret_cast<const Cx2<float>*>(d.grid), d.fx, d.fy, d.fz, (int64_t)(blocks_for(d.nslots, BS)));
    return hipGetLastError();
}

hipError_t launch_gather_terms(const Dev& d, hipStream_t s)
{
    if (!d.fbx) return hipSuccess;
    const int BS = 256;
    k_gather<float, true><<<grid_for(blocks_for(d.nslots, BS)), BS, 0, s>>>(
        d.sys, d.st, d.slot_sys, d.nslots, d.c_start, d.c_mid, d.contrib, d.tf, d.x, d.y, d.z,
        d.q, reinterpret_cast<const Cx2<float>*>(d.grid), d.fbx, d.fby, d.fbz,
        (int64_t)(blocks_for(d.nslots, BS)));
    return hipGetLastError();
}

hipError_t launch_kick_drift(const Dev& d, hipStream_t s, int pre)
{
    const int BS = 256;
    // the mixed path fuses its frames and the Verlet check into the drift
    const DriftOut o{d.org, d.precision == DMDX_FP64 ? nullptr : d.lpos, d.slot_of, d.xref,
                     d.p_disp};
    k_kick_drift<<<grid_for(blocks_for(d.nslots, BS)), BS, 0, s>>>(
        d.sys, d.st, d.slot_sys, d.nslots, d.invm, d.fx, d.fy, d.fz, d.vx, d.vy, d.vz, d.x, d.y,
        d.z, o, pre, (int64_t)(blocks_for(d.nslots, BS)));
    return hipGetLastError();
}

hipError_t launch_kick_ke(const Dev& d, hipStream_t s)
{
    const int BS = 256;
    k_kick_ke<<<grid_for(blocks_for(d.nslots, BS)), BS, 0, s>>>(d.sys, d.st, d.slot_sys, d.nslots, d.invm,
                                                      d.m, d.fx, d.fy, d.fz, d.fbx, d.fby, d.fbz,
                                                      d.vx, d.vy, d.vz, d.p_ke,
                                                      (int64_t)(blocks_for(d.nslots, BS)));
    return hipGetLastError();
}

hipError_t launch_finalize(const Dev& d, hipStream_t s, int with_kinetic)
{
    int tpa = 64, w = 0, ctpa = 64;
    if (d.precision != DMDX_FP64) {
        pair_shape(d, tpa, w, false);
        pair_shape(d, ctpa, w, true);
    }
    Parts P{d.p_lj, d.p_coul, d.p_bond, d.p_angle, d.p_dih, d.p_imp, d.p_pme, d.p_excl, d.p_ke,
            64 / tpa, 64 / ctpa,
            d.precision != DMDX_FP64 && d.pair_rows != 0 ? d.p_rows : nullptr};
    k_finalize<<<d.nsys, FIN_THREADS, 0, s>>>(d.sys, d.st, P, with_kinetic);
    return hipGetLastError();
}

hipError_t launch_barostat(const Dev& d, hipStream_t s)
{
    if (!d.any_baro) return hipSuccess;
    const int BS = 256;
    k_baro_ke_virial<<<grid_for(blocks_for(d.nslots, BS)), BS, 0, s>>>(d.sys, d.st, d.slot_sys, d.nslots,
                                                             d.m, d.vx, d.vy, d.vz, d.x, d.y,
                                                             d.z, d.fx, d.fy, d.fz, d.p_ke,
                                                             d.p_w, (int64_t)(blocks_for(d.nslots, BS)));
    k_baro_finalize<<<d.nsys, 256, 0, s>>>(d.sys, d.st, d.p_ke, d.p_w);
    k_scale_positions<<<grid_for(blocks_for(d.nslots, BS)), BS, 0, s>>>(d.sys, d.st, d.slot_sys, d.nslots,
                                                              d.x, d.y, d.z, d.org, (int64_t)(blocks_for(d.nslots, BS)));
    return hipGetLastError();
}

hipError_t launch_materialize(const Dev& d, hipStream_t s)
{
    const int BS = 256;
    k_materialize<<<grid_for(blocks_for(d.nslots, BS)), BS, 0, s>>>(d.sys, d.st, d.slot_sys, d.nslots,
                                                          d.vx, d.vy, d.vz, (int64_t)(blocks_for(d.nslots, BS)));
    k_lambda_reset<<<blocks_for(d.nsys, 64), 64, 0, s>>>(d.st, d.nsys);
    return hipGetLastError();
}

}  // namespace dmdx
