// pair_rows.hip -- mixed-path Lennard-Jones pair forces on row tiles staged
// in LDS (LJ terms of src/force/lennard_jones.cpp:58-89, full neighbour
// rows as the mixed list build writes them).
//
// A tile is one x-row of cells of one system: the cells (x, y, z) for every
// x at a fixed (y, z).  Its list entries only name atoms of the rows
// (y + oy, z + oz) for the offsets the list build scans (oy, oz within
// +-cdiv, or every row once when an edge has fewer than 2 cdiv + 1 cells).
// Each of those j-rows is a contiguous range of cell-order slots, so the
// workgroup copies them, one coalesced range per row, into LDS and every list
// entry then reads its neighbour from LDS instead of gathering it from
// global memory:
//     staged index = slot + (row base in LDS - row's first slot),
// where the difference depends only on the entry's offset code and sits in
// the w field of the code table beside the cell-offset shift.
// The rows' padding (code NBR_PAD_CODE over the atom's own slot) maps past
// the staged rows and is clamped to a sentinel LDS entry 1e15 nm away, so a
// padding entry adds exactly zero.
//
// Work split.  The tile's atoms are taken up to PR_ATOMS at a time; their
// rows are cut into blocks of W entries (rows are padded to multiples of 8),
// and the chunk's blocks, in atom order, are dealt out to the 256 lanes as
// contiguous ranges of equal length, whatever the row lengths.  A lane sums
// its blocks atom by atom (W entries in FP32, block sums in FP64).  An atom
// whose blocks all lie in one lane is written by that lane; the partial sums
// of atoms cut by lane boundaries go to LDS (the lane's first atom, begun in
// an earlier lane, as `head`; its last atom, continued in a later lane, as
// `tail`) and are added by the atom's own thread in lane order.  Everything
// is a fixed order set by the system alone: results are the same alone, in
// any batch and on every run.
#include "dmdx_internal.h"

#include <type_traits>

namespace dmdx {

constexpr int PR_THREADS = 256;
constexpr int PR_CAP = 1536;   // staged positions per tile (24 KB)
constexpr int PR_ATOMS = 256;  // tile atoms per chunk
constexpr int PR_ROWS = 25;    // j-rows of a tile: 3 x 3, or 5 x 5 with half-width cells

struct PrShared {
    float4 stage[PR_CAP + 1];  // staged rows; [PR_CAP] is the far sentinel
    float4 tab[NBR_CODES];     // per offset code: shift (x, y, z), index delta (w bits);
                               // only the codes of the tile's offsets are written
    float4 head[PR_THREADS];   // per lane: partial sum of its first atom (begun earlier)
    float4 tail[PR_THREADS];   // per lane: partial sum of its last atom (continued later)
    int bpre[PR_ATOMS + 1];    // running block counts of the chunk's atoms
    int nnl[PR_ATOMS];         // row lengths of the chunk's atoms
    int rs0[PR_ROWS];          // first cell-order slot of each j-row (relative)
    int rcnt[PR_ROWS];         // atoms of each j-row
    int rbase[PR_ROWS + 1];    // staged index of each j-row's first atom
    int wsum[PR_THREADS / 64];
};

// one atom's finished sums: forces into the atom's slot, half the row energy
// (rows hold both orders of a pair) into the slot's energy partial
__device__ __forceinline__ void pr_store(int64_t t, int i, double sx, double sy, double sz,
                                         double se, double* __restrict__ fx,
                                         double* __restrict__ fy, double* __restrict__ fz,
                                         double* __restrict__ e_lj_part, int accumulate)
{
    if (accumulate) {
        sx += fx[i];
        sy += fy[i];
        sz += fz[i];
    }
    fx[i] = sx;
    fy[i] = sy;
    fz[i] = sz;
    e_lj_part[t] = 0.5 * se;
}

// index of the last atom a < A with bpre[a] <= b (b < bpre[A]): the atom of block b
__device__ __forceinline__ int pr_atom_of(const int* __restrict__ bpre, int A, int b)
{
    int lo = 0, hi = A - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bpre[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int W, bool WRAP, bool TYPED, bool STAGED>
__device__ __forceinline__ void pr_blocks(PrShared& sh, const SysDev& S, int R0, int R1, int A,
                                          int64_t tchunk, int ibase, const int* __restrict__ rows,
                                          const int* __restrict__ order,
                                          const float4* __restrict__ lps,
                                          const int* __restrict__ tys, const float* __restrict__ c12t,
                                          const float* __restrict__ c6t, double* __restrict__ fx,
                                          double* __restrict__ fy, double* __restrict__ fz,
                                          double* __restrict__ e_lj_part, int accumulate)
{
    // FP32 arithmetic of this kernel may fuse into FMAs (the library is built
    // with contraction off for the FP64 parity kernels)
#pragma clang fp contract(fast)
    const int M = S.max_nbr;
    const int nt = S.ntypes;
    const float c12_0 = c12t[0], c6_0 = c6t[0];
    const float c12x12 = 12.0f * c12_0, c6x6 = 6.0f * c6_0;
    const float cut2 = (float)S.rc_lj2;
    const float Lx = (float)S.L[0], Ly = (float)S.L[1], Lz = (float)S.L[2];
    const float ix = (float)S.iminv[0], iy = (float)S.iminv[1], iz = (float)S.iminv[2];
    const int* __restrict__ bpre = sh.bpre;

    int a = pr_atom_of(bpre, A, R0);
    int bend = bpre[a + 1];  // first block of the next atom
    // the current atom: frame position (and type), row, row length
    float4 li;
    int ti = 0;
    auto enter = [&](int aa) {
        if constexpr (STAGED) {
            li = sh.stage[ibase + aa];
            if constexpr (TYPED) ti = __float_as_int(li.w);
        } else {
            li = lps[tchunk + aa];
            if constexpr (TYPED) ti = tys[tchunk + aa];
        }
    };
    enter(a);
    const int* row = rows + (int64_t)(tchunk + a) * M;
    int cnt = WRAP ? sh.nnl[a] : 0;
    int kb = (R0 - bpre[a]) * W;  // entry offset of the block within its row
    double ax = 0.0, ay = 0.0, az = 0.0, ae = 0.0;
    auto load_block = [&](const int* rp, int k, uint32_t(&e)[W]) {
        const int4 ja = *reinterpret_cast<const int4*>(rp + k);
        e[0] = (uint32_t)ja.x, e[1] = (uint32_t)ja.y, e[2] = (uint32_t)ja.z;
        e[3] = (uint32_t)ja.w;
        if constexpr (W == 8) {
            const int4 jb = *reinterpret_cast<const int4*>(rp + k + 4);
            e[4] = (uint32_t)jb.x, e[5] = (uint32_t)jb.y, e[6] = (uint32_t)jb.z;
            e[7] = (uint32_t)jb.w;
        }
    };
    uint32_t ev[W];
    load_block(row, kb, ev);
    for (int b = R0; b < R1; ++b) {
        // the next block (of this atom, or of the next atom with blocks:
        // atoms without neighbours have none) loads while this one computes
        const int bn = b + 1;
        const bool next_atom = bn == bend && bn < R1;
        int an = a;
        const int* rown = row;
        int kbn = kb + W;
        if (next_atom) {
            do {
                ++an;
            } while (bpre[an + 1] == bn);
            rown = rows + (int64_t)(tchunk + an) * M;
            kbn = 0;
        }
        uint32_t en[W];
        if (bn < R1) load_block(rown, kbn, en);
        // the W entries of a block are summed in FP32 (from the first
        // entry's terms), then added in FP64; an entry outside the cutoff
        // (the padding's sentinel included) or, with the minimum image, past
        // the row end computes on its own r2 and adds zero through a select.
        // Entries are fetched from LDS four at a time (fewer live registers).
        float px = 0.0f, py = 0.0f, pz = 0.0f, pel = 0.0f;
#pragma unroll
        for (int h = 0; h < W; h += 4) {
            float4 pj[4], shv[4];
            int tj[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t e = ev[h + q];
                shv[q] = sh.tab[e >> NBR_SLOT_BITS];
                if constexpr (STAGED)
                    pj[q] = sh.stage[min(e + (uint32_t)__float_as_int(shv[q].w), (uint32_t)PR_CAP)];
                else pj[q] = lps[e & NBR_SLOT_MASK];
                if constexpr (TYPED)
                    tj[q] = STAGED ? __float_as_int(pj[q].w) : tys[e & NBR_SLOT_MASK];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float dx = (pj[q].x - li.x) + shv[q].x;
                float dy = (pj[q].y - li.y) + shv[q].y;
                float dz = (pj[q].z - li.z) + shv[q].z;
                bool valid = true;
                if constexpr (WRAP) {
                    valid = kb + h + q < cnt;
                    dx -= rintf(dx * ix) * Lx;
                    dy -= rintf(dy * iy) * Ly;
                    dz -= rintf(dz * iz) * Lz;
                }
                const float r2 = dx * dx + dy * dy + dz * dz;
                const bool in_lj = valid && r2 < cut2;
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
                if (h == 0 && q == 0) {
                    px = fs * dx;
                    py = fs * dy;
                    pz = fs * dz;
                    pel = el;
                } else {
                    px += fs * dx;
                    py += fs * dy;
                    pz += fs * dz;
                    pel += el;
                }
            }
        }
        ax -= (double)px;
        ay -= (double)py;
        az -= (double)pz;
        ae += (double)pel;
        if (bn == bend || bn == R1) {
            // the lane's segment of atom a ends here
            const bool began_before = bpre[a] < R0;
            const bool continues = bend > R1;
            if (!began_before && !continues) {
                const int64_t t = tchunk + a;
                pr_store(t, order[t], ax, ay, az, ae, fx, fy, fz, e_lj_part, accumulate);
            } else if (began_before) {
                sh.head[threadIdx.x] = make_float4((float)ax, (float)ay, (float)az, (float)ae);
            } else {
                sh.tail[threadIdx.x] = make_float4((float)ax, (float)ay, (float)az, (float)ae);
            }
            ax = ay = az = ae = 0.0;
        }
        if (next_atom) {
            a = an;
            bend = bpre[a + 1];
            enter(a);
            if constexpr (WRAP) cnt = sh.nnl[a];
        }
        row = rown;
        kb = kbn;
#pragma unroll
        for (int q = 0; q < W; ++q) ev[q] = en[q];
    }
}

template <int W>
__global__ void __launch_bounds__(PR_THREADS) __attribute__((amdgpu_waves_per_eu(4))) k_pair_rows(
    const SysDev* __restrict__ sys, const SysState* __restrict__ st,
    const int* __restrict__ row_sys, int nrows, const int* __restrict__ ccount,
    const int* __restrict__ cstart, const int* __restrict__ order, const int* __restrict__ nn,
    const int* __restrict__ nbr, const float4* __restrict__ lpos, const int* __restrict__ typ_s,
    const float* __restrict__ lj_c12, const float* __restrict__ lj_c6, double* __restrict__ fx,
    double* __restrict__ fy, double* __restrict__ fz, double* __restrict__ e_lj_part,
    int accumulate)
{
    __shared__ PrShared sh;
    const int tid = threadIdx.x;
    DMDX_FOR_UNITS(u, nrows) {
        // everything below down to the chunk loop is uniform over the workgroup
        const int s = row_sys[u];
        if (s < 0) continue;
        const SysDev& S = sys[s];
        const SysState& T = st[s];
        if (S.precision == DMDX_FP64 || !S.has_list || S.coul_pairs || T.reuse || !T.active)
            continue;
        const int nx = T.nub[0], ny = T.nub[1], nz = T.nub[2];
        const int lr = (int)(u - S.row_off);
        const int ry = lr % S.nc[1], rz = lr / S.nc[1];
        if (ry >= ny || rz >= nz) continue;
        // j-rows: the (y, z) offsets the list build scans, in its order
        const int reach = S.cdiv, full = 2 * reach + 1;
        auto lo_of = [&](int nu) { return nu >= full ? -reach : -((nu - 1) / 2); };
        auto hi_of = [&](int nu) { return nu >= full ? reach : nu - 1 - (nu - 1) / 2; };
        const int lox = lo_of(nx), hix = hi_of(nx);
        const int loy = lo_of(ny), hiy = hi_of(ny);
        const int loz = lo_of(nz), hiz = hi_of(nz);
        const int eyn = hiy - loy + 1;
        const int nrow = eyn * (hiz - loz + 1);  // <= PR_ROWS
        const bool wrap = nx < full || ny < full || nz < full;
        if (tid < 64) {  // the first wave: rows' slot ranges and their running sizes
            int s0 = 0, cnt = 0;
            if (tid < nrow) {
                const int oy = loy + tid % eyn, oz = loz + tid / eyn;
                int yy = ry + oy, zz = rz + oz;
                yy = yy < 0 ? yy + ny : (yy >= ny ? yy - ny : yy);
                zz = zz < 0 ? zz + nz : (zz >= nz ? zz - nz : zz);
                const int c0 = S.cell_off + (zz * ny + yy) * nx, c1 = c0 + nx - 1;
                s0 = cstart[c0];
                cnt = cstart[c1] + ccount[c1] - s0;
            }
            int incl = cnt;
            for (int o = 1; o < 32; o <<= 1) {
                const int v = __shfl_up(incl, o, 64);
                if (tid >= o) incl += v;
            }
            if (tid < nrow) {
                sh.rs0[tid] = s0;
                sh.rcnt[tid] = cnt;
                sh.rbase[tid] = incl - cnt;
            }
            if (tid == nrow - 1) sh.rbase[nrow] = incl;
        }
        __syncthreads();
        const int total = sh.rbase[nrow];
        const bool staged = total <= PR_CAP;
        const int64_t a0 = S.atom_off;
        const float wx = (float)(S.L[0] / nx), wy = (float)(S.L[1] / ny), wz = (float)(S.L[2] / nz);
        // code table: shift of the neighbour cell's frame, and the staged
        // index delta of its row (entries hold the code in their top bits,
        // so the delta subtracts it)
        if (tid < 125) {  // the offsets -2..2 along each edge
            const int ox = tid % 5 - 2, oy = (tid / 5) % 5 - 2, oz = tid / 25 - 2;
            if (ox >= lox && ox <= hix && oy >= loy && oy <= hiy && oz >= loz && oz <= hiz) {
                const int r = (oy - loy) + eyn * (oz - loz);
                const uint32_t code = nbr_code(ox, oy, oz);
                const uint32_t delta = (uint32_t)(sh.rbase[r] - sh.rs0[r]) -
                                       (code << NBR_SLOT_BITS);
                sh.tab[code] = make_float4((float)ox * wx, (float)oy * wy, (float)oz * wz,
                                           __int_as_float((int)delta));
            }
        } else if (tid == 125) {
            // padding entries (the atom's own slot): past the staged rows,
            // clamped to the sentinel, when staged; else a shift beyond every
            // cutoff
            const uint32_t delta = (uint32_t)PR_CAP - (NBR_PAD_CODE << NBR_SLOT_BITS);
            sh.tab[NBR_PAD_CODE] =
                make_float4(staged ? 0.0f : 1e15f, 0.0f, 0.0f, __int_as_float((int)delta));
        }
        const bool typed = S.use_lj && S.ntypes > 1;
        if (staged) {
            // four positions in flight per thread
            constexpr int U = 4;
            int r = 0;  // the j-row of position k (k only grows)
            for (int k0 = tid; k0 < total; k0 += U * PR_THREADS) {
                float4 p[U];
                int ty[U];
#pragma unroll
                for (int q = 0; q < U; ++q) {
                    const int k = k0 + q * PR_THREADS;
                    if (k < total) {
                        while (r + 1 < nrow && sh.rbase[r + 1] <= k) ++r;
                        const int64_t slot = a0 + sh.rs0[r] + (k - sh.rbase[r]);
                        p[q] = lpos[slot];
                        ty[q] = typed ? typ_s[slot] : 0;
                    }
                }
#pragma unroll
                for (int q = 0; q < U; ++q) {
                    const int k = k0 + q * PR_THREADS;
                    if (k < total) {
                        p[q].w = __int_as_float(ty[q]);
                        sh.stage[k] = p[q];
                    }
                }
            }
            if (tid == 0) sh.stage[PR_CAP] = make_float4(1e15f, 0.0f, 0.0f, 0.0f);
        }
        // the tile's own row: offsets (0, 0)
        const int r0 = (0 - loy) + eyn * (0 - loz);
        const int own0 = sh.rs0[r0], own_n = sh.rcnt[r0], own_base = sh.rbase[r0];
        const int* rows = nbr + S.nbr_off;
        const float4* lps = lpos + a0;
        const int* tys = typ_s + a0;
        const float* c12t = lj_c12 + S.lj_off;
        const float* c6t = lj_c6 + S.lj_off;
        for (int c0 = 0; c0 < own_n; c0 += PR_ATOMS) {
            const int A = min(PR_ATOMS, own_n - c0);
            const int64_t tchunk = own0 + c0;  // relative cell-order slot of the chunk's first atom
            // blocks per atom, exclusive running count over the workgroup
            int nb = 0;
            if (tid < A) {
                const int len = nn[a0 + tchunk + tid];
                sh.nnl[tid] = len;
                nb = (len + W - 1) / W;
            }
            int incl = nb;
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(incl, o, 64);
                if ((tid & 63) >= o) incl += v;
            }
            if ((tid & 63) == 63) sh.wsum[tid >> 6] = incl;
            __syncthreads();
            int wpre = 0;
            for (int w = 0; w < (tid >> 6); ++w) wpre += sh.wsum[w];
            if (tid < A) sh.bpre[tid] = wpre + incl - nb;
            if (tid == A - 1) sh.bpre[A] = wpre + incl;
            __syncthreads();
            const int B = sh.bpre[A];
            const int Q = (B + PR_THREADS - 1) / PR_THREADS;
            const int R0 = min(tid * Q, B), R1 = min(R0 + Q, B);
            if (R0 < R1) {
                const int ib = own_base + c0;
                // frame positions are relative to the system (lps); the
                // chunk's atoms follow tchunk
#define DMDX_PR_ARGS                                                                         \
    sh, S, R0, R1, A, tchunk, ib, rows, order + a0, lps, tys, c12t, c6t, fx, fy, fz,              \
        e_lj_part + a0, accumulate
                if (staged) {
                    if (wrap) {
                        if (typed) pr_blocks<W, true, true, true>(DMDX_PR_ARGS);
                        else pr_blocks<W, true, false, true>(DMDX_PR_ARGS);
                    } else {
                        if (typed) pr_blocks<W, false, true, true>(DMDX_PR_ARGS);
                        else pr_blocks<W, false, false, true>(DMDX_PR_ARGS);
                    }
                } else {
                    if (wrap) {
                        if (typed) pr_blocks<W, true, true, false>(DMDX_PR_ARGS);
                        else pr_blocks<W, true, false, false>(DMDX_PR_ARGS);
                    } else {
                        if (typed) pr_blocks<W, false, true, false>(DMDX_PR_ARGS);
                        else pr_blocks<W, false, false, false>(DMDX_PR_ARGS);
                    }
                }
#undef DMDX_PR_ARGS
            }
            __syncthreads();
            // atoms cut by lane boundaries (and atoms without neighbours):
            // the atom's own thread adds its lanes' partial sums in lane order
            if (tid < A) {
                const int b0 = sh.bpre[tid], b1 = sh.bpre[tid + 1];
                const int64_t t = a0 + tchunk + tid;
                if (b1 == b0) {
                    pr_store(t, order[t], 0.0, 0.0, 0.0, 0.0, fx, fy, fz, e_lj_part, accumulate);
                } else {
                    const int l1 = b0 / Q, l2 = (b1 - 1) / Q;
                    if (l1 != l2) {
                        const float4 p = sh.tail[l1];
                        double sx = p.x, sy = p.y, sz = p.z, se = p.w;
                        for (int l = l1 + 1; l <= l2; ++l) {
                            const float4 h = sh.head[l];
                            sx += h.x;
                            sy += h.y;
                            sz += h.z;
                            se += h.w;
                        }
                        pr_store(t, order[t], sx, sy, sz, se, fx, fy, fz, e_lj_part, accumulate);
                    }
                }
            }
            __syncthreads();
        }
        __syncthreads();  // the next tile rewrites the shared tables
    }
}

hipError_t launch_pair_rows(const Dev& d, hipStream_t s, int accumulate)
{
    if (d.nrows == 0) return hipSuccess;
    const unsigned g = grid_for(d.nrows);
#define DMDX_ROWS_ARGS                                                                      \
    d.sys, d.st, d.row_sys, d.nrows, d.ccount, d.cstart, d.order, d.nn, d.nbr, d.lpos, d.typ_s, \
        d.lj_c12, d.lj_c6, d.fx, d.fy, d.fz, d.p_rows, accumulate
    if (d.pair_rows == 4) k_pair_rows<4><<<g, PR_THREADS, 0, s>>>(DMDX_ROWS_ARGS);
    else k_pair_rows<8><<<g, PR_THREADS, 0, s>>>(DMDX_ROWS_ARGS);
#undef DMDX_ROWS_ARGS
    return hipGetLastError();
}

}  // namespace dmdx
