// andersen.hip -- Andersen thermostat (src/thermostat/andersen.cpp:14-32)
// with the reference's own random stream.
//
// Each apply() call k of a system draws from a std::mt19937_64 seeded with
// 42 + k (the unsigned seed_ member, incremented after every call with a
// positive collision probability).  For atom i in order: one uniform u in
// [0, 1); if u < frequency * dt the atom collides and its three velocity
// components become N(0, 1) * sqrt(KB * T0 / m_i).  Uniforms follow
// libstdc++ (GCC 11) std::generate_canonical<double, 53> over a 64-bit
// engine: one engine output u64, u = double(u64) * 2^-64, clamped below 1.
// Normals follow libstdc++'s std::normal_distribution<double>: the
// Marsaglia polar method, x, y = 2u - 1 until 0 < x^2 + y^2 <= 1, returning
// y * m and keeping x * m for the next call, m = sqrt(-2 log(r2) / r2); the
// distribution object (and so its kept value) is new in every apply() call.
//
// The engine is the C++ standard's mt19937_64 ([rand.predef]: w 64, n 312,
// m 156, r 31, a 0xB5026F5AA96619E9, u 29, d 0x5555555555555555, s 17,
// b 0x71D67FFFEDA60000, t 37, c 0xFFF7EEE000000000, l 43, f
// 6364136223846793005).
//
// One 64-lane workgroup per system.  The 312-word state lives in LDS and is
// regenerated in parallel (two dependent half-twists and the last word);
// tempering is per word.  The consumer walk is serial in the reference, but
// an atom that does not collide consumes exactly one output, so each wave
// pass tests the next 64 atoms against the next 64 outputs at once and a
// ballot finds the first collision; only collisions (a fraction
// frequency * dt of the atoms) are resolved one at a time, by every lane
// redundantly so control flow stays uniform.

#include "dmdx_internal.h"

namespace dmdx {

namespace {

constexpr int MT_N = 312, MT_M = 156;
constexpr uint64_t MT_A = 0xB5026F5AA96619E9ull;
constexpr uint64_t MT_UPPER = 0xFFFFFFFF80000000ull;  // top w - r bits
constexpr uint64_t MT_LOWER = 0x000000007FFFFFFFull;  // low r bits
constexpr uint64_t MT_F = 6364136223846793005ull;

__device__ __forceinline__ uint64_t mt_mix(uint64_t a, uint64_t b)
{
    const uint64_t x = (a & MT_UPPER) | (b & MT_LOWER);
    return (x >> 1) ^ ((x & 1ull) ? MT_A : 0ull);
}

__device__ __forceinline__ uint64_t mt_temper(uint64_t y)
{
    y ^= (y >> 29) & 0x5555555555555555ull;
    y ^= (y << 17) & 0x71D67FFFEDA60000ull;
    y ^= (y << 37) & 0xFFF7EEE000000000ull;
    y ^= y >> 43;
    return y;
}

// Next 312 outputs: twist the state in place, temper into out.  Called by
// all 64 lanes together.
__device__ void mt_refill(uint64_t* mt, uint64_t* out)
{
    const int t = threadIdx.x;
    uint64_t v[3];
    // words [0, 156): old[i + 156] ^ mix(old[i], old[i + 1])
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int i = t + 64 * r;
        if (i < MT_N - MT_M) v[r] = mt[i + MT_M] ^ mt_mix(mt[i], mt[i + 1]);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int i = t + 64 * r;
        if (i < MT_N - MT_M) mt[i] = v[r];
    }
    __syncthreads();
    // words [156, 311): new[i - 156] ^ mix(old[i], old[i + 1])
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int i = MT_N - MT_M + t + 64 * r;
        if (i < MT_N - 1) v[r] = mt[i - (MT_N - MT_M)] ^ mt_mix(mt[i], mt[i + 1]);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const int i = MT_N - MT_M + t + 64 * r;
        if (i < MT_N - 1) mt[i] = v[r];
    }
    __syncthreads();
    // word 311: new[155] ^ mix(old[311], new[0])
    if (t == 0) mt[MT_N - 1] = mt[MT_M - 1] ^ mt_mix(mt[MT_N - 1], mt[0]);
    __syncthreads();
    for (int i = t; i < MT_N; i += 64) out[i] = mt_temper(mt[i]);
    __syncthreads();
}

// std::generate_canonical<double, 53> of one 64-bit output
__device__ __forceinline__ double canonical(uint64_t u)
{
    double r = (double)u * 0x1p-64;
    if (r >= 1.0) r = 0x1.fffffffffffffp-1;  // nextafter(1.0, 0.0)
    return r;
}

struct Stream {
    uint64_t* mt;
    uint64_t* out;
    int p;  // next output index in out; MT_N: refill first

    __device__ uint64_t next()
    {
        if (p == MT_N) {
            mt_refill(mt, out);
            p = 0;
        }
        return out[p++];
    }
};

struct Normal {  // std::normal_distribution<double>(0, 1), one per apply()
    double saved = 0.0;
    bool available = false;

    __device__ double draw(Stream& g)
    {
        double ret;
        if (available) {
            available = false;
            ret = saved;
        } else {
            double x, y, r2;
            do {
                x = 2.0 * canonical(g.next()) - 1.0;
                y = 2.0 * canonical(g.next()) - 1.0;
                r2 = x * x + y * y;
            } while (r2 > 1.0 || r2 == 0.0);
            const double mult = sqrt(-2.0 * log(r2) / r2);
            saved = x * mult;
            available = true;
            ret = y * mult;
        }
        return ret * 1.0 + 0.0;  // stddev 1, mean 0
    }
};

}  // namespace

__global__ void __launch_bounds__(64) k_andersen(const SysDev* __restrict__ sys,
                                                 SysState* __restrict__ st,
                                                 const double* __restrict__ m,
                                                 double* __restrict__ vx,
                                                 double* __restrict__ vy,
                                                 double* __restrict__ vz)
{
    __shared__ uint64_t mt[MT_N];
    __shared__ uint64_t out[MT_N];
    const int s = blockIdx.x;
    const SysDev& S = sys[s];
    if (S.thermostat != DMDX_THERMO_ANDERSEN || !st[s].active) return;
    const double prob = S.and_prob;
    if (prob <= 0.0) return;  // andersen.cpp:16 (no seed increment either)
    const int t = threadIdx.x;
    if (t == 0) {
        // seed_ is an unsigned int: 42 + calls, modulo 2^32
        const uint32_t seed = 42u + (uint32_t)st[s].and_calls;
        uint64_t x = seed;
        mt[0] = x;
        for (int i = 1; i < MT_N; ++i) {
            x = MT_F * (x ^ (x >> 62)) + (uint64_t)i;
            mt[i] = x;
        }
    }
    __syncthreads();
    Stream g{mt, out, MT_N};
    Normal nd;
    const int n = S.n;
    const int64_t base = S.atom_off;
    const double kT = KB * S.T0;
    int i = 0;
    while (i < n) {
        if (g.p == MT_N) {
            mt_refill(mt, out);
            g.p = 0;
        }
        const int w = min(min(64, MT_N - g.p), n - i);
        bool hit = false;
        if (t < w) hit = canonical(out[g.p + t]) < prob;
        const unsigned long long mask = __ballot(hit);
        if (mask == 0ull) {
            i += w;
            g.p += w;
            continue;
        }
        const int k = __ffsll((long long)mask) - 1;
        i += k;
        g.p += k + 1;  // atom i's uniform
        const double sigma = sqrt(kT / m[base + i]);
        const double a = nd.draw(g) * sigma;
        const double b = nd.draw(g) * sigma;
        const double c = nd.draw(g) * sigma;
        if (t == 0) {
            vx[base + i] = a;
            vy[base + i] = b;
            vz[base + i] = c;
        }
        ++i;
    }
    if (t == 0) st[s].and_calls += 1;
}

hipError_t launch_andersen(const Dev& d, hipStream_t s)
{
    if (!d.any_andersen) return hipSuccess;
    k_andersen<<<d.nsys, 64, 0, s>>>(d.sys, d.st, d.m, d.vx, d.vy, d.vz);
    return hipGetLastError();
}

}  // namespace dmdx
