// dev_common.h -- device-side helpers shared by the gfx950 kernels of the 2D Delaunay +
// alpha-shape pipeline: mesh encodings, FP64 semi-static predicate filters, and an exact
// fixed-capacity big-integer path for the cases a filter cannot certify.
//
// Mesh encoding follows the reference ABI (gpudt/gpudt.h:137-244): a triangle stores
// vtx[3] counter-clockwise and tri[i] = (t << 2) | j, the neighbour t across the edge
// opposite vtx[i], with t.vtx[j] opposite the same edge; -1 marks a hull edge.
//
// Every translation unit that includes this is compiled with -ffp-contract=off: the
// filter error bounds (Shewchuk 1997, the constants in gpudt/predicates.cpp:28) assume
// separately rounded products and sums.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GA_INLINE __device__ __forceinline__

// ---- mesh encodings --------------------------------------------------------------
GA_INLINE int enc(int t, int j) { return (t << 2) | j; }
GA_INLINE int enc_t(int e) { return e >> 2; }
GA_INLINE int enc_j(int e) { return e & 3; }
GA_INLINE int p1(int i) { return i == 2 ? 0 : i + 1; }
GA_INLINE int m1(int i) { return i == 0 ? 2 : i - 1; }

// Dead triangle marker in tv[].x
#define GA_DEAD (-2)
#define GA_NONE (-1)

// ---- FP64 semi-static filters -------------------------------------------------------
// Error-bound constants of the published method (epsilon = 2^-53).
#define GA_EPS 1.1102230246251565e-16
#define GA_CCWERRA ((3.0 + 16.0 * GA_EPS) * GA_EPS)
#define GA_ICCERRA ((10.0 + 96.0 * GA_EPS) * GA_EPS)

// returns +1/-1 when certain, 0 when the filter cannot decide
GA_INLINE int orient_filter(double2 a, double2 b, double2 c)
{
    double detleft = (a.x - c.x) * (b.y - c.y);
    double detright = (a.y - c.y) * (b.x - c.x);
    double det = detleft - detright;
    double detsum;
    if (detleft > 0.0) {
        if (detright <= 0.0) return det > 0 ? 1 : (det < 0 ? -1 : 0);
        detsum = detleft + detright;
    } else if (detleft < 0.0) {
        if (detright >= 0.0) return det > 0 ? 1 : (det < 0 ? -1 : 0);
        detsum = -detleft - detright;
    } else {
        // detleft == 0: det = -detright, exact in sign
        return det > 0 ? 1 : (det < 0 ? -1 : 0);
    }
    double eb = GA_CCWERRA * detsum;
    if (det >= eb && det != 0.0) return 1;
    if (-det >= eb && det != 0.0) return -1;
    return 0;
}

GA_INLINE int incircle_filter(double2 a, double2 b, double2 c, double2 d)
{
    double adx = a.x - d.x, bdx = b.x - d.x, cdx = c.x - d.x;
    double ady = a.y - d.y, bdy = b.y - d.y, cdy = c.y - d.y;
    double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    double alift = adx * adx + ady * ady;
    double cdxady = cdx * ady, adxcdy = adx * cdy;
    double blift = bdx * bdx + bdy * bdy;
    double adxbdy = adx * bdy, bdxady = bdx * ady;
    double clift = cdx * cdx + cdy * cdy;
    double det = alift * (bdxcdy - cdxbdy) + blift * (cdxady - adxcdy) + clift * (adxbdy - bdxady);
    double perm = (fabs(bdxcdy) + fabs(cdxbdy)) * alift + (fabs(cdxady) + fabs(adxcdy)) * blift
                + (fabs(adxbdy) + fabs(bdxady)) * clift;
    double eb = GA_ICCERRA * perm;
    if (det > eb) return 1;
    if (-det > eb) return -1;
    return 0;
}

// ---- exact path: fixed-capacity signed integers -------------------------------------
// All coordinates that take part in one exact evaluation are written as integers
// m * 2^(e - s) over the common unit 2^s (s = smallest ulp exponent among them), so the
// predicate polynomials evaluate exactly in integer arithmetic.  NL 32-bit limbs.
template <int NL>
struct BigInt {
    uint32_t l[NL];
    int neg;
    int n;  // used limbs
};

template <int NL>
GA_INLINE void bi_zero(BigInt<NL> &a) { a.n = 0; a.neg = 0; }

template <int NL>
GA_INLINE void bi_norm(BigInt<NL> &a)
{
    while (a.n > 0 && a.l[a.n - 1] == 0) a.n--;
    if (a.n == 0) a.neg = 0;
}

// |a| vs |b|
template <int NL>
GA_INLINE int bi_cmpmag(const BigInt<NL> &a, const BigInt<NL> &b)
{
    if (a.n != b.n) return a.n < b.n ? -1 : 1;
    for (int i = a.n - 1; i >= 0; i--)
        if (a.l[i] != b.l[i]) return a.l[i] < b.l[i] ? -1 : 1;
    return 0;
}

template <int NL>
GA_INLINE void bi_addmag(BigInt<NL> &r, const BigInt<NL> &a, const BigInt<NL> &b, int *ovf)
{
    int n = a.n > b.n ? a.n : b.n;
    uint64_t carry = 0;
    for (int i = 0; i < n; i++) {
        uint64_t s = carry + (i < a.n ? a.l[i] : 0u) + (i < b.n ? b.l[i] : 0u);
        r.l[i] = (uint32_t)s;
        carry = s >> 32;
    }
    if (carry) {
        if (n < NL) r.l[n++] = (uint32_t)carry;
        else *ovf = 1;
    }
    r.n = n;
}

// r = |a| - |b|, requires |a| >= |b|
template <int NL>
GA_INLINE void bi_submag(BigInt<NL> &r, const BigInt<NL> &a, const BigInt<NL> &b)
{
    int64_t borrow = 0;
    for (int i = 0; i < a.n; i++) {
        int64_t s = (int64_t)a.l[i] - (i < b.n ? (int64_t)b.l[i] : 0) - borrow;
        borrow = s < 0;
        r.l[i] = (uint32_t)(s + (borrow << 32));
    }
    r.n = a.n;
    bi_norm(r);
}

template <int NL>
GA_INLINE void bi_add(BigInt<NL> &r, const BigInt<NL> &a, const BigInt<NL> &b, int *ovf)
{
    if (a.neg == b.neg) {
        int s = a.neg;
        bi_addmag(r, a, b, ovf);
        r.neg = s;
        bi_norm(r);
        return;
    }
    int c = bi_cmpmag(a, b);
    if (c == 0) { bi_zero(r); return; }
    if (c > 0) { int s = a.neg; bi_submag(r, a, b); r.neg = s; }
    else { int s = b.neg; bi_submag(r, b, a); r.neg = s; }
    bi_norm(r);
}

template <int NL>
GA_INLINE void bi_sub(BigInt<NL> &r, const BigInt<NL> &a, const BigInt<NL> &b, int *ovf)
{
    BigInt<NL> nb = b;
    if (nb.n) nb.neg = !nb.neg;
    bi_add(r, a, nb, ovf);
}

template <int NL>
GA_INLINE void bi_mul(BigInt<NL> &r, const BigInt<NL> &a, const BigInt<NL> &b, int *ovf)
{
    BigInt<NL> t;
    int n = a.n + b.n;
    if (a.n == 0 || b.n == 0) { bi_zero(r); return; }
    if (n > NL) { *ovf = 1; bi_zero(r); return; }
    for (int i = 0; i < n; i++) t.l[i] = 0;
    for (int i = 0; i < a.n; i++) {
        uint64_t carry = 0;
        uint64_t ai = a.l[i];
        for (int j = 0; j < b.n; j++) {
            uint64_t cur = t.l[i + j] + ai * b.l[j] + carry;
            t.l[i + j] = (uint32_t)cur;
            carry = cur >> 32;
        }
        int k = i + b.n;
        while (carry && k < n) {
            uint64_t cur = (uint64_t)t.l[k] + carry;
            t.l[k] = (uint32_t)cur;
            carry = cur >> 32;
            k++;
        }
    }
    t.n = n;
    t.neg = a.neg ^ b.neg;
    bi_norm(t);
    r = t;
}

template <int NL>
GA_INLINE int bi_sign(const BigInt<NL> &a) { return a.n == 0 ? 0 : (a.neg ? -1 : 1); }

// exponent of the unit in the last place of x (x != 0), for the common scale
GA_INLINE int ulp_exp(double x)
{
    int e;
    (void)frexp(x, &e);   // x = f * 2^e, 0.5 <= |f| < 1
    int ue = e - 53;       // ulp of a normal double
    if (ue < -1074) ue = -1074;
    return ue;
}

// x = m * 2^(ue) with m integer (|m| < 2^53); returns bigint m * 2^(ue - s), s <= ue
template <int NL>
GA_INLINE void bi_from_double(BigInt<NL> &r, double x, int s, int *ovf)
{
    bi_zero(r);
    if (x == 0.0) return;
    int ue = ulp_exp(x);
    double mf = ldexp(fabs(x), -ue);         // exact integer < 2^53
    uint64_t m = (uint64_t)mf;
    int sh = ue - s;
    int limb = sh >> 5, bit = sh & 31;
    if (limb + 3 > NL) { *ovf = 1; return; }
    for (int i = 0; i < limb + 3; i++) r.l[i] = 0;
    uint64_t vlo = m << bit, vhi = bit ? (m >> (64 - bit)) : 0;
    r.l[limb] = (uint32_t)vlo;
    r.l[limb + 1] = (uint32_t)(vlo >> 32);
    r.l[limb + 2] = (uint32_t)vhi;
    r.n = limb + 3;
    r.neg = x < 0;
    bi_norm(r);
}

GA_INLINE int min_ulp_exp(const double *v, int n)
{
    int s = 100000;
    for (int i = 0; i < n; i++)
        if (v[i] != 0.0) {
            int u = ulp_exp(v[i]);
            if (u < s) s = u;
        }
    return s == 100000 ? 0 : s;
}

// exact orientation sign (ovf set when the capacity would be exceeded)
GA_INLINE int orient_exact(double2 a, double2 b, double2 c, int *ovf)
{
    double v[6] = {a.x, a.y, b.x, b.y, c.x, c.y};
    int s = min_ulp_exp(v, 6);
    BigInt<12> ax, ay, bx, by, cx, cy, t1, t2, t3, t4, r;
    bi_from_double(ax, a.x, s, ovf); bi_from_double(ay, a.y, s, ovf);
    bi_from_double(bx, b.x, s, ovf); bi_from_double(by, b.y, s, ovf);
    bi_from_double(cx, c.x, s, ovf); bi_from_double(cy, c.y, s, ovf);
    // (ax-cx)(by-cy) - (ay-cy)(bx-cx)
    bi_sub(t1, ax, cx, ovf); bi_sub(t2, by, cy, ovf);
    bi_sub(t3, ay, cy, ovf); bi_sub(t4, bx, cx, ovf);
    bi_mul(t1, t1, t2, ovf);
    bi_mul(t3, t3, t4, ovf);
    bi_sub(r, t1, t3, ovf);
    return bi_sign(r);
}

GA_INLINE int orient_sign(double2 a, double2 b, double2 c, int *ovf)
{
    int s = orient_filter(a, b, c);
    if (s != 0) return s;
    // exact zero is decided by the exact path too (the filter returns 0 for det == 0)
    return orient_exact(a, b, c, ovf);
}

// exact incircle sign: det3 of rows (dx, dy, dx^2 + dy^2) relative to d
GA_INLINE int incircle_exact(double2 a, double2 b, double2 c, double2 d, int *ovf)
{
    double v[8] = {a.x, a.y, b.x, b.y, c.x, c.y, d.x, d.y};
    int s = min_ulp_exp(v, 8);
    BigInt<24> X[4], Y[4];
    double2 P[4] = {a, b, c, d};
    for (int k = 0; k < 4; k++) {
        bi_from_double(X[k], P[k].x, s, ovf);
        bi_from_double(Y[k], P[k].y, s, ovf);
    }
    BigInt<24> dx[3], dy[3], lf[3], t, u, m, r;
    for (int k = 0; k < 3; k++) {
        bi_sub(dx[k], X[k], X[3], ovf);
        bi_sub(dy[k], Y[k], Y[3], ovf);
        bi_mul(t, dx[k], dx[k], ovf);
        bi_mul(u, dy[k], dy[k], ovf);
        bi_add(lf[k], t, u, ovf);
    }
    bi_zero(r);
    // lift_a*(bdx*cdy - cdx*bdy) + lift_b*(cdx*ady - adx*cdy) + lift_c*(adx*bdy - bdx*ady)
    for (int k = 0; k < 3; k++) {
        int i = (k + 1) % 3, j = (k + 2) % 3;
        bi_mul(t, dx[i], dy[j], ovf);
        bi_mul(u, dx[j], dy[i], ovf);
        bi_sub(m, t, u, ovf);
        bi_mul(t, lf[k], m, ovf);
        bi_add(r, r, t, ovf);
    }
    return bi_sign(r);
}

// Tie-break for exactly co-circular quadruples (lifted heights perturbed by eps^(index+1),
// smaller index more significant).  Same rule as the oracle.
GA_INLINE int incircle_sos_tie(double2 a, int ia, double2 b, int ib, double2 c, int ic,
                               double2 d, int id, int *ovf)
{
    int idx[4] = {ia, ib, ic, id};
    int used = 0;
    for (int r = 0; r < 4; r++) {
        int k = -1;
        for (int q = 0; q < 4; q++)
            if (!(used & (1 << q)) && (k < 0 || idx[q] < idx[k])) k = q;
        used |= 1 << k;
        int o;
        switch (k) {
        case 0: o = orient_sign(b, c, d, ovf); break;
        case 1: o = -orient_sign(a, c, d, ovf); break;
        case 2: o = orient_sign(a, b, d, ovf); break;
        default: o = -orient_sign(a, b, c, ovf); break;
        }
        if (o != 0) return o;
    }
    return 0;
}

// Exact zero test of the incircle determinant by residues: 1 when it is zero, 0 when it is not,
// -1 when the coordinates need more than 60 bits at their common scale (then the big-integer
// path decides).  Each coordinate is m * 2^e with m odd; dropping the common factor 2^t (t the
// smallest e) changes neither the sign nor the zero-ness, and leaves integers X below 2^B.
// Then |det| < 2^(4B + 8), and a determinant that vanishes modulo primes whose product exceeds
// that bound is zero (the primes lie just below 2^31, so residue products fit 64 bits).  Snapped
// coordinates are multiples of the snapping unit, so B <= 52 and eight primes suffice.
// Co-circular lattice quadruples, the exact path's common case, skip the big integers.
__host__ __device__ __forceinline__ int incircle_zero_res(double2 a, double2 b, double2 c, double2 d)
{
    const double v[8] = {a.x, a.y, b.x, b.y, c.x, c.y, d.x, d.y};
    unsigned long long mm[8];
    int ex[8], t = 1 << 30;
    for (int i = 0; i < 8; i++) {
        mm[i] = 0;
        ex[i] = 0;
        if (v[i] == 0.0) continue;
        int e;
        const double f = frexp(fabs(v[i]), &e);            // |v| = f * 2^e, 0.5 <= f < 1
        unsigned long long m = (unsigned long long)ldexp(f, 53);   // exact, below 2^53
        int x = e - 53;
        const int tz = __builtin_ctzll(m);
        m >>= tz;
        x += tz;
        mm[i] = m;
        ex[i] = x;
        t = x < t ? x : t;
    }
    if (t == (1 << 30)) return 1;   // every coordinate zero
    long long X[8];
    int B = 1;
    for (int i = 0; i < 8; i++) {
        X[i] = 0;
        if (mm[i] == 0) continue;
        const int k = ex[i] - t;
        const int bl = 64 - __builtin_clzll(mm[i]) + k;
        if (bl > 60) return -1;
        X[i] = (long long)(mm[i] << k);
        if (v[i] < 0.0) X[i] = -X[i];
        B = bl > B ? bl : B;
    }
    const int np = (4 * B + 8 + 29) / 30;   // primes > 2^30 each
    const long long dx[3] = {X[0] - X[6], X[2] - X[6], X[4] - X[6]};
    const long long dy[3] = {X[1] - X[7], X[3] - X[7], X[5] - X[7]};
    constexpr long long P[9] = {2147483647LL, 2147483629LL, 2147483587LL, 2147483579LL, 2147483563LL,
                                2147483549LL, 2147483543LL, 2147483497LL, 2147483489LL};
#pragma unroll
    for (int q = 0; q < 9; q++) {
        if (q >= np) break;
        const long long p = P[q];
        unsigned long long rx[3], ry[3], lf[3];
        for (int k = 0; k < 3; k++) {
            long long r = dx[k] % p;
            rx[k] = (unsigned long long)(r < 0 ? r + p : r);
            r = dy[k] % p;
            ry[k] = (unsigned long long)(r < 0 ? r + p : r);
            lf[k] = (rx[k] * rx[k] + ry[k] * ry[k]) % (unsigned long long)p;
        }
        unsigned long long det = 0;
        for (int k = 0; k < 3; k++) {
            const int i = (k + 1) % 3, j = (k + 2) % 3;
            // minor = rx[i] * ry[j] - rx[j] * ry[i] (mod p)
            const unsigned long long t1 = (rx[i] * ry[j]) % (unsigned long long)p;
            const unsigned long long t2 = (rx[j] * ry[i]) % (unsigned long long)p;
            const unsigned long long mn = t1 >= t2 ? t1 - t2 : t1 + (unsigned long long)p - t2;
            det = (det + lf[k] * mn) % (unsigned long long)p;
        }
        if (det != 0) return 0;
    }
    return 1;
}

GA_INLINE int incircle_exact_sos(double2 a, int ia, double2 b, int ib, double2 c, int ic,
                                 double2 d, int id, int *ovf)
{
    // a determinant the residues show to be zero skips the big integers
    int s = incircle_zero_res(a, b, c, d) == 1 ? 0 : incircle_exact(a, b, c, d, ovf);
    if (s != 0) return s;
    return incircle_sos_tie(a, ia, b, ib, c, ic, d, id, ovf);
}

// ---- wave-aggregated counters -------------------------------------------------------------
// Every active lane calls this at the same site; lanes with pred reserve one slot each.  One
// atomic per wave instead of one per lane (single-address atomics serialise in L2).
GA_INLINE int wave_reserve(int *counter, bool pred)
{
    unsigned long long active = __ballot(1);
    unsigned long long m = __ballot(pred);
    int lane = __lane_id();
    int leader = __ffsll((long long)active) - 1;
    int base = 0;
    if (lane == leader && m) base = atomicAdd(counter, (int)__popcll(m));
    base = __shfl(base, leader, 64);
    return base + (int)__popcll(m & ((1ull << lane) - 1ull));
}

// sum of v over a fully active wave (call only where all 64 lanes are active, e.g. after a
// grid-stride loop), one atomic per wave
GA_INLINE void wave_count_full(int *counter, int v)
{
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    if (__lane_id() == 0 && v) atomicAdd(counter, v);
}

// Atomics on one address from many waves serialise at the memory side (about 11 ns each on
// MI355X), so hot paths reserve list space per wave or per block, and statistics go to striped
// counters: stat k of stripe s lives at stats[s * GA_STRIPE_INTS + k], the host sums the stripes.
#ifndef GA_STRIPES
#define GA_STRIPES 16
#define GA_STRIPE_INTS 1024
#endif
GA_INLINE int *stat_slot(int *stats, int k)
{
    return stats + ((blockIdx.x * 4 + (threadIdx.x >> 6)) & (GA_STRIPES - 1)) * GA_STRIPE_INTS + k;
}
// statistics: count the active lanes with pred (any active subset of the wave)
GA_INLINE void stat_inc(int *stats, int k, bool pred)
{
    unsigned long long active = __ballot(1), m = __ballot(pred);
    if (m && __lane_id() == __ffsll((long long)active) - 1) atomicAdd(stat_slot(stats, k), (int)__popcll(m));
}
// statistics: sum of v over a fully active wave
GA_INLINE void stat_add_full(int *stats, int k, int v)
{
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    if (__lane_id() == 0 && v) atomicAdd(stat_slot(stats, k), v);
}

// k consecutive slots per lane with pred (any active subset of the wave)
GA_INLINE int wave_reserve_mult(int *counter, bool pred, int k)
{
    unsigned long long active = __ballot(1), m = __ballot(pred);
    int lane = __lane_id(), leader = __ffsll((long long)active) - 1;
    int base = 0;
    if (lane == leader && m) base = atomicAdd(counter, k * (int)__popcll(m));
    base = __shfl(base, leader, 64);
    return base + k * (int)__popcll(m & ((1ull << lane) - 1ull));
}

// one slot for each of pa, pb (any active subset of the wave); one atomic per wave
GA_INLINE void wave_reserve2(int *counter, bool pa, bool pb, int *qa, int *qb)
{
    unsigned long long active = __ballot(1), ma = __ballot(pa), mb = __ballot(pb);
    int lane = __lane_id(), leader = __ffsll((long long)active) - 1;
    int tot = (int)__popcll(ma) + (int)__popcll(mb);
    int base = 0;
    if (lane == leader && tot) base = atomicAdd(counter, tot);
    base = __shfl(base, leader, 64);
    unsigned long long lo = (1ull << lane) - 1ull;
    *qa = base + (int)__popcll(ma & lo);
    *qb = base + (int)__popcll(ma) + (int)__popcll(mb & lo);
}

// c slots per lane, 0 <= c < 2^NB (any active subset of the wave): the wave prefix sum is built
// from one ballot per bit of c, so inactive lanes contribute nothing
template <int NB>
GA_INLINE int wave_reserve_n(int *counter, int c)
{
    unsigned long long active = __ballot(1);
    int lane = __lane_id(), leader = __ffsll((long long)active) - 1;
    unsigned long long lo = (1ull << lane) - 1ull;
    int tot = 0, pre = 0;
#pragma unroll
    for (int b = 0; b < NB; b++) {
        unsigned long long m = __ballot((c >> b) & 1);
        tot += (int)__popcll(m) << b;
        pre += (int)__popcll(m & lo) << b;
    }
    int base = 0;
    if (lane == leader && tot) base = atomicAdd(counter, tot);
    return __shfl(base, leader, 64) + pre;
}

// c slots per thread for a whole block (up to 1024 threads): every thread of the block calls it at
// the same site (block-uniform control flow); one atomic per block
GA_INLINE int block_reserve(int *counter, int c)
{
    __shared__ int s_w[16];
    __shared__ int s_base;
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int inc = c;
    for (int d = 1; d < 64; d <<= 1) {
        int y = __shfl_up(inc, d, 64);
        if (lane >= d) inc += y;
    }
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) tot += s_w[i];
        s_base = tot ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    int off = s_base + inc - c;
    for (int i = 0; i < w; i++) off += s_w[i];
    __syncthreads();
    return off;
}

// ---- misc ----------------------------------------------------------------------------
GA_INLINE uint32_t hash32(uint32_t x)
{
    x ^= x >> 16; x *= 0x7feb352dU;
    x ^= x >> 15; x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// error codes written to the device status word
#define GA_ERR_EXACT_OVERFLOW 1
#define GA_ERR_WALK 2
#define GA_ERR_RECON 4
#define GA_ERR_SCRATCH 8
#define GA_ERR_CAPACITY 16
#define GA_ERR_ISLANDS 32
#define GA_ERR_PATCH 64
// GA_ERR_CONS_* (constraint insertion): kernels.h
