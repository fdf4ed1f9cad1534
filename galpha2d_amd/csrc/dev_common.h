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

// ---- exact zero tests by residues --------------------------------------------------------
// Coordinates as integers: each is m * 2^e with m odd; dropping the common factor 2^t (t the
// smallest e) changes neither the sign nor the zero-ness of a homogeneous polynomial in them, and
// leaves integers X below 2^B (B returned).  False when some X would need more than maxb bits.
#define RES_NP 20   // primes just below 2^31 (each above 2^30): residue products fit 64 bits
__host__ __device__ __forceinline__ long long res_prime(int q)
{
    constexpr long long P[RES_NP] = {
        2147483647LL, 2147483629LL, 2147483587LL, 2147483579LL, 2147483563LL, 2147483549LL, 2147483543LL,
        2147483497LL, 2147483489LL, 2147483477LL, 2147483423LL, 2147483399LL, 2147483353LL, 2147483323LL,
        2147483269LL, 2147483249LL, 2147483237LL, 2147483179LL, 2147483171LL, 2147483137LL};
    return P[q];
}
__host__ __device__ __forceinline__ bool res_ints(const double *v, int n, long long *X, int *Bout, int maxb,
                                             int *tout = nullptr)
{
    // two passes (the smallest exponent, then the integers) keep few values live
    int t = 1 << 30;
    for (int i = 0; i < n; i++) {
        if (v[i] == 0.0) continue;
        int e;
        const double f = frexp(fabs(v[i]), &e);            // |v| = f * 2^e, 0.5 <= f < 1
        const unsigned long long m = (unsigned long long)ldexp(f, 53);   // exact, below 2^53
        const int x = e - 53 + __builtin_ctzll(m);
        t = x < t ? x : t;
    }
    int B = 1;
    for (int i = 0; i < n; i++) {
        X[i] = 0;
        if (v[i] == 0.0) continue;
        int e;
        const double f = frexp(fabs(v[i]), &e);
        unsigned long long m = (unsigned long long)ldexp(f, 53);
        const int tz = __builtin_ctzll(m);
        m >>= tz;
        const int k = e - 53 + tz - t;
        const int bl = 64 - __builtin_clzll(m) + k;
        if (bl > maxb) return false;
        X[i] = (long long)(m << k);
        if (v[i] < 0.0) X[i] = -X[i];
        B = bl > B ? bl : B;
    }
    *Bout = B;
    if (tout) *tout = t;   // v[i] = X[i] 2^t
    return true;
}
// Reduction by Barrett's method (no division): with m = floor((2^64 - 1) / p), q = floor(a m / 2^64)
// is floor(a / p) or at most two less for a < 2^63, so a - q p needs at most two subtractions.
struct ResMod {
    unsigned long long p, m;
};
__host__ __device__ __forceinline__ ResMod res_mod(int q)
{
    constexpr unsigned long long M[RES_NP] = {
        8589934596ULL, 8589934668ULL, 8589934836ULL, 8589934868ULL,
        8589934932ULL, 8589934988ULL, 8589935012ULL, 8589935196ULL,
        8589935228ULL, 8589935276ULL, 8589935492ULL, 8589935588ULL,
        8589935772ULL, 8589935892ULL, 8589936108ULL, 8589936188ULL,
        8589936236ULL, 8589936468ULL, 8589936500ULL, 8589936636ULL};
    return ResMod{(unsigned long long)res_prime(q), M[q]};
}
__host__ __device__ __forceinline__ unsigned long long res_red(unsigned long long a, const ResMod &r)   // a < 2^63
{
    const unsigned long long q = (unsigned long long)(((unsigned __int128)a * r.m) >> 64);
    unsigned long long x = a - q * r.p;
    if (x >= r.p) x -= r.p;
    if (x >= r.p) x -= r.p;
    return x;
}
__host__ __device__ __forceinline__ unsigned long long res_of(long long x, const ResMod &r)   // |x| < 2^63
{
    const unsigned long long a = res_red((unsigned long long)(x < 0 ? -x : x), r);
    return x < 0 && a ? r.p - a : a;
}
__host__ __device__ __forceinline__ unsigned long long res_mul(unsigned long long a, unsigned long long b,
                                                                const ResMod &r)
{
    return res_red(a * b, r);
}
__host__ __device__ __forceinline__ unsigned long long res_sub(unsigned long long a, unsigned long long b,
                                                                const ResMod &r)
{
    return a >= b ? a - b : a + r.p - b;
}

// bit length of |x| (0 for 0)
__host__ __device__ __forceinline__ int res_bits(long long x)
{
    const unsigned long long u = (unsigned long long)(x < 0 ? -x : x);
    return u ? 64 - __builtin_clzll(u) : 0;
}

// Exact zero test of the incircle determinant by residues: 1 when it is zero, 0 when it is not,
// -1 when the coordinates need more than 60 bits at their common scale (then the big-integer
// path decides).  The determinant is a polynomial in the differences to d; with those below 2^D,
// |det| < 3 * 2^(4D + 2) < 2^(4D + 4), and a determinant that vanishes modulo primes whose product
// exceeds that bound is zero.  Co-circular lattice quadruples, the exact path's common case, skip
// the big integers; their differences are far shorter than their coordinates.
__host__ __device__ __forceinline__ int incircle_zero_res(double2 a, double2 b, double2 c, double2 d)
{
    const double v[8] = {a.x, a.y, b.x, b.y, c.x, c.y, d.x, d.y};
    long long X[8];
    int B;
    if (!res_ints(v, 8, X, &B, 60)) return -1;
    const long long dx[3] = {X[0] - X[6], X[2] - X[6], X[4] - X[6]};
    const long long dy[3] = {X[1] - X[7], X[3] - X[7], X[5] - X[7]};
    int D = 0;
    for (int k = 0; k < 3; k++) D = max(D, max(res_bits(dx[k]), res_bits(dy[k])));
    const int np = (4 * D + 4 + 29) / 30;   // primes above 2^30 whose product exceeds 2^(4D + 4)
    for (int q = 0; q < np; q++) {
        const ResMod p = res_mod(q);
        unsigned long long rx[3], ry[3], lf[3];
        for (int k = 0; k < 3; k++) {
            rx[k] = res_of(dx[k], p);
            ry[k] = res_of(dy[k], p);
            lf[k] = res_red(rx[k] * rx[k] + ry[k] * ry[k], p);
        }
        unsigned long long det = 0;
        for (int k = 0; k < 3; k++) {
            const int i = (k + 1) % 3, j = (k + 2) % 3;
            const unsigned long long mn = res_sub(res_mul(rx[i], ry[j], p), res_mul(rx[j], ry[i], p), p);
            det = res_red(det + lf[k] * mn, p);
        }
        if (det != 0) return 0;
    }
    return 1;
}

// Exact equality of two alpha-shape values (squared circumradius of a triangle, squared half
// length of an edge, in the order of rho_cmp_exact) by residues (res_ints): 1 equal, 0 different, -1 unknown.  value =
// num / den with edge num = |pq|^2, den = 1 and triangle num = l1 l2 l3, den = ar^2, all in the
// differences of an entity's points; with those below 2^D, num < 2^(6D + 3) and den < 2^(4D + 2),
// so |num_a den_b - num_b den_a| < 2^(10D + 5) and primes whose product exceeds that decide it.  The lattice's equal circumradii, the exact comparisons' common case, skip the
// big integers.
// the points of two alpha-value entities (st 0: edge of 2 points, 1: triangle of 3) as integers at
// their common scale, X = x0 y0 x1 y1 x2 y2 of each (unused slots 0), and D, the largest bit length
// of the differences between points of one entity; false when a coordinate needs more than 61 bits
__host__ __device__ __forceinline__ bool rho_pair_ints(int sta, const double2 *pa, int stb, const double2 *pb,
                                                       long long *X, int *Dout)
{
    double v[12];
    for (int q = 0; q < 12; q++) v[q] = 0.0;
    const int na = sta ? 3 : 2, nb = stb ? 3 : 2;
    for (int i = 0; i < na; i++) { v[2 * i] = pa[i].x; v[2 * i + 1] = pa[i].y; }
    for (int i = 0; i < nb; i++) { v[6 + 2 * i] = pb[i].x; v[6 + 2 * i + 1] = pb[i].y; }
    int B;
    if (!res_ints(v, 12, X, &B, 61)) return false;   // differences below 2^62
    int D = 0;
    for (int e = 0; e < 2; e++) {
        const long long *P = X + 6 * e;
        const int npt = (e ? stb : sta) ? 3 : 2;
        for (int i = 0; i < npt; i++) {
            const int j = i + 1 == npt ? 0 : i + 1;
            D = max(D, max(res_bits(P[2 * i] - P[2 * j]), res_bits(P[2 * i + 1] - P[2 * j + 1])));
        }
    }
    *Dout = D;
    return true;
}

//
// key_equal: the two values are known to share their 256-bit key (rho_key_fixed: the same E and
// q = floor(rho 2^-E) in [2^255, 2^256)), so |rho_a - rho_b| < 2^E <= 2^-255 rho_a and the
// cross difference is below num_a den_b 2^-255 < 2^(10D - 250): fewer primes, and none when
// 10D <= 250 (an integer below 1 in magnitude is 0).
__host__ __device__ __forceinline__ int rho_equal_res(int sta, const double2 *pa, int stb, const double2 *pb,
                                                      int key_equal = 0)
{
    long long X[12];
    int D;
    if (!rho_pair_ints(sta, pa, stb, pb, X, &D)) return -1;
    const int stv[2] = {sta, stb};
    const int bound = key_equal ? 10 * D - 250 : 10 * D + 5;
    if (bound <= 0) return 1;
    const int np = (bound + 29) / 30;
    if (np > RES_NP) return -1;
    for (int q = 0; q < np; q++) {
        const ResMod p = res_mod(q);
        unsigned long long num[2], den[2];
        for (int e = 0; e < 2; e++) {
            const long long *P = X + 6 * e;   // x0 y0 x1 y1 x2 y2
            if (!stv[e]) {
                const unsigned long long dx = res_of(P[0] - P[2], p), dy = res_of(P[1] - P[3], p);
                num[e] = res_red(dx * dx + dy * dy, p);
                den[e] = 1;
            } else {
                unsigned long long l[3];
                for (int i = 0; i < 3; i++) {
                    const int j = i == 2 ? 0 : i + 1;
                    const unsigned long long dx = res_of(P[2 * i] - P[2 * j], p);
                    const unsigned long long dy = res_of(P[2 * i + 1] - P[2 * j + 1], p);
                    l[i] = res_red(dx * dx + dy * dy, p);
                }
                num[e] = res_mul(res_mul(l[0], l[1], p), l[2], p);
                const unsigned long long ux = res_of(P[2] - P[0], p), uy = res_of(P[3] - P[1], p);
                const unsigned long long wx = res_of(P[4] - P[0], p), wy = res_of(P[5] - P[1], p);
                const unsigned long long ar = res_sub(res_mul(ux, wy, p), res_mul(uy, wx, p), p);
                den[e] = res_mul(ar, ar, p);
            }
        }
        if (res_mul(num[0], den[1], p) != res_mul(num[1], den[0], p)) return 0;
    }
    return 1;
}

// ---- alpha values in fixed-width integers ----------------------------------------------------
// Limbs of 64 bits, least significant first.  Every size is a compile-time constant and every loop
// is unrolled, so the arrays stay in registers (the big integers above live in scratch memory).
template <int NA, int NB>
__host__ __device__ __forceinline__ void mw_mul(const unsigned long long *a, const unsigned long long *b,
                                                unsigned long long *r)   // r[NA + NB] = a b
{
#pragma unroll
    for (int i = 0; i < NA + NB; i++) r[i] = 0;
#pragma unroll
    for (int i = 0; i < NA; i++) {
        unsigned long long c = 0;
#pragma unroll
        for (int j = 0; j < NB; j++) {
            // a b + r + c <= (2^64 - 1)^2 + 2 (2^64 - 1) = 2^128 - 1
            const unsigned __int128 t = (unsigned __int128)a[i] * b[j] + r[i + j] + c;
            r[i + j] = (unsigned long long)t;
            c = (unsigned long long)(t >> 64);
        }
        r[i + NB] = c;
    }
}
template <int N>
__host__ __device__ __forceinline__ int mw_bitlen(const unsigned long long *a)
{
    int L = 0;
#pragma unroll
    for (int i = 0; i < N; i++)
        if (a[i]) L = 64 * i + 64 - __builtin_clzll(a[i]);
    return L;
}
// a <<= sh (0 <= sh < 64 N, bits shifted out are lost): whole limbs in three conditional stages of
// 1, 2 and 4 limbs (register moves, no indexing by a variable), then the bits
template <int N>
__host__ __device__ __forceinline__ void mw_shl(unsigned long long *a, int sh)
{
    const int ws = sh >> 6, bs = sh & 63;
#pragma unroll
    for (int st = 1; st < N && st <= 4; st <<= 1) {
        if (ws & st) {
#pragma unroll
            for (int i = N - 1; i >= 0; i--) a[i] = i >= st ? a[i - st] : 0;
        }
    }
    if (bs) {
#pragma unroll
        for (int i = N - 1; i > 0; i--) a[i] = (a[i] << bs) | (a[i - 1] >> (64 - bs));
        a[0] <<= bs;
    }
}

// num and den of an entity's value rho = num / (4 den) (times the squared unit) from its integer
// points P (x0 y0 x1 y1 x2 y2): edge num = |pq|^2, den = 1; triangle num = l1 l2 l3, den = ar^2.
// With differences below 2^58: l < 2^117, num < 2^351 (6 limbs), |ar| < 2^117, den < 2^234 (4 limbs).
__host__ __device__ __forceinline__ void rho_numden(int st, const long long *P, unsigned long long *num,
                                                    unsigned long long *den)
{
    unsigned long long l[3][2];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int j = i == 2 ? 0 : i + 1;
        const __int128 dx = (__int128)(P[2 * i] - P[2 * j]), dy = (__int128)(P[2 * i + 1] - P[2 * j + 1]);
        const unsigned __int128 q = (unsigned __int128)(dx * dx) + (unsigned __int128)(dy * dy);
        l[i][0] = (unsigned long long)q;
        l[i][1] = (unsigned long long)(q >> 64);
    }
#pragma unroll
    for (int i = 0; i < 6; i++) num[i] = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) den[i] = 0;
    if (!st) {   // edge P0 P1: its length is l[0]
        num[0] = l[0][0];
        num[1] = l[0][1];
        den[0] = 1;
        return;
    }
    unsigned long long t[4];
    mw_mul<2, 2>(l[0], l[1], t);
    mw_mul<4, 2>(t, l[2], num);
    const __int128 ar = (__int128)(P[2] - P[0]) * (P[5] - P[1]) - (__int128)(P[3] - P[1]) * (P[4] - P[0]);
    const unsigned __int128 aa = (unsigned __int128)(ar < 0 ? -ar : ar);
    const unsigned long long A[2] = {(unsigned long long)aa, (unsigned long long)(aa >> 64)};
    mw_mul<2, 2>(A, A, den);
}

// Exact sign of rho_a - rho_b in fixed-width integers (num_a den_b against num_b den_a, 10 limbs
// each); 2 when the differences need more than 58 bits (then the big integers decide).
__host__ __device__ __forceinline__ int rho_cmp_fixed(int sta, const double2 *pa, int stb, const double2 *pb)
{
    long long X[12];
    int D;
    if (!rho_pair_ints(sta, pa, stb, pb, X, &D) || D > 58) return 2;
    unsigned long long na[6], da[4], nb[6], db[4], l[10], r[10];
    rho_numden(sta, X, na, da);
    rho_numden(stb, X + 6, nb, db);
    mw_mul<6, 4>(na, db, l);
    mw_mul<6, 4>(nb, da, r);
    int c = 0;
#pragma unroll
    for (int i = 0; i < 10; i++)
        if (l[i] != r[i]) c = l[i] > r[i] ? 1 : -1;   // the highest differing limb decides
    return c;
}

// Knuth's algorithm D with 32-bit digits and fixed sizes: qd[0..8] = floor(U / V) for U = un[0..16]
// (un[16] = 0) and V = vn[0..7] with its top bit set; un is left holding the remainder.  Each
// quotient digit is estimated from the top two remainder digits, corrected at most twice,
// multiplied back, and added back on an overdraft.
__host__ __device__ __forceinline__ void mw_div_17_8(uint32_t *un, const uint32_t *vn, uint32_t *qd)
{
    const uint64_t b = 1ull << 32;
#pragma unroll
    for (int j = 8; j >= 0; j--) {
        const uint64_t top = ((uint64_t)un[j + 8] << 32) | un[j + 7];
        uint64_t qhat = top / vn[7], rhat = top - qhat * vn[7];
        while (qhat >= b || qhat * vn[6] > ((rhat << 32) | un[j + 6])) {
            qhat--;
            rhat += vn[7];
            if (rhat >= b) break;
        }
        int64_t tt = 0, kk = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint64_t pr = qhat * vn[i];
            tt = (int64_t)un[i + j] - kk - (int64_t)(pr & 0xffffffffull);
            un[i + j] = (uint32_t)tt;
            kk = (int64_t)(pr >> 32) - (tt >> 32);
        }
        tt = (int64_t)un[j + 8] - kk;
        un[j + 8] = (uint32_t)tt;
        qd[j] = (uint32_t)qhat;
        if (tt < 0) {
            qd[j]--;
            uint64_t c = 0;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const uint64_t s2 = (uint64_t)un[i + j] + vn[i] + c;
                un[i + j] = (uint32_t)s2;
                c = s2 >> 32;
            }
            un[j + 8] = (uint32_t)((uint64_t)un[j + 8] + c);
        }
    }
}

// The 256-bit key of an alpha value: q = floor(rho 2^-E) with q in [2^255, 2^256), a function of
// the value alone (equal values give equal keys) and monotone in it.  With num, den at the unit
// 2^t (rho = num / (4 den) 2^(2t)) and k = 256 - bitlen(num) + bitlen(den), floor(num 2^k / den)
// has 256 or 257 bits; E = 2t - 2 - k.  The division runs on U = num 2^(512 - Ln) (16 digits of 32
// bits) by V = den 2^(256 - Ld) (8 digits, top bit set): U / V = num 2^k / den.  Knuth's algorithm D
// with all sizes fixed: 9 quotient digits, each estimated from the top two remainder digits,
// corrected at most twice, multiplied back and added back on an overdraft.  False when the
// differences need more than 58 bits or the entity is degenerate (then the big integers decide).
__host__ __device__ __forceinline__ bool rho_key_fixed(int st, const double2 *p, unsigned long long *q4, int *E)
{
    double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int npt = st ? 3 : 2;
    for (int i = 0; i < npt; i++) { v[2 * i] = p[i].x; v[2 * i + 1] = p[i].y; }
    long long X[6];
    int B, t;
    if (!res_ints(v, 6, X, &B, 61, &t)) return false;   // (an edge's unused slots are 0)
    int D = 0;
    for (int i = 0; i < npt; i++) {
        const int j = i + 1 == npt ? 0 : i + 1;
        D = max(D, max(res_bits(X[2 * i] - X[2 * j]), res_bits(X[2 * i + 1] - X[2 * j + 1])));
    }
    if (D > 58) return false;
    unsigned long long num[8], den[4];
    rho_numden(st, X, num, den);
    num[6] = num[7] = 0;
    const int Ln = mw_bitlen<6>(num), Ld = mw_bitlen<4>(den);
    if (Ln == 0 || Ld == 0) return false;
    mw_shl<8>(num, 512 - Ln);
    mw_shl<4>(den, 256 - Ld);
    uint32_t un[17], vn[8], qd[9];
#pragma unroll
    for (int i = 0; i < 8; i++) { un[2 * i] = (uint32_t)num[i]; un[2 * i + 1] = (uint32_t)(num[i] >> 32); }
    un[16] = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) { vn[2 * i] = (uint32_t)den[i]; vn[2 * i + 1] = (uint32_t)(den[i] >> 32); }
    mw_div_17_8(un, vn, qd);
    int k = 256 - Ln + Ld;
    unsigned long long q[5];
#pragma unroll
    for (int w = 0; w < 4; w++) q[w] = (unsigned long long)qd[2 * w] | ((unsigned long long)qd[2 * w + 1] << 32);
    q[4] = qd[8];
    if (q[4] & 1ull) {   // 257 bits: one more halving (floor of a floor is the floor)
#pragma unroll
        for (int w = 0; w < 4; w++) q[w] = (q[w] >> 1) | (q[w + 1] << 63);
        k -= 1;
    }
#pragma unroll
    for (int w = 0; w < 4; w++) q4[w] = q[w];
    *E = 2 * t - 2 - k;
    return true;
}

// Exact incircle sign in fixed width: with the differences to d below 2^D (D <= 58), each term
// lift_k * minor_k is below 2^(4D + 2) <= 2^234 and the determinant below 2^(4D + 4) <= 2^236, so a
// 256-bit two's complement sum cannot overflow and its top bit is the sign.  The same determinant
// as incircle_exact; 2 when the differences need more bits (then the big integers decide).
__host__ __device__ __forceinline__ int incircle_fixed(double2 a, double2 b, double2 c, double2 d)
{
    const double v[8] = {a.x, a.y, b.x, b.y, c.x, c.y, d.x, d.y};
    long long X[8];
    int B;
    if (!res_ints(v, 8, X, &B, 61)) return 2;
    const long long dx[3] = {X[0] - X[6], X[2] - X[6], X[4] - X[6]};
    const long long dy[3] = {X[1] - X[7], X[3] - X[7], X[5] - X[7]};
    int D = 0;
    for (int k = 0; k < 3; k++) D = max(D, max(res_bits(dx[k]), res_bits(dy[k])));
    if (D > 58) return 2;
    unsigned long long acc[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int i = k == 2 ? 0 : k + 1, j = k == 0 ? 2 : k - 1;   // (k + 1) % 3, (k + 2) % 3
        const unsigned __int128 lf = (unsigned __int128)((__int128)dx[k] * dx[k]) +
                                     (unsigned __int128)((__int128)dy[k] * dy[k]);
        const __int128 mn = (__int128)dx[i] * dy[j] - (__int128)dx[j] * dy[i];
        const unsigned __int128 am = (unsigned __int128)(mn < 0 ? -mn : mn);
        const unsigned long long L[2] = {(unsigned long long)lf, (unsigned long long)(lf >> 64)};
        const unsigned long long M[2] = {(unsigned long long)am, (unsigned long long)(am >> 64)};
        unsigned long long P[4];
        mw_mul<2, 2>(L, M, P);
        if (mn < 0) {   // two's complement: invert and add one
            unsigned long long c = 1;
#pragma unroll
            for (int w = 0; w < 4; w++) {
                const unsigned long long x = ~P[w] + c;
                c = (c && x == 0) ? 1 : 0;
                P[w] = x;
            }
        }
        unsigned long long c = 0;
#pragma unroll
        for (int w = 0; w < 4; w++) {
            const unsigned __int128 sum = (unsigned __int128)acc[w] + P[w] + c;
            acc[w] = (unsigned long long)sum;
            c = (unsigned long long)(sum >> 64);
        }
    }
    if ((acc[0] | acc[1] | acc[2] | acc[3]) == 0) return 0;
    return (acc[3] >> 63) ? -1 : 1;
}

// Attachment in 128-bit integers: |i-j|^2 - |i+j-2k|^2 >= 0 (k in or on the circle with diameter
// ij) with the coordinates as integers below 2^60 (res_ints): its terms stay below 2^125.  1
// attached, 0 not, -1 when the coordinates need more bits (then the big integers decide).
__host__ __device__ __forceinline__ int attached_i128(double2 i, double2 j, double2 k)
{
    const double v[6] = {i.x, i.y, j.x, j.y, k.x, k.y};
    long long X[6];
    int B;
    if (!res_ints(v, 6, X, &B, 60)) return -1;
    const __int128 dx = X[0] - X[2], dy = X[1] - X[3];
    const __int128 sx = (__int128)X[0] + X[2] - 2 * (__int128)X[4], sy = (__int128)X[1] + X[3] - 2 * (__int128)X[5];
    return dx * dx + dy * dy >= sx * sx + sy * sy ? 1 : 0;
}

GA_INLINE int incircle_exact_sos(double2 a, int ia, double2 b, int ib, double2 c, int ic,
                                 double2 d, int id, int *ovf)
{
    // differences within 58 bits: the sign in fixed-width registers; beyond, a determinant the
    // residues show to be zero skips the big integers
    int s = incircle_fixed(a, b, c, d);
    if (s == 2) s = incircle_zero_res(a, b, c, d) == 1 ? 0 : incircle_exact(a, b, c, d, ovf);
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
