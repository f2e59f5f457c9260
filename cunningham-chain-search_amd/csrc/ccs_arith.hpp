// Exact modular and u128 arithmetic shared by the GPU kernels (ccs_kernels.hip) and the
// host confirmation engine (ccs_confirm.cpp): the same code decides primality on both
// sides.  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ccs_internal.hpp"

#define CCS_HD __host__ __device__ __forceinline__
#define CCS_HDN __host__ __device__ inline

namespace ccs {

// ---------------------------------------------------------------------------
// exact residues of small moduli
// ---------------------------------------------------------------------------

// y mod q for y < 2^52, q < 2^18: double reciprocal estimate, corrected by +-q
CCS_HD u32 mod52(u64 y, u32 q, double rcp) {
    u64 qt = (u64)((double)y * rcp);
    int64_t r = (int64_t)(y - qt * (u64)q);
    if (r < 0) r += q;
    if (r < 0) r += q;
    if (r >= (int64_t)q) r -= q;
    if (r >= (int64_t)q) r -= q;
    return (u32)r;
}

// (hi:lo) mod q via limb folding with c_k = 2^(32k) mod q
CCS_HD u32 mod128(u64 lo, u64 hi, const PrimeMod& pm) {
    u64 y = (lo & 0xffffffffull) + (lo >> 32) * (u64)pm.c32 + (hi & 0xffffffffull) * (u64)pm.c64 +
            (hi >> 32) * (u64)pm.c96;
    return mod52(y, pm.q, pm.rcp);
}

// high 64 bits of a 64x64-bit product
CCS_HD u64 umulhi64(u64 a, u64 b) { return (u64)(((unsigned __int128)a * b) >> 64); }

// ---------------------------------------------------------------------------
// u128 helpers
// ---------------------------------------------------------------------------

struct U2 {
    u64 lo, hi;
};

CCS_HD U2 mk(u64 lo, u64 hi) { U2 r; r.lo = lo; r.hi = hi; return r; }
CCS_HD bool ge(U2 a, U2 b) { return a.hi > b.hi || (a.hi == b.hi && a.lo >= b.lo); }
CCS_HD bool eq(U2 a, U2 b) { return a.hi == b.hi && a.lo == b.lo; }
CCS_HD bool iszero(U2 a) { return (a.lo | a.hi) == 0; }
CCS_HD U2 add(U2 a, U2 b) { U2 r; r.lo = a.lo + b.lo; r.hi = a.hi + b.hi + (r.lo < a.lo); return r; }
CCS_HD U2 sub(U2 a, U2 b) { U2 r; r.lo = a.lo - b.lo; r.hi = a.hi - b.hi - (a.lo < b.lo); return r; }
CCS_HD U2 shr1(U2 a) { return mk((a.lo >> 1) | (a.hi << 63), a.hi >> 1); }
CCS_HD U2 shl1(U2 a) { return mk(a.lo << 1, (a.hi << 1) | (a.lo >> 63)); }
CCS_HD int bitlen(U2 a) {
    return a.hi ? 128 - __builtin_clzll(a.hi) : (a.lo ? 64 - __builtin_clzll(a.lo) : 0);
}
CCS_HD bool bit(U2 a, int i) { return i >= 64 ? ((a.hi >> (i - 64)) & 1) : ((a.lo >> i) & 1); }
// low 128 bits of a*b
CCS_HD U2 mullo(U2 a, U2 b) {
    return mk(a.lo * b.lo, umulhi64(a.lo, b.lo) + a.lo * b.hi + a.hi * b.lo);
}
// a (u128) * b (u64), low 128 bits
CCS_HD U2 mul_u64(U2 a, u64 b) { return mk(a.lo * b, umulhi64(a.lo, b) + a.hi * b); }

// ---------------------------------------------------------------------------
// u128 Montgomery arithmetic, R = 2^128, n odd, n < 2^127
// ---------------------------------------------------------------------------

struct Mont {
    U2 n, ni;     // ni = -n^-1 mod 2^128
    U2 one;       // R mod n
    double ninv;  // 1/n in f64 (quotient estimates of mul_small)
};

CCS_HD void mont_init(Mont& c, U2 n) {
    c.n = n;
    U2 x = n;   // n*n = 1 (mod 8)
    const U2 two = mk(2, 0);
#pragma unroll
    for (int i = 0; i < 7; ++i) x = mullo(x, sub(two, mullo(n, x)));
    c.ni = sub(mk(0, 0), x);
    // R mod n: start from 2^(b-1) < n and double up to 2^128
    const int b = bitlen(n);
    U2 r = b - 1 >= 64 ? mk(0, 1ull << (b - 1 - 64)) : mk(1ull << (b - 1), 0);
    for (int i = b - 1; i < 128; ++i) {
        r = shl1(r);
        if (ge(r, n)) r = sub(r, n);
    }
    c.one = r;
    c.ninv = 1.0 / ((double)n.hi * 18446744073709551616.0 + (double)n.lo);
}

// a*b*2^-128 mod n by CIOS over four 32-bit limbs: each step adds a[i]*b, then m*n with
// m = t0 * (-n^-1 mod 2^32), and drops the low limb.  The products are 32x32 -> 64-bit
// multiply-adds (v_mad_u64_u32 on the GPU), 32 of them instead of the 58 multiplies that
// 64-bit limbs take; t stays below 2n < 2^128, so one final subtraction reduces it.
// L = 3 is the same product for n < 2^96: a, b < n then have a zero top limb too, and
// with the three top limbs written as constant zeros the compiler drops their products
// (21 multiply-adds); R stays 2^128, so both forms give the same value.
template <int L>
CCS_HD U2 mont_mul(const Mont& c, U2 a, U2 b) {
    const u32 A[4] = {(u32)a.lo, (u32)(a.lo >> 32), (u32)a.hi, L == 3 ? 0u : (u32)(a.hi >> 32)};
    const u32 B[4] = {(u32)b.lo, (u32)(b.lo >> 32), (u32)b.hi, L == 3 ? 0u : (u32)(b.hi >> 32)};
    const u32 N[4] = {(u32)c.n.lo, (u32)(c.n.lo >> 32), (u32)c.n.hi, L == 3 ? 0u : (u32)(c.n.hi >> 32)};
    const u32 ni = (u32)c.ni.lo;
    u32 t0 = 0, t1 = 0, t2 = 0, t3 = 0, t4 = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        u64 p = (u64)A[i] * B[0] + t0;
        t0 = (u32)p;
        u64 cc = p >> 32;
        p = (u64)A[i] * B[1] + t1 + cc;
        t1 = (u32)p;
        cc = p >> 32;
        p = (u64)A[i] * B[2] + t2 + cc;
        t2 = (u32)p;
        cc = p >> 32;
        p = (u64)A[i] * B[3] + t3 + cc;
        t3 = (u32)p;
        cc = p >> 32;
        u64 s = (u64)t4 + cc;
        t4 = (u32)s;
        const u32 t5 = (u32)(s >> 32);
        const u32 m = t0 * ni;
        p = (u64)m * N[0] + t0;   // low limb becomes 0
        cc = p >> 32;
        p = (u64)m * N[1] + t1 + cc;
        t0 = (u32)p;
        cc = p >> 32;
        p = (u64)m * N[2] + t2 + cc;
        t1 = (u32)p;
        cc = p >> 32;
        p = (u64)m * N[3] + t3 + cc;
        t2 = (u32)p;
        cc = p >> 32;
        s = (u64)t4 + cc;
        t3 = (u32)s;
        t4 = t5 + (u32)(s >> 32);
    }
    U2 r = mk((u64)t0 | ((u64)t1 << 32), (u64)t2 | ((u64)t3 << 32));
    if (t4 || ge(r, c.n)) r = sub(r, c.n);
    return r;
}

CCS_HD U2 madd(const Mont& c, U2 a, U2 b) { U2 r = add(a, b); if (ge(r, c.n)) r = sub(r, c.n); return r; }
CCS_HD U2 msub(const Mont& c, U2 a, U2 b) { return ge(a, b) ? sub(a, b) : add(sub(a, b), c.n); }
CCS_HD U2 mhalf(const Mont& c, U2 a) { if (a.lo & 1) a = add(a, c.n); return shr1(a); }

// x*m mod n for x < n and a small m (< 2^16); keeps the Montgomery form of x.  t = x*m has
// three limbs (< 2^143); its quotient by n (< 2^16) is estimated in f64 with relative error
// ~2^-50, so the estimate is exact or one off, and t - q*n is corrected once either way.
CCS_HD U2 mul_small(const Mont& c, U2 x, u32 m) {
    const u64 t0 = x.lo * m, c0 = umulhi64(x.lo, m);
    const u64 t1a = x.hi * m, c1 = umulhi64(x.hi, m);
    const u64 t1 = t1a + c0;
    const u64 t2 = c1 + (t1 < t1a);
    const double tf = ((double)t2 * 18446744073709551616.0 + (double)t1) * 18446744073709551616.0 + (double)t0;
    const u64 q = (u64)(u32)(tf * c.ninv);
    const u64 p0 = c.n.lo * q, pc0 = umulhi64(c.n.lo, q);
    const u64 p1a = c.n.hi * q, pc1 = umulhi64(c.n.hi, q);
    const u64 p1 = p1a + pc0;
    const u64 p2 = pc1 + (p1 < p1a);
    // r = t - q*n over three limbs
    const u64 r0 = t0 - p0, b0 = t0 < p0;
    const u64 r1 = t1 - p1 - b0, b1 = (t1 < p1) | ((t1 - p1) < b0);
    const u64 r2 = t2 - p2 - b1;
    U2 r = mk(r0, r1);
    if ((int64_t)r2 < 0) r = add(r, c.n);                      // q was one too large
    else if (r2 != 0 || ge(r, c.n)) r = sub(r, c.n);           // q was one too small
    return r;
}

// x*v mod n for a small signed v (|v| < 2^16)
CCS_HD U2 mul_small_signed(const Mont& c, U2 x, int v) {
    const U2 r = mul_small(c, x, (u32)(v < 0 ? -v : v));
    return v < 0 ? msub(c, mk(0, 0), r) : r;
}

// Montgomery form of a small constant x: x*R mod n by double-and-add of R mod n
CCS_HD U2 mont_small(const Mont& c, u32 x) {
    U2 r = mk(0, 0);
    for (int i = 31; i >= 0; --i) {
        r = madd(c, r, r);
        if ((x >> i) & 1) r = madd(c, r, c.one);
    }
    return r;
}

template <int L>
CCS_HD U2 mont_pow(const Mont& c, U2 bm, U2 e) {
    U2 r = c.one;
    for (int i = bitlen(e) - 1; i >= 0; --i) {
        r = mont_mul<L>(c, r, r);
        if (bit(e, i)) r = mont_mul<L>(c, r, bm);
    }
    return r;
}

// 2^e in Montgomery form: the multiply by the base of each set bit is a doubling (an add)
template <int L>
CCS_HD U2 mont_pow2(const Mont& c, U2 e) {
    U2 r = c.one;
    for (int i = bitlen(e) - 1; i >= 0; --i) {
        r = mont_mul<L>(c, r, r);
        if (bit(e, i)) r = madd(c, r, r);
    }
    return r;
}

template <int L>
CCS_HDN bool strong_mr(const Mont& c, u32 a) {
    U2 d = sub(c.n, mk(1, 0));
    int s = 0;
    while (!(d.lo & 1)) { d = shr1(d); ++s; }
    const U2 one = c.one, mone = sub(c.n, c.one);
    U2 x = a == 2 ? mont_pow2<L>(c, d) : mont_pow<L>(c, mont_small(c, a), d);
    if (eq(x, one) || eq(x, mone)) return true;
    for (int i = 1; i < s; ++i) {
        x = mont_mul<L>(c, x, x);
        if (eq(x, mone)) return true;
        if (eq(x, one)) return false;
    }
    return false;
}

// n mod m for a small odd m (< 2^16)
CCS_HD u32 mod_small(U2 n, u32 m) {
    PrimeMod pm;
    pm.q = m;
    u64 c32 = (0xffffffffull % m + 1) % m;
    pm.c32 = (u32)c32;
    pm.c64 = (u32)(c32 * c32 % m);
    pm.c96 = (u32)((u64)pm.c64 * c32 % m);
    pm.rcp = 1.0 / (double)m;
    return mod128(n.lo, n.hi, pm);
}

// Jacobi symbol (a/m), m odd positive, both small
CCS_HD int jacobi_small(u32 a, u32 m) {
    int res = 1;
    a %= m;
    while (a) {
        while (!(a & 1)) {
            a >>= 1;
            const u32 r8 = m & 7;
            if (r8 == 3 || r8 == 5) res = -res;
        }
        const u32 t = a; a = m; m = t;
        if ((a & 3) == 3 && (m & 3) == 3) res = -res;
        a %= m;
    }
    return m == 1 ? res : 0;
}

// (D/n) for odd D, |D| small, n odd large
CCS_HD int jacobi_D(int D, U2 n) {
    const u32 ad = (u32)(D < 0 ? -D : D);
    int res = 1;
    if (D < 0 && (n.lo & 3) == 3) res = -res;          // (-1/n)
    // (|D|/n) = (n/|D|) * (-1)^(((|D|-1)/2) * ((n-1)/2))
    if ((ad & 3) == 3 && (n.lo & 3) == 3) res = -res;
    return res * jacobi_small(mod_small(n, ad), ad);
}

// strong Lucas-Selfridge test (P = 1, Q = (1-D)/4, D = 5, -7, 9, ... first with (D/n) = -1).
// `ok` is cleared when no parameter was found in range (never happens for the chain
// values tested here: they are 2 mod 3, so never perfect squares).
template <int L>
CCS_HDN bool strong_lucas(const Mont& c, bool& ok) {
    int D = 5;
    int tries = 0;
    for (;;) {
        const int j = jacobi_D(D, c.n);
        if (j == -1) break;
        if (j == 0) {
            const u32 ad = (u32)(D < 0 ? -D : D);
            if (!(c.n.hi == 0 && c.n.lo == ad)) return false;
        }
        D = D > 0 ? -(D + 2) : -(D - 2);
        if (++tries > 200) { ok = false; return false; }
    }
    const int Q = (1 - D) / 4;
    // D*U and Q^k*Q multiply by small constants (mul_small, not a Montgomery product);
    // with Q = -1 (D = 5) Q^k is +-1, so it is squared to 1 and negated without products
    const bool qunit = Q == -1;
    const U2 mone = msub(c, mk(0, 0), c.one);
    U2 d = add(c.n, mk(1, 0));
    int s = 0;
    while (!(d.lo & 1)) { d = shr1(d); ++s; }
    U2 U = c.one, V = c.one, Qk = qunit ? mone : mul_small_signed(c, c.one, Q);
    for (int i = bitlen(d) - 2; i >= 0; --i) {
        U = mont_mul<L>(c, U, V);
        V = msub(c, mont_mul<L>(c, V, V), madd(c, Qk, Qk));
        Qk = qunit ? c.one : mont_mul<L>(c, Qk, Qk);
        if (bit(d, i)) {
            const U2 nu = mhalf(c, madd(c, U, V));
            const U2 nv = mhalf(c, madd(c, mul_small_signed(c, U, D), V));
            U = nu;
            V = nv;
            Qk = qunit ? msub(c, mk(0, 0), Qk) : mul_small_signed(c, Qk, Q);
        }
    }
    if (iszero(U) || iszero(V)) return true;
    for (int r = 1; r < s; ++r) {
        V = msub(c, mont_mul<L>(c, V, V), madd(c, Qk, Qk));
        if (iszero(V)) return true;
        Qk = qunit ? c.one : mont_mul<L>(c, Qk, Qk);
    }
    return false;
}

// n < 2^96, so the three-limb products apply.  On the GPU the choice is made for the whole
// wave (its active lanes): one lane above 2^96 sends the wave to the four-limb products,
// so lanes of one wave never run both forms one after the other.
CCS_HD bool below96(U2 n) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __all((n.hi >> 32) == 0);
#else
    return (n.hi >> 32) == 0;
#endif
}

// Baillie-PSW in two parts, for callers that gather the Lucas tests of many values:
// bpsw_pre divides by the primes 2..61 (small18: their PrimeMod records) and runs the
// strong MR to base 2, returning 0 (composite), 1 (prime: n is small) or 2 (a base-2
// strong probable prime, decided by bpsw_lucas).
CCS_HDN int bpsw_pre(U2 n, const PrimeMod* small18) {
    if (n.hi == 0 && n.lo < 2) return 0;
#pragma unroll 1
    for (int i = 0; i < 18; ++i) {
        const u32 p = small18[i].q;
        if (n.hi == 0 && n.lo == p) return 1;
        if (mod128(n.lo, n.hi, small18[i]) == 0) return 0;
    }
    if (n.hi == 0 && n.lo < 61ull * 61ull) return 1;
    Mont c;
    mont_init(c, n);
    return (below96(n) ? strong_mr<3>(c, 2) : strong_mr<4>(c, 2)) ? 2 : 0;
}

CCS_HDN bool bpsw_lucas(U2 n, bool& ok) {
    Mont c;
    mont_init(c, n);
    return below96(n) ? strong_lucas<3>(c, ok) : strong_lucas<4>(c, ok);
}

// Baillie-PSW: the decision of GMP's mpz_probab_prime_p(n, 1) (v13.cu:1841)
CCS_HDN bool is_prime_bpsw(U2 n, const PrimeMod* small18, bool& ok) {
    const int pre = bpsw_pre(n, small18);
    return pre == 1 || (pre == 2 && bpsw_lucas(n, ok));
}

// strong Miller-Rabin over the given bases
CCS_HDN bool is_prime_mr(U2 n, const u32* bases, int n_bases) {
    if (n.hi == 0 && n.lo < 2) return false;
    if (n.hi == 0 && n.lo < 4) return true;
    if (!(n.lo & 1)) return false;
    Mont c;
    mont_init(c, n);
    const bool small = below96(n);
    for (int i = 0; i < n_bases; ++i) {
        const u32 b = bases[i];
        if (n.hi == 0 && (u64)b % n.lo == 0) continue;
        if (!(small ? strong_mr<3>(c, b) : strong_mr<4>(c, b))) return false;
    }
    return true;
}


}  // namespace ccs
