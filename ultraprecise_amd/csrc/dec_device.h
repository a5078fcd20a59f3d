// dec_device.h -- decimal limb arithmetic for gfx950 (CDNA4), header-only.
//
// Compiled twice: by hipcc into the static kernels of libupdec.so, and by
// hipRTC into every per-query expression module (the text of this file is
// embedded in the library and handed to hipRTC as a header).
//
// Values are sign-magnitude, as the reference's Decimal
// (src/datatypes/decimal.hxx:130-141): a base-2^32 little-endian magnitude
// plus a sign flag.  Unlike the reference's fixed 32-limb array, every value
// here has a limb count fixed at code-generation time from a bit-width bound,
// so one lane holds exactly the limbs its value can need and all limb
// indices are compile-time constants (the arrays live in VGPRs, never in
// scratch).  Scales are compile-time too: alignment is a multiply by a
// constant 10^n row.
#ifndef UPDEC_DEC_DEVICE_H
#define UPDEC_DEC_DEVICE_H

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

#define UPD_DEV __device__ __forceinline__

namespace updec {

typedef unsigned char u8;
typedef unsigned int u32;
typedef int i32;
typedef unsigned long long u64;
typedef long long i64;
typedef __SIZE_TYPE__ usize;

template <int W> struct Mag { u32 w[W]; };
template <int W> struct Dec { Mag<W> m; u32 neg; };

// ---------------------------------------------------------------- constants

// Exact 10^n as 32-bit words, built at compile time (n <= 460).
struct BigConst {
    u32 w[48];
    int n;
    constexpr BigConst(int p) : w(), n(0)
    {
        w[0] = 1;
        for (int k = 0; k < p; k++) {
            u64 c = 0;
            for (int i = 0; i < 48; i++) {
                u64 t = (u64)w[i] * 10u + c;
                w[i] = (u32)t;
                c = t >> 32;
            }
        }
        int t = 47;
        while (t > 0 && w[t] == 0) t--;
        n = t + 1;
    }
    constexpr int bits() const
    {
        int b = 0;
        u32 x = w[n - 1];
        while (x) { b++; x >>= 1; }
        return 32 * (n - 1) + b;
    }
};

template <int N> struct P10 {
    static constexpr BigConst c = BigConst(N);
    static constexpr int W = c.n;
    static constexpr int BITS = c.bits();
};

// The reference's estimate of bitlen(10^shift) used to pick the uint64
// division path (decimal.cu:2252-2256).
constexpr int est_shift_bits(int s) { return (s / 3) * 10 + (s % 3 == 1 ? 4 : (s % 3 == 2 ? 7 : 0)); }

// ---------------------------------------------------------------- magnitude helpers

template <int WO, int WI> UPD_DEV Mag<WO> resize(const Mag<WI>& a)
{
    Mag<WO> r;
#pragma unroll
    for (int i = 0; i < WO; i++) r.w[i] = (i < WI) ? a.w[i] : 0u;
    return r;
}

template <int W> UPD_DEV Mag<W> mzero()
{
    Mag<W> r;
#pragma unroll
    for (int i = 0; i < W; i++) r.w[i] = 0u;
    return r;
}

template <int W> UPD_DEV bool is_zero(const Mag<W>& a)
{
    u32 o = 0;
#pragma unroll
    for (int i = 0; i < W; i++) o |= a.w[i];
    return o == 0;
}

template <int W> UPD_DEV u32 bitlen(const Mag<W>& a)
{
    u32 b = 0;
#pragma unroll
    for (int i = 0; i < W; i++) b = a.w[i] ? (u32)(32 * i + 32 - __builtin_clz(a.w[i])) : b;
    return b;
}

// r = a + b, returns carry out
template <int W> UPD_DEV u32 add_to(Mag<W>& r, const Mag<W>& a, const Mag<W>& b)
{
    u64 c = 0;
#pragma unroll
    for (int i = 0; i < W; i++) {
        u64 t = (u64)a.w[i] + b.w[i] + c;
        r.w[i] = (u32)t;
        c = t >> 32;
    }
    return (u32)c;
}

// r = a - b, returns borrow out (1 when a < b)
template <int W> UPD_DEV u32 sub_to(Mag<W>& r, const Mag<W>& a, const Mag<W>& b)
{
    u64 br = 0;
#pragma unroll
    for (int i = 0; i < W; i++) {
        u64 t = (u64)a.w[i] - b.w[i] - br;
        r.w[i] = (u32)t;
        br = (t >> 63) & 1u;
    }
    return (u32)br;
}

// -1, 0, 1
template <int W> UPD_DEV int cmp(const Mag<W>& a, const Mag<W>& b)
{
    int c = 0;
#pragma unroll
    for (int i = 0; i < W; i++) c = (a.w[i] != b.w[i]) ? (a.w[i] > b.w[i] ? 1 : -1) : c;
    return c;
}

// low WO limbs of a * b (schoolbook, 32x32->64 multiply-accumulate)
template <int WO, int WA, int WB> UPD_DEV Mag<WO> mul_lo(const Mag<WA>& a, const Mag<WB>& b)
{
    Mag<WO> r = mzero<WO>();
#pragma unroll
    for (int i = 0; i < WA; i++) {
        if (i >= WO) break;
        u64 c = 0;
#pragma unroll
        for (int j = 0; j < WB; j++) {
            if (i + j < WO) {
                u64 t = (u64)a.w[i] * b.w[j] + r.w[i + j] + c;
                r.w[i + j] = (u32)t;
                c = t >> 32;
            }
        }
        if (i + WB < WO) r.w[i + WB] = (u32)c;
    }
    return r;
}

// low WO limbs of a * 10^N (constant row, zero limbs of 10^N skipped)
template <int WO, int WA, int N> UPD_DEV Mag<WO> mul_p10(const Mag<WA>& a)
{
    if constexpr (N == 0) {
        return resize<WO>(a);
    } else {
        constexpr BigConst p = P10<N>::c;
        constexpr int WP = P10<N>::W;
        Mag<WO> r = mzero<WO>();
#pragma unroll
        for (int i = 0; i < WA; i++) {
            if (i >= WO) break;
            u64 c = 0;
#pragma unroll
            for (int j = 0; j < WP; j++) {
                if (i + j < WO) {
                    u64 t = (u64)a.w[i] * p.w[j] + r.w[i + j] + c;
                    r.w[i + j] = (u32)t;
                    c = t >> 32;
                }
            }
            if (i + WP < WO) r.w[i + WP] = (u32)c;
        }
        return r;
    }
}

// shift left by k bits (0 <= k < 32*W), bits shifted past the top are lost.
// Word moves are a log-depth select network so all indices stay static.
template <int W> UPD_DEV Mag<W> shl_var(const Mag<W>& a, u32 k)
{
    u32 ws = k >> 5, bs = k & 31u;
    Mag<W> t = a;
#pragma unroll
    for (int st = 1; st < W; st <<= 1) {
        bool on = (ws & (u32)st) != 0;
        Mag<W> u;
#pragma unroll
        for (int i = 0; i < W; i++) u.w[i] = (i >= st) ? t.w[i - st] : 0u;
#pragma unroll
        for (int i = 0; i < W; i++) t.w[i] = on ? u.w[i] : t.w[i];
    }
    Mag<W> r;
#pragma unroll
    for (int i = 0; i < W; i++) {
        u64 hi = t.w[i], lo = (i > 0) ? t.w[i - 1] : 0u;
        r.w[i] = (u32)((((hi << 32) | lo) << bs) >> 32);
    }
    return r;
}

// shift right by k bits (0 <= k < 32*W)
template <int W> UPD_DEV Mag<W> shr_var(const Mag<W>& a, u32 k)
{
    u32 ws = k >> 5, bs = k & 31u;
    Mag<W> t = a;
#pragma unroll
    for (int st = 1; st < W; st <<= 1) {
        bool on = (ws & (u32)st) != 0;
        Mag<W> u;
#pragma unroll
        for (int i = 0; i < W; i++) u.w[i] = (i + st < W) ? t.w[i + st] : 0u;
#pragma unroll
        for (int i = 0; i < W; i++) t.w[i] = on ? u.w[i] : t.w[i];
    }
    Mag<W> r;
#pragma unroll
    for (int i = 0; i < W; i++) {
        u64 lo = t.w[i], hi = (i + 1 < W) ? t.w[i + 1] : 0u;
        r.w[i] = (u32)(((hi << 32) | lo) >> bs);
    }
    return r;
}

// Reciprocal of a normalized divisor word d (top bit set):
// v = floor((2^64 - 1) / d) - 2^32  (Moller & Granlund, "Improved division
// by invariant integers", IEEE Trans. Computers 60(2), 2011, section 3).
// A correctly rounded double quotient 2^64 / d (in (2^32, 2^33], relative
// error below 2^-53, so within 2^-20 of the true value, and 2^64 / d exceeds
// (2^64 - 1) / d by less than 2^-31) truncates to V - 1, V or V + 1 for
// V = floor((2^64 - 1) / d); the remainder 2^64 - 1 - c d, formed modulo
// 2^64, is negative (top bit set: it lies in [-d, 0)) for V + 1 and at least
// d for V - 1.  This replaces a 64-bit software division.
UPD_DEV u32 recip_word(u32 d)
{
    u64 c = (u64)(18446744073709551616.0 / (double)d);
    const u64 r = ~0ull - c * (u64)d;
    c = (i64)r < 0 ? c - 1u : (r >= (u64)d ? c + 1u : c);
    return (u32)(c - (1ull << 32));
}

// 2-by-1 division with a precomputed reciprocal (Moller & Granlund 2011,
// Algorithm 4).  Requires u1 < d, d normalized.  q = floor((u1*2^32+u0)/d).
UPD_DEV u32 div_2by1(u32 u1, u32 u0, u32 d, u32 v, u32& r)
{
    u64 q = (u64)v * u1;
    q += ((u64)(u1 + 1u) << 32) | u0;
    u32 q1 = (u32)(q >> 32), q0 = (u32)q;
    u32 rr = u0 - q1 * d;
    if (rr > q0) { q1 -= 1u; rr += d; }
    if (rr >= d) { q1 += 1u; rr -= d; }
    r = rr;
    return q1;
}

// Q = floor(X / B), R = X mod B, for B != 0.  Knuth, TAOCP vol. 2, 4.3.1,
// Algorithm D, with the divisor normalized to fill all NB limbs (so the
// quotient loop has a static trip count of NX+1 digits) and the digit
// estimate from div_2by1 refined by the second divisor limb.
// One quotient digit of Knuth's Algorithm D (TAOCP 4.3.1): digit J of the
// quotient from the window un[J .. J+NB] and the normalized divisor vn.
template <int J, int NB, int NU>
UPD_DEV u32 divmod_digit(Mag<NU>& un, const Mag<NB>& vn, u32 d1, u32 d0, u32 v)
{
    const u32 u2 = un.w[J + NB], u1 = un.w[J + NB - 1];
    // window below the divisor's top limb: the digit is 0 and nothing
    // changes (the top digits of a dividend narrower than its static width;
    // a wave-uniform branch in practice)
    if (u2 == 0u && u1 < d1) return 0u;
    u32 qhat;
    u64 rhat;
    if (u2 >= d1) {
        qhat = 0xffffffffu;
        rhat = (u64)u1 + d1;
    } else {
        u32 r32;
        qhat = div_2by1(u2, u1, d1, v, r32);
        rhat = r32;
    }
    if constexpr (NB >= 2) {
        const u32 u0 = un.w[J + NB - 2];
#pragma unroll
        for (int it = 0; it < 2; it++) {
            bool fix = (rhat < (1ull << 32)) && ((u64)qhat * d0 > ((rhat << 32) | u0));
            qhat -= fix ? 1u : 0u;
            rhat += fix ? d1 : 0u;
        }
    }
    // un[J .. J+NB] -= qhat * vn.  A zero digit changes nothing: the top
    // digits of a dividend shorter than its static width are all zero, and
    // skipping them (a wave-uniform branch when every lane's digit is zero)
    // removes most of the work of a wide division
    if (qhat != 0u) {
        u64 kk = 0;
        i64 t;
#pragma unroll
        for (int i = 0; i < NB; i++) {
            u64 p = (u64)qhat * vn.w[i];
            t = (i64)un.w[i + J] - (i64)kk - (i64)(p & 0xffffffffull);
            un.w[i + J] = (u32)t;
            kk = (p >> 32) - (u64)(t >> 32);
        }
        t = (i64)un.w[J + NB] - (i64)kk;
        un.w[J + NB] = (u32)t;
        // rare add-back when the estimate was one too large
        if (t < 0) {
            qhat -= 1u;
            u64 c = 0;
#pragma unroll
            for (int i = 0; i < NB; i++) {
                u64 s = (u64)un.w[i + J] + vn.w[i] + c;
                un.w[i + J] = (u32)s;
                c = s >> 32;
            }
            un.w[J + NB] += (u32)c;
        }
    }
    return qhat;
}

// digits J, J-1, ..., 0, each its own instantiation: every limb index is a
// compile-time constant whatever the unroller decides, so the window stays
// in registers (a loop over J that the compiler leaves rolled indexes the
// window at run time and moves it to scratch memory)
template <int J, int NB, int NU, int NQ>
UPD_DEV void divmod_digits(Mag<NU>& un, const Mag<NB>& vn, u32 d1, u32 d0, u32 v, Mag<NQ>& Q)
{
    Q.w[J] = divmod_digit<J, NB, NU>(un, vn, d1, d0, v);
    if constexpr (J > 0) divmod_digits<J - 1, NB, NU, NQ>(un, vn, d1, d0, v, Q);
}

template <int NX, int NB> UPD_DEV void divmod(const Mag<NX>& X, const Mag<NB>& B, Mag<NX + 1>& Q, Mag<NB>& R)
{
    constexpr int NU = NX + NB + 1;
    const u32 k = 32u * NB - bitlen(B);
    const Mag<NB> vn = shl_var<NB>(B, k);
    Mag<NU> un = shl_var<NU>(resize<NU>(X), k);
    const u32 d1 = vn.w[NB - 1];
    const u32 d0 = (NB >= 2) ? vn.w[NB - 2] : 0u;
    const u32 v = recip_word(d1);
    divmod_digits<NX, NB, NU, NX + 1>(un, vn, d1, d0, v, Q);
    R = shr_var<NB>(resize<NB>(un), k);
}

// ---------------------------------------------------------------- signed operations

template <int WO, int WI> UPD_DEV Dec<WO> dresize(const Dec<WI>& a)
{
    Dec<WO> r;
    r.m = resize<WO>(a.m);
    r.neg = a.neg;
    return r;
}

// unary minus flips the sign, also of zero (decimal.cu:578-586)
template <int W> UPD_DEV Dec<W> dneg(const Dec<W>& a)
{
    Dec<W> r = a;
    r.neg = a.neg ^ 1u;
    return r;
}

// a * 10^N at the same sign (scale alignment, decimal.cu:2905-2920)
template <int WO, int WI, int N> UPD_DEV Dec<WO> dscale(const Dec<WI>& a)
{
    Dec<WO> r;
    r.m = mul_p10<WO, WI, N>(a.m);
    r.neg = a.neg;
    return r;
}

// ---- truncating division by 10^N (operator>>, decimal.cu:2003-2026)
//
// floor(a / 10^N) = floor(floor(a / 2^N) / 5^N): a static bit shift, then
// passes of one-word division by 5^13 (the largest power of five below
// 2^31) with a compile-time Moller-Granlund reciprocal.  Each pass drops at
// least 30 bits from the static bound, so later passes run on fewer limbs.

constexpr u32 pow5_u32(int k) { return k == 0 ? 1u : 5u * pow5_u32(k - 1); }
constexpr int clz_u32(u32 x) { return x >> 31 ? 0 : 1 + clz_u32(x << 1); }
constexpr int floor_log2_u32(u32 x) { return 31 - clz_u32(x); }

template <u32 D> struct WordDivisor {
    static constexpr int S = clz_u32(D);
    static constexpr u32 DN = D << S;
    static constexpr u32 V = (u32)(~0ull / DN - (1ull << 32));
};

// (r * 2^32 + w) / D with r < D: returns the quotient word, r becomes the remainder
template <u32 D> UPD_DEV u32 div_word_const(u32& r, u32 w)
{
    using C = WordDivisor<D>;
    const u32 u1 = C::S ? ((r << C::S) | (w >> ((32 - C::S) & 31))) : r;
    const u32 u0 = w << C::S;
    u32 rr;
    const u32 q = div_2by1(u1, u0, C::DN, C::V, rr);
    r = rr >> C::S;
    return q;
}

// floor(a / 2^N) for a static N
template <int W, int N> UPD_DEV Mag<W> shr_static(const Mag<W>& a)
{
    Mag<W> r;
    constexpr int WS = N / 32, BS = N % 32;
#pragma unroll
    for (int i = 0; i < W; i++) {
        const u32 lo = i + WS < W ? a.w[i + WS] : 0u;
        const u32 hi = i + WS + 1 < W ? a.w[i + WS + 1] : 0u;
        r.w[i] = BS ? ((lo >> BS) | (hi << ((32 - BS) & 31))) : lo;
    }
    return r;
}

// one pass: x = floor(x / D)
template <int W, u32 D> UPD_DEV void div_pass(Mag<W>& x)
{
    u32 r = 0;
#pragma unroll
    for (int i = W - 1; i >= 0; i--) x.w[i] = div_word_const<D>(r, x.w[i]);
}

// floor(a / 5^N), a < 2^(32 W), by passes of one-word division.  The full
// 5^13 passes run as a loop (one copy of the unrolled limb body keeps
// generated code, and the hipRTC compile time of long expressions, small);
// the last partial pass follows.
template <int W, int N> UPD_DEV Mag<W> div_pow5_passes(Mag<W> x)
{
#pragma unroll 1
    for (int p = 0; p < N / 13; p++) div_pass<W, pow5_u32(13)>(x);
    if constexpr (N % 13 != 0) div_pass<W, pow5_u32(N % 13)>(x);
    return x;
}

// 5^n as 32-bit words, built at compile time (n <= 660)
struct Pow5Const {
    u32 w[48];
    int n;
    constexpr Pow5Const(int p) : w(), n(0)
    {
        w[0] = 1;
        for (int k = 0; k < p; k++) {
            u64 c = 0;
            for (int i = 0; i < 48; i++) {
                u64 t = (u64)w[i] * 5u + c;
                w[i] = (u32)t;
                c = t >> 32;
            }
        }
        int t = 47;
        while (t > 0 && w[t] == 0) t--;
        n = t + 1;
    }
    constexpr int bits() const
    {
        int b = 0;
        u32 x = w[n - 1];
        while (x) { b++; x >>= 1; }
        return 32 * (n - 1) + b;
    }
};

template <int N> struct P5 {
    static constexpr Pow5Const c = Pow5Const(N);
    static constexpr int W = c.n;
    static constexpr int BITS = c.bits();
};

// mu = floor(2^T / 5^N), the Barrett multiplier, by restoring division one
// bit at a time at compile time (T < 32 * 64)
struct Pow5Recip {
    u32 w[64];
    constexpr Pow5Recip(int T, int N) : w()
    {
        const Pow5Const d(N);
        const int n = d.n + 1;  // limbs of the running remainder (< 2 * 5^N)
        u32 r[49] = {};
        for (int i = T; i >= 0; i--) {
            // r = 2 r + (bit i of 2^T)
            u32 c = i == T ? 1u : 0u;
            for (int k = 0; k < n; k++) {
                const u32 t = (r[k] << 1) | c;
                c = r[k] >> 31;
                r[k] = t;
            }
            int ge = 1;
            for (int k = n - 1; k >= 0; k--) {
                const u32 dk = k < d.n ? d.w[k] : 0u;
                if (r[k] != dk) {
                    ge = r[k] > dk;
                    break;
                }
            }
            if (ge) {
                u64 b = 0;
                for (int k = 0; k < n; k++) {
                    const u64 dk = k < d.n ? d.w[k] : 0u;
                    const u64 t = (u64)r[k] - dk - b;
                    r[k] = (u32)t;
                    b = (t >> 32) & 1u;
                }
                w[i >> 5] |= 1u << (i & 31);
            }
        }
    }
};

template <int T, int N> struct P5Recip {
    static constexpr Pow5Recip c = Pow5Recip(T, N);
};

// floor(x / 5^N) for x < 2^BI by Barrett reduction (Barrett, CRYPTO '86;
// Menezes et al., Handbook of Applied Cryptography, 14.42) with a
// compile-time multiplier: with 2^S <= 5^N < 2^(S+1), E = BI - S,
// a = floor(x / 2^S) and mu = floor(2^BI / 5^N), both below 2^E, the estimate
// q' = floor(a mu / 2^E) satisfies q - 2 <= q' <= q, since
// a mu > (x / 2^S - 1)(2^BI / 5^N - 1) > 2^E (x / 5^N - 2).  The remainder
// x - q' 5^N < 3 * 5^N is formed in its low limbs only, and at most two
// conditional subtractions of 5^N finish the quotient.
template <int WM, int BI, int N> UPD_DEV Mag<WM> div_pow5_barrett(const Mag<WM>& x)
{
    constexpr int S = P5<N>::BITS - 1;
    if constexpr (BI <= S) {
        return mzero<WM>();
    } else {
        constexpr int E = BI - S;
        constexpr int LQ = (E + 31) / 32;
        constexpr Pow5Recip mu = P5Recip<BI, N>::c;
        constexpr Pow5Const d = P5<N>::c;
        constexpr int WD = P5<N>::W;
        const Mag<LQ> a = resize<LQ>(shr_static<WM, S>(x));
        // a * mu, then the bits from E up
        Mag<2 * LQ> p = mzero<2 * LQ>();
#pragma unroll
        for (int i = 0; i < LQ; i++) {
            u64 c = 0;
#pragma unroll
            for (int j = 0; j < LQ; j++) {
                const u64 t = (u64)a.w[i] * mu.w[j] + p.w[i + j] + c;
                p.w[i + j] = (u32)t;
                c = t >> 32;
            }
            p.w[i + LQ] = (u32)c;
        }
        Mag<LQ> q = resize<LQ>(shr_static<2 * LQ, E>(p));
        // remainder in WR limbs (it is below 3 * 5^N < 2^(S + 3))
        constexpr int WR0 = (S + 3 + 31) / 32;
        constexpr int WR = WR0 < WM ? WR0 : WM;
        Mag<WR> qd = mzero<WR>();
#pragma unroll
        for (int i = 0; i < LQ; i++) {
            if (i >= WR) break;
            u64 c = 0;
#pragma unroll
            for (int j = 0; j < WD; j++) {
                if (i + j < WR) {
                    const u64 t = (u64)q.w[i] * d.w[j] + qd.w[i + j] + c;
                    qd.w[i + j] = (u32)t;
                    c = t >> 32;
                }
            }
            if (i + WD < WR) qd.w[i + WD] = (u32)c;
        }
        Mag<WR> r;
        {
            u64 b = 0;
#pragma unroll
            for (int k = 0; k < WR; k++) {
                const u64 t = (u64)x.w[k] - qd.w[k] - b;
                r.w[k] = (u32)t;
                b = (t >> 32) & 1u;
            }
        }
#pragma unroll
        for (int pass = 0; pass < 2; pass++) {
            // r >= 5^N: subtract it and count one more
            u64 b = 0;
            Mag<WR> t;
#pragma unroll
            for (int k = 0; k < WR; k++) {
                const u64 v = (u64)r.w[k] - (k < WD ? d.w[k] : 0u) - b;
                t.w[k] = (u32)v;
                b = (v >> 32) & 1u;
            }
            const bool ge = b == 0;
#pragma unroll
            for (int k = 0; k < WR; k++) r.w[k] = ge ? t.w[k] : r.w[k];
            u32 cy = ge ? 1u : 0u;
#pragma unroll
            for (int k = 0; k < LQ; k++) {
                const u32 s = q.w[k] + cy;
                cy = s < cy ? 1u : 0u;
                q.w[k] = s;
            }
        }
        return resize<WM>(q);
    }
}

// instructions of the two forms, roughly: a pass costs about 12 per limb; a
// Barrett reduction about 3 per limb product (LQ^2 for the estimate, about
// WR * WD for the remainder) plus the corrections
constexpr bool barrett_pays(int W, int BI, int N)
{
    if (N <= 13 || BI >= 32 * 64) return false;
    const int S = Pow5Const(N).bits() - 1;
    if (BI <= S) return true;
    const int LQ = (BI - S + 31) / 32, WD = Pow5Const(N).n;
    const int WR = (S + 34) / 32 < W ? (S + 34) / 32 : W;
    return 3 * (LQ * LQ + WR * WD) + 8 * WR + 4 * LQ < 12 * W * ((N + 12) / 13);
}

// floor(a / 5^N) for a < 2^BI (a held in W limbs): Barrett reduction or
// passes of one-word division, whichever issues fewer instructions
template <int W, int BI, int N> UPD_DEV Mag<W> div_pow5(const Mag<W>& x)
{
    if constexpr (barrett_pays(W, BI, N)) return div_pow5_barrett<W, BI, N>(x);
    else return div_pow5_passes<W, N>(x);
}

// cast to a smaller scale (decimal.cu:471-533, scale-down branch): the
// magnitude is floor-divided by 10^N and the sign is kept, so a negative
// value that truncates to zero stays -0
template <int WO, int WI, int N> UPD_DEV Dec<WO> dshr10(const Dec<WI>& a)
{
    constexpr int BI = 32 * WI - N;
    constexpr int WM = BI <= 32 ? 1 : (BI + 31) / 32;
    const Mag<WM> t = resize<WM>(shr_static<WI, N>(a.m));
    Dec<WO> r;
    r.m = resize<WO>(div_pow5<WM, (BI < 1 ? 1 : BI), N>(t));
    r.neg = a.neg;
    return r;
}

// truncate(p) for p <= 0 when the argument's type has no more than -p
// integer digits (decimal.cu:561-565): the magnitude is zeroed, the sign kept
template <int WO, int WI> UPD_DEV Dec<WO> dtrunc_zero(const Dec<WI>& a)
{
    Dec<WO> r;
    r.m = mzero<WO>();
    r.neg = a.neg;
    return r;
}

// ---- run-time scale variants (the precompiled interpreter, where the
// shift counts are data).  Same results as the compile-time forms above;
// every value is held at one register width W.

// 10^n for n <= 9
UPD_DEV u32 pow10_word(u32 n)
{
    u32 p = 1u;
    for (u32 k = 0; k < n; k++) p *= 10u;
    return p;
}

// a = a * m mod 2^(32 W)
template <int W> UPD_DEV void mul_word(Mag<W>& a, u32 m)
{
    u64 c = 0;
#pragma unroll
    for (int i = 0; i < W; i++) {
        u64 t = (u64)a.w[i] * m + c;
        a.w[i] = (u32)t;
        c = t >> 32;
    }
}

// a * 10^n mod 2^(32 W), n at run time (chunks of 10^9: the product mod
// 2^(32 W) does not depend on how 10^n is split)
template <int W> UPD_DEV Mag<W> mul_p10_rt(Mag<W> a, u32 n)
{
    for (; n >= 9u; n -= 9u) mul_word<W>(a, 1000000000u);
    if (n) mul_word<W>(a, pow10_word(n));
    return a;
}

// x = floor(x / 10^r), 1 <= r <= 8, with a run-time divisor word
template <int W> UPD_DEV void div_small_rt(Mag<W>& x, u32 r)
{
    const u32 d = pow10_word(r);
    const int s = __builtin_clz(d);
    const u32 dn = d << s;
    const u32 v = recip_word(dn);
    u32 rem = 0;
#pragma unroll
    for (int i = W - 1; i >= 0; i--) {
        const u32 u1 = (rem << s) | (x.w[i] >> ((32 - s) & 31));
        const u32 u0 = x.w[i] << s;
        u32 rr;
        x.w[i] = div_2by1(u1, u0, dn, v, rr);
        rem = rr >> s;
    }
}

// floor(a / 10^n), n at run time
template <int W> UPD_DEV Mag<W> div_p10_rt(Mag<W> a, u32 n)
{
    for (; n >= 9u; n -= 9u) div_pass<W, 1000000000u>(a);
    if (n) div_small_rt<W>(a, n);
    return a;
}

// DivOrMod at a run-time S (see ddiv).  The caller guarantees that
// |a| * 10^S fits in W limbs.
template <int W> UPD_DEV Dec<W> ddiv_rt(const Dec<W>& a, const Dec<W>& b, u32 S, bool faith)
{
    const bool azero = is_zero(a.m);
    const bool bzero = is_zero(b.m);
    Mag<W> bs = b.m;
    if (bzero) bs.w[0] = 1u;
    const Mag<W> X = mul_p10_rt<W>(a.m, S);
    Mag<W + 1> Q;
    Mag<W> R;
    divmod<W, W>(X, bs, Q, R);
    Mag<W> T;
    sub_to<W>(T, bs, R);
    u32 rnd = cmp<W>(R, T) >= 0 ? 1u : 0u;
    if (faith) {
        const u32 bla = bitlen(a.m), blb = bitlen(bs);
        const bool path_i = (bla + (u32)est_shift_bits((int)S) - 1u <= 64u) && blb <= 64u;
        bool top0 = true;
#pragma unroll
        for (int i = 1; i < W; i++) top0 = top0 && bs.w[i] == 0u;
        const bool path_ii = !path_i && top0;
        const u32 blx = bitlen(X);
        const bool path_iii_one = !path_i && !path_ii && blx < blb && blb - blx <= 64u;
        if (path_ii) rnd = ((u32)(R.w[0] * 2u) >= bs.w[0]) ? 1u : 0u;
        if (path_iii_one) {
            Q = mzero<W + 1>();
            Q.w[0] = 1u;
            rnd = 0u;
        }
        if (path_i) {
            Mag<2> dvs = mul_p10_rt<2>(resize<2>(a.m), S);
            Mag<2> dvt = resize<2>(bs);
            Mag<3> q3;
            Mag<2> r2;
            divmod<2, 2>(dvs, dvt, q3, r2);
            Mag<2> r2d;
            r2d.w[0] = r2.w[0] << 1;
            r2d.w[1] = (r2.w[1] << 1) | (r2.w[0] >> 31);
            rnd = cmp<2>(r2d, dvt) >= 0 ? 1u : 0u;
            Q = resize<W + 1>(resize<2>(q3));
        }
    }
    u64 c = rnd;
#pragma unroll
    for (int i = 0; i < W + 1; i++) {
        u64 s = (u64)Q.w[i] + c;
        Q.w[i] = (u32)s;
        c = s >> 32;
    }
    Dec<W> r;
    if (azero) {
        r.m = mzero<W>();
        r.neg = 0u;
    } else if (bzero) {
        r.m = a.m;
        r.neg = a.neg ^ b.neg;
    } else {
        r.m = resize<W>(Q);
        r.neg = a.neg ^ b.neg;
    }
    return r;
}

// Decode one compact row of B bytes (B at run time) into W limbs.  The
// caller guarantees the magnitude fits W limbs; when B > 4 W only the low
// 4 W bytes and the sign byte are read.
template <int W> UPD_DEV Dec<W> load_compact_rt(const u8* __restrict__ col, usize row, u32 B, usize colbytes)
{
    const usize off = row * (usize)B;
    const u32 nb = B < 4u * W ? B : 4u * W;
    const usize d0 = off >> 2;
    const u32 sh = (u32)(off & 3u);
    const u32 nd = (sh + nb + 3u) >> 2;   // dwords covering the magnitude bytes
    u32 dw[W + 2];
    const bool fast = (d0 + nd) * 4u <= colbytes;
    const u32* p = reinterpret_cast<const u32*>(col) + d0;
#pragma unroll
    for (int i = 0; i < W + 1; i++) {
        u32 x = 0;
        if ((u32)i < nd) {
            if (fast) {
                x = __builtin_nontemporal_load(p + i);
            } else {
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    const usize o = (d0 + (usize)i) * 4u + (usize)b;
                    x |= (o < colbytes ? (u32)col[o] : 0u) << (8 * b);
                }
            }
        }
        dw[i] = x;
    }
    dw[W + 1] = 0u;
    Dec<W> r;
#pragma unroll
    for (int i = 0; i < W; i++) {
        u32 x = __builtin_amdgcn_alignbyte(dw[i + 1], dw[i], sh);
        // keep bytes below nb only
        const u32 lo = 4u * (u32)i;
        if (lo >= nb) x = 0u;
        else if (lo + 4u > nb) x &= (1u << (8u * (nb - lo))) - 1u;
        r.m.w[i] = x;
    }
    // sign: bit 7 of byte B - 1
    if (B <= 4u * W) {
        const u32 sb = 8u * (B - 1u);
#pragma unroll
        for (int i = 0; i < W; i++) {
            if ((u32)i == (sb >> 5)) {
                r.neg = (r.m.w[i] >> (sb & 31u) >> 7) & 1u;
                r.m.w[i] &= ~(0x80u << (sb & 31u));
            }
        }
    } else {
        r.neg = (u32)col[off + B - 1u] >> 7;
    }
    return r;
}

// the compact round trip of a SUM input (see `stored` in the generated
// kernels): keep 8 B - 1 magnitude bits, OR the bit at 8 B - 1 into the sign
template <int W> UPD_DEV Dec<W> stored_rt(const Dec<W>& v, u32 B)
{
    Dec<W> r = v;
    const u32 top = 8u * B - 1u;
#pragma unroll
    for (int i = 0; i < W; i++) {
        const u32 lo = 32u * (u32)i;
        if (lo > top) {
            r.m.w[i] = 0u;
        } else if (lo + 32u > top) {
            const u32 bit = top - lo;
            r.neg |= (v.m.w[i] >> bit) & 1u;
            r.m.w[i] &= bit == 0u ? 0u : (0xffffffffu >> (32u - bit));
        }
    }
    return r;
}

// Add of equally scaled operands with the reference's sign rules
// (decimal.cu:1284-1323): same sign -> magnitudes add and the sign is kept
// (so -0 + -0 = -0); different signs -> |larger| - |smaller| with the sign
// of the larger, and +0 on a tie.
template <int WO, int WA, int WB> UPD_DEV Dec<WO> dadd(const Dec<WA>& a, const Dec<WB>& b)
{
    const Mag<WO> x = resize<WO>(a.m);
    const Mag<WO> y = resize<WO>(b.m);
    const bool same = a.neg == b.neg;
    const u32 m = same ? 0u : 0xffffffffu;
    Mag<WO> t;
    u64 c = same ? 0u : 1u;
#pragma unroll
    for (int i = 0; i < WO; i++) {
        u64 s = (u64)x.w[i] + (y.w[i] ^ m) + c;
        t.w[i] = (u32)s;
        c = s >> 32;
    }
    // different signs and no carry out: x < y, t = x - y + 2^(32 WO) -> negate
    const bool flip = !same && c == 0;
    const u32 fm = flip ? 0xffffffffu : 0u;
    Dec<WO> r;
    u64 cc = flip ? 1u : 0u;
    u32 any = 0;
#pragma unroll
    for (int i = 0; i < WO; i++) {
        u64 s = (u64)(t.w[i] ^ fm) + cc;
        r.m.w[i] = (u32)s;
        cc = s >> 32;
        any |= r.m.w[i];
    }
    r.neg = same ? a.neg : (flip ? b.neg : (any ? a.neg : 0u));
    return r;
}

// a - b == a + (-b) under these rules (decimal.cu:1575-1613)
template <int WO, int WA, int WB> UPD_DEV Dec<WO> dsub(const Dec<WA>& a, const Dec<WB>& b)
{
    return dadd<WO, WA, WB>(a, dneg<WB>(b));
}

// MulDecimalOnCpu (decimal.cu:1325-1351): sign xor (so -0 is possible), magnitude product
template <int WO, int WA, int WB> UPD_DEV Dec<WO> dmul(const Dec<WA>& a, const Dec<WB>& b)
{
    Dec<WO> r;
    r.m = mul_lo<WO, WA, WB>(a.m, b.m);
    r.neg = a.neg ^ b.neg;
    return r;
}

// DivOrMod, isMod = false (decimal.cu:2201-2333): q = round-half-up of
// |a| * 10^S / |b| with S = scale(b) + 4; result scale = scale(a) + 4 and
// sign = xor; 0 / x = +0; x / 0 keeps |a| (the reference flags
// ERR_DIV_BY_ZERO in a field the compact layout does not store).
// FAITH selects the reference's path quirks (see od_div in the oracle):
//   (i)  bitlen(a) + est(10^S) - 1 <= 64 and bitlen(b) <= 64: uint64 numerator
//        and doubled remainder, both wrapping mod 2^64;
//   (ii) |b| < 2^32: the round-up test is (uint32)(2R) >= |b|;
//   (iii) otherwise (DivInt's bisection, decimal.cu:2125-2191): its lower
//        bound is 2^max(0, bitlen(X) - bitlen(b) - 1), so when X is 1 to 64
//        bits shorter than b the search starts and ends at 1 and the
//        negative residual blocks the round-up: the result is 1.  A gap
//        above 64 bits stores outside the reference's qmax (no defined
//        result): the exact rule is kept.
template <int WO, int WA, int WB, int S, int FAITH>
UPD_DEV Dec<WO> ddiv(const Dec<WA>& a, const Dec<WB>& b, u32* err = nullptr)
{
    constexpr int WX = WA + P10<S>::W;
    const bool azero = is_zero(a.m);
    const bool bzero = is_zero(b.m);
    Mag<WB> bs = b.m;
    if (bzero) bs.w[0] = 1u;
    const Mag<WX> X = mul_p10<WX, WA, S>(a.m);
    Mag<WX + 1> Q;
    Mag<WB> R;
    divmod<WX, WB>(X, bs, Q, R);
    Mag<WB> T;
    sub_to<WB>(T, bs, R);
    u32 rnd = cmp<WB>(R, T) >= 0 ? 1u : 0u;
    if constexpr (FAITH) {
        const u32 bla = bitlen(a.m), blb = bitlen(bs);
        const bool path_i = (bla + (u32)est_shift_bits(S) - 1u <= 64u) && blb <= 64u;
        bool top0 = true;
#pragma unroll
        for (int i = 1; i < WB; i++) top0 = top0 && bs.w[i] == 0u;
        const bool path_ii = !path_i && top0;
        const u32 blx = bitlen(X);
        const bool path_iii_one = !path_i && !path_ii && blx < blb && blb - blx <= 64u;
        if (path_ii) rnd = ((u32)(R.w[0] * 2u) >= bs.w[0]) ? 1u : 0u;
        if (path_iii_one) {
            Q = mzero<WX + 1>();
            Q.w[0] = 1u;
            rnd = 0u;
        }
        if (path_i) {
            // uint64 emulation of DivByInt64 (decimal.cu:2431-2446)
            Mag<2> a2 = resize<2>(a.m);
            Mag<2> dvs = mul_p10<2, 2, S>(a2);
            Mag<2> dvt = resize<2>(bs);
            Mag<3> q3;
            Mag<2> r2;
            divmod<2, 2>(dvs, dvt, q3, r2);
            Mag<2> r2d;
            r2d.w[0] = r2.w[0] << 1;
            r2d.w[1] = (r2.w[1] << 1) | (r2.w[0] >> 31);
            rnd = cmp<2>(r2d, dvt) >= 0 ? 1u : 0u;
            Q = resize<WX + 1>(resize<2>(q3));
        }
    }
    // Q += rnd
    u64 c = rnd;
#pragma unroll
    for (int i = 0; i < WX + 1; i++) {
        u64 s = (u64)Q.w[i] + c;
        Q.w[i] = (u32)s;
        c = s >> 32;
    }
    Dec<WO> r;
    if (azero) {
        r.m = mzero<WO>();
        r.neg = 0u;
    } else if (bzero) {
        r.m = resize<WO>(a.m);
        r.neg = a.neg ^ b.neg;
        if (err) *err = 2u;
    } else {
        r.m = resize<WO>(Q);
        r.neg = a.neg ^ b.neg;
    }
    return r;
}

// ModCalc (decimal.cu:2621-2716), the % of two decimals: both operands are
// taken as integers (their scales are dropped), the result has scale 0 and
// sign a.sign xor b.sign, and 0 % x = +0.  By bit lengths la = bitlen(|a|),
// lb = bitlen(|b|):
//   la <= 64 and lb <= 64 (ModByInt64), or |b| < 2^32 (ModByInt): |a| mod |b|;
//   otherwise DivInt's bisection (decimal.cu:2125-2191), whose quotient
//   search starts at [2^(la-lb-1), 2^(la-lb+1)]:
//     la > lb: |a| mod |b|;
//     lb - 64 <= la <= lb: the search starts at quotient 1 and stops there
//       (for la <= lb - 2 its upper bound is a negative shift, 0 on PTX, or
//       a store into qmax's own header bytes), and the remainder |a| - |b|
//       is kept without its sign: ||a| - |b|| (FAITH);
//     la < lb - 64: the reference stores outside qmax (no defined result);
//       |a| mod |b| = |a| here.
// x % 0: the reference scans below the divisor's limbs (no defined result);
// |a| is kept, as x / 0 keeps it.  Without FAITH: |a| mod |b| throughout.
template <int WO, int WA, int WB, int FAITH> UPD_DEV Dec<WO> dmod(const Dec<WA>& a, const Dec<WB>& b)
{
    Dec<WO> r;
    if (is_zero(a.m)) {
        r.m = mzero<WO>();
        r.neg = 0u;
        return r;
    }
    r.neg = a.neg ^ b.neg;
    if (is_zero(b.m)) {
        r.m = resize<WO>(a.m);
        return r;
    }
    Mag<WA + 1> Q;
    Mag<WB> R;
    divmod<WA, WB>(a.m, b.m, Q, R);
    r.m = resize<WO>(R);
    if constexpr (FAITH) {
        const u32 la = bitlen(a.m), lb = bitlen(b.m);
        bool small_b = true;
#pragma unroll
        for (int i = 1; i < WB; i++) small_b = small_b && b.m.w[i] == 0u;
        const bool bisect = !(la <= 64u && lb <= 64u) && !small_b;
        if (bisect && la <= lb && la + 64u >= lb) {
            constexpr int WM = WA > WB ? WA : WB;
            const Mag<WM> x = resize<WM>(a.m), y = resize<WM>(b.m);
            Mag<WM> d;
            if (cmp<WM>(x, y) >= 0) sub_to<WM>(d, x, y);
            else sub_to<WM>(d, y, x);
            r.m = resize<WO>(d);
        }
    }
    return r;
}

// ---------------------------------------------------------------- compact codec

// Decode one compact DECIMAL row (decimal.cu:168-176): B little-endian bytes
// at row*B, sign in bit 7 of the last byte.  Loads are the aligned dwords
// covering the row; the last rows of a column, whose covering dwords would
// run past the column, take a byte path.
// Decode the B-byte compact value at byte offset `off` of a 4-byte aligned
// buffer of `colbytes` bytes (load_compact below: off = row * B).
template <int B> UPD_DEV Dec<(B + 3) / 4> load_compact_off(const u8* __restrict__ col, usize off, usize colbytes)
{
    constexpr int ND = (B + 6) / 4;
    constexpr int NW = (B + 3) / 4;
    const usize d0 = off >> 2;
    const u32 sh = (u32)(off & 3u);
    u32 dw[ND + 1];
    if ((d0 + ND) * 4u <= colbytes) {
        const u32* p = reinterpret_cast<const u32*>(col) + d0;
        // the last dword is read only when the row reaches into it: a row
        // that ends on a dword boundary would otherwise touch the next
        // row's dword (and, at the end of a block's range, a cache line
        // another block on another XCD fetches again)
        const u32 nd = (sh + (u32)B + 3u) >> 2;
#pragma unroll
        for (int i = 0; i < ND; i++) dw[i] = (i < NW || (u32)i < nd) ? __builtin_nontemporal_load(p + i) : 0u;
    } else {
#pragma unroll
        for (int i = 0; i < ND; i++) {
            u32 x = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                usize o = (d0 + i) * 4u + b;
                u32 byte = (o < colbytes) ? (u32)col[o] : 0u;
                x |= byte << (8 * b);
            }
            dw[i] = x;
        }
    }
    dw[ND] = 0u;
    Dec<NW> r;
#pragma unroll
    for (int i = 0; i < NW; i++) r.m.w[i] = __builtin_amdgcn_alignbyte(dw[i + 1], dw[i], sh);
    if constexpr ((B % 4) != 0) r.m.w[NW - 1] &= (1u << (8 * (B % 4))) - 1u;
    constexpr int SB = 8 * ((B - 1) % 4) + 7;
    r.neg = (r.m.w[NW - 1] >> SB) & 1u;
    r.m.w[NW - 1] &= ~(1u << SB);
    return r;
}

template <int B> UPD_DEV Dec<(B + 3) / 4> load_compact(const u8* __restrict__ col, usize row, usize colbytes)
{
    return load_compact_off<B>(col, row * (usize)B, colbytes);
}

// ---- wave-staged loads
//
// A wave's 64 consecutive rows of a B-byte compact column are 64 B
// contiguous bytes (4 B 16-byte vectors).  Loading them with 16-byte vector
// loads into LDS and extracting each lane's row there reads the column at
// the streaming rate; the per-lane covering-dword loads above spread one
// instruction over 64 rows B bytes apart and reach about 75% of it.

// words of a wave's staging area for one column (2 words of slack for the
// covering reads of the last lane)
template <int B> constexpr int stage_words() { return 16 * B + 2; }

typedef unsigned int u32x4 __attribute__((vector_size(16)));
typedef unsigned int u32x2 __attribute__((vector_size(8)));

// stage rows [base, base + 64) of the column; base is a multiple of 64 and
// the column start is 16-byte aligned (checked by the caller), so the span
// starts on a 16-byte boundary
template <int B> UPD_DEV void stage_wave(const u8* __restrict__ col, usize base, u32* lds, u32 lane)
{
    constexpr int NV = 4 * B;   // 64 B bytes / 16
    const u32x4* src = reinterpret_cast<const u32x4*>(col + base * (usize)B);
#pragma unroll
    for (int k = 0; k < (NV + 63) / 64; k++) {
        const u32 i = lane + 64u * (u32)k;
        if (i < (u32)NV) {
            const u32x4 v = __builtin_nontemporal_load(src + i);
            lds[4 * i] = v[0];
            lds[4 * i + 1] = v[1];
            lds[4 * i + 2] = v[2];
            lds[4 * i + 3] = v[3];
        }
    }
}

// this lane's row from the staged bytes (same decoding as load_compact)
template <int B> UPD_DEV Dec<(B + 3) / 4> load_staged(const u32* lds, u32 lane)
{
    constexpr int ND = (B + 6) / 4;
    constexpr int NW = (B + 3) / 4;
    const u32 off = lane * (u32)B;
    const u32 d0 = off >> 2;
    const u32 sh = off & 3u;
    u32 dw[ND + 1];
#pragma unroll
    for (int i = 0; i < ND; i++) dw[i] = lds[d0 + i];
    dw[ND] = 0u;
    Dec<NW> r;
#pragma unroll
    for (int i = 0; i < NW; i++) r.m.w[i] = __builtin_amdgcn_alignbyte(dw[i + 1], dw[i], sh);
    if constexpr ((B % 4) != 0) r.m.w[NW - 1] &= (1u << (8 * (B % 4))) - 1u;
    constexpr int SB = 8 * ((B - 1) % 4) + 7;
    r.neg = (r.m.w[NW - 1] >> SB) & 1u;
    r.m.w[NW - 1] &= ~(1u << SB);
    return r;
}

// ---- row quads
//
// Four consecutive rows of a B-byte compact column are 4 B bytes: B whole
// dwords, dword aligned when the column is.  A lane that owns rows
// [4q, 4q + 4) loads those B dwords with 16-byte vector loads (consecutive
// lanes 4 B bytes apart, every byte fetched belongs to the lane) and
// decodes its four rows from registers at byte offsets fixed at compile
// time once the caller's loop over the four rows is unrolled.

typedef unsigned int u32x4u __attribute__((vector_size(16), aligned(4)));
typedef unsigned int u32x2u __attribute__((vector_size(8), aligned(4)));

template <int B> struct Quad {
    u32 w[B];
};

// dwords [q B, q B + B) of a column of colbytes bytes; in the column's last,
// partial quad the dwords past its end read as zero and the covering dword
// of its last bytes is assembled byte by byte
template <int B> UPD_DEV Quad<B> load_quad(const u8* __restrict__ col, usize q, usize colbytes)
{
    Quad<B> r;
    const usize d0 = q * (usize)B;
    const u32* p = reinterpret_cast<const u32*>(col) + d0;
    if ((d0 + (usize)B) * 4u <= colbytes) {
#pragma unroll
        for (int i = 0; i + 4 <= B; i += 4) {
            const u32x4u v = __builtin_nontemporal_load(reinterpret_cast<const u32x4u*>(p + i));
            r.w[i] = v[0];
            r.w[i + 1] = v[1];
            r.w[i + 2] = v[2];
            r.w[i + 3] = v[3];
        }
        if constexpr ((B & 3) >= 2) {
            const u32x2u v = __builtin_nontemporal_load(reinterpret_cast<const u32x2u*>(p + (B & ~3)));
            r.w[B & ~3] = v[0];
            r.w[(B & ~3) + 1] = v[1];
        }
        if constexpr ((B & 1) != 0) r.w[B - 1] = __builtin_nontemporal_load(p + (B - 1));
    } else {
#pragma unroll
        for (int i = 0; i < B; i++) {
            u32 x = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const usize o = (d0 + (usize)i) * 4u + (usize)b;
                x |= (o < colbytes ? (u32)col[o] : 0u) << (8 * b);
            }
            r.w[i] = x;
        }
    }
    return r;
}

// row k (0..3) of a quad of a B-byte column (decoding as load_compact)
template <int B> UPD_DEV Dec<(B + 3) / 4> quad_row(const Quad<B>& q, u32 k)
{
    constexpr int ND = (B + 6) / 4;
    constexpr int NW = (B + 3) / 4;
    const u32 off = k * (u32)B;
    const u32 d0 = off >> 2;
    const u32 sh = off & 3u;
    u32 dw[ND + 1];
#pragma unroll
    for (int i = 0; i < ND; i++) dw[i] = (d0 + (u32)i < (u32)B) ? q.w[d0 + i] : 0u;
    dw[ND] = 0u;
    Dec<NW> r;
#pragma unroll
    for (int i = 0; i < NW; i++) r.m.w[i] = __builtin_amdgcn_alignbyte(dw[i + 1], dw[i], sh);
    if constexpr ((B % 4) != 0) r.m.w[NW - 1] &= (1u << (8 * (B % 4))) - 1u;
    constexpr int SB = 8 * ((B - 1) % 4) + 7;
    r.neg = (r.m.w[NW - 1] >> SB) & 1u;
    r.m.w[NW - 1] &= ~(1u << SB);
    return r;
}

// byte k of the quad of a byte column
UPD_DEV u32 quad_byte(const Quad<1>& q, u32 k) { return (q.w[0] >> (8u * k)) & 255u; }

// Encode (decimal.cu:337-341): the low B bytes of the magnitude with the sign
// OR-ed into bit 7 of the last byte.  Returns the bytes as words.
template <int B, int W> UPD_DEV Mag<(B + 3) / 4> encode_compact(const Dec<W>& a)
{
    constexpr int NW = (B + 3) / 4;
    Mag<NW> r;
#pragma unroll
    for (int i = 0; i < NW; i++) r.w[i] = (i < W) ? a.m.w[i] : 0u;
    if constexpr ((B % 4) != 0) r.w[NW - 1] &= (1u << (8 * (B % 4))) - 1u;
    constexpr int SW = (B - 1) / 4, SB = 8 * ((B - 1) % 4) + 7;
    r.w[SW] |= a.neg << SB;
    return r;
}

// ---------------------------------------------------------------- two's complement accumulators

// acc += (-1)^neg * m, sign-extended to AW limbs
template <int AW, int W> UPD_DEV void acc_add(Mag<AW>& acc, const Dec<W>& v)
{
    const u32 m = v.neg ? 0xffffffffu : 0u;
    u64 c = v.neg ? 1u : 0u;
#pragma unroll
    for (int i = 0; i < AW; i++) {
        u32 x = (i < W) ? (v.m.w[i] ^ m) : m;
        u64 s = (u64)acc.w[i] + x + c;
        acc.w[i] = (u32)s;
        c = s >> 32;
    }
}

template <int AW> UPD_DEV void acc_merge(Mag<AW>& acc, const Mag<AW>& o)
{
    add_to<AW>(acc, acc, o);
}

// wave64 butterfly reduction of a two's complement multi-limb value
template <int AW> UPD_DEV void wave_reduce(Mag<AW>& acc)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        Mag<AW> o;
#pragma unroll
        for (int i = 0; i < AW; i++) o.w[i] = (u32)__shfl_xor((int)acc.w[i], off, 64);
        add_to<AW>(acc, acc, o);
    }
}

// Convert a two's complement sum into the 133-byte nullable_type<Decimal>
// image {flag, sign, prec, frac, error, v[32]} (aries_types.hxx:298-310,
// decimal.hxx:130-141).  Zero is +0 (a +0 seed never yields -0).
// sign-magnitude value of a two's complement accumulator
template <int AW> UPD_DEV Dec<AW> sum_value(const Mag<AW>& acc)
{
    Dec<AW> d;
    const bool neg = (acc.w[AW - 1] >> 31) != 0;
    const u32 fm = neg ? 0xffffffffu : 0u;
    u64 c = neg ? 1u : 0u;
#pragma unroll
    for (int i = 0; i < AW; i++) {
        u64 s = (u64)(acc.w[i] ^ fm) + c;
        d.m.w[i] = (u32)s;
        c = s >> 32;
    }
    d.neg = neg ? 1u : 0u;
    return d;
}

// 133-byte nullable_type<Decimal> image of a sign-magnitude value of AW
// limbs (limbs past AW are zero)
template <int AW> UPD_DEV void write_dec_image(u8* out, const Dec<AW>& d, u32 prec, u32 frac, u32 error)
{
    out[0] = 1;
    out[1] = (u8)d.neg;
    out[2] = (u8)prec;
    out[3] = (u8)frac;
    out[4] = (u8)error;
#pragma unroll
    for (int i = 0; i < 32; i++) {
        u32 x = (i < AW) ? d.m.w[i] : 0u;
        out[5 + 4 * i + 0] = (u8)x;
        out[5 + 4 * i + 1] = (u8)(x >> 8);
        out[5 + 4 * i + 2] = (u8)(x >> 16);
        out[5 + 4 * i + 3] = (u8)(x >> 24);
    }
}

template <int AW> UPD_DEV void write_sum_image(u8* out, const Mag<AW>& acc, u32 prec, u32 frac, u32 error)
{
    write_dec_image<AW>(out, sum_value<AW>(acc), prec, frac, error);
}

// AVG = SUM / COUNT of one group (kernel_util.hxx:427-438, do_div):
// Decimal(sum) /= Decimal((int64)count) by the division rule, scale S + 4,
// written as a 133-byte nullable_type<Decimal> image with the given flag
// byte.  a is the sum's sign-magnitude value (W limbs of the 32).  Sums that
// fit 8 limbs divide at that width (the quotient does not depend on the
// operand widths).
template <int FAITH, int W>
UPD_DEV void div_dec_int64(const Dec<W>& a, u32 prec, u32 frac, u32 error, u8 flag, i64 cnt, u8* o)
{
    Dec<2> b;
    const u64 uc = cnt < 0 ? (u64)(-(cnt + 1)) + 1u : (u64)cnt;
    b.neg = cnt < 0 ? 1u : 0u;
    b.m.w[0] = (u32)uc;
    b.m.w[1] = (u32)(uc >> 32);
    u32 err = error;
    Dec<32> q;
    if constexpr (W <= 8) {
        q = dresize<32>(ddiv<10, 8, 2, 4, FAITH>(dresize<8>(a), b, &err));
    } else {
        bool narrow = true;
#pragma unroll
        for (int w = 8; w < W; w++) narrow = narrow && a.m.w[w] == 0u;
        if (narrow) q = dresize<32>(ddiv<10, 8, 2, 4, FAITH>(dresize<8>(a), b, &err));
        else q = ddiv<32, 32, 2, 4, FAITH>(dresize<32>(a), b, &err);
    }
    o[0] = flag;
    o[1] = (u8)q.neg;
    o[2] = is_zero(a.m) ? (u8)(frac + 4u + 1u) : (u8)prec;
    o[3] = (u8)(frac + 4u);
    o[4] = (u8)err;
#pragma unroll
    for (int w = 0; w < 32; w++) {
        o[5 + 4 * w] = (u8)q.m.w[w];
        o[6 + 4 * w] = (u8)(q.m.w[w] >> 8);
        o[7 + 4 * w] = (u8)(q.m.w[w] >> 16);
        o[8 + 4 * w] = (u8)(q.m.w[w] >> 24);
    }
}

// the same from a 133-byte SUM image s
template <int FAITH> UPD_DEV void div_image_int64(const u8* s, i64 cnt, u8* o)
{
    Dec<32> a;
    a.neg = s[1];
#pragma unroll
    for (int w = 0; w < 32; w++)
        a.m.w[w] = (u32)s[5 + 4 * w] | ((u32)s[6 + 4 * w] << 8) | ((u32)s[7 + 4 * w] << 16) | ((u32)s[8 + 4 * w] << 24);
    div_dec_int64<FAITH, 32>(a, s[2], s[3], s[4], s[0], cnt, o);
}

} // namespace updec

#endif
