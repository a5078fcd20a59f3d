// dec_mlane.h -- decimal limb arithmetic with one value spread over a
// cooperative group of lanes (the reference's multi-lane XMP form,
// src/AriesEngine/AriesExprCalcNode.cpp:1320-1400 with its CGBN core under
// src/datatypes/xmp_utils, re-designed for wave64).
//
// A value of W = L * TPI 32-bit limbs is held by TPI consecutive lanes of a
// wave (a group, TPI a power of two up to 64): the lane of rank r = lane mod
// TPI holds limbs [r L, r L + L), every lane of the group holds the sign.
// Carries and borrows between the lanes of a group are resolved with one
// wave ballot (carry lookahead on the group's bits of the ballot mask),
// multi-bit carries move one lane up with a DPP wave shift, and a limb is
// broadcast to the group with a shuffle (a scalar lane read when the group
// is the whole wave).  Every group of a wave runs the same instruction
// stream on its own value; all group-wide decisions are uniform within a
// group, so groups of one wave may diverge without splitting a group.
//
// Used by the interpreter's wide register classes (vm.hip): W = 64, 128, 256
// and 512 limbs, where one value per lane no longer fits the register file;
// and by the generated multi-lane calc kernels (planner.cpp, gen_calc_ml).
#ifndef UPDEC_DEC_MLANE_H
#define UPDEC_DEC_MLANE_H
#include "dec_device.h"

namespace updec {
namespace ml {

template <int L> struct MMag { u32 w[L]; };
template <int L> struct MDec { MMag<L> m; u32 neg; };

// lane position in its group
template <int TPI> struct Grp {
    u32 rank, base;
    UPD_DEV Grp()
    {
        const u32 lane = (u32)__lane_id();
        rank = lane & (TPI - 1u);
        base = lane & ~(TPI - 1u);
    }
    // this group's bits of a wave ballot, rank r at bit r
    UPD_DEV u64 bits(u64 ballot) const
    {
        if constexpr (TPI == 64) return ballot;
        else return (ballot >> base) & ((1ull << TPI) - 1ull);
    }
    // v of the group's lane src_rank (any source per lane)
    UPD_DEV u32 from(u32 v, u32 src_rank) const { return (u32)__shfl((int)v, (int)(base + src_rank), 64); }
    // v of lane src_rank, src_rank the same on every lane of the group: a
    // scalar read of the lane when the group is the whole wave
    UPD_DEV u32 bcast(u32 v, u32 src_rank) const
    {
        if constexpr (TPI == 64) return (u32)__builtin_amdgcn_readlane((int)v, (int)src_rank);
        else return from(v, src_rank);
    }
    // v of the next lower lane of the wave (DPP wave_shr:1); rank 0 of a
    // group reads the group below (callers ignore it)
    UPD_DEV u32 below(u32 v) const { return (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xf, 0xf, false); }
    // v of the next higher lane of the wave (DPP wave_shl:1); the top rank of
    // a group reads the group above (callers ignore it)
    UPD_DEV u32 above(u32 v) const { return (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xf, 0xf, false); }
    UPD_DEV u64 below64(u64 v) const { return ((u64)below((u32)(v >> 32)) << 32) | below((u32)v); }
};

template <int L> UPD_DEV MMag<L> mzero()
{
    MMag<L> r;
#pragma unroll
    for (int i = 0; i < L; i++) r.w[i] = 0u;
    return r;
}

// limb t of a lane's slice with t uniform in the group but not a
// compile-time constant
template <int L> UPD_DEV u32 sel(const MMag<L>& a, u32 t)
{
    // masks rather than a select chain: the compiler turns a chain of
    // selects over the array into one load at a run-time index, which moves
    // the array to scratch
    u32 v = 0u;
#pragma unroll
    for (int i = 0; i < L; i++) v |= a.w[i] & (0u - (u32)(t == (u32)i));
    return v;
}

// carry into each lane from per-lane generate (a carry leaves the lane) and
// propagate (a carry entering the lane leaves it) flags: with G and P the
// group's masks (disjoint), the carries into the lanes are (G + (G | P)) ^ P
template <int TPI> UPD_DEV u32 carry_in(const Grp<TPI>& g, bool gen, bool prop)
{
    const u64 G = g.bits(__ballot(gen ? 1 : 0));
    const u64 P = g.bits(__ballot(prop ? 1 : 0));
    const u64 C = (G + (G | P)) ^ P;
    return (u32)((C >> g.rank) & 1ull);
}

// a += b + c0 (c0 enters at rank 0), modulo 2^(32 W); returns the carry out
// of the group (on every lane)
template <int L, int TPI> UPD_DEV u32 add_mag(MMag<L>& a, const MMag<L>& b, u32 c0, const Grp<TPI>& g)
{
    u64 c = g.rank == 0 ? c0 : 0u;
    u32 all1 = 0xffffffffu;
#pragma unroll
    for (int i = 0; i < L; i++) {
        const u64 s = (u64)a.w[i] + b.w[i] + c;
        a.w[i] = (u32)s;
        c = s >> 32;
        all1 &= a.w[i];
    }
    const bool gen = c != 0, prop = !gen && all1 == 0xffffffffu;
    const u32 cin = carry_in<TPI>(g, gen, prop);
    u32 cc = cin;
#pragma unroll
    for (int i = 0; i < L; i++) {
        const u64 s = (u64)a.w[i] + cc;
        a.w[i] = (u32)s;
        cc = (u32)(s >> 32);
    }
    const u32 out = (gen || (prop && cin)) ? 1u : 0u;
    return g.bcast(out, TPI - 1u);
}

// a -= b (modulo 2^(32 W)); returns the borrow out of the group
template <int L, int TPI> UPD_DEV u32 sub_mag(MMag<L>& a, const MMag<L>& b, const Grp<TPI>& g)
{
    MMag<L> nb;
#pragma unroll
    for (int i = 0; i < L; i++) nb.w[i] = ~b.w[i];
    return add_mag<L, TPI>(a, nb, 1u, g) ^ 1u;
}

template <int L, int TPI> UPD_DEV bool is_zero(const MMag<L>& a, const Grp<TPI>& g)
{
    u32 any = 0;
#pragma unroll
    for (int i = 0; i < L; i++) any |= a.w[i];
    return g.bits(__ballot(any != 0 ? 1 : 0)) == 0ull;
}

// compare magnitudes: -1, 0, 1 (on every lane)
template <int L, int TPI> UPD_DEV int cmp(const MMag<L>& a, const MMag<L>& b, const Grp<TPI>& g)
{
    int loc = 0;
#pragma unroll
    for (int i = L - 1; i >= 0; i--)
        if (loc == 0 && a.w[i] != b.w[i]) loc = a.w[i] > b.w[i] ? 1 : -1;
    const u64 nz = g.bits(__ballot(loc != 0 ? 1 : 0));
    if (nz == 0ull) return 0;
    const u32 top = 63u - (u32)__clzll(nz);
    return (int)g.bcast((u32)loc, top);
}

// number of limbs up to the top nonzero one (0 for zero), on every lane
template <int L, int TPI> UPD_DEV u32 limb_len(const MMag<L>& a, const Grp<TPI>& g)
{
    u32 loc = 0;
#pragma unroll
    for (int i = 0; i < L; i++)
        if (a.w[i]) loc = (u32)i + 1u;
    const u64 nz = g.bits(__ballot(loc != 0 ? 1 : 0));
    if (nz == 0ull) return 0u;
    const u32 top = 63u - (u32)__clzll(nz);
    return top * (u32)L + g.bcast(loc, top);
}

// lanes of the group up to the highest one holding a nonzero limb (0 for
// zero), on every lane
template <int L, int TPI> UPD_DEV u32 rank_len(const MMag<L>& a, const Grp<TPI>& g)
{
    u32 o = 0;
#pragma unroll
    for (int i = 0; i < L; i++) o |= a.w[i];
    const u64 nz = g.bits(__ballot(o != 0u ? 1 : 0));
    return nz == 0ull ? 0u : 64u - (u32)__clzll(nz);
}

// limb k of the group's value (k uniform in the group; 0 past the value)
template <int L, int TPI> UPD_DEV u32 limb(const MMag<L>& a, u32 k, const Grp<TPI>& g)
{
    if (k >= (u32)(L * TPI)) return 0u;
    return g.bcast(sel<L>(a, k % (u32)L), k / (u32)L);
}

// add a multi-bit carry word cw (< 2^(32 L - 1)) from the lane below into
// every lane's slice and resolve the one-bit carries that follows
template <int L, int TPI> UPD_DEV void carry_up(MMag<L>& a, u64 cw_out, const Grp<TPI>& g)
{
    const u64 below = g.below64(cw_out);
    u64 c = g.rank == 0 ? 0ull : below;
    u32 all1 = 0xffffffffu;
#pragma unroll
    for (int i = 0; i < L; i++) {
        const u64 s = (u64)a.w[i] + (c & 0xffffffffull);
        a.w[i] = (u32)s;
        c = (c >> 32) + (s >> 32);
        all1 &= a.w[i];
    }
    const bool gen = c != 0, prop = !gen && all1 == 0xffffffffu;
    u32 cc = carry_in<TPI>(g, gen, prop);
#pragma unroll
    for (int i = 0; i < L; i++) {
        const u64 s = (u64)a.w[i] + cc;
        a.w[i] = (u32)s;
        cc = (u32)(s >> 32);
    }
}

// a *= m (one word), modulo 2^(32 W); returns the word carried out of the
// group (on every lane)
template <int L, int TPI> UPD_DEV u32 mul_word(MMag<L>& a, u32 m, const Grp<TPI>& g)
{
    u64 c = 0;
#pragma unroll
    for (int i = 0; i < L; i++) {
        const u64 t = (u64)a.w[i] * m + c;
        a.w[i] = (u32)t;
        c = t >> 32;
    }
    // the top lane's word leaves the group, plus a one-bit carry the
    // resolution below may push out of it
    const u32 top_c = g.bcast((u32)c, TPI - 1u);
    const u64 below = g.below64(c);
    u64 cw = g.rank == 0 ? 0ull : below;
    u32 all1 = 0xffffffffu;
#pragma unroll
    for (int i = 0; i < L; i++) {
        const u64 s = (u64)a.w[i] + (cw & 0xffffffffull);
        a.w[i] = (u32)s;
        cw = (cw >> 32) + (s >> 32);
        all1 &= a.w[i];
    }
    const bool gen = cw != 0, prop = !gen && all1 == 0xffffffffu;
    const u32 cin = carry_in<TPI>(g, gen, prop);
    u32 cc = cin;
#pragma unroll
    for (int i = 0; i < L; i++) {
        const u64 s = (u64)a.w[i] + cc;
        a.w[i] = (u32)s;
        cc = (u32)(s >> 32);
    }
    const u32 top_out = (gen || (prop && cin)) ? 1u : 0u;
    return top_c + g.bcast(top_out, TPI - 1u);
}

// a * 10^n modulo 2^(32 W), n at run time (uniform in the group)
template <int L, int TPI> UPD_DEV void mul_p10(MMag<L>& a, u32 n, const Grp<TPI>& g)
{
    for (; n >= 9u; n -= 9u) (void)mul_word<L, TPI>(a, 1000000000u, g);
    if (n) (void)mul_word<L, TPI>(a, pow10_word(n), g);
}

// a = floor(a / d) for one word d < 2^30: the remainder entering a lane
// from the lane above is the composition of the affine maps
// x -> (x * 2^(32 L) + slice) mod d of the lanes above it, found with a
// suffix scan over the group; each lane then divides its slice.
template <int L, int TPI> UPD_DEV void div_word(MMag<L>& a, u32 d, const Grp<TPI>& g)
{
    const u64 b32 = (1ull << 32) % d;
    u64 M = 1;
#pragma unroll
    for (int i = 0; i < L; i++) M = (M * b32) % d;
    u64 c = 0;
#pragma unroll
    for (int i = L - 1; i >= 0; i--) c = ((c << 32) | a.w[i]) % d;
    // H_r = f_r o f_{r+1} o ... o f_{TPI-1}: (mul, add)
    u64 hm = M, ha = c;
#pragma unroll
    for (u32 o = 1; o < (u32)TPI; o <<= 1) {
        const u32 src = (g.rank + o) & (TPI - 1u);
        const u64 om = g.from((u32)hm, src), oa = g.from((u32)ha, src);
        if (g.rank + o < (u32)TPI) {
            ha = (hm * oa + ha) % d;
            hm = (hm * om) % d;
        }
    }
    const u32 above = g.above((u32)ha);
    u64 rem = g.rank + 1u < (u32)TPI ? above : 0u;
#pragma unroll
    for (int i = L - 1; i >= 0; i--) {
        const u64 x = (rem << 32) | a.w[i];
        a.w[i] = (u32)(x / d);
        rem = x % d;
    }
}

// floor(a / 10^n), n at run time
template <int L, int TPI> UPD_DEV void div_p10(MMag<L>& a, u32 n, const Grp<TPI>& g)
{
    for (; n >= 9u; n -= 9u) div_word<L, TPI>(a, 1000000000u, g);
    if (n) div_word<L, TPI>(a, pow10_word(n), g);
}

// product modulo 2^(32 W) by column sums: the lane of rank r forms output
// limbs k = r L + t from every a_i b_(k-i), a's limbs broadcast rank by rank
// and b's two source lanes shuffled in, then normalizes its columns and
// passes the multi-bit carry one lane up
template <int L, int TPI> UPD_DEV MMag<L> mul(const MMag<L>& a, const MMag<L>& b, const Grp<TPI>& g)
{
    u64 lo[L];
    u32 hi[L];
#pragma unroll
    for (int t = 0; t < L; t++) {
        lo[t] = 0ull;
        hi[t] = 0u;
    }
    // the operand with fewer nonzero lanes is broadcast, and only its
    // nonzero lanes: the ranks above them add nothing to any column
    const u32 ra = rank_len<L, TPI>(a, g), rb = rank_len<L, TPI>(b, g);
    const bool swap = rb < ra;
    const u32 n = swap ? rb : ra;
    MMag<L> x, y;
#pragma unroll
    for (int i = 0; i < L; i++) {
        x.w[i] = swap ? b.w[i] : a.w[i];
        y.w[i] = swap ? a.w[i] : b.w[i];
    }
    for (u32 ir = 0; ir < n; ir++) {
        u32 av[L], b0[L], b1[L];
        const u32 s0 = (g.rank - ir) & (TPI - 1u), s1 = (g.rank - ir - 1u) & (TPI - 1u);
#pragma unroll
        for (int i = 0; i < L; i++) {
            av[i] = g.bcast(x.w[i], ir);
            b0[i] = g.from(y.w[i], s0);
            b1[i] = g.from(y.w[i], s1);
        }
        const bool v0 = g.rank >= ir, v1 = g.rank >= ir + 1u;
#pragma unroll
        for (int il = 0; il < L; il++) {
#pragma unroll
            for (int t = 0; t < L; t++) {
                const u32 bj = t >= il ? (v0 ? b0[t - il] : 0u) : (v1 ? b1[t - il + L] : 0u);
                const u64 p = (u64)av[il] * bj;
                const u64 s = lo[t] + p;
                hi[t] += s < p ? 1u : 0u;
                lo[t] = s;
            }
        }
    }
    MMag<L> r;
    u64 c = 0;   // < 2^42 (column sums of at most 256 products)
#pragma unroll
    for (int t = 0; t < L; t++) {
        const u64 s = lo[t] + c;
        const u32 h = hi[t] + (s < c ? 1u : 0u);
        r.w[t] = (u32)s;
        c = ((u64)h << 32) | (s >> 32);
    }
    carry_up<L, TPI>(r, c, g);
    return r;
}

// shift left by s bits (0 <= s < 32, uniform); returns the bits shifted out
// of the group's top limb (on every lane)
template <int L, int TPI> UPD_DEV u32 shl_bits(MMag<L>& a, u32 s, const Grp<TPI>& g)
{
    const u32 topw = g.bcast(a.w[L - 1], TPI - 1u);
    const u32 out = s ? topw >> (32u - s) : 0u;
    const u32 below = g.below(a.w[L - 1]);
    u32 prev = g.rank == 0 ? 0u : below;
    if (s) {
#pragma unroll
        for (int i = 0; i < L; i++) {
            const u32 w = a.w[i];
            a.w[i] = (w << s) | (prev >> (32u - s));
            prev = w;
        }
    }
    return out;
}

// shift right by s bits (0 <= s < 32, uniform); `in` enters the top limb
template <int L, int TPI> UPD_DEV void shr_bits(MMag<L>& a, u32 s, u32 in, const Grp<TPI>& g)
{
    const u32 above = g.above(a.w[0]);
    u32 next = g.rank + 1u < (u32)TPI ? above : in;
    if (s) {
#pragma unroll
        for (int i = L - 1; i >= 0; i--) {
            const u32 w = a.w[i];
            a.w[i] = (w >> s) | (next << (32u - s));
            next = w;
        }
    }
}

// a << (32 n) (n limbs, uniform), limbs past the top dropped
template <int L, int TPI> UPD_DEV MMag<L> shl_limbs(const MMag<L>& a, u32 n, const Grp<TPI>& g)
{
    MMag<L> r;
#pragma unroll
    for (int t = 0; t < L; t++) {
        // source limb i = r L + t - n; its local index is the same in every lane
        const u32 loc = (u32)(t + L - (int)(n % L)) % (u32)L;
        const u32 v = sel<L>(a, loc);
        const int i = (int)(g.rank * L) + t - (int)n;
        const u32 src = i >= 0 ? (u32)i / (u32)L : 0u;
        const u32 got = g.from(v, src);
        r.w[t] = i >= 0 ? got : 0u;
    }
    return r;
}

// a >> 32 (one limb right)
template <int L, int TPI> UPD_DEV void shr_limb1(MMag<L>& a, const Grp<TPI>& g)
{
    const u32 above = g.above(a.w[0]);
#pragma unroll
    for (int i = 0; i < L - 1; i++) a.w[i] = a.w[i + 1];
    a.w[L - 1] = g.rank + 1u < (u32)TPI ? above : 0u;
}

// limb k of a set to v (k uniform): only the owning lane writes
template <int L, int TPI> UPD_DEV void set_limb(MMag<L>& a, u32 k, u32 v, const Grp<TPI>& g)
{
    if (k / (u32)L != g.rank) return;
#pragma unroll
    for (int i = 0; i < L; i++)
        if ((u32)i == k % (u32)L) a.w[i] = v;
}

// Q = floor(X / V), R = X mod V for V != 0: Knuth's Algorithm D (TAOCP
// 4.3.1) with the divisor normalized, each quotient digit estimated from the
// top words of the window (broadcast to the group) and applied by one
// group multiply-subtract; the numerator carries one extra word above the
// group's W limbs.
template <int L, int TPI>
UPD_DEV void divmod(const MMag<L>& X, const MMag<L>& V, MMag<L>& Q, MMag<L>& R, const Grp<TPI>& g)
{
    Q = mzero<L>();
    const u32 n = limb_len<L, TPI>(X, g), m = limb_len<L, TPI>(V, g);
    if (n < m) {
        R = X;
        return;
    }
    const u32 s = (u32)__clz((int)limb<L, TPI>(V, m - 1u, g));
    MMag<L> vn = V;
    (void)shl_bits<L, TPI>(vn, s, g);
    MMag<L> un = X;
    u32 ux = shl_bits<L, TPI>(un, s, g);   // numerator limb W
    const u32 v1 = limb<L, TPI>(vn, m - 1u, g);
    const u32 v0 = m >= 2u ? limb<L, TPI>(vn, m - 2u, g) : 0u;
    const u32 rv = recip_word(v1);
    MMag<L> vs = shl_limbs<L, TPI>(vn, n - m, g);   // vn << 32 j
    constexpr u32 W = (u32)(L * TPI);
    for (int j = (int)(n - m); j >= 0; j--) {
        const u32 top = (u32)j + m;
        const u32 u2 = top >= W ? ux : limb<L, TPI>(un, top, g);
        const u32 u1 = limb<L, TPI>(un, top - 1u, g);
        const u32 u0 = top >= 2u ? limb<L, TPI>(un, top - 2u, g) : 0u;
        // estimate (Knuth D3): q = min(floor((u2 b + u1) / v1), b - 1), then
        // at most two corrections against v0
        u32 qh;
        u64 rh;
        if (u2 >= v1) {
            qh = 0xffffffffu;
            rh = (u64)u1 + v1 + ((u64)(u2 - v1) << 32);
        } else {
            u32 r32;
            qh = div_2by1(u2, u1, v1, rv, r32);
            rh = r32;
        }
        while (rh < (1ull << 32) && (u64)qh * v0 > ((rh << 32) | u0)) {
            qh--;
            rh += v1;
        }
        // window -= qh * vs; the product's word above the group leaves into ux
        MMag<L> t = vs;
        const u32 tc = mul_word<L, TPI>(t, qh, g);
        const u32 bo = sub_mag<L, TPI>(un, t, g);
        const u64 need = (u64)tc + bo;
        const bool neg = (u64)ux < need;
        ux = (u32)((u64)ux - need);
        if (neg) {
            // add back: the estimate was one too large
            qh--;
            const u32 co = add_mag<L, TPI>(un, vs, 0u, g);
            ux += co;
        }
        set_limb<L, TPI>(Q, (u32)j, qh, g);
        shr_limb1<L, TPI>(vs, g);
    }
    R = un;
    shr_bits<L, TPI>(R, s, ux, g);
}

// ---------------------------------------------------------------- signed values

// sign rules of dadd (decimal.cu:1284-1323): same signs add and keep the
// sign; different signs subtract the smaller from the larger and take its
// sign, +0 on a tie
template <int L, int TPI> UPD_DEV MDec<L> add(const MDec<L>& a, const MDec<L>& b, const Grp<TPI>& g)
{
    const bool same = a.neg == b.neg;
    MMag<L> t = a.m, y;
#pragma unroll
    for (int i = 0; i < L; i++) y.w[i] = same ? b.m.w[i] : ~b.m.w[i];
    const u32 c = add_mag<L, TPI>(t, y, same ? 0u : 1u, g);
    const bool flip = !same && c == 0u;
    MDec<L> r;
    MMag<L> z = mzero<L>();
    if (flip) {
#pragma unroll
        for (int i = 0; i < L; i++) t.w[i] = ~t.w[i];
    }
    (void)add_mag<L, TPI>(t, z, flip ? 1u : 0u, g);
    r.m = t;
    const bool any = !is_zero<L, TPI>(t, g);
    r.neg = same ? a.neg : (flip ? b.neg : (any ? a.neg : 0u));
    return r;
}

template <int L, int TPI> UPD_DEV MDec<L> sub(const MDec<L>& a, const MDec<L>& b, const Grp<TPI>& g)
{
    MDec<L> nb = b;
    nb.neg ^= 1u;
    return add<L, TPI>(a, nb, g);
}

// DivOrMod, extended rule (ddiv_rt with faith = false): round half up of
// |a| 10^S / |b|, sign xor; 0 / x = +0; x / 0 keeps |a|.  |a| 10^S fits W limbs.
template <int L, int TPI> UPD_DEV MDec<L> div(const MDec<L>& a, const MDec<L>& b, u32 S, const Grp<TPI>& g)
{
    const bool azero = is_zero<L, TPI>(a.m, g), bzero = is_zero<L, TPI>(b.m, g);
    MDec<L> r;
    r.neg = a.neg ^ b.neg;
    if (azero) {
        r.m = mzero<L>();
        r.neg = 0u;
        return r;
    }
    if (bzero) {
        r.m = a.m;
        return r;
    }
    MMag<L> X = a.m;
    mul_p10<L, TPI>(X, S, g);
    MMag<L> Q, R;
    divmod<L, TPI>(X, b.m, Q, R, g);
    MMag<L> T = b.m;
    (void)sub_mag<L, TPI>(T, R, g);
    const u32 rnd = cmp<L, TPI>(R, T, g) >= 0 ? 1u : 0u;
    MMag<L> z = mzero<L>();
    (void)add_mag<L, TPI>(Q, z, rnd, g);
    r.m = Q;
    return r;
}

// ---------------------------------------------------------------- columns and sums

// one compact row of B bytes (B <= 4 W) at `row`: the lane of rank r reads
// the bytes of its limbs; the sign is bit 7 of byte B - 1
template <int L, int TPI>
UPD_DEV MDec<L> load_compact(const u8* __restrict__ col, usize row, u32 B, usize colbytes, const Grp<TPI>& g)
{
    const usize off = row * (usize)B;
    MDec<L> r;
#pragma unroll
    for (int i = 0; i < L; i++) {
        const u32 b0 = (g.rank * (u32)L + (u32)i) * 4u;   // first byte of this limb
        u32 x = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const u32 bi = b0 + (u32)k;
            const usize o = off + bi;
            if (bi < B && o < colbytes) x |= (u32)col[o] << (8 * k);
        }
        r.m.w[i] = x;
    }
    // the sign bit: bit 8 (B - 1) + 7 of the magnitude's bytes
    const u32 sb = 8u * (B - 1u) + 7u;
    const u32 owner = sb / (32u * (u32)L);
    const u32 loc = (sb / 32u) % (u32)L;
    const u32 w = g.bcast(sel<L>(r.m, loc), owner);
    r.neg = (w >> (sb & 31u)) & 1u;
    if (g.rank == owner) {
#pragma unroll
        for (int i = 0; i < L; i++)
            if ((u32)i == loc) r.m.w[i] &= ~(1u << (sb & 31u));
    }
    return r;
}

// load_compact with the row width B fixed at compile time (the generated
// multi-lane kernels): rows of a whole number of dwords are read as aligned
// dword loads (the column base is 4-byte aligned), others byte by byte
template <int L, int TPI, int B>
UPD_DEV MDec<L> load_compact_b(const u8* __restrict__ col, usize row, const Grp<TPI>& g)
{
    MDec<L> r;
    if constexpr (B % 4 == 0) {
        const u32* p = reinterpret_cast<const u32*>(col + row * (usize)B);
#pragma unroll
        for (int i = 0; i < L; i++) {
            const u32 k = g.rank * (u32)L + (u32)i;
            r.m.w[i] = k < (u32)(B / 4) ? p[k] : 0u;
        }
        constexpr u32 sb = 8u * (B - 1u) + 7u;
        constexpr u32 owner = sb / (32u * (u32)L), loc = (sb / 32u) % (u32)L;
        r.neg = (g.bcast(r.m.w[loc], owner) >> (sb & 31u)) & 1u;
        if (g.rank == owner) r.m.w[loc] &= ~(1u << (sb & 31u));
    } else {
        r = load_compact<L, TPI>(col, row, (u32)B, (row + 1) * (usize)B, g);
    }
    return r;
}

// a compact output row of B bytes (encode_compact): the magnitude's low B
// bytes, the sign ORed into bit 7 of byte B - 1; the lane of rank r writes
// the bytes of its limbs (whole dwords when B is a multiple of 4)
template <int L, int TPI, int B>
UPD_DEV void store_compact_b(u8* __restrict__ out, usize row, const MDec<L>& v, const Grp<TPI>& g)
{
    constexpr u32 sb = 8u * (B - 1u) + 7u;
    if constexpr (B % 4 == 0) {
        u32* p = reinterpret_cast<u32*>(out + row * (usize)B);
#pragma unroll
        for (int i = 0; i < L; i++) {
            const u32 k = g.rank * (u32)L + (u32)i;
            if (k < (u32)(B / 4)) p[k] = v.m.w[i] | (k == sb / 32u ? v.neg << (sb & 31u) : 0u);
        }
    } else {
        u8* o = out + row * (usize)B;
#pragma unroll
        for (int i = 0; i < L; i++) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const u32 bi = (g.rank * (u32)L + (u32)i) * 4u + (u32)k;
                if (bi < (u32)B) o[bi] = (u8)(((v.m.w[i] >> (8 * k)) & 0xffu) | (bi == (u32)B - 1u ? v.neg << 7 : 0u));
            }
        }
    }
}

// the compact round trip of a SUM input (stored_rt): keep 8 B - 1 magnitude
// bits, OR the bit at 8 B - 1 into the sign
template <int L, int TPI> UPD_DEV MDec<L> stored(const MDec<L>& v, u32 B, const Grp<TPI>& g)
{
    MDec<L> r = v;
    const u32 top = 8u * B - 1u;
    u32 bit = 0;
#pragma unroll
    for (int i = 0; i < L; i++) {
        const u32 lo = 32u * (g.rank * (u32)L + (u32)i);
        if (lo > top) {
            r.m.w[i] = 0u;
        } else if (lo + 32u > top) {
            const u32 k = top - lo;
            bit = (v.m.w[i] >> k) & 1u;
            r.m.w[i] &= k == 0u ? 0u : (0xffffffffu >> (32u - k));
        }
    }
    const u64 any = g.bits(__ballot(bit ? 1 : 0));
    r.neg |= any ? 1u : 0u;
    return r;
}

// two's complement accumulator of W + 2 limbs: the group's W limbs plus two
// words above them, kept (identically) on every lane of the group
template <int L> struct Acc {
    MMag<L> m;
    u32 x0, x1;
};

template <int L> UPD_DEV Acc<L> acc_zero()
{
    Acc<L> a;
    a.m = mzero<L>();
    a.x0 = a.x1 = 0u;
    return a;
}

// acc += (-1)^neg * |v|
template <int L, int TPI> UPD_DEV void acc_add(Acc<L>& acc, const MDec<L>& v, const Grp<TPI>& g)
{
    MMag<L> y;
    const u32 fm = v.neg ? 0xffffffffu : 0u;
#pragma unroll
    for (int i = 0; i < L; i++) y.w[i] = v.m.w[i] ^ fm;
    const u32 c = add_mag<L, TPI>(acc.m, y, v.neg ? 1u : 0u, g);
    const u64 s0 = (u64)acc.x0 + fm + c;
    acc.x0 = (u32)s0;
    acc.x1 = acc.x1 + fm + (u32)(s0 >> 32);
}

}  // namespace ml
}  // namespace updec
#endif
