/*
 * updec_oracle.c -- TEST INFRASTRUCTURE ONLY (parity oracle + CPU baseline).
 *
 * A plain-C restatement of the reference's host Decimal rules
 * (src/datatypes/decimal.cu).  Each function cites the reference lines it
 * follows.  The HIP product never links or loads this file; tests, the
 * package smoke check and bench.py's cpu_baseline leg do.
 *
 * Parity pins for every function live in tests/test_oracle_*.py: the
 * reference's own known-answer values (test/testcases/datatypes/
 * test_decimal.cu), an independent Python-integer restatement of the
 * closed forms, and brute force on tiny inputs.
 */
#include "updec_oracle.h"
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <pthread.h>

#define FW 32 /* faithful limb count: NUM_TOTAL_DIG (decimal.hxx:47) */

static int NW(int mode) { return mode == OD_FAITHFUL ? FW : OD_MAXW; }

static __thread int g_overflow = 0;
static int g_overflow_any = 0;
int od_overflow_flag(void) { return g_overflow || g_overflow_any; }
void od_clear_overflow(void) { g_overflow = 0; g_overflow_any = 0; }
static void flag_overflow(int mode) { if (mode == OD_EXTENDED) { g_overflow = 1; g_overflow_any = 1; } }

/* B(P): __PREC_BYTE_ARRAY (decimal.cu:92), a format constant of the compact
 * column layout, indexed by precision 0..299.  Past the table the closed form
 * ceil((bitlen(10^P - 1) + 1) / 8) is used (extended mode only). */
static const uint8_t PREC_BYTES[300] = {
0,1,1,2,2,3,3,4,4,4,5,5,6,6,6,7,7,8,8,9,9,9,10,10,11,11,11,12,12,13,13,14,14,14,15,15,16,16,16,17,
17,18,18,19,19,19,20,20,21,21,21,22,22,23,23,24,24,24,25,25,26,26,26,27,27,28,28,29,29,29,30,30,31,31,31,32,
32,33,33,34,34,34,35,35,36,36,36,37,37,38,38,38,39,39,40,40,41,41,41,42,42,43,43,43,44,44,45,45,46,46,
46,47,47,48,48,48,49,49,50,50,51,51,51,52,52,53,53,53,54,54,55,55,56,56,56,57,57,58,58,58,59,59,60,60,
61,61,61,62,62,63,63,63,64,64,65,65,66,66,66,67,67,68,68,68,69,69,70,70,71,71,71,72,72,73,73,73,74,74,
75,75,76,76,76,77,77,78,78,78,79,79,80,80,81,81,81,82,82,83,83,83,84,84,85,85,86,86,86,87,87,88,88,88,
89,89,90,90,91,91,91,92,92,93,93,93,94,94,95,95,96,96,96,97,97,98,98,98,99,99,100,100,101,101,101,102,102,103,
103,103,104,104,105,105,105,106,106,107,107,108,108,108,109,109,110,110,110,111,111,112,112,113,113,113,114,114,115,115,
115,116,116,117,117,118,118,118,119,119,120,120,120,121,121,122,122,123,123,123,124,124,125,125};

static int bl32(uint32_t x) { return x ? 32 - __builtin_clz(x) : 0; }
static int top_word(const uint32_t* v, int nw) { int i = nw - 1; while (i >= 0 && v[i] == 0) i--; return i; }
static int bitlen_w(const uint32_t* v, int nw) { int t = top_word(v, nw); return t < 0 ? 0 : 32 * t + bl32(v[t]); }

/* closed form past the table: bytes for bitlen(10^P - 1) magnitude bits plus
 * the sign bit.  10^P - 1 has the bit length of 10^P for P >= 1 (10^P is not
 * a power of two), so one pass of exact limb arithmetic fills every P. */
#define OD_MAX_DIGITS 20000
#define OD_TABLE_WORDS 2100   /* 10^20000 < 2^66440 */
static int g_bytes_past[OD_MAX_DIGITS + 1];
static pthread_once_t g_bytes_once = PTHREAD_ONCE_INIT;
static void fill_bytes_past(void)
{
    static uint32_t v[OD_TABLE_WORDS];
    int nw = OD_TABLE_WORDS;
    memset(v, 0, sizeof v);
    v[0] = 1;
    int top = 1;   /* words in use */
    for (int P = 1; P <= OD_MAX_DIGITS; P++) {
        uint64_t c = 0;
        for (int i = 0; i < top + 1 && i < nw; i++) { uint64_t t = (uint64_t)v[i] * 10u + c; v[i] = (uint32_t)t; c = t >> 32; }
        if (top < nw && v[top]) top++;
        g_bytes_past[P] = (bitlen_w(v, top) + 1 + 7) / 8;
    }
}

int od_compact_bytes(int P)
{
    if (P < 0) return 0;
    if (P < 300) return PREC_BYTES[P];
    if (P > OD_MAX_DIGITS) return 0;
    pthread_once(&g_bytes_once, fill_bytes_past);
    return g_bytes_past[P];
}

void od_zero(od_dec* d) { memset(d, 0, sizeof *d); }
int od_is_zero(const od_dec* d) { for (int i = 0; i < OD_MAXW; i++) if (d->v[i]) return 0; return 1; }
int od_bitlen(const od_dec* d) { return bitlen_w(d->v, OD_MAXW); }

/* ---------------- limb helpers ---------------- */

/* v = v * m + a (mod 2^(32nw)); returns carry out */
static uint32_t mul_small(uint32_t* v, int nw, uint32_t m, uint32_t a)
{
    uint64_t c = a;
    for (int i = 0; i < nw; i++) { uint64_t t = (uint64_t)v[i] * m + c; v[i] = (uint32_t)t; c = t >> 32; }
    return (uint32_t)c;
}

static int top_word(const uint32_t* v, int nw);

/* v = floor(v / d); returns remainder (words above the top nonzero one stay 0) */
static uint32_t div_small(uint32_t* v, int nw, uint32_t d)
{
    uint64_t r = 0;
    for (int i = top_word(v, nw); i >= 0; i--) { uint64_t t = (r << 32) | v[i]; v[i] = (uint32_t)(t / d); r = t % d; }
    return (uint32_t)r;
}

/* div_small by the constant 10^9 (the compiler turns it into a multiply) */
static uint32_t div_1e9(uint32_t* v, int nw)
{
    uint64_t r = 0;
    for (int i = top_word(v, nw); i >= 0; i--) {
        uint64_t t = (r << 32) | v[i];
        v[i] = (uint32_t)(t / 1000000000u);
        r = t % 1000000000u;
    }
    return (uint32_t)r;
}

static const uint32_t POW10_9[10] = {1u,10u,100u,1000u,10000u,100000u,1000000u,10000000u,100000000u,1000000000u};

/* v = v * 10^n mod 2^(32nw).  The reference multiplies by a multi-word 10^n
 * row (decimal.cu:2028-2058); the product mod 2^(32nw) is the same when
 * computed as a chain of word multiplies. */
static void mul10n_w(uint32_t* v, int nw, int n, int mode)
{
    while (n > 0) {
        int k = n > 9 ? 9 : n;
        /* words above top + 1 stay zero: a word product carries at most one word up */
        int used = top_word(v, nw) + 2;
        if (used > nw) used = nw;
        if (mul_small(v, used, POW10_9[k], 0) && used == nw) flag_overflow(mode);
        n -= k;
    }
}

static int cmp_w(const uint32_t* a, const uint32_t* b, int nw)
{
    for (int i = nw - 1; i >= 0; i--) if (a[i] != b[i]) return a[i] > b[i] ? 1 : -1;
    return 0;
}

static uint32_t add_w(uint32_t* r, const uint32_t* a, const uint32_t* b, int nw)
{
    uint64_t c = 0;
    for (int i = 0; i < nw; i++) { uint64_t t = (uint64_t)a[i] + b[i] + c; r[i] = (uint32_t)t; c = t >> 32; }
    return (uint32_t)c;
}

/* r = a - b (a >= b assumed by callers) */
static void sub_w(uint32_t* r, const uint32_t* a, const uint32_t* b, int nw)
{
    uint64_t br = 0;
    for (int i = 0; i < nw; i++) { uint64_t t = (uint64_t)a[i] - b[i] - br; r[i] = (uint32_t)t; br = (t >> 63) & 1; }
}

/* r[0..na+nb) = a * b, schoolbook base 2^32 (decimal.cu:1325-1351) */
static void mul_full(uint32_t* r, const uint32_t* a, int na, const uint32_t* b, int nb)
{
    memset(r, 0, sizeof(uint32_t) * (size_t)(na + nb));
    for (int i = 0; i < na; i++) {
        if (!a[i]) continue;
        uint64_t c = 0;
        for (int j = 0; j < nb; j++) {
            uint64_t t = (uint64_t)a[i] * b[j] + r[i + j] + c;
            r[i + j] = (uint32_t)t; c = t >> 32;
        }
        r[i + nb] = (uint32_t)c;
    }
}

/* Knuth, TAOCP vol. 2, 4.3.1 Algorithm D (long division, base 2^32).
 * q[0..m-n] and r[0..n-1]; u has m words, v has n words with v[n-1] != 0, m >= n. */
static void knuth_d(uint32_t* q, uint32_t* r, const uint32_t* u, int m, const uint32_t* v, int n)
{
    if (n == 1) {
        uint64_t k = 0;
        for (int j = m - 1; j >= 0; j--) { uint64_t t = (k << 32) | u[j]; q[j] = (uint32_t)(t / v[0]); k = t % v[0]; }
        r[0] = (uint32_t)k;
        return;
    }
    uint32_t* un = (uint32_t*)calloc((size_t)m + 1, sizeof(uint32_t));
    uint32_t* vn = (uint32_t*)calloc((size_t)n, sizeof(uint32_t));
    int s = __builtin_clz(v[n - 1]);
    for (int i = n - 1; i > 0; i--) vn[i] = (v[i] << s) | (s ? v[i - 1] >> (32 - s) : 0);
    vn[0] = v[0] << s;
    un[m] = s ? u[m - 1] >> (32 - s) : 0;
    for (int i = m - 1; i > 0; i--) un[i] = (u[i] << s) | (s ? u[i - 1] >> (32 - s) : 0);
    un[0] = u[0] << s;
    const uint64_t B = 1ull << 32;
    for (int j = m - n; j >= 0; j--) {
        uint64_t num = ((uint64_t)un[j + n] << 32) | un[j + n - 1];
        uint64_t qhat = num / vn[n - 1];
        uint64_t rhat = num % vn[n - 1];
        while (qhat >= B || qhat * vn[n - 2] > ((rhat << 32) | un[j + n - 2])) {
            qhat--; rhat += vn[n - 1];
            if (rhat >= B) break;
        }
        int64_t t; uint64_t k = 0;
        for (int i = 0; i < n; i++) {
            uint64_t p = qhat * vn[i];
            t = (int64_t)un[i + j] - (int64_t)k - (int64_t)(p & 0xffffffffull);
            un[i + j] = (uint32_t)t;
            k = (p >> 32) - (uint64_t)(t >> 32);
        }
        t = (int64_t)un[j + n] - (int64_t)k;
        un[j + n] = (uint32_t)t;
        q[j] = (uint32_t)qhat;
        if (t < 0) {
            q[j]--;
            uint64_t c = 0;
            for (int i = 0; i < n; i++) { uint64_t tt = (uint64_t)un[i + j] + vn[i] + c; un[i + j] = (uint32_t)tt; c = tt >> 32; }
            un[j + n] += (uint32_t)c;
        }
    }
    for (int i = 0; i < n - 1; i++) r[i] = (un[i] >> s) | (s ? un[i + 1] << (32 - s) : 0);
    r[n - 1] = un[n - 1] >> s;
    free(un); free(vn);
}

/* ---------------- codec ---------------- */

/* Decimal(const CompactDecimal*, P, S) (decimal.cu:168-176): copy B(P) bytes
 * little-endian into the magnitude, take the sign from bit 7 of the last
 * byte and clear it. */
void od_from_compact(od_dec* d, const uint8_t* buf, int P, int S)
{
    od_zero(d);
    d->prec = P; d->frac = S;
    int len = od_compact_bytes(P);
    uint8_t* mv = (uint8_t*)d->v;
    int keep = len < (int)sizeof d->v ? len : (int)sizeof d->v;
    memcpy(mv, buf, (size_t)keep);
    if (len > 0) {
        d->sign = (buf[len - 1] >> 7) & 1;
        if (len == keep) mv[len - 1] &= 0x7f;
    }
}

/* ToCompactDecimal (decimal.cu:337-341): low `len` bytes, then OR the sign
 * into bit 7 of the last byte (a magnitude bit there is kept and merged). */
void od_to_compact(const od_dec* d, uint8_t* buf, int len)
{
    int keep = len < (int)sizeof d->v ? len : (int)sizeof d->v;
    memcpy(buf, (const uint8_t*)d->v, (size_t)keep);
    if (len > keep) memset(buf + keep, 0, (size_t)(len - keep));
    if (len > 0) buf[len - 1] |= (uint8_t)(d->sign << 7);
}

/* v += a (a sign-extended) mod 2^(32 nw) */
static void add_signed_small(uint32_t* v, int nw, int32_t a)
{
    const uint32_t ext = a < 0 ? 0xffffffffu : 0u;
    uint64_t c = 0;
    for (int i = 0; i < nw; i++) {
        uint64_t t = (uint64_t)v[i] + (i == 0 ? (uint32_t)a : ext) + c;
        v[i] = (uint32_t)t;
        c = t >> 32;
    }
}

/* StringToDecimal (decimal.cu:2980-3025): every character c after an
 * optional leading '-' is accumulated as v = v*10 + (c - '0') in base 2^32,
 * mod 2^(32nw), with (c - '0') signed (a character that is not a digit
 * still shifts and offsets the value, as in the reference); a '.' sets the
 * fraction flag and the character after it is taken as a digit whatever it
 * is; every character after the first '.' counts one fractional digit;
 * prec = characters taken.  A '.' that ends the string ends it: the
 * reference goes on to read its terminator as a digit and past it. */
int od_from_string(od_dec* d, const char* s, int mode)
{
    int nw = NW(mode);
    od_zero(d);
    int t = 0, flag = 0, dotflag = 0, dot = 0, i;
    if (s[0] == '-') { flag = 1; t = 1; }
    for (i = t; s[i] != '\0'; i++) {
        if (s[i] == '.') { dotflag = 1; i++; if (s[i] == '\0') break; }
        if (dotflag) dot++;
        if (mul_small(d->v, nw, 10u, 0u)) flag_overflow(mode);
        add_signed_small(d->v, nw, (int32_t)(signed char)s[i] - '0');
    }
    d->sign = flag;
    d->prec = i - flag - dotflag;
    d->frac = dot;
    return 1;
}

/* cast (decimal.cu:471-533): scale up multiplies by 10^n, scale down is a
 * truncating floor division by 10^n; the sign is kept. */
void od_cast(od_dec* d, int P, int S, int mode)
{
    int nw = NW(mode);
    if (S >= d->frac) mul10n_w(d->v, nw, S - d->frac, mode);
    else od_shr10(d, d->frac - S, mode);
    d->prec = P; d->frac = S;
}

/* Decimal::truncate(p) (decimal.cu:541-570), with prec and frac the value's:
 *   p > 0: to scale p, floor-dividing by 10^(frac - p) or multiplying by
 *          10^(p - frac); prec moves by the same count; frac = p;
 *   p <= 0, q = -p, intg = prec - frac: when intg > q the magnitude becomes
 *          floor(M / 10^(frac + q)) * 10^q and prec = intg; otherwise it is
 *          zeroed (the sign is kept) and prec = 1; frac = 0. */
void od_truncate(od_dec* d, int p, int mode)
{
    if (p > 0) {
        if (d->frac >= p) {
            od_shr10(d, d->frac - p, mode);
            d->prec -= d->frac - p;
        } else {
            od_shl10(d, p - d->frac, mode);
            d->prec += p - d->frac;
        }
        d->frac = p;
    } else {
        int q = -p;
        int intg = d->prec - d->frac;
        if (intg > q) {
            od_shr10(d, d->frac + q, mode);
            od_shl10(d, q, mode);
            d->prec = intg;
        } else {
            memset(d->v, 0, sizeof d->v);
            d->prec = 1;
        }
        d->frac = 0;
    }
}

/* Decimal(int32_t t, int32_t scale, ...) (decimal.cu:230-260): |t| * 10^scale, frac = scale */
void od_from_int32_scale(od_dec* d, int32_t t, int32_t scale, int mode)
{
    od_zero(d);
    int prec = 10;
    if (scale > prec) prec += scale;
    d->prec = prec; d->frac = scale;
    int64_t tt = t;
    if (tt < 0) { d->sign = 1; tt = -tt; }
    d->v[0] = (uint32_t)tt;
    mul10n_w(d->v, NW(mode), scale, mode);
}

/* Decimal(int64_t) (decimal.cu:301-309) */
void od_from_int64(od_dec* d, int64_t t)
{
    od_zero(d);
    d->prec = 19; d->frac = 0;
    uint64_t u;
    if (t < 0) { d->sign = 1; u = (uint64_t)(-(t + 1)) + 1u; } else u = (uint64_t)t;
    d->v[0] = (uint32_t)u; d->v[1] = (uint32_t)(u >> 32);
}

/* Decimal(Decimal t, int32_t scale, ...) (decimal.cu:263-268): t << (scale - t.frac), frac = scale */
void od_rescale_up(od_dec* d, int scale, int mode)
{
    int gap = scale - d->frac;
    if (gap > 0) mul10n_w(d->v, NW(mode), gap, mode);
    d->frac = scale;
    d->prec = d->prec + gap;
}

void od_shl10(od_dec* d, int n, int mode) { if (n > 0) mul10n_w(d->v, NW(mode), n, mode); }

/* operator>> (decimal.cu:2003-2026): floor(M / 10^n) */
void od_shr10(od_dec* d, int n, int mode)
{
    int nw = NW(mode);
    while (n > 9) { div_1e9(d->v, nw); n -= 9; }
    if (n > 0) div_small(d->v, nw, POW10_9[n]);
}

/* operator-() (decimal.cu:578-586): flips the sign, also for zero */
void od_neg(od_dec* d) { d->sign = d->sign ? 0 : 1; }

/* AlignAddSubData (decimal.cu:2905-2920): the operand with the smaller
 * scale is multiplied up to the larger one. */
static void align_pair(od_dec* x, od_dec* y, int mode)
{
    int nw = NW(mode);
    if (x->frac == y->frac) return;
    if (x->frac < y->frac) { mul10n_w(x->v, nw, y->frac - x->frac, mode); x->frac = y->frac; }
    else { mul10n_w(y->v, nw, x->frac - y->frac, mode); y->frac = x->frac; }
}

/* AddDecimalOnCpu (decimal.cu:1284-1323): align; same sign -> add magnitudes
 * (carry out of the top limb dropped); different signs -> larger minus
 * smaller, sign of the larger, +0 on a tie. */
void od_add(od_dec* a, const od_dec* b, int mode)
{
    int nw = NW(mode);
    od_dec t = *b;
    align_pair(&t, a, mode);
    /* words above both operands' top words are zero and stay zero, except
     * for one carry word */
    int ta = top_word(a->v, nw), tb = top_word(t.v, nw);
    int used = (ta > tb ? ta : tb) + 2;
    if (used > nw) used = nw;
    if (a->sign == t.sign) {
        if (add_w(a->v, a->v, t.v, used) && used == nw) flag_overflow(mode);
    } else {
        int r = cmp_w(a->v, t.v, used);
        if (r >= 0) sub_w(a->v, a->v, t.v, used);
        else sub_w(a->v, t.v, a->v, used);
        a->sign = (r > 0 && !b->sign) || (r < 0 && b->sign);
    }
}

/* operator-= (decimal.cu:1575-1613): identical to adding the negated operand */
void od_sub(od_dec* a, const od_dec* b, int mode)
{
    od_dec nb = *b;
    od_neg(&nb);
    od_add(a, &nb, mode);
}

/* MulDecimalOnCpu (decimal.cu:1325-1351): sign xor, frac add, low limbs of the product */
void od_mul(od_dec* a, const od_dec* b, int mode)
{
    int nw = NW(mode);
    a->sign = b->sign ^ a->sign;
    a->frac = a->frac + b->frac;
    int na = top_word(a->v, nw) + 1, nb = top_word(b->v, nw) + 1;
    if (na == 0 || nb == 0) { memset(a->v, 0, sizeof a->v); return; }
    uint32_t r[2 * OD_MAXW];
    mul_full(r, a->v, na, b->v, nb);
    for (int i = nw; i < na + nb; i++) if (r[i]) { flag_overflow(mode); break; }
    memset(a->v, 0, sizeof a->v);
    int nc = na + nb < nw ? na + nb : nw;
    memcpy(a->v, r, sizeof(uint32_t) * (size_t)nc);
}

/* compare (decimal.cu:638-745): signs first, then aligned magnitudes.
 * Returns -1, 0, 1 for a<b, a==b, a>b in the reference's own sense
 * (+0 and -0 compare by sign bit, as operator< does). */
int od_cmp(const od_dec* a, const od_dec* b, int mode)
{
    int nw = NW(mode);
    if (a->sign != b->sign) return a->sign == 0 ? 1 : -1;
    od_dec l = *a, r = *b;
    align_pair(&l, &r, mode);
    int c = cmp_w(l.v, r.v, nw);
    return a->sign == 0 ? c : -c;
}

/* the reference's estimate of bitlen(10^shift) (decimal.cu:2252-2256) */
static uint32_t hit_shift_bits(uint32_t shift)
{
    uint32_t h = (shift / 3) * 10;
    if (shift % 3 == 1) h += 4;
    else if (shift % 3 == 2) h += 7;
    return h;
}

/* DivOrMod, isMod = false (decimal.cu:2201-2333) with DivByInt64
 * (:2431-2446), DivByInt (:2335-2429) and DivInt (:2125-2191).
 *   result frac = Sa + 4, sign = xor; 0 / x -> +0; x / 0 -> error = 2, |x| kept.
 *   Q = floor(|A| * 10^(Sb+4) / |B|), R = remainder; Q += 1 when 2R >= |B|.
 * Faithful mode keeps the three path quirks:
 *   (i)  bitlen(A) + est(10^s) - 1 <= 64 and bitlen(B) <= 64: the numerator
 *        and the doubled remainder are uint64 and wrap mod 2^64;
 *   (ii) |B| < 2^32: the round-up test is (uint32)(2R) >= |B| (wraps for R >= 2^31);
 *   (iii) otherwise, binary-search division (DivInt): the lower bound of the
 *        search is 2^max(0, bitlen(X) - bitlen(B) - 1), so when X is shorter
 *        than B by 1 to 64 bits the search starts and ends at 1 and the
 *        negative residual X - B blocks the round-up: the result is 1.  (The
 *        upper bound is then a shift by a negative count, which PTX clamps to
 *        0, or for a gap of 33 to 64 bits a store into qmax's own header
 *        bytes; neither moves the lower bound.)  A gap above 64 bits stores
 *        outside qmax: no defined result; the exact rule is kept. */
void od_div(od_dec* a, const od_dec* b, int mode)
{
    int nw = NW(mode);
    od_dec A = *a;
    a->sign = b->sign ^ a->sign;
    a->frac += 4;
    if (od_is_zero(a)) { a->sign = 0; a->prec = a->frac + 1; return; }
    if (od_is_zero(b)) { a->error = 2; return; }
    int s = b->frac + 4;
    int na = top_word(A.v, nw) + 1, nb = top_word(b->v, nw) + 1;
    uint32_t hitDt = (uint32_t)bitlen_w(A.v, nw), hitDs = (uint32_t)bitlen_w(b->v, nw);
    if (mode == OD_FAITHFUL && hitDt + hit_shift_bits((uint32_t)s) - 1 <= 64 && hitDs <= 64) {
        uint64_t dvs = (uint64_t)A.v[0] | ((uint64_t)A.v[1] << 32);
        int sh = s;
        while (sh > 9) { dvs *= 1000000000ull; sh -= 9; }
        dvs *= POW10_9[sh];
        uint64_t dvt = (uint64_t)b->v[0] | ((uint64_t)b->v[1] << 32);
        uint64_t res = dvs / dvt + ((((dvs % dvt) << 1) >= dvt) ? 1 : 0);
        memset(a->v, 0, sizeof a->v);
        a->v[0] = (uint32_t)res; a->v[1] = (uint32_t)(res >> 32);
        return;
    }
    /* X = |A| * 10^s, exact (faithful inputs are classified so that X < 2^1024) */
    int xw = na + (s * 4) / 38 + 2; /* 10^s < 2^(3.33 s) -> at most s/9.6 + 1 extra words */
    if (xw > 2 * OD_MAXW) xw = 2 * OD_MAXW;
    uint32_t X[2 * OD_MAXW + 2];
    memset(X, 0, sizeof X);
    memcpy(X, A.v, sizeof(uint32_t) * (size_t)na);
    mul10n_w(X, xw, s, OD_FAITHFUL);
    int nx = top_word(X, xw) + 1;
    uint32_t Q[2 * OD_MAXW + 2], R[OD_MAXW];
    memset(Q, 0, sizeof Q); memset(R, 0, sizeof R);
    int bx = bitlen_w(X, nx), bb = bitlen_w(b->v, nb);
    if (nx < nb) { memcpy(R, X, sizeof(uint32_t) * (size_t)nx); }
    else knuth_d(Q, R, X, nx, b->v, nb);
    int roundup;
    if (mode == OD_FAITHFUL && nb == 1) {
        roundup = (uint32_t)(R[0] * 2u) >= b->v[0];
    } else if (mode == OD_FAITHFUL && bx < bb && bb - bx <= 64) {
        memset(Q, 0, sizeof Q); Q[0] = 1; roundup = 0;
    } else {
        /* 2R >= B  <=>  R >= B - R */
        uint32_t T[OD_MAXW];
        sub_w(T, b->v, R, nb);
        roundup = cmp_w(R, T, nb) >= 0;
    }
    if (roundup) {
        for (int i = 0; i < 2 * OD_MAXW + 2; i++) { if (++Q[i]) break; }
    }
    int nq = top_word(Q, 2 * OD_MAXW + 2) + 1;
    if (nq > nw) flag_overflow(mode);
    memset(a->v, 0, sizeof a->v);
    memcpy(a->v, Q, sizeof(uint32_t) * (size_t)(nq < nw ? nq : nw));
}

void od_to_nullable_image(const od_dec* d, uint8_t* out)
{
    out[0] = 1;
    out[1] = (uint8_t)d->sign;
    out[2] = (uint8_t)d->prec;
    out[3] = (uint8_t)d->frac;
    out[4] = (uint8_t)d->error;
    memcpy(out + 5, d->v, 32 * sizeof(uint32_t));
}

static void from_nullable_image(od_dec* d, const uint8_t* in)
{
    od_zero(d);
    d->sign = in[1]; d->prec = in[2]; d->frac = in[3]; d->error = in[4];
    memcpy(d->v, in + 5, 32 * sizeof(uint32_t));
}

/* ModCalc (decimal.cu:2621-2716), operator %= of two decimals:
 *   sign = xor, frac = 0, prec = b.prec; 0 % x -> +0.  The scales are
 *   dropped: A = |a| and B = |b| are taken as integers.  Faithful paths:
 *   (i)  bitlen(A) <= 64 and bitlen(B) <= 64: ModByInt64 (:2682-2693), A % B in uint64;
 *   (ii) B < 2^32: ModByInt (:2695-2714), one-limb long division;
 *   (iii) DivInt (:2125-2191): bisection for the quotient between
 *        2^max(0, la-lb-1) and 2^(la-lb+1) with the midpoint
 *        floor((lo+hi)/2) (HalfIntDecimal :2093-2119), products and
 *        comparisons in the 32-limb arithmetic, then the magnitude of the
 *        signed residual A - q*B (CopyValue keeps no sign).  For
 *        2 <= lb - la <= 64 the search starts and ends at q = 1 (the upper
 *        bound is a negative shift, 0 on PTX, or a store into qmax's own
 *        header bytes), so the result is B - A.  For lb - la > 64 the
 *        reference stores outside qmax: no defined result; here the
 *        remainder A.
 *   x % 0: the reference scans below the divisor's limbs: no defined
 *   result; here A is kept, as x / 0 keeps it.
 * Extended mode: A % B exactly (x % 0 -> A), same sign and scale rules. */
void od_mod(od_dec* a, const od_dec* b, int mode)
{
    int nw = NW(mode);
    od_dec A = *a;
    a->sign = b->sign ^ a->sign;
    a->prec = b->prec;
    a->frac = 0;
    if (od_is_zero(&A)) { a->sign = 0; a->prec = 1; return; }
    if (od_is_zero(b)) return;
    int la = bitlen_w(A.v, nw), lb = bitlen_w(b->v, nw);
    int na = top_word(A.v, nw) + 1, nb = top_word(b->v, nw) + 1;
    uint32_t R[OD_MAXW];
    memset(R, 0, sizeof R);
    if (mode == OD_FAITHFUL && la <= 64 && lb <= 64) {
        uint64_t dvs = (uint64_t)A.v[0] | ((uint64_t)A.v[1] << 32);
        uint64_t dvt = (uint64_t)b->v[0] | ((uint64_t)b->v[1] << 32);
        uint64_t res = dvs % dvt;
        R[0] = (uint32_t)res; R[1] = (uint32_t)(res >> 32);
    } else if (mode == OD_FAITHFUL && nb == 1) {
        uint32_t rem = 0;
        for (int i = nw - 1; i >= 0; i--) {
            if (A.v[i] || rem) {
                uint64_t t = (uint64_t)A.v[i] + (uint64_t)rem * 4294967296ull;
                rem = (uint32_t)(t % b->v[0]);
            }
        }
        R[0] = rem;
    } else if (mode == OD_FAITHFUL && la + 1 < lb && lb - la <= 64) {
        /* DivInt with qmin = 1 and no bisection step: |A - B| = B - A */
        sub_w(R, b->v, A.v, nw);
    } else if (mode == OD_FAITHFUL && la + 1 >= lb) {
        od_dec ds = A, dt = *b, qmin, qmax, qmid, t;
        ds.sign = dt.sign = 0; ds.frac = dt.frac = 0;
        int pmax = la - lb + 1, pmin = la - lb - 1;
        if (pmin < 0) pmin = 0;
        od_zero(&qmin); od_zero(&qmax);
        qmax.v[pmax / 32] = 1u << (pmax % 32);
        qmin.v[pmin / 32] = 1u << (pmin % 32);
        while (cmp_w(qmin.v, qmax.v, nw) < 0) {
            od_zero(&qmid);
            add_w(qmid.v, qmin.v, qmax.v, nw);   /* HalfIntDecimal: (qmin + qmax) / 2, 32 limbs */
            uint32_t carry = 0;
            for (int i = nw - 1; i >= 0; i--) {
                uint64_t x = ((uint64_t)carry << 32) | qmid.v[i];
                qmid.v[i] = (uint32_t)(x >> 1);
                carry = (uint32_t)(x & 1u);
            }
            if (cmp_w(qmid.v, qmin.v, nw) == 0) break;
            t = qmid;
            od_mul(&t, &dt, OD_FAITHFUL);
            int c = cmp_w(t.v, ds.v, nw);
            if (c == 0) { qmin = qmid; break; }
            if (c < 0) qmin = qmid; else qmax = qmid;
        }
        t = qmin;
        od_mul(&t, &dt, OD_FAITHFUL);
        if (cmp_w(ds.v, t.v, nw) >= 0) sub_w(R, ds.v, t.v, nw);
        else sub_w(R, t.v, ds.v, nw);
    } else if (na < nb) {
        memcpy(R, A.v, sizeof(uint32_t) * (size_t)na);
    } else {
        uint32_t Q[OD_MAXW + 1];
        memset(Q, 0, sizeof Q);
        knuth_d(Q, R, A.v, na, b->v, nb);
    }
    memset(a->v, 0, sizeof a->v);
    memcpy(a->v, R, sizeof(uint32_t) * (size_t)nw);
}

/* GetDecimal (decimal.cu:362-397): '+' or '-', then 8 upper-case hex
 * digits per limb from the top nonzero limb of the reference's 32 down to
 * limb 0 (none for a zero magnitude), then "(" three decimal digits of frac
 * ")".  out needs 8 * 32 + 7 bytes; returns the length without the NUL. */
int od_format_hex(const od_dec* d, char* out)
{
    static const char dict[] = "0123456789ABCDEF";
    int idx = 0, flag = 0;
    out[idx] = d->sign ? '-' : '+';
    for (int i = 31; i >= 0; i--) {
        if (d->v[i] != 0) flag = 1;
        if (flag) {
            uint32_t t = d->v[i];
            idx += 8;
            for (int j = 0; j < 8; j++) { out[idx - j] = dict[t & 0xF]; t >>= 4; }
        }
    }
    idx++;
    out[idx++] = '(';
    out[idx++] = (char)(d->frac / 100 + '0');
    out[idx++] = (char)(d->frac / 10 % 10 + '0');
    out[idx++] = (char)(d->frac % 10 + '0');
    out[idx++] = ')';
    out[idx] = '\0';
    return idx;
}

/* Bulk ingest of one text column (importCsvFile.cpp:640-660): row r is
 * chars[offsets[r] .. offsets[r+1]), cut at its first NUL; each row is
 * Decimal(P, S, text) (decimal.cu:154-165: StringToDecimal, then cast) in
 * the reference's 32 limbs, written with ToCompactDecimal
 * (decimal.cu:337-341) as B(P) bytes at out + r * B(P). */
void od_parse_compact(const char* chars, const int64_t* offsets, int64_t rows, int P, int S, uint8_t* out)
{
    const int B = od_compact_bytes(P);
    char* buf = NULL;
    size_t cap = 0;
    od_dec d;
    for (int64_t r = 0; r < rows; r++) {
        const size_t len = (size_t)(offsets[r + 1] - offsets[r]);
        if (len + 1 > cap) { cap = len + 64; buf = (char*)realloc(buf, cap); }
        memcpy(buf, chars + offsets[r], len);
        buf[len] = '\0';
        od_from_string(&d, buf, OD_FAITHFUL);
        od_cast(&d, P, S, OD_FAITHFUL);
        od_to_compact(&d, out + (size_t)r * (size_t)B, B);
    }
    free(buf);
}

/* Bulk output of one compact column: Decimal(compact, P, S)
 * (decimal.cu:167-176) of each row, formatted by GetDecimal (hex = 1,
 * decimal.cu:362-397) or as decimal text (hex = 0, od_to_string), into
 * stride-byte slots, NUL-terminated and zero-filled. */
void od_format_compact(const uint8_t* col, int64_t rows, int P, int S, int hex, char* out, int64_t stride)
{
    const int B = od_compact_bytes(P);
    od_dec d;
    char tmp[OD_MAXW * 10 + 16];
    for (int64_t r = 0; r < rows; r++) {
        od_from_compact(&d, col + (size_t)r * (size_t)B, P, S);
        int n = hex ? od_format_hex(&d, tmp) : od_to_string(&d, tmp, (int)sizeof tmp);
        char* o = out + (size_t)r * (size_t)stride;
        memset(o, 0, (size_t)stride);
        memcpy(o, tmp, (size_t)(n < stride ? n : stride - 1));
    }
}

int od_to_string(const od_dec* d, char* buf, int cap)
{
    /* decimal digits of the magnitude, then insert the point */
    uint32_t t[OD_MAXW];
    memcpy(t, d->v, sizeof t);
    char dig[OD_MAXW * 10 + 8];
    int n = 0;
    int nw = top_word(t, OD_MAXW) + 1;
    while (nw > 0) {
        uint32_t r = div_small(t, nw, 1000000000u);
        for (int k = 0; k < 9; k++) { dig[n++] = (char)('0' + r % 10); r /= 10; }
        nw = top_word(t, nw) + 1;
    }
    while (n > 1 && dig[n - 1] == '0') n--;
    if (n == 0) dig[n++] = '0';
    while (n <= d->frac) dig[n++] = '0';
    int o = 0;
    if (d->sign && o < cap) buf[o++] = '-';
    for (int i = n - 1; i >= 0; i--) {
        if (o < cap) buf[o++] = dig[i];
        if (i == d->frac && d->frac > 0 && o < cap) buf[o++] = '.';
    }
    if (o < cap) buf[o] = 0; else if (cap > 0) buf[cap - 1] = 0;
    return o;
}

/* ---------------- plan ---------------- */

enum { N_COL, N_CONST, N_NEG, N_ADD, N_SUB, N_MUL, N_DIV, N_CAST, N_MOD, N_TRUNC };

typedef struct pnode {
    int kind, P, S, col, a, b;
    int tp;            /* N_TRUNC: the truncate argument p */
    od_dec k;
} pnode;

struct od_plan {
    pnode* n;
    int count, cap, root;
};

static int pnew(od_plan* p)
{
    if (p->count == p->cap) {
        p->cap = p->cap ? p->cap * 2 : 16;
        p->n = (pnode*)realloc(p->n, sizeof(pnode) * (size_t)p->cap);
    }
    memset(&p->n[p->count], 0, sizeof(pnode));
    return p->count++;
}

static const char* skip_ws(const char* s) { while (*s == ' ') s++; return s; }

static const char* parse_node(od_plan* p, const char* s, int* out)
{
    s = skip_ws(s);
    if (*s == 'c') {
        int id = pnew(p);
        int col, P, S, used = 0;
        if (sscanf(s, "c%d:%d:%d%n", &col, &P, &S, &used) != 3) return NULL;
        p->n[id].kind = N_COL; p->n[id].col = col; p->n[id].P = P; p->n[id].S = S;
        *out = id;
        return s + used;
    }
    if (*s == 'k') {
        int id = pnew(p);
        pnode* nd = &p->n[id];
        nd->kind = N_CONST;
        s++;
        int neg = 0;
        if (*s == '-') { neg = 1; s++; } else if (*s == '+') s++;
        od_zero(&nd->k);
        while ((*s >= '0' && *s <= '9') || (*s >= 'a' && *s <= 'f') || (*s >= 'A' && *s <= 'F')) {
            uint32_t hv = (*s <= '9') ? (uint32_t)(*s - '0') : (uint32_t)((*s | 0x20) - 'a' + 10);
            mul_small(nd->k.v, OD_MAXW, 16u, hv);
            s++;
        }
        if (*s != ':') return NULL;
        s++;
        int S = (int)strtol(s, (char**)&s, 10);
        nd->k.sign = neg; nd->k.frac = S; nd->S = S; nd->P = 0;
        *out = id;
        return s;
    }
    if (*s != '(') return NULL;
    s++;
    char op[8] = {0};
    int k = 0;
    while (*s >= 'A' && *s <= 'Z' && k < 7) op[k++] = *s++;
    int id = pnew(p);
    if (!strcmp(op, "NEG")) {
        int c;
        s = parse_node(p, s, &c);
        if (!s) return NULL;
        p->n[id].kind = N_NEG; p->n[id].a = c; p->n[id].P = p->n[c].P; p->n[id].S = p->n[c].S;
    } else if (!strcmp(op, "CAST")) {
        int c;
        int P = (int)strtol(s, (char**)&s, 10);
        int S = (int)strtol(s, (char**)&s, 10);
        s = parse_node(p, s, &c);
        if (!s) return NULL;
        p->n[id].kind = N_CAST; p->n[id].a = c; p->n[id].P = P; p->n[id].S = S;
    } else if (!strcmp(op, "TRUNC")) {
        int c;
        int P = (int)strtol(s, (char**)&s, 10);
        int S = (int)strtol(s, (char**)&s, 10);
        int tp = (int)strtol(s, (char**)&s, 10);
        s = parse_node(p, s, &c);
        if (!s) return NULL;
        p->n[id].kind = N_TRUNC; p->n[id].a = c; p->n[id].P = P; p->n[id].S = S; p->n[id].tp = tp;
    } else {
        int kind = !strcmp(op, "ADD") ? N_ADD : !strcmp(op, "SUB") ? N_SUB : !strcmp(op, "MUL") ? N_MUL
                 : !strcmp(op, "DIV") ? N_DIV : !strcmp(op, "MOD") ? N_MOD : -1;
        if (kind < 0) return NULL;
        int P = (int)strtol(s, (char**)&s, 10);
        int S = (int)strtol(s, (char**)&s, 10);
        int a, b;
        s = parse_node(p, s, &a); if (!s) return NULL;
        s = parse_node(p, s, &b); if (!s) return NULL;
        p->n[id].kind = kind; p->n[id].P = P; p->n[id].S = S; p->n[id].a = a; p->n[id].b = b;
    }
    s = skip_ws(s);
    if (*s != ')') return NULL;
    *out = id;
    return s + 1;
}

od_plan* od_plan_parse(const char* text)
{
    od_plan* p = (od_plan*)calloc(1, sizeof(od_plan));
    int root;
    const char* e = parse_node(p, text, &root);
    if (!e) { od_plan_free(p); return NULL; }
    p->root = root;
    return p;
}

void od_plan_free(od_plan* p) { if (p) { free(p->n); free(p); } }

int od_plan_result(const od_plan* p, int* P, int* S) { *P = p->n[p->root].P; *S = p->n[p->root].S; return 0; }

typedef struct evctx {
    const od_plan* p;
    const uint8_t* const* cols;
    const int* colP;
    const int* colS;
    int mode;
} evctx;

static void eval_node(const evctx* c, int id, int64_t row, od_dec* out)
{
    const pnode* nd = &c->p->n[id];
    switch (nd->kind) {
    case N_COL: {
        int P = c->colP[nd->col], S = c->colS[nd->col];
        int len = od_compact_bytes(P);
        od_from_compact(out, c->cols[nd->col] + (size_t)row * (size_t)len, P, S);
        return;
    }
    case N_CONST: *out = nd->k; return;
    case N_NEG: eval_node(c, nd->a, row, out); od_neg(out); return;
    case N_CAST: eval_node(c, nd->a, row, out); od_cast(out, nd->P, nd->S, c->mode); return;
    case N_TRUNC:
        /* the argument's prec and frac are its plan type */
        eval_node(c, nd->a, row, out);
        out->prec = c->p->n[nd->a].P;
        out->frac = c->p->n[nd->a].S;
        od_truncate(out, nd->tp, c->mode);
        return;
    default: break;
    }
    od_dec rb;
    eval_node(c, nd->a, row, out);
    eval_node(c, nd->b, row, &rb);
    switch (nd->kind) {
    case N_ADD: od_add(out, &rb, c->mode); break;
    case N_SUB: od_sub(out, &rb, c->mode); break;
    case N_MUL: od_mul(out, &rb, c->mode); break;
    case N_DIV: od_div(out, &rb, c->mode); break;
    case N_MOD: od_mod(out, &rb, c->mode); break;
    }
}

typedef struct job {
    evctx c;
    int64_t lo, hi;
    uint8_t* out; int outlen;
    const uint8_t* mask;
    const int32_t* gid; int ngroups;
    od_dec* acc; int64_t* counts;
    int64_t n;
} job;

static void* run_eval(void* arg)
{
    job* j = (job*)arg;
    od_dec r;
    for (int64_t i = j->lo; i < j->hi; i++) {
        eval_node(&j->c, j->c.p->root, i, &r);
        od_to_compact(&r, j->out + (size_t)i * (size_t)j->outlen, j->outlen);
    }
    return NULL;
}

/* SUM sees the materialized result column: the value is stored as compact
 * (P_root, S_root) bytes and decoded again before it is added
 * (AriesExprCalcNode.cpp:1623-1700 materializes the child expression). */
static void eval_stored(const evctx* c, int64_t row, od_dec* r)
{
    const pnode* root = &c->p->n[c->p->root];
    eval_node(c, c->p->root, row, r);
    int len = od_compact_bytes(root->P);
    /* wider than the value array: every magnitude bit survives, the round
     * trip only keeps the sign */
    if (len > OD_MAXW * 4) return;
    uint8_t buf[OD_MAXW * 4];
    od_to_compact(r, buf, len);
    od_from_compact(r, buf, root->P, root->S);
}

static void* run_sum(void* arg)
{
    job* j = (job*)arg;
    od_dec r;
    j->n = 0;
    for (int64_t i = j->lo; i < j->hi; i++) {
        if (j->mask && !j->mask[i]) continue;
        eval_stored(&j->c, i, &r);
        od_add(j->acc, &r, j->c.mode);
        j->n++;
    }
    return NULL;
}

static void* run_gsum(void* arg)
{
    job* j = (job*)arg;
    od_dec r;
    for (int64_t i = j->lo; i < j->hi; i++) {
        int g = j->gid[i];
        if (g < 0) continue;
        eval_stored(&j->c, i, &r);
        od_add(&j->acc[g], &r, j->c.mode);
        j->counts[g]++;
    }
    return NULL;
}

static void run_threads(job* jobs, int nt, void* (*fn)(void*))
{
    if (nt == 1) { fn(&jobs[0]); return; }
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nt);
    for (int t = 0; t < nt; t++) pthread_create(&th[t], NULL, fn, &jobs[t]);
    for (int t = 0; t < nt; t++) pthread_join(th[t], NULL);
    free(th);
}

static int clamp_threads(int threads, int64_t rows)
{
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if (rows < threads) threads = rows > 0 ? (int)rows : 1;
    return threads;
}

int od_plan_eval(const od_plan* p, const uint8_t* const* cols, const int* colP, const int* colS,
                 int64_t rows, uint8_t* out, int outlen, int mode, int threads)
{
    int nt = clamp_threads(threads, rows);
    job* jobs = (job*)calloc((size_t)nt, sizeof(job));
    for (int t = 0; t < nt; t++) {
        jobs[t].c = (evctx){p, cols, colP, colS, mode};
        jobs[t].lo = rows * t / nt; jobs[t].hi = rows * (t + 1) / nt;
        jobs[t].out = out; jobs[t].outlen = outlen;
    }
    run_threads(jobs, nt, run_eval);
    free(jobs);
    return 0;
}

/* SUM seed: nullable_type<Decimal>() = +0; each row added in order with
 * AddDecimalOnCpu; thread partials are combined in thread order, which is
 * the same sum since the exact sum is order-independent and a +0 seed never
 * produces -0. */
int64_t od_plan_sum(const od_plan* p, const uint8_t* const* cols, const int* colP, const int* colS,
                    int64_t rows, const uint8_t* mask, uint8_t* out133, int mode, int threads)
{
    int nt = clamp_threads(threads, rows);
    job* jobs = (job*)calloc((size_t)nt, sizeof(job));
    od_dec* acc = (od_dec*)calloc((size_t)nt, sizeof(od_dec));
    for (int t = 0; t < nt; t++) {
        jobs[t].c = (evctx){p, cols, colP, colS, mode};
        jobs[t].lo = rows * t / nt; jobs[t].hi = rows * (t + 1) / nt;
        jobs[t].mask = mask; jobs[t].acc = &acc[t];
    }
    run_threads(jobs, nt, run_sum);
    od_dec total;
    od_zero(&total);
    int64_t n = 0;
    for (int t = 0; t < nt; t++) { od_add(&total, &acc[t], mode); n += jobs[t].n; }
    int P, S;
    od_plan_result(p, &P, &S);
    total.frac = S; total.prec = P;
    od_to_nullable_image(&total, out133);
    free(acc); free(jobs);
    return n;
}

int od_plan_group_sum(const od_plan* p, const uint8_t* const* cols, const int* colP, const int* colS,
                      int64_t rows, const int32_t* gid, int ngroups, uint8_t* out133, int64_t* counts,
                      int mode, int threads)
{
    int nt = clamp_threads(threads, rows);
    job* jobs = (job*)calloc((size_t)nt, sizeof(job));
    od_dec* acc = (od_dec*)calloc((size_t)nt * (size_t)ngroups, sizeof(od_dec));
    int64_t* cnt = (int64_t*)calloc((size_t)nt * (size_t)ngroups, sizeof(int64_t));
    for (int t = 0; t < nt; t++) {
        jobs[t].c = (evctx){p, cols, colP, colS, mode};
        jobs[t].lo = rows * t / nt; jobs[t].hi = rows * (t + 1) / nt;
        jobs[t].gid = gid; jobs[t].ngroups = ngroups;
        jobs[t].acc = &acc[(size_t)t * (size_t)ngroups]; jobs[t].counts = &cnt[(size_t)t * (size_t)ngroups];
    }
    run_threads(jobs, nt, run_gsum);
    int P, S;
    od_plan_result(p, &P, &S);
    for (int g = 0; g < ngroups; g++) {
        od_dec total;
        od_zero(&total);
        int64_t n = 0;
        for (int t = 0; t < nt; t++) { od_add(&total, &acc[(size_t)t * (size_t)ngroups + (size_t)g], mode); n += cnt[(size_t)t * (size_t)ngroups + (size_t)g]; }
        total.frac = S; total.prec = P;
        od_to_nullable_image(&total, out133 + (size_t)g * 133u);
        if (counts) counts[g] = n;
    }
    free(acc); free(cnt); free(jobs);
    return 0;
}

typedef struct segjob {
    const uint8_t* col; int P, S, len, mode;
    const int32_t* assoc; const int32_t* groups; int ngroups; int64_t rows;
    int g0, g1; uint8_t* out;
} segjob;

static void* run_seg(void* arg)
{
    segjob* j = (segjob*)arg;
    for (int g = j->g0; g < j->g1; g++) {
        int64_t lo = j->groups[g];
        int64_t hi = g + 1 < j->ngroups ? j->groups[g + 1] : j->rows;
        od_dec acc, r;
        od_zero(&acc);
        for (int64_t i = lo; i < hi; i++) {
            int64_t row = j->assoc ? j->assoc[i] : i;
            od_from_compact(&r, j->col + (size_t)row * (size_t)j->len, j->P, j->S);
            od_add(&acc, &r, j->mode);
        }
        acc.frac = j->S; acc.prec = j->P;
        od_to_nullable_image(&acc, j->out + (size_t)g * 133u);
    }
    return NULL;
}

int od_segsum_compact(const uint8_t* col, int P, int S, int64_t rows, const int32_t* associated,
                      const int32_t* groups, int ngroups, uint8_t* out133, int mode, int threads)
{
    int nt = clamp_threads(threads, ngroups);
    segjob* jobs = (segjob*)calloc((size_t)nt, sizeof(segjob));
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nt);
    for (int t = 0; t < nt; t++) {
        jobs[t] = (segjob){col, P, S, od_compact_bytes(P), mode, associated, groups, ngroups, rows,
                           (int)((int64_t)ngroups * t / nt), (int)((int64_t)ngroups * (t + 1) / nt), out133};
    }
    if (nt == 1) run_seg(&jobs[0]);
    else {
        for (int t = 0; t < nt; t++) pthread_create(&th[t], NULL, run_seg, &jobs[t]);
        for (int t = 0; t < nt; t++) pthread_join(th[t], NULL);
    }
    free(th); free(jobs);
    return 0;
}

/* AVG: Decimal(sum) /= Decimal((int64_t)count) via DivOrMod (kernel_util.hxx:427-438) */
void od_avg_from_sum(const uint8_t* sum133, int64_t count, uint8_t* out133, int mode)
{
    od_dec s, c;
    from_nullable_image(&s, sum133);
    od_from_int64(&c, count);
    od_div(&s, &c, mode);
    od_to_nullable_image(&s, out133);
}
