// planner.cpp -- expression parsing, the reference's AlignOptimize rewrite,
// result typing, limb-width analysis and HIP source generation.
#include "planner.h"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <queue>
#include <sstream>

namespace updec {

// ------------------------------------------------------------------ format constants

// __PREC_BYTE_ARRAY (decimal.cu:92): compact bytes per precision, a format
// constant of the column layout, valid for P <= 299.
static const uint8_t kPrecBytes[300] = {
    0,   1,   1,   2,   2,   3,   3,   4,   4,   4,   5,   5,   6,   6,   6,   7,   7,   8,   8,   9,   9,   9,   10,  10,  11,
    11,  11,  12,  12,  13,  13,  14,  14,  14,  15,  15,  16,  16,  16,  17,  17,  18,  18,  19,  19,  19,  20,  20,  21,  21,
    21,  22,  22,  23,  23,  24,  24,  24,  25,  25,  26,  26,  26,  27,  27,  28,  28,  29,  29,  29,  30,  30,  31,  31,  31,
    32,  32,  33,  33,  34,  34,  34,  35,  35,  36,  36,  36,  37,  37,  38,  38,  38,  39,  39,  40,  40,  41,  41,  41,  42,
    42,  43,  43,  43,  44,  44,  45,  45,  46,  46,  46,  47,  47,  48,  48,  48,  49,  49,  50,  50,  51,  51,  51,  52,  52,
    53,  53,  53,  54,  54,  55,  55,  56,  56,  56,  57,  57,  58,  58,  58,  59,  59,  60,  60,  61,  61,  61,  62,  62,  63,
    63,  63,  64,  64,  65,  65,  66,  66,  66,  67,  67,  68,  68,  68,  69,  69,  70,  70,  71,  71,  71,  72,  72,  73,  73,
    73,  74,  74,  75,  75,  76,  76,  76,  77,  77,  78,  78,  78,  79,  79,  80,  80,  81,  81,  81,  82,  82,  83,  83,  83,
    84,  84,  85,  85,  86,  86,  86,  87,  87,  88,  88,  88,  89,  89,  90,  90,  91,  91,  91,  92,  92,  93,  93,  93,  94,
    94,  95,  95,  96,  96,  96,  97,  97,  98,  98,  98,  99,  99,  100, 100, 101, 101, 101, 102, 102, 103, 103, 103, 104, 104,
    105, 105, 105, 106, 106, 107, 107, 108, 108, 108, 109, 109, 110, 110, 110, 111, 111, 112, 112, 113, 113, 113, 114, 114, 115,
    115, 115, 116, 116, 117, 117, 118, 118, 118, 119, 119, 120, 120, 120, 121, 121, 122, 122, 123, 123, 123, 124, 124, 125, 125};

// bitlen(10^n), from a table grown on demand (exact limb arithmetic)
static int pow10_bits(int n)
{
    if (n <= 0) return n == 0 ? 1 : 0;
    static std::mutex mu;
    static std::vector<int> table{1};
    static std::vector<uint32_t> v{1u};   // 10^(table.size() - 1)
    std::lock_guard<std::mutex> g(mu);
    while ((int)table.size() <= n) {
        uint64_t c = 0;
        for (auto& w : v) { uint64_t t = (uint64_t)w * 10u + c; w = (uint32_t)t; c = t >> 32; }
        if (c) v.push_back((uint32_t)c);
        table.push_back(32 * ((int)v.size() - 1) + (32 - __builtin_clz(v.back())));
    }
    return table[n];
}

int compact_bytes(int P)
{
    if (P < 0) return 0;
    if (P < 300) return kPrecBytes[P];
    // ceil((bitlen(10^P - 1) + 1) / 8); 10^P is never a power of two, so
    // bitlen(10^P - 1) == bitlen(10^P)
    return (pow10_bits(P) + 1 + 7) / 8;
}

// ------------------------------------------------------------------ HostDec

static const uint32_t kPow9[10] = {1u, 10u, 100u, 1000u, 10000u, 100000u, 1000000u, 10000000u, 100000000u, 1000000000u};

static void hd_reduce(HostDec& d, Mode m)
{
    if (m == Mode::Faithful)
        for (int i = 32; i < HostDec::NW; i++) d.v[i] = 0;
}

static uint32_t hd_mul_small(uint32_t* v, uint32_t k, uint32_t add)
{
    uint64_t c = add;
    for (int i = 0; i < HostDec::NW; i++) { uint64_t t = (uint64_t)v[i] * k + c; v[i] = (uint32_t)t; c = t >> 32; }
    return (uint32_t)c;
}

static void hd_mul10n(HostDec& d, int n, Mode m)
{
    while (n > 0) { int k = n > 9 ? 9 : n; hd_mul_small(d.v, kPow9[k], 0); n -= k; }
    hd_reduce(d, m);
}

HostDec HostDec::from_string(const std::string& s, Mode m)
{
    HostDec d;
    size_t i = 0;
    int dot = 0, digits = 0;
    bool dotflag = false;
    if (!s.empty() && s[0] == '-') { d.sign = 1; i = 1; }
    for (; i < s.size(); i++) {
        if (s[i] == '.') { dotflag = true; continue; }
        if (dotflag) dot++;
        digits++;
        hd_mul_small(d.v, 10u, (uint32_t)(s[i] - '0'));
    }
    d.prec = digits;
    d.frac = dot;
    hd_reduce(d, m);
    return d;
}

HostDec HostDec::from_int(int64_t t, int prec)
{
    HostDec d;
    d.prec = prec;
    uint64_t u = t < 0 ? (uint64_t)(-(t + 1)) + 1u : (uint64_t)t;
    if (t < 0) d.sign = 1;
    d.v[0] = (uint32_t)u;
    d.v[1] = (uint32_t)(u >> 32);
    return d;
}

HostDec HostDec::from_int_scale(int32_t t, int32_t scale, Mode m)
{
    HostDec d;
    int prec = 10;
    if (scale > prec) prec += scale;
    d.prec = prec;
    d.frac = scale;
    int64_t tt = t;
    if (tt < 0) { d.sign = 1; tt = -tt; }
    d.v[0] = (uint32_t)tt;
    hd_mul10n(d, scale, m);
    return d;
}

HostDec HostDec::rescaled_up(int scale, Mode m) const
{
    HostDec d = *this;
    int gap = scale - frac;
    if (gap > 0) hd_mul10n(d, gap, m);
    d.frac = scale;
    d.prec = prec + gap;
    return d;
}

static int hd_cmp(const uint32_t* a, const uint32_t* b)
{
    for (int i = HostDec::NW - 1; i >= 0; i--)
        if (a[i] != b[i]) return a[i] > b[i] ? 1 : -1;
    return 0;
}

void HostDec::add(const HostDec& d, Mode m)
{
    HostDec t = d;
    if (t.frac < frac) { hd_mul10n(t, frac - t.frac, m); t.frac = frac; }
    else if (t.frac > frac) { hd_mul10n(*this, t.frac - frac, m); frac = t.frac; }
    if (sign == t.sign) {
        uint64_t c = 0;
        for (int i = 0; i < NW; i++) { uint64_t s = (uint64_t)v[i] + t.v[i] + c; v[i] = (uint32_t)s; c = s >> 32; }
    } else {
        int r = hd_cmp(v, t.v);
        const uint32_t* x = r >= 0 ? v : t.v;
        const uint32_t* y = r >= 0 ? t.v : v;
        uint32_t out[NW];
        uint64_t br = 0;
        for (int i = 0; i < NW; i++) { uint64_t s = (uint64_t)x[i] - y[i] - br; out[i] = (uint32_t)s; br = (s >> 63) & 1; }
        memcpy(v, out, sizeof out);
        sign = (r > 0 && !d.sign) || (r < 0 && d.sign);
    }
    hd_reduce(*this, m);
}

void HostDec::mul(const HostDec& d, Mode m)
{
    sign = d.sign ^ sign;
    frac = frac + d.frac;
    uint32_t r[NW] = {0};
    for (int i = 0; i < NW; i++) {
        if (!v[i]) continue;
        uint64_t c = 0;
        for (int j = 0; i + j < NW; j++) {
            uint64_t t = (uint64_t)v[i] * d.v[j] + r[i + j] + c;
            r[i + j] = (uint32_t)t;
            c = t >> 32;
        }
    }
    memcpy(v, r, sizeof r);
    hd_reduce(*this, m);
}

bool HostDec::is_zero() const
{
    for (int i = 0; i < NW; i++)
        if (v[i]) return false;
    return true;
}

// Decimal == Decimal(1): sign 0 and |M| == 10^frac after alignment
bool HostDec::equals_one() const
{
    if (sign) return false;
    HostDec one;
    one.v[0] = 1;
    hd_mul10n(one, frac, Mode::Extended);
    return hd_cmp(v, one.v) == 0;
}

int HostDec::bitlen() const
{
    for (int i = NW - 1; i >= 0; i--)
        if (v[i]) return 32 * i + 32 - __builtin_clz(v[i]);
    return 0;
}

// CheckAndSetRealPrecision (decimal.cu:2060-2088), double arithmetic included
void HostDec::check_and_set_real_precision()
{
    int i = 31;
    for (; i >= 0; i--)
        if (v[i] != 0) break;
    if (i == -1) { prec = 1; frac = 0; return; }
    prec = i * 9;
    double maxUint32 = 4.294967295;
    double tmpMul = 1;
    for (int j = 0; j < i; j++) tmpMul *= maxUint32;
    tmpMul *= v[i];
    int tt = (int)tmpMul;
    while (tt > 0) { tt /= 10; prec++; }
}

std::string HostDec::hex() const
{
    int t = NW - 1;
    while (t > 0 && !v[t]) t--;
    char buf[16];
    std::string s;
    snprintf(buf, sizeof buf, "%x", v[t]);
    s += buf;
    for (int i = t - 1; i >= 0; i--) { snprintf(buf, sizeof buf, "%08x", v[i]); s += buf; }
    return s;
}

// ------------------------------------------------------------------ parser

namespace {
struct Lexer {
    const std::string& s;
    size_t i = 0;
    explicit Lexer(const std::string& x) : s(x) {}
    void ws() { while (i < s.size() && isspace((unsigned char)s[i])) i++; }
    bool eat(char c) { ws(); if (i < s.size() && s[i] == c) { i++; return true; } return false; }
    char peek() { ws(); return i < s.size() ? s[i] : 0; }
};

struct Parser {
    Lexer lx;
    const std::vector<std::string>& names;
    const std::vector<ColumnType>& types;
    Mode mode;
    std::string err;
    Parser(const std::string& t, const std::vector<std::string>& n, const std::vector<ColumnType>& ty, Mode m)
        : lx(t), names(n), types(ty), mode(m) {}

    ExprP calc(Op op, ExprP a, ExprP b)
    {
        auto e = std::make_shared<Expr>();
        e->kind = NK::Calc;
        e->op = op;
        e->ch = {a, b};
        return e;
    }
    ExprP expr()
    {
        ExprP l = term();
        if (!l) return nullptr;
        for (;;) {
            char c = lx.peek();
            if (c == '+' || c == '-') {
                lx.i++;
                ExprP r = term();
                if (!r) return nullptr;
                l = calc(c == '+' ? Op::Add : Op::Sub, l, r);
            } else return l;
        }
    }
    ExprP term()
    {
        ExprP l = unary();
        if (!l) return nullptr;
        for (;;) {
            char c = lx.peek();
            if (c == '*' || c == '/' || c == '%') {
                lx.i++;
                ExprP r = unary();
                if (!r) return nullptr;
                l = calc(c == '*' ? Op::Mul : c == '/' ? Op::Div : Op::Mod, l, r);
            } else return l;
        }
    }
    ExprP unary()
    {
        if (lx.eat('-')) {
            ExprP e = unary();
            if (!e) return nullptr;
            if (e->kind == NK::Int) e->ival = -e->ival;
            else if (e->kind == NK::Dec) e->dval.sign ^= 1u;
            else e->reverse = !e->reverse;
            return e;
        }
        return primary();
    }
    ExprP primary()
    {
        char c = lx.peek();
        if (c == '(') {
            lx.i++;
            ExprP e = expr();
            if (!e) return nullptr;
            if (!lx.eat(')')) { err = "expected ')'"; return nullptr; }
            return e;
        }
        if (isdigit((unsigned char)c) || c == '.') {
            size_t st = lx.i;
            bool dot = false;
            while (lx.i < lx.s.size() && (isdigit((unsigned char)lx.s[lx.i]) || lx.s[lx.i] == '.')) {
                if (lx.s[lx.i] == '.') dot = true;
                lx.i++;
            }
            std::string lit = lx.s.substr(st, lx.i - st);
            auto e = std::make_shared<Expr>();
            if (dot) {
                e->kind = NK::Dec;
                e->dval = HostDec::from_string(lit, mode);
            } else {
                // an integer literal that int64 cannot hold is a DECIMAL
                // literal of its digits (MySQL's rule for integer literals
                // past BIGINT)
                size_t nz = lit.find_first_not_of('0');
                const std::string digits = nz == std::string::npos ? "0" : lit.substr(nz);
                if (digits.size() < 19 || (digits.size() == 19 && digits <= "9223372036854775807")) {
                    e->kind = NK::Int;
                    e->ival = 0;
                    for (char d : digits) e->ival = e->ival * 10 + (d - '0');
                } else {
                    e->kind = NK::Dec;
                    e->dval = HostDec::from_string(digits, mode);
                }
            }
            return e;
        }
        if (isalpha((unsigned char)c) || c == '_') {
            size_t st = lx.i;
            while (lx.i < lx.s.size() && (isalnum((unsigned char)lx.s[lx.i]) || lx.s[lx.i] == '_')) lx.i++;
            std::string id = lx.s.substr(st, lx.i - st);
            if (id == "cast" && lx.peek() == '(') {
                lx.i++;
                ExprP arg = expr();
                if (!arg) return nullptr;
                long P = 0, S = 0;
                auto num = [&](long* v) -> bool {
                    if (!lx.eat(',')) return false;
                    lx.ws();
                    size_t b = lx.i;
                    while (lx.i < lx.s.size() && isdigit((unsigned char)lx.s[lx.i])) lx.i++;
                    if (b == lx.i) return false;
                    if (lx.i - b > 6) return false;   // P and S have at most a few digits
                    *v = std::stol(lx.s.substr(b, lx.i - b));
                    return true;
                };
                if (!num(&P) || !num(&S) || !lx.eat(')') || S > P) {
                    err = "cast expects cast(expr, P, S) with S <= P";
                    return nullptr;
                }
                auto e = std::make_shared<Expr>();
                e->kind = NK::Cast;
                e->P = (int)P;
                e->S = (int)S;
                e->ch = {arg};
                return e;
            }
            if (id == "truncate" && lx.peek() == '(') {
                lx.i++;
                ExprP arg = expr();
                if (!arg) return nullptr;
                bool ok = lx.eat(',');
                bool neg = ok && lx.eat('-');
                lx.ws();
                size_t b = lx.i;
                while (lx.i < lx.s.size() && isdigit((unsigned char)lx.s[lx.i])) lx.i++;
                const size_t nd = lx.i - b;
                ok = ok && nd > 0 && nd <= 6 && lx.eat(')');
                if (!ok) {
                    err = "truncate expects truncate(expr, p) with an integer p";
                    return nullptr;
                }
                long p = std::stol(lx.s.substr(b, nd));
                auto e = std::make_shared<Expr>();
                e->kind = NK::Trunc;
                e->ival = neg ? -p : p;
                e->ch = {arg};
                return e;
            }
            for (size_t k = 0; k < names.size(); k++) {
                if (names[k] == id) {
                    auto e = std::make_shared<Expr>();
                    e->kind = NK::Column;
                    e->col = (int)k;
                    e->P = types[k].P;
                    e->S = types[k].S;
                    return e;
                }
            }
            err = "unknown column '" + id + "'";
            return nullptr;
        }
        err = std::string("unexpected character '") + (c ? c : '?') + "'";
        return nullptr;
    }
};
}  // namespace

ExprP parse_expr(const std::string& text, const std::vector<std::string>& names, const std::vector<ColumnType>& types,
                 Mode mode, std::string* err)
{
    Parser p(text, names, types, mode);
    ExprP e = p.expr();
    if (e && p.lx.peek() != 0) { p.err = "trailing input"; e = nullptr; }
    if (!e && err) *err = p.err.empty() ? "parse error" : p.err;
    if (e) infer_frontend_types(*e, mode);
    return e;
}

void trunc_type(int P, int S, int p, int* outP, int* outS)
{
    if (p > 0) {
        *outP = P - S + p;
        *outS = p;
    } else {
        *outP = P - S > -p ? P - S : 1;
        *outS = 0;
    }
}

static int digits_of(int64_t v)
{
    HostDec d = HostDec::from_int(v, 10);
    d.check_and_set_real_precision();
    return d.prec;
}

void infer_frontend_types(Expr& e, Mode mode)
{
    switch (e.kind) {
    case NK::Column: return;
    case NK::Int: e.P = digits_of(e.ival); e.S = 0; return;
    case NK::Dec:
        // the literal's own type is StringToDecimal's (digit count, leading
        // zeros included, and the digits after the point), as the engine
        // bridge gives it (AriesExprBridge.cpp:381-399)
        e.P = e.dval.prec;
        e.S = e.dval.frac;
        return;
    case NK::Cast: infer_frontend_types(*e.ch[0], mode); return;
    case NK::Trunc:
        infer_frontend_types(*e.ch[0], mode);
        trunc_type(e.ch[0]->P, e.ch[0]->S, (int)e.ival, &e.P, &e.S);
        return;
    case NK::Calc: break;
    }
    for (auto& c : e.ch) infer_frontend_types(*c, mode);
    // the frontend types a CALC from GetDecimalValue of its operands
    // (CommonBiaodashi.cpp:2990-3030): a decimal literal there has its
    // CheckAndSetRealPrecision precision (and scale 0 when it is zero)
    auto fe = [](const Expr& c, int* P, int* S) {
        if (c.kind == NK::Dec) {
            HostDec t = c.dval;
            t.check_and_set_real_precision();
            *P = t.prec;
            *S = t.frac;
        } else {
            *P = c.P;
            *S = c.S;
        }
    };
    int Pa, Sa, Pb, Sb;
    fe(*e.ch[0], &Pa, &Sa);
    fe(*e.ch[1], &Pb, &Sb);
    int P = 0, S = 0;
    switch (e.op) {
    case Op::Add:
    case Op::Sub: P = std::max(Pa - Sa, Pb - Sb) + 1; S = std::max(Sa, Sb); P += S; break;
    case Op::Mul: S = Sa + Sb; P = Pa + Pb; break;
    case Op::Div: P = Pa - Pb + Sb + 4 + 1; S = Sa + 4; break;
    case Op::Mod: P = Pb; S = 0; break;
    }
    if (mode == Mode::Faithful) { P &= 0xff; S &= 0xff; }  // Decimal::prec/frac are uint8
    else P = std::max(P, 1);
    e.P = P;
    e.S = S;
}

// ------------------------------------------------------------------ AlignOptimize

namespace {

struct BT;
struct MT;
struct Arena {
    std::vector<std::unique_ptr<BT>> bts;
    std::vector<std::unique_ptr<MT>> mts;
    BT* bt();
    MT* mt();
};

struct BT {
    NK kind = NK::Column;
    Op op = Op::Add;
    int col = -1;
    int64_t ival = 0;
    HostDec dval;
    int P = 0, S = 0;
    bool reverse = false;
    ExprP fn;   // Cast: the cast node, its argument already rewritten
    std::vector<BT*> ch;
};

struct MT {
    NK kind = NK::Column;
    Op op = Op::Add;
    int col = -1;
    int64_t ival = 0;
    HostDec dval;
    int P = 0, S = 0;
    bool reverse = false;
    ExprP fn;
    BT* root = nullptr;
    std::vector<MT*> ch;
};

BT* Arena::bt() { bts.emplace_back(new BT()); return bts.back().get(); }
MT* Arena::mt() { mts.emplace_back(new MT()); return mts.back().get(); }

// A cast is a leaf of the rewrite: its argument is rewritten as a tree of
// its own first (AlignOptimize rewrites every CALC subtree found below a
// non-CALC node separately, AriesEngineShell.cpp:646-685).
BT* clone_bt(Arena& A, const Expr& e, Mode mode, bool* ok)
{
    BT* b = A.bt();
    b->kind = e.kind; b->op = e.op; b->col = e.col; b->ival = e.ival; b->dval = e.dval;
    b->P = e.P; b->S = e.S; b->reverse = e.reverse;
    if (e.kind == NK::Cast || e.kind == NK::Trunc) {
        b->fn = align_optimize(std::make_shared<Expr>(e), mode);
        if (!b->fn) *ok = false;
        else { b->P = b->fn->P; b->S = b->fn->S; }   // truncate's type follows its rewritten argument
        return b;
    }
    for (auto& c : e.ch) b->ch.push_back(clone_bt(A, *c, mode, ok));
    return b;
}

MT* mt_from_bt(Arena& A, BT* b)
{
    MT* m = A.mt();
    m->root = b; m->kind = b->kind; m->op = b->op; m->col = b->col; m->ival = b->ival; m->dval = b->dval;
    m->P = b->P; m->S = b->S; m->reverse = b->reverse; m->fn = b->fn;
    return m;
}

// ModifySubOperation (AriesEngineShell.cpp:244-281)
void modify_sub(BT* t)
{
    std::queue<BT*> q;
    q.push(t);
    while (!q.empty()) {
        BT* n = q.front();
        if (n->kind == NK::Calc) {
            if (n->op == Op::Mul && n->reverse) n->ch[0]->reverse = !n->ch[0]->reverse;
            else if (n->op == Op::Div && n->reverse) n->ch[0]->reverse = !n->ch[0]->reverse;
            else if (n->op == Op::Add && n->reverse) {
                n->ch[0]->reverse = !n->ch[0]->reverse;
                n->ch[1]->reverse = !n->ch[1]->reverse;
            } else if (n->op == Op::Sub) {
                if (!n->reverse) n->ch[1]->reverse = !n->ch[1]->reverse;
                if (n->reverse) n->ch[0]->reverse = !n->ch[0]->reverse;
                n->op = Op::Add;
            }
            if (n->ch.size() == 2) { q.push(n->ch[0]); q.push(n->ch[1]); }
        }
        q.pop();
    }
}

// getChild (AriesEngineShell.cpp:284-322)
std::vector<MT*> get_child(Arena& A, BT* b, Op sym)
{
    std::vector<MT*> out;
    if (sym == Op::Div) {
        out.push_back(mt_from_bt(A, b->ch[0]));
        out.push_back(mt_from_bt(A, b->ch[1]));
        return out;
    }
    for (int k = 0; k < 2; k++) {
        BT* c = b->ch[k];
        if (c->kind != NK::Calc) out.push_back(mt_from_bt(A, c));
        else if (c->op == sym) {
            auto nx = get_child(A, c, sym);
            out.insert(out.end(), nx.begin(), nx.end());
        } else out.push_back(mt_from_bt(A, c));
    }
    return out;
}

// BinToMultiTree (AriesEngineShell.cpp:325-342)
void bin_to_multi(Arena& A, BT* b, MT* m)
{
    m->ch = get_child(A, b, m->root->op);
    for (MT* c : m->ch)
        if (c->root && c->root->kind == NK::Calc && c->root->ch.size() == 2) bin_to_multi(A, c->root, c);
}

// CmpByScale (AriesEngineShell.cpp:344-358)
bool cmp_by_scale(const MT* a, const MT* b)
{
    if (a->kind == NK::Int) return true;
    if (b->kind == NK::Int) return false;
    if (a->kind == NK::Dec && b->kind == NK::Dec) return a->S < b->S;
    if (a->kind == NK::Dec) return true;
    if (b->kind == NK::Dec) return false;
    return a->S < b->S;
}

// std::sort on <= 16 elements is an insertion sort (libstdc++
// __insertion_sort); reproduced so that equal-scale operands keep the order
// the reference gives them.
void insertion_sort(std::vector<MT*>& v)
{
    if (v.size() > 16) { std::stable_sort(v.begin(), v.end(), cmp_by_scale); return; }
    for (size_t i = 1; i < v.size(); i++) {
        MT* val = v[i];
        if (cmp_by_scale(val, v[0])) {
            for (size_t k = i; k > 0; k--) v[k] = v[k - 1];
            v[0] = val;
        } else {
            size_t k = i;
            while (cmp_by_scale(val, v[k - 1])) { v[k] = v[k - 1]; k--; }
            v[k] = val;
        }
    }
}

bool is_decimal_calc(const MT* m)
{
    // only CALC nodes of decimal result type are rewritten; a tree with a
    // division or a decimal operand anywhere below is decimal-typed
    std::function<bool(const BT*)> has_dec = [&](const BT* b) -> bool {
        if (b->kind == NK::Column || b->kind == NK::Dec || b->kind == NK::Cast || b->kind == NK::Trunc) return true;
        if (b->kind == NK::Calc && b->op == Op::Div) return true;
        for (auto* c : b->ch)
            if (has_dec(c)) return true;
        return false;
    };
    return m->kind == NK::Calc && m->root && has_dec(m->root);
}

// AlignOptMultiTree (AriesEngineShell.cpp:360-597)
bool align_opt(Arena& A, MT* top, Mode mode, std::string* err)
{
    std::queue<MT*> q;
    q.push(top);
    while (!q.empty()) {
        MT* node = q.front();
        if (node->kind == NK::Calc && is_decimal_calc(node)) {
            if (node->op == Op::Add) {
                insertion_sort(node->ch);
                int nodeInt = 0;
                HostDec nodeDec = HostDec::from_int(0, 10);
                size_t index = 0;
                for (index = 0; index < node->ch.size(); index++) {
                    MT* c = node->ch[index];
                    if (c->kind == NK::Int) {
                        int t = (int)c->ival;
                        if (c->reverse) t = -t;
                        nodeInt += t;
                    } else if (c->kind == NK::Dec) {
                        HostDec d = c->dval;
                        if (c->reverse) d.sign = 1;
                        nodeDec.add(d, mode);
                    } else break;
                }
                if (index == 0)
                    for (MT* c : node->ch)
                        if (c->kind == NK::Calc) q.push(c);
                std::vector<MT*> nl;
                if (nodeInt == 0 && nodeDec.is_zero()) {
                    if (index == node->ch.size()) { *err = "addition of constants only"; return false; }
                    if (index == node->ch.size() - 1) {
                        // CopyNodeInfo: the single remaining operand replaces the node
                        MT* c = node->ch[index];
                        node->kind = c->kind; node->op = c->op; node->col = c->col; node->ival = c->ival;
                        node->dval = c->dval; node->P = c->P; node->S = c->S; node->reverse = c->reverse;
                        node->fn = c->fn;
                        node->ch = c->ch;
                        node->root = c->root;
                    } else {
                        for (; index < node->ch.size(); index++) {
                            nl.push_back(node->ch[index]);
                            if (node->ch[index]->kind == NK::Calc) q.push(node->ch[index]);
                        }
                        node->ch = nl;
                    }
                } else {
                    if (nodeDec.is_zero()) {
                        if (index >= node->ch.size()) { *err = "addition of constants only"; return false; }
                        MT* nn = A.mt();
                        int nearScale = node->ch[index]->S;
                        nn->kind = NK::Dec;
                        nn->S = nearScale;
                        nn->P = 0;
                        nn->dval = HostDec::from_int_scale(nodeInt, nearScale, mode);
                        nl.push_back(nn);
                        for (; index < node->ch.size(); index++) {
                            nl.push_back(node->ch[index]);
                            if (node->ch[index]->kind == NK::Calc) q.push(node->ch[index]);
                        }
                    } else {
                        bool optInsert = true;
                        if (nodeInt != 0) nodeDec.add(HostDec::from_int(nodeInt, 10), mode);
                        for (; index < node->ch.size(); index++) {
                            if (node->ch[index]->S > nodeDec.frac && optInsert) {
                                optInsert = false;
                                HostDec d = nodeDec.rescaled_up(node->ch[index]->S, mode);
                                MT* nn = A.mt();
                                nn->kind = NK::Dec;
                                nn->S = d.frac;
                                nn->P = 0;
                                nn->dval = d;
                                nl.push_back(nn);
                            }
                            nl.push_back(node->ch[index]);
                            if (node->ch[index]->kind == NK::Calc) q.push(node->ch[index]);
                        }
                        if (optInsert) {
                            MT* nn = A.mt();
                            nn->kind = NK::Dec;
                            nn->P = nodeDec.frac;
                            nn->S = nodeDec.frac;
                            nn->dval = nodeDec;
                            nl.push_back(nn);
                        }
                    }
                    node->ch = nl;
                }
            } else if (node->op == Op::Mul) {
                int nodeInt = 1;
                HostDec nodeDec = HostDec::from_int(1, 10);
                std::vector<MT*> nl;
                for (MT* c : node->ch) {
                    if (c->kind == NK::Int) nodeInt *= (int)c->ival;
                    else if (c->kind == NK::Dec) nodeDec.mul(c->dval, mode);
                    else nl.push_back(c);
                }
                if (nodeInt != 1 || !nodeDec.equals_one()) {
                    if (nodeInt == 0 || nodeDec.is_zero()) {
                        node->kind = NK::Int;
                        node->ival = 0;
                    } else {
                        nodeDec.mul(HostDec::from_int(nodeInt, 10), mode);
                        if (!nodeDec.equals_one()) {
                            MT* nn = A.mt();
                            nn->kind = NK::Dec;
                            nn->S = nodeDec.frac;
                            nn->P = 0;
                            nn->dval = nodeDec;
                            nl.push_back(nn);
                        }
                        node->ch = nl;
                    }
                }
                if (node->kind == NK::Calc)
                    for (MT* c : node->ch)
                        if (c->kind == NK::Calc) q.push(c);
            } else if (node->op == Op::Div || node->op == Op::Mod) {
                if (node->ch[0]->kind == NK::Int && node->ch[0]->ival == 0) {
                    node->kind = NK::Int;
                    node->ival = 0;
                } else {
                    for (size_t i = 0; i < node->ch.size(); i++) {
                        MT* c = node->ch[i];
                        if (c->kind == NK::Calc) q.push(c);
                        if (c->kind == NK::Int) {
                            bool fits32 = c->ival >= INT32_MIN && c->ival <= INT32_MAX;
                            HostDec d = HostDec::from_int(c->ival, fits32 ? 10 : 19);
                            MT* nn = A.mt();
                            nn->kind = NK::Dec;
                            nn->P = d.prec;
                            nn->S = d.frac;
                            nn->dval = d;
                            node->ch[i] = nn;
                        }
                    }
                }
            } else {
                for (MT* c : node->ch)
                    if (c->kind == NK::Calc) q.push(c);
            }
        }
        q.pop();
    }
    return true;
}

// RemoveZeroNode (AriesEngineShell.cpp:184-219)
MT* remove_zero(MT* m)
{
    std::vector<MT*> nl;
    for (size_t i = 0; i < m->ch.size(); i++) {
        MT* c = m->ch[i];
        if (c->kind == NK::Int) {
            if (m->op == Op::Mul) { m->kind = NK::Int; m->ival = 0; return m; }
        } else if (c->kind == NK::Calc) {
            m->ch[i] = remove_zero(c);
            if (m->ch[i]->kind == NK::Int) {
                if (m->op == Op::Mul) { m->kind = NK::Int; m->ival = 0; return m; }
            } else nl.push_back(m->ch[i]);
        } else nl.push_back(c);
    }
    m->ch = nl;
    return m;
}

BT* leaf_from_mt(Arena& A, const MT* m)
{
    BT* b = A.bt();
    b->kind = m->kind; b->op = m->op; b->col = m->col; b->ival = m->ival; b->dval = m->dval;
    b->P = m->P; b->S = m->S; b->reverse = m->reverse; b->fn = m->fn;
    return b;
}

// ListToBinTree (AriesEngineShell.cpp:90-144): left-deep chain with the
// per-node result types.  Faithful mode keeps the reference's uint16 type
// arithmetic (a DIV whose rule gives a negative precision wraps to ~65500);
// extended mode computes the same rules in plain integers and keeps the
// precision at least 1, so the materialized width stays a real type.
static int ty(int v, Mode mode) { return mode == Mode::Faithful ? (int)(uint16_t)v : v; }

BT* list_to_bin(Arena& A, std::vector<BT*>& L, int index, Op op, Mode mode)
{
    if (index == 0) return L[0];
    BT* l = list_to_bin(A, L, index - 1, op, mode);
    BT* r = L[index];
    BT* t = A.bt();
    t->kind = NK::Calc;
    int P0 = l->P, S0 = l->S, P1 = r->P, S1 = r->S;
    if (op == Op::Add) {
        if (r->reverse) { r->reverse = false; t->op = Op::Sub; }
        else t->op = Op::Add;
        int P = ty(std::max(P0 - S0, P1 - S1) + 1, mode);
        int S = std::max(S0, S1);
        t->P = ty(P + S, mode);
        t->S = S;
    } else if (op == Op::Mul) {
        t->op = Op::Mul;
        int P = ty((P0 - S0) + (P1 - S1), mode);
        int S = ty(S0 + S1, mode);
        t->P = ty(P + S, mode);
        t->S = S;
    } else if (op == Op::Div) {
        t->op = Op::Div;
        t->P = ty(P0 - P1 + S1 + 4 + 1, mode);
        t->S = ty(S0 + 4, mode);
    } else {
        t->op = Op::Mod;
        t->P = P1;
        t->S = 0;
    }
    if (mode == Mode::Extended) t->P = std::max(t->P, 1);
    t->ch = {l, r};
    return t;
}

// MultiToBinTree (AriesEngineShell.cpp:146-182)
BT* multi_to_bin(Arena& A, MT* m, Mode mode)
{
    std::vector<BT*> tmp;
    for (MT* c : m->ch) {
        if (c->kind != NK::Calc) tmp.push_back(leaf_from_mt(A, c));
        else tmp.push_back(multi_to_bin(A, c, mode));
    }
    if (m->ch.empty()) return leaf_from_mt(A, m);
    if (m->kind == NK::Calc && m->ch.size() == 1) {
        if (m->ch[0]->kind == NK::Calc) return list_to_bin(A, tmp, (int)tmp.size() - 1, m->op, mode);
        return leaf_from_mt(A, m->ch[0]);
    }
    return list_to_bin(A, tmp, (int)tmp.size() - 1, m->op, mode);
}

ExprP to_expr(const BT* b)
{
    auto e = std::make_shared<Expr>();
    if (b->kind == NK::Cast || b->kind == NK::Trunc) {
        *e = *b->fn;
        e->reverse = b->reverse;
        return e;
    }
    e->kind = b->kind; e->op = b->op; e->col = b->col; e->ival = b->ival; e->dval = b->dval;
    e->P = b->P; e->S = b->S; e->reverse = b->reverse;
    for (auto* c : b->ch) e->ch.push_back(to_expr(c));
    return e;
}

}  // namespace

ExprP align_optimize(const ExprP& e, Mode mode)
{
    if (e && (e->kind == NK::Cast || e->kind == NK::Trunc)) {
        ExprP arg = align_optimize(e->ch[0], mode);
        if (!arg) return nullptr;
        auto c = std::make_shared<Expr>(*e);
        c->ch = {arg};
        if (e->kind == NK::Trunc) trunc_type(arg->P, arg->S, (int)e->ival, &c->P, &c->S);
        return c;
    }
    if (!e || e->kind != NK::Calc) return e;
    Arena A;
    bool ok = true;
    BT* bt = clone_bt(A, *e, mode, &ok);
    if (!ok) return nullptr;
    modify_sub(bt);
    MT* mt = mt_from_bt(A, bt);
    bin_to_multi(A, bt, mt);
    std::string err;
    if (!align_opt(A, mt, mode, &err)) return nullptr;
    mt = remove_zero(mt);
    if (mt->kind == NK::Int) return nullptr;
    BT* out = multi_to_bin(A, mt, mode);
    // AlignOptimizeNode (AriesEngineShell.cpp:628-644): root precision capped at 288
    if (mode == Mode::Faithful && out->P > 288) out->P = 288;
    return to_expr(out);
}

// ------------------------------------------------------------------ plan

static int words_for(int bits) { return std::max(1, (bits + 31) / 32); }

// Magnitude bound of an operation node from its operands' bounds ba, bb
// (value < 2^bits).  Leaves carry their own bound.
static int op_bits(const PNode& n, const PNode& A, const PNode* B, int ba, int bb)
{
    switch (n.kind) {
    case PK::Neg: return ba;
    case PK::Add:
    case PK::Sub: {
        int xa = ba + (n.S > A.S ? pow10_bits(n.S - A.S) : 0);
        int xb = bb + (n.S > B->S ? pow10_bits(n.S - B->S) : 0);
        return std::max(xa, xb) + 1;
    }
    case PK::Mul: return ba + bb;
    case PK::Div: return std::max(64, ba + pow10_bits(B->S + 4) + 1);
    // every path of the reference's % leaves a value below |b| when b != 0
    // (a remainder, or ||a| - |b|| with bitlen(a) <= bitlen(b)); x % 0 keeps |a|
    case PK::Mod:
        if (B->kind == PK::Const && B->k.bitlen() > 0) return std::max(1, bb);
        return std::max(1, std::max(ba, bb));
    case PK::Cast:
        if (n.S > A.S) return ba + pow10_bits(n.S - A.S);
        if (n.S < A.S) return std::max(1, ba - (pow10_bits(A.S - n.S) - 1));
        return ba;
    case PK::Trunc:
        if (n.tp > 0) {
            if (n.S > A.S) return ba + pow10_bits(n.S - A.S);
            if (n.S < A.S) return std::max(1, ba - (pow10_bits(A.S - n.S) - 1));
            return ba;
        }
        if (A.P - A.S <= -n.tp) return 1;
        return std::max(1, ba - (pow10_bits(A.S - n.tp) - 1)) + pow10_bits(-n.tp);
    default: return n.bits;
    }
}

bool build_plan(const ExprP& rw, const std::vector<ColumnType>& cols, Mode mode, Plan* out, std::string* err)
{
    Plan p;
    p.mode = mode;
    // identical subtrees (same operation, type and operand nodes) become one
    // node, so a repeated term such as x*x in a series is computed once
    std::map<std::string, int> seen;
    auto intern = [&](const PNode& n) -> int {
        std::string key = std::to_string((int)n.kind) + "," + std::to_string(n.P) + "," + std::to_string(n.S) + "," +
                          std::to_string(n.col) + "," + std::to_string(n.a) + "," + std::to_string(n.b) + "," +
                          std::to_string(n.tp);
        if (n.kind == PK::Const) key += std::string(",") + (n.k.sign ? "-" : "+") + n.k.hex() + ":" + std::to_string(n.k.frac);
        auto it = seen.find(key);
        if (it != seen.end()) return it->second;
        p.n.push_back(n);
        seen[key] = (int)p.n.size() - 1;
        return (int)p.n.size() - 1;
    };
    std::function<int(const Expr&)> rec = [&](const Expr& e) -> int {
        PNode n{};
        int id = -1;
        if (e.kind == NK::Column) {
            n.kind = PK::Col;
            n.col = e.col;
            n.P = cols[e.col].P;
            n.S = cols[e.col].S;
            n.bits = 8 * compact_bytes(n.P) - 1;
            if (cols[e.col].bits > 0 && cols[e.col].bits < n.bits) n.bits = cols[e.col].bits;
        } else if (e.kind == NK::Dec) {
            n.kind = PK::Const;
            n.k = e.dval;
            n.P = e.P;
            n.S = e.dval.frac;
            if (e.S != e.dval.frac) { *err = "decimal constant whose scale differs from its type"; return -1; }
            n.bits = std::max(1, e.dval.bitlen());
        } else if (e.kind == NK::Int) {
            *err = "integer operand left in a decimal calculation after rewrite";
            return -1;
        } else if (e.kind == NK::Cast) {
            int a = rec(*e.ch[0]);
            if (a < 0) return -1;
            n.kind = PK::Cast;
            n.a = a;
            n.P = e.P;
            n.S = e.S;
        } else if (e.kind == NK::Trunc) {
            int a = rec(*e.ch[0]);
            if (a < 0) return -1;
            n.kind = PK::Trunc;
            n.a = a;
            n.tp = (int)e.ival;
            trunc_type(p.n[a].P, p.n[a].S, n.tp, &n.P, &n.S);
            if (n.P != e.P || n.S != e.S) { *err = "internal: truncate type differs from its node"; return -1; }
        } else {
            // RemoveZeroNode drops the integer operands of an integer-typed
            // calculation below a decimal one, leaving a calculation node
            // without operands: the reference has no defined result for it
            if (e.ch.size() != 2) {
                *err = "integer-typed calculation inside a decimal expression";
                return -1;
            }
            int a = rec(*e.ch[0]);
            if (a < 0) return -1;
            int b = rec(*e.ch[1]);
            if (b < 0) return -1;
            const PNode& A = p.n[a];
            const PNode& B = p.n[b];
            n.a = a;
            n.b = b;
            n.P = e.P;
            switch (e.op) {
            case Op::Add:
            case Op::Sub:
                n.kind = e.op == Op::Add ? PK::Add : PK::Sub;
                n.S = std::max(A.S, B.S);
                break;
            case Op::Mul:
                n.kind = PK::Mul;
                n.S = A.S + B.S;
                break;
            case Op::Div:
                n.kind = PK::Div;
                n.S = A.S + 4;
                break;
            case Op::Mod:
                n.kind = PK::Mod;
                n.S = 0;
                break;
            }
            if (n.S != e.S) { *err = "internal: runtime scale differs from node type"; return -1; }
        }
        if (n.a >= 0) {
            const PNode& A = p.n[n.a];
            const PNode* B = n.b >= 0 ? &p.n[n.b] : nullptr;
            n.bits = op_bits(n, A, B, A.bits, B ? B->bits : 0);
            n.ubits = op_bits(n, A, B, A.ubits, B ? B->ubits : 0);
        } else {
            n.ubits = n.bits;
        }
        if (mode == Mode::Faithful) n.bits = std::min(n.bits, 1024);
        n.W = words_for(n.bits);
        id = intern(n);
        if (e.reverse) {
            PNode g{};
            g.kind = PK::Neg;
            g.a = id;
            g.P = n.P;
            g.S = n.S;
            g.bits = n.bits;
            g.ubits = n.ubits;
            g.W = n.W;
            id = intern(g);
        }
        return id;
    };
    if (!rw) { *err = "expression rewrites to a constant or integer"; return false; }
    int r = rec(*rw);
    if (r < 0) return false;
    p.root = r;
    for (PNode& x : p.n)
        if (x.kind == PK::Const) {
            x.koff = p.kwords;
            p.kwords += x.W + 1;
        }
    p.outP = rw->P;
    p.outS = rw->S;
    p.outB = compact_bytes(p.outP);
    *out = std::move(p);
    return true;
}

std::vector<uint32_t> Plan::const_data() const
{
    std::vector<uint32_t> d((size_t)kwords, 0u);
    for (const PNode& x : n)
        if (x.kind == PK::Const) {
            d[(size_t)x.koff] = x.k.sign ? 1u : 0u;
            for (int w = 0; w < x.W; w++) d[(size_t)x.koff + 1 + (size_t)w] = x.k.v[w];
        }
    return d;
}

std::string Plan::text(bool values) const
{
    std::function<std::string(int)> rec = [&](int id) -> std::string {
        const PNode& x = n[id];
        char buf[64];
        switch (x.kind) {
        case PK::Col: snprintf(buf, sizeof buf, "c%d:%d:%d", x.col, x.P, x.S); return buf;
        case PK::Const:
            if (!values) return "k@" + std::to_string(x.koff) + ":" + std::to_string(x.S);
            return std::string("k") + (x.k.sign ? "-" : "+") + x.k.hex() + ":" + std::to_string(x.S);
        case PK::Neg: return "(NEG " + rec(x.a) + ")";
        case PK::Cast: snprintf(buf, sizeof buf, "(CAST %d %d ", x.P, x.S); return buf + rec(x.a) + ")";
        case PK::Trunc: snprintf(buf, sizeof buf, "(TRUNC %d %d %d ", x.P, x.S, x.tp); return buf + rec(x.a) + ")";
        default: break;
        }
        const char* op = x.kind == PK::Add   ? "ADD"
                         : x.kind == PK::Sub ? "SUB"
                         : x.kind == PK::Mul ? "MUL"
                         : x.kind == PK::Mod ? "MOD"
                                             : "DIV";
        snprintf(buf, sizeof buf, "(%s %d %d ", op, x.P, x.S);
        return buf + rec(x.a) + " " + rec(x.b) + ")";
    };
    return rec(root);
}

bool plan_is_faithful_envelope(const Plan& p, const std::vector<ColumnType>& cols)
{
    for (auto& c : cols)
        if (c.P > 299 || c.P < 0 || c.S > 255) return false;
    // the materialized result needs a row of the 300-entry byte table
    if (p.outP > 299) return false;
    for (const PNode& x : p.n) {
        if (x.S > 255) return false;
        if (x.kind == PK::Add || x.kind == PK::Sub) {
            if (x.S - p.n[x.a].S > 38 || x.S - p.n[x.b].S > 38) return false;
        }
        // operator<< reads one row of the 38-row 10^n table (decimal.cu:2028-2058)
        if (x.kind == PK::Cast && x.S - p.n[x.a].S > 38) return false;
        if (x.kind == PK::Trunc && (x.tp > 0 ? x.tp - p.n[x.a].S : -x.tp) > 38) return false;
        // the reference's MOD bisection multiplies quotient candidates by the
        // divisor modulo 2^1024: exact only while the dividend is below 2^1023
        if (x.kind == PK::Mod && p.n[x.a].bits > 1023) return false;
        if (x.kind == PK::Div) {
            int s = p.n[x.b].S + 4;
            if (s > 38) return false;
            // the reference divides in one chunk when the dividend leaves room
            // for s more decimal digits (decimal.cu:2283-2300)
            int bitsA = std::min(p.n[x.a].bits, 1024);
            int dtz = (bitsA - 1) / 32;
            if ((32 - dtz) * 9 - 10 < s) return false;
        }
    }
    return true;
}

static bool plan_expression_impl(const std::string& text, const std::vector<std::string>& names,
                                 const std::vector<ColumnType>& types, int mode, Plan* out, std::string* err)
{
    for (int attempt = 0; attempt < 2; attempt++) {
        Mode m;
        if (mode == 1) m = Mode::Faithful;
        else if (mode == 2) m = Mode::Extended;
        else m = attempt == 0 ? Mode::Faithful : Mode::Extended;
        ExprP e = parse_expr(text, names, types, m, err);
        if (!e) return false;
        ExprP rw = align_optimize(e, m);
        if (!rw) { *err = "expression rewrites to a constant or integer"; return false; }
        Plan p;
        if (!build_plan(rw, types, m, &p, err)) return false;
        bool inside = plan_is_faithful_envelope(p, types);
        if (mode == 0 && m == Mode::Faithful && !inside) continue;
        if (mode == 1 && !inside) {
            *err = "expression leaves the reference's defined envelope; use extended mode";
            return false;
        }
        *out = std::move(p);
        return true;
    }
    return false;
}

bool plan_expression(const std::string& text, const std::vector<std::string>& names,
                     const std::vector<ColumnType>& types, int mode, Plan* out, std::string* err)
{
    // nothing thrown below may reach the C ABI
    try {
        return plan_expression_impl(text, names, types, mode, out, err);
    } catch (const std::exception& ex) {
        *err = std::string("planner: ") + ex.what();
        return false;
    }
}

// ------------------------------------------------------------------ code generation

namespace {

struct Emit {
    std::ostringstream os;
    int tmp = 0;
    // column-iterator inputs (IterArgs): column loads report NULL in a flag
    // and every plan keeps its own NULL flag <prefix>nn
    bool iter = false;
    // constants read into the kernel's kcv<N> variables before its row loop
    // (emit_const_prologue); else each use reads the constant array
    bool kvars = false;
    std::string fresh(const char* p) { return std::string(p) + std::to_string(tmp++); }
};

// bit c set when column c appears in the subtree of node i
static uint32_t subtree_cols(const Plan& p, int i)
{
    const PNode& x = p.n[i];
    uint32_t m = x.kind == PK::Col ? (1u << x.col) : 0u;
    if (x.a >= 0) m |= subtree_cols(p, x.a);
    if (x.b >= 0) m |= subtree_cols(p, x.b);
    return m;
}

// Emit straight-line code for one plan; returns the variable of the root.
// Column loads are shared across plans through `loaded`.
std::string emit_plan(Emit& E, const Plan& p, const std::vector<ColumnType>& cols, std::vector<std::string>& loaded,
                      const std::string& ind, const std::string& prefix)
{
    const bool faith = p.mode == Mode::Faithful;
    const std::string nnv = prefix + "nn";
    if (E.iter) E.os << ind << "u32 " << nnv << " = 1u;\n";
    std::vector<std::string> var(p.n.size());
    for (size_t i = 0; i < p.n.size(); i++) {
        const PNode& x = p.n[i];
        std::string v = prefix + "n" + std::to_string(i);
        var[i] = v;
        switch (x.kind) {
        case PK::Col: {
            int c = x.col;
            int B = compact_bytes(cols[c].P);
            if (loaded[c].empty()) {
                loaded[c] = "c" + std::to_string(c);
                if (E.iter) {
                    // NULL position, NULL row or position past the column: flag 0
                    E.os << ind << "u32 n" << loaded[c] << " = 1u;\n";
                    E.os << ind << "const Dec<" << (B + 3) / 4 << "> " << loaded[c] << " = col_load<" << B << ">(a, " << c
                         << ", row, n" << loaded[c] << ");\n";
                } else {
                    E.os << ind << "const Dec<" << (B + 3) / 4 << "> " << loaded[c] << " = col_load<" << B << ">(a, " << c
                         << ", row);\n";
                }
            }
            if (E.iter) E.os << ind << nnv << " &= n" << loaded[c] << ";\n";
            E.os << ind << "const Dec<" << x.W << "> " << v << " = dresize<" << x.W << ">(" << loaded[c] << ");\n";
            break;
        }
        case PK::Const: {
            // sign and limbs from the launch's constant array (kernel data)
            const int kb = p.kbase + x.koff;
            auto kw = [&](int i) {
                return E.kvars ? "kcv" + std::to_string(i) : "kld(a.kc, " + std::to_string(i) + ")";
            };
            E.os << ind << "Dec<" << x.W << "> " << v << ";\n";
            E.os << ind << v << ".neg = " << kw(kb) << ";\n";
            for (int w = 0; w < x.W; w++) E.os << ind << v << ".m.w[" << w << "] = " << kw(kb + 1 + w) << ";\n";
            break;
        }
        case PK::Neg:
            E.os << ind << "const Dec<" << x.W << "> " << v << " = dneg<" << x.W << ">(dresize<" << x.W << ">(" << var[x.a]
                 << "));\n";
            break;
        case PK::Add:
        case PK::Sub: {
            const PNode& A = p.n[x.a];
            const PNode& B = p.n[x.b];
            auto aligned = [&](const PNode& c, const std::string& cv) -> std::pair<std::string, int> {
                if (c.S == x.S) return {cv, c.W};
                int sh = x.S - c.S;
                int bits = c.bits + pow10_bits(sh);
                if (faith) bits = std::min(bits, 1024);
                int w = words_for(bits);
                std::string t = E.fresh((prefix + "al").c_str());
                E.os << ind << "const Dec<" << w << "> " << t << " = dscale<" << w << ", " << c.W << ", " << sh << ">(" << cv
                     << ");\n";
                return {t, w};
            };
            auto pa = aligned(A, var[x.a]);
            auto pb = aligned(B, var[x.b]);
            E.os << ind << "const Dec<" << x.W << "> " << v << " = " << (x.kind == PK::Add ? "dadd" : "dsub") << "<" << x.W
                 << ", " << pa.second << ", " << pb.second << ">(" << pa.first << ", " << pb.first << ");\n";
            break;
        }
        case PK::Mul:
            E.os << ind << "const Dec<" << x.W << "> " << v << " = dmul<" << x.W << ", " << p.n[x.a].W << ", " << p.n[x.b].W
                 << ">(" << var[x.a] << ", " << var[x.b] << ");\n";
            break;
        case PK::Div:
            E.os << ind << "const Dec<" << x.W << "> " << v << " = ddiv<" << x.W << ", " << p.n[x.a].W << ", " << p.n[x.b].W
                 << ", " << (p.n[x.b].S + 4) << ", " << (faith ? 1 : 0) << ">(" << var[x.a] << ", " << var[x.b] << ");\n";
            break;
        case PK::Mod:
            E.os << ind << "const Dec<" << x.W << "> " << v << " = dmod<" << x.W << ", " << p.n[x.a].W << ", " << p.n[x.b].W
                 << ", " << (faith ? 1 : 0) << ">(" << var[x.a] << ", " << var[x.b] << ");\n";
            break;
        case PK::Cast: {
            const PNode& A = p.n[x.a];
            E.os << ind << "const Dec<" << x.W << "> " << v << " = ";
            if (x.S > A.S) E.os << "dscale<" << x.W << ", " << A.W << ", " << (x.S - A.S) << ">(";
            else if (x.S < A.S) E.os << "dshr10<" << x.W << ", " << A.W << ", " << (A.S - x.S) << ">(";
            else E.os << "dresize<" << x.W << ">(";
            E.os << var[x.a] << ");\n";
            break;
        }
        case PK::Trunc: {
            const PNode& A = p.n[x.a];
            E.os << ind << "const Dec<" << x.W << "> " << v << " = ";
            if (x.tp > 0) {
                if (x.S > A.S) E.os << "dscale<" << x.W << ", " << A.W << ", " << (x.S - A.S) << ">(";
                else if (x.S < A.S) E.os << "dshr10<" << x.W << ", " << A.W << ", " << (A.S - x.S) << ">(";
                else E.os << "dresize<" << x.W << ">(";
                E.os << var[x.a] << ");\n";
            } else if (A.P - A.S > -x.tp) {
                // floor(M / 10^(S + q)) * 10^q, the sign kept
                E.os << "dscale<" << x.W << ", " << x.W << ", " << -x.tp << ">(dshr10<" << x.W << ", " << A.W << ", "
                     << (A.S - x.tp) << ">(" << var[x.a] << "));\n";
            } else {
                // no integer digits survive: zero magnitude, the sign kept
                E.os << "dtrunc_zero<" << x.W << ", " << A.W << ">(" << var[x.a] << ");\n";
            }
            break;
        }
        }
    }
    if (E.iter) {
        // a division or MOD with a nullable operand is NULL when its divisor
        // is zero (nullable_type operator/ and %, aries_types.hxx:626-663)
        for (size_t i = 0; i < p.n.size(); i++) {
            const PNode& x = p.n[i];
            if (x.kind != PK::Div && x.kind != PK::Mod) continue;
            const uint32_t m = subtree_cols(p, (int)i);
            E.os << ind << "if ((a.nullcols & " << m << "u) && is_zero<" << p.n[x.b].W << ">(" << var[x.b] << ".m)) " << nnv
                 << " = 0u;\n";
        }
    }
    return var[p.root];
}

// Plans with many operations (a long series) or wide limb arithmetic (a
// division of 500-bit operands) are evaluated by an out-of-line
// __noinline__ function: the loop around the row evaluation stays small,
// which keeps LLVM's loop passes, and with them the hipRTC compile time,
// close to linear in the plan size (inline in the aggregation loop, a
// 16-limb division alone took more than 15 minutes to compile).  Short,
// narrow plans stay inline.
constexpr int kOutOfLineOps = 16;
constexpr int kOutOfLineCost = 600;   // limb products, roughly

int plan_cost(const Plan& p)
{
    int cost = 0;
    for (const PNode& x : p.n) {
        const int wa = x.a >= 0 ? p.n[x.a].W : 0, wb = x.b >= 0 ? p.n[x.b].W : 0;
        switch (x.kind) {
        case PK::Mul: cost += wa * wb; break;
        case PK::Div: cost += 4 * (wa + words_for(pow10_bits(p.n[x.b].S + 4)) + 1) * wb; break;
        case PK::Mod: cost += 4 * (wa + 1) * wb; break;
        case PK::Add:
        case PK::Sub: cost += x.W; break;
        case PK::Cast: cost += wa * (1 + std::abs(x.S - p.n[x.a].S) / 9); break;
        case PK::Trunc: cost += wa * (2 + (p.n[x.a].S + std::abs(x.tp)) / 9); break;
        default: break;
        }
    }
    return cost;
}

static int plan_ops(const Plan& p)
{
    int ops = 0;
    for (const PNode& x : p.n) ops += (x.kind != PK::Col && x.kind != PK::Const) ? 1 : 0;
    return ops;
}

bool plan_out_of_line(const Plan& p) { return plan_ops(p) > kOutOfLineOps || plan_cost(p) > kOutOfLineCost; }

}  // namespace

// Algorithmic work per row in 32 x 32 -> 64 limb products, at the plan's
// static widths: a product Wa Wb; a division the scaling of the dividend by
// 10^S (Wa W(10^S)) and Knuth D's (WX - Wb + 1) Wb; a MOD (Wa - Wb + 1) Wb;
// an operand aligned up by 10^n W W(10^n); a cast or truncation down by 10^n
// one word division per limb and 9 digits.
int64_t plan_limb_products(const Plan& p)
{
    int64_t m = 0;
    auto w10 = [](int n) { return (int64_t)words_for(pow10_bits(n) + 1); };
    for (const PNode& x : p.n) {
        const int64_t wa = x.a >= 0 ? p.n[x.a].W : 0, wb = x.b >= 0 ? p.n[x.b].W : 0;
        switch (x.kind) {
        case PK::Mul: m += wa * wb; break;
        case PK::Div: {
            const int S = p.n[x.b].S + 4;
            const int64_t wx = words_for(p.n[x.a].bits + pow10_bits(S));
            m += wa * w10(S) + std::max<int64_t>(1, wx - wb + 1) * wb;
            break;
        }
        case PK::Mod: m += std::max<int64_t>(1, wa - wb + 1) * wb; break;
        case PK::Add:
        case PK::Sub: {
            const PNode& A = p.n[x.a];
            const PNode& B = p.n[x.b];
            if (x.S > A.S) m += A.W * w10(x.S - A.S);
            if (x.S > B.S) m += B.W * w10(x.S - B.S);
            break;
        }
        case PK::Cast:
        case PK::Trunc: {
            const PNode& A = p.n[x.a];
            if (x.S > A.S) m += wa * w10(x.S - A.S);
            else if (x.S < A.S) m += wa * ((A.S - x.S + 8) / 9);
            break;
        }
        default: break;
        }
    }
    return m;
}

namespace {


// Read the constants of the inline plans into kernel-scope variables
// kcv<N> before the row loop, when there are few of them (they then stay in
// scalar registers); a kernel with more reads them at each use.  The reads
// go through the constant address space (kld), so they are scalar loads.
constexpr int kMaxHoistedConstWords = 32;
void emit_const_prologue(Emit& E, const std::vector<Plan>& plans, bool all_inline = false)
{
    auto inl = [&](const Plan& p) { return all_inline || !plan_out_of_line(p); };
    int words = 0;
    for (const Plan& p : plans)
        if (inl(p)) words += p.kwords;
    E.kvars = words <= kMaxHoistedConstWords;
    if (!E.kvars) return;
    for (const Plan& p : plans) {
        if (!inl(p)) continue;
        for (int w = 0; w < p.kwords; w++)
            E.os << "  const u32 kcv" << p.kbase + w << " = kld(a.kc, " << p.kbase + w << ");\n";
    }
}

// definition of the out-of-line evaluator `fname` (any args struct whose
// first members are col[] and colbytes[])
void emit_plan_function(Emit& E, const Plan& p, const std::vector<ColumnType>& cols, const std::string& fname)
{
    const int W = p.n[p.root].W;
    E.os << "template <class A> __device__ __noinline__ Dec<" << W << "> " << fname << "(const A& a, usize row"
         << (E.iter ? ", u32& nn_out" : "") << ")\n{\n";
    std::vector<std::string> loaded(cols.size());
    const bool kv = E.kvars;
    E.kvars = false;   // the kernel's kcv<N> are not visible here
    std::string rv = emit_plan(E, p, cols, loaded, "  ", "");
    E.kvars = kv;
    if (E.iter) E.os << "  nn_out = nn;\n";
    E.os << "  return " << rv << ";\n}\n";
}

// evaluate plan p for `row`: inline, or a call of the out-of-line evaluator
std::string emit_plan_any(Emit& E, const Plan& p, const std::vector<ColumnType>& cols, std::vector<std::string>& loaded,
                          const std::string& ind, const std::string& prefix, const std::string& fname)
{
    if (!plan_out_of_line(p)) return emit_plan(E, p, cols, loaded, ind, prefix);
    const std::string v = prefix + "r";
    if (E.iter) E.os << ind << "u32 " << prefix << "nn = 1u;\n";
    E.os << ind << "const Dec<" << p.n[p.root].W << "> " << v << " = " << fname << "(a, row"
         << (E.iter ? ", " + prefix + "nn" : std::string()) << ");\n";
    return v;
}

std::string plan_fn_name(int i) { return "upd_plan" + std::to_string(i); }

void emit_plan_functions(Emit& E, const std::vector<Plan>& plans, const std::vector<ColumnType>& cols)
{
    for (int i = 0; i < (int)plans.size(); i++)
        if (plan_out_of_line(plans[i])) emit_plan_function(E, plans[i], cols, plan_fn_name(i));
}

// Blocks are dealt to the 8 XCDs round-robin (block b runs on XCD b mod 8)
// and each XCD has its own L2.  xcd_block() renumbers the blocks so that XCD
// x owns the logical range [x G/8, (x+1) G/8): the row ranges of one
// grid-stride step are then contiguous per XCD, and a cache line that
// straddles two blocks' ranges is fetched into one L2, not two.
const char* kArgsStruct =
    // constants are read through the constant address space: the compiler
    // may then use scalar loads for them
    "__device__ __forceinline__ u32 kld(const u32* p, int i)\n"
    "{ return ((const __attribute__((address_space(4))) u32*)p)[i]; }\n"
    "__device__ __forceinline__ usize xcd_block() {\n"
    "  const u32 g = gridDim.x, b = blockIdx.x;\n"
    "  return (g & 7u) ? (usize)b : (usize)((b & 7u) * (g >> 3) + (b >> 3));\n"
    "}\n"
    "#define UPD_MAXCOLS 16\n"
    // kc: the plans' constants (sign, limbs per constant), the reference's
    // constValues (AriesExprCalcNode.cpp:346-349)
    "struct ColArgs { const u8* col[UPD_MAXCOLS]; usize colbytes[UPD_MAXCOLS]; usize rows; u8* out;\n"
    "                 const u32* __restrict__ kc; };\n"
    "struct AggArgs { const u8* col[UPD_MAXCOLS]; usize colbytes[UPD_MAXCOLS]; usize rows;\n"
    "                 const i32* fcol; const u8* key0; const u8* key1; const u8* lut; u32* partials;\n"
    "                 const u32* __restrict__ kc; };\n"
    // column iterators (AriesColumnDataIterator, AriesColumnDataIterator.hxx:8-41,
    // operator[] in AriesColumnDataIterator.cu:1-117): column c is nblocks
    // data blocks starting at positions starts[k] (positions in all), rows of
    // `stride` bytes (a flag byte first when flags bit 0 marks nullable data);
    // with an index (flags bits 1-2: 1 int32 with -1 = NULL, 2 nullable_type<int32>
    // of 5 packed bytes) launch row i reads the position the index holds at i,
    // from one array (idx) or from index blocks starting at rows istarts[k]
    "struct ColIter { const u8* const* blocks; const i64* starts; usize positions; i32 nblocks; u32 stride;\n"
    "                 const void* idx; const void* const* iblocks; const i64* istarts; i32 niblocks; u32 flags; };\n"
    "struct IterArgs { ColIter it[UPD_MAXCOLS]; usize rows; const i32* fcol; const u8* key0; const u8* key1;\n"
    "                  const u8* lut; u32* partials; u8* out; u8* valid; const u32* __restrict__ kc;\n"
    "                  u32 outnull; u32 nullcols; u32 skipnull; u32 pad; };\n"
    "template <int B> __device__ __forceinline__ Dec<(B + 3) / 4> col_load(const ColArgs& a, int c, usize row)\n"
    "{ return load_compact<B>(a.col[c], row, a.colbytes[c]); }\n"
    "template <int B> __device__ __forceinline__ Dec<(B + 3) / 4> col_load(const AggArgs& a, int c, usize row)\n"
    "{ return load_compact<B>(a.col[c], row, a.colbytes[c]); }\n"
    // the value of column c at launch row i; nn = 0 (and a zero value) for a
    // NULL index entry, a position outside the column or a NULL row
    "template <int B> __device__ __forceinline__ Dec<(B + 3) / 4> col_load(const IterArgs& a, int c, usize i, u32& nn)\n"
    "{\n"
    "  const ColIter& t = a.it[c];\n"
    "  Dec<(B + 3) / 4> z;\n"
    "  z.m = mzero<(B + 3) / 4>();\n"
    "  z.neg = 0u;\n"
    "  i64 pos = (i64)i;\n"
    "  const u32 kind = (t.flags >> 1) & 3u;\n"
    "  if (kind) {\n"
    "    const u8* ib = reinterpret_cast<const u8*>(t.idx);\n"
    "    usize off = i;\n"
    "    if (t.niblocks > 1) {\n"
    "      int lo = 0, hi = t.niblocks;\n"
    "      while (hi - lo > 1) { const int mid = (lo + hi) >> 1; if ((usize)t.istarts[mid] <= i) lo = mid; else hi = mid; }\n"
    "      ib = reinterpret_cast<const u8*>(t.iblocks[lo]);\n"
    "      off = i - (usize)t.istarts[lo];\n"
    "    }\n"
    "    if (kind == 1u) {\n"
    "      pos = (i64)reinterpret_cast<const i32*>(ib)[off];\n"
    "    } else {\n"
    "      const u8* e = ib + off * 5u;\n"
    "      pos = e[0] ? (i64)(i32)((u32)e[1] | ((u32)e[2] << 8) | ((u32)e[3] << 16) | ((u32)e[4] << 24)) : (i64)-1;\n"
    "    }\n"
    "  }\n"
    "  if (pos < 0 || (usize)pos >= t.positions) { nn = 0u; return z; }\n"
    "  int lo = 0, hi = t.nblocks;\n"
    "  while (hi - lo > 1) { const int mid = (lo + hi) >> 1; if ((usize)t.starts[mid] <= (usize)pos) lo = mid; else hi = mid; }\n"
    "  const usize first = (usize)t.starts[lo];\n"
    "  const usize end = lo + 1 < t.nblocks ? (usize)t.starts[lo + 1] : t.positions;\n"
    "  const usize base = (usize)t.blocks[lo];\n"
    "  const u8* al = reinterpret_cast<const u8*>(base & ~(usize)3);\n"
    "  const usize sh = base & 3u;\n"
    "  const usize off = sh + ((usize)pos - first) * t.stride;\n"
    "  const usize lim = sh + (end - first) * t.stride;\n"
    "  if (t.flags & 1u) {\n"
    "    if (!al[off]) { nn = 0u; return z; }\n"
    "    return load_compact_off<B>(al, off + 1u, lim);\n"
    "  }\n"
    "  return load_compact_off<B>(al, off, lim);\n"
    "}\n";

// Columns of narrow rows (B <= 32) are read through a per-wave LDS stage
// with 16-byte loads (stage_wave / load_staged), up to kStageWordsPerWave
// words per wave; wider ones and partial waves use per-lane loads.
constexpr int kStageWordsPerWave = 1024;

struct StagePlan {
    std::vector<int> col, off;
    int words = 0;   // per wave
};

static std::vector<int> inline_columns(const AggSpec& s)
{
    std::vector<int> out;
    for (const Plan& p : s.plans) {
        if (plan_out_of_line(p)) continue;
        for (const PNode& x : p.n)
            if (x.kind == PK::Col && std::find(out.begin(), out.end(), x.col) == out.end()) out.push_back(x.col);
    }
    return out;
}

static StagePlan stage_plan(const AggSpec& s, const std::vector<ColumnType>& cols)
{
    StagePlan sp;
    for (int c : inline_columns(s)) {
        const int B = compact_bytes(cols[c].P);
        const int w = 16 * B + 2;
        if (B > 32 || sp.words + w > kStageWordsPerWave) continue;
        sp.col.push_back(c);
        sp.off.push_back(sp.words);
        sp.words += w;
    }
    return sp;
}

// declarations before the row loop (needs tid, wave, lane)
static void emit_stage_decl(Emit& E, const StagePlan& sp, int nwave)
{
    if (sp.col.empty()) return;
    E.os << "  __shared__ u32 stg[" << nwave << " * " << sp.words << "];\n";
    E.os << "  u32* const wstg = stg + wave * " << sp.words << "u;\n";
    E.os << "  const bool stg_aligned = ((";
    for (size_t k = 0; k < sp.col.size(); k++) E.os << (k ? " | " : "") << "(usize)a.col[" << sp.col[k] << "]";
    E.os << ") & 15u) == 0;\n";
}

static void emit_row_preload(Emit& E, const AggSpec& s, const std::vector<ColumnType>& cols, const StagePlan& sp,
                             std::vector<std::string>& loaded, const std::string& ind)
{
    if (!sp.col.empty()) {
        // wave-uniform: the wave's 64 rows are all in range and the columns aligned
        E.os << ind << "const usize wb = row - lane;\n";
        E.os << ind << "const bool wfull = stg_aligned && wb + 64u <= a.rows;\n";
        E.os << ind << "if (wfull) {\n";
        for (size_t k = 0; k < sp.col.size(); k++)
            E.os << ind << "  stage_wave<" << compact_bytes(cols[sp.col[k]].P) << ">(a.col[" << sp.col[k] << "], wb, wstg + "
                 << sp.off[k] << "u, lane);\n";
        E.os << ind << "}\n";
        // LDS accesses of one wave are performed in order; the fences keep
        // the compiler from moving the reads above the writes
        E.os << ind << "__builtin_amdgcn_fence(__ATOMIC_RELEASE, \"wavefront\");\n";
        E.os << ind << "__builtin_amdgcn_wave_barrier();\n";
        E.os << ind << "__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, \"wavefront\");\n";
    }
    for (int c : inline_columns(s)) {
        if (!loaded[c].empty()) continue;
        const int B = compact_bytes(cols[c].P);
        loaded[c] = "c" + std::to_string(c);
        auto it = std::find(sp.col.begin(), sp.col.end(), c);
        E.os << ind << "const Dec<" << (B + 3) / 4 << "> " << loaded[c] << " = ";
        if (it != sp.col.end())
            E.os << "wfull ? load_staged<" << B << ">(wstg + " << sp.off[it - sp.col.begin()] << "u, lane) : ";
        E.os << "col_load<" << B << ">(a, " << c << ", row);\n";
    }
}

// quad registers of the inline columns, -1 when one is too wide
static int quad_words(const std::vector<int>& incols, const std::vector<ColumnType>& cols)
{
    int w = 0;
    for (int c : incols) {
        const int B = compact_bytes(cols[c].P);
        if (B > kQuadMaxBytes) return -1;
        w += B;
    }
    return w;
}

// Opens the row loop of a generated kernel (needs tid, lblk, and the stage
// declarations when sp stages columns).  The body that follows sees `row`,
// `fv` (filter value), `key` (when keys) and the inline columns named in
// `loaded`, and may `continue` to drop its row.  rps = 4: the lane's row
// quad, columns loaded as quads and decoded per row.  Returns the body's
// indentation.
static std::string emit_row_loop_open(Emit& E, const AggSpec& s, const std::vector<ColumnType>& cols, int rps,
                                      const StagePlan& sp, bool keys, std::vector<std::string>& loaded,
                                      bool blocks = false)
{
    if (blocks) {
        // column iterators: the plans load their columns at launch row
        // `row` through the iterators (NULL flags per plan); the filter and
        // key columns are given per launch row
        E.os << "  for (usize row = lblk * T + tid; row < a.rows; row += (usize)gridDim.x * T) {\n";
        if (s.filter_col >= 0) E.os << "    const i32 fv = a.fcol[row];\n";
        if (keys) {
            if (s.nkeys == 2) E.os << "    const u32 key = ((u32)a.key0[row] << 8) | a.key1[row];\n";
            else E.os << "    const u32 key = a.key0[row];\n";
        }
        return "    ";
    }
    if (rps == 1) {
        E.os << "  for (usize row = lblk * T + tid; row < a.rows; row += (usize)gridDim.x * T) {\n";
        if (s.filter_col >= 0) E.os << "    const i32 fv = a.fcol[row];\n";
        if (keys) {
            if (s.nkeys == 2) E.os << "    const u32 key = ((u32)a.key0[row] << 8) | a.key1[row];\n";
            else E.os << "    const u32 key = a.key0[row];\n";
        }
        emit_row_preload(E, s, cols, sp, loaded, "    ");
        return "    ";
    }
    E.os << "  for (usize q = lblk * T + tid; q * 4u < a.rows; q += (usize)gridDim.x * T) {\n";
    if (s.filter_col >= 0)
        E.os << "    const Quad<4> fq = load_quad<4>(reinterpret_cast<const u8*>(a.fcol), q, a.rows * 4u);\n";
    if (keys) {
        E.os << "    const Quad<1> kq0 = load_quad<1>(a.key0, q, a.rows);\n";
        if (s.nkeys == 2) E.os << "    const Quad<1> kq1 = load_quad<1>(a.key1, q, a.rows);\n";
    }
    const std::vector<int> inl = inline_columns(s);
    for (int c : inl) {
        const int B = compact_bytes(cols[c].P);
        E.os << "    const Quad<" << B << "> q" << c << " = load_quad<" << B << ">(a.col[" << c << "], q, a.colbytes[" << c
             << "]);\n";
    }
    E.os << "#pragma unroll\n    for (u32 k = 0; k < 4u; k++) {\n";
    E.os << "      const usize row = q * 4u + k;\n";
    E.os << "      if (row >= a.rows) continue;\n";
    if (s.filter_col >= 0) E.os << "      const i32 fv = (i32)fq.w[k];\n";
    if (keys) {
        if (s.nkeys == 2) E.os << "      const u32 key = (quad_byte(kq0, k) << 8) | quad_byte(kq1, k);\n";
        else E.os << "      const u32 key = quad_byte(kq0, k);\n";
    }
    for (int c : inl) {
        const int B = compact_bytes(cols[c].P);
        loaded[c] = "c" + std::to_string(c);
        E.os << "      const Dec<" << (B + 3) / 4 << "> " << loaded[c] << " = quad_row<" << B << ">(q" << c << ", k);\n";
    }
    return "      ";
}

static void emit_row_loop_close(Emit& E, int rps) { E.os << (rps == 4 ? "    }\n  }\n" : "  }\n"); }

}  // namespace

// A costly plan with few operations (wide limbs, not a long series) runs in
// the calc kernel inline, one row per thread and no row loop: the loop
// passes that made inline bodies slow to compile have no loop to work on
// (2-3 s for a P = 300 division), and the evaluator's arguments and result
// no longer pass through scratch (P = 150 sweep: 2.34 -> 1.87 ms).
bool calc_one_row(const Plan& p) { return plan_out_of_line(p) && plan_ops(p) <= kOutOfLineOps; }

// The calc kernel assembles rows in an LDS tile of block x B(P_out) bytes;
// for rows wider than 64 bytes a 256-row tile takes enough LDS to cap the
// blocks per CU, so those kernels run 64-row blocks (P = 150 sweep: 1.86 ->
// 1.63 ms per 10 M rows; P = 300: 16.7 -> 16.2 ms).
int calc_block(const Plan& p) { return p.outB > 64 ? 64 : 256; }

// Widest limb array a plan keeps live: a node's own limbs, and for a
// division the window of its long division (quotient + divisor + 1 limbs).
static int plan_peak_limbs(const Plan& p)
{
    int peak = 0;
    for (const PNode& n : p.n) {
        peak = std::max(peak, n.W);
        if ((n.kind == PK::Div || n.kind == PK::Mod) && n.b >= 0) peak = std::max(peak, n.W + p.n[n.b].W + 1);
    }
    return peak;
}

// A calc kernel whose plan keeps 48 or more limbs live runs one wave per
// SIMD with the whole register file (512 VGPRs): at the default occupancy
// target the compiler keeps those limb arrays partly in scratch (P = 300
// sweep: 388 B per lane, 16.3 ms per 10 M rows; one wave per SIMD: no
// scratch, 8.6 ms).  Narrower plans keep the default (P = 150: 1.6 ms, one
// wave per SIMD would be 2.5 ms).
static bool calc_one_wave(const Plan& p) { return plan_peak_limbs(p) >= 48; }

std::string gen_calc_kernel(const Plan& p, const std::vector<ColumnType>& cols, const std::string& name, const GenOptions& o)
{
    Emit E;
    E.iter = o.blocks;
    const int T = o.block;
    const int OB = p.outB;
    E.os << "#include \"dec_device.h\"\nusing namespace updec;\n" << kArgsStruct;
    E.os << "// plan: " << p.text(false) << "\n";
    const bool one_row = calc_one_row(p);
    std::vector<std::string> loaded(cols.size());
    // emit_plan_any calls the out-of-line evaluator; one_row inlines instead
    auto body = [&](const std::string& ind) {
        return one_row ? emit_plan(E, p, cols, loaded, ind, "") : emit_plan_any(E, p, cols, loaded, ind, "", plan_fn_name(0));
    };
    if (OB == 0) {
        // a result of precision 0 has no bytes (B(0) = 0, decimal.cu:92):
        // nothing is stored; the iterator form still reports which rows
        // are NULL (and writes the flag byte of a nullable output)
        E.os << "extern \"C\" __global__ void __launch_bounds__(" << T << ") " << name << "("
             << (o.blocks ? "IterArgs" : "ColArgs") << " a)\n{\n";
        if (o.blocks) {
            E.os << "  for (usize row = (usize)blockIdx.x * " << T << " + threadIdx.x; row < a.rows; row += (usize)gridDim.x * "
                 << T << ") {\n";
            emit_plan(E, p, cols, loaded, "    ", "");
            E.os << "    if (a.valid) a.valid[row] = (u8)nn;\n";
            E.os << "    if (a.outnull) a.out[row] = (u8)nn;\n  }\n";
        }
        E.os << "}\n";
        return E.os.str();
    }
    if (!one_row) emit_plan_functions(E, {p}, cols);
    // Rows whose image is 4, 8 or a multiple of 16 bytes go straight to HBM
    // as dword / dwordx2 / dwordx4 stores (coalesced: lane i writes bytes
    // [i*OB, (i+1)*OB) of the block's span); other widths, and every row of
    // the iterator form (whose rows may carry a flag byte), are assembled in
    // an LDS tile and copied out in dwords.
    const int VB = o.blocks ? 0 : OB % 16 == 0 ? 16 : (OB == 8 || OB == 4) ? OB : 0;
    // output row stride: OB, or OB + 1 for the iterator form's nullable
    // output (flag byte, then the compact row: GenSaveDataCode,
    // AriesExprCalcNode.cpp:76-108)
    const std::string OS = o.blocks ? "os_" : std::to_string(OB);
    E.os << "extern \"C\" __global__ void __launch_bounds__(" << T << ") "
         << (calc_one_wave(p) ? "__attribute__((amdgpu_waves_per_eu(1, 1))) " : "") << name << "("
         << (o.blocks ? "IterArgs" : "ColArgs") << " a)\n{\n";
    emit_const_prologue(E, {p}, one_row);
    E.os << "  __shared__ u32 tile[(" << T << " * " << (o.blocks ? OB + 1 : OB) << " + 3) / 4];\n";
    E.os << "  u8* t8 = reinterpret_cast<u8*>(tile);\n";
    if (o.blocks) E.os << "  const usize os_ = " << OB << "u + (a.outnull ? 1u : 0u);\n";
    if (VB) E.os << "  const bool direct = ((usize)a.out & " << VB - 1 << "u) == 0;\n";
    else E.os << "  const bool direct = false;\n";
    if (one_row)   // one block per T rows; `continue` leaves the do-while
        E.os << "  do {\n    const usize base = (usize)blockIdx.x * " << T << ";\n    if (base >= a.rows) break;\n";
    else
        E.os << "  for (usize base = xcd_block() * " << T << "; base < a.rows; base += (usize)gridDim.x * " << T << ") {\n";
    E.os << "    const usize row = base + threadIdx.x;\n";
    E.os << "    if (row < a.rows) {\n";
    if (o.blocks) {
        // launch row `row` reads every column through its iterator; a NULL
        // value gives a zero row, flag 0 and valid 0
        std::string rv = body("      ");
        E.os << "      const Mag<" << (OB + 3) / 4 << "> enc = nn ? encode_compact<" << OB << ", " << p.n[p.root].W << ">(" << rv
             << ") : mzero<" << (OB + 3) / 4 << ">();\n";
        E.os << "      if (a.valid) a.valid[row] = (u8)nn;\n";
        E.os << "      u8* tr = t8 + threadIdx.x * os_;\n";
        E.os << "      if (a.outnull) *tr++ = (u8)nn;\n";
        E.os << "#pragma unroll\n      for (int b = 0; b < " << OB << "; b++) tr[b] = (u8)(enc.w[b >> 2] >> (8 * (b & 3)));\n";
    } else {
        std::string rv = body("      ");
        E.os << "      const Mag<" << (OB + 3) / 4 << "> enc = encode_compact<" << OB << ", " << p.n[p.root].W << ">(" << rv
             << ");\n";
        if (VB) {
            E.os << "      if (direct) {\n";
            if (VB == 4) {
                E.os << "        reinterpret_cast<u32*>(a.out)[row] = enc.w[0];\n";
            } else if (VB == 8) {
                E.os << "        u32x2 v; v[0] = enc.w[0]; v[1] = enc.w[1];\n";
                E.os << "        reinterpret_cast<u32x2*>(a.out)[row] = v;\n";
            } else {
                E.os << "        u32x4* d4 = reinterpret_cast<u32x4*>(a.out + row * " << OB << ");\n";
                for (int q = 0; q < OB / 16; q++)
                    E.os << "        { u32x4 v; v[0] = enc.w[" << 4 * q << "]; v[1] = enc.w[" << 4 * q + 1 << "]; v[2] = enc.w["
                         << 4 * q + 2 << "]; v[3] = enc.w[" << 4 * q + 3 << "]; d4[" << q << "] = v; }\n";
            }
            E.os << "      } else {\n";
        }
        E.os << "#pragma unroll\n      for (int b = 0; b < " << OB << "; b++) t8[threadIdx.x * " << OB
             << " + b] = (u8)(enc.w[b >> 2] >> (8 * (b & 3)));\n";
        if (VB) E.os << "      }\n";
    }
    E.os << "    }\n    if (direct) continue;\n    __syncthreads();\n";
    E.os << "    const usize n = (a.rows - base < " << T << " ? a.rows - base : " << T << ") * " << OS << ";\n";
    E.os << "    u8* dst = a.out + base * " << OS << ";\n";
    E.os << "    for (usize i = threadIdx.x; i < n / 4; i += " << T << ") reinterpret_cast<u32*>(dst)[i] = tile[i];\n";
    E.os << "    for (usize i = (n / 4) * 4 + threadIdx.x; i < n; i += " << T << ") dst[i] = t8[i];\n";
    E.os << "    __syncthreads();\n  }" << (one_row ? " while (0);" : "") << "\n}\n";
    return E.os.str();
}

static int agg_value_bits(const AggSpec& s)
{
    int vbits = 1;
    for (const Plan& p : s.plans) vbits = std::max(vbits, std::min(p.n[p.root].bits, 8 * p.outB - 1));
    return vbits;
}

int agg_acc_words(const AggSpec& s) { return words_for(agg_value_bits(s) + 40 + 1); }

int agg_rows_per_step(const AggSpec& s, const std::vector<ColumnType>& cols, const GenOptions& o)
{
    if (!o.row_quads || o.blocks) return 1;
    const std::vector<int> inl = inline_columns(s);
    const int w = quad_words(inl, cols);
    const int extra = (s.filter_col >= 0 ? 4 : 0) + s.nkeys;
    if (w < 0 || w + extra == 0 || w + extra > kQuadMaxWords) return 1;
    return 4;
}

// rows that can reach one slot copy: 2^14 per lane times the lanes sharing it
static int share_log_for(int split) { return kAggRowsPerLaneLog2 + (split >= 64 ? 0 : 6 - __builtin_ctz((unsigned)split)); }

// Accumulator layout of the grouped kernels: u64 words [copy][slot][plan
// words] where copy = (wave, lane group).  AWR = 1: one word per plan holds
// the signed 64-bit sum directly.  Otherwise plan i has W_i + 1 words (W_i
// the limbs of its stored value, at most AW2) holding the 32-bit limbs of
// the two's complement value separately ("redundant" limbs: each word
// collects its limb's column sum, carries are resolved once at the end), so
// every update is a plain ds_add_u64 with no carry chain; a block's sum of
// values below 2^(32 W_i) over at most 2^20 rows fits W_i + 1 words.
struct AccCfg {
    int T = 256, SPLIT = 4, AWR = 1;
};

static int plan_slot_words(const Plan& p, int AWR, int AW2) { return AWR == 1 ? 1 : std::min(AW2, p.n[p.root].W + 1); }

// words of one slot over all plans, and plan i's first word
static int slot_words_total(const AggSpec& s, int AWR, int AW2)
{
    int w = 0;
    for (const Plan& p : s.plans) w += plan_slot_words(p, AWR, AW2);
    return w;
}
static int slot_word_offset(const AggSpec& s, int i, int AWR, int AW2)
{
    int w = 0;
    for (int k = 0; k < i; k++) w += plan_slot_words(s.plans[k], AWR, AW2);
    return w;
}

static bool choose_acc(const AggSpec& s, int nslots, const GenOptions& o, AccCfg* c)
{
    const int vbits = agg_value_bits(s), AW2 = agg_acc_words(s), NP = (int)s.plans.size();
    const int cands[][2] = {{o.block, o.slot_split}, {o.block, 2}, {o.block, 1}, {128, 1}, {64, 1}};
    for (auto& cd : cands) {
        const int T = cd[0], SP = std::max(1, std::min(64, cd[1]));
        const int AWR = (vbits + share_log_for(SP) + 1 <= 64) ? 1 : AW2;
        // multi-word slots take at most two copies per wave: more copies
        // cost more LDS than the contention they save (Q1 shape without
        // column statistics: 0.539 against 0.564 ms at four)
        if (AWR > 1 && SP > 2) continue;
        const long bytes = (long)(T / 64) * SP * nslots * ((long)slot_words_total(s, AWR, AW2) * 8 + 4);
        if (bytes <= 60 * 1024) {
            c->T = T;
            c->SPLIT = SP;
            c->AWR = AWR;
            return true;
        }
    }
    return false;
}

bool agg_onepass_ok(const AggSpec& s, const GenOptions& o)
{
    AccCfg c;
    return s.nkeys > 0 && choose_acc(s, s.maxslots, o, &c);
}

int agg_onepass_record_words(const AggSpec& s)
{
    return s.maxslots * (1 + (int)s.plans.size() * agg_acc_words(s) + 2);
}

static void emit_agg_prelude(Emit& E, const AggSpec& s, const AccCfg& c, int nslots)
{
    const int NP = (int)s.plans.size();
    E.os << "#include \"dec_device.h\"\nusing namespace updec;\n" << kArgsStruct;
    for (int i = 0; i < NP; i++) E.os << "// plan " << i << ": " << s.plans[i].text(false) << "\n";
    E.os << "#define NP " << NP << "\n#define AW2 " << agg_acc_words(s) << "\n#define T " << c.T << "\n#define SPLIT "
         << c.SPLIT << "\n#define NWAVE (T / 64)\n#define NCOPY (NWAVE * SPLIT)\n#define NS " << nslots
         << "\n#define AWR " << c.AWR << "\n#define SLOTW " << slot_words_total(s, c.AWR, agg_acc_words(s)) << "\n";
    // plan pp's first word and word count in a slot, at run time and as
    // compile-time constants
    E.os << "UPD_DEV u32 slot_off(u32 pp) {\n  u32 r = 0u;\n";
    for (int i = 0; i < NP; i++)
        E.os << "  if (pp == " << i << "u) r = " << slot_word_offset(s, i, c.AWR, agg_acc_words(s)) << "u;\n";
    E.os << "  return r;\n}\nUPD_DEV u32 slot_words(u32 pp) {\n  u32 r = 1u;\n";
    for (int i = 0; i < NP; i++)
        E.os << "  if (pp == " << i << "u) r = " << plan_slot_words(s.plans[i], c.AWR, agg_acc_words(s)) << "u;\n";
    E.os << "  return r;\n}\n";
    for (int i = 0; i < NP; i++)
        E.os << "#define SLOT_OFF" << i << " " << slot_word_offset(s, i, c.AWR, agg_acc_words(s)) << "\n#define SLOT_WORDS" << i
             << " " << plan_slot_words(s.plans[i], c.AWR, agg_acc_words(s)) << "\n";
    // round trip through the materialized compact column (the reference sums
    // the stored compact result, AriesExprCalcNode.cpp:1623-1700)
    E.os << "template <int B, int W> UPD_DEV Dec<W> stored(const Dec<W>& v) {\n"
            "  Dec<W> r = v;\n"
            "  constexpr int TOPBIT = 8 * B - 1;\n"
            "#pragma unroll\n"
            "  for (int i = 0; i < W; i++) {\n"
            "    if (32 * i >= TOPBIT + 1) r.m.w[i] = 0u;\n"
            "    else if (32 * i + 32 > TOPBIT) {\n"
            "      const int bit = TOPBIT - 32 * i;\n"
            "      r.neg |= (v.m.w[i] >> bit) & 1u;\n"
            "      r.m.w[i] &= (bit == 0) ? 0u : (0xffffffffu >> (32 - bit));\n"
            "    }\n"
            "  }\n"
            "  return r;\n"
            "}\n";
    // add one signed value into a slot of AWR u64 words
    E.os << "template <int W, int N> UPD_DEV void slot_add(u64* sl, const Dec<W>& v) {\n"
            "  if constexpr (AWR == 1) {\n"
            "    u64 m = v.m.w[0];\n"
            "    if constexpr (W > 1) m |= (u64)v.m.w[1] << 32;\n"
            "    atomicAdd(sl, v.neg ? (u64)(-(i64)m) : m);\n"
            "  } else {\n"
            "    const u32 x = v.neg ? 0xffffffffu : 0u;\n"
            "#pragma unroll\n"
            "    for (int i = 0; i < N; i++) {\n"
            "      const u64 add = (u64)((i < W ? v.m.w[i] : 0u) ^ x) + (i == 0 ? (u64)(v.neg & 1u) : 0ull);\n"
            "      if (add) atomicAdd(sl + i, add);\n"
            "    }\n"
            "  }\n"
            "}\n";
    // slot -> AW2-limb two's complement value
    E.os << "UPD_DEV Mag<AW2> slot_value(const u64* sl, u32 n) {\n"
            "  Mag<AW2> r;\n"
            "  if constexpr (AWR == 1) {\n"
            "    const u64 x = sl[0];\n"
            "    const u32 ext = (x >> 63) ? 0xffffffffu : 0u;\n"
            "#pragma unroll\n"
            "    for (int w = 0; w < AW2; w++) r.w[w] = w == 0 ? (u32)x : (w == 1 ? (u32)(x >> 32) : ext);\n"
            "  } else {\n"
            "    // n words: the value modulo 2^(32 n), then its sign extended\n"
            "    u64 carry = 0;\n"
            "    u32 ext = 0u;\n"
            "#pragma unroll\n"
            "    for (int w = 0; w < AW2; w++) {\n"
            "      if ((u32)w < n) {\n"
            "        const u64 lo = (sl[w] & 0xffffffffull) + (carry & 0xffffffffull);\n"
            "        r.w[w] = (u32)lo;\n"
            "        carry = (sl[w] >> 32) + (carry >> 32) + (lo >> 32);\n"
            "        ext = (r.w[w] >> 31) ? 0xffffffffu : 0u;\n"
            "      } else {\n"
            "        r.w[w] = ext;\n"
            "      }\n"
            "    }\n"
            "  }\n"
            "  return r;\n"
            "}\n";
}

// per-row body shared by both forms: evaluate every plan and add it to the
// slots at `sbase` (u64 index), then count the row
// Issue every column load of the row (filter value, keys, the decimal
// columns of the inline plans) before anything depends on one of them, so
// a row costs one memory round trip instead of a chain of filter -> key ->
// value loads.  Rows the filter rejects have their lines read anyway.
static void emit_agg_row(Emit& E, const AggSpec& s, const std::vector<ColumnType>& cols,
                         std::vector<std::string>& loaded, const std::string& ind)
{
    for (int i = 0; i < (int)s.plans.size(); i++) {
        const Plan& p = s.plans[i];
        const int W = p.n[p.root].W;
        std::string rv = emit_plan_any(E, p, cols, loaded, ind, "p" + std::to_string(i) + "_", plan_fn_name(i));
        // the iterator form skips a NULL value (agg_sum_t of nullable_type,
        // operators.hxx:1175-1195)
        E.os << ind << (E.iter ? "if (p" + std::to_string(i) + "_nn) " : std::string()) << "slot_add<" << W
             << ", SLOT_WORDS" << i << ">(&aslot[sbase + SLOT_OFF" << i << "], stored<" << p.outB << ", " << W << ">(" << rv
             << "));\n";
    }
}

// the iterator form's row count: every row (COUNT(*)), or with skipnull
// (the single shared index of updec_agg_run_blocks) only rows whose values
// are not NULL
static std::string iter_count_expr(const AggSpec& s)
{
    std::string all;
    for (int i = 0; i < (int)s.plans.size(); i++) all += (i ? " & p" : "p") + std::to_string(i) + "_nn";
    return "(a.skipnull ? (" + all + ") : 1u)";
}

static void emit_agg_final_prec(Emit& E, const AggSpec& s)
{
    E.os << "      u32 prec = 0, frac = 0;\n";
    for (int i = 0; i < (int)s.plans.size(); i++)
        E.os << "      if (pp == " << i << ") { prec = " << (s.plans[i].outP & 0xff) << "u; frac = " << (s.plans[i].outS & 0xff)
             << "u; }\n";
}

static void emit_agg_final_image(Emit& E, const AggSpec& s, const std::string& target)
{
    emit_agg_final_prec(E, s);
    E.os << "      write_sum_image<AW2>(" << target << ", t, prec, frac, 0u);\n";
}

// block sum of the AW2-word accumulators sm of `waves` waves into t (in
// thread 0, which the text leaves inside an open `if (tid == 0) {`)
static std::string block_reduce_aw2(int waves)
{
    std::ostringstream os;
    os << "    wave_reduce<AW2>(sm);\n"
       << "    if ((tid & 63u) == 0) {\n#pragma unroll\n      for (int w = 0; w < AW2; w++) red[(tid >> 6) * AW2 + w] = sm.w[w];\n    }\n"
       << "    __syncthreads();\n"
       << "    if (tid == 0) {\n      Mag<AW2> t = mzero<AW2>();\n      for (int wv = 0; wv < " << waves << "; wv++) {\n        Mag<AW2> o;\n"
       << "#pragma unroll\n        for (int w = 0; w < AW2; w++) o.w[w] = red[wv * AW2 + w];\n        add_to<AW2>(t, t, o);\n      }\n";
    return os.str();
}

static const char* kBlockReduceCount =
    "#pragma unroll\n    for (int off = 32; off >= 1; off >>= 1) c += (u64)__shfl_xor((long long)c, off, 64);\n"
    "    if ((tid & 63u) == 0) { red[2 * (tid >> 6)] = (u32)c; red[2 * (tid >> 6) + 1] = (u32)(c >> 32); }\n"
    "    __syncthreads();\n";

// One-pass grouped aggregation.  agg_main: each block keeps an LDS hash
// table from the 16-bit key to a dense slot (assigned on first sight) and
// 64-bit atomic accumulators per (lane group, slot, plan); at the end it
// writes one record per slot {key+1, sums, count} and ORs its keys into a
// global key bitmap.  agg_final: every block numbers the groups in
// ascending key order from that bitmap (prefix popcounts), then block
// (g, p) sums the records of group g (and divides an AVG expression's sum
// by the group's count).
static std::string gen_agg_onepass(const AggSpec& s, const std::vector<ColumnType>& cols, const std::string& name,
                                   const GenOptions& o, int* acc_words, int* block_out)
{
    Emit E;
    AccCfg c;
    if (!choose_acc(s, s.maxslots, o, &c)) return std::string();
    const int NP = (int)s.plans.size();
    const int AW2 = agg_acc_words(s);
    const int MAXS = s.maxslots;
    const int EW = 1 + NP * AW2 + 2;
    *acc_words = AW2;
    *block_out = c.T;
    const char* fops[] = {"<=", "<", ">=", ">", "==", "!="};
    emit_agg_prelude(E, s, c, MAXS);
    emit_plan_functions(E, s.plans, cols);
    E.os << "// accumulation: one pass, LDS key hash + 64-bit atomic slots\n";
    E.os << "#define MAXS " << MAXS << "\n#define HT 64\n#define EW " << EW << "\n";
    // key -> slot: open addressing over HT entries {key+1 << 8 | slot};
    // slot 0xff marks an entry whose slot number is being assigned.  Table
    // reads and writes are relaxed workgroup-scope atomics, so a waiting lane
    // re-reads the entry; unlike volatile accesses they do not make the wave
    // wait for its outstanding global loads.
    E.os << "UPD_DEV u32 ht_get(u32* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }\n";
    E.os << "UPD_DEV u32 slot_of(u32 key, u32* htab, u32* nslot, u32* skey) {\n"
            "  u32 h = (key * 2654435761u) >> 26;\n"
            "  for (int probe = 0; probe < 4 * HT; probe++) {\n"
            "    u32 e = ht_get(&htab[h]);\n"
            "    if (e == 0u) {\n"
            "      const u32 old = atomicCAS(&htab[h], 0u, ((key + 1u) << 8) | 0xffu);\n"
            "      if (old == 0u) {\n"
            "        u32 sl = atomicAdd(nslot, 1u);\n"
            "        if (sl >= MAXS) sl = 0xfeu;\n"
            "        else skey[sl] = key + 1u;\n"
            "        __hip_atomic_store(&htab[h], ((key + 1u) << 8) | sl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);\n"
            "        return sl;\n"
            "      }\n"
            "      e = old;\n"
            "    }\n"
            "    if ((e >> 8) == key + 1u) {\n"
            "      if ((e & 0xffu) != 0xffu) return e & 0xffu;\n"
            "      continue;\n"
            "    }\n"
            "    h = (h + 1u) & (HT - 1);\n"
            "  }\n"
            "  return 0xfeu;\n"
            "}\n";
    E.os << "extern \"C\" __global__ void __launch_bounds__(" << c.T << ") " << name
         << "_main(AggArgs a, u32* __restrict__ overflow)\n{\n";
    emit_const_prologue(E, s.plans);
    E.os << "  const u32 tid = threadIdx.x;\n  const u32 wave = tid >> 6, lane = tid & 63u;\n";
    // the one-pass form accumulates with LDS atomics and reads its row
    // quads straight from memory.  Staging whole chunks through LDS (16-byte
    // coalesced loads one chunk ahead in registers) reads the Q1 columns
    // alone in 0.29 ms against 0.33 ms for the quads, but with this
    // kernel's work it ran 0.49-0.50 ms against 0.38 ms (two blocks per CU,
    // two barriers per chunk); loading the next quad ahead changed nothing.
    const StagePlan sp;
    emit_stage_decl(E, sp, c.T / 64);
    E.os << "  __shared__ u32 htab[HT];\n  __shared__ u32 nslot;\n  __shared__ u32 skey[MAXS];\n";
    E.os << "  __shared__ u64 aslot[NCOPY * NS * SLOTW];\n  __shared__ u32 acnt[NCOPY * NS];\n";
    E.os << "  for (u32 i = tid; i < HT; i += T) htab[i] = 0u;\n";
    E.os << "  for (u32 i = tid; i < MAXS; i += T) skey[i] = 0u;\n";
    E.os << "  if (tid == 0) nslot = 0u;\n";
    E.os << "  for (u32 i = tid; i < NCOPY * NS * SLOTW; i += T) aslot[i] = 0ull;\n";
    E.os << "  for (u32 i = tid; i < NCOPY * NS; i += T) acnt[i] = 0u;\n";
    E.os << "  __syncthreads();\n";
    E.os << "  const u32 copy = wave * SPLIT + lane / (64 / SPLIT);\n";
    E.os << "  u32 ovf = 0;\n";
    E.os << "  const usize lblk = xcd_block();\n";
    std::vector<std::string> loaded(cols.size());
    const int rps = agg_rows_per_step(s, cols, o);
    const std::string bi = emit_row_loop_open(E, s, cols, rps, sp, true, loaded);
    if (s.filter_col >= 0) E.os << bi << "if (!(fv " << fops[s.filter_op] << " (i64)" << s.filter_val << "LL)) continue;\n";
    E.os << bi << "const u32 g = slot_of(key, htab, &nslot, skey);\n";
    E.os << bi << "if (g >= MAXS) { ovf = 1u; continue; }\n";
    E.os << bi << "const u32 sbase = (copy * NS + g) * SLOTW;\n";
    emit_agg_row(E, s, cols, loaded, bi);
    E.os << bi << "atomicAdd(&acnt[copy * NS + g], 1u);\n";
    emit_row_loop_close(E, rps);
    E.os << "  if (ovf) atomicOr(overflow, 1u);\n";
    E.os << "  __syncthreads();\n";
    // records are stored word-major: word f of slot sl of block b at
    // partials[(sl * EW + f) * nblocks + b], so the final kernel's scans
    // over the blocks read consecutive words
    E.os << "  const usize nb = gridDim.x;\n";
    E.os << "  const u32 ns = nslot < MAXS ? nslot : MAXS;\n";
    // overflow[1]: the most slots any block used (the final kernel scans only those)
    E.os << "  if (tid == 0 && ns) atomicMax(overflow + 1, ns);\n";
    // overflow[2..2050): bitmap of the keys seen by any block (bit k for
    // key k), from which the final kernel orders the groups
    E.os << "  if (tid < ns) { const u32 k = skey[tid] - 1u; atomicOr(overflow + 2 + (k >> 5), 1u << (k & 31u)); }\n";
    E.os << "  for (u32 k = tid; k < MAXS * (NP + 1); k += T) {\n";
    E.os << "    const u32 sl = k / (NP + 1), pp = k % (NP + 1);\n";
    E.os << "    u32* e = a.partials + (usize)sl * EW * nb + blockIdx.x;\n";
    E.os << "    if (pp == 0) e[0] = sl < ns ? skey[sl] : 0u;\n";
    E.os << "    if (pp < NP) {\n      Mag<AW2> sm = mzero<AW2>();\n";
    E.os << "      if (sl < ns) for (u32 cc = 0; cc < NCOPY; cc++) add_to<AW2>(sm, sm, slot_value(&aslot[(cc * NS + sl) * SLOTW + slot_off(pp)], slot_words(pp)));\n";
    E.os << "#pragma unroll\n      for (int w = 0; w < AW2; w++) e[(usize)(1 + pp * AW2 + w) * nb] = sm.w[w];\n";
    E.os << "    } else {\n      u64 cnt = 0;\n";
    E.os << "      if (sl < ns) for (u32 cc = 0; cc < NCOPY; cc++) cnt += acnt[cc * NS + sl];\n";
    E.os << "      e[(usize)(1 + NP * AW2) * nb] = (u32)cnt; e[(usize)(2 + NP * AW2) * nb] = (u32)(cnt >> 32);\n    }\n  }\n}\n";

    // final: stat[0] = 1 when a block's key table overflowed, stat[1] = the
    // most slots any block used, stat[2..2050) = the key bitmap; out (host
    // memory): meta[0] = ngroups, meta[1] = 1 when ngroups > maxg, meta[2] =
    // stat[0], meta[4..] = keys in ascending order.  Block (g, p) sums
    // expression p (p = NP: COUNT) over the records of group g.  Block 0
    // also clears `next`, the state the next run accumulates into.
    const int FT = kAggFinalThreads, FW = FT / 64, BW = 2048 / FT;
    const bool avg = s.avg_mask != 0;
    int navg = 0;
    for (int k = 0; k < NP; k++) navg += (s.avg_mask >> k) & 1u;
    E.os << "UPD_DEV usize rec_at(u32 i, u32 nblocks) { const u32 sl = i / nblocks; return (usize)sl * EW * nblocks + (i - sl * nblocks); }\n";
    E.os << "extern \"C\" __global__ void __launch_bounds__(" << FT << ") " << name
         << "_final(const u32* __restrict__ rec, u32 nblocks, const u32* __restrict__ stat, u32* __restrict__ next, u32 maxg, u8* __restrict__ images, u64* __restrict__ counts, u32* __restrict__ meta, u8* __restrict__ avgs)\n{\n";
    E.os << "  const u32 tid = threadIdx.x;\n";
    E.os << "  __shared__ u32 bm[2048];\n  __shared__ u32 pre[2049];\n  __shared__ u32 part[" << FW << "];\n";
    E.os << "  __shared__ u32 red[" << FW << " * AW2 + " << 2 * FW << "];\n  __shared__ u32 gkey;\n";
    if (avg) E.os << "  __shared__ u64 cred[" << FW << "];\n";
    // prefix popcounts of the bitmap: thread t owns words [BW t, BW t + BW);
    // scan in each wave, then over the wave totals
    E.os << "  u32 loc = 0;\n";
    E.os << "#pragma unroll\n  for (int j = 0; j < " << BW << "; j++) { const u32 x = stat[2 + tid * " << BW << " + j]; bm[tid * " << BW
         << " + j] = x; loc += __popc(x); }\n";
    E.os << "  u32 inc = loc;\n";
    E.os << "#pragma unroll\n  for (int off = 1; off < 64; off <<= 1) { const u32 y = __shfl_up(inc, off, 64); if ((tid & 63u) >= (u32)off) inc += y; }\n";
    E.os << "  if ((tid & 63u) == 63u) part[tid >> 6] = inc;\n  __syncthreads();\n";
    E.os << "  u32 wbase = 0;\n  for (u32 wv = 0; wv < (tid >> 6); wv++) wbase += part[wv];\n";
    E.os << "  if (tid == " << FT - 1 << "u) pre[2048] = wbase + inc;\n";
    E.os << "  { u32 acc = wbase + inc - loc;\n#pragma unroll\n    for (int j = 0; j < " << BW << "; j++) { pre[tid * " << BW
         << " + j] = acc; acc += __popc(bm[tid * " << BW << " + j]); } }\n";
    E.os << "  __syncthreads();\n";
    E.os << "  const u32 ngroups = pre[2048];\n";
    E.os << "  if (blockIdx.x == 0) {\n";
    E.os << "    if (tid == 0) { meta[0] = ngroups; meta[1] = ngroups > maxg ? 1u : 0u; meta[2] = stat[0]; }\n";
    E.os << "    for (u32 i = tid; i < " << kAggOnepassStatWords << "u; i += " << FT << ") next[i] = 0u;\n";
    E.os << "    for (u32 w = tid; w < 2048; w += " << FT << ") { u32 x = bm[w]; u32 r = pre[w];\n";
    E.os << "      while (x) { const u32 b = __ffs(x) - 1u; x &= x - 1u; if (r < maxg) meta[4 + r] = w * 32u + b; r++; } }\n";
    E.os << "  }\n";
    E.os << "  const u32 g = blockIdx.x / (NP + 1), pp = blockIdx.x % (NP + 1);\n";
    E.os << "  if (g >= ngroups || g >= maxg) return;\n";
    E.os << "  for (u32 w = tid; w < 2048; w += " << FT << ") { u32 x = bm[w]; u32 r = pre[w];\n";
    E.os << "    while (x) { const u32 b = __ffs(x) - 1u; x &= x - 1u; if (r == g) gkey = w * 32u + b + 1u; r++; } }\n";
    E.os << "  __syncthreads();\n  const u32 want = gkey;\n";
    // Record i of the used slots * nblocks is slot i / nblocks of block
    // i % nblocks.  Every record of a batch is loaded whole, matching or not,
    // so that the batch's loads are all in flight at once (the records sit
    // in other XCDs' caches); the key test selects after.
    E.os << "  const u32 nrec = nblocks * stat[1];\n";
    const int RW = 1 + AW2 + (avg ? 2 : 0);
    const int D = std::max(2, std::min(16, 96 / RW));
    E.os << "  if (pp < NP) {\n    Mag<AW2> sm = mzero<AW2>();\n";
    if (avg) E.os << "    const bool avg = ((" << s.avg_mask << "u >> pp) & 1u) != 0u;\n    u64 c = 0;\n";
    E.os << "    for (u32 i0 = tid; i0 < nrec; i0 += " << FT << "u * " << D << "u) {\n      u32 kk[" << D << "];\n      Mag<AW2> o[" << D << "];\n";
    if (avg) E.os << "      u64 oc[" << D << "];\n";
    E.os << "#pragma unroll\n      for (int j = 0; j < " << D << "; j++) {\n        const u32 i = i0 + " << FT << "u * j;\n";
    E.os << "        const usize at = i < nrec ? rec_at(i, nblocks) : 0;\n        kk[j] = i < nrec ? rec[at] : 0u;\n";
    E.os << "#pragma unroll\n        for (int w = 0; w < AW2; w++) o[j].w[w] = i < nrec ? rec[at + (usize)(1 + pp * AW2 + w) * nblocks] : 0u;\n";
    if (avg)
        E.os << "        oc[j] = (avg && i < nrec) ? ((u64)rec[at + (usize)(1 + NP * AW2) * nblocks] | ((u64)rec[at + (usize)(2 + NP * AW2) * nblocks] << 32)) : 0ull;\n";
    E.os << "      }\n";
    E.os << "#pragma unroll\n      for (int j = 0; j < " << D << "; j++) {\n        if (kk[j] != want) continue;\n";
    E.os << "        add_to<AW2>(sm, sm, o[j]);\n";
    if (avg) E.os << "        c += oc[j];\n";
    E.os << "      }\n    }\n";
    if (avg) {
        E.os << "#pragma unroll\n    for (int off = 32; off >= 1; off >>= 1) c += (u64)__shfl_xor((long long)c, off, 64);\n";
        E.os << "    if ((tid & 63u) == 0) cred[tid >> 6] = c;\n";
    }
    E.os << block_reduce_aw2(FW);
    emit_agg_final_prec(E, s);
    E.os << "      const Dec<AW2> v = sum_value<AW2>(t);\n";
    E.os << "      write_dec_image<AW2>(images + ((usize)g * NP + pp) * 133u, v, prec, frac, 0u);\n";
    // a SUM block of an AVG expression also counted its group: it divides
    if (avg) {
        E.os << "      if (avg) {\n        u32 aj = 0;\n";
        for (int k = 0, j = 0; k < NP; k++)
            if ((s.avg_mask >> k) & 1u) E.os << "        if (pp == " << k << ") aj = " << j++ << "u;\n";
        E.os << "        u64 n = 0;\n        for (int wv = 0; wv < " << FW << "; wv++) n += cred[wv];\n";
        E.os << "        div_dec_int64<" << (s.avg_faithful ? 1 : 0) << ", AW2>(v, prec, frac, 0u, (u8)1, (i64)n, avgs + ((usize)g * "
             << navg << " + aj) * 133u);\n";
        E.os << "      }\n";
    }
    E.os << "    }\n  } else {\n    u64 c = 0;\n";
    E.os << "    for (u32 i0 = tid; i0 < nrec; i0 += " << FT << "u * 16u) {\n      u32 kk[16];\n      u64 oc[16];\n";
    E.os << "#pragma unroll\n      for (int j = 0; j < 16; j++) {\n        const u32 i = i0 + " << FT << "u * j;\n";
    E.os << "        const usize at = i < nrec ? rec_at(i, nblocks) : 0;\n        kk[j] = i < nrec ? rec[at] : 0u;\n";
    E.os << "        oc[j] = i < nrec ? ((u64)rec[at + (usize)(1 + NP * AW2) * nblocks] | ((u64)rec[at + (usize)(2 + NP * AW2) * nblocks] << 32)) : 0ull;\n";
    E.os << "      }\n";
    E.os << "#pragma unroll\n      for (int j = 0; j < 16; j++)\n        if (kk[j] == want) c += oc[j];\n";
    E.os << "    }\n";
    E.os << kBlockReduceCount;
    E.os << "    if (tid == 0) { u64 t = 0; for (int wv = 0; wv < " << FW
         << "; wv++) t += (u64)red[2 * wv] | ((u64)red[2 * wv + 1] << 32); counts[g] = t; }\n  }\n}\n";
    return E.os.str();
}

std::string gen_agg_kernels(const AggSpec& s, const std::vector<ColumnType>& cols, const std::string& name,
                            const GenOptions& o, int* acc_words, int* block_out)
{
    if (s.onepass) return gen_agg_onepass(s, cols, name, o, acc_words, block_out);
    Emit E;
    E.iter = o.blocks;
    const int NP = (int)s.plans.size();
    const int NG = s.ngroups;
    const int AW2 = agg_acc_words(s);
    *acc_words = AW2;
    const bool reg = NG == 1;
    AccCfg c;
    if (!reg && !choose_acc(s, NG, o, &c)) return std::string();
    if (reg) c.T = o.block;
    *block_out = c.T;
    const int PW = NP * NG * AW2 + 2 * NG;  // partial words per block
    const char* fops[] = {"<=", "<", ">=", ">", "==", "!="};
    emit_agg_prelude(E, s, c, NG);
    emit_plan_functions(E, s.plans, cols);
    E.os << "// accumulation: " << (reg ? "registers" : "LDS 64-bit atomic slots") << "\n";
    E.os << "#define NG " << NG << "\n#define PW " << PW << "\n";
    // a lane accumulates at most 2^14 rows (kAggRowsPerLaneLog2)
    const int vbits = agg_value_bits(s);
    const int AW1 = words_for(vbits + kAggRowsPerLaneLog2 + 1);
    E.os << "#define AW1 " << AW1 << "\n";
    E.os << "extern \"C\" __global__ void __launch_bounds__(" << c.T << ") " << name << "_main("
         << (o.blocks ? "IterArgs" : "AggArgs") << " a)\n{\n";
    emit_const_prologue(E, s.plans);
    E.os << "  const u32 tid = threadIdx.x;\n  const u32 wave = tid >> 6, lane = tid & 63u;\n";
    const StagePlan sp = (reg && !o.blocks && agg_rows_per_step(s, cols, o) == 1) ? stage_plan(s, cols) : StagePlan();
    emit_stage_decl(E, sp, c.T / 64);
    if (!reg) {
        E.os << "  __shared__ u64 aslot[NCOPY * NS * SLOTW];\n  __shared__ u32 acnt[NCOPY * NS];\n";
        E.os << "  for (u32 i = tid; i < NCOPY * NS * SLOTW; i += T) aslot[i] = 0ull;\n";
        E.os << "  for (u32 i = tid; i < NCOPY * NS; i += T) acnt[i] = 0u;\n";
        E.os << "  __syncthreads();\n";
        E.os << "  const u32 copy = wave * SPLIT + lane / (64 / SPLIT);\n";
    } else {
        for (int i = 0; i < NP; i++) E.os << "  Mag<AW1> acc" << i << " = mzero<AW1>();\n";
        E.os << "  u32 cnt = 0;\n";
    }
    E.os << "  const usize lblk = xcd_block();\n";
    std::vector<std::string> loaded(cols.size());
    const int rps = agg_rows_per_step(s, cols, o);
    const std::string bi = emit_row_loop_open(E, s, cols, rps, sp, !reg, loaded, o.blocks);
    if (s.filter_col >= 0) E.os << bi << "if (!(fv " << fops[s.filter_op] << " (i64)" << s.filter_val << "LL)) continue;\n";
    if (reg) {
        for (int i = 0; i < NP; i++) {
            const Plan& p = s.plans[i];
            const int W = p.n[p.root].W;
            std::string rv = emit_plan_any(E, p, cols, loaded, bi, "p" + std::to_string(i) + "_", plan_fn_name(i));
            E.os << bi << (o.blocks ? "if (p" + std::to_string(i) + "_nn) " : std::string()) << "acc_add<AW1, " << W
                 << ">(acc" << i << ", stored<" << p.outB << ", " << W << ">(" << rv << "));\n";
        }
        E.os << bi << "cnt += " << (o.blocks ? iter_count_expr(s) : std::string("1u")) << ";\n";
    } else {
        E.os << bi << "const u32 g = a.lut[key];\n";
        E.os << bi << "const u32 sbase = (copy * NS + g) * SLOTW;\n";
        emit_agg_row(E, s, cols, loaded, bi);
        E.os << bi << "atomicAdd(&acnt[copy * NS + g], " << (o.blocks ? iter_count_expr(s) : std::string("1u")) << ");\n";
    }
    emit_row_loop_close(E, rps);
    E.os << "  __shared__ u32 red[NWAVE * PW];\n";
    if (!reg) {
        // combine the slot copies of the block: one thread per (group, plan)
        E.os << "  __syncthreads();\n";
        E.os << "  for (u32 k = tid; k < NP * NG + NG; k += T) {\n";
        E.os << "    u32* dst = a.partials + (usize)blockIdx.x * PW;\n";
        E.os << "    if (k < NP * NG) {\n      const u32 gg = k / NP, pp = k % NP;\n      Mag<AW2> sm = mzero<AW2>();\n";
        E.os << "      for (u32 cc = 0; cc < NCOPY; cc++) add_to<AW2>(sm, sm, slot_value(&aslot[(cc * NS + gg) * SLOTW + slot_off(pp)], slot_words(pp)));\n";
        E.os << "#pragma unroll\n      for (int w = 0; w < AW2; w++) dst[k * AW2 + w] = sm.w[w];\n";
        E.os << "    } else {\n      const u32 gg = k - NP * NG;\n      u64 cn = 0;\n";
        E.os << "      for (u32 cc = 0; cc < NCOPY; cc++) cn += acnt[cc * NS + gg];\n";
        E.os << "      dst[NP * NG * AW2 + 2 * gg] = (u32)cn; dst[NP * NG * AW2 + 2 * gg + 1] = (u32)(cn >> 32);\n    }\n  }\n}\n";
    } else {
        // block reduction: each wave reduces, wave leaders combine through LDS
        E.os << "#pragma unroll\n  for (int pp = 0; pp < NP; pp++) {\n";
        E.os << "    Mag<AW1> s1 = mzero<AW1>();\n";
        for (int i = 0; i < NP; i++) E.os << "    if (pp == " << i << ") s1 = acc" << i << ";\n";
        E.os << "    Mag<AW2> s2;\n    const u32 ext = (s1.w[AW1 - 1] >> 31) ? 0xffffffffu : 0u;\n";
        E.os << "#pragma unroll\n    for (int w = 0; w < AW2; w++) s2.w[w] = w < AW1 ? s1.w[w] : ext;\n";
        E.os << "    wave_reduce<AW2>(s2);\n";
        E.os << "    if (lane == 0) {\n#pragma unroll\n      for (int w = 0; w < AW2; w++) red[wave * PW + pp * AW2 + w] = s2.w[w];\n    }\n";
        E.os << "  }\n";
        E.os << "  {\n    u64 c = cnt;\n";
        E.os << "#pragma unroll\n    for (int off = 32; off >= 1; off >>= 1) c += (u64)__shfl_xor((long long)c, off, 64);\n";
        E.os << "    if (lane == 0) { red[wave * PW + NP * AW2] = (u32)c; red[wave * PW + NP * AW2 + 1] = (u32)(c >> 32); }\n";
        E.os << "  }\n  __syncthreads();\n";
        E.os << "  for (u32 k = tid; k < NP + 1; k += T) {\n";
        E.os << "    u32* dst = a.partials + (usize)blockIdx.x * PW;\n";
        E.os << "    if (k < NP) {\n      Mag<AW2> sm = mzero<AW2>();\n";
        E.os << "      for (int wv = 0; wv < NWAVE; wv++) {\n        Mag<AW2> o;\n#pragma unroll\n        for (int w = 0; w < AW2; w++) o.w[w] = red[wv * PW + k * AW2 + w];\n        add_to<AW2>(sm, sm, o);\n      }\n";
        E.os << "#pragma unroll\n      for (int w = 0; w < AW2; w++) dst[k * AW2 + w] = sm.w[w];\n";
        E.os << "    } else {\n      u64 cn = 0;\n";
        E.os << "      for (int wv = 0; wv < NWAVE; wv++) cn += (u64)red[wv * PW + NP * AW2] | ((u64)red[wv * PW + NP * AW2 + 1] << 32);\n";
        E.os << "      dst[NP * AW2] = (u32)cn; dst[NP * AW2 + 1] = (u32)(cn >> 32);\n    }\n  }\n}\n";
    }

    // final: one block per (group, plan) sum and one per group count
    E.os << "extern \"C\" __global__ void __launch_bounds__(256) " << name
         << "_final(const u32* __restrict__ partials, u32 nblocks, u8* __restrict__ images, u64* __restrict__ counts)\n{\n";
    E.os << "  const u32 k = blockIdx.x, tid = threadIdx.x;\n";
    E.os << "  __shared__ u32 red[4 * AW2 + 8];\n";
    E.os << "  if (k < NP * NG) {\n    const u32 pp = k % NP;\n    Mag<AW2> sm = mzero<AW2>();\n";
    E.os << "    for (u32 b = tid; b < nblocks; b += 256) {\n      Mag<AW2> o;\n#pragma unroll\n      for (int w = 0; w < AW2; w++) o.w[w] = partials[(usize)b * PW + k * AW2 + w];\n      add_to<AW2>(sm, sm, o);\n    }\n";
    E.os << block_reduce_aw2(4);
    emit_agg_final_image(E, s, "images + (usize)k * 133u");
    E.os << "    }\n  } else {\n    const u32 gg = k - NP * NG;\n    u64 c = 0;\n";
    E.os << "    for (u32 b = tid; b < nblocks; b += 256) c += (u64)partials[(usize)b * PW + NP * NG * AW2 + 2 * gg] | ((u64)partials[(usize)b * PW + NP * NG * AW2 + 2 * gg + 1] << 32);\n";
    E.os << kBlockReduceCount;
    E.os << "    if (tid == 0) { u64 t = 0; for (int wv = 0; wv < 4; wv++) t += (u64)red[2 * wv] | ((u64)red[2 * wv + 1] << 32); counts[gg] = t; }\n  }\n}\n";
    return E.os.str();
}

// ------------------------------------------------------------------ interpreter programs

namespace {

struct VmBuilder {
    const Plan& p;
    const std::vector<ColumnType>& cols;
    VmProgram& out;
    std::string err;
    bool used[updec_vm::kMaxRegs] = {};
    std::map<int, int> const_index;   // plan node -> constant slot
    std::vector<int> need_memo;

    VmBuilder(const Plan& pl, const std::vector<ColumnType>& c, VmProgram& o) : p(pl), cols(c), out(o) {}

    // registers needed to evaluate node `id` into a register (leaves: 0,
    // they are read in place)
    int need(int id)
    {
        if (need_memo[id] >= 0) return need_memo[id];
        const PNode& x = p.n[id];
        int r = 0;
        switch (x.kind) {
        case PK::Col:
        case PK::Const: r = 0; break;
        case PK::Neg:
        case PK::Cast:
        case PK::Trunc: r = std::max(1, need(x.a)); break;
        default: {
            int l = need(x.a), q = need(x.b);
            r = l == q ? std::max(1, l + (l > 0 ? 1 : 0)) : std::max(l, q);
            break;
        }
        }
        need_memo[id] = r;
        return r;
    }

    int alloc()
    {
        for (int r = 0; r < updec_vm::kMaxRegs; r++)
            if (!used[r]) { used[r] = true; out.nregs = std::max(out.nregs, r + 1); return r; }
        err = "expression needs more than 8 interpreter registers";
        return -1;
    }
    void release(const updec_vm::Operand& o)
    {
        if (o.kind == updec_vm::K_REG) used[o.idx] = false;
    }

    bool constant(int id, updec_vm::Operand* o)
    {
        auto it = const_index.find(id);
        int k;
        if (it != const_index.end()) {
            k = it->second;
        } else {
            k = (int)const_index.size();
            if (k >= 256) { err = "too many constants"; return false; }
            const_index[id] = k;
            const HostDec& d = p.n[id].k;
            out.consts.push_back(d.sign);
            for (int w = 0; w < out.W; w++) out.consts.push_back(w < HostDec::NW ? d.v[w] : 0u);
        }
        *o = {updec_vm::K_CONST, (uint8_t)k, 0};
        return true;
    }

    // a constant +0 (keyed apart from every plan node)
    bool zero_constant(updec_vm::Operand* o)
    {
        auto it = const_index.find(-1);
        int k;
        if (it != const_index.end()) {
            k = it->second;
        } else {
            k = (int)const_index.size();
            if (k >= 256) { err = "too many constants"; return false; }
            const_index[-1] = k;
            out.consts.push_back(0u);
            for (int w = 0; w < out.W; w++) out.consts.push_back(0u);
        }
        *o = {updec_vm::K_CONST, (uint8_t)k, 0};
        return true;
    }

    bool emit(updec_vm::Op op, int dst, int arg, updec_vm::Operand a, updec_vm::Operand b)
    {
        if ((int)out.ins.size() >= updec_vm::kMaxIns) { err = "interpreter program too long"; return false; }
        updec_vm::Ins I{};
        I.op = op;
        I.dst = (uint8_t)dst;
        I.arg = (uint16_t)arg;
        I.a = a;
        I.b = b;
        out.ins.push_back(I);
        return true;
    }

    // evaluate node `id`; the result is a register, a column or a constant
    bool gen(int id, updec_vm::Operand* res)
    {
        const PNode& x = p.n[id];
        if (x.kind == PK::Col) {
            if (x.col >= updec_vm::kMaxCols) { err = "too many columns for the interpreter"; return false; }
            *res = {updec_vm::K_COL, (uint8_t)x.col, 0};
            return true;
        }
        if (x.kind == PK::Const) return constant(id, res);
        // b stays a (never read) constant operand for the unary operations
        updec_vm::Operand a{}, b{updec_vm::K_CONST, 0, 0};
        const bool unary = x.kind == PK::Neg || x.kind == PK::Cast || x.kind == PK::Trunc;
        if (unary) {
            if (!gen(x.a, &a)) return false;
        } else if (need(x.b) > need(x.a)) {
            if (!gen(x.b, &b) || !gen(x.a, &a)) return false;
        } else {
            if (!gen(x.a, &a) || !gen(x.b, &b)) return false;
        }
        release(a);
        if (!unary) release(b);
        const int dst = alloc();
        if (dst < 0) return false;
        const PNode& A = p.n[x.a];
        bool ok = true;
        switch (x.kind) {
        case PK::Neg: ok = emit(updec_vm::OP_NEG, dst, 0, a, {}); break;
        case PK::Cast:
            if (x.S > A.S) ok = emit(updec_vm::OP_UP, dst, x.S - A.S, a, {});
            else if (x.S < A.S) ok = emit(updec_vm::OP_DOWN, dst, A.S - x.S, a, {});
            else ok = emit(updec_vm::OP_MOV, dst, 0, a, {});
            break;
        case PK::Add:
        case PK::Sub: {
            const PNode& B = p.n[x.b];
            a.shift = (uint16_t)(x.S - A.S);
            b.shift = (uint16_t)(x.S - B.S);
            ok = emit(x.kind == PK::Add ? updec_vm::OP_ADD : updec_vm::OP_SUB, dst, 0, a, b);
            break;
        }
        case PK::Mul: ok = emit(updec_vm::OP_MUL, dst, 0, a, b); break;
        case PK::Div: ok = emit(updec_vm::OP_DIV, dst, p.n[x.b].S + 4, a, b); break;
        case PK::Trunc: {
            const updec_vm::Operand self{updec_vm::K_REG, (uint8_t)dst, 0};
            if (x.tp > 0) {
                if (x.S > A.S) ok = emit(updec_vm::OP_UP, dst, x.S - A.S, a, {});
                else if (x.S < A.S) ok = emit(updec_vm::OP_DOWN, dst, A.S - x.S, a, {});
                else ok = emit(updec_vm::OP_MOV, dst, 0, a, {});
            } else if (A.P - A.S > -x.tp) {
                ok = emit(updec_vm::OP_DOWN, dst, A.S - x.tp, a, {}) && emit(updec_vm::OP_UP, dst, -x.tp, self, {});
            } else {
                // zero magnitude with the argument's sign: a times +0
                updec_vm::Operand z;
                ok = zero_constant(&z) && emit(updec_vm::OP_MUL, dst, 0, a, z);
            }
            break;
        }
        case PK::Mod: err = "MOD runs on the JIT path only (no interpreter instruction)"; return false;
        default: err = "internal: unexpected plan node"; return false;
        }
        *res = {updec_vm::K_REG, (uint8_t)dst, 0};
        return ok;
    }
};

}  // namespace

bool build_vm_program(const Plan& p, const std::vector<ColumnType>& cols, VmProgram* out, std::string* err)
{
    VmProgram prog;
    // Extended mode: the register width holds every node's exact bound, so
    // no value wraps and no intermediate (aligned operand, scaled dividend)
    // is cut.  Faithful mode: the reference's 32-limb values wrap, as the
    // JIT kernels' widths do (node bounds capped at 1024 bits); a division
    // whose scaled dividend could pass 1024 bits stays on the JIT path.
    const bool faithful = p.mode == Mode::Faithful;
    int words = 1;
    for (const PNode& x : p.n) {
        words = std::max(words, words_for(faithful ? x.bits : x.ubits));
        if (x.S > 65535) { *err = "scale out of the interpreter's range"; return false; }
        if (faithful && x.kind == PK::Div && p.n[x.a].bits + pow10_bits(p.n[x.b].S + 4) > 1024) {
            *err = "faithful division whose scaled dividend may pass 1024 bits (JIT path only)";
            return false;
        }
    }
    for (const PNode& x : p.n)
        if (x.kind == PK::Col) words = std::max(words, words_for(x.bits));
    prog.wclass = -1;
    for (int c = 0; c < updec_vm::kNumClasses; c++)
        if (updec_vm::kClassW[c] >= words) { prog.wclass = c; break; }
    if (prog.wclass < 0) {
        *err = "exact intermediates exceed the interpreter's 512-limb registers (" + std::to_string(words) +
               " limbs needed)";
        return false;
    }
    prog.W = updec_vm::kClassW[prog.wclass];
    prog.faithful = faithful;
    prog.outP = p.outP;
    prog.outS = p.outS;
    prog.outB = p.outB;
    VmBuilder b(p, cols, prog);
    b.need_memo.assign(p.n.size(), -1);
    updec_vm::Operand r{};
    if (!b.gen(p.root, &r)) { *err = b.err; return false; }
    if (r.kind != updec_vm::K_REG) {
        const int dst = b.alloc();
        if (dst < 0 || !b.emit(updec_vm::OP_MOV, dst, 0, r, {})) { *err = b.err; return false; }
        r = {updec_vm::K_REG, (uint8_t)dst, 0};
    }
    prog.root = r.idx;
    // forwarding: a result read only by the next instruction (or, for the
    // last instruction, only as the root) stays in VGPRs and is not stored
    auto& I = prog.ins;
    for (size_t i = 0; i < I.size(); i++) {
        const uint8_t d = I[i].dst;
        std::vector<std::pair<size_t, int>> reads;   // (instruction, operand 0/1)
        bool root_read = false;
        size_t k = i + 1;
        for (; k < I.size(); k++) {
            const bool binary = I[k].op >= updec_vm::OP_ADD && I[k].op <= updec_vm::OP_DIV;
            if (I[k].a.kind == updec_vm::K_REG && I[k].a.idx == d) reads.push_back({k, 0});
            if (binary && I[k].b.kind == updec_vm::K_REG && I[k].b.idx == d) reads.push_back({k, 1});
            if (I[k].dst == d) break;   // overwritten
        }
        if (k == I.size() && prog.root == d) root_read = true;
        bool only_next = true;
        for (auto& rd : reads) only_next = only_next && rd.first == i + 1;
        if (root_read && i + 1 != I.size()) only_next = false;
        if (!only_next || (reads.empty() && !root_read)) continue;
        for (auto& rd : reads) {
            updec_vm::Operand& o = rd.second == 0 ? I[rd.first].a : I[rd.first].b;
            o.kind = updec_vm::K_LAST;
            o.idx = 0;
        }
        if (root_read) prog.root = updec_vm::kNoStore;
        I[i].dst = updec_vm::kNoStore;
    }
    if (prog.nregs > updec_vm::kClassRegs[prog.wclass]) {
        *err = "expression needs more interpreter registers than its width class holds";
        return false;
    }
    *out = std::move(prog);
    return true;
}

// Multi-lane calc kernel (the reference's XMP form of calc_expr_<id>,
// AriesExprCalcNode.cpp:1320-1400, re-designed for wave64): the register
// program of an extended-mode plan whose exact intermediates need 64 or
// more limbs, emitted straight-line with every shift count, scale and row
// width as a compile-time constant.  A group of W / L lanes evaluates one
// row, L limbs per lane (dec_mlane.h); the groups of a block take
// consecutive rows.  Each column the program reads is loaded once per row;
// constants ({sign, W limbs} each, in a.kc) are loaded once per thread.
std::string gen_calc_ml(const VmProgram& vp, const std::vector<ColumnType>& cols, int L, const std::string& name)
{
    using namespace updec_vm;
    const int W = vp.W, TPI = W / L;
    std::ostringstream os;
    os << "#include \"dec_mlane.h\"\nusing namespace updec;\n";
    os << "struct MlArgs { const u8* col[" << kMaxCols << "]; usize rows; u8* out; const u32* kc; };\n";
    os << "// program (" << vp.ins.size() << " instructions, " << W << " limbs, " << TPI << " lanes of " << L
       << " limbs per row)\n";
    {
        std::istringstream lines(vp.text());
        std::string ln;
        while (std::getline(lines, ln)) os << "//   " << ln << "\n";
    }
    os << "#define L_ " << L << "\n#define TPI_ " << TPI << "\n#define NG_ (256 / TPI_)\n";
    os << "extern \"C\" __global__ void __launch_bounds__(256) " << name << "(MlArgs a)\n{\n";
    os << "  const ml::Grp<TPI_> g;\n";
    std::vector<bool> used_col(cols.size(), false);
    int nconst = 0;
    auto note = [&](const Operand& o) {
        if (o.kind == K_COL) used_col[o.idx] = true;
        if (o.kind == K_CONST) nconst = std::max(nconst, o.idx + 1);
    };
    for (const Ins& I : vp.ins) {
        note(I.a);
        if (I.op >= OP_ADD && I.op <= OP_DIV) note(I.b);
    }
    for (int k = 0; k < nconst; k++) {
        os << "  ml::MDec<L_> k" << k << ";\n  k" << k << ".neg = a.kc[" << k * (W + 1) << "];\n";
        os << "#pragma unroll\n  for (int w = 0; w < L_; w++) k" << k << ".m.w[w] = a.kc[" << k * (W + 1)
           << " + 1 + g.rank * L_ + w];\n";
    }
    os << "  const usize step = (usize)gridDim.x * NG_;\n";
    os << "  for (usize row = (usize)blockIdx.x * NG_ + threadIdx.x / TPI_; row < a.rows; row += step) {\n";
    for (size_t c = 0; c < cols.size(); c++)
        if (used_col[c])
            os << "    const ml::MDec<L_> c" << c << " = ml::load_compact_b<L_, TPI_, " << compact_bytes(cols[c].P)
               << ">(a.col[" << c << "], row, g);\n";
    os << "    ml::MDec<L_> last";
    for (int r = 0; r < vp.nregs; r++) os << ", r" << r;
    os << ";\n";
    auto fetch = [&](const Operand& o, const char* v) {
        std::string src = o.kind == K_LAST ? "last"
                          : (o.kind == K_REG ? "r" : o.kind == K_COL ? "c" : "k") + std::to_string(o.idx);
        os << "      ml::MDec<L_> " << v << " = " << src << ";\n";
        if (o.shift) os << "      ml::mul_p10<L_, TPI_>(" << v << ".m, " << o.shift << "u, g);\n";
    };
    for (const Ins& I : vp.ins) {
        os << "    {\n";
        fetch(I.a, "A");
        if (I.op >= OP_ADD && I.op <= OP_DIV) fetch(I.b, "B");
        os << "      ml::MDec<L_> R;\n";
        switch (I.op) {
        case OP_MOV: os << "      R = A;\n"; break;
        case OP_NEG: os << "      R = A;\n      R.neg ^= 1u;\n"; break;
        case OP_UP: os << "      R = A;\n      ml::mul_p10<L_, TPI_>(R.m, " << I.arg << "u, g);\n"; break;
        case OP_DOWN: os << "      R = A;\n      ml::div_p10<L_, TPI_>(R.m, " << I.arg << "u, g);\n"; break;
        case OP_ADD: os << "      R = ml::add<L_, TPI_>(A, B, g);\n"; break;
        case OP_SUB: os << "      R = ml::sub<L_, TPI_>(A, B, g);\n"; break;
        case OP_MUL: os << "      R.m = ml::mul<L_, TPI_>(A.m, B.m, g);\n      R.neg = A.neg ^ B.neg;\n"; break;
        default: os << "      R = ml::div<L_, TPI_>(A, B, " << I.arg << "u, g);\n"; break;
        }
        if (I.dst != kNoStore) os << "      r" << (int)I.dst << " = R;\n";
        os << "      last = R;\n    }\n";
    }
    os << "    ml::store_compact_b<L_, TPI_, " << vp.outB << ">(a.out, row, "
       << (vp.root == kNoStore ? std::string("last") : "r" + std::to_string(vp.root)) << ", g);\n";
    os << "  }\n}\n";
    return os.str();
}

std::string VmProgram::text() const
{
    static const char* names[] = {"MOV", "NEG", "ADD", "SUB", "MUL", "DIV", "UP", "DOWN"};
    auto opnd = [](const updec_vm::Operand& o) {
        std::string s = o.kind == updec_vm::K_LAST ? std::string("acc")
                        : (o.kind == updec_vm::K_REG ? "r" : o.kind == updec_vm::K_COL ? "c" : "k") + std::to_string(o.idx);
        if (o.shift) s += "<<" + std::to_string(o.shift);
        return s;
    };
    std::ostringstream os;
    os << "W" << W << " regs " << nregs << " root " << (root == updec_vm::kNoStore ? std::string("acc") : "r" + std::to_string(root))
       << "\n";
    for (const auto& I : ins) {
        os << names[I.op] << " " << (I.dst == updec_vm::kNoStore ? std::string("-") : "r" + std::to_string(I.dst)) << " "
           << opnd(I.a);
        if (I.op >= updec_vm::OP_ADD && I.op <= updec_vm::OP_DIV) os << " " << opnd(I.b);
        if (I.op == updec_vm::OP_DIV || I.op == updec_vm::OP_UP || I.op == updec_vm::OP_DOWN) os << " #" << I.arg;
        os << "\n";
    }
    return os.str();
}

}  // namespace updec
