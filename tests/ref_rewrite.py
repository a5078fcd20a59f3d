"""Independent restatement of the reference's expression rewrite and result
typing, for checking the product planner's plan text (tests only).

It plans from the expression text, as the product does, and renders the
result in the plan-text notation the oracle reads (`c<col>:<P>:<S>`,
`k<sign><hex>:<S>`, `(NEG x)`, `(CAST P S x)`, `(<OP> P S a b)`).  Sources:

  StringToDecimal              decimal.cu:2980-3025   literal value; prec = digit count, frac
  CheckAndSetRealPrecision     decimal.cu:2060-2088   frontend literal precision (double arithmetic)
  Decimal(int32) / (int64)     decimal.cu:281-309     prec 10 / 19
  Decimal(int32, scale, ...)   decimal.cu:229-260
  Decimal(Decimal, scale, ...) decimal.cu:263-268
  AlignAddSubData              decimal.cu:2905-2920
  AddDecimalOnCpu              decimal.cu:1284-1323
  MulDecimalOnCpu              decimal.cu:1325-1351
  operator!=(Decimal, int32)   decimal.cu:998-1001, operator== :728-743
  Calc*TargetPrecision         decimal.cu:1353-1358, 1568-1572, 1783-1786, 1997-2000, 2610-2613
  frontend CALC typing         CommonBiaodashi.cpp:3338-3377 (Decimal prec/frac are uint8)
  literal bridging             AriesExprBridge.cpp:357-396 (DECIMAL leaf type = StringToDecimal's prec/frac)
  ModifySubOperation           AriesEngineShell.cpp:244-281
  getChild / BinToMultiTree    AriesEngineShell.cpp:284-342
  CmpByScale + std::sort       AriesEngineShell.cpp:344-358, libstdc++ introsort / insertion sort
  AlignOptMultiTree            AriesEngineShell.cpp:360-597
  RemoveZeroNode               AriesEngineShell.cpp:184-219
  MultiToBinTree/ListToBinTree AriesEngineShell.cpp:90-182 (uint16 type arithmetic)
  AlignOptimizeNode/-Optimize  AriesEngineShell.cpp:628-685 (root precision capped at 288)

Grammar conventions the reference's SQL front end does not fix, shared with
the product: a unary minus on a literal negates the literal itself; an
integer literal is INT when it fits 32 bits, else LONG_INT; cast(e, P, S)
is a function node whose argument is rewritten as a tree of its own (the
reference's MultiToBinTree rebuilds function nodes without their
arguments, which would drop it).  Extended mode (the product's rule outside
the reference's envelope) computes the ListToBinTree rules in plain
integers with the precision kept >= 1, applies no 288 cap and no uint8 wrap
of frontend types, and does not reduce constants mod 2^1024.
"""
import math

FAITHFUL, EXTENDED = 1, 2
MOD1024 = 1 << 1024
M32 = 0xFFFFFFFF


class Undefined(Exception):
    """the reference leaves this case undefined (or the product rejects it
    by design)"""


# ------------------------------------------------------------------ host Decimal

class HD:
    __slots__ = ("sign", "prec", "frac", "mag")

    def __init__(self, mag=0, sign=0, prec=0, frac=0):
        self.sign, self.prec, self.frac, self.mag = sign, prec, frac, mag

    def copy(self):
        return HD(self.mag, self.sign, self.prec, self.frac)


def _wrap(m, mode):
    return m % MOD1024 if mode == FAITHFUL else m


def string_to_decimal(text):
    """StringToDecimal over a literal without sign: the digits accumulate
    into the magnitude, prec = characters minus the '.', frac = digits after it"""
    dot = text.find(".")
    digits = text.replace(".", "")
    d = HD(int(digits) % MOD1024 if digits else 0)
    d.prec = len(text) - (1 if dot >= 0 else 0)
    d.frac = len(text) - dot - 1 if dot >= 0 else 0
    return d


def _cvttsd2si32(x):
    """(int)double on x86-64: truncation, 0x80000000 when out of range"""
    if x != x or x >= 2 ** 31 or x < -(2 ** 31):
        return -(2 ** 31)
    return int(x)


def real_precision(d):
    """CheckAndSetRealPrecision: returns (prec, frac)"""
    limbs = [(d.mag >> (32 * k)) & M32 for k in range(32)]
    i = 31
    while i >= 0 and limbs[i] == 0:
        i -= 1
    if i == -1:
        return 1, 0
    prec = i * 9
    t = 1.0
    for _ in range(i):
        t *= 4.294967295
    t *= limbs[i]
    tt = _cvttsd2si32(t)
    while tt > 0:
        tt //= 10
        prec += 1
    return prec, d.frac


def dec_int(t, prec):
    """Decimal(int32 t) (prec 10) or Decimal(int64 t) (prec 19)"""
    return HD(abs(t), 1 if t < 0 else 0, prec, 0)


def dec_int_scale(t, scale, mode):
    """Decimal(int32 t, int32 scale, bool, bool, bool)"""
    prec = 10 + (scale if scale > 10 else 0)
    return HD(_wrap(abs(t) * 10 ** scale, mode), 1 if t < 0 else 0, prec, scale)


def shl10(d, n, mode):
    if n <= 0:
        return d
    if mode == FAITHFUL and n > 38:
        raise Undefined("operator<< past __POW10_ARRAY")
    d.mag = _wrap(d.mag * 10 ** n, mode)
    return d


def dec_rescaled(t, scale, mode):
    """Decimal(Decimal t, int32 scale, bool x4): t << (scale - frac)"""
    gap = scale - t.frac
    r = shl10(t.copy(), gap, mode)
    r.frac = scale
    r.prec = t.prec + gap
    return r


def align_add_sub(x, d, mode):
    """x.AlignAddSubData(d)"""
    if x.frac == d.frac:
        return
    if x.frac < d.frac:
        shl10(x, d.frac - x.frac, mode)
        x.frac = d.frac
    else:
        shl10(d, x.frac - d.frac, mode)
        d.frac = x.frac


def add_on_cpu(x, d, mode):
    added = d.copy()
    align_add_sub(added, x, mode)
    if x.sign == added.sign:
        x.mag = _wrap(x.mag + added.mag, mode)
    else:
        r = (x.mag > added.mag) - (x.mag < added.mag)
        x.mag = x.mag - added.mag if r >= 0 else added.mag - x.mag
        x.sign = 1 if (r > 0 and not d.sign) or (r < 0 and d.sign) else 0


def mul_on_cpu(x, d, mode):
    x.sign ^= d.sign
    x.frac += d.frac
    x.mag = _wrap(x.mag * d.mag, mode)


def equals_int(x, i, mode):
    """x == Decimal(int32 i): signs, then limbs after aligning scales"""
    d = dec_int(i, 10)
    if x.sign != d.sign:
        return False
    a, b = x.copy(), d.copy()
    align_add_sub(a, b, mode)
    return a.mag == b.mag


# ------------------------------------------------------------------ expression tree

class N:
    """One node: kind in {"col", "int", "dec", "calc", "cast"}; op in
    {"+", "-", "*", "/", "%"}; P, S its value type."""
    __slots__ = ("kind", "op", "col", "ival", "dval", "P", "S", "reverse", "ch", "decimal_calc", "fn")

    def __init__(self, kind, **kw):
        self.kind = kind
        self.op = kw.get("op")
        self.col = kw.get("col", -1)
        self.ival = kw.get("ival", 0)
        self.dval = kw.get("dval")
        self.P = kw.get("P", 0)
        self.S = kw.get("S", 0)
        self.reverse = False
        self.ch = kw.get("ch", [])
        self.decimal_calc = False
        self.fn = None

    def clone(self):
        n = N(self.kind, op=self.op, col=self.col, ival=self.ival, dval=self.dval.copy() if self.dval else None,
              P=self.P, S=self.S, ch=[c.clone() for c in self.ch])
        n.reverse, n.decimal_calc = self.reverse, self.decimal_calc
        return n


class Parser:
    def __init__(self, text, cols):
        self.s, self.i = text, 0
        self.names = [c[0] for c in cols]
        self.types = [(c[1], c[2]) for c in cols]

    def peek(self):
        while self.i < len(self.s) and self.s[self.i].isspace():
            self.i += 1
        return self.s[self.i] if self.i < len(self.s) else ""

    def eat(self, c):
        if self.peek() == c:
            self.i += 1
            return True
        return False

    def expr(self):
        left = self.term()
        while self.peek() in ("+", "-"):
            op = self.s[self.i]
            self.i += 1
            left = N("calc", op=op, ch=[left, self.term()])
        return left

    def term(self):
        left = self.unary()
        while self.peek() in ("*", "/", "%"):
            op = self.s[self.i]
            self.i += 1
            left = N("calc", op=op, ch=[left, self.unary()])
        return left

    def unary(self):
        if self.eat("-"):
            e = self.unary()
            if e.kind == "int":
                e.ival = -e.ival
            elif e.kind == "dec":
                e.dval.sign ^= 1
            else:
                e.reverse = not e.reverse
            return e
        return self.primary()

    def primary(self):
        c = self.peek()
        if c == "(":
            self.i += 1
            e = self.expr()
            if not self.eat(")"):
                raise ValueError("expected )")
            return e
        if c.isdigit() or c == ".":
            st = self.i
            while self.i < len(self.s) and (self.s[self.i].isdigit() or self.s[self.i] == "."):
                self.i += 1
            lit = self.s[st:self.i]
            if "." in lit:
                return N("dec", dval=string_to_decimal(lit))
            v = int(lit)
            if v < 2 ** 63:
                return N("int", ival=v)
            return N("dec", dval=string_to_decimal(str(v)))
        st = self.i
        while self.i < len(self.s) and (self.s[self.i].isalnum() or self.s[self.i] == "_"):
            self.i += 1
        ident = self.s[st:self.i]
        if ident == "cast" and self.eat("("):
            arg = self.expr()
            self.eat(",")
            P = self._num()
            self.eat(",")
            S = self._num()
            self.eat(")")
            return N("cast", P=P, S=S, ch=[arg])
        k = self.names.index(ident)
        return N("col", col=k, P=self.types[k][0], S=self.types[k][1])

    def _num(self):
        self.peek()
        st = self.i
        while self.i < len(self.s) and self.s[self.i].isdigit():
            self.i += 1
        return int(self.s[st:self.i])


# ---- frontend result types (CommonBiaodashi.cpp:3338-3377)

def frontend_types(e, mode):
    """sets P, S on every node as the frontend types it; returns whether the
    node is decimal-typed (a DIV promotes to DECIMAL; ints stay INT)"""
    if e.kind == "col":
        return True
    if e.kind == "int":
        p, _ = real_precision(dec_int(e.ival, 10 if -2 ** 31 <= e.ival < 2 ** 31 else 19))
        e.P, e.S = p, 0
        return False
    if e.kind == "dec":
        # AriesExprBridge: the leaf's type is StringToDecimal's prec / frac
        e.P, e.S = e.dval.prec, e.dval.frac
        return True
    if e.kind == "cast":
        frontend_types(e.ch[0], mode)
        return True
    da = frontend_types(e.ch[0], mode)
    db = frontend_types(e.ch[1], mode)
    e.decimal_calc = da or db or e.op == "/"

    def fe(n):
        """GetDecimalValue: literals by CheckAndSetRealPrecision, others by type"""
        if n.kind == "dec":
            return real_precision(n.dval)
        return n.P, n.S

    (pa, sa), (pb, sb) = fe(e.ch[0]), fe(e.ch[1])
    if e.op in "+-":
        P = max(pa - sa, pb - sb) + 1
        S = max(sa, sb)
        P += S
    elif e.op == "*":
        S, P = sa + sb, pa + pb
    elif e.op == "/":
        P, S = pa - pb + sb + 4 + 1, sa + 4
    else:
        P, S = pb, 0
    if mode == FAITHFUL:
        P, S = P & 0xFF, S & 0xFF
    else:
        P = max(P, 1)
    e.P, e.S = P, S
    return e.decimal_calc


# ---- AlignOptimize

def modify_sub(t):
    """ModifySubOperation: a - b -> a + (-b), pending negations pushed down"""
    q = [t]
    while q:
        n = q.pop(0)
        if n.kind == "calc":
            if n.op == "*" and n.reverse:
                n.ch[0].reverse = not n.ch[0].reverse
            elif n.op == "/" and n.reverse:
                n.ch[0].reverse = not n.ch[0].reverse
            elif n.op == "+" and n.reverse:
                n.ch[0].reverse = not n.ch[0].reverse
                n.ch[1].reverse = not n.ch[1].reverse
            elif n.op == "-":
                if not n.reverse:
                    n.ch[1].reverse = not n.ch[1].reverse
                else:
                    n.ch[0].reverse = not n.ch[0].reverse
                n.op = "+"
            if len(n.ch) == 2:
                q.append(n.ch[0])
                q.append(n.ch[1])


class MT:
    """MultipleTree node: a copy of a binary node's fields plus its flat
    child list; `root` is the binary node it was made from"""
    __slots__ = ("kind", "op", "col", "ival", "dval", "P", "S", "reverse", "ch", "root", "decimal_calc")

    def __init__(self, b=None):
        self.ch = []
        self.root = b
        if b is not None:
            self.kind, self.op, self.col, self.ival = b.kind, b.op, b.col, b.ival
            self.dval, self.P, self.S, self.reverse = b.dval, b.P, b.S, b.reverse
            self.decimal_calc = b.decimal_calc
        else:
            self.kind, self.op, self.col, self.ival, self.dval = "dec", None, -1, 0, None
            self.P = self.S = 0
            self.reverse = False
            self.decimal_calc = False


def get_child(b, sym):
    if sym == "/":
        return [MT(b.ch[0]), MT(b.ch[1])]
    out = []
    for c in b.ch:
        if c.kind != "calc":
            out.append(MT(c))
        elif c.op == sym:
            out.extend(get_child(c, sym))
        else:
            out.append(MT(c))
    return out


def bin_to_multi(b, m):
    m.ch = get_child(b, m.root.op)
    for c in m.ch:
        if c.root is not None and c.root.kind == "calc" and len(c.root.ch) == 2:
            bin_to_multi(c.root, c)


def cmp_by_scale(a, b):
    if a.kind == "int":
        return True
    if b.kind == "int":
        return False
    if a.kind == "dec" and b.kind == "dec":
        return a.S < b.S
    if a.kind == "dec":
        return True
    if b.kind == "dec":
        return False
    return a.S < b.S


def std_sort(v, comp):
    """libstdc++ std::sort: introsort to runs of <= 16, then insertion sort.
    The comparator is not a strict weak order (comp(int, int) is true), so
    the exact algorithm fixes the order; an unguarded scan that would leave
    the range is Undefined."""

    def unguarded_linear_insert(last):
        val = v[last]
        nxt = last - 1
        while comp(val, v[nxt]):
            v[last] = v[nxt]
            last = nxt
            nxt -= 1
            if nxt < 0:
                raise Undefined("unguarded insertion leaves the range")
        v[last] = val

    def insertion_sort(first, last):
        if first == last:
            return
        for i in range(first + 1, last):
            if comp(v[i], v[first]):
                val = v[i]
                v[first + 1:i + 1] = v[first:i]
                v[first] = val
            else:
                unguarded_linear_insert(i)

    def move_median_to_first(result, a, b, c):
        if comp(v[a], v[b]):
            if comp(v[b], v[c]):
                x = b
            elif comp(v[a], v[c]):
                x = c
            else:
                x = a
        elif comp(v[a], v[c]):
            x = a
        elif comp(v[b], v[c]):
            x = c
        else:
            x = b
        v[result], v[x] = v[x], v[result]

    def unguarded_partition(first, last, pivot):
        while True:
            while comp(v[first], v[pivot]):
                first += 1
                if first >= len(v):
                    raise Undefined("partition leaves the range")
            last -= 1
            while comp(v[pivot], v[last]):
                last -= 1
                if last < 0:
                    raise Undefined("partition leaves the range")
            if not first < last:
                return first
            v[first], v[last] = v[last], v[first]
            first += 1

    def introsort_loop(first, last, depth):
        while last - first > 16:
            if depth == 0:
                raise Undefined("introsort falls back to heap sort")
            depth -= 1
            mid = first + (last - first) // 2
            move_median_to_first(first, first + 1, mid, last - 1)
            cut = unguarded_partition(first + 1, last, first)
            introsort_loop(cut, last, depth)
            last = cut

    n = len(v)
    if n <= 1:
        return
    introsort_loop(0, n, 2 * int(math.log2(n)))
    if n > 16:
        insertion_sort(0, 16)
        for i in range(16, n):
            unguarded_linear_insert(i)
    else:
        insertion_sort(0, n)


def _int_to_dec_node(c):
    """DIV / MOD operands: an integer becomes Decimal(int32 or int64)"""
    d = dec_int(c.ival, 10 if -2 ** 31 <= c.ival < 2 ** 31 else 19)
    n = MT()
    n.kind, n.dval, n.P, n.S = "dec", d, d.prec, d.frac
    return n


def _c_int(x):
    """C int arithmetic (32-bit wrap)"""
    x &= 0xFFFFFFFF
    return x - (1 << 32) if x >= 1 << 31 else x


def align_opt_multi(top, mode):
    """AlignOptMultiTree: breadth-first over decimal CALC nodes"""
    q = [top]
    while q:
        node = q[0]
        if node.kind == "calc" and node.decimal_calc:
            if node.op == "+":
                std_sort(node.ch, cmp_by_scale)
                node_int = 0
                node_dec = dec_int(0, 10)
                index = 0
                while index < len(node.ch):
                    c = node.ch[index]
                    if c.kind == "int":
                        t = -c.ival if c.reverse else c.ival
                        node_int = _c_int(node_int + t)
                    elif c.kind == "dec":
                        d = c.dval.copy()
                        if c.reverse:
                            d.sign = 1
                        add_on_cpu(node_dec, d, mode)
                    else:
                        break
                    index += 1
                if index == 0:
                    for c in node.ch:
                        if c.kind == "calc":
                            q.append(c)
                nl = []
                if node_int == 0 and node_dec.mag == 0:
                    if index == len(node.ch):
                        raise Undefined("an addition of constants only")
                    if index == len(node.ch) - 1:
                        c = node.ch[index]        # CopyNodeInfo: the operand replaces the node
                        node.kind, node.op, node.col, node.ival = c.kind, c.op, c.col, c.ival
                        node.dval, node.P, node.S, node.reverse = c.dval, c.P, c.S, c.reverse
                        node.ch, node.root, node.decimal_calc = c.ch, c.root, c.decimal_calc
                    else:
                        while index < len(node.ch):
                            nl.append(node.ch[index])
                            if node.ch[index].kind == "calc":
                                q.append(node.ch[index])
                            index += 1
                        node.ch = nl
                else:
                    if node_dec.mag == 0:
                        if index >= len(node.ch):
                            raise Undefined("an addition of constants only")
                        near = node.ch[index].S
                        n = MT()
                        n.kind, n.S, n.P = "dec", near, 0
                        n.dval = dec_int_scale(node_int, near, mode)
                        nl.append(n)
                        while index < len(node.ch):
                            nl.append(node.ch[index])
                            if node.ch[index].kind == "calc":
                                q.append(node.ch[index])
                            index += 1
                    else:
                        insert = True
                        if node_int != 0:
                            add_on_cpu(node_dec, dec_int(node_int, 10), mode)
                        while index < len(node.ch):
                            if node.ch[index].S > node_dec.frac and insert:
                                insert = False
                                d = dec_rescaled(node_dec, node.ch[index].S, mode)
                                n = MT()
                                n.kind, n.S, n.P, n.dval = "dec", d.frac, 0, d
                                nl.append(n)
                            nl.append(node.ch[index])
                            if node.ch[index].kind == "calc":
                                q.append(node.ch[index])
                            index += 1
                        if insert:
                            n = MT()
                            n.kind, n.P, n.S, n.dval = "dec", node_dec.frac, node_dec.frac, node_dec
                            nl.append(n)
                    node.ch = nl
            elif node.op == "*":
                node_int = 1
                node_dec = dec_int(1, 10)
                nl = []
                for c in node.ch:
                    if c.kind == "int":
                        node_int = _c_int(node_int * c.ival)
                    elif c.kind == "dec":
                        mul_on_cpu(node_dec, c.dval, mode)
                    else:
                        nl.append(c)
                if node_int != 1 or not equals_int(node_dec, 1, mode):
                    if node_int == 0 or node_dec.mag == 0:
                        node.kind, node.ival = "int", 0
                    else:
                        mul_on_cpu(node_dec, dec_int(node_int, 10), mode)
                        if not equals_int(node_dec, 1, mode):
                            n = MT()
                            n.kind, n.S, n.P, n.dval = "dec", node_dec.frac, 0, node_dec
                            nl.append(n)
                        node.ch = nl
                for c in node.ch:
                    if c.kind == "calc":
                        q.append(c)
            elif node.op in "/%":
                if node.ch[0].kind == "int" and node.ch[0].ival == 0:
                    node.kind, node.ival = "int", 0
                else:
                    for i, c in enumerate(node.ch):
                        if c.kind == "calc":
                            q.append(c)
                        if c.kind == "int":
                            node.ch[i] = _int_to_dec_node(c)
            else:
                for c in node.ch:
                    if c.kind == "calc":
                        q.append(c)
        q.pop(0)


def remove_zero(m):
    nl = []
    for i, c in enumerate(m.ch):
        if c.kind == "int":
            if m.op == "*":
                m.kind, m.ival = "int", 0
                return m
        elif c.kind == "calc":
            m.ch[i] = remove_zero(c)
            if m.ch[i].kind == "int":
                if m.op == "*":
                    m.kind, m.ival = "int", 0
                    return m
            else:
                nl.append(m.ch[i])
        else:
            nl.append(c)
    m.ch = nl
    return m


class BT:
    """rewritten binary node"""
    __slots__ = ("kind", "op", "col", "dval", "P", "S", "reverse", "ch", "fn", "ival")

    def __init__(self, kind, op=None, P=0, S=0):
        self.kind, self.op, self.P, self.S = kind, op, P, S
        self.col, self.dval, self.reverse, self.ch, self.fn, self.ival = -1, None, False, [], None, 0


def leaf(m):
    b = BT(m.kind, m.op, m.P, m.S)
    b.col, b.dval, b.reverse, b.ival = m.col, m.dval, m.reverse, m.ival
    if m.kind == "cast":
        b.fn = m.root.fn if m.root is not None else None
    return b


def _u16(v, mode):
    return v & 0xFFFF if mode == FAITHFUL else v


def list_to_bin(L, index, op, mode):
    if index == 0:
        return L[0]
    left = list_to_bin(L, index - 1, op, mode)
    r = L[index]
    p0, s0, p1, s1 = left.P, left.S, r.P, r.S
    if op == "+":
        t = BT("calc", "+")
        if r.reverse:
            r.reverse = False
            t.op = "-"
        P = _u16(max(p0 - s0, p1 - s1) + 1, mode)
        S = max(s0, s1)
        t.P, t.S = _u16(P + S, mode), S
    elif op == "*":
        t = BT("calc", "*")
        P = _u16((p0 - s0) + (p1 - s1), mode)
        S = _u16(s0 + s1, mode)
        t.P, t.S = _u16(P + S, mode), S
    elif op == "/":
        t = BT("calc", "/")
        t.P, t.S = _u16(p0 - p1 + s1 + 4 + 1, mode), _u16(s0 + 4, mode)
    else:
        t = BT("calc", "%", p1, 0)
    if mode == EXTENDED:
        t.P = max(t.P, 1)
    t.ch = [left, r]
    return t


def multi_to_bin(m, mode):
    tmp = [leaf(c) if c.kind != "calc" else multi_to_bin(c, mode) for c in m.ch]
    if not m.ch:
        return leaf(m)
    if m.kind == "calc" and len(m.ch) == 1:
        if m.ch[0].kind == "calc":
            return list_to_bin(tmp, len(tmp) - 1, m.op, mode)
        return leaf(m.ch[0])
    return list_to_bin(tmp, len(tmp) - 1, m.op, mode)


def align_optimize(e, mode):
    """AlignOptimize of one tree: a CALC root is rewritten; a cast's argument
    is rewritten on its own.  Returns a BT."""
    if e.kind == "cast":
        b = BT("cast", P=e.P, S=e.S)
        b.reverse = e.reverse
        b.fn = align_optimize(e.ch[0], mode)
        return b
    if e.kind != "calc":
        m = MT(e)
        return leaf(m)
    if not e.decimal_calc:
        raise Undefined("integer-typed calculation")

    def prep(n):
        """function nodes below are leaves whose argument is rewritten first"""
        if n.kind == "cast":
            n.fn = align_optimize(n.ch[0], mode)
            return
        for c in n.ch:
            prep(c)

    prep(e)
    modify_sub(e)
    m = MT(e)
    bin_to_multi(e, m)
    align_opt_multi(m, mode)
    m = remove_zero(m)
    if m.kind == "int":
        raise Undefined("the expression rewrites to an integer")
    out = multi_to_bin(m, mode)
    if mode == FAITHFUL and out.P > 288:
        out.P = 288
    return out


# ------------------------------------------------------------------ rendering

def render(b, cols):
    if b.kind == "col":
        t = "c%d:%d:%d" % (b.col, cols[b.col][1], cols[b.col][2])
    elif b.kind == "dec":
        t = "k%s%x:%d" % ("-" if b.dval.sign else "+", b.dval.mag, b.S)
    elif b.kind == "int":
        raise Undefined("integer operand left after the rewrite")
    elif b.kind == "cast":
        t = "(CAST %d %d %s)" % (b.P, b.S, render(b.fn, cols))
    else:
        if len(b.ch) != 2:
            raise Undefined("a calculation node left without operands")
        op = {"+": "ADD", "-": "SUB", "*": "MUL", "/": "DIV", "%": "MOD"}[b.op]
        t = "(%s %d %d %s %s)" % (op, b.P, b.S, render(b.ch[0], cols), render(b.ch[1], cols))
    if b.reverse:
        t = "(NEG %s)" % t
    return t


def plan(text, cols, mode):
    """plan text, result P and S for `text` over cols [(name, P, S), ...]"""
    e = Parser(text, cols).expr()
    frontend_types(e, mode)
    b = align_optimize(e, mode)
    if b.kind == "int":
        raise Undefined("integer result")
    return render(b, cols), b.P, b.S
