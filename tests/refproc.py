"""Literal restatement of the reference's device division procedures, step
for step, over the 32-limb Decimal (tests only).

The C oracle (oracle/updec_oracle.c: od_div, od_mod) computes division in
closed form: an exact quotient plus predicates on bit lengths that select
the reference's path behaviours.  This module instead walks the reference's
procedures as written, so that a misreading of either is caught by
comparing the two:

  DivOrMod        decimal.cu:2201-2333   path selection and the chunk loop
  DivByInt64      decimal.cu:2431-2446   uint64 numerator and round-up test
  DivByInt        decimal.cu:2335-2429   one-word divisor, uint32 round-up
  DivInt          decimal.cu:2125-2191   bisection for the quotient
  HalfIntDecimal  decimal.cu:2093-2122   midpoint of the bisection
  ModCalc         decimal.cu:2621-2680   with ModByInt64 :2682-2693, ModByInt :2695-2714
  operator<<      decimal.cu:2028-2058   x 10^n over the __POW10_ARRAY rows
  operator+= -=   decimal.cu:1362-1396, 1575-1613 (device sign rules)
  compare         decimal.cu:638-745

Machine behaviour that the C++ source leaves to the hardware is modelled as
the reference's target executes it and is named where it is used:
  - `x <<= n` on a uint32 with a negative int n compiles to PTX shl.b32,
    which reads n as unsigned and clamps counts >= 32 to 32: the result is 0;
  - DivInt's `qmax.v[-1]` store (pmax in [-63, -32]) lands in the four bytes
    before v[], which are qmax's own sign, prec, frac and error bytes
    (struct Decimal, decimal.hxx:130-141, packed).
Cases with no defined result raise `Undefined` and are left out of every
comparison: a shift past the 38 rows of __POW10_ARRAY, a store outside the
qmax object (pmax < -63), DivByInt's limb scan of a zero remainder, and a
DivOrMod chunk with a non-positive digit count.

One deliberate choice, the parity contract's: `qmid * dt` inside DivInt is
the exact product mod 2^1024 (host MulDecimalOnCpu, decimal.cu:1325-1351),
not the device operator*= carry sequence (decimal.cu:1808-1831).
"""

NT = 32                 # NUM_TOTAL_DIG
LAST = NT - 1           # INDEX_LAST_DIG
M32 = 0xFFFFFFFF
M64 = (1 << 64) - 1
MOD = 1 << (32 * NT)
DIG_PER_INT32 = 9
DIV_FIX_EX_FRAC = 4
# __ARRAY_SCALE[n-1] / __POW10_ARRAY[n-1]: 10^n as limbs, n = 1..38 (decimal.cu:39-90)
POW10_ROWS = []
for _n in range(1, 39):
    _p, _row = 10 ** _n, []
    while _p:
        _row.append(_p & M32)
        _p >>= 32
    POW10_ROWS.append(_row)


class Undefined(Exception):
    """the reference's behaviour is not defined for this input"""


class Dec:
    __slots__ = ("sign", "prec", "frac", "error", "v")

    def __init__(self, mag=0, sign=0, frac=0, prec=0):
        self.sign, self.prec, self.frac, self.error = sign, prec, frac, 0
        self.v = [(mag >> (32 * i)) & M32 for i in range(NT)]

    def copy(self):
        d = Dec()
        d.sign, d.prec, d.frac, d.error, d.v = self.sign, self.prec, self.frac, self.error, list(self.v)
        return d

    @property
    def mag(self):
        m = 0
        for i in reversed(range(NT)):
            m = (m << 32) | self.v[i]
        return m

    def set_mag(self, m):
        m %= MOD
        self.v = [(m >> (32 * i)) & M32 for i in range(NT)]

    def is_zero(self):
        return not any(self.v)


def top_index(d):
    """`while (v[z] == 0) z--` from INDEX_LAST_DIG, on a uint32 index: a zero
    value has no top limb and the scan runs off the array"""
    z = LAST
    while d.v[z] == 0:
        z -= 1
        if z < 0:
            raise Undefined("limb scan of a zero value runs below v[0]")
    return z


def hit_bit(d):
    """bfind of the top limb, plus one, plus 32 per limb below it"""
    z = top_index(d)
    return d.v[z].bit_length() + 32 * z


def shl10(d, n):
    """operator<<(n), decimal.cu:2028-2058: x 10^n, limbs past v[31] dropped"""
    if n <= 0:
        return d
    if n > 38:
        raise Undefined("operator<< reads past __POW10_ARRAY (n > 38)")
    row = POW10_ROWS[n - 1]
    res = [0] * (NT + 1)    # res[32] is written by the i + j == 32 term and never read
    for i in range(NT):
        carry = 0
        for j in range(len(row)):
            if i + j > NT:
                break
            t = d.v[i] * row[j] + res[i + j] + carry
            carry = t >> 32
            res[i + j] = t & M32
        if i + len(row) < NT:
            res[i + len(row)] = carry
    d.v = res[:NT]
    return d


def align(a, b):
    """AlignAddSubData: the operand of smaller scale is raised to the larger"""
    if a.frac < b.frac:
        shl10(a, b.frac - a.frac)
        a.frac = b.frac
    elif b.frac < a.frac:
        shl10(b, a.frac - b.frac)
        b.frac = a.frac


def lt(l_, r_):
    """operator< (decimal.cu:687-722): sign first, then aligned limbs"""
    if l_.sign != r_.sign:
        return l_.sign != 0
    l, r = l_.copy(), r_.copy()
    if l.frac != r.frac:
        align(l, r)
    for i in range(LAST, -1, -1):
        t = l.v[i] - r.v[i]
        if t:
            return t < 0 if l_.sign == 0 else t > 0
    return False


def eq(l_, r_):
    """operator== (decimal.cu:728-743)"""
    if l_.sign != r_.sign:
        return False
    l, r = l_.copy(), r_.copy()
    if l.frac != r.frac:
        align(l, r)
    return l.v == r.v


def iadd(a, d):
    """device operator+= (decimal.cu:1362-1396)"""
    added = d.copy()
    align(added, a)
    if a.sign == added.sign:
        a.set_mag(a.mag + added.mag)
    else:
        r = 0
        for i in range(LAST, -1, -1):
            r = a.v[i] - added.v[i]
            if r:
                break
        a.set_mag(a.mag - added.mag if r >= 0 else added.mag - a.mag)
        a.sign = 1 if (r > 0 and not d.sign) or (r < 0 and d.sign) else 0
    return a


def isub(a, d):
    """device operator-= (decimal.cu:1575-1613)"""
    added = d.copy()
    align(added, a)
    if added.sign != a.sign:
        a.set_mag(a.mag + added.mag)
    else:
        r = 0
        for i in range(LAST, -1, -1):
            r = a.v[i] - added.v[i]
            if r:
                break
        a.set_mag(a.mag - added.mag if r >= 0 else added.mag - a.mag)
        a.sign = 1 if (r > 0 and added.sign) or (r < 0 and not added.sign) else 0
    return a


def mul(a, b):
    """qmid * dt: sign xor, frac sum, exact product mod 2^1024 (the contract's
    product; see the module docstring)"""
    r = a.copy()
    r.sign = a.sign ^ b.sign
    r.frac = a.frac + b.frac
    r.set_mag(a.mag * b.mag)
    return r


def half_int(d1, d2):
    """HalfIntDecimal (decimal.cu:2093-2122): (d1 + d2) / 2 word by word,
    the odd bit of each word carried down as 2^32 (PER_DEC_MAX_SCALE)"""
    tmp = iadd(d1.copy(), d2)
    rds = 0
    t = list(tmp.v)
    for i in range(LAST, -1, -1):
        if rds:
            t[i] += rds * (1 << 32)
        if t[i]:
            rds = t[i] % 2
            t[i] //= 2
    tmp.v = [x & M32 for x in t]
    return tmp


def _pow2_store(pexp):
    """qmax/qmin construction (decimal.cu:2157-2163): v[p/32] = 1 << (p % 32),
    C truncating division and remainder.  Returns the Decimal."""
    q = Dec(0)
    idx = int(pexp / 32)
    sh = pexp - 32 * idx          # C's % (sign of the dividend)
    word = 0 if sh < 0 else (1 << sh)   # PTX shl.b32 clamps a negative count: 0
    if idx >= 0:
        q.v[idx] = word
    elif idx == -1:
        # the four bytes before v[]: sign, prec, frac, error (little-endian)
        q.sign, q.prec, q.frac, q.error = word & 0xFF, (word >> 8) & 0xFF, (word >> 16) & 0xFF, word >> 24
    else:
        raise Undefined("DivInt stores outside qmax (pmax < -63)")
    return q


def div_int(ds, dt, dt_hit):
    """DivInt (decimal.cu:2125-2191): returns (quotient, residual)"""
    if ds.is_zero():
        return ds.copy(), Dec(0)
    ds_hit = hit_bit(ds)
    pmax = ds_hit - dt_hit + 1
    pmin = ds_hit - dt_hit - 1
    if pmin < 0:
        pmin = 0
    qmax = _pow2_store(pmax)
    qmin = _pow2_store(pmin)
    steps = 0
    while lt(qmin, qmax):
        qmid = half_int(qmax, qmin)
        if eq(qmid, qmin):
            break
        restmp = mul(qmid, dt)
        if eq(restmp, ds):
            qmin = qmid
            break
        elif lt(restmp, ds):
            qmin = qmid
        else:
            qmax = qmid
        steps += 1
        assert steps < 4096
    residual = isub(ds.copy(), mul(qmin, dt))
    return qmin, residual


def digits_room(d, dtz):
    """(INDEX_LAST_DIG - dtz + 1) * 9 minus the decimal digits of v[dtz]"""
    n = (LAST - dtz + 1) * DIG_PER_INT32
    t = d.v[dtz]
    while t > 0:
        t //= 10
        n -= 1
    return n


def div_by_int64(x, divisor, shift, is_mod):
    """DivByInt64 (decimal.cu:2431-2446), uint64 arithmetic throughout"""
    dvs = (x.v[0] | (x.v[1] << 32))
    while shift > DIG_PER_INT32:
        dvs = dvs * 10 ** 9 & M64
        shift -= DIG_PER_INT32
    dvs = dvs * 10 ** shift & M64
    dvt = divisor.v[0] | (divisor.v[1] << 32)
    if is_mod:
        res = dvs % dvt
    else:
        res = (dvs // dvt + (1 if ((dvs % dvt) << 1) & M64 >= dvt else 0)) & M64
    x.v[1] = res >> 32
    x.v[0] = res & M32
    return x


def div_by_int(x, d, shift, is_mod):
    """DivByInt (decimal.cu:2335-2429): one-word divisor, chunked shifts,
    round-up test `remainder * 2 >= dvs` in uint32"""
    dvs = d.v[0]
    remainder = 0
    tmp_ex = shift
    dtz = top_index(x)
    n = digits_room(x, dtz)
    if n < 0:
        n = 0
    if n > tmp_ex:
        n = tmp_ex
    tmp_ex -= n
    shl10(x, n)
    for i in range(LAST, -1, -1):
        if x.v[i] or remainder:
            t = x.v[i] + remainder * (1 << 32)
            x.v[i] = t // dvs
            remainder = t % dvs
    while tmp_ex > 0:
        tmp_res = Dec(remainder)
        dtz = top_index(tmp_res)
        n = digits_room(tmp_res, dtz)
        if n > tmp_ex:
            n = tmp_ex
        tmp_ex -= n
        shl10(tmp_res, n)
        remainder = 0
        for i in range(LAST, -1, -1):
            if tmp_res.v[i] or remainder:
                t = tmp_res.v[i] + remainder * (1 << 32)
                tmp_res.v[i] = t // dvs
                remainder = t % dvs
        shl10(x, n)
        iadd(x, tmp_res)
    if not is_mod and (remainder * 2) & M32 >= dvs:
        x.set_mag(x.mag + 1)
    return x


def div_or_mod(a, d):
    """a / d by DivOrMod(d, isMod=false), decimal.cu:2201-2333.  Returns
    (result Decimal, path, chunks) with path in {"zero", "div0", "i", "ii", "iii"}."""
    this = a.copy()
    divitend, divisor = a.copy(), d.copy()
    this.sign = d.sign ^ this.sign
    this.frac += DIV_FIX_EX_FRAC
    if this.is_zero():
        this.sign = 0
        this.prec = this.frac + 1
        return this, "zero", 0
    if d.is_zero():
        this.error = 2
        return this, "div0", 0
    shift = divisor.frac + DIV_FIX_EX_FRAC
    divitend.frac = divisor.frac = 0
    divitend.sign = divisor.sign = 0
    res = Dec(0)
    dtz = top_index(divitend)
    dsz = top_index(divisor)
    hit_dt = divitend.v[dtz].bit_length() + dtz * 32
    hit_shift = (shift // 3) * 10 + {0: 0, 1: 4, 2: 7}[shift % 3]
    hit_ds = divisor.v[dsz].bit_length() + dsz * 32
    chunks = 1
    if hit_dt + hit_shift - 1 <= 64 and hit_ds <= 64:
        res = div_by_int64(divitend, divisor, shift, False)
        path = "i"
    elif dsz == 0:
        res = div_by_int(divitend, divisor, shift, False)
        path = "ii"
    else:
        path = "iii"
        tmp_ex = shift
        chunks = 0
        while tmp_ex > 0:
            chunks += 1
            n = digits_room(divitend, dtz)      # dtz is not recomputed per chunk
            if n > tmp_ex:
                n = tmp_ex
            if n <= 0:
                raise Undefined("DivOrMod chunk with a non-positive digit count")
            tmp_ex -= n
            shl10(divitend, n)
            tmp_res, divitend = div_int(divitend, divisor, hit_ds)
            if not res.is_zero():
                shl10(res, n)
            iadd(res, tmp_res)
        double = iadd(divitend.copy(), divitend)
        if not lt(double, divisor):
            res.set_mag(res.mag + 1)
    this.v = list(res.v)
    return this, path, chunks


def mod_calc(a, d):
    """a % d by ModCalc (decimal.cu:2621-2680).  Returns (result, path)."""
    this = a.copy()
    divitend, divisor = a.copy(), d.copy()
    this.sign = d.sign ^ this.sign
    this.prec = d.prec
    this.frac = 0
    if this.is_zero():
        this.sign = 0
        this.prec = this.frac + 1
        return this, "zero"
    divitend.frac = divisor.frac = 0
    divitend.sign = divisor.sign = 0
    dtz = top_index(divitend)
    if d.is_zero():
        raise Undefined("ModCalc scans the limbs of a zero divisor")
    dsz = top_index(divisor)
    hit_dt = divitend.v[dtz].bit_length() + dtz * 32
    hit_ds = divisor.v[dsz].bit_length() + dsz * 32
    if hit_dt <= 64 and hit_ds <= 64:
        res = div_by_int64(divitend, divisor, 0, True)      # ModByInt64: dvs % dvt
        path = "i"
    elif dsz == 0:
        dvs, remainder = divisor.v[0], 0                     # ModByInt
        for i in range(LAST, -1, -1):
            if divitend.v[i] or remainder:
                t = divitend.v[i] + remainder * (1 << 32)
                divitend.v[i] = t // dvs
                remainder = t % dvs
        res = Dec(remainder)
        path = "ii"
    else:
        _, res = div_int(divitend, divisor, hit_ds)
        path = "iii"
    this.v = list(res.v)          # CopyValue: limbs only, the residual's sign is not kept
    return this, path
