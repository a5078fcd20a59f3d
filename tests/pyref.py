"""Independent Python-integer restatement of the reference's host Decimal
rules, written for cross-checking the C oracle (tests only).

A value is (sign, magnitude, scale).  Rules (decimal.cu):
  add/sub (1284-1323, 1575-1613): align to the larger scale; same sign ->
      magnitudes add, sign kept; different signs -> larger minus smaller with
      the sign of the larger, +0 on a tie.  a - b == a + (-b).
  mul (1325-1351): sign xor, scale sum, magnitude product.
  div (2201-2333): scale Sa + 4; 0/x -> +0; x/0 -> |x| kept (error flag);
      q = round-half-up(|a| * 10^(Sb+4) / |b|); faithful mode keeps the
      uint64 and uint32 path behaviours, and the bisection's quotient 1
      for a numerator 1 to 64 bits shorter than the divisor.
Faithful mode reduces magnitudes mod 2^1024 after every operation.
"""
MOD1024 = 1 << 1024


def wrap(m, faithful):
    return m % MOD1024 if faithful else m


def parse(text):
    """StringToDecimal (decimal.cu:2980-3025)"""
    sign = 0
    if text.startswith("-"):
        sign, text = 1, text[1:]
    if "." in text:
        i, f = text.split(".")
    else:
        i, f = text, ""
    return sign, int(i + f or "0"), len(f)


def cast(v, S):
    """cast to scale S: multiply up, truncate down (decimal.cu:471-533)"""
    sign, m, s = v
    if S >= s:
        return sign, m * 10 ** (S - s), S
    return sign, m // 10 ** (s - S), S


def decimal(P, S, text):
    return cast(parse(text), S)


def neg(v):
    return (v[0] ^ 1, v[1], v[2])


def add(a, b, faithful=True):
    sa, ma, xa = a
    sb, mb, xb = b
    S = max(xa, xb)
    ma = wrap(ma * 10 ** (S - xa), faithful)
    mb = wrap(mb * 10 ** (S - xb), faithful)
    if sa == sb:
        return sa, wrap(ma + mb, faithful), S
    if ma > mb:
        return sa, ma - mb, S
    if ma < mb:
        return sb, mb - ma, S
    return 0, 0, S


def sub(a, b, faithful=True):
    return add(a, neg(b), faithful)


def mul(a, b, faithful=True):
    return a[0] ^ b[0], wrap(a[1] * b[1], faithful), a[2] + b[2]


def _est_shift_bits(s):
    return (s // 3) * 10 + {0: 0, 1: 4, 2: 7}[s % 3]


def div(a, b, faithful=True):
    """returns (sign, magnitude, scale, error)"""
    sa, ma, xa = a
    sb, mb, xb = b
    S = xa + 4
    if ma == 0:
        return 0, 0, S, 0
    if mb == 0:
        return sa ^ sb, ma, S, 2
    s = xb + 4
    sign = sa ^ sb
    if faithful and ma.bit_length() + _est_shift_bits(s) - 1 <= 64 and mb.bit_length() <= 64:
        M64 = (1 << 64) - 1
        dvs = (ma & M64) * 10 ** s & M64
        q, r = divmod(dvs, mb)
        q = (q + (1 if ((r << 1) & M64) >= mb else 0)) & M64
        return sign, q, S, 0
    X = ma * 10 ** s
    q, r = divmod(X, mb)
    if faithful and mb < (1 << 32):
        up = ((2 * r) & 0xFFFFFFFF) >= mb
    elif faithful and 1 <= mb.bit_length() - X.bit_length() <= 64:
        q, up = 1, False
    else:
        up = 2 * r >= mb
    return sign, wrap(q + (1 if up else 0), faithful), S, 0


def to_text(v):
    sign, m, s = v[0], v[1], v[2]
    t = str(m).rjust(s + 1, "0")
    if s:
        t = t[:-s] + "." + t[-s:]
    return ("-" if sign else "") + t


def compact_encode(v, nbytes):
    sign, m = v[0], v[1]
    b = bytearray((m % (1 << (8 * nbytes))).to_bytes(nbytes, "little"))
    b[-1] |= sign << 7
    return bytes(b)


def compact_decode(buf, S):
    b = bytearray(buf)
    sign = b[-1] >> 7
    b[-1] &= 0x7F
    return sign, int.from_bytes(bytes(b), "little"), S


def string_to_decimal(raw):
    """StringToDecimal (decimal.cu:2980-3025) over the bytes of one row, cut
    at its first NUL: v = v*10 + (c - '0') mod 2^1024 for every byte taken,
    (c - '0') signed (bytes are C chars, signed); after a '.', the next byte
    is taken whatever it is; bytes after the first '.' are fractional.  A
    '.' that ends the row ends it."""
    raw = bytes(raw).split(b"\0", 1)[0]
    sign, i, v, frac, dot = 0, 0, 0, 0, False
    if raw[:1] == b"-":
        sign, i = 1, 1
    while i < len(raw):
        c = raw[i]
        if c == ord("."):
            dot = True
            i += 1
            if i >= len(raw):
                break
            c = raw[i]
        if dot:
            frac += 1
        d = (c - 256 if c >= 128 else c) - ord("0")
        v = (v * 10 + d) % MOD1024
        i += 1
    return sign, v, frac


def decimal_of_bytes(P, S, raw, nbytes):
    """Decimal(P, S, text) (decimal.cu:154-165) then ToCompactDecimal (:337-341)"""
    sign, m, s = string_to_decimal(raw)
    if S >= s:
        m = (m * 10 ** (S - s)) % MOD1024
    else:
        m = m // 10 ** (s - S)
    return compact_encode((sign, m, S), nbytes)


def get_decimal_hex(v):
    """GetDecimal (decimal.cu:362-397): sign, 8 hex digits per limb from the
    top nonzero one of 32, '(' three scale digits ')'"""
    sign, m, s = v[0], v[1], v[2]
    m %= MOD1024
    limbs = [(m >> (32 * i)) & 0xFFFFFFFF for i in range(32)]
    top = max((i for i in range(32) if limbs[i]), default=-1)
    body = "".join("%08X" % limbs[i] for i in range(top, -1, -1))
    return ("-" if sign else "+") + body + "(" + "%d%d%d" % (s // 100, s // 10 % 10, s % 10) + ")"


def mod(a, b, faithful=True):
    """ModCalc (decimal.cu:2621-2716) in closed form: scales dropped, sign
    xor, 0 % x = +0, x % 0 keeps |x|; faithful: when the reference bisects
    (not both <= 64 bits, |b| >= 2^32) and bitlen(|a|) is bitlen(|b|) down
    to bitlen(|b|) - 64, the result is ||a| - |b||"""
    sa, A = a[0], a[1]
    sb, B = b[0], b[1]
    if A == 0:
        return 0, 0, 0
    sign = sa ^ sb
    if B == 0:
        return sign, A, 0
    la, lb = A.bit_length(), B.bit_length()
    if faithful and not (la <= 64 and lb <= 64) and B >= 1 << 32 and lb - 64 <= la <= lb:
        return sign, abs(A - B), 0
    return sign, A % B, 0


def truncate(v, P, p):
    """Decimal::truncate(p) (decimal.cu:541-570) in closed form on a value of
    type (P, S): returns (sign, magnitude, scale, precision).  p > 0: to scale
    p by truncation or padding; p <= 0: the integer part with its last -p
    digits zeroed when the type has more than -p integer digits, else 0 (the
    sign is kept)."""
    sign, m, S = v
    if p > 0:
        m = m // 10 ** (S - p) if S >= p else m * 10 ** (p - S)
        return sign, m, p, P - S + p
    q = -p
    if P - S > q:
        return sign, m // 10 ** (S + q) * 10 ** q, 0, P - S
    return sign, 0, 0, 1
