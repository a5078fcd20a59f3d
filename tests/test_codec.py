"""CPU: the compact codec at the edges -- bulk ingest of decimal text
(Decimal(P, S, text) then ToCompactDecimal) and output formatting (GetDecimal
hex, decimal text) -- in the oracle, pinned against

1. the reference's own known answers (tests/golden/known_answers.json, from
   test/testcases/datatypes/test_decimal.cu): parse then print;
2. an independent Python-integer restatement of the character rules
   (tests/pyref.py: string_to_decimal, get_decimal_hex) on seeded random rows
   that include non-digit bytes, repeated and trailing dots, NULs, signs,
   bytes >= 0x80 and rows longer than 308 digits (the mod 2^1024 wrap);
3. closed forms: the hex form is the magnitude's base-16 digits, the text
   form parses back to the same value.
"""
import json
import os
import random

import numpy as np
import pytest

import pyref

HERE = os.path.dirname(os.path.abspath(__file__))
KA = json.load(open(os.path.join(HERE, "golden", "known_answers.json")))


def random_rows(rng, n, P):
    rows = []
    for i in range(n):
        k = i % 10
        if k < 4:  # ordinary decimals
            ip = str(rng.randrange(0, 10 ** rng.randint(1, max(1, P - 3))))
            fp = "".join(rng.choice("0123456789") for _ in range(rng.randint(0, 12)))
            t = ("-" if rng.random() < 0.4 else "") + ip + ("." + fp if fp or rng.random() < 0.2 else "")
            rows.append(t.encode())
        elif k == 4:  # long digit strings, past 2^1024 sometimes
            rows.append(("".join(rng.choice("0123456789") for _ in range(rng.randint(100, 420)))).encode())
        elif k == 5:  # long fractions
            rows.append(("0." + "".join(rng.choice("0123456789") for _ in range(rng.randint(20, 340)))).encode())
        elif k == 6:  # bytes that are not digits
            rows.append(bytes(rng.choice(b"0123456789.-+ eE,x") for _ in range(rng.randint(0, 30))))
        elif k == 7:  # any byte, NULs included
            rows.append(bytes(rng.randrange(256) for _ in range(rng.randint(0, 25))))
        elif k == 8:  # dots: repeated, doubled, trailing
            rows.append(rng.choice([b"1..5", b"12.", b".", b"-.", b"1.2.3", b"..", b"-0.00", b"-", b"", b"0"]))
        else:  # huge scale-ups
            rows.append(str(rng.randrange(10 ** 30)).encode())
    return rows


@pytest.mark.parametrize("P,S", [(1, 0), (9, 2), (15, 2), (18, 4), (38, 10), (76, 20), (150, 50), (300, 100), (300, 0)])
def test_parse_matches_python_restatement(O, P, S):
    rng = random.Random(P * 1000 + S)
    rows = random_rows(rng, 300, P)
    B = O.lib.od_compact_bytes(P)
    got = O.parse_compact(rows, P, S).reshape(len(rows), B)
    for r, raw in enumerate(rows):
        assert bytes(got[r]) == pyref.decimal_of_bytes(P, S, raw, B), (raw, P, S)


def test_parse_known_answers(O):
    """the reference's Decimal(P, S, text) known answers print back as written"""
    seen = 0
    for c in KA["compact_round_trip"]:
        P, S, text = c["P"], c["S"], c["text"]
        B = O.lib.od_compact_bytes(P)
        buf = O.parse_compact([text], P, S)
        d = O.from_compact(bytes(buf), P, S)
        assert O.to_str(d) == O.to_str(O.dec_cast(text, P, S))
        seen += 1
    assert seen > 0
    for case in KA["binary"]:
        for P, S, text in (case["a"], case["b"]):
            B = O.lib.od_compact_bytes(P)
            buf = O.parse_compact([text], P, S)
            want = pyref.decimal_of_bytes(P, S, text.encode(), B)
            assert bytes(buf) == want, text


def test_hex_form_closed_form(O):
    rng = random.Random(7)
    for _ in range(500):
        bits = rng.choice([0, 1, 31, 32, 33, 64, 100, 500, 1023])
        m = rng.getrandbits(bits) if bits else 0
        d = O.OdDec()
        d.sign = rng.randint(0, 1)
        d.frac = rng.randint(0, 300)
        d.set_magnitude(m)
        h = O.format_hex(d)
        assert h == pyref.get_decimal_hex((d.sign, m, d.frac))
        # base-16 digits of the magnitude, whole limbs from the top nonzero one
        body = h[1:h.index("(")]
        assert len(body) % 8 == 0 and int(body or "0", 16) == m
        assert body == "" or body[:8] != "00000000"


@pytest.mark.parametrize("P,S", [(1, 0), (15, 2), (38, 10), (76, 20), (150, 50), (300, 100)])
def test_format_compact_text_and_hex(O, P, S):
    """every compact row, including -0 and the largest magnitude, prints as
    text that parses back to it, and as GetDecimal hex"""
    rng = random.Random(P)
    B = O.lib.od_compact_bytes(P)
    rows = 200
    col = np.frombuffer(bytes(rng.randrange(256) for _ in range(rows * B)), dtype=np.uint8).copy()
    col[:B] = 0
    col[B - 1] = 0x80                    # -0
    col[B:2 * B] = 0xFF                  # largest magnitude, negative
    for hex_form in (False, True):
        stride = 8 * 32 + 8 if hex_form else 1 + 310 + 2 + 3
        out = O.format_compact(col, rows, P, S, hex_form, stride).reshape(rows, stride)
        for r in range(rows):
            v = pyref.compact_decode(bytes(col[r * B:(r + 1) * B]), S)
            slot = bytes(out[r])
            text = slot.split(b"\0", 1)[0].decode()
            assert not slot[len(text):].strip(b"\0"), "slot not zero-filled"
            if hex_form:
                assert text == pyref.get_decimal_hex(v)
            else:
                assert text == pyref.to_text(v)
                back = pyref.string_to_decimal(text.encode())
                assert (back[0], back[1]) == (v[0], v[1]) and back[2] == S
