"""GPU parity of the compact codec at the edges, through the C ABI, bit-exact
against the oracle: bulk ingest of text rows (updec_parse_compact) and output
formatting (updec_format_compact, GetDecimal hex and decimal text)."""
import random

import numpy as np
import pytest

from test_codec import random_rows

pytestmark = pytest.mark.gpu

CASES = [(1, 0), (9, 2), (15, 2), (18, 4), (38, 10), (76, 20), (150, 50), (300, 100), (300, 0)]


def text_column(U, rows):
    chars = b"".join(rows)
    offs = np.zeros(len(rows) + 1, dtype=np.int64)
    offs[1:] = np.cumsum([len(r) for r in rows])
    dch = U.DeviceBuffer.from_numpy(np.frombuffer(chars + b"\0", dtype=np.uint8))
    doff = U.DeviceBuffer.from_numpy(offs)
    return dch, doff


@pytest.mark.parametrize("P,S", CASES)
@pytest.mark.parametrize("n", [1, 63, 64, 65, 2000])
def test_parse_bit_exact(U, O, gpu, P, S, n):
    rng = random.Random(P * 7 + S + n)
    rows = random_rows(rng, n, P)
    B = U.compact_bytes(P)
    dch, doff = text_column(U, rows)
    out = U.DeviceBuffer(n * B + 8).zero()
    U.parse_compact(dch, doff, n, P, S, out)
    got = out.numpy(count=n * B + 8)
    want = O.parse_compact(rows, P, S)
    assert np.array_equal(got[: n * B], want), f"first mismatch row {np.argmax(got[: n * B] != want) // B}"
    assert not got[n * B:].any(), "wrote past the last row"


def test_parse_unaligned_output(U, O, gpu):
    rng = random.Random(3)
    rows = random_rows(rng, 300, 15)
    B = U.compact_bytes(15)
    dch, doff = text_column(U, rows)
    buf = U.DeviceBuffer(300 * B + 16).zero()
    U.parse_compact(dch, doff, 300, 15, 2, U.DeviceView(buf, 1, 300 * B))
    got = buf.numpy()
    assert np.array_equal(got[1: 1 + 300 * B], O.parse_compact(rows, 15, 2))
    assert got[0] == 0 and not got[1 + 300 * B:].any()


@pytest.mark.parametrize("P,S", CASES)
@pytest.mark.parametrize("form", ["hex", "text"])
def test_format_bit_exact(U, O, gpu, P, S, form):
    rng = np.random.default_rng(P + S)
    n = 1000
    B = U.compact_bytes(P)
    col = rng.integers(0, 256, size=n * B, dtype=np.uint8)
    col[:B] = 0
    col[B - 1] = 0x80            # -0
    col[B:2 * B] = 0xFF          # largest magnitude, negative
    col[2 * B:3 * B] = 0         # +0
    f = U.FORMAT_HEX if form == "hex" else U.FORMAT_TEXT
    for extra in (0, 8):
        stride = U.format_stride(P, S, f) + extra
        dcol = U.DeviceBuffer.from_numpy(col)
        out = U.DeviceBuffer(n * stride + 16).zero()
        U.format_compact(dcol, n, P, S, f, out, stride)
        got = out.numpy()
        want = O.format_compact(col, n, P, S, form == "hex", stride)
        assert np.array_equal(got[: n * stride], want), \
            f"first mismatch row {np.argmax(got[: n * stride] != want) // stride}"
        assert not got[n * stride:].any()


@pytest.mark.parametrize("P,S", [(15, 2), (38, 10), (300, 100)])
def test_text_round_trip_on_device(U, O, gpu, P, S):
    """format as text, parse the text back: the same compact bytes"""
    rng = np.random.default_rng(11)
    n = 5000
    B = U.compact_bytes(P)
    col = rng.integers(0, 256, size=n * B, dtype=np.uint8)
    stride = U.format_stride(P, S, U.FORMAT_TEXT)
    dcol = U.DeviceBuffer.from_numpy(col)
    txt = U.DeviceBuffer(n * stride).zero()
    U.format_compact(dcol, n, P, S, U.FORMAT_TEXT, txt, stride)
    t = txt.numpy().reshape(n, stride)
    rows = [bytes(r).split(b"\0", 1)[0] for r in t]
    dch, doff = text_column(U, rows)
    back = U.DeviceBuffer(n * B).zero()
    U.parse_compact(dch, doff, n, P, S, back)
    assert np.array_equal(back.numpy(), col)


def test_codec_argument_errors(U, gpu):
    out = U.DeviceBuffer(64)
    with pytest.raises(U.UpdecError, match="precision or scale"):
        U.parse_compact(out, out, 1, 10, 11, out)
    with pytest.raises(U.UpdecError, match="128 bytes"):
        U.parse_compact(out, out, 1, 400, 2, out)
    with pytest.raises(U.UpdecError, match="stride"):
        U.format_compact(out, 1, 15, 2, U.FORMAT_TEXT, out, 4)
