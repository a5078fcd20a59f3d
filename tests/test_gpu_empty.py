"""GPU: every entry point with empty inputs -- zero rows, zero groups --
returns without touching its outputs beyond what the empty result defines
(SUM of no rows is +0 with count 0; grouped runs find no groups)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C15 = [("p", 15, 2), ("d", 15, 2)]


def test_calc_zero_rows(U, gpu):
    e = U.Expr("p*(1-d)", C15)
    col = U.DeviceBuffer(64).zero()
    out = U.DeviceBuffer.from_numpy(np.full(64, 0xAB, dtype=np.uint8))
    e.calc([col, col], 0, out)
    U.device_sync()
    assert (out.numpy() == 0xAB).all()


def test_agg_zero_rows(U, O, gpu):
    col = U.DeviceBuffer(64).zero()
    agg = U.Agg(["p*(1-d)", "p"], C15)
    images = U.DeviceBuffer(2 * 133).zero()
    counts = U.DeviceBuffer(8).zero()
    ng, keys = agg.run([col, col], 0, images, counts, max_groups=1)
    assert ng == 1
    assert int(counts.numpy("int64")[0]) == 0
    img = images.numpy().reshape(2, 133)
    for k in range(2):
        sign, _, _, _, mag = O.image_value(bytes(img[k]))
        assert (sign, mag) == (0, 0)   # +0


def test_grouped_zero_rows(U, gpu):
    col = U.DeviceBuffer(64).zero()
    f1 = U.DeviceBuffer(64).zero()
    ship = U.DeviceBuffer(64).zero()
    agg = U.Agg(["p"], C15, filter=("<=", 10), nkeys=2)
    images = U.DeviceBuffer(16 * 133).zero()
    counts = U.DeviceBuffer(16 * 8).zero()
    ng, keys = agg.run([col, col], 0, images, counts, filter_col=ship, key0=f1, key1=f1, max_groups=16)
    assert ng == 0 and list(keys) == []


def test_blocks_and_codec_zero_rows(U, gpu):
    host = np.zeros(7 * 4, dtype=np.uint8)
    b = U.BlockColumn([U.DeviceBuffer.from_numpy(host)], [4])
    e = U.Expr("p*2", [("p", 15, 2)])
    out = U.DeviceBuffer.from_numpy(np.full(32, 0xAB, dtype=np.uint8))
    e.calc_blocks([b], 0, out)
    agg = U.Agg(["p"], [("p", 15, 2)])
    images = U.DeviceBuffer(133).zero()
    counts = U.DeviceBuffer.from_numpy(np.array([5], dtype=np.int64))
    agg.run_blocks([b], 0, images, counts)
    U.device_sync()
    assert int(counts.numpy("int64")[0]) == 0
    U.parse_compact(out, out, 0, 15, 2, out)
    U.format_compact(out, 0, 15, 2, U.FORMAT_TEXT, out, U.format_stride(15, 2, U.FORMAT_TEXT))
    U.div_int64(images, 133, counts, images, 0)
    U.device_sync()
    assert (out.numpy() == 0xAB).all()


def test_batch_with_empty_query(U, O, gpu):
    v = U.VmProgram("p*(1-d)", C15)
    col = U.DeviceBuffer(64).zero()
    b = U.Batch(max_queries=4, max_items=64)
    image = U.DeviceBuffer(133).zero()
    count = U.DeviceBuffer.from_numpy(np.array([7], dtype=np.int64))
    assert b.add(v, [col, col], 0, image, count) is not None
    b.launch()
    U.device_sync()
    assert int(count.numpy("int64")[0]) == 0
    sign, _, _, _, mag = O.image_value(bytes(image.numpy()))
    assert (sign, mag) == (0, 0)
