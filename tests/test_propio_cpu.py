"""Property files (INC read/write, GSLIB write) through the library's C ABI.

The readers are checked against the reference's own pure-Python readers
(geo.py:168-212, restated in hpgl_amd/geo_bsd/geo.py as _load_prop_*_slow)
on the property files the reference's test scripts load
(tests/golden/inc/, copied unchanged); the writers by their exact text and
by round trips.  Host code only: these tests run without a GPU.
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
INC = os.path.join(HERE, "golden", "inc")


@pytest.fixture(scope="module")
def geo():
    import hpgl_amd.geo_bsd as g
    return g


def test_read_reference_float_file_matches_python_reader(geo):
    f = os.path.join(INC, "NEW_TEST_PROP.INC")
    fast = geo.load_cont_property(f, -99, size=(55, 52, 1))
    slow = geo.load_cont_property(f, -99)
    assert fast.data.shape == (55, 52, 1)
    assert np.array_equal(fast.data.ravel(order="F"), slow.data)
    assert np.array_equal(fast.mask.ravel(order="F"), slow.mask)
    assert 0 < int(fast.mask.sum()) < fast.mask.size  # both informed and undefined cells


def test_read_reference_indicator_file_matches_python_reader(geo):
    f = os.path.join(INC, "NEW_TEST_PROP_01.INC")
    fast = geo.load_ind_property(f, -99, [0, 1], size=(55, 52, 1))
    slow = geo.load_ind_property(f, -99, [0, 1])
    m = slow.mask == 1
    assert np.array_equal(fast.mask.ravel(order="F"), slow.mask)
    assert np.array_equal(fast.data.ravel(order="F")[m], slow.data[m])
    assert set(np.unique(fast.data.ravel(order="F")[m])) <= {0, 1}
    # indicator values map to their index: [1, 0] swaps the categories
    swapped = geo.load_ind_property(f, -99, [1, 0], size=(55, 52, 1))
    assert np.array_equal(swapped.data.ravel(order="F")[m], 1 - slow.data[m])


def test_read_reference_mean_file(geo):
    f = os.path.join(INC, "mp_0_8_52_55_1.inc")
    p = geo.load_cont_property(f, -99, size=(52, 55, 1))
    slow = geo.load_cont_property(f, -99)
    assert np.array_equal(p.data.ravel(order="F"), slow.data)
    assert np.all(p.data == np.float32(0.8))


def test_write_inc_float_text_and_round_trip(geo, tmp_path):
    rng = np.random.default_rng(5)
    data = rng.normal(size=(6, 5, 4)).astype(np.float32)
    mask = (rng.random((6, 5, 4)) < 0.7).astype(np.uint8)
    data = np.asfortranarray(data)
    mask = np.asfortranarray(mask)
    path = str(tmp_path / "p.inc")
    geo.write_property((data, mask), path, "PORO", -99)
    lines = open(path).read().split("\n")
    flat_d, flat_m = data.ravel(order="F"), mask.ravel(order="F")
    expect = ["PORO"] + ["%E" % (float(v) if m else -99.0) for v, m in zip(flat_d, flat_m)] + ["/", ""]
    assert lines == expect
    back = geo.load_cont_property(path, -99, size=(6, 5, 4))
    want = np.array([np.float32(float("%E" % v)) for v in flat_d])  # %E keeps 7 significant digits
    assert np.array_equal(back.mask.ravel(order="F"), flat_m)
    assert np.array_equal(back.data.ravel(order="F")[flat_m == 1], want[flat_m == 1])


def test_write_inc_byte_remap_and_round_trip(geo, tmp_path):
    rng = np.random.default_rng(6)
    data = np.asfortranarray(rng.integers(0, 3, size=(5, 4, 3)).astype(np.uint8))
    mask = np.asfortranarray((rng.random((5, 4, 3)) < 0.8).astype(np.uint8))
    path = str(tmp_path / "f.inc")
    geo.write_property((data, mask, 3), path, "FACIES", 200, indicator_values=[10, 20, 130])
    lines = open(path).read().split("\n")
    code = {0: "10", 1: "20", 2: str(130 - 256)}  # bytes are printed as signed chars
    flat_d, flat_m = data.ravel(order="F"), mask.ravel(order="F")
    expect = ["FACIES"] + [code[int(v)] if m else str(200 - 256) for v, m in zip(flat_d, flat_m)] + ["/", ""]
    assert lines == expect
    # the reader compares the int value with the undefined value, then maps
    # the value as a byte: "-126" is byte 130
    back = geo.load_ind_property(path, -56, [10, 20, 130], size=(5, 4, 3))
    assert np.array_equal(back.mask.ravel(order="F"), flat_m)
    assert np.array_equal(back.data.ravel(order="F")[flat_m == 1], flat_d[flat_m == 1])


def test_write_gslib(geo, tmp_path):
    data = np.asfortranarray(np.arange(12, dtype=np.float32).reshape(3, 2, 2))
    mask = np.asfortranarray(np.ones((3, 2, 2), np.uint8))
    mask[0, 0, 0] = 0
    path = str(tmp_path / "g.dat")
    geo.write_gslib_property((data, mask), path, "VAL", -999.0)
    lines = open(path).read().split("\n")
    flat = data.ravel(order="F")
    assert lines[:3] == ["HPGL saved GSLIB file", "1", "VAL"]
    assert lines[3] == "%E" % -999.0
    assert lines[4:-1] == ["%E" % float(v) for v in flat[1:]]
    ind = np.asfortranarray(np.array([0, 1, 1, 0], np.uint8).reshape(4, 1, 1))
    path2 = str(tmp_path / "g2.dat")
    geo.write_gslib_property((ind, np.asfortranarray(np.ones((4, 1, 1), np.uint8)), 2), path2, "IND", -99,
                             indicator_values=[5, 7])
    assert open(path2).read().split("\n")[3:7] == ["5", "7", "7", "5"]


def test_reader_errors(geo, tmp_path):
    with pytest.raises(geo.HpglError, match="Error opening file"):
        geo.read_inc_file_float(str(tmp_path / "missing.inc"), -99, (2, 1, 1))
    p = tmp_path / "short.inc"
    p.write_text("NAME\n1 2\n/\n")
    with pytest.raises(geo.HpglError, match="Unexpected end of data"):
        geo.read_inc_file_float(str(p), -99, (3, 1, 1))
    p.write_text("-- only a comment\n")
    with pytest.raises(geo.HpglError, match="Property name not found"):
        geo.read_inc_file_float(str(p), -99, (1, 1, 1))
    p.write_text("NAME\n1 x 2\n")
    with pytest.raises(geo.HpglError, match="Error parsing 'x'"):
        geo.read_inc_file_float(str(p), -99, (3, 1, 1))


def test_reader_tokens(geo, tmp_path):
    """Comments inside the data, a name after non-letter lines, prefixes
    converted as scanf would (3*0.5 reads as 3), a leading empty line."""
    p = tmp_path / "t.inc"
    p.write_text("\n12 not a name\n  -- comment\nPERMX a b\n1.5 -- rest ignored 9 9\n 2e1 3*0.5\n-99 /\n")
    r = geo.read_inc_file_float(str(p), -99, (4, 1, 1))
    assert np.array_equal(r.data.ravel(order="F"), np.array([1.5, 20.0, 3.0, -99.0], np.float32))
    assert np.array_equal(r.mask.ravel(order="F"), [1, 1, 1, 0])
