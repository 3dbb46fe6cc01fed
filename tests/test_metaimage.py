"""MetaImage round trips and the header layout of util/io/mhd/writeMetaImage_mhd.m."""
import numpy as np
import pytest

from ctsim_amd.metaimage import read_metaimage, write_metaimage


@pytest.mark.parametrize("dt", [np.float32, np.float64, np.int16, np.uint16, np.uint8, np.int8])
def test_round_trip_keeps_values_order_and_type(tmp_path, dt):
    rng = np.random.default_rng(0)
    a = (rng.uniform(0, 100, (5, 4, 3))).astype(dt)
    mhd = write_metaimage(str(tmp_path / "vol"), a, spacing=(0.5, 0.7, 1.0), origin=(1, 2, 3))
    b, meta = read_metaimage(mhd)
    assert b.dtype == a.dtype and b.shape == a.shape
    np.testing.assert_array_equal(a, b)
    assert meta["ElementSpacing"] == [0.5, 0.7, 1.0] and meta["Position"] == [1.0, 2.0, 3.0]


def test_header_order_and_first_dimension_fastest(tmp_path):
    a = np.arange(6, dtype=np.float32).reshape((3, 2), order="F")  # MATLAB [3 2]
    mhd = write_metaimage(str(tmp_path / "img.mhd"), a)
    text = open(mhd).read().splitlines()
    keys = [line.split("=")[0].strip() for line in text]
    assert keys == ["NDims", "DimSize", "BinaryDataByteOrderMSB", "ElementSpacing", "ElementType",
                    "ElementByteOrderMSB", "ElementDataFile"]
    assert text[1] == "DimSize = 3 2" and text[4] == "ElementType = MET_FLOAT"
    raw = np.fromfile(str(tmp_path / "img.raw"), "<f4")
    np.testing.assert_array_equal(raw, np.arange(6, dtype=np.float32))


def test_reader_takes_the_last_bytes_and_rejects_big_endian(tmp_path):
    a = np.arange(4, dtype=np.int16)
    (tmp_path / "v.raw").write_bytes(b"HEADERJUNK" + a.astype("<i2").tobytes())
    (tmp_path / "v.mhd").write_text("NDims = 1\nDimSize = 4\nElementType = MET_SHORT\nElementDataFile = v.raw\n")
    b, _ = read_metaimage(str(tmp_path / "v.mhd"))
    np.testing.assert_array_equal(b, a)
    (tmp_path / "w.mhd").write_text("NDims = 1\nDimSize = 4\nElementType = MET_SHORT\n"
                                    "ElementByteOrderMSB = True\nElementDataFile = v.raw\n")
    with pytest.raises(ValueError):
        read_metaimage(str(tmp_path / "w.mhd"))
