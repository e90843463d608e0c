"""PDL::IO::FastRaw (lib/PDL/IO/FastRaw.pm) restated in pdl_amd/fastraw.py:
the scenarios of the reference's t/fastraw.t on the host halves (no GPU),
the header layout of FastRaw.pm:221-229, and gluefraw's refusals
(FastRaw.pm:327-340).  A device round trip is in the GPU part."""
import os

import numpy as np
import pytest

from pdl_amd import fastraw as F
from pdl_amd._lib import PDLError


def startdata():
    # t/fastraw.t:16: pdl [2,3],[4,5],[6,7] -- a double (2,3) ndarray
    return np.array([[2, 3], [4, 5], [6, 7]], dtype=np.float64)


def test_write_creates_data_and_header_and_reads_back(tmp_path):
    """t/fastraw.t:19-26"""
    name = str(tmp_path / "tmp0")
    x = startdata()
    F.write_host(name, 10, [2, 3], x)
    assert os.path.isfile(name) and os.path.isfile(name + ".hdr")
    with open(name + ".hdr") as h:
        assert h.read() == "10\n2\n2 3\n"
    assert os.path.getsize(name) == 6 * 8
    code, dims, y = F.read_host(name)
    assert code == 10 and dims == [2, 3]
    np.testing.assert_array_equal(y, x)


def test_gluefraw_appends_along_the_last_dim(tmp_path):
    """t/fastraw.t:32-38: writefraw then gluefraw pdl [10,11]"""
    name = str(tmp_path / "tmp0.g")
    F.write_host(name, 10, [2, 3], startdata())
    F.glue_host(name, 10, [2], np.array([10, 11], dtype=np.float64))
    code, dims, y = F.read_host(name)
    assert dims == [2, 4]
    np.testing.assert_array_equal(y, [[2, 3], [4, 5], [6, 7], [10, 11]])


def test_custom_header_file(tmp_path):
    """t/fastraw.t:42-50"""
    name, header = str(tmp_path / "tmp0"), str(tmp_path / "headerfile")
    F.write_host(name, 10, [2, 3], startdata(), header=header)
    assert os.path.isfile(header) and not os.path.exists(name + ".hdr")
    np.testing.assert_array_equal(F.read_host(name, header=header)[2], startdata())


def test_mapfraw_reads_the_data(tmp_path):
    """t/fastraw.t:54-64 (read side)"""
    name = str(tmp_path / "tmp0")
    F.write_host(name, 10, [2, 3], startdata())
    np.testing.assert_array_equal(np.asarray(F.mapfraw(name)), startdata())


@pytest.mark.parametrize("code,dt", [(0, np.int8), (1, np.uint8), (2, np.int16), (3, np.uint16), (4, np.int32),
                                     (5, np.uint32), (6, np.int64), (7, np.uint64), (8, np.int64),
                                     (9, np.float32), (10, np.float64)])
def test_every_real_type_round_trips(tmp_path, code, dt):
    rng = np.random.default_rng(code)
    x = (rng.standard_normal((4, 5, 3)) * 100).astype(dt)
    name = str(tmp_path / ("t%d" % code))
    F.write_host(name, code, [3, 5, 4], x)
    c, dims, y = F.read_host(name)
    assert c == code and dims == [3, 5, 4] and y.dtype == np.dtype(dt)
    assert y.tobytes() == x.tobytes()


def test_short_file_reads_as_zero_filled(tmp_path):
    """readfraw starts from zeroes() and reads what the file holds (FastRaw.pm:297-309)"""
    name = str(tmp_path / "short")
    F.write_host(name, 4, [6], np.arange(6, dtype=np.int32))
    with open(name, "r+b") as f:
        f.truncate(3 * 4)
    np.testing.assert_array_equal(F.read_host(name)[2], [0, 1, 2, 0, 0, 0])


def test_header_and_glue_errors(tmp_path):
    name = str(tmp_path / "e")
    with pytest.raises(PDLError, match="Couldn't open '.*e.hdr' for reading"):
        F.read_header(name)
    with open(name + ".hdr", "w") as h:
        h.write("10\n3\n2 3\n")
    with pytest.raises(PDLError, match="Format error reading fraw header file"):
        F.read_header(name)
    F.write_host(name, 10, [2, 3], startdata())
    with pytest.raises(PDLError, match="gluefraw: ndarray has type 'float' but file has type '10'"):
        F.glue_host(name, 9, [2], np.zeros(2, np.float32))
    with pytest.raises(PDLError, match=r"gluefraw: incompatible lower dims, ndarray \(3\) vs header \(2\)"):
        F.glue_host(name, 10, [3], np.zeros(3))
    F.write_host(name, 10, [6], np.zeros(6))
    with pytest.raises(PDLError, match=r"gluefraw: header dims needs at least 2 dims, got \(6\)"):
        F.glue_host(name, 10, [6], np.zeros(6))
    with pytest.raises(PDLError, match="must be real filename"):
        F.glue_host(str(tmp_path / "nope"), 10, [2], np.zeros(2))


@pytest.mark.gpu
def test_device_ndarrays_round_trip(tmp_path):
    """writefraw of a device ndarray (a transposed view: its elements in the
    view's own order) and readfraw back into HBM"""
    from pdl_amd import core as C
    C.init(0)
    x = np.arange(24, dtype=np.int16).reshape(4, 6)
    d = C.ndarray.from_numpy(x).xchg(0, 1)
    name = str(tmp_path / "dev")
    F.writefraw(d, name)
    assert F.read_header(name) == {"Type": 2, "Dims": [4, 6], "NDims": 2}
    back = F.readfraw(name)
    assert back.dims == [4, 6] and back.type() == "short"
    np.testing.assert_array_equal(back.numpy(), x.T)
    F.gluefraw(C.ndarray.from_numpy(np.array([100, 101, 102, 103], np.int16)), name)
    np.testing.assert_array_equal(F.readfraw(name).numpy()[-1], [100, 101, 102, 103])
