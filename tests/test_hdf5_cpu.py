"""The package's HDF5 reader and writer (pyphasefield_amd/hdf5.py) against
files written and read by the HDF5 library itself (tests/golden/hdf5,
made by make_fixtures.py with h5py 3.3 / HDF5 1.10.6).  No GPU.

* every dataset of the fixtures -- the meshio time-series layout, the
  default file format (contiguous, compact, chunked with deflate / shuffle /
  fletcher32, partial and unwritten chunks, big-endian integers, nested and
  large groups) and libver="latest" (version-2 object headers, layout
  version 4 chunk indexes, dense link storage) -- equals what h5py read;
* files from the package's writer hash to the digests of files h5py read
  back unchanged;
* XDMF thermal histories with HDF heavy data load frame by frame.
"""

import hashlib
import json
import os
import shutil

import numpy as np
import pytest

from pyphasefield_amd import hdf5, thermal

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hdf5")


def _expected():
    exp = np.load(os.path.join(HERE, "expected.npz"))
    out = {}
    for key in exp.files:
        fname, *parts = key.split("|")
        out[(fname, "/".join(parts))] = exp[key]
    return out


EXPECTED = _expected()
SUPPORTED = sorted(k for k in EXPECTED if k != ("latest.h5", "unlimited"))


@pytest.mark.parametrize("fname,path", SUPPORTED, ids=["%s:%s" % k for k in SUPPORTED])
def test_datasets_equal_what_libhdf5_read(fname, path):
    want = EXPECTED[(fname, path)]
    got = hdf5.read_dataset(os.path.join(HERE, fname), path)
    assert got.shape == want.shape
    assert got.dtype.newbyteorder("=") == want.dtype.newbyteorder("=")
    assert np.array_equal(got, want)


def test_unsupported_chunk_index_is_named():
    with pytest.raises(hdf5.HDF5Error, match="extensible array"):
        hdf5.read_dataset(os.path.join(HERE, "latest.h5"), "unlimited")


def test_group_listing():
    with hdf5.File(os.path.join(HERE, "earliest.h5")) as f:
        assert f.keys("many") == ["x%02d" % k for k in range(20)]
        assert f.keys("group/sub") == ["deep"]
    with hdf5.File(os.path.join(HERE, "latest.h5")) as f:
        assert f.keys("dense") == ["frame%02d" % k for k in range(12)]
        with pytest.raises(KeyError):
            f.read("dense/frame99")


def test_not_hdf5_and_truncated(tmp_path):
    bad = tmp_path / "x.h5"
    bad.write_bytes(b"\0" * 4096)
    with pytest.raises(hdf5.HDF5Error, match="not an HDF5 file"):
        hdf5.File(str(bad))
    src = os.path.join(HERE, "earliest.h5")
    cut = tmp_path / "cut.h5"
    cut.write_bytes(open(src, "rb").read()[:3000])
    with pytest.raises(hdf5.HDF5Error):
        hdf5.read_dataset(str(cut), "gzip_shuffle_fletcher")


def _writer_cases():
    """make_fixtures.writer_cases (the script imports h5py at module level,
    so only that function's source is compiled here)."""
    src = open(os.path.join(HERE, "make_fixtures.py")).read()
    start = src.index("def writer_cases():")
    end = src.index("\ndef _package_hdf5")
    ns = {"np": np}
    exec(compile(src[start:end], "make_fixtures.writer_cases", "exec"), ns)
    return ns["writer_cases"]()


def test_writer_bytes_are_those_libhdf5_read(tmp_path):
    digests = json.load(open(os.path.join(HERE, "writer.json")))
    cases = _writer_cases()
    assert sorted(cases) == sorted(digests)
    for case, arrays in cases.items():
        path = str(tmp_path / (case + ".h5"))
        hdf5.write_datasets(path, arrays)
        assert hashlib.sha256(open(path, "rb").read()).hexdigest() == digests[case], case
        with hdf5.File(path) as f:
            assert f.keys() == sorted(arrays)
            for name, a in arrays.items():
                got = f.read(name)
                assert got.dtype.newbyteorder("=") == a.dtype.newbyteorder("=") and np.array_equal(got, a)


def test_meshio_time_series_loads(tmp_path):
    """The meshio layout: a mesh grid (data0) first, frames data1..data12
    named "meshio_T.h5:/dataK", times in the Time elements."""
    for name in ("meshio_T.xdmf", "meshio_T.h5"):
        shutil.copy(os.path.join(HERE, name), tmp_path / name)
    times = np.load(os.path.join(HERE, "meshio_T_times.npy"))
    ld = thermal.XDMFLoader(str(tmp_path / "meshio_T.xdmf"))
    assert ld.num_steps == 12
    for k in range(12):
        t, a = ld.data(k)
        assert t == times[k]
        assert np.array_equal(a, EXPECTED[("meshio_T.h5", "data%d" % (k + 1))])
    ld.close()
    hist = thermal.ThermalHistory(str(tmp_path / "meshio_T.xdmf"), (9, 13))
    t = 0.5 * (times[3] + times[4])
    hist.advance_to(t)
    r = (t - times[3]) / (times[4] - times[3])
    want = EXPECTED[("meshio_T.h5", "data4")] * (1 - r) + EXPECTED[("meshio_T.h5", "data5")] * r
    assert np.allclose(hist.at(t), want, rtol=1e-15, atol=0)


def test_csv_to_xdmf_writes_hdf_by_default(tmp_path):
    csv = tmp_path / "dump.csv"
    lines = ["thermal solver export", "origin", "0,0", "spacing", "1e-6,1e-6", "numpoints", "3,2",
             "time, temperature", "0.0,1.0,1.1,1.2,1.3,1.4,1.5", "2.5,1.5,1.6,1.7,1.8,1.9,2.0"]
    csv.write_text("\n".join(lines) + "\n")
    out = str(tmp_path / "T.xdmf")
    thermal.CSVtoXDMF(str(csv), xdmf_path=out)
    assert (tmp_path / "T.h5").exists()
    series = thermal.read_xdmf_series(out)
    assert [t for t, _ in series] == [0.0, 2.5 * 0.000001]  # the reference scales microseconds so
    np.testing.assert_array_equal(series[1][1], np.transpose(1000 * np.array([[1.5, 1.6], [1.7, 1.8],
                                                                              [1.9, 2.0]])))
