"""Write the HDF5 fixtures of tests/test_hdf5_cpu.py with the HDF5 library
itself (h5py), so that the pure-Python reader is checked against files that
libhdf5 wrote, not against a writer sharing the reader's reading of the
format.

Run with an interpreter that has h5py (it is not a dependency of the
package or of the tests):

    python3.9 tests/golden/hdf5/make_fixtures.py

Outputs, next to this script:
  meshio_T.h5 / meshio_T.xdmf  a thermal history in the layout of meshio's
                               XDMF TimeSeriesWriter (its default HDF heavy
                               data): root datasets data0 (mesh points),
                               data1..data12 (frames), the XDMF items naming
                               "meshio_T.h5:/dataK";
  earliest.h5                  default file format: contiguous, compact and
                               chunked datasets (v1 B-tree chunk index) with
                               deflate / shuffle / fletcher32 filters,
                               partial edge chunks, unwritten chunks (fill
                               value), integer and big-endian types, nested
                               groups;
  latest.h5                    libver="latest": version-2 object headers,
                               layout version 4 (single-chunk, implicit and
                               fixed-array chunk indexes), a group with 12
                               links (dense link storage: fractal heap and
                               version-2 B-tree name index), and one dataset
                               with an unlimited dimension (extensible-array
                               index, which the reader refuses);
  expected.npz                 every dataset as h5py reads it, keyed
                               "<file>|<path with / as |>";
  writer.json                  SHA-256 of files written by the package's own
                               HDF5 writer (pyphasefield_amd/hdf5.py,
                               write_datasets) from writer_cases(), after
                               h5py has read every dataset of them back
                               unchanged: the test rewrites the same cases
                               and compares the digests, so the bytes it
                               produces are bytes libhdf5 has read.
"""

import hashlib
import importlib.util
import json
import os
import xml.etree.ElementTree as ET

import h5py
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _key(fname, path):
    return fname + "|" + path.strip("/").replace("/", "|")


def meshio_series(expected):
    rng = np.random.default_rng(1)
    fname = "meshio_T.h5"
    times = [0.0] + list(np.cumsum(rng.uniform(1e-6, 3e-6, 11)))
    frames = [1500.0 + 300.0 * rng.random((9, 13)) for _ in times]
    xdmf = ET.Element("Xdmf", Version="3.0")
    domain = ET.SubElement(xdmf, "Domain")
    with h5py.File(os.path.join(HERE, fname), "w") as h5:
        points = np.array([[[0, 0], [0, 0]], [[0, 0], [0, 0]]])
        h5.create_dataset("data0", data=points)
        expected[_key(fname, "data0")] = points
        mesh = ET.SubElement(domain, "Grid", Name="mesh", GridType="Uniform")
        geo = ET.SubElement(mesh, "Geometry", GeometryType="XY")
        ET.SubElement(geo, "DataItem", DataType="Int", Dimensions="2 2 2", Format="HDF",
                      Precision="8").text = fname + ":/data0"
        coll = ET.SubElement(domain, "Grid", Name="TimeSeries_meshio", GridType="Collection",
                             CollectionType="Temporal")
        for k, (t, T) in enumerate(zip(times, frames), start=1):
            name = "data%d" % k
            h5.create_dataset(name, data=T)
            expected[_key(fname, name)] = T
            g = ET.SubElement(coll, "Grid")
            ET.SubElement(g, "{http://www.w3.org/2003/XInclude}include",
                          xpointer='xpointer(//Grid[@Name="mesh"]/*[self::Topology or self::Geometry])')
            ET.SubElement(g, "Time", Value=repr(float(t)))
            at = ET.SubElement(g, "Attribute", Name="T", AttributeType="Matrix", Center="Node")
            ET.SubElement(at, "DataItem", DataType="Float", Dimensions="9 13", Format="HDF",
                          Precision="8").text = fname + ":/" + name
    ET.ElementTree(xdmf).write(os.path.join(HERE, "meshio_T.xdmf"), xml_declaration=True, encoding="utf-8")
    np.save(os.path.join(HERE, "meshio_T_times.npy"), np.array(times))


def earliest(expected):
    rng = np.random.default_rng(2)
    fname = "earliest.h5"
    with h5py.File(os.path.join(HERE, fname), "w") as h5:
        def put(path, data, **kw):
            ds = h5.create_dataset(path, data=data, **kw)
            expected[_key(fname, path)] = ds[()]

        put("contiguous_f8", rng.random((12, 17)))
        put("contiguous_f4_1d", rng.random(40).astype(np.float32))
        put("contiguous_i4_be", rng.integers(-1000, 1000, (6, 7)).astype(">i4"))
        put("chunked_plain", rng.random((12, 17)), chunks=(5, 7))
        put("gzip4", rng.random((12, 17)), chunks=(5, 7), compression="gzip", compression_opts=4)
        put("gzip_shuffle_f4", rng.random((10, 11, 9)).astype(np.float32), chunks=(4, 4, 4),
            compression="gzip", shuffle=True)
        put("gzip_shuffle_fletcher", rng.random((33, 20)), chunks=(8, 8), compression="gzip",
            shuffle=True, fletcher32=True)
        put("group/sub/deep", rng.random((3, 4)))
        # unwritten chunks read as the fill value
        ds = h5.create_dataset("sparse", shape=(20, 20), dtype="f8", chunks=(6, 6), fillvalue=-2.5)
        ds[0:6, 0:6] = rng.random((6, 6))
        ds[13:20, 7:11] = rng.random((7, 4))
        expected[_key(fname, "sparse")] = ds[()]
        # compact layout through the low-level interface
        space = h5py.h5s.create_simple((3, 5))
        dcpl = h5py.h5p.create(h5py.h5p.DATASET_CREATE)
        dcpl.set_layout(h5py.h5d.COMPACT)
        dsid = h5py.h5d.create(h5.id, b"compact", h5py.h5t.NATIVE_DOUBLE, space, dcpl=dcpl)
        vals = rng.random((3, 5))
        dsid.write(h5py.h5s.ALL, h5py.h5s.ALL, vals)
        expected[_key(fname, "compact")] = vals
        for k in range(20):  # enough names for several symbol-table nodes
            put("many/x%02d" % k, np.full((2, 3), float(k)))


def latest(expected):
    rng = np.random.default_rng(3)
    fname = "latest.h5"
    with h5py.File(os.path.join(HERE, fname), "w", libver="latest") as h5:
        def put(path, data, **kw):
            ds = h5.create_dataset(path, data=data, **kw)
            expected[_key(fname, path)] = ds[()]

        put("contiguous_f8", rng.random((12, 17)))
        put("single_chunk_gzip", rng.random((12, 17)), chunks=(12, 17), compression="gzip")
        put("fixed_array_gzip_shuffle", rng.random((30, 41)), chunks=(7, 9), compression="gzip", shuffle=True)
        for k in range(12):
            put("dense/frame%02d" % k, rng.random((4, 5)) + k)
        # implicit chunk index: no filters, fixed size, space allocated early
        space = h5py.h5s.create_simple((16, 10))
        dcpl = h5py.h5p.create(h5py.h5p.DATASET_CREATE)
        dcpl.set_chunk((4, 5))
        dcpl.set_alloc_time(h5py.h5d.ALLOC_TIME_EARLY)
        dsid = h5py.h5d.create(h5.id, b"implicit", h5py.h5t.NATIVE_DOUBLE, space, dcpl=dcpl)
        vals = rng.random((16, 10))
        dsid.write(h5py.h5s.ALL, h5py.h5s.ALL, vals)
        expected[_key(fname, "implicit")] = vals
        ds = h5.create_dataset("unlimited", data=rng.random((6, 4)), maxshape=(None, 4), chunks=(2, 4))
        expected[_key(fname, "unlimited")] = ds[()]


def writer_cases():
    """{case: {name: array}}; the values are exact in binary so that every
    numpy version makes the same bytes."""
    many = {"data%d" % k: (np.arange(12, dtype=np.float64).reshape(3, 4) + 0.25 * k) for k in range(41)}
    many["data0"] = np.array([[[0, 0], [0, 0]], [[0, 0], [0, 0]]])
    mixed = {"f8": np.arange(30, dtype=np.float64).reshape(5, 6) / 8.0 - 1.5,
             "f4": np.arange(7, dtype=np.float32) * 0.5,
             "i4": np.arange(-3, 9, dtype=np.int32).reshape(3, 4),
             "u2": np.arange(5, dtype=np.uint16),
             "be_f8": (np.arange(6, dtype=np.float64) * 0.125).astype(">f8"),
             "three_d": np.arange(24, dtype=np.float64).reshape(2, 3, 4)}
    return {"many": many, "mixed": mixed, "empty": {}}


def _package_hdf5():
    path = os.path.join(HERE, "..", "..", "..", "pyphasefield_amd", "hdf5.py")
    spec = importlib.util.spec_from_file_location("ppf_hdf5", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def writer_digests():
    mod = _package_hdf5()
    out = {}
    for case, arrays in writer_cases().items():
        path = os.path.join("/tmp", "ppf_writer_%s.h5" % case)
        mod.write_datasets(path, arrays)
        with h5py.File(path, "r") as f:
            assert sorted(f.keys()) == sorted(arrays), case
            for name, a in arrays.items():
                got = f[name][()]
                same_type = got.dtype.newbyteorder("=") == a.dtype.newbyteorder("=")
                assert got.shape == a.shape and same_type and np.array_equal(got, a), (case, name)
        with open(path, "rb") as f:
            out[case] = hashlib.sha256(f.read()).hexdigest()
        os.remove(path)
    with open(os.path.join(HERE, "writer.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def main():
    expected = {}
    meshio_series(expected)
    earliest(expected)
    latest(expected)
    np.savez(os.path.join(HERE, "expected.npz"), **expected)
    writer_digests()
    print(h5py.version.info)


if __name__ == "__main__":
    main()
