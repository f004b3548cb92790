"""Thermal histories (XDMF_FILE temperature): the XDMF reader/writer, the
CSV conversion of ``ppf_utils.py:63-110``, the oracle's time interpolation
(``simulation.py:157-177,316-330``) and the host side of the Simulation
wiring.  No GPU."""

import numpy as np
import pytest

from oracle import engines as oe
from pyphasefield_amd import thermal


def _frames(shape, n, seed=0):
    rng = np.random.default_rng(seed)
    return [1500.0 + 50.0 * rng.random(shape) for _ in range(n)]


@pytest.mark.parametrize("fmt", ["XML", "Binary", "HDF"])
def test_xdmf_round_trip(tmp_path, fmt):
    times = [0.0, 1.25e-7, 3.5e-7]
    frames = _frames((4, 6), 3)
    path = str(tmp_path / "T.xdmf")
    thermal.write_xdmf_series(path, times, frames, data_format=fmt)
    ld = thermal.XDMFLoader(path)
    assert ld.num_steps == 3
    for k in range(3):
        t, a = ld.data(k)
        assert t == times[k]
        assert a.shape == (4, 6)
        assert np.array_equal(a, frames[k])  # repr() text and raw doubles are exact


def test_xdmf_reader_meshio_layout(tmp_path):
    """The layout meshio's TimeSeriesWriter emits: a mesh grid first, an
    xi:include in each step grid, Float/4 XML data and Binary data with an
    explicit byte order and offset; singleton axes are squeezed away."""
    big = (np.arange(6.0).reshape(1, 2, 3) * 0.5 + 1000.0).astype(">f8")
    raw = b"\x00" * 16 + big.tobytes()
    (tmp_path / "T1.bin").write_bytes(raw)
    text = """<?xml version="1.0"?>
<Xdmf Version="3.0" xmlns:xi="http://www.w3.org/2001/XInclude">
  <Domain>
    <Grid Name="mesh" GridType="Uniform">
      <Geometry GeometryType="XYZ"><DataItem DataType="Float" Dimensions="1 3" Format="XML" Precision="8">0 0 0</DataItem></Geometry>
    </Grid>
    <Grid Name="TimeSeries_meshio" GridType="Collection" CollectionType="Temporal">
      <Grid>
        <xi:include xpointer="xpointer(//Grid[@Name=&quot;mesh&quot;]/*[self::Topology or self::Geometry])" />
        <Time Value="0.0" />
        <Attribute Name="T" AttributeType="Scalar" Center="Node">
          <DataItem DataType="Float" Dimensions="1 2 3" Format="XML" Precision="4">1 2 3 4 5 6.5</DataItem>
        </Attribute>
      </Grid>
      <Grid>
        <xi:include xpointer="xpointer(//Grid[@Name=&quot;mesh&quot;]/*[self::Topology or self::Geometry])" />
        <Time Value="2e-06" />
        <Attribute Name="other" AttributeType="Scalar" Center="Node">
          <DataItem DataType="Float" Dimensions="1" Format="XML" Precision="8">7</DataItem>
        </Attribute>
        <Attribute Name="T" AttributeType="Scalar" Center="Node">
          <DataItem DataType="Float" Dimensions="1 2 3" Format="Binary" Precision="8" Endian="Big" Seek="16">T1.bin</DataItem>
        </Attribute>
      </Grid>
    </Grid>
  </Domain>
</Xdmf>
"""
    path = tmp_path / "T.xdmf"
    path.write_text(text)
    ld = thermal.XDMFLoader(str(path))
    assert ld.num_steps == 2
    t0, a0 = ld.data(0)
    assert t0 == 0.0 and a0.dtype == np.float32
    assert np.array_equal(a0, np.array([[1, 2, 3], [4, 5, 6.5]], dtype=np.float32))
    t1, a1 = ld.data(1)
    assert t1 == 2e-06 and a1.dtype == np.float64
    assert np.array_equal(a1, np.array([[1000.0, 1000.5, 1001.0], [1001.5, 1002.0, 1002.5]]))


def test_xdmf_hdf_heavy_data_missing_file(tmp_path):
    """An HDF item whose file is absent fails when its frame is read."""
    path = tmp_path / "T.xdmf"
    path.write_text("""<Xdmf Version="3.0"><Domain><Grid GridType="Collection" CollectionType="Temporal">
<Grid><Time Value="0"/><Attribute Name="T"><DataItem DataType="Float" Precision="8" Dimensions="2"
Format="HDF">T.h5:/data0</DataItem></Attribute></Grid></Grid></Domain></Xdmf>""")
    ld = thermal.XDMFLoader(str(path))
    assert ld.num_steps == 1
    with pytest.raises(FileNotFoundError):
        ld.data(0)


def _write_csv(path, rows, dims="3,2"):
    lines = ["thermal solver export", "origin", "0,0", "spacing", "1e-6,1e-6", "numpoints", dims,
             "time, temperature"]
    lines += rows
    path.write_text("\n".join(lines) + "\n")


def test_csv_to_xdmf_layout_and_units(tmp_path):
    """Point counts "3,2": the 6 values fill a (3, 2) array row by row, which
    is transposed to (2, 3); values in kK become K, times in us become s."""
    csv = tmp_path / "T.csv"
    _write_csv(csv, ["0.5, 1,2,3,4,5,6", "1.5, 2,2,2,2,2,2,"])
    out = str(tmp_path / "T.xdmf")
    thermal.CSVtoXDMF(str(csv), xdmf_path=out)
    ld = thermal.XDMFLoader(out)
    assert ld.num_steps == 2
    t, a = ld.data(0)
    assert t == pytest.approx(0.5e-6, rel=1e-15)
    assert np.array_equal(a, np.array([[1000.0, 3000.0, 5000.0], [2000.0, 4000.0, 6000.0]]))
    t, a = ld.data(1)
    assert t == pytest.approx(1.5e-6, rel=1e-15)
    assert np.array_equal(a, np.full((2, 3), 2000.0))


def test_csv_to_xdmf_cutoffs_and_reflection(tmp_path):
    """With cutoffs the series starts (re-timed to 0) at the first row whose
    maximum exceeds starting_T and stops before the first later row whose
    maximum falls below ending_T; rows after that are never read.  The
    reflection prepends each frame's mirror image along its second axis."""
    csv = tmp_path / "T.csv"
    rows = ["1, 0.1,0.1,0.1,0.1,0.1,0.1",      # max 100 K: before the start
            "2, 0.1,0.1,0.1,0.1,0.1,2.0",      # max 2000 K: start, t = 0
            "3, 1.0,1.0,1.0,1.0,1.0,1.5",      # 1500 K: kept
            "4, 0.2,0.2,0.2,0.2,0.2,0.2",      # 200 K < ending_T: stop
            "5, 3.0,3.0,3.0,3.0,3.0,3.0"]      # never written
    _write_csv(csv, rows)
    out = str(tmp_path / "T.xdmf")
    thermal.CSVtoXDMF(str(csv), T_cutoffs=True, starting_T=1800.0, ending_T=1000.0, reflect_X=True,
                      xdmf_path=out, data_format="Binary")
    ld = thermal.XDMFLoader(out)
    assert ld.num_steps == 2
    t0, a0 = ld.data(0)
    t1, a1 = ld.data(1)
    assert t0 == 0.0
    assert t1 == pytest.approx(1e-6, rel=1e-12)
    base = np.array([[100.0, 100.0, 100.0], [100.0, 100.0, 2000.0]])
    assert np.array_equal(a0, np.concatenate([base[:, ::-1], base], axis=1))
    assert a0.shape == (2, 6)
    assert a1[1, 0] == 1500.0 and a1[1, -1] == 1500.0


# ---------------------------------------------------------- oracle pinned --
@pytest.mark.parametrize("form", ["kernel", "host"])
def test_oracle_history_hits_frames_and_reproduces_linear_histories(form):
    """Closed forms: at a frame's time the interpolation returns that frame;
    a history sampled from T(t) = A + B*t (any frame spacing) is reproduced
    at every t in between (a swapped weight or a wrong denominator is not);
    the window before the end is used up to and including the last frame
    time, and a time past it is an error."""
    rng = np.random.default_rng(3)
    A, B = 1500.0 + rng.random((3, 4)), rng.random((3, 4)) * 1e6
    times = [0.0, 1e-6, 1.7e-6, 4e-6, 4.1e-6]
    frames = [A + B * t for t in times]
    for k, t in enumerate(times):
        assert np.allclose(oe.temperature_history(times, frames, t, form), frames[k], rtol=0, atol=1e-12)
    for t in np.linspace(0.0, 4.1e-6, 37):
        assert np.allclose(oe.temperature_history(times, frames, t, form), A + B * t, rtol=1e-14, atol=0)
    with pytest.raises(ValueError):
        oe.temperature_history(times, frames, 4.2e-6, form)


@pytest.mark.parametrize("form", ["kernel", "host"])
def test_oracle_history_uses_the_bracketing_window(form):
    """A non-linear history: between t_k and t_k+1 only those two frames
    contribute (a planted window shift would mix in a third)."""
    times = [0.0, 1.0, 2.0, 3.0]
    frames = [np.full((2, 2), v) for v in (0.0, 10.0, 0.0, 40.0)]
    assert oe.temperature_history(times, frames, 0.5, form)[0, 0] == 5.0
    assert oe.temperature_history(times, frames, 1.0, form)[0, 0] == 10.0
    assert oe.temperature_history(times, frames, 1.25, form)[0, 0] == 7.5
    assert oe.temperature_history(times, frames, 2.75, form)[0, 0] == 30.0


def test_oracle_history_forms_agree_to_rounding():
    times = [0.0, 3e-7, 1.1e-6]
    frames = _frames((6, 7), 3, seed=4)
    for t in np.linspace(0.0, 1.1e-6, 29):
        k = oe.temperature_history(times, frames, t, "kernel")
        h = oe.temperature_history(times, frames, t, "host")
        assert np.abs(k - h).max() <= 4 * np.spacing(2000.0)


def test_thermal_history_window_matches_oracle(tmp_path):
    times = [0.0, 2e-7, 3e-7, 7e-7, 8e-7]
    frames = _frames((5, 3), 5, seed=1)
    path = str(tmp_path / "T.xdmf")
    thermal.write_xdmf_series(path, times, frames)
    h = thermal.ThermalHistory(path, (5, 3))
    for t in np.linspace(0.0, 8e-7, 41):
        h.advance_to(t)
        want = oe.temperature_history(times, frames, t)
        assert np.array_equal(h.at(t), want)
    lo, hi = h.range()
    assert lo == min(f.min() for f in frames) and hi == max(f.max() for f in frames)
    with pytest.raises(ValueError, match="ends"):
        h.advance_to(9e-7)
    with pytest.raises(ValueError, match="shape"):
        thermal.ThermalHistory(path, (3, 5))


def test_simulation_initial_temperature_from_history(tmp_path):
    """simulation.py:157-177: the default path is T.xdmf; the field holds the
    interpolation at counter*dt, halos by the boundary rule."""
    from pyphasefield_amd import Simulation
    times = [0.0, 1e-6, 2e-6, 3e-6]
    frames = _frames((4, 5), 4, seed=2)
    thermal.write_xdmf_series(str(tmp_path / "T.xdmf"), times, frames)
    import os
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        sim = Simulation(dimensions=[4, 5], dt=1.3e-7, initial_time_step=9, temperature_type="XDMF_FILE",
                         boundary_conditions="NEUMANN")
        sim.init_temperature_field()
    finally:
        os.chdir(cwd)
    assert sim._temperature_path == "T.xdmf"
    want = oe.temperature_history(times, frames, 9 * 1.3e-7)
    T = sim.temperature._data
    assert np.array_equal(T[1:-1, 1:-1], want)
    assert np.array_equal(T[0, 1:-1], T[1, 1:-1]) and np.array_equal(T[1:-1, -1], T[1:-1, -2])
    assert sim._thermal.bounds == [1e-6, 2e-6]
    assert sim._graph_chunk() == 0


# ------------------------------------------------- linear gradients, 2D/3D --
@pytest.mark.parametrize("shape", [(5, 7), (4, 5, 7)])
def test_oracle_linear_gradient_closed_form(shape):
    """T(k, j, i) = T0 + dx*(i*dTdx + j*dTdy + k*dTdz) + step*dt*dTdt, with
    every axis given a different slope so that a swapped axis shows."""
    dx, dt, step = 2e-6, 3e-7, 11
    dTdx, dTdy, dTdz, dTdt = 1.0e5, -3.0e5, 7.0e5, -2.0e6
    T = oe.temperature_linear(shape, 1600.0, dTdx, dTdy, dx, dTdt, dt, step, dTdz=dTdz)
    idx = np.indices(shape).astype(np.float64)
    want = 1600.0 + dx * (idx[-1] * dTdx + idx[-2] * dTdy) + step * dt * dTdt
    if len(shape) == 3:
        want = want + dx * idx[0] * dTdz
    assert T.shape == shape
    assert np.allclose(T, want, rtol=1e-14, atol=0)


def test_simulation_3d_linear_gradient_host_field():
    """simulation.py:134-156 in 3D: the host field follows the same formula."""
    from pyphasefield_amd import Simulation
    sim = Simulation(dimensions=[4, 5, 6], dx=2e-6, dt=3e-7, initial_time_step=5, temperature_type="LINEAR_GRADIENT",
                     initial_T=1600.0, dTdx=1.0e5, dTdy=-3.0e5, dTdz=7.0e5, dTdt=-2.0e6,
                     boundary_conditions="NEUMANN")
    sim.init_temperature_field()
    want = oe.temperature_linear((4, 5, 6), 1600.0, 1.0e5, -3.0e5, 2e-6, -2.0e6, 3e-7, 5, dTdz=7.0e5)
    assert np.array_equal(sim.temperature._data[1:-1, 1:-1, 1:-1], want)
