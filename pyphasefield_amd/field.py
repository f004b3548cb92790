"""Field: a halo-padded array (field.py:4-73 of the reference).

Storage layout is the reference's: +1 halo cell per axis, axis order
(z, y, x), ``data`` is the full padded array and ``get_cells()`` the interior
view.  While a simulation runs, the authoritative copy lives on the GPU; the
``data`` property downloads it on first access after a ``simulate`` call and
marks the host copy as possibly modified, so the next ``simulate`` uploads
it -- the reference re-uploads host data on every ``simulate`` call
(simulation.py:259-261 with ppf_gpu_utils.py:396-420), this keeps that
semantics without the per-call transfer when nobody touched the data.
"""

import numpy as np


class Field:
    def __init__(self, data=None, simulation=None, name=None, colormap="GnBu", dtype=None):
        data = np.asarray(data)
        dim = [s + 2 for s in data.shape]
        self._slice = tuple(slice(1, -1) for _ in data.shape)
        full = np.zeros(dim, dtype=dtype or np.float64)
        full[self._slice] += data
        self._data = full
        self.name = name
        self._simulation = simulation
        self.colormap = colormap

    # -- host/device coherence -------------------------------------------
    @property
    def data(self):
        sim = self._simulation
        if sim is not None:
            sim._field_host_access(self)
        return self._data

    @data.setter
    def data(self, value):
        self._data = np.asarray(value)
        sim = self._simulation
        if sim is not None:
            sim._field_host_access(self)

    def __str__(self):
        return self.data.__str__()

    # -- reference helpers (field.py:26-73) --------------------------------
    def gradient_cell(self):
        d = self.data
        inv2 = 0.5 / self._simulation.get_cell_spacing()
        return np.stack([(np.roll(d, -1, i) - np.roll(d, 1, i)) * inv2 for i in range(d.ndim)])

    def gradient_face_left(self):
        d = self.data
        inv = 1. / self._simulation.get_cell_spacing()
        return np.stack([(d - np.roll(d, 1, i)) * inv for i in range(d.ndim)])

    def gradient_face_right(self):
        d = self.data
        inv = 1. / self._simulation.get_cell_spacing()
        return np.stack([(np.roll(d, -1, i) - d) * inv for i in range(d.ndim)])

    def laplacian(self):
        d = self.data
        idx2 = (1. / self._simulation.get_cell_spacing()) ** 2
        r = (np.roll(d, -1, 0) + np.roll(d, 1, 0) - 2 * d) * idx2
        for i in range(1, d.ndim):
            r += (np.roll(d, -1, i) + np.roll(d, 1, i) - 2 * d) * idx2
        return r

    def get_cells(self):
        return self.data[self._slice]

    def get_all_cells(self):
        return self.data
