"""DiffusionGPU engine (class DiffusionGPU, Engines/DiffusionGPU.py:76-117 of
the reference): explicit diffusion of one field c, D = user_data["D"]
(default 0.1), c = 1 in the middle half of every axis.

The reference's kernels (DiffusionGPU.py:18-75) omit dt, use -4 instead of
-6 for the 3D centre weight and read one cell past the array; they cannot
run.  This class runs the explicit step of ``Diffusion`` (c <- c + dt*D*lap c,
Diffusion.py:198-208) through the same fused HIP stencil, so the reference
scripts that construct ``DiffusionGPU`` work unchanged."""

from .Diffusion import Diffusion


class DiffusionGPU(Diffusion):
    _DEFAULT_USER_DATA = {"D": 0.1}

    def __init__(self, **kwargs):
        super().__init__(**kwargs)
        self._framework = "GPU_SERIAL"  # DiffusionGPU.py:82 (uses_gpu = True)

    def _graph_chunk(self):
        return super(Diffusion, self)._graph_chunk()

    def simulation_loop(self):
        if self.user_data.get("solver", "explicit") != "explicit":
            raise ValueError("DiffusionGPU is explicit only (DiffusionGPU.py:116-117); use Diffusion for the "
                             "implicit and Crank-Nicolson solvers")
        super().simulation_loop()
