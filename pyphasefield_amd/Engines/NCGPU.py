"""The legacy N-component engine (class NCGPU, Engines/NCGPU.py:329 of the
reference) is not provided: its helper kernel takes one derivative, for
component 0 only, and applies it to every component (NCGPU.py:287-303), and
its RNG states are indexed racily (:42, :253).  NCGPU_new is the maintained
engine with the same fields and parameters."""

from ..simulation import Simulation


class NCGPU(Simulation):
    def __init__(self, **kwargs):
        raise NotImplementedError("the legacy NCGPU engine (Engines/NCGPU.py) is not provided: it applies the "
                                  "component-0 free-energy derivative to every component (NCGPU.py:287-303); "
                                  "use NCGPU_new, which takes the same user_data")
