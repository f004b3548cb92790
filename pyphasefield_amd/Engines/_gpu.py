"""Shared plumbing of the engine classes: one call into libppf_hip.so per
step on torch's current stream, then a buffer swap (the reference swaps its
numba device arrays the same way, Engines/NCGPU_new.py:553-554)."""

from .. import _lib


def stencil_step(sim, fn_name, *scalars, nactive=None):
    """in -> out with the engine kernel, carry passive fields, swap."""
    L = _lib.lib()
    fn = getattr(L, fn_name)
    rc = fn(sim._grid, _lib.ptr(sim._fields_gpu_device), _lib.ptr(sim._fields_out_gpu_device),
            _lib.ptr(sim._boundary_conditions_gpu_device), *scalars, _lib.stream_handle())
    _lib.check(rc, fn_name)
    carry_passive(sim, nactive)
    sim._swap_buffers()


def carry_passive(sim, nactive):
    """Fields an engine does not update are copied so they survive the swap."""
    F = sim._fields_gpu_device.shape[0]
    if nactive is not None and F > nactive:
        sim._fields_out_gpu_device[nactive:].copy_(sim._fields_gpu_device[nactive:])
