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


def multistep(sim, fn_name, nsteps, *scalars, nactive=None, mx=None):
    """nsteps steps with the engine's several-steps-per-launch kernel, in an
    even number of launches (the two buffers end where they started) of at
    most mx steps each (default: ppf_multistep_max).  False (nothing
    launched) when the grid does not allow it."""
    L = _lib.lib()
    if mx is None:
        mx = L.ppf_multistep_max(sim._grid, _lib.ptr(sim._boundary_conditions_gpu_device))
    if mx < 2:
        return False
    launches = -(-nsteps // mx)
    launches += launches % 2
    if launches > nsteps:
        return False
    fn = getattr(L, fn_name)
    base, extra = divmod(nsteps, launches)
    for i in range(launches):
        k = base + (1 if i < extra else 0)
        rc = fn(sim._grid, _lib.ptr(sim._fields_gpu_device), _lib.ptr(sim._fields_out_gpu_device),
                _lib.ptr(sim._boundary_conditions_gpu_device), *scalars, k, _lib.stream_handle())
        _lib.check(rc, fn_name)
        carry_passive(sim, nactive)
        sim._swap_buffers()
    return True


def carry_passive(sim, nactive):
    """Fields an engine does not update are copied so they survive the swap."""
    F = sim._fields_gpu_device.shape[0]
    if nactive is not None and F > nactive:
        sim._fields_out_gpu_device[nactive:].copy_(sim._fields_gpu_device[nactive:])
