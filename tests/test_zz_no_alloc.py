"""ppf_reduce_noalloc (partial sums in a file-scope device array, no
allocation) against ppf_reduce (caller's workspace): FP64 mass and
Cahn-Hilliard free energy, two calls in a row, the second on a smaller
input, and the rejection of inputs beyond the array's size."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _stack(nf, shape, dtype, seed):
    import torch
    from pyphasefield_amd import _lib
    t = _lib.alloc_stack(nf, shape, dtype, "cuda")
    rng = np.random.default_rng(seed)
    t.copy_(torch.from_numpy(rng.random(tuple(t.shape))))
    return t


def _reduce(entry, t, bc, kind, eps):
    import torch
    from pyphasefield_amd import _lib
    L = _lib.lib()
    g = _lib.make_grid(t, t.dim() - 1, bc, 0.7)
    nf = t.shape[0] if kind == 0 else 1
    out = torch.zeros(nf, dtype=torch.float64, device="cuda")
    s = _lib.stream_handle()
    if entry == "noalloc":
        rc = L.ppf_reduce_noalloc(g, _lib.ptr(t), None, kind, eps, _lib.ptr(out), s)
    else:
        work = torch.empty(int(L.ppf_reduce_work_bytes(g)), dtype=torch.uint8, device="cuda")
        rc = L.ppf_reduce(g, _lib.ptr(t), None, kind, eps, _lib.ptr(out), _lib.ptr(work), s)
    _lib.check(rc, entry)
    return out.cpu().numpy()


@pytest.mark.parametrize("dtype_name", ["float64", "float32"])
def test_noalloc_equals_workspace_entry_point(gpu, dtype_name):
    import torch
    dtype = getattr(torch, dtype_name)
    big = _stack(5, [64, 96, 40], dtype, 1)
    small = _stack(3, [17, 33], dtype, 2)
    for t, bc in ((big, ["PERIODIC", "NEUMANN", "DIRCHLET"]), (small, "NEUMANN")):
        for kind in (0, 1):
            want = _reduce("workspace", t, bc, kind, 1.3)
            got = _reduce("noalloc", t, bc, kind, 1.3)
            # same kernels, same partial-sum order: equal bit for bit
            np.testing.assert_array_equal(got, want)
    # a first call leaves partial sums of the big input behind; the small
    # call after it must not see them
    assert np.all(np.isfinite(_reduce("noalloc", small, "PERIODIC", 0, 1.0)))


def test_noalloc_rejects_too_many_fields(gpu):
    import torch
    from pyphasefield_amd import _lib
    L = _lib.lib()
    t = _stack(17, [8, 8], torch.float64, 3)
    g = _lib.make_grid(t, 2, "PERIODIC", 1.0)
    out = torch.zeros(17, dtype=torch.float64, device="cuda")
    rc = L.ppf_reduce_noalloc(g, _lib.ptr(t), None, 0, 1.0, _lib.ptr(out), _lib.stream_handle())
    assert rc != 0
    assert b"at most 16 fields" in L.ppf_last_error()
    # the free energy reads field 0 only: allowed on the same stack
    rc = L.ppf_reduce_noalloc(g, _lib.ptr(t), None, 1, 1.0, _lib.ptr(out), _lib.stream_handle())
    assert rc == 0
