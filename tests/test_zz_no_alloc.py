"""td3_direct_noalloc, the direct sum for callers that cannot allocate: through the C ABI,
its results equal td3_direct's (same kernels and chunking, so bit for bit) on two calls in
a row, the second smaller, with the library's own scratch in place of a workspace; sizes
beyond that scratch are refused before any launch."""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _inputs(m, n, seed):
    rng = np.random.default_rng(seed)
    pos = torch.tensor(rng.uniform(-1, 1, (2, n)), device="cuda")
    q = torch.tensor(rng.choice([-1.0, 1.0], n), device="cuda")
    targets = torch.tensor(rng.uniform(-1, 1, (2, m)), device="cuda")
    exclude = torch.tensor(rng.integers(-1, n, m), device="cuda")
    return targets, exclude, pos, q


def _call(abi, name, kernel, prec, targets, exclude, pos, q, eps, workspace=None):
    m, n = targets.shape[1], pos.shape[1]
    phi = torch.full((m,), np.nan, dtype=torch.float64, device="cuda")
    field = torch.full((2, m), np.nan, dtype=torch.float64, device="cuda")
    status = torch.zeros(1, dtype=torch.int64, device="cuda")
    args = [kernel, prec, abi.ptr(targets), m, abi.ptr(exclude), abi.ptr(pos), n, abi.ptr(q),
            eps, abi.ptr(phi), abi.ptr(field), abi.ptr(status)]
    if workspace is not None:
        args += [abi.ptr(workspace), workspace.numel()]
    code = getattr(abi.lib(), name)(*args, abi.stream())
    torch.cuda.synchronize()
    return code, phi, field, status


def test_noalloc_direct_equals_direct_on_two_calls(td3):
    from tree_d_3d_amd import _abi

    scratch = _abi.lib().td3_direct_noalloc_scratch_bytes()
    assert scratch == 16 << 20
    # the first call uses many source chunks (the scratch holds 128 x 4096 partials), the
    # second fewer targets and sources
    for case, (m, n, seed) in enumerate(((4096, 300_000, 1), (1000, 20_000, 2))):
        need = _abi.lib().td3_direct_workspace_bytes(m, n)
        assert 0 < need <= scratch
        targets, exclude, pos, q = _inputs(m, n, seed)
        for kernel in (_abi.LOG, _abi.COULOMB):
            for prec in (_abi.FP64, _abi.MIXED):
                print(f"noalloc case {case} m={m} n={n} kernel={kernel} prec={prec}",
                      flush=True)
                ws = torch.empty(need, dtype=torch.uint8, device="cuda")
                code, phi, field, status = _call(_abi, "td3_direct", kernel, prec, targets,
                                                 exclude, pos, q, 1e-3, ws)
                assert code == _abi.TD3_OK
                code2, phi2, field2, status2 = _call(_abi, "td3_direct_noalloc", kernel, prec,
                                                     targets, exclude, pos, q, 1e-3)
                assert code2 == _abi.TD3_OK, _abi.last_error()
                assert int(status2[0]) == int(status[0]) == 0
                assert torch.equal(phi2, phi) and torch.equal(field2, field)


def test_noalloc_direct_single_chunk_and_empty(td3):
    """A call that needs no scratch (one chunk) and an empty source set."""
    from tree_d_3d_amd import _abi

    m, n = 600_000, 700
    assert _abi.lib().td3_direct_workspace_bytes(m, n) == 0
    targets, exclude, pos, q = _inputs(m, n, 3)
    print(f"noalloc single chunk m={m} n={n}", flush=True)
    code, phi, field, _ = _call(_abi, "td3_direct", _abi.LOG, _abi.FP64, targets, exclude, pos,
                                q, 1e-3, torch.empty(8, dtype=torch.uint8, device="cuda"))
    code2, phi2, field2, _ = _call(_abi, "td3_direct_noalloc", _abi.LOG, _abi.FP64, targets,
                                   exclude, pos, q, 1e-3)
    assert code == code2 == _abi.TD3_OK
    assert torch.equal(phi2, phi) and torch.equal(field2, field)
    empty = pos[:, :0].contiguous()
    code3, phi3, field3, _ = _call(_abi, "td3_direct_noalloc", _abi.LOG, _abi.FP64,
                                   targets[:, :5].contiguous(), exclude[:5].contiguous(), empty,
                                   q[:0].contiguous(), 1e-3)
    assert code3 == _abi.TD3_OK
    assert torch.equal(phi3, torch.zeros_like(phi3)) and torch.equal(field3,
                                                                     torch.zeros_like(field3))


def test_noalloc_direct_refuses_sizes_beyond_its_scratch(td3):
    from tree_d_3d_amd import _abi

    m, n = 400_000, 5000  # two source chunks of 400,000 targets: 19.2 MB of partials
    assert _abi.lib().td3_direct_workspace_bytes(m, n) > _abi.lib().td3_direct_noalloc_scratch_bytes()
    targets, exclude, pos, q = _inputs(m, n, 4)
    print(f"noalloc refusal m={m} n={n}", flush=True)
    code, phi, _, _ = _call(_abi, "td3_direct_noalloc", _abi.COULOMB, _abi.FP64, targets, exclude,
                            pos, q, 1e-3)
    assert code == _abi.TD3_ERR_CAPACITY
    assert "scratch" in _abi.last_error()
    assert bool(torch.isnan(phi).all())  # nothing was launched
