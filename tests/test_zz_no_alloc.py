"""hipns_evidence_scan_noalloc: the evidence scan with the library's fixed device scratch instead of a
caller workspace, against hipns_evidence_scan with a workspace (the same kernels: bit-identical), on
two calls in a row, the second smaller, a dynamic stop, and the size limit."""
import numpy as np
import pytest

from jaxns_amd import _lib
from jaxns_amd._lib import check, lib, ptr

pytestmark = pytest.mark.gpu


def _inputs(M, seed):
    rng = np.random.default_rng(seed)
    logL = np.sort(rng.normal(size=M).cumsum() * 0.01 + rng.normal(size=M) * 0.1)
    logL[: min(3, M)] = -np.inf
    n = rng.integers(1, 300, size=M).astype(np.float64)
    return logL, n


def _scan(logL, n, stop, noalloc):
    import torch
    M = logL.size
    L = torch.from_numpy(logL).cuda()
    nn = torch.from_numpy(n).cuda()
    per = torch.full((8, M), 7.0, dtype=torch.float64, device="cuda")
    fin = torch.full((8,), 7.0, dtype=torch.float64, device="cuda")
    if noalloc:
        check(lib().hipns_evidence_scan_noalloc(ptr(L), ptr(nn), M, stop, ptr(per), ptr(fin), _lib.stream_ptr()),
              "hipns_evidence_scan_noalloc")
    else:
        ws = torch.empty(int(lib().hipns_evidence_scan_workspace_size(M)), dtype=torch.uint8, device="cuda")
        check(lib().hipns_evidence_scan(ptr(L), ptr(nn), M, stop, ptr(per), ptr(fin), ptr(ws), ws.numel(),
                                        _lib.stream_ptr()), "hipns_evidence_scan")
    torch.cuda.synchronize()
    return per.cpu().numpy(), fin.cpu().numpy()


def test_noalloc_scan_matches_workspace_scan():
    mx = int(lib().hipns_evidence_scan_noalloc_max())
    print("noalloc max samples", mx, flush=True)
    assert mx >= 4_000_000
    for M, stop, seed in ((1_500_001, -1, 1), (70_001, 40_000, 2), (5, -1, 3)):
        print(f"noalloc scan M={M} stop={stop}", flush=True)
        logL, n = _inputs(M, seed)
        a = _scan(logL, n, stop, noalloc=True)
        b = _scan(logL, n, stop, noalloc=False)
        assert np.array_equal(a[0].view(np.int64), b[0].view(np.int64))
        assert np.array_equal(a[1].view(np.int64), b[1].view(np.int64))


def test_noalloc_scan_rejects_sizes_beyond_its_scratch():
    import torch
    mx = int(lib().hipns_evidence_scan_noalloc_max())
    d = torch.zeros(8, dtype=torch.float64, device="cuda")
    rc = lib().hipns_evidence_scan_noalloc(ptr(d), ptr(d), mx + 1, -1, None, ptr(d), _lib.stream_ptr())
    assert rc == -4 and "fixed scan workspace" in _lib.last_error()
