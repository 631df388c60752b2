"""GPU: the streaming kernels that scale the profiler's byte counters read
and write exactly the buffer they are given (ragged lengths included)."""
import numpy as np
import pytest

from hadros_amd import abi, ffi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [0, 1, 7, 4096, 1 << 20, (1 << 20) + 3])
def test_stream_probe_fill_and_read(n):
    import torch

    x = torch.zeros(n + 2, dtype=torch.float64, device="cuda")
    body = x[:n] if n else x[:0]
    ffi.stream_probe(1, body, value=0.5)
    torch.cuda.synchronize()
    h = x.cpu().numpy()
    assert np.all(h[:n] == 0.5) and np.all(h[n:] == 0.0)      # nothing past the buffer is written
    x[:n] = torch.arange(n, dtype=torch.float64, device="cuda")
    part = torch.full((abi.STREAM_GRID,), np.nan, dtype=torch.float64, device="cuda")
    ffi.stream_probe(0, body, part)
    torch.cuda.synchronize()
    assert float(part.sum()) == n * (n - 1) / 2                # every element read once


def test_stream_probe_rejects_misaligned():
    import torch

    x = torch.zeros(16, dtype=torch.float64, device="cuda")
    with pytest.raises(Exception):
        ffi.stream_probe(1, x[1:])
