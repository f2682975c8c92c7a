"""GPU: the allocation-free log det entry point (isle_logdet_noalloc) against isle_logdet.

Both run the same LU kernels on the same input, so their results are equal bit for bit; two
calls in a row, the second with a smaller matrix (n = 300, then 37), check that the static
scratch carries nothing from one call to the next.  The log dets also equal numpy's
slogdet to 1e-12 (a textbook routine), and n = 513 is rejected with ValueError."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

isle = pytest.importorskip("isle_amd")
from isle_amd import _lib  # noqa: E402


def _logdet(fn, a):
    out = np.zeros(2)
    _lib.check(fn(_lib.ptr(a), a.shape[0], _lib.dptr(out), _lib.stream_ptr()))
    return out


def test_noalloc_logdet_matches_allocating_entry_point():
    lib = _lib.load()
    rng = np.random.RandomState(5)
    for n in (300, 37):
        print(f"noalloc logdet, n={n}", flush=True)
        host = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        a = torch.as_tensor(host, device="cuda")
        want = _logdet(lib.isle_logdet, a)
        got = _logdet(lib.isle_logdet_noalloc, a)
        assert np.array_equal(got, want), (n, got, want)
        sign, mag = np.linalg.slogdet(host)
        assert abs(got[0] - mag) < 1e-12 * abs(mag)
        assert abs(np.exp(1j * got[1]) - sign) < 1e-12
    with pytest.raises(ValueError, match="exceeds 512"):
        _logdet(lib.isle_logdet_noalloc, torch.zeros((513, 513), dtype=torch.complex128, device="cuda"))
