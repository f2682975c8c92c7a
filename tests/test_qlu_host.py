"""Host logic of the QLU interface (no GPU): block-count consistency (the reference's
QLU::isConsistent, hubbardFermiMatrixExp.cpp:290-303), the nt >= 3 requirement of
reconstruct, and the argument checks of logdetQ / solveQ, all before any device work."""
import pytest
import torch

import isle_amd as isle


def _lu(nt, nx=2, du=None, dv=None):
    z = lambda k: torch.zeros((max(k, 0), nx, nx), dtype=torch.complex128)  # noqa: E731
    return isle.QLU(z(nt), z(nt - 1 if du is None else du), z(nt - 1 if du is None else du),
                    z(nt - 2 if dv is None else dv), z(nt - 2 if dv is None else dv), [0j] * nt)


def test_consistency():
    for nt in (1, 2, 3, 8):
        assert _lu(nt).isConsistent()
    assert not _lu(0).isConsistent()
    assert not _lu(4, du=2).isConsistent()
    assert not _lu(4, dv=1).isConsistent()


def test_reconstruct_needs_three_slices():
    for nt in (1, 2):
        with pytest.raises(ValueError):
            _lu(nt).reconstruct()
    with pytest.raises(RuntimeError):
        _lu(4, dv=1).reconstruct()


def test_argument_checks():
    with pytest.raises(TypeError):
        isle.logdetQ(object())
    with pytest.raises(TypeError):
        isle.solveQ(object(), None)
    with pytest.raises(TypeError):
        isle.solveQ(_lu(3))
    with pytest.raises(RuntimeError):
        isle.logdetQ(_lu(4, du=2))
