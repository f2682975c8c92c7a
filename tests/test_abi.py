"""CPU checks of the boundary: the library loads, exports every symbol that
include/isle_hfa.h declares, and the Python binding declares exactly those."""
import pathlib
import re

from isle_amd import _lib

HEADER = pathlib.Path(__file__).resolve().parent.parent / "include" / "isle_hfa.h"


def declared():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(isle_[a-zA-Z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    lib = _lib.load()
    names = declared()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name


def test_binding_matches_header():
    assert sorted(_lib.SIGNATURES) == declared()


def test_no_gpu_raises_cleanly():
    import torch
    if torch.cuda.is_available():
        return
    import pytest
    with pytest.raises(RuntimeError):
        _lib.device()
