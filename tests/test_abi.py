"""The C-ABI library loads on a host without a GPU, exports every symbol the header
declares, and maps argument errors to the reference's exception types."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols(name="cipc_gpu.h", prefix="cipc_gpu_"):
    text = open(os.path.join(ROOT, "include", name)).read()
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from cipc import gpu
    L = gpu.load()
    declared = header_symbols()
    assert declared, "no declarations found"
    for name in declared:
        assert hasattr(L, name), name
    assert set(declared) == set(gpu.EXPORTS)


def test_invalid_arguments_raise_value_error():
    from cipc import gpu
    import numpy as np
    with pytest.raises(ValueError):
        gpu.eval_make_pd(np.zeros((1, 5, 5)))
    with pytest.raises(ValueError):
        gpu._check(gpu.load().cipc_gpu_eval_accd(1, 7, None, None, 0.2, 0.0, None))


def test_library_exports_every_host_symbol():
    from cipc import gpu
    L = gpu.load()
    declared = header_symbols("cipc_host.h", "cipc_host_")
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(L, name), name
