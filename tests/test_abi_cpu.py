"""CPU-side checks of the C-ABI library: it loads, exports every function include/cqbem.h declares,
computes the same CQ frequencies as the oracle, validates arguments, and reports a missing GPU
as an error instead of falling back."""
import ctypes
import os
import re

import numpy as np
import pytest

from surfaceonlydeformables_amd import abi
from surfaceonlydeformables_amd import mesh as M

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "cqbem.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*\**(cqbem_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = abi.lib()
    names = declared_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(abi.EXPORTED) == names


@pytest.mark.parametrize("L,bdf", [(32, 2), (31, 1), (16, 2)])
def test_frequencies_match_oracle(oracle, L, bdf):
    R = 1e-10 ** (1 / (2 * L))
    a = abi.frequencies(L, R, 1e-3, bdf)
    b = oracle.frequencies(L, R, 1e-3, bdf)
    assert np.array_equal(a, b)


def test_argument_validation():
    V, F = M.icosphere(0, 1.0)
    mesh = abi.make_mesh(V, F)
    good = abi.make_params(2.0, 1.0, 1000.0, 0.3, 1e-3, 0.7, 8)
    assert abi.job_footprint(mesh, good) > 0
    for field, val in [("L", 1), ("bdf", 3), ("order", 6), ("subdivision", 0), ("c2", -1.0), ("precision", 7),
                       ("which", 64)]:
        bad = abi.make_params(2.0, 1.0, 1000.0, 0.3, 1e-3, 0.7, 8)
        setattr(bad, field, val)
        with pytest.raises(abi.CQBEMError) as e:
            abi.job_footprint(mesh, bad)
        assert e.value.code == abi.ERR_ARG
    Fbad = F.copy(); Fbad[3, 1] = len(V)
    with pytest.raises(abi.CQBEMError):
        abi.job_footprint(abi.make_mesh(V, Fbad), good)
    assert abi.lib().cqbem_frequencies(1, 0.7, 1e-3, 2, None) == abi.ERR_ARG


def test_footprint_scales_with_outputs():
    V, F = M.icosphere(2, 1.0)
    mesh = abi.make_mesh(V, F)
    p = abi.make_params(2.0, 1.0, 1000.0, 0.3, 1e-3, 0.7, 32)
    dev = abi.job_footprint(mesh, p, outputs_on_device=True)
    host = abi.job_footprint(mesh, p, outputs_on_device=False)
    n3 = 3 * len(V)
    assert host - dev >= 2 * 32 * n3 * n3 * 8


def test_context_without_gpu_is_an_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(abi.CQBEMError):
        abi.Context(0)
