"""C-ABI boundary checks that need no GPU.

The library must load, export every entry point the headers declare, and
fail loudly (TNR_ERR_NODEVICE) instead of computing anywhere but on a
gfx950 device.  Host-only entry points (free energy, error strings) are
checked against the oracle.
"""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from oracle import schemes as osch


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names.update(re.findall(r"\b(tnr_\w+)\s*\(", txt))
    return sorted(names)


def test_headers_declare_the_boundary():
    names = declared_symbols()
    for n in ("tnr_create", "tnr_step", "tnr_finalize", "tnr_run", "tnr_free_energy",
              "tnr_destroy", "tnr_run_batch", "tnr_server_submit", "tnr_kernel_gemm"):
        assert n in names


def test_library_exports_every_declared_symbol(tk):
    lib = tk.lib()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_signatures_cover_tnr_h(tk):
    from importlib import import_module
    sig = import_module("tnrkit_jl_amd._lib").SIGNATURES
    txt = open(os.path.join(ROOT, "include", "tnr.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    for n in re.findall(r"\b(tnr_\w+)\s*\(", txt):
        assert n in sig, n


def test_free_energy_entry_point_matches_oracle(tk):
    rng = np.random.default_rng(0)
    data = list(rng.uniform(0.5, 3.0, 31))
    for sf, init in ((2.0, 1.0), (4.0, 1.0), (2.0, 2.0)):
        f = tk.free_energy(data, 0.44, scalefactor=sf, initial_size=init)
        g = osch.free_energy(data, 0.44, scalefactor=sf, initial_size=init)
        assert abs(f / g - 1) < 1e-15


def test_invalid_arguments_are_rejected(tk):
    lib = tk.lib()
    f = C.c_double()
    d = (C.c_double * 3)(1.0, 2.0, 3.0)
    assert lib.tnr_free_energy(d, 3, 0.5, 1.0, 1.0, C.byref(f)) == -1      # scalefactor 1
    assert lib.tnr_free_energy(None, 3, 0.5, 2.0, 1.0, C.byref(f)) == -1
    assert lib.tnr_strerror(-7).decode() == "no usable gfx950 device"


def test_no_cpu_fallback(tk):
    """Without a GPU every compute entry point must fail with NODEVICE."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    T = tk.classical_ising(symmetry="Z2")
    with pytest.raises(tk.TNRError) as e:
        tk.BTRG(T)
    assert e.value.status == -7
