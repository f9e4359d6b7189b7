"""The C-ABI library loads and exports every function include/nuosc_hip.h
declares; host-only entry points work without a GPU (no compute calls)."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            txt = open(os.path.join(inc, fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(nuosc_\w+)\s*\(", txt))
    return names


def test_every_declared_symbol_is_exported(nuosc_mod):
    L = nuosc_mod.lib()
    names = declared_functions()
    assert len(names) >= 25
    for n in sorted(names):
        assert hasattr(L, n), n
    # the ctypes binding declares a signature for each of them
    assert names == set(nuosc_mod.SIGNATURES), names ^ set(nuosc_mod.SIGNATURES)


def test_identity(nuosc_mod):
    L = nuosc_mod.lib()
    assert L.nuosc_abi_version() == 5
    assert b"gfx950" in L.nuosc_build_info()


def test_host_vgrid_matches_oracle(nuosc_mod, oracle_mod):
    for nn in (1, 3, 6):
        for a, b in zip(nuosc_mod.vgrid(nn), oracle_mod.vgrid(nn)):
            assert np.array_equal(a, b)


def test_host_vgrid_rejects_bad_input(nuosc_mod):
    import ctypes as C
    L = nuosc_mod.lib()
    buf = (C.c_double * 16)()
    p = C.cast(buf, C.POINTER(C.c_double))
    assert L.nuosc_vgrid(0, 0, p, p, p, p) == -1
    assert L.nuosc_vgrid(1, 7, p, p, p, p) == -1
    assert L.nuosc_vgrid(1, 0, None, p, p, p) == -1


def test_gpu_kernels_are_gfx950_code_objects():
    """The shared library embeds gfx950 device code (offload bundle)."""
    data = open(os.path.join(ROOT, "nuosc_amd", "libnuosc_hip.so"), "rb").read()
    assert b"gfx950" in data
    assert b"k_stage" in data


def test_estimate_refuses_unaddressable_boxes(nuosc_mod):
    """Host-only sizing (no GPU): the stage kernels address one block's x rows
    through 32-bit buffer offsets, so a box whose stored x row is too long is
    refused up front, for both flavour counts; one x row of the same length is
    accepted with 3 flavours (that kernel addresses with 64-bit pointers)."""
    N = nuosc_mod
    # 642 directions, FP64: a stored x row is nz * nf * 648 * 8 bytes
    ok2 = N.estimate_device_bytes(nv=8, x0=0.0, x1=0.8, z0=0.0, z1=409.6, dx=0.1, dz=0.1, nflavor=2)   # nz 4096
    assert ok2 > 0
    with pytest.raises(N.NuOscError):   # 12 rows x 16384 x 8 x 648 x 8 B > 4 GiB
        N.estimate_device_bytes(nv=8, x0=0.0, x1=0.8, z0=0.0, z1=1638.4, dx=0.1, dz=0.1, nflavor=2)
    ok3 = N.estimate_device_bytes(nv=8, x0=0.0, x1=0.8, z0=0.0, z1=204.8, dx=0.1, dz=0.1, nflavor=3)   # nz 2048
    assert ok3 > 0
    with pytest.raises(N.NuOscError):   # 12 rows x 4096 x 18 x 648 x 8 B > 4 GiB
        N.estimate_device_bytes(nv=8, x0=0.0, x1=0.8, z0=0.0, z1=409.6, dx=0.1, dz=0.1, nflavor=3)
    # the same z length in one x row: no block span limit for 3 flavours
    assert N.estimate_device_bytes(nv=8, x0=0.0, x1=0.1, z0=0.0, z1=409.6, dx=0.1, dz=0.1, nflavor=3) > 0
