"""CPU checks of the C ABI: every symbol include/geon.h declares is exported by
libgeon_hip.so, and the ctypes mirror has the C struct sizes."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "geon.h")
LIB = os.path.join(ROOT, "physsim_amd", "libgeon_hip.so")


def declared_functions():
    src = open(HDR).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(geon_[a-z0-9_]+)\s*\(", src, re.M)))


@pytest.fixture(scope="module")
def lib_built():
    if not os.path.exists(LIB):
        from physsim_amd import build
        build.build()
    return LIB


def test_exports_every_declared_symbol(lib_built):
    out = subprocess.run(["nm", "-D", "--defined-only", lib_built], capture_output=True, text=True, check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    decl = declared_functions()
    assert len(decl) >= 20
    missing = [d for d in decl if d not in exported]
    assert not missing, missing


def test_library_loads_and_binds(lib_built):
    from physsim_amd import load_library, _abi as A
    lib = load_library()
    assert lib.geon_abi_version() == A.ABI_VERSION
    assert set(A.PROTOTYPES) == set(declared_functions())
    p = A.GeonParams()
    lib.geon_default_params(p)
    assert p.theta == 0.5 and p.softening == 8.0 and p.max_substeps == 32
    assert p.toggles0 & A.BARNES_HUT


def test_struct_sizes_match_c(tmp_path):
    from physsim_amd import _abi as A
    import ctypes as C
    prog = tmp_path / "sz.c"
    prog.write_text('#include "%s"\n#include <stdio.h>\nint main(){printf("%%zu %%zu %%zu %%zu %%zu %%zu %%zu %%zu\\n",'
                    'sizeof(GeonParams),sizeof(GeonParticleInit),sizeof(GeonSoAView),sizeof(GeonForceView),'
                    'sizeof(GeonStats),sizeof(GeonCounters),sizeof(GeonTreeNode),sizeof(GeonHandle));}' % HDR)
    exe = tmp_path / "sz"
    subprocess.run(["gcc", str(prog), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    want = [C.sizeof(t) for t in (A.GeonParams, A.GeonParticleInit, A.GeonSoAView, A.GeonForceView,
                                  A.GeonStats, A.GeonCounters, A.GeonTreeNode, C.c_void_p)]
    assert got == want


def test_create_without_gpu_fails_loudly(lib_built):
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from physsim_amd import HipPhysics, GeonError
    with pytest.raises(GeonError):
        HipPhysics(max_particles=16)


def test_upload_rejects_ragged_columns():
    """Every upload column must hold as many entries as x: the C ABI reads n
    entries from each, so a shorter column would be read past its end."""
    from physsim_amd.geon import _check_lengths
    import numpy as np
    names = ("x", "y", "mass")
    _check_lengths(4, [np.zeros(4), np.zeros(4), None], names)
    with pytest.raises(ValueError, match="column y has 3 entries"):
        _check_lengths(4, [np.zeros(4), np.zeros(3), np.zeros(4)], names)
