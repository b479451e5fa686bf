"""The C-ABI library builds for gfx950, loads and exports every symbol that
include/creep_hip.h declares; the ctypes structure layouts match the header.
CPU only (no compute calls)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "creep_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from creep_modeling_amd import lib

    L = lib.load()
    names = declared_functions()
    assert names, "no declarations parsed"
    for n in names:
        assert hasattr(L, n), n
    assert sorted(lib.EXPORTS) == names
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib.LIBPATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (cg_[a-z0-9_]+)", out))
    assert set(names) <= exported


def test_library_targets_gfx950():
    from creep_modeling_amd import lib

    out = subprocess.run(["/opt/rocm/bin/roc-obj-ls", str(lib.LIBPATH)], capture_output=True, text=True)
    if out.returncode != 0:
        pytest.skip("roc-obj-ls unavailable")
    assert "gfx950" in out.stdout


def _sizeof_from_hipcc(struct):
    src = '#include "%s"\n#include <stdio.h>\nint main(){printf("%%zu\\n", sizeof(%s));return 0;}\n' % (HEADER, struct)
    import tempfile

    with tempfile.TemporaryDirectory() as d:
        p = Path(d) / "t.c"
        p.write_text(src)
        exe = Path(d) / "t"
        subprocess.run(["gcc", "-o", str(exe), str(p)], check=True)
        return int(subprocess.run([str(exe)], capture_output=True, text=True).stdout)


@pytest.mark.parametrize("name,cls", [("cg_process", "CgProcess"), ("cg_iter_state", "CgIterState"),
                                      ("cg_iter_out", "CgIterOut"), ("cg_incr_out", "CgIncrOut"),
                                      ("cg_stats", "CgStats")])
def test_struct_layouts_match_header(name, cls):
    from creep_modeling_amd import lib

    assert C.sizeof(getattr(lib, cls)) == _sizeof_from_hipcc(name)


def test_cg_create_without_gpu_reports_error():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from creep_modeling_amd import lib

    L = lib.load()
    h = C.c_void_p()
    rc = L.cg_create(8, 8, 8, 1, 0.01, 0, 1, 0, C.byref(h))
    assert rc != 0 and not h.value
