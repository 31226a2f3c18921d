"""C-ABI checks that need no GPU: libmagritte_hip.so loads and exports every
entry point include/magritte_hip.h declares; the Python view struct matches
the C layout of mgr_model_view; the binding table covers the header."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from magritte_amd import hip
from magritte_amd.model import ConvStats, ModelView

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "magritte_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mgr_[A-Za-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = hip.load()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name


def test_binding_table_matches_header():
    assert set(hip.SIGNATURES) == set(declared_functions())


def test_no_cpu_fallback_when_library_missing(tmp_path, monkeypatch):
    monkeypatch.setattr(hip, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(hip, "_LIB", None)
    with pytest.raises(hip.HipError):
        hip.load()


def _c_layout(struct, fields):
    src = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void){",
           f'printf("%zu\\n", sizeof({struct}));']
    for f in fields:
        src.append(f'printf("%zu\\n", offsetof({struct}, {f}));')
    src.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write("\n".join(src))
        subprocess.run(["gcc", "-std=c99", "-o", exe, c], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    return [int(x) for x in out]


@pytest.mark.parametrize("pystruct,cname", [(ModelView, "mgr_model_view"), (ConvStats, "mgr_conv")])
def test_struct_layout_matches_header(pystruct, cname):
    fields = [f[0] for f in pystruct._fields_]
    layout = _c_layout(cname, fields)
    assert layout[0] == C.sizeof(pystruct)
    for f, off in zip(fields, layout[1:]):
        assert getattr(pystruct, f).offset == off, f


def test_create_without_gpu_fails_cleanly():
    """Here there is no GPU: mgr_create must return an error code, not crash."""
    if hip.device_count() > 0:
        pytest.skip("a GPU is visible")
    from magritte_amd.setup import synthetic
    with pytest.raises(hip.HipError):
        hip.Context(synthetic.regular_grid_model(n=4, nquads=3))
