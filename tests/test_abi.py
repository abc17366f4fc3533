"""The C-ABI library loads and exports every symbol include/*.h declares (no GPU calls)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for hdr in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(hdr).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src)
        names.update(re.findall(r"\b(xc_[a-z0-9_]+)\s*\(", src))
    return names


def test_library_exports_all_declared_symbols():
    import xcore_amd
    lib = ctypes.CDLL(xcore_amd.LIB_PATH)
    names = _declared()
    assert len(names) > 10
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_declares_all_symbols():
    import xcore_amd
    xcore_amd.load()
    assert _declared() <= set(xcore_amd.declared_symbols()) | {"xc_scheme_halo"}


def test_no_cpu_fallback(tmp_path):
    # the product path fails loudly when the HIP library is missing
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r); import xcore_amd\n"
            "try:\n    xcore_amd.load()\nexcept RuntimeError as e:\n"
            "    print('RAISED', 'no CPU fallback' in str(e))\n") % os.path.join(ROOT, "xcore.jl_amd")
    env = dict(os.environ, XCORE_HIP_LIB=str(tmp_path / "missing.so"))
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert "RAISED True" in r.stdout, r.stdout + r.stderr
