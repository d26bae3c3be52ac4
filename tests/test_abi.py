"""The C ABI: libkma.so loads without a GPU and exports every function that
the host headers (include/*.h with an extern "C" block) declare; the Python
binding lists exactly those.  No device call is made."""
import ctypes
import glob
import os
import re

import kma_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^[A-Za-z_][\w \t\*]*?\b(kma_\w+)\s*\(", re.M)


def declared():
    names = {}
    for path in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))):
        text = open(path).read()
        if 'extern "C"' not in text:
            continue  # device-only header
        body = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        body = "\n".join(l for l in body.splitlines() if not l.lstrip().startswith("#"))
        for m in DECL.finditer(body):
            line = body[m.start():body.find("\n", m.start())]
            if "typedef" in line or "(*" in line:
                continue
            names[m.group(1)] = os.path.basename(path)
    return names


def test_headers_declare_what_binding_lists():
    d = declared()
    assert len(d) >= 40
    assert set(d) == set(kma_amd.EXPORTS), (set(d) ^ set(kma_amd.EXPORTS))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(kma_amd.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert missing == []


def test_no_undeclared_kma_exports():
    """Every exported kma_* function is declared in a public header (the
    host-side helpers and layout functions are hidden)."""
    import shutil
    import subprocess

    if not shutil.which("nm"):
        import pytest
        pytest.skip("nm not available")
    out = subprocess.run(["nm", "-D", "--defined-only", kma_amd.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.split()[-2:-1] == ["T"]}
    kma = {n for n in exported if n.startswith("kma_")}
    assert kma == set(declared()), kma ^ set(declared())
