"""The C-ABI library loads on a host without a GPU and exports every symbol the
public header declares; the ctypes signatures agree with the header's
parameter counts."""
import os
import re

from sod2d_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _prototypes():
    txt = open(os.path.join(ROOT, "include", "sod2d_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    protos = {}
    for m in re.finditer(r"\b(sodhip_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", txt):
        args = m.group(2).strip()
        n = 0 if args in ("", "void") else args.count(",") + 1
        protos[m.group(1)] = n
    return protos


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    protos = _prototypes()
    assert protos, "no prototypes parsed"
    missing = [s for s in protos if not hasattr(lib, s)]
    assert not missing, missing


def test_ctypes_signatures_match_header():
    protos = _prototypes()
    for name, (res, args) in _lib._SIGS.items():
        assert name in protos, name
        assert len(args) == protos[name], (name, len(args), protos[name])
