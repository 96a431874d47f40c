"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls, CPU only)."""

from __future__ import annotations

import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "yggdrax_amd", "libyggdrax_amd.so")


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(ygg_[a-z0-9_]+)\s*\(", text))
    return names


@pytest.fixture(scope="module")
def lib_path():
    if not os.path.exists(LIB):
        from yggdrax_amd import build
        build.build()
    return LIB


def test_library_exports_every_declared_symbol(lib_path):
    declared = _declared()
    assert declared, "no declarations found"
    out = subprocess.run(["nm", "-D", "--defined-only", lib_path], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = sorted(declared - exported)
    assert not missing, f"declared but not exported: {missing}"


def test_ctypes_binding_covers_declarations(lib_path):
    from yggdrax_amd import _lib
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name)
    assert set(_lib.declared_symbols()) <= _declared()


def _declared_arity():
    """Parameter count of every ygg_* function declared in include/*.h."""
    out = {}
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for name, args in re.findall(r"\b(ygg_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S):
            args = args.strip()
            out[name] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_ctypes_argument_counts_match_header():
    from yggdrax_amd import _lib
    arity = _declared_arity()
    bad = {n: (len(a), arity[n]) for n, (_, a) in _lib._SIGS.items() if len(a) != arity.get(n, -1)}
    assert not bad, f"ctypes argtypes vs header parameter count: {bad}"


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of the header's structs have the C compiler's layout."""
    import ctypes
    from yggdrax_amd import _lib
    structs = {"YggWalkParams": _lib.WalkParams, "YggRadixTopology": _lib.RadixTopologyPtrs,
               "YggOctreeMeta": _lib.OctreeMetaPtrs, "YggFlatWalkParams": _lib.FlatWalkParams,
               "YggBatch": _lib.Batch, "YggPolicyDesc": _lib.PolicyDesc}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "yggdrax_amd.h"',
             "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = dict(ln.split() for ln in subprocess.run([str(exe)], capture_output=True, text=True,
                                                     check=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(out[cname]) == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(out[f"{cname}.{fname}"]) == getattr(py, fname).offset, (cname, fname)


def test_gpu_entry_points_fail_loudly_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("device present")
    import numpy as np
    import yggdrax_amd as ygg
    with pytest.raises(RuntimeError, match="GPU only"):
        ygg.build_tree(np.zeros((8, 3), np.float32), None, leaf_size=2)
