"""The octree and tie-break checks are sensitive: each one-character
mutation of the oracle's source below, built into its own library on the
CPU, fails at least one of them (no GPU)."""

from __future__ import annotations

import concurrent.futures
import ctypes
import os
import shutil
import subprocess

import pytest

from oracle import oracle as O
from tests import oracle_checks as C
from tests.test_oracle_pins_octree import CASES

SRC = os.path.join(os.path.dirname(O.__file__), "ygg_oracle.cpp")

# name -> (text in the oracle source, replacement, checks expected to catch it)
MUTANTS = {
    "range_first_side": ("int64_t f = jax_searchsorted(lc, L, oct_node_codes[k], true);",
                         "int64_t f = jax_searchsorted(lc, L, oct_node_codes[k], false);",
                         "ranges"),
    "range_last_side": ("int64_t l = jax_searchsorted(lc, L, endc, false) - 1;",
                        "int64_t l = jax_searchsorted(lc, L, endc, true) - 1;", "ranges"),
    "root_shift_clamp": ("std::min<uint64_t>(bs, 62ull)", "std::min<uint64_t>(bs, 63ull)",
                         "ranges"),
    "radix_map_side": ("oct_address(code, dep), true);", "oct_address(code, dep), false);",
                       "radix_map"),
    "level_bins": ("for (int k = 0; k < 22; ++k) oct_level_offsets",
                   "for (int k = 0; k < 21; ++k) oct_level_offsets", "levels"),
    "num_levels": ("scalars[0] = maxd + 1;", "scalars[0] = maxd + 2;", "levels"),
    "tie_target_strict": ("et >= es", "et > es", "tie"),
    "tie_source_nonstrict": ("es > et", "es >= et", "tie"),
}


def _build(name, tmpdir):
    old, new, _ = MUTANTS[name]
    src = open(SRC).read()
    assert src.count(old) == 1, name
    path = os.path.join(tmpdir, f"ygg_oracle_{name}.cpp")
    with open(path, "w") as f:
        f.write(src.replace(old, new))
    so = os.path.join(tmpdir, f"libygg_oracle_{name}.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off",
                    "-shared", "-o", so, path], check=True)
    return name, so


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = str(tmp_path_factory.mktemp("oracle_mutants"))
    with concurrent.futures.ThreadPoolExecutor(max_workers=4) as ex:
        built = dict(ex.map(lambda n: _build(n, d), MUTANTS))
    return {n: O.configure(ctypes.CDLL(p)) for n, p in built.items()}


def _run_checks(kind):
    if kind == "tie":
        C.run_tie_break_on_oracle()
        return
    check = {"ranges": C.check_oct_node_ranges, "radix_map": C.check_radix_node_to_oct,
             "levels": C.check_oct_levels}[kind]
    for pos, leaf in CASES.values():
        check(O.build_octree(pos, None, leaf_size=leaf))


def test_checks_pass_on_the_oracle():
    for kind in ("ranges", "radix_map", "levels", "tie"):
        _run_checks(kind)


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_mutant_fails_its_check(name, mutant_libs, monkeypatch):
    monkeypatch.setattr(O, "lib", mutant_libs[name])
    with pytest.raises(AssertionError):
        _run_checks(MUTANTS[name][2])
