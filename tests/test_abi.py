"""The C-ABI library loads and exports every entry point include/bz.h declares
(no device calls: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "bz.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bz_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("bz_ctx_create", "bz_push_mcc", "bz_deposit", "bz_field_step", "bz_bin_sort",
                 "bz_resample", "bz_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from boltzmann_amd import _lib
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_host_philox_matches_oracle():
    from boltzmann_amd import _lib
    from oracle import oracle as O
    rng = np.random.default_rng(0)
    for _ in range(50):
        c = tuple(int(v) for v in rng.integers(0, 2 ** 32, 4))
        k = tuple(int(v) for v in rng.integers(0, 2 ** 32, 2))
        assert _lib.philox(c, k) == O.philox(c, k)


def test_error_path_without_device():
    """A null context is rejected with a message, not a crash."""
    from boltzmann_amd import _lib
    lib = _lib.load()
    rc = lib.bz_push_mcc(None, None, None, None, None, None, None, None, None)
    assert rc != 0
    assert b"null context" in lib.bz_last_error()


def test_struct_layouts_match_the_header(tmp_path):
    """Every ctypes mirror has the size and field offsets the C compiler gives
    the header's struct (a host program built from include/bz.h with gcc)."""
    import subprocess

    from boltzmann_amd import _lib
    pairs = {"bz_params_t": _lib.Params, "bz_xsec_t": _lib.XSec,
             "bz_field_params_t": _lib.FieldParams, "bz_mlp_t": _lib.Mlp,
             "bz_adam_t": _lib.Adam, "bz_resample_policy_t": _lib.ResamplePolicy,
             "bz_state_t": _lib.State}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "bz.h"', "int main(void) {"]
    for cname, cls in pairs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True,
                               text=True).stdout.splitlines():
        cname, field, val = line.split()
        got[(cname, field)] = int(val)
    for cname, cls in pairs.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
