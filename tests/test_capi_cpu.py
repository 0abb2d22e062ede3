"""CPU-side checks of the HIP library: it loads, exports every entry point that
include/mpres_hip.h declares, and its host-computed constant tables equal the
oracle's (which are pinned by test_oracle_consts.py). No GPU is touched."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def mp():
    import importlib.util

    spec = importlib.util.spec_from_file_location("mpres", os.path.join(ROOT, "mpres-blas_amd", "mpres.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def declared_symbols():
    names = set()
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if not hdr.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", hdr)).read()
        names |= set(re.findall(r"\b(mpres_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_all_declared_symbols(mp):
    L = mp.lib()
    syms = declared_symbols()
    assert len(syms) > 25
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_moduli_sets_listed(mp):
    names = mp.set_names()
    assert names == po.set_names()
    assert mp.lib().mpres_moduli_set_id(b"32_8double") == names.index("32_8double")
    assert mp.lib().mpres_moduli_set_id(b"nope") == -1


@pytest.mark.parametrize("name", ["8_2double", "16_4double", "32_8double", "8_32bit", "16_32bit", "32_32bit",
                                  "8_16bit", "16_16bit", "32_16bit"])
def test_host_constants_match_oracle(mp, oracle_lib, name):
    L = mp.lib()
    sid = L.mpres_moduli_set_id(name.encode())
    o = po.Oracle(name)
    n = o.n

    def table(what, count, dtype):
        out = np.zeros(count, dtype)
        mp.check(L.mpres_moduli_set_table(sid, what, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    assert (table(0, n, np.int32) == o.moduli()).all()
    assert (table(1, n, np.int32) == o.part_inverse()).all()
    assert (table(2, (o.log2m + 1) * n, np.int32).reshape(-1, n) == o.pow2_table()).all()
    assert (table(3, 30, np.int32) == o.pow2_residues()).all()
    assert (table(4, 30 * n, np.int32).reshape(30, n) == o.part_pow2_residues()).all()
    assert (table(5, 30 * n, np.int32).reshape(30, n) == o.pow2_inverse()).all()
    assert (table(6, n * n, np.int32).reshape(n, n) == o.mrc_inverse()).all()
    d = table(7, 2 * n + 1, np.float64)
    rd, ru = o.recip()
    ec = o.eval_consts()
    assert (d[:n].view(np.uint64) == rd.view(np.uint64)).all()
    assert (d[n:2 * n].view(np.uint64) == ru.view(np.uint64)).all()
    assert d[2 * n] == ec["acc"]
    u = table(8, 4, mp.ER_DTYPE)
    for rec, key in zip(u, ("unit_low", "unit_upp", "inv_unit_low", "inv_unit_upp")):
        assert rec["frac"] == ec[key]["frac"] and rec["exp"] == ec[key]["exp"], key
    iv = table(9, 5, np.int32)
    assert tuple(iv[:3]) == o.mp_consts()
    assert iv[3] == ec["ref_factor"] and iv[4] == o.log2m


def test_errors_without_device(mp):
    L = mp.lib()
    h = C.c_void_p()
    rc = L.mpres_ctx_create(C.byref(h), 999, 0)
    assert rc == -1  # unknown moduli set -> MPRES_ERR_ARG
    assert b"unknown moduli set" in L.mpres_last_error()


def test_compat_header_compiles_and_links(tmp_path):
    """The header-only layer with the reference's cuda::mp_* names forwards to
    the C ABI: a program using each name compiles and links."""
    import subprocess

    out = tmp_path / "compat_calls"
    cmd = ["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "host", "compat_calls.cpp"), "-o", str(out),
           "-L", os.path.join(ROOT, "mpres-blas_amd"), "-lmpres_hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
