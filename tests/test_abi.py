"""The C-ABI library loads, exports every entry point include/huntian_abi.h declares, its
struct layouts match the ctypes mirror, and the host-only entry points agree with the
oracle.  No GPU needed."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from huntian_hip import abi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "huntian_abi.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hnt_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    names = declared_functions()
    assert len(names) >= 14
    assert set(names) == set(abi.EXPORTED)
    for n in names:
        assert hasattr(lib, n), n
    assert lib.hnt_abi_version() == 3


def test_struct_layouts_match_the_header(tmp_path):
    structs = {"hnt_shell_tables": abi.ShellTables, "hnt_ctab": abi.CTab,
               "hnt_peaks": abi.Peaks, "hnt_fields": abi.Fields, "hnt_batch": abi.Batch,
               "hnt_params": abi.Params, "hnt_results": abi.Results, "hnt_mlp": abi.Mlp,
               "hnt_filter_bank": abi.FilterBank, "hnt_cosmo_tables": abi.CosmoTables,
               "hnt_pack_params": abi.PackParams, "hnt_pk_table": abi.PkTable,
               "hnt_transfer": abi.Transfer}
    prog = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"',
            'int main(void) {']
    for cname, py in structs.items():
        prog.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            prog.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    prog.append("return 0; }")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(prog))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(c)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = dict(line.split() for line in out.strip().splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)


@pytest.mark.parametrize("rmax", [2, 8, 15, 30])
def test_shell_tables_match_oracle(oracle, rmax):
    from huntian_hip.api import build_shell_tables

    t = build_shell_tables(rmax)
    o = oracle.shell_tables(rmax)
    for k in ("di", "dj", "dk", "shell_start", "shell_count", "shell_r2"):
        assert np.array_equal(getattr(t, k), o[k]), k


def test_fsc_of_z_matches_oracle(oracle, websky, fortran_table):
    from huntian_hip.api import fsc_of_z

    for ct in (websky, fortran_table):
        for z in (0.0, 0.5, 2.0):
            assert fsc_of_z(ct, z) == oracle.fsc_of_z(ct, z)  # same FP64 steps, bit for bit


def test_errors_are_reported_not_raised_across_the_abi():
    lib = abi.load()
    nc, ns = C.c_int64(), C.c_int32()
    assert lib.hnt_shell_table_size(-3, C.byref(nc), C.byref(ns)) == abi.HNT_ERR_INVALID
    assert "rmax" in abi.last_error(lib)
    # NULL context
    assert lib.hnt_eig3_symmetric_batch(None, None, 1, None) == abi.HNT_ERR_INVALID
    assert "context" in abi.last_error(lib)
    assert lib.hnt_ctx_destroy(None) == abi.HNT_OK
