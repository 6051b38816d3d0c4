"""The C-ABI library loads without a GPU and exports every declared symbol."""
import os
import re

import pytest

import mgcfd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = ["mgcfd_op2.h", "mgcfd_mesh.h"]


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"#[^\n]*", "", text)
    names = []
    for stmt in text.split(";"):
        stmt = " ".join(stmt.split())
        if stmt.startswith("typedef") or "{" in stmt or "}" in stmt:
            stmt = stmt.split("}")[-1].strip()
            if not stmt or stmt.startswith("typedef"):
                continue
        m = re.match(r"^(?:extern \"C\" )?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", stmt)
        if m:
            names.append(m.group(1))
    return names


@pytest.mark.parametrize("header", HEADERS)
def test_exports(header):
    lib = mgcfd.lib()
    names = declared_functions(header)
    assert len(names) >= 4
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{header}: not exported: {missing}"


def test_every_reference_loop_has_an_entry_point():
    # the 22 op_par_loop kernels of euler3d.cpp (SURVEY inventory of loop sites)
    names = declared_functions("mgcfd_op2.h")
    for k in mgcfd.LOOPS:
        assert "op_par_loop_" + k in names


def test_loop_without_init_fails_cleanly():
    """An entry point called before op_init returns an error code, no crash."""
    lib = mgcfd.lib()
    rc = lib.op_par_loop_dampen_ewt(b"dampen_ewt", None, mgcfd.OpArg())
    assert rc != 0
    assert "op_init" in mgcfd.last_error()
