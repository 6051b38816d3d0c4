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


# The kernels of euler3d.cpp's 28 op_par_loop sites (three kernels have two
# sites each).
REFERENCE_LOOPS = [
    "initialize_variables_kernel", "zero_5d_array_kernel", "zero_1d_array_kernel", "calculate_cell_volumes",
    "dampen_ewt", "copy_double_kernel", "calculate_dt_kernel", "get_min_dt_kernel",
    "compute_step_factor_kernel", "compute_flux_edge_kernel", "compute_bnd_node_flux_kernel", "time_step_kernel",
    "unstructured_stream_kernel", "residual_kernel", "calc_rms_kernel", "count_bad_vals", "up_pre_kernel",
    "up_kernel", "up_post_kernel", "down_v2_kernel_pre", "down_v2_kernel", "down_v2_kernel_post", "down_kernel",
    "identify_differences", "count_non_zeros",
]


def test_every_reference_loop_has_an_entry_point():
    names = declared_functions("mgcfd_op2.h")
    assert len(REFERENCE_LOOPS) == 25
    assert sorted(mgcfd.LOOPS) == sorted(REFERENCE_LOOPS)
    for k in REFERENCE_LOOPS:
        assert "op_par_loop_" + k in names
        assert hasattr(mgcfd.lib(), "op_par_loop_" + k)


def test_loop_without_init_fails_cleanly():
    """An entry point called before op_init returns an error code, no crash."""
    lib = mgcfd.lib()
    rc = lib.op_par_loop_dampen_ewt(b"dampen_ewt", None, mgcfd.OpArg())
    assert rc != 0
    assert "op_init" in mgcfd.last_error()
