"""The C-ABI library loads and exports every entry point include/frpa.h declares (CPU; no GPU calls)."""
import ctypes
import os
import re

import pytest

from fast_rpa_eval_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "frpa.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"defined", "if", "sizeof"}
    return sorted({n for n in names if n not in skip and (n.startswith("frpa_") or n.endswith("_") or
                                                            n.startswith("init_dfpt") or n.startswith("free_dfpt"))})


def test_header_declares_reference_and_frpa_entries():
    fns = declared_functions()
    for must in ("evaluate_first_order_rho_reduce_memory_c_v3_batches_atoms_cu_host_",
                 "sum_up_whole_potential_c_v3_atoms_full_points_j_atom_tile_cu_host_",
                 "init_dfpt_device_info", "free_dfpt_device_info", "init_dfpt_device_data", "free_dfpt_device_data",
                 "generate_first_order_rho_bias_part2_", "frpa_rho1_forward", "frpa_rho1_adjoint",
                 "frpa_rho1_bias_part2", "frpa_far_forward", "frpa_far_adjoint", "frpa_far_plan_set_near",
                 "frpa_near_forward"):
        assert must in fns


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_error_reporting_without_gpu_work():
    lib = _lib.lib()
    assert lib.frpa_version().startswith(b"fast_rpa_eval_amd")
    # a null plan is rejected before any device call
    rc = lib.frpa_rho1_forward(None, 1, 1, 0, None, None, None, None, None, None, None, None)
    assert rc == 1 and b"plan is null" in lib.frpa_last_error()
    rc = lib.frpa_far_forward(None, 1, 1, 1, 0, 0, None, None, None, None, None, None, None, None)
    assert rc == 1
