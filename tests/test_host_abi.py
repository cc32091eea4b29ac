"""The C-ABI library without a device: it loads, exports every symbol the header
declares, its host-side planning agrees with the oracle and the mesh module, and calls
that need a device fail with a status instead of crashing."""
import ctypes
import subprocess

import numpy as np
import pytest

from tests import helpers as T
from wavebem_amd import _lib as L
from wavebem_amd import bem as B
from wavebem_amd import mesh as M


def test_library_exports_every_header_symbol():
    lib = L.lib()
    declared = L.header_symbols()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(declared) <= exported
    assert lib.wb_version().decode().startswith("wavebem_amd")


def test_config_defaults_match_reference_parameters():
    c = L.WbConfig()
    L.lib().wb_config_default(ctypes.byref(c))
    # prm-files/default.prm: quadrature gauss 4, singular order 5; AdditionalData(100)
    assert c.quadrature_order == 4 and c.singular_quadrature_order == 5
    assert c.gmres_max_tmp_vectors == 100 and c.solver_tolerance == 1e-10


def test_null_context_is_rejected():
    lib = L.lib()
    assert lib.wb_assemble(None) == L.WB_E_ARG
    assert lib.wb_set_masks(None, None, None) == L.WB_E_ARG
    assert lib.wb_last_error(None) == b"null context"


def test_invalid_config_is_rejected():
    c = L.WbConfig()
    L.lib().wb_config_default(ctypes.byref(c))
    c.quadrature_order = 0
    ctx = ctypes.c_void_p()
    assert L.lib().wb_create(ctypes.byref(ctx), ctypes.byref(c)) == L.WB_E_ARG
    assert not ctx.value


@pytest.mark.parametrize("kind", ["wigley_small", "box"])
def test_mesh_plan_double_nodes_and_permutation(kind):
    m = T.wigley_small() if kind == "wigley_small" else M.box_mesh(n=5)
    plan = B.host_mesh_plan(m)
    ptr, idx = M.double_nodes_csr(m.support_points)
    assert np.array_equal(plan["dn_ptr"], ptr) and np.array_equal(plan["dn_idx"], idx)
    assert np.array_equal(np.sort(plan["perm"]), np.arange(m.n_dofs))
    # singular pairs: for every row, the cells having a vertex coincident with it
    cls = np.array([idx[ptr[i]] for i in range(m.n_dofs)])  # smallest coincident dof
    cell_cls = cls[m.cell_dofs]
    expect = sum(int(np.any(cell_cls == cls[i], axis=1).sum()) for i in range(m.n_dofs))
    assert plan["n_singular_pairs"] == expect
    # every cell is listed once per tile it touches: at least once, at most 4 times
    assert m.n_cells <= plan["n_tile_cells"] <= 4 * m.n_cells


def test_device_calls_fail_cleanly_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    ctx = ctypes.c_void_p()
    st = L.lib().wb_create(ctypes.byref(ctx), None)
    assert st in (L.WB_E_HIP, L.WB_OK)
    if st == L.WB_OK:
        L.lib().wb_destroy(ctx)
