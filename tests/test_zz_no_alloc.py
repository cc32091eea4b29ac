"""wb_constrained_vmult_noalloc (the product on device pointers with a fixed static work
space) against wb_constrained_vmult on the same contexts: two calls in a row, the second on
a smaller system, so the second call works on the work space the first one left behind.
The two entry points run the same kernels, so the products agree bit for bit."""
import ctypes

import numpy as np
import pytest

from wavebem_amd import _lib as L
from wavebem_amd import mesh as M
from wavebem_amd.bem import BEMProblem, Config

pytestmark = pytest.mark.gpu


def test_noalloc_vmult_matches_constrained_vmult(gpu_available):
    import torch

    cases = (("2.5k", "fp64", M.wigley_preset("2.5k")),
             ("small", "fp32_ir", M.wigley_tank(nh=12, nt=4, nu=6, nd=8, ny=8, nzd=3)))
    sizes = []
    for name, prec, m in cases:
        ptr, idx = M.double_nodes_csr(m.support_points)
        bc, _ = M.wigley_bc(m, seed=11)
        x = np.random.default_rng(11).standard_normal(m.n_dofs)
        print(f"[noalloc] {name} n={m.n_dofs} {prec}", flush=True)
        with BEMProblem(m, Config(precision=prec), double_nodes=(ptr, idx)) as p:
            p.compute_constraints(bc)
            p.assemble_system()
            ref = p.constrained_vmult(x)
            xd = torch.tensor(x, dtype=torch.float64, device="cuda")
            yd = torch.full_like(xd, float("nan"))
            torch.cuda.synchronize()
            st = L.lib().wb_constrained_vmult_noalloc(p._ctx, ctypes.c_void_p(xd.data_ptr()),
                                                      ctypes.c_void_p(yd.data_ptr()))
            assert st == L.WB_OK, L.lib().wb_last_error(p._ctx)
            y = yd.cpu().numpy()
        assert np.array_equal(y, ref), float(np.abs(y - ref).max())
        sizes.append(m.n_dofs)
    assert sizes[1] < sizes[0]


def test_noalloc_vmult_rejects_null_and_unassembled(gpu_available):
    m = M.wigley_tank(nh=12, nt=4, nu=6, nd=8, ny=8, nzd=3)
    with BEMProblem(m, Config()) as p:
        st = L.lib().wb_constrained_vmult_noalloc(p._ctx, None, None)
        assert st == L.WB_E_STATE  # not assembled
        p.assemble_system()
        st = L.lib().wb_constrained_vmult_noalloc(p._ctx, None, None)
        assert st == L.WB_E_ARG
