"""cg_kinematic_substep_noalloc against cg_kinematic_substep on the same fields.

Two calls in a row on different contexts' states (a 24^3 RVE, then a smaller 16^3 one), each
at two step sizes (one that needs no sub-stepping and one that does), through the C-ABI with
device buffers from torch: the results agree bit for bit with the context entry point, and
nsub follows kinematic_substep's rule (various_functions.f90:2759-2760).  A grid above the
scratch capacity is rejected before anything is launched.
"""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_kinematic_matches_the_context_entry_point():
    import torch

    torch.cuda.init()
    from creep_modeling_amd import driver, inputs, lib, parse

    L = lib.load()
    stream = torch.cuda.current_stream().cuda_stream
    for n in (24, 16):
        print("no-alloc kinematic check, %d^3" % n, flush=True)
        ms, mat, procs = inputs.config_case(0, n=n, ngrains=6, nsteps=2)
        p = parse.problem_from_case(ms, mat, procs)
        g = driver.Solver(p)
        g.step()
        V = p.nvox
        vg = g.ctx.field("velgrad").reshape(V, 9)
        J = g.ctx.field("detF")
        ld = V + 37  # any column stride >= V
        dL = torch.zeros(9 * ld, dtype=torch.float64, device="cuda")
        dL.view(9, ld)[:, :V] = torch.as_tensor(vg.T.copy(), device="cuda")
        dJ = torch.as_tensor(J, device="cuda")
        for tdot in (1.0, 400.0):
            ref = (C.c_double * 4)()
            nsub = C.c_int()
            assert L.cg_kinematic_substep(g.h, tdot, ref, C.byref(nsub)) == 0
            out = torch.full((5,), float("nan"), dtype=torch.float64, device="cuda")
            rc = L.cg_kinematic_substep_noalloc(V, dL.data_ptr(), ld, dJ.data_ptr(), tdot, out.data_ptr(), stream)
            assert rc == 0, L.cg_last_error(None)
            torch.cuda.synchronize()
            o = out.cpu().numpy()
            print("tdot %g: %s nsub %d" % (tdot, o, nsub.value), flush=True)
            assert np.array_equal(o[:4], np.array(list(ref)))
            assert o[4] == nsub.value
            worst = max(1.0, ref[0] / 5.0e-2, ref[1] / 2.0, ref[2] / 5.0e-2)
            assert nsub.value == max(1, math.ceil(worst))
        assert nsub.value > 1  # the long step needs sub-stepping
        g.close()
    rc = L.cg_kinematic_substep_noalloc((1 << 24) + 1, dL.data_ptr(), (1 << 24) + 1, dJ.data_ptr(), 1.0,
                                        out.data_ptr(), stream)
    assert rc == -6 and b"scratch capacity" in L.cg_last_error(None)
