"""Allocation failure paths: a context whose HBM budget is too small refuses the work with
EDM_ENOMEM and gives back everything it had allocated for it (no device memory leaks), for the
entry points that allocate several buffers: all-sector enumeration, the stored matrix, the
block eigensolver and the block matvec scratch."""
import numpy as np
import pytest

import momentumed as me
from momentumed import device as dev
from momentumed import workloads as W
from momentumed._lib import EDM_ENOMEM, EDMError

pytestmark = pytest.mark.gpu


def _system(para, N_each):
    return dev.system_struct(para.k_list, para.Gk, para.Nc_mix, para.Nc_conserve, N_each)


def test_budget_exhaustion_releases_everything(gpu_ctx):
    print("[budget] enumerate_all / csr / block eigsolve under a small HBM budget", flush=True)
    w = W.medium()
    sysd = _system(w.para, w.N_each)
    ctx = dev.Context(0, hbm_budget_bytes=24 << 20)
    try:
        small = dev.DeviceBasis.enumerate(ctx, _system(W.small().para, (5,)), 0)
        small.close()
        used0, budget = ctx.memory()
        assert budget == 24 << 20
        # 27 sectors of ~173k states: 37.5 MB of states alone
        with pytest.raises(EDMError) as e:
            dev.DeviceBasis.enumerate_all(ctx, sysd)
        assert e.value.code == EDM_ENOMEM
        assert ctx.memory()[0] == used0
        # one sector (1.4 MB) fits; its stored matrix (~720 MB) does not
        b = dev.DeviceBasis.enumerate(ctx, sysd, 0)
        op = dev.DeviceOperator.from_mboperator(ctx, w.operator(), w.para.bits)
        used1 = ctx.memory()[0]
        with pytest.raises(EDMError) as e:
            dev.DeviceCSR(op, b)
        assert e.value.code == EDM_ENOMEM
        assert ctx.memory()[0] == used1
        # a block Lanczos workspace of 44 vectors (122 MB) does not fit either
        lm = dev.ROCmLinearMap(op, b)
        with pytest.raises(EDMError) as e:
            dev.krylov_map_solve(lm, 6, krylovdim=40, blocksize=4, want_vectors=False)
        assert e.value.code == EDM_ENOMEM
        assert ctx.memory()[0] == used1
        op.close()
        b.close()
        assert ctx.memory()[0] == used0
    finally:
        ctx.close()
