"""Device-side admission: a context budget refuses allocations before they are
made (the reference admits BEM work before assembly, admission.rs:17-30,
bem_solver.rs:1767-1772), and requests beyond the device's free memory are
refused the same way without a budget."""

import numpy as np
import pytest

import sonium_amd as sa
from sonium_amd.mesh import icosphere

pytestmark = pytest.mark.gpu


def test_device_budget_refuses_before_allocating():
    ctx = sa.Context(0)
    assert ctx.device_budget() == (0, 0)
    n = 1280
    sys_bytes = 16 * (n * n + n) + 8 * n * ((n + 255) // 256 + 1)  # matrix, rhs, row maxima
    ctx.set_device_budget(sys_bytes + 1024)
    s1 = sa.TbemSystem(ctx, n)
    assert ctx.device_budget() == (sys_bytes + 1024, sys_bytes)
    with pytest.raises(sa.SbhError) as e:
        sa.TbemSystem(ctx, n)
    print(e.value)
    assert e.value.status == 4 and "admission refused" in str(e.value) and "budget" in str(e.value)
    assert ctx.device_budget()[1] == sys_bytes  # the refused call holds nothing
    # a solver workspace that does not fit is refused before its basis is allocated
    m = icosphere(1.0, 3)
    tm = sa.TbemMesh(ctx, m)
    ph = sa.PhysicsParams.from_wavenumber(2.0)
    ws = sa.TbemAssemblyWorkspace(ctx, tm)
    ws.assemble(ph, ph.burton_miller_beta(), s1)
    op = sa.DenseMatrixOperator(s1)
    B = np.ones((n, 8), dtype=np.complex128)
    with pytest.raises(sa.SbhError) as e:
        sa.block_gmres(op, B, sa.BlockGmresConfig(restart=50, tolerance=1e-8))
    assert e.value.status == 4 and "block Krylov basis" in str(e.value)
    ctx.set_device_budget(0)
    sol = sa.block_gmres(op, B, sa.BlockGmresConfig(restart=50, tolerance=1e-8))
    assert sol.converged
    for o in (op, ws, tm):
        o.close()
    s1.close()
    assert ctx.device_budget() == (0, 0)
    # beyond the device's free memory (1.44 TB): refused without a budget
    with pytest.raises(sa.SbhError) as e:
        sa.TbemSystem(ctx, 300000)
    assert e.value.status == 4 and "free of" in str(e.value)
    ctx.close()
