"""Failure injection at the C ABI (the reference's CudaFailureInjectionPoint and
its test cuda_tbem_launch_failure_is_typed_and_releases_transient_buffers,
tbem.rs:6740-6794): an armed one-shot failure makes the call return a typed
status before the boundary takes effect, publishes nothing, releases the
call's transient device buffers, and leaves the context usable."""

import ctypes
import gc

import numpy as np
import pytest

import sonium_amd as sa
from sonium_amd.mesh import icosphere

pytestmark = pytest.mark.gpu


def test_injected_failures_are_typed_and_release_transient_buffers():
    ctx = sa.Context(0)
    m = icosphere(1.0, 2)
    tm = sa.TbemMesh(ctx, m)
    ph = sa.PhysicsParams.from_wavenumber(1.0)
    beta = ph.burton_miller_beta()
    system = sa.TbemSystem(ctx, m.num_elements)
    gc.collect()  # objects of earlier tests release their device memory now, not during the checks
    baseline = sa.live_device_bytes()

    # launch: the one-shot assembly builds a workspace (transient routes), fails before its kernels
    def build():  # sbh_build_tbem_system: workspace (routes), assembly, workspace released
        sa._check(sa._lib().sbh_build_tbem_system(ctx.handle, tm.handle, ctypes.byref(ph._p), system.handle))

    ctx.inject_failure("launch")
    with pytest.raises(sa.SbhError) as e:
        build()
    assert e.value.status == 3 and "injected launch failure" in str(e.value)
    assert sa.live_device_bytes() == baseline, "transient assembly buffers must be released"
    # one shot: the same call succeeds next
    build()
    A = system.matrix
    assert np.all(np.isfinite(A)) and np.abs(A).max() > 0

    # allocation: refused before hipMalloc
    ctx.inject_failure("allocation")
    with pytest.raises(sa.SbhError) as e:
        sa.TbemSystem(ctx, 640)
    assert e.value.status == 4 and "injected allocation failure" in str(e.value)
    assert sa.live_device_bytes() == baseline

    op = sa.DenseMatrixOperator(system)
    x = np.ones(m.num_elements, dtype=np.complex128)
    y_ref = op.apply(x)
    base_op = sa.live_device_bytes()
    for point, text in (("h2d", "host-to-device"), ("d2h", "device-to-host"), ("synchronization", "synchronization")):
        ctx.inject_failure(point)
        with pytest.raises(sa.SbhError) as e:
            op.apply(x)
        assert e.value.status == 3 and text in str(e.value)
        assert sa.live_device_bytes() == base_op
        assert np.array_equal(op.apply(x), y_ref)  # the context recovers
    # a solver whose first product fails releases its Krylov workspace
    ctx.inject_failure("launch")
    with pytest.raises(sa.SbhError):
        sa.gmres(op, A @ x, sa.GmresConfig(restart=30, tolerance=1e-8))
    assert sa.live_device_bytes() == base_op
    sol = sa.gmres(op, A @ x, sa.GmresConfig(restart=30, tolerance=1e-8))
    assert sol.converged
    for o in (op, system, tm):
        o.close()
    ctx.close()
