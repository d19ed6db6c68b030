"""Concurrent use of one context from several host threads.

The reference's LinearOperator is Send + Sync and its callers apply one
operator from rayon workers (traits.rs:2949-3290); the C ABI serialises every
call on a context with the context's operation lock, so concurrent calls
return exactly the bits of the same calls made one at a time.  ctypes
releases the GIL during each foreign call, so the threads below really do
enter the library together.  The transposed product grows the context's
shared device scratch and the solves use its pinned staging buffer: the two
pieces of shared state the lock protects.
"""

from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import sonium_amd as sa
from sonium_amd.mesh import icosphere

pytestmark = pytest.mark.gpu


def test_concurrent_calls_on_one_context_match_serial_bits():
    ctx = sa.Context(0)
    m = icosphere(1.0, 3)
    tm = sa.TbemMesh(ctx, m)
    ph = sa.PhysicsParams.from_wavenumber(2.0)
    beta = ph.burton_miller_beta()
    system = sa.TbemAssemblyWorkspace(ctx, tm).assemble(ph, beta)
    sa.IncidentField.plane_wave_z().add_to_system(system, tm, ph, beta)
    op = sa.DenseMatrixOperator(system)
    n = op.num_rows()
    rng = np.random.default_rng(4)
    xs = [rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(4)]
    b = system.rhs
    cfg = sa.GmresConfig(max_iterations=10, restart=50, tolerance=1e-8)
    small = sa.TbemAssemblyWorkspace(ctx, sa.TbemMesh(ctx, icosphere(1.0, 1))).assemble(ph, beta)
    op_small = sa.DenseMatrixOperator(small)
    xs_small = rng.standard_normal(80) + 0j

    def jobs():
        out = []
        for i in range(4):
            out.append(("apply", i))
            out.append(("transpose", i))
            out.append(("hermitian", i))
        out.append(("gmres", 0))
        out.append(("transpose_small", 0))
        return out

    def run(job):
        kind, i = job
        if kind == "apply":
            return op.apply(xs[i])
        if kind == "transpose":
            return op.apply_transpose(xs[i])
        if kind == "hermitian":
            return op.apply_hermitian(xs[i])
        if kind == "transpose_small":
            return op_small.apply_transpose(xs_small)
        return sa.gmres(op, b, cfg).x

    serial = {job: run(job) for job in jobs()}
    work = jobs() * 6
    rng.shuffle(work)
    with ThreadPoolExecutor(max_workers=6) as pool:
        results = list(pool.map(run, work))
    for job, y in zip(work, results):
        assert np.array_equal(y, serial[job]), f"{job}: concurrent result differs from the serial one"
    op_small.close()
    small.close()
    op.close()
    system.close()
    ctx.close()
