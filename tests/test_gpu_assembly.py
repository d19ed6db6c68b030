"""GPU parity of the Burton–Miller assembly against the CPU oracle.

The oracle restates build_tbem_system_with_beta_parallel
(crates/sonium-bem/src/core/assembly/tbem.rs:1284-1466).  The bar follows the
north star: relative Frobenius error <= 1e-12 for the FP64 matrix and rhs
(the reference's own CUDA-vs-CPU gate is 1e-9, tbem.rs:6618-6625).
"""

import numpy as np
import pytest

import oracle
import sonium_amd as sa
from sonium_amd.mesh import ellipsoid, icosphere

from helpers import rel_fro

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _assemble_gpu(ctx, m, k, beta, **bc):
    tm = sa.TbemMesh(ctx, m, **bc)
    ph = sa.PhysicsParams.from_wavenumber(k)
    sys = sa.build_tbem_system_with_beta(ctx, tm, ph, beta)
    return sys.matrix, sys.rhs, tm, ph


def _assemble_cpu(m, k, beta, **bc):
    return oracle.build_tbem_system(m.nodes, m.triangles, k, beta, **bc)


def test_device_sincos_accuracy(ctx):
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(-60.0, 60.0, 20000), np.arange(-40, 41) * (np.pi / 2),
                        np.array([0.0, 1e-12, -1e-9, 3.0, 1e5])])
    s, c = sa.device_sincos(ctx, x)
    # absolute error against the C library's sin/cos: a few ulp of 1
    assert np.max(np.abs(s - np.sin(x))) < 4e-16 * max(1.0, 1.0)
    assert np.max(np.abs(c - np.cos(x))) < 4e-16


def test_device_rsqrt_accuracy(ctx):
    """Hardware estimate plus one third-order step: within about one ulp over
    the distances squared the kernels see (1e-30 .. 1e8)."""
    rng = np.random.default_rng(8)
    x = np.concatenate([10.0 ** rng.uniform(-30.0, 8.0, 50000), np.array([1.0, 4.0, 0.25, 2.0, 3.0])])
    y = sa.device_rsqrt(ctx, x)
    ref = 1.0 / np.sqrt(x)  # correctly rounded sqrt and division: within 1 ulp of 1/sqrt
    assert np.max(np.abs(y - ref) / ref) < 3.4e-16


@pytest.mark.parametrize("subdiv,k", [(2, 1.0), (3, 2.0)])
def test_rigid_sphere_matrix_matches_oracle(ctx, subdiv, k):
    """Config C1 (1280 panels, k=2, beta=i/k) and the reference's s=2, k=1 CUDA fixture."""
    m = icosphere(1.0, subdiv)
    beta = 1j / k
    A, b, _, _ = _assemble_gpu(ctx, m, k, beta)
    Ar, br = _assemble_cpu(m, k, beta)
    print(f"N={m.num_elements} k={k}: rel Frobenius {rel_fro(A, Ar):.3e}")
    assert rel_fro(A, Ar) <= TOL
    assert np.array_equal(b, br)  # rigid: zero BC rhs


def test_vibrating_panel_rhs_matches_oracle(ctx):
    """The reference's CUDA parity fixture: icosphere s=2, k=1, elements[0] velocity 1 (tbem.rs:6561-6572)."""
    m = icosphere(1.0, 2)
    k = 1.0
    beta = 1j / k
    bv = np.zeros(m.num_elements, dtype=np.complex128)
    bv[0] = 1.0
    A, b, _, _ = _assemble_gpu(ctx, m, k, beta, bc_value=bv)
    Ar, br = _assemble_cpu(m, k, beta, bc_value=bv)
    assert rel_fro(A, Ar) <= TOL
    assert rel_fro(b, br) <= TOL


def test_pressure_and_admittance_columns_match_oracle(ctx):
    m = icosphere(1.0, 2)
    n = m.num_elements
    rng = np.random.default_rng(3)
    k = 3.0
    beta = 0.2j / k
    bt = (rng.random(n) < 0.3).astype(np.int32)
    bv = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) * (rng.random(n) < 0.5)
    adm = np.where(rng.random(n) < 0.3, 0.01 * (rng.standard_normal(n) + 1j), np.nan)
    adm = np.where(bt == 0, adm, np.nan)
    A, b, _, _ = _assemble_gpu(ctx, m, k, beta, bc_type=bt, bc_value=bv, admittance=adm)
    Ar, br = _assemble_cpu(m, k, beta, bc_type=bt, bc_value=bv, admittance=adm)
    assert rel_fro(A, Ar) <= TOL
    assert rel_fro(b, br) <= TOL


@pytest.mark.parametrize("k", [0.5, 4.0, 8.0])
def test_self_panel_tiers_on_coarse_mesh(ctx, k):
    """80-panel icosphere: k*sqrt(area) spans the 4 QuadratureParams::for_ka tiers (singular.rs:51-85)."""
    m = icosphere(1.0, 1)
    beta = 1j / k
    A, _, _, _ = _assemble_gpu(ctx, m, k, beta)
    Ar, _ = _assemble_cpu(m, k, beta)
    assert rel_fro(A, Ar) <= TOL
    # diagonal alone (self integrals)
    assert rel_fro(np.diag(A), np.diag(Ar)) <= TOL


def test_ellipsoid_uneven_panels_match_oracle(ctx):
    m = ellipsoid([1.4, 0.7, 0.55], 3)
    k = 5.0
    beta = 1j / k
    A, _, _, _ = _assemble_gpu(ctx, m, k, beta)
    Ar, _ = _assemble_cpu(m, k, beta)
    assert rel_fro(A, Ar) <= TOL


def test_workspace_reuse_across_frequencies(ctx):
    """CudaTbemAssemblyWorkspace semantics: routes built once, assemble per k."""
    m = icosphere(1.0, 2)
    tm = sa.TbemMesh(ctx, m)
    ws = sa.TbemAssemblyWorkspace(ctx, tm)
    sys = sa.TbemSystem(ctx, m.num_elements)
    for k in (0.7, 2.5):
        ph = sa.PhysicsParams.from_wavenumber(k)
        ws.assemble(ph, ph.burton_miller_beta(), sys)
        Ar, _ = _assemble_cpu(m, k, 1j / k)
        assert rel_fro(sys.matrix, Ar) <= TOL
    st = ws.stats
    assert st["near_pairs"] > 0 and st["leaves"] >= st["near_pairs"]


def test_plane_wave_rhs_matches_oracle(ctx):
    m = icosphere(1.0, 3)
    k = 2.0
    beta = 1j / k
    tm = sa.TbemMesh(ctx, m)
    ph = sa.PhysicsParams.from_wavenumber(k)
    c, n, _ = m.element_geometry()
    dirs = np.array([[0.0, 0.0, 1.0], [1.0, 2.0, -0.5], [0.3, -0.2, 0.9]])
    amps = np.array([1.0, 0.5 - 0.25j, 2.0])
    inc = sa.IncidentField(dirs, amps)
    got = inc.compute_rhs_with_beta(tm, ph, beta)
    for col in range(3):
        d = dirs[col] / np.linalg.norm(dirs[col])
        ref = oracle.plane_wave_rhs(c, n, k, beta, direction=d, amplitude=amps[col])
        assert np.max(np.abs(got[:, col] - ref)) <= 1e-14 * np.max(np.abs(ref))
