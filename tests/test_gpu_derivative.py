"""GPU tests of the parameter-derivative assembly
(build_tbem_system_stabilized_*_derivative, tbem.rs:1734-2019;
CudaTbemAssemblyWorkspace::assemble_wavenumber_derivative, tbem.rs:4399-4591).

* the wavenumber derivative of rigid and general-BC systems (matrix and BC
  right-hand side) agrees with the oracle to 1e-12 relative Frobenius, for
  both signs and a scaled beta;
* the sound-speed, density and impedance variants agree with the oracle;
* at 1280 panels the derivative equals the central difference of the GPU's
  own value assembly (a size-independent check of every term);
* the derivative of the solution, dx/dk = A^-1 (db/dk - dA/dk x), predicts
  the solution at k + dk to second order.
"""

import numpy as np
import pytest

import oracle
import sonium_amd as sa
from sonium_amd.mesh import ellipsoid, icosphere

from helpers import rel_fro, rel_l2

pytestmark = pytest.mark.gpu


def _mixed(n, seed):
    rng = np.random.default_rng(seed)
    bt = (rng.random(n) < 0.3).astype(np.int32)
    bv = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) * (rng.random(n) < 0.5)
    adm = np.where((rng.random(n) < 0.4) & (bt == 0), 0.003 - 0.001j, np.nan)
    return dict(bc_type=bt, bc_value=bv, admittance=adm)


@pytest.mark.parametrize("sign,beta_scale", [(1.0, 1.0), (-1.0, 0.2)])
def test_rigid_wavenumber_derivative_matches_oracle(ctx, sign, beta_scale):
    m = icosphere(1.0, 2)
    k = 2.0
    tm = sa.TbemMesh(ctx, m)
    ws = sa.TbemAssemblyWorkspace(ctx, tm)
    ph = sa.PhysicsParams.from_wavenumber(k)
    d = ws.assemble_wavenumber_derivative(ph, sign=sign, beta_scale=beta_scale)
    beta = ph.burton_miller_beta_scaled(beta_scale)
    dA, db = oracle.build_tbem_derivative(m.nodes, m.triangles, k, beta,
                                          ph.burton_miller_beta_wavenumber_derivative() * beta_scale, sign=sign)
    assert rel_fro(d.matrix, dA) < 1e-12
    assert np.all(d.rhs == 0) and np.all(db == 0)


def test_general_bc_wavenumber_derivative_matches_oracle(ctx):
    m = ellipsoid([1.2, 0.8, 1.0], 2)
    n = m.num_elements
    bc = _mixed(n, 3)
    k = 2.6
    tm = sa.TbemMesh(ctx, m, **bc)
    ws = sa.TbemAssemblyWorkspace(ctx, tm)
    ph = sa.PhysicsParams.from_wavenumber(k)
    d = ws.assemble_wavenumber_derivative(ph, beta_scale=0.7)
    dA, db = oracle.build_tbem_derivative(m.nodes, m.triangles, k, ph.burton_miller_beta_scaled(0.7),
                                          ph.burton_miller_beta_wavenumber_derivative() * 0.7, **bc)
    assert rel_fro(d.matrix, dA) < 1e-12
    assert rel_l2(d.rhs, db) < 1e-12


@pytest.mark.parametrize("kind", ["speed", "density", "impedance_density", "impedance_speed"])
def test_parameter_derivative_variants_match_oracle(ctx, kind):
    m = icosphere(1.0, 2)
    n = m.num_elements
    k = 1.7
    ph = sa.PhysicsParams.from_wavenumber(k)
    rho, c = ph.density, ph.speed_of_sound
    if kind.startswith("impedance"):
        has = np.random.default_rng(5).random(n) < 0.5
        bc = dict(admittance=np.where(has, 1.0 / (rho * c) + 0j, np.nan), bc_value=np.where(has, 0.0, 0.3 - 0.1j))
    else:
        bc = _mixed(n, 7)
    tm = sa.TbemMesh(ctx, m, **bc)
    ws = sa.TbemAssemblyWorkspace(ctx, tm)
    d = getattr(ws, f"assemble_{kind}_derivative")(ph)
    wsc = {"speed": -k / c, "density": 0.0, "impedance_density": 0.0, "impedance_speed": -k / c}[kind]
    pfd = {"speed": 0.0, "density": ph.pressure_factor / rho, "impedance_density": ph.pressure_factor / rho,
           "impedance_speed": 0.0}[kind]
    adm = {"speed": 0.0, "density": 0.0, "impedance_density": -1.0 / rho, "impedance_speed": -1.0 / c}[kind]
    dA, db = oracle.build_tbem_derivative(m.nodes, m.triangles, k, ph.burton_miller_beta(),
                                          ph.burton_miller_beta_wavenumber_derivative() * wsc, wavenumber_scale=wsc,
                                          pressure_factor_derivative=pfd, admittance_scale=adm, **bc)
    # relative to the derivative, or to 1e-3 of the value matrix where the
    # derivative cancels exactly (Y i rho omega is density independent)
    scale = np.linalg.norm(ws.assemble(ph).matrix)
    assert np.linalg.norm(d.matrix - dA) <= 1e-12 * max(np.linalg.norm(dA), 1e-3 * scale)
    assert np.linalg.norm(d.rhs - db) <= 1e-12 * max(np.linalg.norm(db), 1e-300)


def test_wavenumber_derivative_equals_central_difference_of_gpu_assembly(ctx):
    m = icosphere(1.0, 3)
    k, h = 2.0, 2e-4
    tm = sa.TbemMesh(ctx, m)
    ws = sa.TbemAssemblyWorkspace(ctx, tm)
    ph = sa.PhysicsParams.from_wavenumber(k)
    d = ws.assemble_wavenumber_derivative(ph).matrix
    A = []
    for kk in (k + h, k - h):
        p = sa.PhysicsParams.from_wavenumber(kk)
        A.append(ws.assemble(p, p.burton_miller_beta()).matrix)
    fd = (A[0] - A[1]) / (2 * h)
    assert rel_fro(d, fd) < 1e-7


def test_solution_derivative_predicts_neighbouring_frequency(ctx):
    """dx/dk from A dx = db - dA x (the sweep predictor the derivative exists
    for): x(k + dk) - (x + dk dx/dk) is O(dk^2), much smaller than the change."""
    m = icosphere(1.0, 3)
    k, dk = 2.0, 0.02
    tm = sa.TbemMesh(ctx, m)
    ws = sa.TbemAssemblyWorkspace(ctx, tm)
    inc = sa.IncidentField.plane_wave_z()

    def solve(kk):
        p = sa.PhysicsParams.from_wavenumber(kk)
        s = ws.assemble(p, p.burton_miller_beta())
        b = s.rhs + inc.compute_rhs_with_beta(tm, p, p.burton_miller_beta())
        return s.matrix, b, np.linalg.solve(s.matrix, b)

    A, b, x = solve(k)
    ph = sa.PhysicsParams.from_wavenumber(k)
    dA = ws.assemble_wavenumber_derivative(ph).matrix
    # d(incident rhs)/dk by central difference (the incident derivative is not part of this API)
    p1, p0 = sa.PhysicsParams.from_wavenumber(k + 1e-5), sa.PhysicsParams.from_wavenumber(k - 1e-5)
    db = (inc.compute_rhs_with_beta(tm, p1, p1.burton_miller_beta())
          - inc.compute_rhs_with_beta(tm, p0, p0.burton_miller_beta())) / 2e-5
    dx = np.linalg.solve(A, db - dA @ x)
    _, _, x1 = solve(k + dk)
    change = np.linalg.norm(x1 - x)
    pred_err = np.linalg.norm(x1 - (x + dk * dx))
    assert pred_err < 0.05 * change, (pred_err, change)
