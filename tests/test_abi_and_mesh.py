"""CPU checks of the C ABI library and the host-side inputs (no GPU work).

* libsbh.so loads and exports every function declared in include/sbh.h;
* the physics and Burton–Miller helpers follow types.rs:190-218, 314-320, 418-424;
* on a machine without a GPU, context creation fails with a status and a
  message instead of crashing;
* the mesh generators reproduce the reference's panel counts and orientation
  (generators.rs:35-77, 309-427, 779-870).
"""

import math
import os
import re

import numpy as np
import pytest

import oracle
import sonium_amd as sa
from sonium_amd import _ffi
from sonium_amd.mesh import ellipsoid, fibonacci_directions, fibonacci_sphere, icosphere

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "sbh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(sbh_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = _ffi.load()
    declared = _declared_functions()
    assert len(declared) > 40
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    # the Python bindings cover the whole header
    assert declared <= set(_ffi.SIGNATURES), sorted(declared - set(_ffi.SIGNATURES))


def test_physics_params_follow_reference_formulas():
    ph = sa.PhysicsParams(100.0, 343.0, 1.21)
    assert ph.omega == 2.0 * math.pi * 100.0
    assert ph.wave_number == ph.omega / 343.0
    assert ph.wave_length == 343.0 / 100.0
    assert ph.harmonic_factor == 1.0 and ph.tau == 1.0
    assert ph.pressure_factor == 1.21 * ph.omega
    assert ph.burton_miller_beta() == 1j / ph.wave_number
    assert ph.burton_miller_beta_scaled(0.2) == 0.2j / ph.wave_number
    interior = sa.PhysicsParams(100.0, is_internal=True)
    assert interior.tau == -1.0 and interior.burton_miller_beta() == 0
    k = sa.PhysicsParams.from_wavenumber(10.0)
    assert k.wave_number == 10.0
    with pytest.raises(sa.SbhError):
        sa.PhysicsParams(-1.0)


def _has_gpu():
    try:
        import ctypes

        h = ctypes.c_void_p()
        st = _ffi.load().sbh_context_create(0, ctypes.byref(h))
        if st == 0:
            _ffi.load().sbh_context_destroy(h)
            return True
        return False
    except OSError:
        return False


def test_context_without_gpu_reports_an_error():
    if _has_gpu():
        pytest.skip("a GPU is present")
    with pytest.raises(sa.SbhError) as e:
        sa.Context(0)
    assert e.value.status in (1, 3)


def test_icosphere_counts_and_orientation():
    for s, (v, f) in enumerate([(12, 20), (42, 80), (162, 320), (642, 1280)]):
        m = icosphere(1.0, s)
        assert m.nodes.shape[0] == v and m.num_elements == f
        assert np.allclose(np.linalg.norm(m.nodes, axis=1), 1.0)
        p = m.nodes[m.triangles]
        cr = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
        assert np.all(np.einsum("ij,ij->i", cr, p.mean(axis=1)) > 0)  # vertex order is outward
    # C2 size
    assert icosphere(1.0, 5).num_elements == 20480


def test_element_geometry_matches_oracle():
    m = ellipsoid([1.3, 0.8, 0.6], 2)
    c, n, a = m.element_geometry()
    co, no, ao = oracle.element_geometry(m.nodes, m.triangles)
    assert np.array_equal(c, co) and np.array_equal(n, no) and np.array_equal(a, ao)
    # polyhedral area lies between the spheres on the smallest and largest semi-axis
    assert 4 * math.pi * 0.6 ** 2 < a.sum() < 4 * math.pi * 1.3 ** 2


@pytest.mark.parametrize("nt,nphi,count", [(65, 80, 10240), (129, 160, 40960)])
def test_fibonacci_hull_counts(nt, nphi, count):
    m = fibonacci_sphere(1.0, nt, nphi)
    v = (nt - 1) * nphi + 2
    assert m.nodes.shape[0] == v
    assert m.num_elements == count == 2 * v - 4
    p = m.nodes[m.triangles]
    cr = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    assert np.all(np.einsum("ij,ij->i", cr, p.mean(axis=1)) > 0)


def test_fibonacci_directions_are_unit_and_spread():
    d = fibonacci_directions(64)
    assert d.shape == (64, 3)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0)
    assert np.allclose(d[:, 2], 1.0 - (2 * np.arange(64) + 1) / 64)
    assert np.linalg.norm(d.mean(axis=0)) < 0.05


def test_mesh_rejects_bad_input():
    with pytest.raises(ValueError):
        sa.Mesh(np.zeros((4, 3)), np.array([[0, 1, 7]]))
    with pytest.raises(ValueError):
        ellipsoid([1.0, -1.0, 1.0], 1)
