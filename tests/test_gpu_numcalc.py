"""GPU runs of NumCalc projects (run_project, crates/sonium-bem/src/numcalc/runner.rs:51-125).

* mesh2hrtf_basic (20-panel head, one vibrating element, 6 evaluation
  elements, 3 frequencies): surface pressure and evaluation-grid pressure agree
  with the oracle pipeline (assembly with beta = 4i/k, direct solve,
  constant_cauchy_data + field) to 1e-8, and satisfy the reference's own
  pipeline assertions (tests/mesh2hrtf_pipeline.rs:37-84);
* the frequency range is clamped as the reference's runner tests require
  (runner.rs:196-221).
"""

import os

import numpy as np
import pytest

import oracle
from sonium_amd import numcalc as nc

from helpers import rel_l2

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SAMPLE = os.path.join(HERE, "golden", "numcalc", "sample_project", "NC.inp")
BASIC = os.path.join(HERE, "golden", "numcalc", "mesh2hrtf_basic", "NC.inp")


def _oracle_sweep(project, freqs):
    mesh, bt, bv, adm = project.surface_arrays()
    c, nn, _ = oracle.element_geometry(mesh.nodes, mesh.triangles)
    pts = project.evaluation_centers()
    xs, fields = [], []
    for f in freqs:
        k = project.physics_at(f).wave_number
        A, rhs = oracle.build_tbem_system(mesh.nodes, mesh.triangles, k, 4j / k, bc_type=bt, bc_value=bv,
                                          admittance=adm, density=project.density,
                                          speed_of_sound=project.speed_of_sound)
        x = np.linalg.solve(A, rhs)  # no incident wave: the vibrating element drives the system
        p, v = oracle.cauchy_data(c, nn, x, k, bc_type=bt, bc_value=bv, admittance=adm, density=project.density,
                                  speed_of_sound=project.speed_of_sound)
        fields.append(oracle.field(pts, mesh.nodes, mesh.triangles, p, k, v_surf=v, density=project.density,
                                   speed_of_sound=project.speed_of_sound))
        xs.append(x)
    return np.array(xs), np.array(fields)


def test_mesh2hrtf_basic_sweep_matches_oracle(ctx):
    project = nc.load_project(BASIC)
    out = nc.run_project(project, ctx, tolerance=1e-11)
    assert list(out.frequencies) == [500.0, 1000.0, 2000.0]
    assert out.surface_pressure.shape == (3, 20) and out.eval_pressure.shape == (3, 6)
    xs, fields = _oracle_sweep(project, out.frequencies)
    for i in range(3):
        assert rel_l2(out.surface_pressure[i], xs[i]) < 1e-8
        assert rel_l2(out.eval_pressure[i], fields[i]) < 1e-8
    assert np.all(np.isfinite(out.eval_pressure)) and np.all(np.abs(out.eval_pressure) > 0)
    # tests/mesh2hrtf_pipeline.rs:63-84: the +z microphone hears the +z "ear" best
    mag = np.abs(out.eval_pressure[0])
    assert mag[4] > mag[:4].mean()
    # every solve (warm-started from the previous frequency) meets the tolerance
    assert all(r <= 1e-11 for r in out.residuals)


def test_run_project_frequency_ranges(ctx):
    project = nc.load_project(SAMPLE)
    out = nc.run_project(project, ctx, 1, 3, max_iterations=50, tolerance=1e-4)
    assert list(out.frequencies) == list(project.frequencies[1:3])
    assert out.surface_pressure.shape == (2, 4)
    for lo, hi in ((99, 200), (2, 0)):
        empty = nc.run_project(project, ctx, lo, hi, max_iterations=50, tolerance=1e-4)
        assert len(empty.frequencies) == 0 and empty.surface_pressure.size == 0


def test_point_source_project_matches_oracle(ctx, tmp_path):
    """A project whose incident field is a POINT SOURCES record (pick_incident,
    runner.rs:160-180) on the mesh2hrtf_basic head, rigid."""
    import shutil
    src_dir = os.path.dirname(BASIC)
    for name in os.listdir(src_dir):
        shutil.copy(os.path.join(src_dir, name), tmp_path / name)
    text = open(BASIC).read().replace("ELEM 0 TO 0 VELO 1.0 -1 0.0 -1\n", "")
    text = text.replace("RETU\n##\nEND", "RETU\n##\nPOINT SOURCES\n1 0.0 0.3 0.2 2.0 -1 0.5 -1\n##\nEND")
    (tmp_path / "NC.inp").write_text(text)
    project = nc.load_project(str(tmp_path / "NC.inp"))
    assert project.config.point_sources[0].position == (0.0, 0.3, 0.2)
    out = nc.run_project(project, ctx, tolerance=1e-11)
    mesh, bt, bv, adm = project.surface_arrays()
    c, nn, _ = oracle.element_geometry(mesh.nodes, mesh.triangles)
    pts = project.evaluation_centers()
    src = dict(positions=[[0.0, 0.3, 0.2]], strengths=[2.0 + 0.5j])
    for i, f in enumerate(out.frequencies):
        k = project.physics_at(f).wave_number
        A, rhs = oracle.build_tbem_system(mesh.nodes, mesh.triangles, k, 4j / k)
        x = np.linalg.solve(A, rhs + oracle.source_rhs(c, nn, k, 4j / k, **src))
        assert rel_l2(out.surface_pressure[i], x) < 1e-8
        ref = oracle.field(pts, mesh.nodes, mesh.triangles, x, k, **src)
        assert rel_l2(out.eval_pressure[i], ref) < 1e-8
