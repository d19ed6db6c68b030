"""NC.inp / NumCalc project input (nc_format.rs, numcalc/project.rs), host only.

Pinned on the reference's own fixtures (tests/golden/numcalc, copies of
crates/sonium-bem/tests/data) and the assertions its tests make on them
(nc_format.rs:836-959, project.rs:240-338, tests/mesh2hrtf_pipeline.rs:21-35),
plus malformed variants of the same text.
"""

import math
import os

import numpy as np
import pytest

from sonium_amd import numcalc as nc

HERE = os.path.dirname(os.path.abspath(__file__))
SAMPLE = os.path.join(HERE, "golden", "numcalc", "sample_project", "NC.inp")
BASIC = os.path.join(HERE, "golden", "numcalc", "mesh2hrtf_basic", "NC.inp")


def _text(path):
    with open(path) as f:
        return f.read()


def test_parse_sample_project():
    cfg = nc.parse_nc_input(SAMPLE)
    assert cfg.version == "Mesh2HRTF 1.0.0"
    assert cfg.description == ""  # the reference's parser never stores it
    assert cfg.control_params_i == [0, 0, 0, 0, 7, 0]
    assert cfg.control_params_ii == [1.0, 1.0, 1e-6, 0.0, 1.0, 0.0, 0.0]
    assert cfg.frequency_curve == [(0.0, 100.0, 0.0), (1e-6, 500.0, 0.0), (2e-6, 1000.0, 0.0)]
    assert cfg.main_params_i == nc.MainParamsI(2, 4, 4, 1, 0, 2, 0, 0, 0)
    assert cfg.main_params_ii == nc.MainParamsII(1, 1, 0, 0.0, 0, 0, 0)
    assert cfg.main_params_iii == [0, 0, 0, 0]
    assert cfg.main_params_iv == nc.MainParamsIV(343.0, 1.21, 1.0, [0.0, 0.0, 0.0, 0.0])
    base = os.path.dirname(SAMPLE)
    assert cfg.node_files == [os.path.join(base, "nodes.txt")]
    assert cfg.element_files == [os.path.join(base, "elements.txt")]
    assert cfg.boundary_conditions == [nc.BoundarySpec(0, 0, "VELO", 1.0, -1, 0.0, -1)]
    assert cfg.plane_waves == [] and cfg.point_sources == [] and cfg.symmetry is None
    assert cfg.bem_method() == "Traditional" and cfg.solver_method() == "BiCgstab"
    ph = cfg.to_physics_params(1000.0)
    assert ph.speed_of_sound == 343.0 and ph.density == 1.21 and ph.frequency == 1000.0


def test_load_sample_and_basic_projects():
    p = nc.load_project(SAMPLE)
    assert p.nodes.shape == (4, 3) and p.connectivity.shape == (4, 4)
    assert list(p.frequencies) == [100.0, 500.0, 1000.0]
    assert p.bc_type.tolist() == [0, 0, 0, 0]
    assert p.bc_value.tolist() == [1.0, 0.0, 0.0, 0.0]
    assert np.all(np.isnan(p.admittance.real))
    q = nc.load_project(BASIC)
    # tests/mesh2hrtf_pipeline.rs:21-35
    assert int(np.count_nonzero(~q.is_evaluation)) == 20
    assert len(q.evaluation_indices) == 6
    assert q.nodes.shape == (30, 3)
    # evaluation elements index the second node file (offset 12)
    assert q.connectivity[q.evaluation_indices, :3].min() == 12
    c = q.evaluation_centers()
    ref = np.array([q.nodes[q.connectivity[e, :3]].mean(axis=0) for e in q.evaluation_indices])
    assert np.allclose(c, ref, rtol=0, atol=1e-15)
    assert np.allclose(np.linalg.norm(c, axis=1), 0.5, atol=1e-3)
    mesh, bt, bv, adm = q.surface_arrays()
    assert mesh.num_elements == 20 and bv[0] == 1.0 and np.all(bv[1:] == 0)


def _mutate(old, new, path=SAMPLE):
    text = _text(path)
    assert old in text
    return text.replace(old, new)


@pytest.mark.parametrize("old,new,kind,same_line", [
    ("2 4 4 1 0 2 0 0 0", "2 not-a-node-count 4 1 0 2 0 0 0", "parse", True),   # nc_format.rs:906-915
    ("2 4 4 1 0 2 0 0 0", "2 4.0 4 1 0 2 0 0 0", "parse", True),
    ("2 4 4 1 0 2 0 0 0", "2 4 4 1 0 2 0 0", "parse", True),                    # nine values required
    ("ELEM 0 TO 0 VELO 1.0 -1 0.0 -1", "ELEM malformed boundary", "parse", True),  # :917-926
    ("0 3\n", "0 4\n", "parse", False),                                        # curve shorter than declared
    ("0 3\n", "0\n", "parse", True),
    ("END\n", "", "missing_section", False),                                   # :928-936
    ("NODES\n", "NODE\n", "missing_section", False),
    ("Mesh2HRTF 1.0.0", "NumCalc 1.0.0", "missing_section", False),
    ("343 1.21 1.0 0.0 0.0 0.0 0.0", "343 1.21", "parse", True),
    ("1 1 0.000001 0.00e+00 1 0 0", "1 1 0x10 0 1 0 0", "parse", True),         # no hexadecimal
])
def test_malformed_records_raise_typed_errors(old, new, kind, same_line):
    text = _text(SAMPLE)
    with pytest.raises(nc.NcParseError) as e:
        nc.parse_nc_input_string(_mutate(old, new))
    assert e.value.kind == kind
    if same_line:  # the error names the 1-based line of the record
        assert e.value.line == text[:text.index(old)].count("\n") + 1
    if kind == "missing_section":
        assert "missing required section" in str(e.value)


def test_rust_number_grammar():
    """Tokens Rust accepts for f64 and i32 parse; others are rejected or, in
    the lenient source lines, dropped (nc_format.rs:650-722)."""
    cfg = nc.parse_nc_input_string(_mutate("1 1 0.000001 0.00e+00 1 0 0", "+1 1. .5 5e-1 INF -infinity NaN"))
    v = cfg.control_params_ii
    assert v[:4] == [1.0, 1.0, 0.5, 0.5] and v[4] == math.inf and v[5] == -math.inf and math.isnan(v[6])
    cfg = nc.parse_nc_input_string(_mutate("0 0 0 0 7 0", "+0 -0 0 0 7 0"))
    assert cfg.control_params_i == [0, 0, 0, 0, 7, 0]
    for bad in ("1e", "e5", "--1", "1.2.3", "0x1p3"):
        with pytest.raises(nc.NcParseError):
            nc.parse_nc_input_string(_mutate("1 1 0.000001 0.00e+00 1 0 0", f"1 1 {bad} 0 1 0 0"))
    # plane-wave records drop unparsable tokens and need 8 numbers
    text = _mutate("RETU\n##\nEND", "RETU\n##\nPLANE WAVES\n1 0.0 -1.0 0.0 1.0 -1 0.0 -1\n2 0 0 1 0.5 x -1 0.25 -1\n##\nEND")
    cfg = nc.parse_nc_input_string(text)
    assert cfg.plane_waves[0] == nc.PlaneWaveSource((0.0, -1.0, 0.0), 1.0, -1, 0.0, -1)
    assert cfg.plane_waves[1] == nc.PlaneWaveSource((0.0, 0.0, 1.0), 0.5, -1, 0.25, -1)  # 'x' dropped
    bad = _mutate("RETU\n##\nEND", "RETU\n##\nPOINT SOURCES\n1 0 0 0 1 -1\n##\nEND")
    with pytest.raises(nc.NcParseError) as e:
        nc.parse_nc_input_string(bad)
    assert e.value.kind == "parse"


def test_symmetry_and_comments():
    text = _mutate("BOUNDARY", "SYMMETRY\n1 0 1\n0.0 0.5 0.0\n##\n# a comment line\nBOUNDARY")
    cfg = nc.parse_nc_input_string(text)
    assert cfg.symmetry == nc.SymmetryConfig((True, False, True), (0.0, 0.5, 0.0))


def _write_project(tmp_path, bc_lines, curve=None):
    src = os.path.dirname(SAMPLE)
    text = _text(SAMPLE).replace("ELEM 0 TO 0 VELO 1.0 -1 0.0 -1", bc_lines)
    if curve is not None:
        text = text.replace("0 3\n0.000000 1.000000e+02 0.0\n0.000001 5.000000e+02 0.0\n0.000002 1.000000e+03 0.0",
                            curve)
    for name in ("nodes.txt", "elements.txt"):
        with open(os.path.join(src, name)) as f, open(os.path.join(tmp_path, name), "w") as g:
            g.write(f.read())
    path = os.path.join(tmp_path, "NC.inp")
    with open(path, "w") as f:
        f.write(text)
    return path


def test_boundary_ranges_and_keywords(tmp_path):
    """apply_boundary_specs (project.rs:171-205): later records overwrite,
    ranges are inclusive, ADMI is velocity 0 with that admittance."""
    p = nc.load_project(_write_project(tmp_path, "ELEM 0 TO 3 PRES 0.5 -1 0.25 -1\nELEM 1 TO 2 ADMI 0.01 -1 -0.02 -1\n"
                                                 "ELEM 3 TO 3 VELO 2.0 -1 1.0 -1"))
    assert p.bc_type.tolist() == [1, 0, 0, 0]
    assert p.bc_value.tolist() == [0.5 + 0.25j, 0, 0, 2 + 1j]
    assert np.isnan(p.admittance[0].real) and np.isnan(p.admittance[3].real)
    assert p.admittance[1] == 0.01 - 0.02j and p.admittance[2] == 0.01 - 0.02j
    for bc, kind in (("ELEM 3 TO 1 VELO 1.0 -1 0.0 -1", "boundary_out_of_range"),
                     ("ELEM 0 TO 9 VELO 1.0 -1 0.0 -1", "boundary_out_of_range"),
                     ("ELEM 0 TO 1 IMPE 1.0 -1 0.0 -1", "unsupported_bc")):
        with pytest.raises(nc.ProjectError) as e:
            nc.load_project(_write_project(tmp_path, bc))
        assert e.value.kind == kind
    with pytest.raises(nc.NcParseError) as e:
        nc.load_project(os.path.join(tmp_path, "missing", "NC.inp"))
    assert e.value.kind == "io"


def test_frequency_curve_is_sorted_unique_positive(tmp_path):
    """expand_frequency_curve (project.rs:207-225)."""
    curve = "0 6\n0 500 0\n0 100 0\n0 -1 0\n0 nan 0\n0 inf 0\n0 500 0"
    p = nc.load_project(_write_project(tmp_path, "ELEM 0 TO 0 VELO 1.0 -1 0.0 -1", curve))
    assert list(p.frequencies) == [100.0, 500.0]


def test_sweep_output_json_layout(tmp_path):
    out = nc.SweepOutput(np.array([100.0]), 343.0, 1.21, 2, np.array([[1 + 2j, 3 - 4j]]), np.zeros((1, 3)),
                         np.array([[0.5j]]), [3], [1e-9])
    path = os.path.join(tmp_path, "out.json")
    out.write_json(path)
    import json
    d = json.load(open(path))
    assert d["surface_pressure"] == [[1.0, 2.0], [3.0, -4.0]] and d["eval_pressure"] == [[0.0, 0.5]]
    assert d["num_surface_dofs"] == 2 and d["frequencies"] == [100.0]
