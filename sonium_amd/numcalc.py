"""Mesh2HRTF / NumCalc projects: NC.inp input and a GPU frequency sweep.

Mirrors crates/sonium-bem/src/core/io/nc_format.rs (NcInputConfig,
parse_nc_input, parse_nc_input_string, load_nc_nodes, load_nc_elements) and
crates/sonium-bem/src/numcalc/project.rs / runner.rs (load_project,
NumCalcProject, run_project, SweepOutput).  Parsing and loading run in the
native library (sbh_nc_*); run_project drives the GPU assembly, GMRES and
field evaluation.
"""

from __future__ import annotations

import ctypes
import json
import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from . import _ffi
from . import (Context, DenseMatrixOperator, GmresConfig, IncidentField, PhysicsParams, SbhError,
               TbemAssemblyWorkspace, TbemMesh, _check, _cptr, _lib, compute_total_field, gmres_with_guess)
from .mesh import Mesh

_VERSION, _DESCRIPTION, _NODE_FILE, _ELEMENT_FILE, _BOUNDARY_TYPE = range(5)
(_CONTROL_I, _CONTROL_II, _MAIN_III, _MAIN_IV_RESERVED, _FREQUENCY_CURVE, _FREQUENCIES, _BOUNDARY_SPECS,
 _PLANE_WAVES, _POINT_SOURCES) = range(9)


class NcParseError(Exception):
    """NcParseError (nc_format.rs:184-203): kind is "parse", "missing_section" or "io"."""

    def __init__(self, kind: str, message: str, line: Optional[int] = None):
        super().__init__(message)
        self.kind = kind
        self.line = line


class ProjectError(Exception):
    """ProjectError (project.rs:18-33): kind is "boundary_out_of_range" or "unsupported_bc"."""

    def __init__(self, kind: str, message: str):
        super().__init__(message)
        self.kind = kind


def _raise_mapped(e: SbhError):
    msg = str(e)
    if "parse error at line " in msg:
        tail = msg.split("parse error at line ", 1)[1]
        raise NcParseError("parse", msg, int(tail.split(":", 1)[0])) from None
    if "missing required section: " in msg:
        raise NcParseError("missing_section", msg) from None
    if "IO error: " in msg:
        raise NcParseError("io", msg) from None
    if "boundary spec covers element" in msg:
        raise ProjectError("boundary_out_of_range", msg) from None
    if "unsupported boundary condition keyword" in msg:
        raise ProjectError("unsupported_bc", msg) from None
    raise e


@dataclass
class MainParamsI:
    element_type: int = 0
    num_nodes: int = 0
    num_elements: int = 0
    num_object_files: int = 0
    num_eval_files: int = 0
    bc_type: int = 0
    solver_method: int = 0
    fmm_method: int = 0
    parallel: int = 0


@dataclass
class MainParamsII:
    preconditioner: int = 0
    iterative_solver: int = 0
    reserved1: int = 0
    reserved2: float = 0.0
    output_level: int = 0
    reserved3: int = 0
    reserved4: int = 0


@dataclass
class MainParamsIV:
    speed_of_sound: float = 343.0
    density: float = 1.21
    reference_pressure: float = 1.0
    reserved: List[float] = field(default_factory=lambda: [0.0] * 4)


@dataclass
class SymmetryConfig:
    flags: Tuple[bool, bool, bool]
    origin: Tuple[float, float, float]


@dataclass
class BoundarySpec:
    elem_start: int
    elem_end: int
    bc_type: str
    value_re: float
    curve_re: int
    value_im: float
    curve_im: int


@dataclass
class PlaneWaveSource:
    direction: Tuple[float, float, float]
    amplitude_re: float
    curve_re: int
    amplitude_im: float
    curve_im: int


@dataclass
class PointSource:
    position: Tuple[float, float, float]
    amplitude_re: float
    curve_re: int
    amplitude_im: float
    curve_im: int


def _as_i32(v: float) -> int:
    """Rust `as i32` on f64 (truncation, saturation, NaN -> 0)."""
    if math.isnan(v):
        return 0
    return int(max(-2147483648.0, min(2147483647.0, math.trunc(v))))


class _Handle:
    def __init__(self, h):
        self.h = h

    def __del__(self):
        try:
            if self.h:
                _lib().sbh_nc_project_destroy(self.h)
                self.h = None
        except Exception:
            pass


def _string(h, which, index=0) -> str:
    buf = ctypes.create_string_buffer(4096)
    _check(_lib().sbh_nc_project_string(h, which, index, buf, len(buf)))
    return buf.value.decode()


def _values(h, which) -> np.ndarray:
    n = ctypes.c_int64()
    _check(_lib().sbh_nc_project_values(h, which, None, 0, ctypes.byref(n)))
    out = np.zeros(n.value)
    if n.value:
        _check(_lib().sbh_nc_project_values(h, which, _cptr(out), n.value, ctypes.byref(n)))
    return out


@dataclass
class NcInputConfig:
    """Parsed NC.inp (nc_format.rs:22-58)."""

    version: str
    description: str
    control_params_i: List[int]
    control_params_ii: List[float]
    frequency_curve: List[Tuple[float, float, float]]
    main_params_i: MainParamsI
    main_params_ii: MainParamsII
    main_params_iii: List[int]
    main_params_iv: MainParamsIV
    node_files: List[str]
    element_files: List[str]
    symmetry: Optional[SymmetryConfig]
    boundary_conditions: List[BoundarySpec]
    plane_waves: List[PlaneWaveSource]
    point_sources: List[PointSource]
    base_dir: str

    @classmethod
    def _from_handle(cls, h, base_dir: str) -> "NcInputConfig":
        info = _ffi.nc_info_t()
        _check(_lib().sbh_nc_project_info(h, ctypes.byref(info)))
        mi = list(info.main_params_i)
        mii = list(info.main_params_ii)
        curve = _values(h, _FREQUENCY_CURVE).reshape(-1, 3)
        specs = _values(h, _BOUNDARY_SPECS).reshape(-1, 6)
        pw = _values(h, _PLANE_WAVES).reshape(-1, 7)
        ps = _values(h, _POINT_SOURCES).reshape(-1, 7)
        return cls(
            version=_string(h, _VERSION),
            description=_string(h, _DESCRIPTION),
            control_params_i=[int(v) for v in _values(h, _CONTROL_I)],
            control_params_ii=list(_values(h, _CONTROL_II)),
            frequency_curve=[tuple(r) for r in curve],
            main_params_i=MainParamsI(*[int(v) for v in mi]),
            main_params_ii=MainParamsII(_as_i32(mii[0]), _as_i32(mii[1]), _as_i32(mii[2]), mii[3], _as_i32(mii[4]),
                                        _as_i32(mii[5]), _as_i32(mii[6])),
            main_params_iii=[int(v) for v in _values(h, _MAIN_III)],
            main_params_iv=MainParamsIV(info.speed_of_sound, info.density, info.reference_pressure,
                                        list(_values(h, _MAIN_IV_RESERVED))),
            node_files=[_string(h, _NODE_FILE, i) for i in range(info.num_node_files)],
            element_files=[_string(h, _ELEMENT_FILE, i) for i in range(info.num_element_files)],
            symmetry=(SymmetryConfig(tuple(bool(f) for f in info.symmetry_flags), tuple(info.symmetry_origin))
                      if info.has_symmetry else None),
            boundary_conditions=[BoundarySpec(int(r[0]), int(r[1]), _string(h, _BOUNDARY_TYPE, i), r[2], int(r[3]),
                                              r[4], int(r[5])) for i, r in enumerate(specs)],
            plane_waves=[PlaneWaveSource(tuple(r[:3]), r[3], int(r[4]), r[5], int(r[6])) for r in pw],
            point_sources=[PointSource(tuple(r[:3]), r[3], int(r[4]), r[5], int(r[6])) for r in ps],
            base_dir=base_dir,
        )

    def to_physics_params(self, frequency: float) -> PhysicsParams:
        """NcInputConfig::to_physics_params (nc_format.rs:797-805)."""
        return PhysicsParams(frequency, self.main_params_iv.speed_of_sound, self.main_params_iv.density)

    def bem_method(self) -> str:
        """0 Traditional, 1 SingleLevelFmm, 2/3 MultiLevelFmm (nc_format.rs:807-815)."""
        return {1: "SingleLevelFmm", 2: "MultiLevelFmm", 3: "MultiLevelFmm"}.get(self.main_params_i.solver_method,
                                                                                  "Traditional")

    def solver_method(self) -> str:
        """0 Cgs, 1 BiCgstab (nc_format.rs:817-823)."""
        return "BiCgstab" if self.main_params_ii.iterative_solver == 1 else "Cgs"


def parse_nc_input_string(content: str, base_dir: str = ".") -> NcInputConfig:
    """parse_nc_input_string (nc_format.rs:215-636)."""
    h = ctypes.c_void_p()
    try:
        _check(_lib().sbh_nc_parse(content.encode(), base_dir.encode(), ctypes.byref(h)))
    except SbhError as e:
        _raise_mapped(e)
    keep = _Handle(h)
    return NcInputConfig._from_handle(keep.h, base_dir)


def parse_nc_input(path: str) -> NcInputConfig:
    """parse_nc_input (nc_format.rs:205-213): base directory = the file's parent."""
    import os
    with open(path, "r") as f:
        content = f.read()
    return parse_nc_input_string(content, os.path.dirname(path) or ".")


@dataclass
class NumCalcProject:
    """A loaded project (project.rs:35-73): all elements in file order with
    their BCs, the evaluation-element indices and the solve frequencies."""

    config: NcInputConfig
    nodes: np.ndarray              # (V, 3)
    connectivity: np.ndarray       # (E, 4), -1 padded
    node_counts: np.ndarray        # (E,)
    is_evaluation: np.ndarray      # (E,) bool
    bc_type: np.ndarray            # (E,) 0 velocity, 1 pressure
    bc_value: np.ndarray           # (E,) complex
    admittance: np.ndarray         # (E,) complex, NaN where absent
    evaluation_indices: np.ndarray
    frequencies: np.ndarray
    speed_of_sound: float
    density: float
    _handle: _Handle = field(repr=False, default=None)

    @property
    def plane_waves(self):
        return self.config.plane_waves

    @property
    def point_sources(self):
        return self.config.point_sources

    def physics_at(self, frequency: float) -> PhysicsParams:
        return PhysicsParams(frequency, self.speed_of_sound, self.density)

    def evaluation_centers(self) -> np.ndarray:
        """Centroids of the evaluation elements (project.rs:65-71)."""
        out = np.zeros((len(self.evaluation_indices), 3))
        if len(out):
            _check(_lib().sbh_nc_project_evaluation_centers(self._handle.h, _cptr(out)))
        return out

    def surface_arrays(self):
        """Surface (object) elements as a Tri3 Mesh plus their BC arrays."""
        surf = ~self.is_evaluation
        if np.any(self.node_counts[surf] != 3):
            raise SbhError(2, "only Tri3 surface elements are supported")
        tris = self.connectivity[surf][:, :3]
        return (Mesh(self.nodes, tris), self.bc_type[surf], self.bc_value[surf], self.admittance[surf])

    def surface_mesh(self, ctx: Context) -> TbemMesh:
        """The object's panels as a TbemMesh with vertex-order normals (runner.rs:63-68)."""
        mesh, bt, bv, adm = self.surface_arrays()
        return TbemMesh(ctx, mesh, bc_type=bt, bc_value=bv, admittance=adm, normals="vertex_order")


def load_project(path: str) -> NumCalcProject:
    """load_project (project.rs:78-105)."""
    import os
    h = ctypes.c_void_p()
    try:
        _check(_lib().sbh_nc_project_load(path.encode(), ctypes.byref(h)))
    except SbhError as e:
        _raise_mapped(e)
    keep = _Handle(h)
    config = NcInputConfig._from_handle(h, os.path.dirname(path) or ".")
    info = _ffi.nc_info_t()
    _check(_lib().sbh_nc_project_info(h, ctypes.byref(info)))
    nn, ne = info.num_nodes, info.num_elements
    nodes = np.zeros((nn, 3))
    conn = np.zeros((ne, 4), dtype=np.int64)
    counts = np.zeros(ne, dtype=np.int32)
    is_eval = np.zeros(ne, dtype=np.uint8)
    bt = np.zeros(ne, dtype=np.int32)
    bv = np.zeros(ne, dtype=np.complex128)
    adm = np.zeros(ne, dtype=np.complex128)
    has = np.zeros(ne, dtype=np.uint8)
    _check(_lib().sbh_nc_project_mesh(h, _cptr(nodes), _cptr(conn), _cptr(counts), _cptr(is_eval), _cptr(bt),
                                      _cptr(bv), _cptr(adm), _cptr(has)))
    adm = np.where(has.astype(bool), adm, np.nan + 0j)
    return NumCalcProject(config=config, nodes=nodes, connectivity=conn, node_counts=counts,
                          is_evaluation=is_eval.astype(bool), bc_type=bt, bc_value=bv, admittance=adm,
                          evaluation_indices=np.flatnonzero(is_eval), frequencies=_values(h, _FREQUENCIES),
                          speed_of_sound=info.speed_of_sound, density=info.density, _handle=keep)


@dataclass
class SweepOutput:
    """SweepOutput (runner.rs:18-37): per-frequency surface and evaluation-grid pressure."""

    frequencies: np.ndarray
    speed_of_sound: float
    density: float
    num_surface_dofs: int
    surface_pressure: np.ndarray   # (F, N) complex
    eval_points: np.ndarray        # (M, 3)
    eval_pressure: np.ndarray      # (F, M) complex total pressure
    iterations: List[int]
    residuals: List[float]

    def write_json(self, path: str):
        """The reference's JSON layout: complex values as [re, im] pairs, arrays flattened row-major."""
        pairs = lambda a: [[float(v.real), float(v.imag)] for v in np.ravel(a)]  # noqa: E731
        doc = {"frequencies": [float(f) for f in self.frequencies], "speed_of_sound": self.speed_of_sound,
               "density": self.density, "num_surface_dofs": self.num_surface_dofs,
               "surface_pressure": pairs(self.surface_pressure),
               "eval_points": [float(v) for v in np.ravel(self.eval_points)],
               "eval_pressure": pairs(self.eval_pressure)}
        with open(path, "w") as f:
            json.dump(doc, f, indent=2)


def pick_incident(project: NumCalcProject) -> IncidentField:
    """pick_incident (runner.rs:160-180): the first plane wave (direction as
    written), else the first point source, else a zero-amplitude plane wave."""
    if project.plane_waves:
        pw = project.plane_waves[0]
        return IncidentField.from_sources([IncidentField.PLANE_WAVE], [pw.direction],
                                          [complex(pw.amplitude_re, pw.amplitude_im)])
    if project.point_sources:
        ps = project.point_sources[0]
        return IncidentField.from_sources([IncidentField.POINT_SOURCE], [ps.position],
                                          [complex(ps.amplitude_re, ps.amplitude_im)])
    return IncidentField.from_sources([IncidentField.PLANE_WAVE], [[0.0, 0.0, 1.0]], [0.0])


def run_project(project: NumCalcProject, ctx: Context, freq_start_idx: Optional[int] = None,
                freq_end_idx: Optional[int] = None, max_iterations: int = 200, tolerance: float = 1e-6,
                restart: int = 30, beta_scale: float = 4.0) -> SweepOutput:
    """run_project (runner.rs:51-125) on the GPU: for each frequency of
    [start, end) (clamped as the reference does) the stabilized system with
    beta = burton_miller_beta_scaled(4) and sign +1 (frequency_sweep.rs:897-916,
    686-700) plus the incident right-hand side, GMRES(30) from the previous
    frequency's solution, and the total field at the evaluation centroids."""
    total = len(project.frequencies)
    start = min(freq_start_idx or 0, total)
    end = max(min(total if freq_end_idx is None else freq_end_idx, total), start)
    freqs = np.asarray(project.frequencies[start:end], dtype=np.float64)
    incident = pick_incident(project)
    eval_pts = project.evaluation_centers()
    n = int(np.count_nonzero(~project.is_evaluation))
    surface = np.zeros((len(freqs), n), dtype=np.complex128)
    field_p = np.zeros((len(freqs), len(eval_pts)), dtype=np.complex128)
    iters, resid = [], []
    if len(freqs):
        tm = project.surface_mesh(ctx)
        ws = TbemAssemblyWorkspace(ctx, tm)
        x_prev = None
        cfg = GmresConfig(max_iterations=max_iterations, restart=restart, tolerance=tolerance)
        for fi, f in enumerate(freqs):
            ph = project.physics_at(float(f))
            beta = ph.burton_miller_beta_scaled(beta_scale)
            system = ws.assemble(ph, beta)
            b = system.rhs + incident.compute_rhs_with_beta(tm, ph, beta).reshape(n)
            sol = gmres_with_guess(DenseMatrixOperator(system), b, x_prev, cfg)
            if not sol.converged:
                raise SbhError(5, f"GMRES did not converge at {f:.3f} Hz (residual {sol.residual:.3e})")
            x_prev = sol.x
            surface[fi] = sol.x
            iters.append(int(sol.iterations))
            resid.append(float(sol.residual))
            if len(eval_pts):
                field_p[fi] = compute_total_field(eval_pts, tm, sol.x, incident, ph).p_total
    return SweepOutput(frequencies=freqs, speed_of_sound=project.speed_of_sound, density=project.density,
                       num_surface_dofs=n, surface_pressure=surface, eval_points=eval_pts, eval_pressure=field_p,
                       iterations=iters, residuals=resid)

