"""sonium_amd — dense Burton–Miller Helmholtz BEM on AMD Instinct MI355X.

Host-side mirror of the reference's interface for this path, over the C ABI
of libsbh (include/sbh.h):

* ``PhysicsParams`` — crates/sonium-bem/src/core/types.rs:190-218, 306-487
* ``TbemAssemblyWorkspace`` (``new``/``assemble``) — CudaTbemAssemblyWorkspace,
  crates/sonium-bem/src/core/assembly/tbem.rs:3779-4389
* ``build_tbem_system`` / ``build_tbem_system_with_beta`` — tbem.rs:877, 1196
* ``TbemSystem`` — tbem.rs:134-141 (device-resident matrix and rhs)
* ``IncidentField`` — crates/sonium-bem/src/core/incident.rs:165-198, 807-834
* ``DenseMatrixOperator`` — LinearOperator<Complex64>,
  crates/sonium-solvers/crates/sonium-solvers-core/src/traits.rs:2949-3290
* ``IdentityPreconditioner`` / ``JacobiPreconditioner`` — traits.rs:3645-3780
* ``GmresConfig`` / ``GmresSolution`` / ``gmres`` and variants —
  crates/sonium-solvers/crates/sonium-solvers-iterative/src/gmres.rs:26-130, 286, 564-626

All compute runs in the HIP library; a missing library raises on first use.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _ffi
from . import mesh as mesh  # noqa: F401  (re-export the mesh generators)
from .mesh import Mesh

MEM_HOST = _ffi.MEM_HOST
MEM_DEVICE = _ffi.MEM_DEVICE

STATUS_NAMES = {0: "ok", 1: "invalid_argument", 2: "unsupported", 3: "backend", 4: "out_of_memory",
                5: "numerical"}
SOLVER_STATUS = {0: "converged", 1: "max_iterations", 2: "breakdown", 3: "numerical_error"}


class SbhError(RuntimeError):
    """A failed C ABI call; ``status`` is the sbh_status code."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


INJECTION_POINTS = {"none": 0, "allocation": 1, "launch": 2, "h2d": 3, "d2h": 4, "synchronization": 5}


def live_device_bytes() -> int:
    """Bytes held by the library's device allocations in this process."""
    return int(_lib().sbh_live_device_bytes())


def _lib():
    return _ffi.load()


def _check(status: int) -> None:
    if status != 0:
        msg = _lib().sbh_last_error()
        raise SbhError(status, msg.decode() if msg else "")


def _cptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c128(a, n=None) -> np.ndarray:
    arr = np.ascontiguousarray(a, dtype=np.complex128)
    if n is not None and arr.size != n:
        raise ValueError(f"expected {n} complex values, got {arr.size}")
    return arr


class PhysicsParams:
    """Physical parameters of one frequency (types.rs:190-218)."""

    def __init__(self, frequency: float, speed_of_sound: float = 343.0, density: float = 1.21,
                 is_internal: bool = False, *, _raw: Optional[_ffi.physics_t] = None):
        if _raw is not None:
            self._p = _raw
            return
        self._p = _ffi.physics_t()
        _check(_lib().sbh_physics_new(frequency, speed_of_sound, density, int(is_internal),
                                      ctypes.byref(self._p)))

    @classmethod
    def from_wavenumber(cls, wave_number: float, speed_of_sound: float = 343.0, density: float = 1.21,
                        is_internal: bool = False) -> "PhysicsParams":
        p = _ffi.physics_t()
        _check(_lib().sbh_physics_from_wavenumber(wave_number, speed_of_sound, density, int(is_internal),
                                                  ctypes.byref(p)))
        return cls(0.0, _raw=p)

    def __getattr__(self, name):
        if name in dict(_ffi.physics_t._fields_):
            return getattr(self._p, name)
        raise AttributeError(name)

    def burton_miller_beta_scaled(self, scale: float) -> complex:
        b = _lib().sbh_burton_miller_beta_scaled(ctypes.byref(self._p), scale)
        return complex(b.re, b.im)

    def burton_miller_beta(self) -> complex:
        return self.burton_miller_beta_scaled(1.0)

    def burton_miller_beta_wavenumber_derivative(self) -> complex:
        """d(i hf / k)/dk = -i hf / k^2 for exterior problems (types.rs:322-333)."""
        if self.tau > 0.0:
            return complex(0.0, -self.harmonic_factor / (self.wave_number * self.wave_number))
        return 0j

    def pressure_factor_wavenumber_derivative(self) -> float:
        """d(rho omega hf)/dk at fixed c and rho (types.rs:340-345)."""
        return self.density * self.speed_of_sound * self.harmonic_factor

    def gamma(self) -> float:
        return 1.0


class Context:
    """One HIP device and stream."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(_lib().sbh_context_create(device, ctypes.byref(h)))
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    def synchronize(self):
        _check(_lib().sbh_context_synchronize(self._h))

    def inject_failure(self, point: str) -> None:
        """Arm a one-shot failure: "allocation", "launch", "h2d", "d2h" or "synchronization"."""
        _check(_lib().sbh_context_inject_failure(self._h, INJECTION_POINTS[point]))

    def set_device_budget(self, max_device_bytes: int) -> None:
        """Device-memory admission budget of this context (0 = none)."""
        _check(_lib().sbh_context_set_device_budget(self._h, int(max_device_bytes)))

    def device_budget(self):
        """(budget, bytes held by this context's systems, copies and batches)."""
        b, h = ctypes.c_int64(), ctypes.c_int64()
        _check(_lib().sbh_context_device_budget(self._h, ctypes.byref(b), ctypes.byref(h)))
        return int(b.value), int(h.value)

    def set_profiling(self, enabled: bool):
        _check(_lib().sbh_context_set_profiling(self._h, int(enabled)))

    def reset_stats(self):
        _check(_lib().sbh_context_reset_stats(self._h))

    def kernel_stats(self, name: str):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        _check(_lib().sbh_context_kernel_stats(self._h, _ffi.KERNEL_CLASSES[name], ctypes.byref(n),
                                               ctypes.byref(ms)))
        return int(n.value), float(ms.value)

    def close(self):
        if getattr(self, "_h", None):
            _lib().sbh_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceVector:
    """Library-owned device vector of complex128 (an n x k block when used as one)."""

    def __init__(self, ctx: Context, length: int):
        self.ctx = ctx
        h = ctypes.c_void_p()
        _check(_lib().sbh_vector_create(ctx.handle, int(length), ctypes.byref(h)))
        self._h = h
        self.length = int(length)

    @property
    def ptr(self):
        return _lib().sbh_vector_data(self._h)

    def upload(self, values, offset: int = 0):
        v = _c128(values).ravel()
        _check(_lib().sbh_vector_upload(self._h, _cptr(v), offset, v.size))

    def download(self, offset: int = 0, count: Optional[int] = None) -> np.ndarray:
        count = self.length - offset if count is None else count
        out = np.empty(count, dtype=np.complex128)
        _check(_lib().sbh_vector_download(self._h, _cptr(out), offset, count))
        return out

    def close(self):
        if getattr(self, "_h", None):
            _lib().sbh_vector_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TbemMesh:
    """Device-side mesh + boundary conditions (types.rs:582-719).

    ``bc_type`` 0 = Velocity, 1 = Pressure; ``bc_value`` the panel BC value
    (zero = rigid); ``admittance`` complex per panel, NaN where absent.
    """

    def __init__(self, ctx: Context, mesh: Mesh, bc_type=None, bc_value=None, admittance=None,
                 normals: str = "outward"):
        """normals: "outward" orients each vertex-order normal away from the
        origin (compute_element_geometry, generators.rs:857-866); "vertex_order"
        keeps (v1-v0)x(v2-v0), as meshes read from files are
        (compute_element_normal, element.rs:368-392)."""
        self.ctx = ctx
        self.mesh = mesh
        n = mesh.num_elements
        self.num_elements = n
        bt = None if bc_type is None else np.ascontiguousarray(bc_type, dtype=np.int32).reshape(n)
        bv = None if bc_value is None else _c128(bc_value, n)
        ad = has = None
        if admittance is not None:
            a = _c128(admittance, n)
            has = np.ascontiguousarray(~np.isnan(a.real), dtype=np.uint8)
            ad = np.nan_to_num(a)
        self._keep = (bt, bv, ad, has)
        rule = {"outward": 0, "vertex_order": 1}[normals]
        h = ctypes.c_void_p()
        _check(_lib().sbh_mesh_create_oriented(
            ctx.handle, _cptr(mesh.nodes), mesh.nodes.shape[0], _cptr(mesh.triangles), n,
            None if bt is None else _cptr(bt), None if bv is None else _cptr(bv),
            None if ad is None else _cptr(ad), None if has is None else _cptr(has), rule, ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def geometry(self):
        n = self.num_elements
        c = np.empty((n, 3))
        nn = np.empty((n, 3))
        a = np.empty(n)
        _check(_lib().sbh_mesh_geometry(self._h, _cptr(c), _cptr(nn), _cptr(a)))
        return c, nn, a

    def close(self):
        if getattr(self, "_h", None):
            _lib().sbh_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TbemSystem:
    """Device-resident dense system (tbem.rs:134-141): row-major N x N matrix and rhs."""

    def __init__(self, ctx: Context, num_dofs: int):
        self.ctx = ctx
        self.num_dofs = int(num_dofs)
        h = ctypes.c_void_p()
        _check(_lib().sbh_system_create(ctx.handle, self.num_dofs, ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def device_pointers(self):
        a = ctypes.c_void_p()
        lda = ctypes.c_int64()
        r = ctypes.c_void_p()
        _check(_lib().sbh_system_device_pointers(self._h, ctypes.byref(a), ctypes.byref(lda), ctypes.byref(r)))
        return a.value, int(lda.value), r.value

    def rhs_pointer(self) -> int:
        """Device address of the right-hand side."""
        r = ctypes.c_void_p()
        _check(_lib().sbh_system_device_pointers(self._h, None, None, ctypes.byref(r)))
        return r.value

    def matrix_rows(self, row_begin: int, row_end: int) -> np.ndarray:
        out = np.empty((row_end - row_begin, self.num_dofs), dtype=np.complex128)
        _check(_lib().sbh_system_download(self._h, row_begin, row_end, _cptr(out), None))
        return out

    @property
    def matrix(self) -> np.ndarray:
        return self.matrix_rows(0, self.num_dofs)

    @property
    def rhs(self) -> np.ndarray:
        out = np.empty(self.num_dofs, dtype=np.complex128)
        _check(_lib().sbh_system_download(self._h, 0, 0, None, _cptr(out)))
        return out

    def set_rhs(self, values):
        v = _c128(values, self.num_dofs)
        _check(_lib().sbh_system_set_rhs(self._h, _cptr(v), MEM_HOST))

    def close(self):
        if getattr(self, "_h", None):
            _lib().sbh_system_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TbemAssemblyWorkspace:
    """Resident assembly workspace: geometry + near-pair routes built once, assemble per frequency."""

    def __init__(self, ctx: Context, mesh: TbemMesh, route_threads: int = 0):
        self.ctx = ctx
        self.mesh = mesh
        h = ctypes.c_void_p()
        _check(_lib().sbh_workspace_create(ctx.handle, mesh.handle, route_threads, ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    @property
    def stats(self):
        a = ctypes.c_int64()
        b = ctypes.c_int64()
        ms = ctypes.c_double()
        _check(_lib().sbh_workspace_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(ms)))
        return {"near_pairs": int(a.value), "leaves": int(b.value), "build_ms": float(ms.value)}

    def assemble(self, physics: PhysicsParams, beta: Optional[complex] = None,
                 system: Optional[TbemSystem] = None) -> TbemSystem:
        if beta is None:
            beta = physics.burton_miller_beta()
        if system is None:
            system = TbemSystem(self.ctx, self.mesh.num_elements)
        b = _ffi.c64(complex(beta).real, complex(beta).imag)
        _check(_lib().sbh_workspace_assemble(self._h, ctypes.byref(physics._p), b, system.handle))
        return system

    def assemble_parameter_derivative(self, physics: PhysicsParams, beta: complex, beta_derivative: complex, *,
                                      sign: float = 1.0, wavenumber_scale: float = 1.0,
                                      pressure_factor_derivative: Optional[float] = None,
                                      admittance_scale: float = 0.0,
                                      system: Optional[TbemSystem] = None) -> TbemSystem:
        """d/dp of the stabilized system (tbem.rs:1858-2019) into a device
        TbemSystem: dk/dp = wavenumber_scale, d(rho omega hf)/dp =
        pressure_factor_derivative (default: the wavenumber derivative's
        rho c hf), dY/dp = admittance_scale * Y; beta_derivative is the total
        derivative of beta."""
        if pressure_factor_derivative is None:
            pressure_factor_derivative = physics.pressure_factor_wavenumber_derivative()
        if system is None:
            system = TbemSystem(self.ctx, self.mesh.num_elements)
        b = _ffi.c64(complex(beta).real, complex(beta).imag)
        bd = _ffi.c64(complex(beta_derivative).real, complex(beta_derivative).imag)
        _check(_lib().sbh_workspace_assemble_derivative(
            self._h, ctypes.byref(physics._p), b, bd, float(sign), float(wavenumber_scale),
            float(pressure_factor_derivative), float(admittance_scale), system.handle))
        return system

    def _derivative(self, physics, sign, beta_scale, ws, pfd, adm, system):
        beta = physics.burton_miller_beta_scaled(beta_scale)
        beta_d = physics.burton_miller_beta_wavenumber_derivative() * beta_scale * ws
        return self.assemble_parameter_derivative(physics, beta, beta_d, sign=sign, wavenumber_scale=ws,
                                                  pressure_factor_derivative=pfd, admittance_scale=adm,
                                                  system=system)

    def assemble_wavenumber_derivative(self, physics: PhysicsParams, sign: float = 1.0, beta_scale: float = 1.0,
                                       system: Optional[TbemSystem] = None) -> TbemSystem:
        """build_tbem_system_stabilized_wavenumber_derivative (tbem.rs:1734-1752)."""
        return self._derivative(physics, sign, beta_scale, 1.0, physics.pressure_factor_wavenumber_derivative(), 0.0,
                                system)

    def assemble_speed_derivative(self, physics: PhysicsParams, sign: float = 1.0, beta_scale: float = 1.0,
                                  system: Optional[TbemSystem] = None) -> TbemSystem:
        """build_tbem_system_stabilized_speed_derivative (tbem.rs:1754-1782): dk/dc = -k/c at fixed omega."""
        return self._derivative(physics, sign, beta_scale, -physics.wave_number / physics.speed_of_sound, 0.0, 0.0,
                                system)

    def assemble_density_derivative(self, physics: PhysicsParams, sign: float = 1.0, beta_scale: float = 1.0,
                                    system: Optional[TbemSystem] = None) -> TbemSystem:
        """build_tbem_system_stabilized_density_derivative (tbem.rs:1784-1805)."""
        return self._derivative(physics, sign, beta_scale, 0.0, physics.pressure_factor / physics.density, 0.0,
                                system)

    def assemble_impedance_density_derivative(self, physics: PhysicsParams, sign: float = 1.0,
                                              beta_scale: float = 1.0,
                                              system: Optional[TbemSystem] = None) -> TbemSystem:
        """build_tbem_system_stabilized_impedance_density_derivative (tbem.rs:1807-1830): Y = 1/(rho c)."""
        return self._derivative(physics, sign, beta_scale, 0.0, physics.pressure_factor / physics.density,
                                -1.0 / physics.density, system)

    def assemble_impedance_speed_derivative(self, physics: PhysicsParams, sign: float = 1.0,
                                            beta_scale: float = 1.0,
                                            system: Optional[TbemSystem] = None) -> TbemSystem:
        """build_tbem_system_stabilized_impedance_speed_derivative (tbem.rs:1832-1856)."""
        return self._derivative(physics, sign, beta_scale, -physics.wave_number / physics.speed_of_sound, 0.0,
                                -1.0 / physics.speed_of_sound, system)

    def close(self):
        if getattr(self, "_h", None):
            _lib().sbh_workspace_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_tbem_system_with_beta(ctx: Context, mesh: TbemMesh, physics: PhysicsParams, beta: complex) -> TbemSystem:
    """tbem.rs:1196-1203 (→ build_tbem_system_with_beta_parallel, :1284-1466)."""
    ws = TbemAssemblyWorkspace(ctx, mesh)
    try:
        sys = ws.assemble(physics, beta)
        ctx.synchronize()
        return sys
    finally:
        ws.close()


def build_tbem_system(ctx: Context, mesh: TbemMesh, physics: PhysicsParams) -> TbemSystem:
    """tbem.rs:877-883: beta = physics.burton_miller_beta()."""
    return build_tbem_system_with_beta(ctx, mesh, physics, physics.burton_miller_beta())


class IncidentField:
    """Incident field (incident.rs:17-39, 165-420): plane waves
    IncidentField::plane_wave(direction, amplitude) (directions normalised, as
    that constructor does) and point sources IncidentField::PointSource
    {position, strength}.  Each source is one right-hand-side column; the
    field evaluation sums them (MultiplePlaneWaves / MultiplePointSources)."""

    PLANE_WAVE, POINT_SOURCE = 0, 1

    def __init__(self, directions, amplitudes=None):
        d = np.ascontiguousarray(directions, dtype=np.float64).reshape(-1, 3)
        self.directions = d
        amps = np.ones(d.shape[0], dtype=np.complex128) if amplitudes is None else _c128(amplitudes, d.shape[0])
        self.amplitudes = amps
        self.kinds = np.zeros(d.shape[0], dtype=np.int32)
        self.normalise = True

    @classmethod
    def plane_wave_z(cls) -> "IncidentField":
        return cls([[0.0, 0.0, 1.0]])

    @classmethod
    def plane_wave(cls, direction, amplitude: float = 1.0) -> "IncidentField":
        return cls([direction], [amplitude])

    @classmethod
    def point_sources(cls, positions, strengths=None) -> "IncidentField":
        """IncidentField::PointSource {position, strength}: p = S e^{ikr}/(4 pi r)."""
        f = cls(positions, strengths)
        f.kinds[:] = cls.POINT_SOURCE
        return f

    @classmethod
    def from_sources(cls, kinds, vectors, amplitudes) -> "IncidentField":
        """Sources exactly as given (the enum variants built directly: plane-wave
        directions are not normalised)."""
        f = cls(vectors, amplitudes)
        f.kinds = np.ascontiguousarray(kinds, dtype=np.int32).reshape(-1)
        f.normalise = False
        return f

    @property
    def num_rhs(self) -> int:
        return int(self.directions.shape[0])

    @property
    def _plane_only(self) -> bool:
        return self.normalise and not np.any(self.kinds != self.PLANE_WAVE)

    def source_array(self):
        """The sources as sbh_incident_source records (plane-wave directions
        normalised like IncidentField::plane_wave unless built from_sources)."""
        arr = (_ffi.incident_source_t * self.num_rhs)()
        for i in range(self.num_rhs):
            v = self.directions[i].copy()
            if self.kinds[i] == self.PLANE_WAVE and self.normalise:
                ln = float(np.sqrt(v @ v))
                v = v / ln if ln > 1e-10 else np.array([0.0, 0.0, -1.0])
            arr[i].kind = int(self.kinds[i])
            arr[i].vec[:] = [float(x) for x in v]
            arr[i].re = float(self.amplitudes[i].real)
            arr[i].im = float(self.amplitudes[i].imag)
        return arr

    def compute_rhs_with_beta(self, mesh: TbemMesh, physics: PhysicsParams, beta: complex,
                              sign: float = 1.0) -> np.ndarray:
        """sign * -(gamma p_inc + beta tau dp_inc/dn) at the panel centres, N x num_rhs."""
        out = np.zeros((mesh.num_elements, self.num_rhs), dtype=np.complex128)
        if self._plane_only:
            self._launch(mesh, physics, beta, sign, _cptr(out), MEM_HOST, accumulate=False)
        else:
            src = self.source_array()
            col = np.zeros(mesh.num_elements, dtype=np.complex128)
            b = _ffi.c64(complex(beta).real, complex(beta).imag)
            for i in range(self.num_rhs):
                _check(_lib().sbh_incident_rhs(mesh.ctx.handle, mesh.handle, ctypes.byref(physics._p), b, float(sign),
                                               ctypes.byref(src[i]), 1, _cptr(col), MEM_HOST, 0))
                out[:, i] = col
        return out if self.num_rhs > 1 else out[:, 0]

    def add_to_system(self, system: TbemSystem, mesh: TbemMesh, physics: PhysicsParams, beta: complex,
                      sign: float = 1.0):
        """add_incident_field_rhs_stabilized (bem_solver.rs:5005-5034) on the device rhs."""
        if self.num_rhs != 1:
            raise ValueError("a TbemSystem holds one right-hand side")
        rhs_ptr = system.rhs_pointer()
        self.write_device_ptr(ctypes.c_void_p(rhs_ptr), mesh, physics, beta, sign, accumulate=True)

    def write_device(self, vec: DeviceVector, mesh: TbemMesh, physics: PhysicsParams, beta: complex,
                     sign: float = 1.0, accumulate: bool = False):
        self.write_device_ptr(ctypes.c_void_p(vec.ptr), mesh, physics, beta, sign, accumulate)

    def write_device_ptr(self, ptr, mesh, physics, beta, sign=1.0, accumulate=False):
        if self._plane_only:
            self._launch(mesh, physics, beta, sign, ptr, MEM_DEVICE, accumulate)
            return
        if self.num_rhs != 1:
            raise ValueError("device right-hand sides with point sources take one source per column")
        b = _ffi.c64(complex(beta).real, complex(beta).imag)
        src = self.source_array()
        _check(_lib().sbh_incident_rhs(mesh.ctx.handle, mesh.handle, ctypes.byref(physics._p), b, float(sign), src, 1,
                                       ptr, MEM_DEVICE, int(accumulate)))

    def _launch(self, mesh, physics, beta, sign, ptr, memory, accumulate):
        b = _ffi.c64(complex(beta).real, complex(beta).imag)
        _check(_lib().sbh_incident_plane_wave_rhs(
            mesh.ctx.handle, mesh.handle, ctypes.byref(physics._p), b, float(sign), _cptr(self.directions),
            _cptr(self.amplitudes), self.num_rhs, ptr, memory, int(accumulate)))


def _field(mesh, solution, surface_velocity, physics, incident, points, with_incident):
    n = mesh.num_elements
    sol = _c128(solution, n)
    vel = None if surface_velocity is None else _c128(surface_velocity, n)
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    m = pts.shape[0]
    scat = np.zeros(m, dtype=np.complex128)
    inc = np.zeros(m, dtype=np.complex128) if with_incident else None
    tot = np.zeros(m, dtype=np.complex128) if with_incident else None
    nw = 0 if incident is None else incident.num_rhs
    outs = (_cptr(pts), m, _cptr(scat), None if inc is None else _cptr(inc), None if tot is None else _cptr(tot),
            MEM_HOST)
    if incident is None or incident._plane_only:
        _check(_lib().sbh_field_pressure(
            mesh.ctx.handle, mesh.handle, ctypes.byref(physics._p), _cptr(sol), None if vel is None else _cptr(vel),
            None if incident is None else _cptr(incident.directions),
            None if incident is None else _cptr(incident.amplitudes), nw, *outs))
    else:
        _check(_lib().sbh_field_pressure_sources(
            mesh.ctx.handle, mesh.handle, ctypes.byref(physics._p), _cptr(sol), None if vel is None else _cptr(vel),
            incident.source_array(), nw, *outs))
    return pts, scat, inc, tot


def compute_scattered_field(points, mesh: TbemMesh, solution, physics: PhysicsParams,
                            surface_velocity=None) -> np.ndarray:
    """Scattered pressure at points off the surface (compute_scattered_field(_checked),
    pressure.rs:250-280, 868-894; integration pressure.rs:1388-1480, 1566-1620,
    1987-2044): constant Tri3 panels, 7-point rule.  surface_velocity None takes
    the panels' Cauchy data from their BCs (constant_cauchy_data,
    pressure.rs:1481-1520) with no incident field."""
    return _field(mesh, solution, surface_velocity, physics, None, points, False)[1]


class FieldPoint:
    """Position and incident/scattered/total pressure (pressure.rs:32-41)."""

    __slots__ = ("position", "p_incident", "p_scattered", "p_total")

    def __init__(self, position, p_incident, p_scattered, p_total):
        self.position = position
        self.p_incident = p_incident
        self.p_scattered = p_scattered
        self.p_total = p_total

    def __repr__(self):
        return f"FieldPoint({self.position}, total={self.p_total})"


class FieldValues:
    """compute_total_field's result as arrays: positions (M x 3), p_incident,
    p_scattered and p_total (M complex each); indexing gives FieldPoint."""

    def __init__(self, positions, p_incident, p_scattered, p_total):
        self.positions = positions
        self.p_incident = p_incident
        self.p_scattered = p_scattered
        self.p_total = p_total

    def __len__(self):
        return self.positions.shape[0]

    def __getitem__(self, i):
        return FieldPoint(self.positions[i].tolist(), complex(self.p_incident[i]), complex(self.p_scattered[i]),
                          complex(self.p_total[i]))


def compute_total_field(points, mesh: TbemMesh, solution, incident: "IncidentField", physics: PhysicsParams,
                        surface_velocity=None) -> FieldValues:
    """Incident plus scattered pressure at points off the surface
    (compute_total_field(_checked), pressure.rs:501-557, 2105-2145): the plane
    waves are the incident field, also for the Cauchy data of pressure panels."""
    pts, scat, inc, tot = _field(mesh, solution, surface_velocity, physics, incident, points, True)
    return FieldValues(pts, inc, scat, tot)


AdjointCapability = {0: "Unknown", 1: "Unsupported", 2: "Exact"}


class DenseMatrixOperator:
    """LinearOperator<Complex64> over a device-resident TbemSystem (traits.rs:2949-3290)."""

    def __init__(self, system: TbemSystem):
        self.system = system
        self.ctx = system.ctx
        h = ctypes.c_void_p()
        _check(_lib().sbh_dense_operator_create(system.handle, ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def num_rows(self) -> int:
        return int(_lib().sbh_operator_num_rows(self._h))

    def num_cols(self) -> int:
        return int(_lib().sbh_operator_num_cols(self._h))

    def transpose_capability(self) -> str:
        return AdjointCapability[_lib().sbh_operator_transpose_capability(self._h)]

    def hermitian_capability(self) -> str:
        return AdjointCapability[_lib().sbh_operator_hermitian_capability(self._h)]

    def _apply(self, fn, x):
        xv = _c128(x, self.num_cols())
        y = np.empty(self.num_rows(), dtype=np.complex128)
        _check(fn(self._h, _cptr(xv), _cptr(y), MEM_HOST))
        return y

    def apply(self, x) -> np.ndarray:
        return self._apply(_lib().sbh_operator_apply, x)

    def apply_into(self, x, y: np.ndarray) -> None:
        y[...] = self.apply(x)

    def apply_transpose(self, x) -> np.ndarray:
        return self._apply(_lib().sbh_operator_apply_transpose, x)

    def apply_hermitian(self, x) -> np.ndarray:
        return self._apply(_lib().sbh_operator_apply_hermitian, x)

    def apply_block(self, X) -> np.ndarray:
        Xa = np.ascontiguousarray(X, dtype=np.complex128)
        if Xa.ndim != 2 or Xa.shape[0] != self.num_cols():
            raise ValueError("X must be n x k")
        Y = np.empty((self.num_rows(), Xa.shape[1]), dtype=np.complex128)
        _check(_lib().sbh_operator_apply_block(self._h, _cptr(Xa), _cptr(Y), Xa.shape[1], MEM_HOST))
        return Y

    def apply_device(self, x: DeviceVector, y: DeviceVector) -> None:
        _check(_lib().sbh_operator_apply(self._h, ctypes.c_void_p(x.ptr), ctypes.c_void_p(y.ptr), MEM_DEVICE))

    def apply_block_device(self, X: DeviceVector, Y: DeviceVector, k: int) -> None:
        """Y = A X for device-resident n x k row-major blocks."""
        if X.length < self.num_cols() * k or Y.length < self.num_rows() * k:
            raise ValueError("device blocks are shorter than n x k")
        _check(_lib().sbh_operator_apply_block(self._h, ctypes.c_void_p(X.ptr), ctypes.c_void_p(Y.ptr), int(k),
                                               MEM_DEVICE))

    def apply_block_noalloc(self, X: DeviceVector, Y: DeviceVector, k: int) -> None:
        """Y = A X like apply_block_device, with no device allocation (the
        split-K partials use the library's fixed per-GPU scratch)."""
        if X.length < self.num_cols() * k or Y.length < self.num_rows() * k:
            raise ValueError("device blocks are shorter than n x k")
        _check(_lib().sbh_operator_apply_block_noalloc(self._h, ctypes.c_void_p(X.ptr), ctypes.c_void_p(Y.ptr),
                                                       int(k)))

    def close(self):
        if getattr(self, "_h", None):
            _lib().sbh_operator_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TbemExactOperator(DenseMatrixOperator):
    """Matrix-free exact operator: y = sign * A(beta) x with the entries of the
    dense assembly, evaluated on the fly (apply_tbem_stabilized_exact,
    hmatrix.rs:2363-2425).  O(N) device memory; each product costs about one
    assembly.  Transpose/Hermitian products are not available."""

    def __init__(self, workspace: "TbemAssemblyWorkspace", physics: PhysicsParams, beta: Optional[complex] = None,
                 sign: float = 1.0):
        self.workspace = workspace
        self.ctx = workspace.ctx
        bv = physics.burton_miller_beta() if beta is None else complex(beta)
        h = ctypes.c_void_p()
        _check(_lib().sbh_exact_operator_create(workspace.handle, ctypes.byref(physics._p),
                                                _ffi.c64(bv.real, bv.imag), float(sign), ctypes.byref(h)))
        self._h = h


_APPLY_FN = _ffi.APPLY_FN


def _host_callback(fn):
    """A C apply callback (host memory) calling fn(x) -> y on complex128 arrays;
    any exception becomes a nonzero status."""

    def tramp(user, xp, yp, n, memory):
        try:
            x = np.ctypeslib.as_array((ctypes.c_double * (2 * n)).from_address(xp)).view(np.complex128)
            y = np.ctypeslib.as_array((ctypes.c_double * (2 * n)).from_address(yp)).view(np.complex128)
            out = np.asarray(fn(x.copy()), dtype=np.complex128).reshape(n)
            if not np.all(np.isfinite(out)):
                return 2
            y[:] = out
            return 0
        except Exception:
            return 1

    return _APPLY_FN(tramp)


class CallbackOperator(DenseMatrixOperator):
    """A caller-supplied LinearOperator<Complex64> (traits.rs:2949-3290):
    ``fn(x) -> A x`` on host arrays of length n, usable with gmres, block GMRES
    and apply/apply_block; adjoint products are Unsupported."""

    def __init__(self, ctx: "Context", n: int, fn):
        self.system = None
        self.ctx = ctx
        self._cb = _host_callback(fn)
        h = ctypes.c_void_p()
        _check(_lib().sbh_operator_create_callback(ctx.handle, int(n), self._cb, None, MEM_HOST, ctypes.byref(h)))
        self._h = h


class _Precond:
    def apply(self, r) -> np.ndarray:
        rv = _c128(r, self.n)
        y = np.empty(self.n, dtype=np.complex128)
        _check(_lib().sbh_precond_apply(self._h, _cptr(rv), _cptr(y), MEM_HOST))
        return y

    def close(self):
        if getattr(self, "_h", None):
            _lib().sbh_precond_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CallbackPreconditioner(_Precond):
    """A caller-supplied Preconditioner<Complex64> (traits.rs:3645-3780): ``fn(r) -> M^-1 r``."""

    def __init__(self, ctx: "Context", n: int, fn):
        self.n = int(n)
        self._cb = _host_callback(fn)
        h = ctypes.c_void_p()
        _check(_lib().sbh_precond_create_callback(ctx.handle, self.n, self._cb, None, MEM_HOST, ctypes.byref(h)))
        self._h = h


class IdentityPreconditioner(_Precond):
    def __init__(self, ctx: Context, n: int):
        self.n = int(n)
        h = ctypes.c_void_p()
        _check(_lib().sbh_precond_identity(ctx.handle, self.n, ctypes.byref(h)))
        self._h = h


class JacobiPreconditioner(_Precond):
    """Inverse-diagonal left preconditioner of a dense operator."""

    def __init__(self, op: DenseMatrixOperator):
        self.n = op.num_rows()
        h = ctypes.c_void_p()
        _check(_lib().sbh_precond_jacobi(op.handle, ctypes.byref(h)))
        self._h = h


@dataclass
class GmresConfig:
    """gmres.rs:26-71; defaults of GmresConfig<f64>::default().

    ``min_restart``/``max_restart`` of None keep the restart fixed; with both
    set, the restart adapts between them (gmres.rs:1013-1029)."""

    max_iterations: int = 100
    restart: int = 30
    tolerance: float = 1e-6
    convergence_check_interval: int = 1
    exact_residual_check_interval: int = 0
    min_restart: Optional[int] = None
    max_restart: Optional[int] = None

    @classmethod
    def for_small_problems(cls) -> "GmresConfig":
        return cls(max_iterations=50, restart=50, tolerance=1e-8)

    @classmethod
    def with_restart(cls, restart: int) -> "GmresConfig":
        return cls(restart=restart)

    def _c(self):
        return _ffi.gmres_config_t(int(self.max_iterations), int(self.restart), float(self.tolerance),
                                   int(self.convergence_check_interval), int(self.exact_residual_check_interval),
                                   int(self.min_restart or 0), int(self.max_restart or 0))


@dataclass
class GmresSolution:
    """gmres.rs:108-130."""

    x: Optional[np.ndarray]
    iterations: int
    restarts: int
    residual: float
    estimated_residual: float
    backward_error: float
    converged: bool
    status: str
    spmv_count: int = 0
    reorthogonalizations: int = 0
    exact_residual_checks: int = 0
    extra: dict = field(default_factory=dict)


def _solution(raw: _ffi.gmres_solution_t, x) -> GmresSolution:
    return GmresSolution(x=x, iterations=int(raw.iterations), restarts=int(raw.restarts),
                         residual=float(raw.residual), estimated_residual=float(raw.estimated_residual),
                         backward_error=float(raw.backward_error), converged=bool(raw.converged),
                         status=SOLVER_STATUS[int(raw.status)], spmv_count=int(raw.spmv_count),
                         reorthogonalizations=int(raw.reorthogonalizations),
                         exact_residual_checks=int(raw.exact_residual_checks))


def gmres_preconditioned_with_guess(op: DenseMatrixOperator, precond, b, x0, config: GmresConfig) -> GmresSolution:
    """gmres.rs:608-626 (host vectors)."""
    n = op.num_rows()
    bv = _c128(b, n)
    x0v = None if x0 is None else _c128(x0, n)
    x = np.zeros(n, dtype=np.complex128)
    raw = _ffi.gmres_solution_t()
    cfg = config._c()
    _check(_lib().sbh_gmres(op.handle, None if precond is None else precond._h, _cptr(bv),
                            None if x0v is None else _cptr(x0v), ctypes.byref(cfg), _cptr(x), MEM_HOST,
                            ctypes.byref(raw)))
    return _solution(raw, x)


def gmres(op: DenseMatrixOperator, b, config: GmresConfig) -> GmresSolution:
    """gmres.rs:286-297."""
    return gmres_preconditioned_with_guess(op, None, b, None, config)


def gmres_with_guess(op: DenseMatrixOperator, b, x0, config: GmresConfig) -> GmresSolution:
    """gmres.rs:564-580."""
    return gmres_preconditioned_with_guess(op, None, b, x0, config)


def gmres_preconditioned(op: DenseMatrixOperator, precond, b, config: GmresConfig) -> GmresSolution:
    """gmres.rs:585-600."""
    return gmres_preconditioned_with_guess(op, precond, b, None, config)


def gmres_device(op: DenseMatrixOperator, b_ptr: int, x: DeviceVector, config: GmresConfig,
                 precond=None, x0: Optional[DeviceVector] = None) -> GmresSolution:
    """GMRES on device-resident b (device pointer) and x; the solution stays in HBM."""
    raw = _ffi.gmres_solution_t()
    cfg = config._c()
    _check(_lib().sbh_gmres(op.handle, None if precond is None else precond._h, ctypes.c_void_p(b_ptr),
                            None if x0 is None else ctypes.c_void_p(x0.ptr), ctypes.byref(cfg),
                            ctypes.c_void_p(x.ptr), MEM_DEVICE, ctypes.byref(raw)))
    return _solution(raw, None)


@dataclass
class GcroDrConfig:
    """GcroDrConfig (gcro_dr.rs:112-157); from_gmres keeps max_iterations,
    restart and tolerance."""

    max_iterations: int = 100
    restart: int = 50
    tolerance: float = 1e-6

    @staticmethod
    def from_gmres(g: "GmresConfig") -> "GcroDrConfig":
        return GcroDrConfig(g.max_iterations, g.restart, g.tolerance)

    def _c(self):
        return _ffi.gcro_dr_config_t(int(self.max_iterations), int(self.restart), float(self.tolerance))


class GcroDrSubspace:
    """GcroDrSubspace (gcro_dr.rs:63-110): up to k recycled vectors of order
    n, kept on the device between solves."""

    def __init__(self, ctx: Context, n: int, k: int):
        self.ctx = ctx
        self.n = int(n)
        self.k = int(k)
        h = ctypes.c_void_p()
        _check(_lib().sbh_gcro_subspace_create(ctx.handle, self.n, self.k, ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def num_vectors(self) -> int:
        return int(_lib().sbh_gcro_subspace_size(self._h))

    def is_populated(self) -> bool:
        return self.num_vectors() > 0

    def target_size(self) -> int:
        return self.k

    def reset(self) -> None:
        _check(_lib().sbh_gcro_subspace_reset(self._h))

    def vectors(self) -> np.ndarray:
        """The recycled vectors, num_vectors() x n."""
        out = np.empty((self.num_vectors(), self.n), dtype=np.complex128)
        if out.size:
            _check(_lib().sbh_gcro_subspace_download(self._h, _cptr(out)))
        return out

    def close(self):
        if getattr(self, "_h", None):
            _lib().sbh_gcro_subspace_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gcro_dr_preconditioned(op: DenseMatrixOperator, precond, b, x0, subspace: GcroDrSubspace,
                           config: GcroDrConfig) -> GmresSolution:
    """try_gcro_dr_preconditioned (gcro_dr.rs:199-290), host vectors; x0 may be None."""
    n = op.num_rows()
    bv = _c128(b, n)
    x0v = None if x0 is None else _c128(x0, n)
    x = np.zeros(n, dtype=np.complex128)
    raw = _ffi.gmres_solution_t()
    cfg = config._c()
    _check(_lib().sbh_gcro_dr(op.handle, None if precond is None else precond._h, _cptr(bv),
                              None if x0v is None else _cptr(x0v), subspace.handle, ctypes.byref(cfg), _cptr(x),
                              MEM_HOST, ctypes.byref(raw)))
    return _solution(raw, x)


def gcro_dr(op: DenseMatrixOperator, b, x0, subspace: GcroDrSubspace, config: GcroDrConfig) -> GmresSolution:
    """try_gcro_dr (gcro_dr.rs:183-197): the identity preconditioner."""
    return gcro_dr_preconditioned(op, None, b, x0, subspace, config)


def gcro_dr_device(op: DenseMatrixOperator, b_ptr: int, x: DeviceVector, subspace: GcroDrSubspace,
                   config: GcroDrConfig, precond=None, x0: Optional[DeviceVector] = None) -> GmresSolution:
    """GCRO-DR on device-resident b (device pointer) and x."""
    raw = _ffi.gmres_solution_t()
    cfg = config._c()
    _check(_lib().sbh_gcro_dr(op.handle, None if precond is None else precond._h, ctypes.c_void_p(b_ptr),
                              None if x0 is None else ctypes.c_void_p(x0.ptr), subspace.handle, ctypes.byref(cfg),
                              ctypes.c_void_p(x.ptr), MEM_DEVICE, ctypes.byref(raw)))
    return _solution(raw, None)


def gmres_block(op: DenseMatrixOperator, B, config: GmresConfig, X0=None):
    """Many right-hand sides (BemSolver::solve_many, bem_solver.rs:2032): B is n x k (k <= 64).

    Returns (X, [GmresSolution per column]); every column follows the
    single-RHS gmres_core, the columns share one block product per step.
    """
    Bv = np.ascontiguousarray(B, dtype=np.complex128)
    n = op.num_rows()
    if Bv.ndim == 1:
        Bv = Bv.reshape(n, 1)
    k = Bv.shape[1]
    X = np.zeros_like(Bv)
    X0v = None if X0 is None else np.ascontiguousarray(X0, dtype=np.complex128).reshape(n, k)
    raws = (_ffi.gmres_solution_t * k)()
    cfg = config._c()
    _check(_lib().sbh_gmres_block(op.handle, _cptr(Bv), None if X0v is None else _cptr(X0v), k,
                                  ctypes.byref(cfg), _cptr(X), MEM_HOST, raws))
    return X, [_solution(raws[c], X[:, c]) for c in range(k)]


def _estimate_dict(e) -> dict:
    return {name: int(getattr(e, name)) for name, _ in _ffi.resource_estimate_t._fields_}


def admission_estimate(active_dofs: int, rhs_count: int = 1, frequency_count: int = 1, restart: int = 30) -> dict:
    """BemResourceRequest::estimate for the dense method (admission.rs:112-200),
    plus this library's device bytes for the same work."""
    e = _ffi.resource_estimate_t()
    _check(_lib().sbh_admission_estimate(int(active_dofs), int(rhs_count), int(frequency_count), int(restart),
                                         ctypes.byref(e)))
    return _estimate_dict(e)


def admission_admit(max_host_bytes: int, max_device_bytes: int, max_active_work: int, active_dofs: int,
                    rhs_count: int = 1, frequency_count: int = 1, restart: int = 30) -> dict:
    """BemResourceBudget::admit (admission.rs:284-312)."""
    b = _ffi.resource_budget_t(int(max_host_bytes), int(max_device_bytes), int(max_active_work))
    e = _ffi.resource_estimate_t()
    _check(_lib().sbh_admission_admit(ctypes.byref(b), int(active_dofs), int(rhs_count), int(frequency_count),
                                      int(restart), ctypes.byref(e)))
    return _estimate_dict(e)


@dataclass
class BlockGmresConfig:
    """BlockGmresConfig (block_gmres.rs:58-89); defaults of BlockGmresConfig<f64>::default()."""

    max_iterations: int = 100
    restart: int = 30
    tolerance: float = 1e-6
    true_residual_check: bool = True
    exact_residual_check_interval: int = 0

    def _c(self):
        return _ffi.block_gmres_config_t(int(self.max_iterations), int(self.restart), float(self.tolerance),
                                         1 if self.true_residual_check else 0, 0,
                                         int(self.exact_residual_check_interval))


@dataclass
class BlockGmresSolution:
    """BlockGmresSolution (block_gmres.rs:117-148): x is n x k (None for
    device-resident solves), iterations counts block Arnoldi steps."""

    x: Optional[np.ndarray]
    iterations: int
    residuals: np.ndarray
    true_residuals: np.ndarray
    backward_errors: np.ndarray
    converged: bool
    status: str
    spmv_count: int
    exact_residual_checks: int
    restart_cycles: int
    max_basis: int
    deflated: int
    cholesky_qr_blocks: int = 0
    chain_qr_blocks: int = 0


def _block_solution(info, x, res, tres, berr) -> BlockGmresSolution:
    return BlockGmresSolution(x=x, iterations=int(info.iterations), residuals=res, true_residuals=tres,
                              backward_errors=berr, converged=bool(info.converged),
                              status=SOLVER_STATUS[int(info.status)], spmv_count=int(info.spmv_count),
                              exact_residual_checks=int(info.exact_residual_checks),
                              restart_cycles=int(info.restart_cycles), max_basis=int(info.max_basis),
                              deflated=int(info.deflated), cholesky_qr_blocks=int(info.cholesky_qr_blocks),
                              chain_qr_blocks=int(info.chain_qr_blocks))


def block_gmres_preconditioned_with_guess(op: DenseMatrixOperator, B, precond=None, X0=None,
                                          config: Optional[BlockGmresConfig] = None) -> BlockGmresSolution:
    """Shared-basis block GMRES (block_gmres.rs:163-215): B is n x k (k <= 64),
    host arrays; precond None = no preconditioner."""
    config = config or BlockGmresConfig()
    n = op.num_rows()
    Bv = np.ascontiguousarray(B, dtype=np.complex128)
    if Bv.ndim == 1:
        Bv = Bv.reshape(n, 1)
    k = Bv.shape[1]
    X = np.zeros_like(Bv)
    X0v = None if X0 is None else np.ascontiguousarray(X0, dtype=np.complex128).reshape(n, k)
    info = _ffi.block_gmres_info_t()
    res, tres, berr = (np.zeros(k) for _ in range(3))
    _check(_lib().sbh_block_gmres(op.handle, None if precond is None else precond._h, _cptr(Bv),
                                  None if X0v is None else _cptr(X0v), k, ctypes.byref(config._c()), _cptr(X),
                                  MEM_HOST, ctypes.byref(info), _cptr(res), _cptr(tres), _cptr(berr)))
    return _block_solution(info, X, res, tres, berr)


def block_gmres(op: DenseMatrixOperator, B, config: Optional[BlockGmresConfig] = None) -> BlockGmresSolution:
    """block_gmres (block_gmres.rs:149-161)."""
    return block_gmres_preconditioned_with_guess(op, B, None, None, config)


def block_gmres_device(op: DenseMatrixOperator, B: DeviceVector, k: int, X: DeviceVector,
                       config: Optional[BlockGmresConfig] = None, X0: Optional[DeviceVector] = None,
                       precond=None) -> BlockGmresSolution:
    """Shared-basis block GMRES on device-resident row-major n x k blocks."""
    config = config or BlockGmresConfig()
    _block_vectors(op, B, X, k, X0)
    info = _ffi.block_gmres_info_t()
    res, tres, berr = (np.zeros(k) for _ in range(3))
    _check(_lib().sbh_block_gmres(op.handle, None if precond is None else precond._h, ctypes.c_void_p(B.ptr),
                                  None if X0 is None else ctypes.c_void_p(X0.ptr), k, ctypes.byref(config._c()),
                                  ctypes.c_void_p(X.ptr), MEM_DEVICE, ctypes.byref(info), _cptr(res), _cptr(tres),
                                  _cptr(berr)))
    return _block_solution(info, None, res, tres, berr)


def _block_vectors(op, B: DeviceVector, X: DeviceVector, k: int, X0: Optional[DeviceVector] = None):
    n = op.num_rows()
    if not 1 <= k <= 64:
        raise ValueError("block width must be in 1..64")
    for v in (B, X) + (() if X0 is None else (X0,)):
        if v.length != n * k:
            raise ValueError(f"device blocks must hold n*k = {n * k} entries (row-major n x k)")


def gmres_block_device(op: DenseMatrixOperator, B: DeviceVector, k: int, X: DeviceVector, config: GmresConfig,
                       X0: Optional[DeviceVector] = None):
    """gmres_block on device-resident row-major n x k blocks; X stays in HBM."""
    _block_vectors(op, B, X, k, X0)
    raws = (_ffi.gmres_solution_t * k)()
    _check(_lib().sbh_gmres_block(op.handle, ctypes.c_void_p(B.ptr), None if X0 is None else ctypes.c_void_p(X0.ptr),
                                  k, ctypes.byref(config._c()), ctypes.c_void_p(X.ptr), MEM_DEVICE, raws))
    return [_solution(raws[c], None) for c in range(k)]


LOWP_FORMATS = {"fp16": 0, "bf16": 1}


class LowpMatrix:
    """Complex FP16/BF16 copy of a TbemSystem's matrix with power-of-two row scaling."""

    def __init__(self, system: TbemSystem, fmt: str = "fp16"):
        self.system = system
        self.format = fmt
        h = ctypes.c_void_p()
        _check(_lib().sbh_lowp_matrix_create(system.handle, LOWP_FORMATS[fmt], ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def update(self):
        _check(_lib().sbh_lowp_matrix_update(self._h))

    def apply(self, X) -> np.ndarray:
        n = self.system.num_dofs
        Xv = np.ascontiguousarray(X, dtype=np.complex128)
        k = 1 if Xv.ndim == 1 else Xv.shape[1]
        Y = np.empty_like(Xv)
        _check(_lib().sbh_lowp_matrix_apply(self._h, _cptr(Xv), _cptr(Y), k, MEM_HOST))
        return Y

    def download(self):
        """(halves as uint16 n x 2*ld, row inverse scales, ld)."""
        n = self.system.num_dofs
        ld = ctypes.c_int64()
        _check(_lib().sbh_lowp_matrix_download(self._h, None, None, ctypes.byref(ld)))
        halves = np.empty((n, 2 * ld.value), dtype=np.uint16)
        inv = np.empty(n)
        _check(_lib().sbh_lowp_matrix_download(self._h, _cptr(halves), _cptr(inv), ctypes.byref(ld)))
        return halves, inv, int(ld.value)

    def dequantized(self) -> np.ndarray:
        halves, inv, ld = self.download()
        n = self.system.num_dofs
        if self.format == "fp16":
            vals = halves.view(np.float16).astype(np.float64)
        else:
            vals = (halves.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        z = vals[:, 0:2 * n:2] + 1j * vals[:, 1:2 * n:2]
        return z * inv[:, None]

    def close(self):
        if getattr(self, "_h", None):
            _lib().sbh_lowp_matrix_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class GmresIrConfig:
    """GmresIrConfig (crates/sonium-compute/src/gmres_ir.rs:540-580); defaults as there."""

    max_outer_iterations: int = 30
    outer_tolerance: float = 1e-10
    max_inner_iterations: int = 50
    inner_tolerance: float = 1e-3
    restart_size: int = 0

    def _c(self):
        return _ffi.gmres_ir_config_t(int(self.max_outer_iterations), float(self.outer_tolerance),
                                      int(self.max_inner_iterations), float(self.inner_tolerance),
                                      int(self.restart_size))


@dataclass
class GmresIrStatus:
    outer_iterations: int
    total_inner_iterations: int
    estimated_residual: float
    final_outer_residual: float
    backward_error: float
    converged: bool
    status: str
    restart_cycles: int
    fp64_products: int
    lowp_products: int


def gmres_ir(op: DenseMatrixOperator, lp: LowpMatrix, B, config: GmresIrConfig):
    """Mixed-precision FGMRES-IR (gmres_ir.rs:665): returns (X, [GmresIrStatus per column])."""
    Bv = np.ascontiguousarray(B, dtype=np.complex128)
    n = op.num_rows()
    if Bv.ndim == 1:
        Bv = Bv.reshape(n, 1)
    k = Bv.shape[1]
    X = np.zeros_like(Bv)
    raws = (_ffi.gmres_ir_status_t * k)()
    cfg = config._c()
    _check(_lib().sbh_gmres_ir(op.handle, lp.handle, _cptr(Bv), k, ctypes.byref(cfg), _cptr(X), MEM_HOST, raws))
    return X, _ir_statuses(raws, k)


def gmres_ir_device(op: DenseMatrixOperator, lp: LowpMatrix, B: DeviceVector, k: int, X: DeviceVector,
                    config: GmresIrConfig):
    """gmres_ir on device-resident row-major n x k blocks; X stays in HBM."""
    _block_vectors(op, B, X, k)
    raws = (_ffi.gmres_ir_status_t * k)()
    _check(_lib().sbh_gmres_ir(op.handle, lp.handle, ctypes.c_void_p(B.ptr), k, ctypes.byref(config._c()),
                               ctypes.c_void_p(X.ptr), MEM_DEVICE, raws))
    return _ir_statuses(raws, k)


def _ir_statuses(raws, k):
    out = []
    for c in range(k):
        r = raws[c]
        out.append(GmresIrStatus(int(r.outer_iterations), int(r.total_inner_iterations), float(r.estimated_residual),
                                 float(r.final_outer_residual), float(r.backward_error), bool(r.converged),
                                 SOLVER_STATUS[int(r.status)], int(r.restart_cycles), int(r.fp64_products),
                                 int(r.lowp_products)))
    return out


BATCH_FORMATS = {"fp64": -1, "fp16_ir": 0, "bf16_ir": 1}


@dataclass
class BatchColumnStatus:
    iterations: int
    cycles: int
    reorthogonalizations: int
    residual: float
    converged: bool
    completion_ms: float


class Batch:
    """Many independent rigid scatterers assembled and solved together.

    ``meshes`` are :class:`TbemMesh` objects (velocity BC 0 on every panel);
    ``set_problem`` gives each one a wavenumber, Burton-Miller beta (default
    i/k) and 1..16 plane-wave directions.  ``assemble`` fills every FP64
    matrix and right-hand side in two persistent launches; ``solve`` runs all
    columns in lock step (restarted GMRES or GMRES iterative refinement on an
    FP16/BF16 copy) with each restart cycle replayed from one HIP graph.
    """

    def __init__(self, ctx: Context, meshes, route_threads: int = 0):
        self.ctx = ctx
        self.meshes = list(meshes)
        arr = (ctypes.c_void_p * len(self.meshes))(*[m.handle.value for m in self.meshes])
        h = ctypes.c_void_p()
        _check(_lib().sbh_batch_create(ctx.handle, len(self.meshes), arr, int(route_threads), ctypes.byref(h)))
        self._h = h
        self.num_rhs = None

    @property
    def handle(self):
        return self._h

    def info(self) -> dict:
        b, c, p, e = (ctypes.c_int64() for _ in range(4))
        r = ctypes.c_double()
        _check(_lib().sbh_batch_info(self._h, ctypes.byref(b), ctypes.byref(c), ctypes.byref(p), ctypes.byref(e),
                                     ctypes.byref(r)))
        return {"instances": b.value, "columns": c.value, "panels": p.value, "matrix_entries": e.value,
                "route_ms": r.value}

    def set_problem(self, wave_numbers, directions, betas=None, sign: float = 1.0):
        nb = len(self.meshes)
        ks = [float(k) for k in wave_numbers]
        if len(ks) != nb or len(directions) != nb:
            raise ValueError("one wavenumber and one direction array per instance")
        phys = (_ffi.physics_t * nb)()
        bet = (_ffi.c64 * nb)()
        for i, k in enumerate(ks):
            ph = PhysicsParams.from_wavenumber(k)
            phys[i] = ph._p
            bv = ph.burton_miller_beta() if betas is None else complex(betas[i])
            bet[i] = _ffi.c64(bv.real, bv.imag)
        dirs = [np.ascontiguousarray(d, dtype=np.float64).reshape(-1, 3) for d in directions]
        counts = np.array([d.shape[0] for d in dirs], dtype=np.int32)
        flat = np.ascontiguousarray(np.concatenate(dirs, axis=0))
        _check(_lib().sbh_batch_set_problem(self._h, phys, bet, _cptr(counts), _cptr(flat), float(sign)))
        self.num_rhs = counts

    def assemble(self):
        _check(_lib().sbh_batch_assemble(self._h))

    def convert(self, fmt: str):
        _check(_lib().sbh_batch_convert(self._h, {"fp16": 0, "bf16": 1}[fmt]))

    def solve(self, fmt: str = "fp64", restart: int = 100, max_cycles: int = 10, tolerance: float = 1e-8,
              inner_tolerance: float = 1e-4, use_graph: bool = True, use_guess: bool = False):
        cfg = _ffi.batch_solve_config_t(BATCH_FORMATS[fmt], 1 if use_graph else 0, int(restart), int(max_cycles),
                                        float(tolerance), float(inner_tolerance), 1 if use_guess else 0, 0)
        ncol = int(self.num_rhs.sum())
        raws = (_ffi.batch_column_status_t * ncol)()
        _check(_lib().sbh_batch_solve(self._h, ctypes.byref(cfg), raws))
        return [BatchColumnStatus(int(r.iterations), int(r.cycles), int(r.reorthogonalizations), float(r.residual),
                                  bool(r.converged), float(r.completion_ms)) for r in raws]

    def set_guess(self, i: int, X0) -> None:
        """Initial guess of instance i (n x num_rhs), used by solve(use_guess=True)."""
        x = np.ascontiguousarray(X0, dtype=np.complex128).reshape(self.meshes[i].num_elements, int(self.num_rhs[i]))
        _check(_lib().sbh_batch_set_guess(self._h, int(i), _cptr(x)))

    def lowp_copy(self, i: int):
        """(halves n x 2*ld as uint16, row scales, ld) of instance i's FP16/BF16 copy."""
        n = self.meshes[i].num_elements
        ld = ctypes.c_int64()
        _check(_lib().sbh_batch_download_lowp(self._h, int(i), None, None, ctypes.byref(ld)))
        halves = np.empty((n, 2 * ld.value), dtype=np.uint16)
        inv = np.empty(n)
        _check(_lib().sbh_batch_download_lowp(self._h, int(i), _cptr(halves), _cptr(inv), ctypes.byref(ld)))
        return halves, inv, int(ld.value)

    def _packed_shapes(self):
        return [(m.num_elements, int(c)) for m, c in zip(self.meshes, self.num_rhs)]

    def _unpack(self, flat):
        res, o = [], 0
        for shp in self._packed_shapes():
            cnt = shp[0] * shp[1]
            res.append(flat[o:o + cnt].reshape(shp))
            o += cnt
        return res

    def set_guesses(self, X0) -> None:
        """Every instance's initial guess (list of n_i x num_rhs[i]) in one upload."""
        flat = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.complex128).reshape(-1) for x in X0]))
        if flat.size != sum(a * b for a, b in self._packed_shapes()):
            raise ValueError("one n_i x num_rhs[i] block per instance")
        _check(_lib().sbh_batch_set_guesses(self._h, _cptr(flat)))

    @staticmethod
    def _pointer_table(ptrs):
        return (ctypes.c_void_p * len(ptrs))(*[None if p is None else int(p) for p in ptrs])

    def set_guesses_device(self, pointers) -> None:
        """Initial guesses copied on the device: pointers[i] is a device
        address of instance i's n_i x num_rhs[i] block (None: zero)."""
        if len(pointers) != len(self.meshes):
            raise ValueError("one pointer per instance")
        _check(_lib().sbh_batch_set_guesses_device(self._h, self._pointer_table(pointers)))

    def store_solutions(self, pointers) -> None:
        """Each instance's solution block copied to the device address pointers[i] (None: skipped)."""
        if len(pointers) != len(self.meshes):
            raise ValueError("one pointer per instance")
        _check(_lib().sbh_batch_store_solutions(self._h, self._pointer_table(pointers)))

    def guess_gram(self, candidates, fmt: str = "fp64"):
        """Minimal-residual data of device-resident candidate guesses.
        candidates[a][i] is the device address of instance i's candidate a
        (None: zero).  Returns (G, h, bb) per column: G[c] = P^H P (m x m),
        h[c] = P^H b, bb[c] = ||b||^2 with P_a = A X_a."""
        m = len(candidates)
        if any(len(c) != len(self.meshes) for c in candidates):
            raise ValueError("one pointer per instance in every candidate slot")
        ncol = int(self.num_rhs.sum())
        G = np.empty((ncol, m, m), dtype=np.complex128)
        h = np.empty((ncol, m), dtype=np.complex128)
        bb = np.empty(ncol)
        table = self._pointer_table([p for slot in candidates for p in slot])
        _check(_lib().sbh_batch_guess_gram(self._h, {"fp64": -1, "fp16": 0, "bf16": 1}[fmt], m, table, _cptr(G),
                                           _cptr(h), _cptr(bb)))
        return G, h, bb

    def guess_combine(self, coefficients) -> None:
        """Initial guesses sum_a coefficients[c, a] X_a per column, from the
        candidates of the last guess_gram."""
        cf = np.ascontiguousarray(coefficients, dtype=np.complex128)
        if cf.ndim != 2 or cf.shape[0] != int(self.num_rhs.sum()):
            raise ValueError("coefficients must be (columns, candidates)")
        _check(_lib().sbh_batch_guess_combine(self._h, cf.shape[1], _cptr(cf)))

    def learned_guesses(self, net, out: Optional[DeviceVector] = None) -> None:
        """The network's guesses for every column, computed on the device, into
        ``out`` (packed layout) or as the batch's initial guesses."""
        _check(_lib().sbh_batch_learned_guesses(self._h, net._h, None if out is None else out.ptr))

    def solutions(self):
        """Every instance's solution block, in one download."""
        out = np.empty(sum(a * b for a, b in self._packed_shapes()), dtype=np.complex128)
        _check(_lib().sbh_batch_download_all(self._h, 0, _cptr(out)))
        return self._unpack(out)

    def rhs_all(self):
        """Every instance's right-hand-side block, in one download."""
        out = np.empty(sum(a * b for a, b in self._packed_shapes()), dtype=np.complex128)
        _check(_lib().sbh_batch_download_all(self._h, 1, _cptr(out)))
        return self._unpack(out)

    def apply(self, X, fmt: str = "fp64"):
        """Products A_i X_i of every instance with its own matrix in one batched
        launch (the stored format: "fp64", or "fp16"/"bf16" after convert).
        X is a list of n_i x num_rhs[i] arrays; returns the list of products."""
        if self.num_rhs is None:
            raise ValueError("set_problem first: the column layout follows the right-hand sides")
        if len(X) != len(self.meshes):
            raise ValueError("one block per instance")
        blocks = [np.ascontiguousarray(x, dtype=np.complex128).reshape(m.num_elements, int(c))
                  for x, m, c in zip(X, self.meshes, self.num_rhs)]
        flat = np.ascontiguousarray(np.concatenate([b.reshape(-1) for b in blocks]))
        out = np.empty_like(flat)
        _check(_lib().sbh_batch_apply(self._h, {"fp64": -1, "fp16": 0, "bf16": 1}[fmt], _cptr(flat), _cptr(out)))
        res, o = [], 0
        for b in blocks:
            res.append(out[o:o + b.size].reshape(b.shape))
            o += b.size
        return res

    def _download(self, fn, i, shape):
        out = np.empty(shape, dtype=np.complex128)
        _check(fn(self._h, int(i), _cptr(out)))
        return out

    def solution(self, i: int) -> np.ndarray:
        return self._download(_lib().sbh_batch_download_solution, i, (self.meshes[i].num_elements, self.num_rhs[i]))

    def rhs(self, i: int) -> np.ndarray:
        return self._download(_lib().sbh_batch_download_rhs, i, (self.meshes[i].num_elements, self.num_rhs[i]))

    def matrix(self, i: int) -> np.ndarray:
        n = self.meshes[i].num_elements
        return self._download(_lib().sbh_batch_download_matrix, i, (n, n))

    def tuning(self) -> dict:
        t = _ffi.batch_tuning_t()
        _check(_lib().sbh_batch_get_tuning(self._h, ctypes.byref(t)))
        return {"rows_per_item": t.rows_per_item, "blocks_per_cu": t.blocks_per_cu, "chunk_elems": t.chunk_elems}

    def set_tuning(self, rows_per_item: int, blocks_per_cu: int, chunk_elems: int):
        t = _ffi.batch_tuning_t(int(rows_per_item), int(blocks_per_cu), int(chunk_elems), 0)
        _check(_lib().sbh_batch_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self):
        """Times the candidate launch shapes on this batch; keeps and returns the fastest."""
        table = np.zeros((64, 4))
        cnt = ctypes.c_int()
        best = _ffi.batch_tuning_t()
        _check(_lib().sbh_batch_autotune(self._h, _cptr(table), 64, ctypes.byref(cnt), ctypes.byref(best)))
        rows = [dict(zip(("rows_per_item", "blocks_per_cu", "chunk_elems", "ms"), r)) for r in table[: cnt.value]]
        return {"rows_per_item": best.rows_per_item, "blocks_per_cu": best.blocks_per_cu,
                "chunk_elems": best.chunk_elems}, rows

    def close(self):
        if getattr(self, "_h", None):
            _lib().sbh_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_sincos(ctx: Context, x) -> tuple:
    """The assembly kernels' sin/cos evaluated on the device (test hook)."""
    xv = np.ascontiguousarray(x, dtype=np.float64)
    s = np.empty_like(xv)
    c = np.empty_like(xv)
    _check(_lib().sbh_debug_device_sincos(ctx.handle, _cptr(xv), xv.size, _cptr(s), _cptr(c)))
    return s, c


def device_rsqrt(ctx: Context, x) -> np.ndarray:
    """The kernels' 1/sqrt evaluated on the device (test hook)."""
    xv = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(xv)
    _check(_lib().sbh_debug_device_rsqrt(ctx.handle, _cptr(xv), xv.size, _cptr(y)))
    return y


def debug_subelements(source_point, element_coords, element_area: float) -> np.ndarray:
    """The host route builder's subdivision of one pair (test hook, no GPU)."""
    src = np.ascontiguousarray(source_point, dtype=np.float64).reshape(3)
    coords = np.ascontiguousarray(element_coords, dtype=np.float64).reshape(9)
    out = np.zeros((110, 6))
    cnt = ctypes.c_int()
    _check(_lib().sbh_debug_subelements(_cptr(src), _cptr(coords), float(element_area), _cptr(out),
                                        ctypes.byref(cnt)))
    return out[: cnt.value].copy()


__all__ = [
    "Context", "DeviceVector", "PhysicsParams", "TbemMesh", "TbemSystem", "TbemAssemblyWorkspace",
    "build_tbem_system", "build_tbem_system_with_beta", "IncidentField", "DenseMatrixOperator",
    "IdentityPreconditioner", "JacobiPreconditioner", "GmresConfig", "GmresSolution", "gmres",
    "gmres_with_guess", "gmres_preconditioned", "gmres_preconditioned_with_guess", "gmres_device",
    "Mesh", "SbhError", "MEM_HOST", "MEM_DEVICE", "gmres_block", "LowpMatrix", "GmresIrConfig",
    "GmresIrStatus", "gmres_ir", "Batch", "BatchColumnStatus", "TbemExactOperator", "gmres_block_device",
    "gmres_ir_device",
]
