"""MI355X (gfx950) BioFVM diffusion-decay: Python host side over the C ABI.

``Microenvironment`` mirrors the part of the reference's
``BioFVM::Microenvironment`` (BioFVM/BioFVM_microenvironment.h:71-373) that the
LOD diffusion-decay path touches: resize_space, diffusion_coefficients /
decay_rates, Dirichlet activation and face values, density access,
simulate_diffusion_decay(dt), simulate_cell_sources_and_sinks(dt) and
apply_dirichlet_boundaries_conditions.  All compute runs in the HIP library
``lib/libbiofvm_amd.so`` built from ``csrc/``; there is no CPU fallback, and a
missing library raises at first use.

Field arrays on the host are float64 in the reference layout
``((i*ny + j)*nz + k)*ns + d`` unless ``layout="soa"`` is passed
(``((d*nz + k)*ny + j)*nx + i``).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libbiofvm_amd.so")

FP64, FP32 = 0, 1
LAYOUT_REF, LAYOUT_SOA = 0, 1
_PREC = {"fp64": FP64, "f64": FP64, "double": FP64, FP64: FP64,
         "fp32": FP32, "f32": FP32, "float": FP32, FP32: FP32}
_LAYOUT = {"ref": LAYOUT_REF, "soa": LAYOUT_SOA, LAYOUT_REF: LAYOUT_REF, LAYOUT_SOA: LAYOUT_SOA}

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_u8 = ctypes.POINTER(ctypes.c_uint8)
_i32 = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p


class BioFVMError(RuntimeError):
    pass


def build(force: bool = False) -> str:
    """Compile lib/libbiofvm_amd.so for gfx950 with hipcc (cross-compiles without a GPU)."""
    cmd = ["make", "-C", HERE]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BioFVMError(f"{LIB_PATH} is missing: build it first (make -C {HERE})")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER(ctypes.c_void_p)
    L.bfa_create.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                             ctypes.c_double, ctypes.c_int]
    L.bfa_destroy.argtypes = [_vp]
    L.bfa_destroy.restype = None
    L.bfa_last_error.argtypes = [_vp]
    L.bfa_last_error.restype = ctypes.c_char_p
    L.bfa_set_coefficients.argtypes = [_vp, _dp, _dp]
    L.bfa_set_dirichlet.argtypes = [_vp, _u8, _u8, _dp]
    L.bfa_set_dirichlet_planes.argtypes = [_vp, _dp]
    L.bfa_set_solver.argtypes = [_vp, ctypes.c_int]
    L.bfa_track_internalized.argtypes = [_vp, ctypes.c_int, _dp, _dp]
    L.bfa_get_internalized.argtypes = [_vp, _dp]
    L.bfa_get_solver.argtypes = [_vp]
    L.bfa_set_dt_caching.argtypes = [_vp, ctypes.c_int]
    L.bfa_upload.argtypes = [_vp, _dp, ctypes.c_int]
    L.bfa_download.argtypes = [_vp, _dp, ctypes.c_int]
    L.bfa_fill.argtypes = [_vp, _dp]
    L.bfa_set_fp32_basis.argtypes = [_vp, _dp, _dp]
    L.bfa_get_fp32_basis.argtypes = [_vp, _dp, _dp]
    L.bfa_set_agents.argtypes = [_vp, ctypes.c_int, _i32, _u8, _dp, _dp, _dp]
    L.bfa_diffusion_decay.argtypes = [_vp, ctypes.c_double]
    L.bfa_secretion_uptake.argtypes = [_vp]
    L.bfa_step.argtypes = [_vp, ctypes.c_double, ctypes.c_int]
    L.bfa_apply_dirichlet.argtypes = [_vp]
    L.bfa_total_mass.argtypes = [_vp, _dp]
    L.bfa_synchronize.argtypes = [_vp]
    L.bfa_time_kernels.argtypes = [_vp, ctypes.c_double, ctypes.c_int, _dp]
    L.bfa_copy_kernel.argtypes = [ctypes.c_size_t, ctypes.c_int, _dp]
    L.bfa_copy_kernel_stream.argtypes = [ctypes.c_size_t, ctypes.c_int, _dp]
    L.bfa_set_tiles.argtypes = [_vp] + [ctypes.c_int] * 6
    L.bfa_get_tiles.argtypes = [_vp, ctypes.POINTER(ctypes.c_int)]
    L.bfa_set_slab.argtypes = [_vp, ctypes.c_int]
    L.bfa_get_slab.argtypes = [_vp]
    L.bfa_batch_create.argtypes = [P, ctypes.c_int]
    L.bfa_batch_destroy.argtypes = [_vp]
    L.bfa_batch_destroy.restype = None
    L.bfa_batch_last_error.argtypes = [_vp]
    L.bfa_batch_last_error.restype = ctypes.c_char_p
    L.bfa_batch_member_bytes.argtypes = [ctypes.c_int] * 5
    L.bfa_batch_member_bytes.restype = ctypes.c_size_t
    L.bfa_batch_add.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_double, _dp, _dp, _u8, _u8, _dp,
                                ctypes.POINTER(ctypes.c_int)]
    L.bfa_batch_remove.argtypes = [_vp, ctypes.c_int]
    L.bfa_batch_upload.argtypes = [_vp, ctypes.c_int, _dp, ctypes.c_int]
    L.bfa_batch_download.argtypes = [_vp, ctypes.c_int, _dp, ctypes.c_int]
    L.bfa_batch_set_agents.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _i32, _u8, _dp, _dp, _dp]
    L.bfa_batch_track_internalized.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp]
    L.bfa_batch_get_internalized.argtypes = [_vp, ctypes.c_int, _dp]
    L.bfa_batch_set_member_steps.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
    L.bfa_batch_fill.argtypes = [_vp, ctypes.c_int, _dp]
    L.bfa_batch_step.argtypes = [_vp, ctypes.c_double, ctypes.c_int]
    L.bfa_batch_synchronize.argtypes = [_vp]
    L.bfa_batch_total_mass.argtypes = [_vp, ctypes.c_int, _dp]
    L.bfa_batch_set_tiles.argtypes = [_vp] + [ctypes.c_int] * 5
    L.bfa_batch_get_tiles.argtypes = [_vp, ctypes.POINTER(ctypes.c_int)]
    L.bfa_set_agents_state.argtypes = [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _u8, ctypes.c_double]
    L.bfa_set_bulk_rates.argtypes = [_vp, _dp, _dp, _dp, ctypes.c_int]
    L.bfa_bulk_sources_and_sinks.argtypes = [_vp, ctypes.c_double]
    L.bfa_time_gradient.argtypes = [_vp, ctypes.c_int, _dp]
    L.bfa_gradients.argtypes = [_vp, _dp, ctypes.c_int]
    L.bfa_gradients_at.argtypes = [_vp, ctypes.c_int, _i32, _dp]
    L.bfa_set_substrate_streams.argtypes = [_vp, ctypes.c_int]
    L.bfa_get_substrate_streams.argtypes = [_vp]
    L.bfa_set_x_grid.argtypes = [_vp, ctypes.c_int]
    L.bfa_get_x_grid.argtypes = [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.bfa_set_col_runs.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
    L.bfa_set_agent_stores.argtypes = [_vp, ctypes.c_int]
    L.bfa_get_agent_stores.argtypes = [_vp]
    L.bfa_get_col_runs.argtypes = [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.bfa_batch_set_fusion.argtypes = [_vp, ctypes.c_int]
    L.bfa_batch_set_col_threads.argtypes = [_vp, ctypes.c_int]
    L.bfa_batch_set_streams.argtypes = [_vp, ctypes.c_int]
    L.bfa_batch_get_streams.argtypes = [_vp]
    L.bfa_batch_get_col_threads.argtypes = [_vp, ctypes.POINTER(ctypes.c_int)]
    L.bfa_batch_fused_members.argtypes = [_vp]
    _lib = L
    return L


def _d(a):
    return a.ctypes.data_as(_dp)


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise BioFVMError(f"expected {n} values, got {a.size}")
    return a


def _flags(a, n):
    """Per-agent activity flags as a contiguous uint8 array of exactly n
    entries (the C side reads n of them), or None for all active."""
    if a is None:
        return None
    a = np.ascontiguousarray(np.asarray(a).ravel(), dtype=np.uint8)
    if a.size != n:
        raise BioFVMError(f"expected {n} activity flags, got {a.size}")
    return a


def copy_kernel_ms(nbytes: int, reps: int = 20) -> float:
    """Mean time of a plain device copy of nbytes (HBM counter calibration)."""
    out = ctypes.c_double()
    rc = lib().bfa_copy_kernel(int(nbytes), int(reps), ctypes.byref(out))
    if rc:
        raise BioFVMError(f"bfa_copy_kernel failed with status {rc}")
    return out.value


def copy_stream_ms(nbytes: int, reps: int = 20) -> float:
    """Mean time of the one-pass streaming device copy of nbytes (the
    achievable-bandwidth reference)."""
    out = ctypes.c_double()
    rc = lib().bfa_copy_kernel_stream(int(nbytes), int(reps), ctypes.byref(out))
    if rc:
        raise BioFVMError(f"bfa_copy_kernel_stream failed with status {rc}")
    return out.value


class Microenvironment:
    """One microenvironment resident on the current GPU.

    ``Microenvironment(nx, ny, nz, ns, dx, precision)`` replaces
    ``resize_space`` (BioFVM/BioFVM_microenvironment.cpp:610-637) for a
    uniform mesh; the field starts at zero.
    """

    def __init__(self, nx, ny, nz, ns, dx, precision="fp64"):
        self._h = None
        L = lib()
        h = ctypes.c_void_p()
        rc = L.bfa_create(ctypes.byref(h), int(nx), int(ny), int(nz), int(ns), float(dx),
                          _PREC[precision])
        if rc:
            raise BioFVMError(f"bfa_create: {L.bfa_last_error(None).decode()}")
        self._h = h
        self.nx, self.ny, self.nz, self.ns = int(nx), int(ny), int(nz), int(ns)
        self.dx = float(dx)
        self.precision = _PREC[precision]
        self.diffusion_coefficients = np.zeros(self.ns)
        self.decay_rates = np.zeros(self.ns)
        self.dirichlet_activation_vector = np.zeros(self.ns, np.uint8)
        self.face_registered = np.zeros(6, np.uint8)
        self.face_values = np.ones((6, self.ns))

    # -- lifecycle ------------------------------------------------------------
    def close(self):
        if self._h is not None:
            lib().bfa_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc:
            msg = lib().bfa_last_error(self._h).decode()
            raise BioFVMError(f"{what}: status {rc}: {msg}")

    @property
    def number_of_voxels(self):
        return self.nx * self.ny * self.nz

    def number_of_densities(self):
        return self.ns

    # -- configuration ----------------------------------------------------------
    def set_coefficients(self, D, decay):
        self.diffusion_coefficients = _f64(D, self.ns).copy()
        self.decay_rates = _f64(decay, self.ns).copy()
        self._check(lib().bfa_set_coefficients(self._h, _d(self.diffusion_coefficients),
                                               _d(self.decay_rates)), "set_coefficients")

    def set_dirichlet(self, active, face_registered, face_values):
        """dirichlet_activation_vector[ns], faces registered [6] (xmin..zmax) and
        their value vectors [6][ns] as add_dirichlet_node would store them."""
        self.dirichlet_activation_vector = np.ascontiguousarray(active, np.uint8).reshape(self.ns)
        self.face_registered = np.ascontiguousarray(face_registered, np.uint8).reshape(6)
        self.face_values = _f64(face_values, 6 * self.ns).reshape(6, self.ns).copy()
        self._check(lib().bfa_set_dirichlet(self._h,
                                            self.dirichlet_activation_vector.ctypes.data_as(_u8),
                                            self.face_registered.ctypes.data_as(_u8),
                                            _d(self.face_values)), "set_dirichlet")

    def set_dirichlet_planes(self, planes):
        """Per-voxel Dirichlet values of the six face planes (the face-voxel
        entries of dirichlet_value_vectors): a sequence of six arrays xmin, xmax
        [ny][nz][ns], ymin, ymax [nx][nz][ns], zmin, zmax [nx][ny][ns], or one
        array of them back to back; None returns to the face values."""
        if planes is None:
            self._check(lib().bfa_set_dirichlet_planes(self._h, None), "set_dirichlet_planes")
            return
        if not isinstance(planes, np.ndarray):
            planes = np.concatenate([np.asarray(p, np.float64).ravel() for p in planes])
        nx, ny, nz, ns = self.nx, self.ny, self.nz, self.ns
        total = 2 * (ny * nz + nx * nz + nx * ny) * ns
        flat = _f64(planes, total)
        self._check(lib().bfa_set_dirichlet_planes(self._h, _d(flat)), "set_dirichlet_planes")

    SOLVERS = {"blocking": 0, "avx256d": 1}

    def set_solver(self, name):
        """Which reference solver the step reproduces: "blocking"
        (LOD_3D_BLOCKING, Dirichlet after the z sweep too) or "avx256d"
        (LOD_3D_AVX256D, no Dirichlet pass after the z sweep)."""
        if name not in self.SOLVERS:
            raise BioFVMError(f"unknown solver {name!r}")
        self._check(lib().bfa_set_solver(self._h, self.SOLVERS[name]), "set_solver")

    def set_dt_caching(self, keep_first):
        """keep_first=True keeps the Thomas coefficients of the first dt for
        later dts, as the reference's solver does (diffusion_solver_setup_done)."""
        self._check(lib().bfa_set_dt_caching(self._h, 1 if keep_first else 0), "set_dt_caching")

    def solver(self):
        v = lib().bfa_get_solver(self._h)
        return {0: "blocking", 1: "avx256d"}[v]

    def set_substrate_dirichlet_activation(self, index, flag):
        """Microenvironment::set_substrate_dirichlet_activation
        (BioFVM/BioFVM_microenvironment.cpp:265-273)."""
        act = self.dirichlet_activation_vector.copy()
        act[index] = 1 if flag else 0
        self.set_dirichlet(act, self.face_registered, self.face_values)

    def set_fp32_basis(self, b, s):
        self._check(lib().bfa_set_fp32_basis(self._h, _d(_f64(b, self.ns)), _d(_f64(s, self.ns))),
                    "set_fp32_basis")

    def fp32_basis(self):
        b = np.zeros(self.ns)
        s = np.zeros(self.ns)
        self._check(lib().bfa_get_fp32_basis(self._h, _d(b), _d(s)), "get_fp32_basis")
        return b, s

    def set_tiles(self, Lx=0, Ly=0, Lz=0, rows_x=0, cols_y=0, cols_z=0):
        """Kernel shapes (see bfa_set_tiles); 0 means automatic."""
        self._check(lib().bfa_set_tiles(self._h, int(Lx), int(Ly), int(Lz), int(rows_x),
                                        int(cols_y), int(cols_z)), "set_tiles")

    def tiles(self):
        """(Lx, Ly, Lz, rows_x, cols_y, cols_z) currently in use."""
        v = (ctypes.c_int * 6)()
        self._check(lib().bfa_get_tiles(self._h, v), "get_tiles")
        return tuple(v)

    def set_slab(self, planes):
        """x/y sweeps in k-slabs of `planes` planes (0: whole-volume sweeps)."""
        self._check(lib().bfa_set_slab(self._h, int(planes)), "set_slab")

    def slab(self):
        return lib().bfa_get_slab(self._h)

    def set_agents_state(self, position, bbox, volume, secretion, uptake, export_rate, saturation,
                         dt, active=None):
        """Agents from their raw state (update_voxel_index and
        set_internal_uptake_constants computed on the device)."""
        pos = _f64(position).reshape(-1, 3)
        n = pos.shape[0]
        ns = self.ns
        args = [_f64(a, n * ns) for a in (secretion, uptake, export_rate, saturation)]
        vol = _f64(volume, n)
        box = _f64(bbox, 6)
        act = None
        if active is not None:
            act = np.ascontiguousarray(np.asarray(active, np.uint8).ravel())
            if act.size != n:
                raise BioFVMError(f"expected {n} activity flags, got {act.size}")
        self._check(lib().bfa_set_agents_state(self._h, n, _d(pos), _d(box), _d(vol), *(_d(a) for a in args),
                                               act.ctypes.data_as(_u8) if act is not None else None,
                                               float(dt)), "set_agents_state")
        self._n_agents = n

    def set_bulk_rates(self, S, T, U, layout="ref"):
        """Per-value supply rate S, supply target T, uptake rate U for
        simulate_bulk_sources_and_sinks; None for all three releases them."""
        if S is None and T is None and U is None:
            self._check(lib().bfa_set_bulk_rates(self._h, None, None, None, _LAYOUT[layout]), "set_bulk_rates")
            return
        n = self.number_of_voxels * self.ns
        S, T, U = (_f64(a, n) for a in (S, T, U))
        self._check(lib().bfa_set_bulk_rates(self._h, _d(S), _d(T), _d(U), _LAYOUT[layout]), "set_bulk_rates")

    def simulate_bulk_sources_and_sinks(self, dt):
        self._check(lib().bfa_bulk_sources_and_sinks(self._h, float(dt)), "bulk_sources_and_sinks")

    def gradients(self, layout="ref"):
        """Gradient field (compute_all_gradient_vectors): array (nvox, ns, 3) in
        reference voxel order, or (ns, 3, nvox) device order for layout="soa"."""
        out = np.empty(self.number_of_voxels * self.ns * 3)
        self._check(lib().bfa_gradients(self._h, _d(out), _LAYOUT[layout]), "gradients")
        if layout == "ref":
            return out.reshape(self.number_of_voxels, self.ns, 3)
        return out.reshape(self.ns, 3, self.number_of_voxels)

    def time_gradient(self, reps=10):
        """Mean device time (ms) of the whole-field gradient kernel."""
        ms = np.zeros(1)
        self._check(lib().bfa_time_gradient(self._h, int(reps), _d(ms)), "time_gradient")
        return float(ms[0])

    def gradients_at(self, voxels):
        """Gradient at the given reference voxel indices: array (n, ns, 3)."""
        v = np.ascontiguousarray(np.asarray(voxels, dtype=np.int32).ravel())
        out = np.empty(v.size * self.ns * 3)
        self._check(lib().bfa_gradients_at(self._h, int(v.size), v.ctypes.data_as(_i32), _d(out)),
                    "gradients_at")
        return out.reshape(v.size, self.ns, 3)

    def set_substrate_streams(self, groups):
        """1: one launch per sweep for all substrates; 2: two substrate halves
        swept on parallel streams (same results)."""
        self._check(lib().bfa_set_substrate_streams(self._h, int(groups)), "set_substrate_streams")

    def substrate_streams(self):
        return lib().bfa_get_substrate_streams(self._h)

    def set_x_grid(self, blocks_per_cu):
        """Blocks per compute unit of the wave-local x sweep (0: as many as fit)."""
        self._check(lib().bfa_set_x_grid(self._h, int(blocks_per_cu)), "set_x_grid")

    def x_grid(self):
        """(requested blocks per CU, blocks per CU that fit the current x shape)."""
        req, res = ctypes.c_int(), ctypes.c_int()
        self._check(lib().bfa_get_x_grid(self._h, ctypes.byref(req), ctypes.byref(res)), "x_grid")
        return req.value, res.value

    def set_col_runs(self, runs_y, runs_z):
        """Tile sub-runs each XCD advances together in the y and z sweeps (1: one run)."""
        self._check(lib().bfa_set_col_runs(self._h, int(runs_y), int(runs_z)), "set_col_runs")

    def col_runs(self):
        """(runs_y, runs_z)."""
        ry, rz = ctypes.c_int(), ctypes.c_int()
        self._check(lib().bfa_get_col_runs(self._h, ctypes.byref(ry), ctypes.byref(rz)), "col_runs")
        return ry.value, rz.value

    def set_agent_stores(self, whole_lines):
        """The agent step stores whole 128-byte lines (True) or single voxels (False, default)."""
        self._check(lib().bfa_set_agent_stores(self._h, 1 if whole_lines else 0), "set_agent_stores")

    def agent_stores(self):
        """True when the agent step stores whole lines."""
        return lib().bfa_get_agent_stores(self._h) == 1

    # -- field transfer -----------------------------------------------------------
    def upload(self, field, layout="ref"):
        f = _f64(field, self.number_of_voxels * self.ns)
        self._check(lib().bfa_upload(self._h, _d(f), _LAYOUT[layout]), "upload")

    def fill(self, values):
        """Every voxel of substrate d set to values[d], on the device."""
        v = _f64(values, self.ns)
        self._check(lib().bfa_fill(self._h, _d(v)), "fill")

    def download(self, layout="ref", out=None):
        """The field as FP64 in the given layout, into `out` when given (a
        C-contiguous float64 array of number_of_voxels * ns values)."""
        size = self.number_of_voxels * self.ns
        if out is None:
            out = np.empty(size)
        elif out.dtype != np.float64 or out.size != size or not out.flags.c_contiguous:
            raise ValueError(f"download: out must be a C-contiguous float64 array of {size} values")
        self._check(lib().bfa_download(self._h, _d(out), _LAYOUT[layout]), "download")
        return out

    # -- agents --------------------------------------------------------------------
    def set_agents(self, voxel, c1, c2, e2, active=None):
        """Per-agent constants in agent order (reference voxel numbering)."""
        voxel = np.ascontiguousarray(voxel, np.int32)
        n = voxel.size
        c1 = _f64(c1, n * self.ns)
        c2 = _f64(c2, n * self.ns)
        e2 = _f64(e2, n * self.ns)
        act = _flags(active, n)
        self._check(lib().bfa_set_agents(self._h, n, voxel.ctypes.data_as(_i32),
                                         None if act is None else act.ctypes.data_as(_u8),
                                         _d(c1), _d(c2), _d(e2)), "set_agents")
        self._n_agents = n

    def track_internalized(self, initial=None, e1=None):
        """Switch on internalised-substrate tracking for the current agents
        (track_internalized_substrates_in_each_agent): initial totals [n][ns]
        (None: zeros) and temp_export1 e1 [n][ns] (None: from
        set_agents_state's export rates, or zeros)."""
        n = getattr(self, "_n_agents", 0)
        cnt = n * self.ns
        ini = None if initial is None else _f64(initial, cnt)
        ee = None if e1 is None else _f64(e1, cnt)
        self._check(lib().bfa_track_internalized(self._h, 1, None if ini is None else _d(ini),
                                                 None if ee is None else _d(ee)), "track_internalized")

    def stop_tracking_internalized(self):
        self._check(lib().bfa_track_internalized(self._h, 0, None, None), "track_internalized")

    def internalized(self):
        """Per-agent internalised totals [n][ns] (caller's agent order)."""
        n = getattr(self, "_n_agents", 0)
        out = np.empty((n, self.ns))
        self._check(lib().bfa_get_internalized(self._h, _d(out)), "get_internalized")
        return out

    # -- stepping ------------------------------------------------------------------
    def simulate_diffusion_decay(self, dt):
        self._check(lib().bfa_diffusion_decay(self._h, float(dt)), "simulate_diffusion_decay")

    def simulate_cell_sources_and_sinks(self, dt=None):
        """Agent secretion/uptake with the constants given to set_agents (they
        already carry dt, as Basic_Agent's cached constants do)."""
        self._check(lib().bfa_secretion_uptake(self._h), "simulate_cell_sources_and_sinks")

    def step(self, dt, nsteps=1):
        self._check(lib().bfa_step(self._h, float(dt), int(nsteps)), "step")

    def apply_dirichlet_boundaries_conditions(self):
        self._check(lib().bfa_apply_dirichlet(self._h), "apply_dirichlet_boundaries_conditions")

    def total_mass(self):
        out = np.zeros(self.ns)
        self._check(lib().bfa_total_mass(self._h, _d(out)), "total_mass")
        return out

    def synchronize(self):
        self._check(lib().bfa_synchronize(self._h), "synchronize")

    def time_kernels(self, dt, reps=10):
        ms = np.zeros(5)
        self._check(lib().bfa_time_kernels(self._h, float(dt), int(reps), _d(ms)), "time_kernels")
        return ms


def member_bytes(nx, ny, nz, ns, precision="fp32"):
    """Device bytes a batch member of this shape occupies."""
    return int(lib().bfa_batch_member_bytes(int(nx), int(ny), int(nz), int(ns), _PREC[precision]))


class Batch:
    """Many independent microenvironments advanced together (one persistent
    kernel per sweep over all members, one HIP graph per run of steps).

    Members are added with the same arguments as a ``Microenvironment`` plus
    its coefficients and Dirichlet state; each member's result equals what a
    ``Microenvironment`` of the same state computes.
    """

    def __init__(self, precision="fp32"):
        self._h = None
        L = lib()
        h = ctypes.c_void_p()
        rc = L.bfa_batch_create(ctypes.byref(h), _PREC[precision])
        if rc:
            raise BioFVMError(f"bfa_batch_create: {L.bfa_last_error(None).decode()}")
        self._h = h
        self.precision = _PREC[precision]
        self.members = {}
        self._n_agents = {}  # agent count of each member's last set_agents call

    def close(self):
        if self._h is not None:
            lib().bfa_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc:
            msg = lib().bfa_batch_last_error(self._h).decode()
            raise BioFVMError(f"{what}: status {rc}: {msg}")

    def add(self, nx, ny, nz, ns, dx, D, decay, active=None, face_registered=None,
            face_values=None):
        """Add a member; returns its id."""
        if active is None:
            active = np.zeros(ns, np.uint8)
        if face_registered is None:
            face_registered = np.zeros(6, np.uint8)
        if face_values is None:
            face_values = np.ones((6, ns))
        act = np.ascontiguousarray(active, np.uint8).reshape(ns)
        reg = np.ascontiguousarray(face_registered, np.uint8).reshape(6)
        fv = _f64(face_values, 6 * ns)
        mid = ctypes.c_int()
        self._check(lib().bfa_batch_add(self._h, int(nx), int(ny), int(nz), int(ns), float(dx),
                                        _d(_f64(D, ns)), _d(_f64(decay, ns)),
                                        act.ctypes.data_as(_u8), reg.ctypes.data_as(_u8), _d(fv),
                                        ctypes.byref(mid)), "batch_add")
        self.members[mid.value] = (int(nx), int(ny), int(nz), int(ns))
        return mid.value

    def remove(self, mid):
        self._check(lib().bfa_batch_remove(self._h, int(mid)), "batch_remove")
        self.members.pop(mid, None)
        self._n_agents.pop(mid, None)

    def upload(self, mid, field, layout="ref"):
        nx, ny, nz, ns = self.members[mid]
        f = _f64(field, nx * ny * nz * ns)
        self._check(lib().bfa_batch_upload(self._h, int(mid), _d(f), _LAYOUT[layout]), "batch_upload")

    def download(self, mid, layout="ref"):
        nx, ny, nz, ns = self.members[mid]
        out = np.empty(nx * ny * nz * ns)
        self._check(lib().bfa_batch_download(self._h, int(mid), _d(out), _LAYOUT[layout]),
                    "batch_download")
        return out

    def set_agents(self, mid, voxel, c1, c2, e2, active=None):
        ns = self.members[mid][3]
        voxel = np.ascontiguousarray(voxel, np.int32)
        n = voxel.size
        act = _flags(active, n)
        self._check(lib().bfa_batch_set_agents(self._h, int(mid), n, voxel.ctypes.data_as(_i32),
                                               None if act is None else act.ctypes.data_as(_u8),
                                               _d(_f64(c1, n * ns)), _d(_f64(c2, n * ns)),
                                               _d(_f64(e2, n * ns))), "batch_set_agents")
        self._n_agents[mid] = n

    def track_internalized(self, mid, initial=None, e1=None):
        """Internalised-substrate tracking for member mid's current agents:
        initial totals [n][ns] (None: zeros), net export e1 [n][ns] (None:
        kept, else zeros)."""
        cnt = self._n_agents.get(mid, 0) * self.members[mid][3]
        ini = None if initial is None else _f64(initial, cnt)
        ee = None if e1 is None else _f64(e1, cnt)
        self._check(lib().bfa_batch_track_internalized(self._h, int(mid), 1, None if ini is None else _d(ini),
                                                       None if ee is None else _d(ee)), "batch_track_internalized")

    def stop_tracking_internalized(self, mid):
        self._check(lib().bfa_batch_track_internalized(self._h, int(mid), 0, None, None),
                    "batch_track_internalized")

    def internalized(self, mid):
        """Member mid's per-agent internalised totals [n][ns] (caller's agent order)."""
        out = np.empty((self._n_agents.get(mid, 0), self.members[mid][3]))
        self._check(lib().bfa_batch_get_internalized(self._h, int(mid), _d(out)), "batch_get_internalized")
        return out

    def fill(self, mid, values):
        """Set every voxel of substrate d to values[d]."""
        ns = self.members[mid][3]
        self._check(lib().bfa_batch_fill(self._h, int(mid), _d(_f64(values, ns))), "batch_fill")

    def set_member_steps(self, mid, steps):
        self._check(lib().bfa_batch_set_member_steps(self._h, int(mid), int(steps)),
                    "batch_set_member_steps")

    def step(self, dt, nsteps=1):
        self._check(lib().bfa_batch_step(self._h, float(dt), int(nsteps)), "batch_step")

    def synchronize(self):
        self._check(lib().bfa_batch_synchronize(self._h), "batch_synchronize")

    def total_mass(self, mid):
        out = np.zeros(self.members[mid][3])
        self._check(lib().bfa_batch_total_mass(self._h, int(mid), _d(out)), "batch_total_mass")
        return out

    def set_tiles(self, Lx=0, Ly=0, Lz=0, cols_y=0, cols_z=0):
        self._check(lib().bfa_batch_set_tiles(self._h, int(Lx), int(Ly), int(Lz), int(cols_y),
                                              int(cols_z)), "batch_set_tiles")

    def set_col_threads(self, threads):
        """Threads per y/z column block (multiple of 64; 0: automatic)."""
        self._check(lib().bfa_batch_set_col_threads(self._h, int(threads)), "batch_set_col_threads")

    def col_threads(self):
        out = (ctypes.c_int * 2)()
        self._check(lib().bfa_batch_get_col_threads(self._h, out), "batch_get_col_threads")
        return tuple(out)

    def set_streams(self, groups):
        """1: one kernel per sweep for all members; 2..4: member groups swept on
        that many streams (same results)."""
        self._check(lib().bfa_batch_set_streams(self._h, int(groups)), "batch_set_streams")

    def streams(self):
        return lib().bfa_batch_get_streams(self._h)

    def tiles(self):
        v = (ctypes.c_int * 5)()
        self._check(lib().bfa_batch_get_tiles(self._h, v), "batch_get_tiles")
        return tuple(v)

    def set_fusion(self, enabled=True):
        """Fused x + y sweeps of k-planes in LDS for the members whose plane
        fits; off (the default): every member takes the separate sweeps."""
        self._check(lib().bfa_batch_set_fusion(self._h, int(bool(enabled))), "batch_set_fusion")

    def fused_members(self):
        return int(lib().bfa_batch_fused_members(self._h))
