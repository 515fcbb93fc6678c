"""lpmcxx_amd -- MI355X-native spherical Biot-Savart / Poisson sum.

Python access to the two native libraries built in-tree:

* ``lib/liblpm_gpu.so`` -- HIP kernels for gfx950 behind the C ABI in
  ``include/lpm_gpu.h`` (direct sum, octree, treecode, RK4, order model).
* ``lib/liblpm.so`` -- the C++ host library (namespace ``Lpm``, the
  reference's API) and the mesh builder C ABI in ``include/lpm_mesh.h``.

There is no CPU fallback: every compute entry point goes through the HIP
library and raises if it is missing or fails.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
LIBDIR = ROOT / "lib"

LPM_K_BIOT_SAVART_SPHERE = 1
LPM_K_GREENS_SPHERE = 2
LPM_F64 = 0
LPM_F32C = 1
LPM_F64F = 2
LPM_SEED_ICOS_TRI = 0
LPM_SERIES_REFERENCE = 0
LPM_SERIES_CARTESIAN = 1
LPM_SEED_CUBED = 1

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_vp = C.c_void_p


class LpmError(RuntimeError):
    """Raised when a native entry point returns a non-zero lpm_status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"lpm status {status}: {message}")
        self.status = status


_gpu = None
_host = None


def _load(name: str) -> C.CDLL:
    path = LIBDIR / name
    if not path.exists():
        raise ImportError(f"{path} is not built; run __graft_entry__.build() or python -m lpmcxx_amd.build")
    return C.CDLL(str(path), mode=C.RTLD_GLOBAL)


def gpu_lib() -> C.CDLL:
    """The HIP library (loaded once)."""
    global _gpu
    if _gpu is None:
        lib = _load("liblpm_gpu.so")
        lib.lpm_create.argtypes = [C.POINTER(_vp), C.c_int]
        lib.lpm_destroy.argtypes = [_vp]
        lib.lpm_destroy.restype = None
        lib.lpm_status_string.argtypes = [C.c_int]
        lib.lpm_status_string.restype = C.c_char_p
        lib.lpm_last_error.argtypes = [_vp]
        lib.lpm_last_error.restype = C.c_char_p
        lib.lpm_synchronize.argtypes = [_vp]
        lib.lpm_stream.argtypes = [_vp]
        lib.lpm_stream.restype = _vp
        ds_args = [_vp, C.c_int, C.c_int, C.c_double, C.c_double, _vp, _vp, _vp, C.c_int64,
                   _vp, _vp, _vp, _vp, C.c_int64, _vp, _vp, _vp, _vp, _vp]
        lib.lpm_direct_sum.argtypes = ds_args
        lib.lpm_direct_sum_dev.argtypes = ds_args
        if hasattr(lib, "lpm_direct_sum_dev_noalloc"):
            lib.lpm_direct_sum_dev_noalloc.argtypes = ds_args
        _declare_extra(lib)
        _gpu = lib
    return _gpu


class SolverConfig(C.Structure):
    """lpm_solver_config (include/lpm_gpu.h): velocity solver of the RK4 step."""
    _fields_ = [("method", C.c_int), ("prec", C.c_int), ("series", C.c_int), ("order", C.c_int),
                ("leaf_max", C.c_int), ("theta", C.c_double), ("delta", C.c_double),
                ("root_box_min", C.c_double * 3), ("root_box_max", C.c_double * 3), ("margin", C.c_int),
                ("u_ref", C.c_double)]

    @classmethod
    def default(cls, **kw) -> "SolverConfig":
        cfg = cls()
        gpu_lib().lpm_solver_config_default(C.byref(cfg))
        for k, v in kw.items():
            setattr(cfg, k, v)
        return cfg


LPM_OPT_BATCH_CELL_TARGETS = 1
LPM_OPT_STEREO, LPM_OPT_STEREO_EXTRA_ORDER, LPM_OPT_STEREO_RATIO = 2, 3, 4
LPM_METHOD_DIRECT = 0
LPM_METHOD_TREE = 1
LPM_METHOD_TREE_BATCH = 2
LPM_METHOD_TREE_LEARNED = 3


def _declare_extra(lib: C.CDLL) -> None:
    """Signatures of the tree / RK4 / order-model entry points (when present)."""
    sigs = {
        "lpm_tree_build": [_vp, _vp, _vp, _vp, C.c_int64, C.c_double, C.c_int, _i64p, C.POINTER(C.c_int)],
        "lpm_tree_build_dev": [_vp, _vp, _vp, _vp, C.c_int64, C.c_double, C.c_int, _i64p, C.POINTER(C.c_int)],
        "lpm_tree_moments": [_vp, _vp, C.c_int, C.c_int],
        "lpm_tree_moments_psi": [_vp, _vp, C.c_int],
        "lpm_tree_moments_psi_dev": [_vp, _vp, C.c_int],
        "lpm_tree_eval_psi": [_vp, C.c_int, C.c_int, C.c_double, _vp, _vp, _vp, C.c_int64, _vp],
        "lpm_tree_eval_psi_dev": [_vp, C.c_int, C.c_int, C.c_double, _vp, _vp, _vp, C.c_int64, _vp],
        "lpm_tree_moments_dev": [_vp, _vp, C.c_int, C.c_int],
        "lpm_tree_eval": [_vp, C.c_int, C.c_int, C.c_double, C.c_double, _vp, _vp, _vp, C.c_int64, _vp, _vp,
                          _vp, _i64p],
        "lpm_tree_eval_dev": [_vp, C.c_int, C.c_int, C.c_double, C.c_double, _vp, _vp, _vp, C.c_int64, _vp, _vp,
                              _vp, _i64p],
        "lpm_tree_info": [_vp, _i64p, C.POINTER(C.c_int), _i64p],
        "lpm_tree_build_box": [_vp, _vp, _vp, _vp, C.c_int64, _vp, _vp, C.c_int, C.c_int, _i64p,
                               C.POINTER(C.c_int)],
        "lpm_tree_export": [_vp] * 9,
        "lpm_tree_export_moments": [_vp, _vp],
        "lpm_tree_last_stats": [_vp, _i64p, _i64p],
        "lpm_tree_stats": [_vp, _vp],
        "lpm_last_kernel_ms": [_vp, _vp],
        "lpm_set_option": [_vp, C.c_int, C.c_double],
        "lpm_tree_eval_batch": [_vp, C.c_int, C.c_int, C.c_double, C.c_double, _vp, _vp, _vp, C.c_int64, _vp,
                                _vp, _vp],
        "lpm_tree_eval_batch_dev": [_vp, C.c_int, C.c_int, C.c_double, C.c_double, _vp, _vp, _vp, C.c_int64,
                                    _vp, _vp, _vp],
        "lpm_tree_eval_learned": [_vp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double, _vp, _vp,
                                  _vp, C.c_int64, _vp, _vp, _vp],
        "lpm_tree_eval_learned_dev": [_vp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double, _vp,
                                      _vp, _vp, C.c_int64, _vp, _vp, _vp],
        "lpm_tree_labels": [_vp, C.c_double, C.c_double, C.c_double, C.c_int64, C.c_double, C.c_int, C.c_uint64,
                            C.c_int64, _vp, _vp, _vp, C.c_int64, _vp, _vp, _vp, _i64p],
        "lpm_model_init": [_vp, C.c_uint64],
        "lpm_model_set_params": [_vp, _vp, _vp, _vp],
        "lpm_model_get_params": [_vp, _vp, _vp, _vp],
        "lpm_model_set_range": [_vp, _vp, _vp],
        "lpm_model_get_range": [_vp, _vp, _vp],
        "lpm_model_gradient": [_vp, _vp, _vp, C.c_int64, _vp, _vp],
        "lpm_model_train": [_vp, _vp, _vp, C.c_int64, C.c_int, C.c_int, C.c_float, C.c_uint64, _vp],
        "lpm_model_forward": [_vp, _vp, C.c_int64, _vp, _vp],
        "lpm_solver_config_default": [_vp],
        "lpm_rk4_step_dev": [_vp, _vp, C.c_double, C.c_double, C.c_int64, _vp, _vp, _vp, _vp, _vp],
        "lpm_rk4_steps": [_vp, _vp, C.c_double, C.c_double, C.c_int, C.c_int64, _vp, _vp, _vp, _vp, _vp],
        "lpm_bve_step_dev": [_vp, _vp, C.c_double, C.c_double, C.c_int64, C.c_int64, _vp, _vp, _vp, _vp, _vp],
        "lpm_bve_steps": [_vp, _vp, C.c_double, C.c_double, C.c_int, C.c_int64, C.c_int64, _vp, _vp, _vp, _vp,
                          _vp],
        "lpm_series_coeffs": [_vp, C.c_int, C.c_int64] + [_vp] * 7,
        "lpm_series_moments": [_vp, C.c_int, C.c_int64] + [_vp] * 5 + [C.c_int, C.c_double, C.c_double,
                                                                      C.c_double, _vp],
        "lpm_direct_sum_dev_noalloc": None,
        "lpm_mesh_counts": [C.c_int, C.c_int, _i64p, _i64p, _i64p],
        "lpm_mesh_generate": [_vp, C.c_int, C.c_int, C.c_double] + [_vp] * 8,
        "lpm_mesh_generate_dev": [_vp, C.c_int, C.c_int, C.c_double] + [_vp] * 8,
    }
    for name, args in sigs.items():
        if args is not None and hasattr(lib, name):
            getattr(lib, name).argtypes = args
    if hasattr(lib, "lpm_solver_config_default"):
        lib.lpm_solver_config_default.restype = None


def host_lib() -> C.CDLL:
    """The C++ host library with the mesh C ABI (loaded once)."""
    global _host
    if _host is None:
        gpu_lib()
        lib = _load("liblpm.so")
        lib.lpm_mesh_create.argtypes = [C.POINTER(_vp), C.c_int, C.c_int, C.c_double]
        lib.lpm_mesh_destroy.argtypes = [_vp]
        lib.lpm_mesh_destroy.restype = None
        lib.lpm_mesh_sizes.argtypes = [_vp, _i64p, _i64p, _i64p]
        lib.lpm_mesh_vertices.argtypes = [_vp, _vp, _vp, _vp]
        lib.lpm_mesh_leaf_faces.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
        lib.lpm_mesh_faces.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
        lib.lpm_mesh_max_radius.argtypes = [_vp]
        lib.lpm_mesh_max_radius.restype = C.c_double
        _host = lib
    return _host


def _ptr(a) -> int | None:
    """Address of a contiguous numpy array or torch tensor (None for None)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags.c_contiguous:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return a.data_ptr()
    raise TypeError(f"unsupported buffer type {type(a)}")


def _as_host(arrays: dict, n: int | None = None, dtype=np.float64) -> list:
    """C-contiguous host copies (numpy, dtype) of named arrays, all of length n
    (or of the first array's length); ValueError before any native call
    otherwise, so the copies to the device never over-read a short array."""
    out = []
    for name, a in arrays.items():
        if a is None:
            out.append(None)
            continue
        if hasattr(a, "device") and getattr(a.device, "type", "cpu") != "cpu":
            raise ValueError(f"{name}: host array expected, got a tensor on {a.device}")
        v = np.ascontiguousarray(a, dtype=dtype).reshape(-1)
        if n is None:
            n = v.size
        if v.size != n:
            raise ValueError(f"{name} has {v.size} entries, expected {n}")
        out.append(v)
    return out


def _numel(t, name: str) -> int:
    """Entry count of a torch tensor (TypeError for anything else)."""
    if not hasattr(t, "numel") or not hasattr(t, "data_ptr"):
        raise TypeError(f"{name}: torch tensor expected, got {type(t).__name__}")
    return t.numel()


def _dev(ctx, tensors: dict, n: int, dtype: str = "float64", optional: tuple = ()) -> None:
    """Checks device tensors before a *_dev call: each a contiguous torch tensor
    of the given dtype with n entries on the context's GPU (names listed in
    `optional` may be None).  Raises ValueError / TypeError otherwise: the
    kernels take n from one argument and would read or write the others out
    of bounds."""
    import torch
    want = {"float64": torch.float64, "int64": torch.int64}[dtype]
    for name, t in tensors.items():
        if t is None:
            if name in optional:
                continue
            raise ValueError(f"{name} is required")
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name}: torch tensor expected, got {type(t).__name__}")
        if t.dtype != want:
            raise ValueError(f"{name}: dtype {t.dtype}, expected {want}")
        if t.device.type != "cuda" or (t.device.index if t.device.index is not None else 0) != ctx.device:
            raise ValueError(f"{name}: on {t.device}, expected cuda:{ctx.device}")
        if t.numel() != n:
            raise ValueError(f"{name} has {t.numel()} entries, expected {n}")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")


# ---------------------------------------------------------------------- mesh
class Mesh:
    """Vertices and leaf faces of a refined spherical mesh (reference index order)."""

    def __init__(self, seed: int, level: int, radius: float = 1.0):
        lib = host_lib()
        h = _vp()
        st = lib.lpm_mesh_create(C.byref(h), seed, level, radius)
        if st != 0:
            raise LpmError(st, f"mesh seed={seed} level={level}")
        try:
            nv, nf, nl = C.c_int64(), C.c_int64(), C.c_int64()
            lib.lpm_mesh_sizes(h, C.byref(nv), C.byref(nf), C.byref(nl))
            self.n_vertices, self.n_faces, self.n_leaves = nv.value, nf.value, nl.value
            self.vx = np.empty(self.n_vertices)
            self.vy = np.empty(self.n_vertices)
            self.vz = np.empty(self.n_vertices)
            lib.lpm_mesh_vertices(h, _ptr(self.vx), _ptr(self.vy), _ptr(self.vz))
            self.cx = np.empty(self.n_leaves)
            self.cy = np.empty(self.n_leaves)
            self.cz = np.empty(self.n_leaves)
            self.area = np.empty(self.n_leaves)
            self.face_index = np.empty(self.n_leaves, dtype=np.int64)
            lib.lpm_mesh_leaf_faces(h, _ptr(self.cx), _ptr(self.cy), _ptr(self.cz), _ptr(self.area),
                                    _ptr(self.face_index))
            self.max_radius = lib.lpm_mesh_max_radius(h)
        finally:
            lib.lpm_mesh_destroy(h)
        self.seed, self.level, self.radius = seed, level, radius

    def targets(self):
        """Targets of meshSolve: all vertices, then all leaf faces."""
        return (np.concatenate([self.vx, self.cx]), np.concatenate([self.vy, self.cy]),
                np.concatenate([self.vz, self.cz]))

    def face_self_skip(self) -> np.ndarray:
        """Per-target source index to exclude (leaf face i is source i); -1 for vertices."""
        return np.concatenate([np.full(self.n_vertices, -1, np.int64), np.arange(self.n_leaves, dtype=np.int64)])


# ---------------------------------------------------------------------- GPU context
class Context:
    """One lpm_ctx: a HIP stream, its workspace and at most one source octree."""

    def __init__(self, device: int = 0):
        self.lib = gpu_lib()
        h = _vp()
        st = self.lib.lpm_create(C.byref(h), device)
        if st != 0:
            raise LpmError(st, self.lib.lpm_status_string(st).decode())
        self.h = h
        self.device = device

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.lpm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, st: int) -> None:
        if st != 0:
            raise LpmError(st, f"{self.lib.lpm_status_string(st).decode()}: {self.lib.lpm_last_error(self.h).decode()}")

    @property
    def stream(self) -> int:
        return self.lib.lpm_stream(self.h)

    def set_option(self, option: int, value: float) -> None:
        self.check(self.lib.lpm_set_option(self.h, option, value))

    def last_kernel_ms(self) -> float:
        """Duration of the last pair / evaluation kernel (HIP events on the context stream)."""
        ms = C.c_float()
        self.check(self.lib.lpm_last_kernel_ms(self.h, C.byref(ms)))
        return ms.value

    def synchronize(self) -> None:
        self.check(self.lib.lpm_synchronize(self.h))

    def after_torch(self) -> None:
        """Make the context stream wait for the work enqueued so far on torch's
        current stream.  The context stream does not synchronize with torch's,
        so device tensors just written by torch (a clone, a copy, an
        elementwise op) would otherwise be read before they are complete.
        Every *_dev wrapper of this package calls it first; it enqueues an
        event wait and does not block the host."""
        import torch
        ext = getattr(self, "_torch_ext", None)
        if ext is None:
            ext = torch.cuda.ExternalStream(self.stream, device=torch.device("cuda", self.device))
            self._torch_ext = ext
        ext.wait_stream(torch.cuda.current_stream(self.device))

    def direct_sum(self, targets, sources, gamma, *, kernels=LPM_K_BIOT_SAVART_SPHERE, prec=LPM_F64,
                   radius=1.0, delta=0.0, skip=None):
        """Host-array direct sum; returns (ux, uy, uz) and/or psi as float64 arrays."""
        tx, ty, tz = _as_host({"tx": targets[0], "ty": targets[1], "tz": targets[2]})
        sx, sy, sz, g = _as_host({"sx": sources[0], "sy": sources[1], "sz": sources[2], "gamma": gamma})
        nt, ns = tx.size, sx.size
        (sk,) = _as_host({"skip": skip}, nt, np.int64)
        vel = bool(kernels & LPM_K_BIOT_SAVART_SPHERE)
        pot = bool(kernels & LPM_K_GREENS_SPHERE)
        ux = np.zeros(nt) if vel else None
        uy = np.zeros(nt) if vel else None
        uz = np.zeros(nt) if vel else None
        psi = np.zeros(nt) if pot else None
        self.check(self.lib.lpm_direct_sum(self.h, kernels, prec, radius, delta, _ptr(tx), _ptr(ty), _ptr(tz), nt,
                                           _ptr(sx), _ptr(sy), _ptr(sz), _ptr(g), ns, _ptr(sk),
                                           _ptr(ux), _ptr(uy), _ptr(uz), _ptr(psi)))
        if vel and pot:
            return (ux, uy, uz), psi
        return (ux, uy, uz) if vel else psi

    def direct_sum_dev(self, tx, ty, tz, sx, sy, sz, gamma, out, *, kernels=LPM_K_BIOT_SAVART_SPHERE,
                       prec=LPM_F64, radius=1.0, delta=0.0, skip=None, psi=None):
        """Device-tensor direct sum enqueued on the context stream (no synchronize)."""
        ux, uy, uz = out if out is not None else (None, None, None)
        nt, ns = _numel(tx, "tx"), _numel(sx, "sx")
        vel, pot = bool(kernels & LPM_K_BIOT_SAVART_SPHERE), bool(kernels & LPM_K_GREENS_SPHERE)
        _dev(self, {"tx": tx, "ty": ty, "tz": tz}, nt)
        _dev(self, {"sx": sx, "sy": sy, "sz": sz, "gamma": gamma}, ns)
        _dev(self, {"ux": ux, "uy": uy, "uz": uz}, nt, optional=() if vel else ("ux", "uy", "uz"))
        _dev(self, {"psi": psi}, nt, optional=() if pot else ("psi",))
        _dev(self, {"skip": skip}, nt, "int64", optional=("skip",))
        self.after_torch()
        self.check(self.lib.lpm_direct_sum_dev(self.h, kernels, prec, radius, delta, _ptr(tx), _ptr(ty), _ptr(tz),
                                               tx.numel(), _ptr(sx), _ptr(sy), _ptr(sz), _ptr(gamma), sx.numel(),
                                               _ptr(skip), _ptr(ux), _ptr(uy), _ptr(uz), _ptr(psi)))


def _rk4_steps(self, cfg, dt, omega, nsteps, x, y, z, zeta_abs, area):
    """nsteps RK4 steps on host arrays; returns new (x, y, z) copies."""
    x, y, z, za, ar = _as_host({"x": x, "y": y, "z": z, "zeta_abs": zeta_abs, "area": area})
    x, y, z = x.copy(), y.copy(), z.copy()
    if cfg.method != LPM_METHOD_DIRECT:
        self.tree_sources = None
    self.check(self.lib.lpm_rk4_steps(self.h, C.byref(cfg), dt, omega, nsteps, x.size, _ptr(x), _ptr(y), _ptr(z),
                                      _ptr(za), _ptr(ar)))
    return x, y, z


def _rk4_step_dev(self, cfg, dt, omega, x, y, z, zeta_abs, area):
    """One RK4 step on device tensors, in place, enqueued on the context stream."""
    _dev(self, {"x": x, "y": y, "z": z, "zeta_abs": zeta_abs, "area": area}, _numel(x, "x"))
    if cfg.method != LPM_METHOD_DIRECT:
        self.tree_sources = None  # the step rebuilds the context's tree
    self.after_torch()
    self.check(self.lib.lpm_rk4_step_dev(self.h, C.byref(cfg), dt, omega, x.numel(), _ptr(x), _ptr(y), _ptr(z),
                                         _ptr(zeta_abs), _ptr(area)))


def _bve_steps(self, cfg, dt, omega, nsteps, src, zeta_abs, area, tracers=None):
    """nsteps barotropic-vorticity RK4 steps of vortex particles src = (x, y, z)
    (with zeta_abs, area) and passive tracers = (x, y, z) on host arrays;
    returns the new (src, tracers) positions."""
    sx, sy, sz, za, ar = _as_host({"src x": src[0], "src y": src[1], "src z": src[2], "zeta_abs": zeta_abs,
                                "area": area})
    ns = sx.size
    if tracers is not None:
        qx, qy, qz = _as_host({"tracer x": tracers[0], "tracer y": tracers[1], "tracer z": tracers[2]})
        nt = qx.size
    else:
        nt = 0
    if nt:
        x, y, z = np.concatenate([sx, qx]), np.concatenate([sy, qy]), np.concatenate([sz, qz])
    else:
        x, y, z = sx.copy(), sy.copy(), sz.copy()
    if cfg.method != LPM_METHOD_DIRECT:
        self.tree_sources = None
    self.check(self.lib.lpm_bve_steps(self.h, C.byref(cfg), dt, omega, nsteps, ns, nt, _ptr(x), _ptr(y), _ptr(z),
                                      _ptr(za), _ptr(ar)))
    return (x[:ns], y[:ns], z[:ns]), ((x[ns:], y[ns:], z[ns:]) if nt else None)


def _bve_step_dev(self, cfg, dt, omega, n_src, x, y, z, zeta_abs, area):
    """One BVE step on device tensors (first n_src entries are vortex particles,
    the rest tracers), in place, enqueued on the context stream."""
    n = _numel(x, "x")
    if not 0 <= n_src <= n:
        raise ValueError(f"n_src = {n_src} outside [0, {n}]")
    _dev(self, {"x": x, "y": y, "z": z}, n)
    _dev(self, {"zeta_abs": zeta_abs, "area": area}, n_src)
    if cfg.method != LPM_METHOD_DIRECT:
        self.tree_sources = None
    self.after_torch()
    self.check(self.lib.lpm_bve_step_dev(self.h, C.byref(cfg), dt, omega, n_src, x.numel() - n_src, _ptr(x), _ptr(y),
                                         _ptr(z), _ptr(zeta_abs), _ptr(area)))


def mesh_counts(seed: int, level: int) -> tuple[int, int, int]:
    """(vertices, faces with parents, leaf faces) of seed refined to level."""
    nv, nf, nl = C.c_int64(), C.c_int64(), C.c_int64()
    st = gpu_lib().lpm_mesh_counts(seed, level, C.byref(nv), C.byref(nf), C.byref(nl))
    if st != 0:
        raise LpmError(st, f"mesh seed={seed} level={level}")
    return nv.value, nf.value, nl.value


class GpuMesh(Mesh):
    """The particle sets of Mesh(seed, level, radius) generated on the GPU
    (lpm_mesh_generate): same attributes, same order; leaf_vertices (n_leaves
    x 3 or 4 corner indices) when asked for."""

    def __init__(self, ctx: "Context", seed: int, level: int, radius: float = 1.0, leaf_vertices: bool = False):
        self.n_vertices, self.n_faces, self.n_leaves = mesh_counts(seed, level)
        nv, nl, nepf = self.n_vertices, self.n_leaves, 4 if seed == LPM_SEED_CUBED else 3
        self.vx, self.vy, self.vz = np.empty(nv), np.empty(nv), np.empty(nv)
        self.cx, self.cy, self.cz, self.area = np.empty(nl), np.empty(nl), np.empty(nl), np.empty(nl)
        self.leaf_vertices = np.empty((nl, nepf), np.int64) if leaf_vertices else None
        ctx.check(ctx.lib.lpm_mesh_generate(ctx.h, seed, level, radius, _ptr(self.vx), _ptr(self.vy), _ptr(self.vz),
                                            _ptr(self.cx), _ptr(self.cy), _ptr(self.cz), _ptr(self.area),
                                            _ptr(self.leaf_vertices)))
        # leaves are the faces of the last level, the last n_leaves face indices
        self.face_index = np.arange(self.n_faces - nl, self.n_faces, dtype=np.int64)
        self.max_radius = float(np.sqrt(np.max(self.vx * self.vx + self.vy * self.vy + self.vz * self.vz)))
        self.seed, self.level, self.radius = seed, level, radius


def _mesh_generate_dev(self, seed, level, radius=1.0, leaf_vertices=False, out=None):
    """Mesh generated into device tensors, enqueued on the context stream:
    dict of vx, vy, vz, cx, cy, cz, area (float64) and leaf_vertices (int64,
    or None); `out` (a dict returned by an earlier call of the same seed and
    level) is filled again instead of allocating."""
    import torch
    nv, _, nl = mesh_counts(seed, level)
    nepf = 4 if seed == LPM_SEED_CUBED else 3
    if out is None:
        dev = torch.device("cuda", self.device)
        out = {k: torch.empty(nv, dtype=torch.float64, device=dev) for k in ("vx", "vy", "vz")}
        out.update({k: torch.empty(nl, dtype=torch.float64, device=dev) for k in ("cx", "cy", "cz", "area")})
        out["leaf_vertices"] = torch.empty((nl, nepf), dtype=torch.int64, device=dev) if leaf_vertices else None
    else:
        sizes = {"vx": nv, "vy": nv, "vz": nv, "cx": nl, "cy": nl, "cz": nl, "area": nl}
        if any(out[k].numel() != n or out[k].dtype != torch.float64 for k, n in sizes.items()):
            raise ValueError("out does not match this seed and level")
        lv = out.get("leaf_vertices")
        if lv is not None and (lv.numel() != nl * nepf or lv.dtype != torch.int64):
            raise ValueError("out['leaf_vertices'] does not match this seed and level")
    self.after_torch()
    self.check(self.lib.lpm_mesh_generate_dev(self.h, seed, level, radius, *(_ptr(out[k]) for k in
                                              ("vx", "vy", "vz", "cx", "cy", "cz", "area", "leaf_vertices"))))
    return out


Context.rk4_steps = _rk4_steps
Context.rk4_step_dev = _rk4_step_dev
Context.bve_steps = _bve_steps
Context.bve_step_dev = _bve_step_dev
Context.mesh_generate_dev = _mesh_generate_dev


def native_libraries() -> list[str]:
    """Paths of the native libraries this process has loaded."""
    out = []
    for lib in (_gpu, _host):
        if lib is not None:
            out.append(lib._name)
    return out


# ---------------------------------------------------------------------- analytic fields
def sphere_harmonic54(x, y, z):
    """zeta = 30 cos(4 lambda) z (z^2-1)^2 (src/LpmAnalyticFunctions.h:87-103)."""
    lon = np.mod(np.arctan2(y, x), 2 * np.pi)
    return 30.0 * np.cos(4.0 * lon) * z * (z * z - 1.0) * (z * z - 1.0)


def rh54_velocity(x, y, z):
    """Exact velocity for sphere_harmonic54 (src/LpmAnalyticFunctions.h:105-120)."""
    lat = np.arctan2(z, np.sqrt(x * x + y * y))
    lon = np.mod(np.arctan2(y, x), 2 * np.pi)
    u = 0.5 * np.cos(4 * lon) * np.cos(lat) ** 3 * (5 * np.cos(2 * lat) - 3)
    v = 4 * np.cos(lat) ** 3 * np.sin(lat) * np.sin(4 * lon)
    return (-u * np.sin(lon) - v * np.sin(lat) * np.cos(lon), u * np.cos(lon) - v * np.sin(lat) * np.sin(lon),
            v * np.cos(lat))


def gaussian_vortex(x, y, z, area, lat0_deg=60.0, lon0_deg=0.0, beta=16.0, amp=4 * np.pi):
    """zeta = amp exp(-beta |x - x0|^2) minus its area-weighted mean."""
    la, lo = np.radians(lat0_deg), np.radians(lon0_deg)
    x0 = np.cos(la) * np.cos(lo), np.cos(la) * np.sin(lo), np.sin(la)
    z_ = amp * np.exp(-beta * ((x - x0[0]) ** 2 + (y - x0[1]) ** 2 + (z - x0[2]) ** 2))
    return z_ - np.sum(z_ * area) / np.sum(area)


def random_vortices(x, y, z, area, n=8, seed=1234):
    """Sum of n Gaussian vortices (centers uniform on the sphere, amplitude U(-2,2),
    width U(4,16)) minus the area-weighted mean; numpy PCG64 seeded with `seed`."""
    rng = np.random.default_rng(seed)
    zeta = np.zeros_like(x)
    for _ in range(n):
        zc = rng.uniform(-1.0, 1.0)
        lam = rng.uniform(0.0, 2 * np.pi)
        a = rng.uniform(-2.0, 2.0)
        b = rng.uniform(4.0, 16.0)
        r = np.sqrt(1 - zc * zc)
        c = (r * np.cos(lam), r * np.sin(lam), zc)
        zeta += a * np.exp(-b * ((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2))
    return zeta - np.sum(zeta * area) / np.sum(area)
