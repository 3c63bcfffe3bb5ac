"""ctypes binding of include/fd3d_tsn.h (lib/libfd3d_tsn.so).

There is no CPU fallback: if the HIP library is missing or no GPU is
visible, :func:`load` raises.  The library is built in-tree by
``__graft_entry__.build()`` (``make -C fd3d_tsn_amd/csrc``).
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from . import cases as C

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("FD3D_TSN_LIB", os.path.join(HERE, "lib", "libfd3d_tsn.so"))
HEADER = os.path.join(os.path.dirname(HERE), "include", "fd3d_tsn.h")

FD3D_SW, FD3D_FVW = 0, 1
NOUT = 5
OUT_NAMES = ("sliprateX", "sliprateZ", "shearstressX", "shearstressZ", "psi")
FIELDS = ("u", "v", "w", "xx", "yy", "zz", "xy", "yz", "xz")
FINAL_NAMES = ("ruptime", "rise", "sliptime", "slipX", "slipZ", "stressdropX", "stressdropZ")
ERRORS = {-1: "FD3D_E_INVALID", -2: "FD3D_E_CFL", -3: "FD3D_E_NOMEM", -4: "FD3D_E_HIP",
          -5: "FD3D_E_STATE", -6: "FD3D_E_NEWTON", -7: "FD3D_E_NONFINITE"}


class FD3DError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [("nxt", ctypes.c_int), ("nyt", ctypes.c_int), ("nzt", ctypes.c_int),
                ("nabc", ctypes.c_int), ("fspace", ctypes.c_int), ("friction", ctypes.c_int),
                ("nucleation", ctypes.c_int), ("nstations", ctypes.c_int),
                ("stations", ctypes.POINTER(ctypes.c_int)),
                ("dh", ctypes.c_float), ("dt", ctypes.c_float), ("damp_s", ctypes.c_float),
                ("pml_vp", ctypes.c_float), ("pml_fact", ctypes.c_float),
                ("Sn", ctypes.c_double), ("uini", ctypes.c_double), ("wini", ctypes.c_double),
                ("f0", ctypes.c_double), ("v0", ctypes.c_double), ("fw", ctypes.c_double),
                ("hx0", ctypes.c_double), ("hz0", ctypes.c_double), ("perturb", ctypes.c_double),
                ("RR2", ctypes.c_double), ("TT2", ctypes.c_double), ("strinixI", ctypes.c_double),
                ("newton_max", ctypes.c_int), ("ring_steps", ctypes.c_int)]


_lib = None
_FP = ctypes.POINTER(ctypes.c_float)
_DP = ctypes.POINTER(ctypes.c_double)


def exported_symbols() -> Tuple[str, ...]:
    """Function names declared in include/fd3d_tsn.h."""
    import re
    text = open(HEADER).read()
    pat = r"^(?:int|void|size_t|double|const\s+char\s*\*)\s*\**\s*(fd3d_[a-z_0-9]+)\s*\("
    return tuple(sorted(set(re.findall(pat, text, flags=re.M))))


def load(path: Optional[str] = None):
    """Load the HIP library; raises if it is not built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"HIP library {p} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(p)
    H = ctypes.c_void_p
    L.fd3d_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(H)]
    L.fd3d_footprint.restype = ctypes.c_size_t
    L.fd3d_footprint.argtypes = [ctypes.POINTER(Config)]
    L.fd3d_upload_medium.argtypes = [H, _FP, _FP, _FP]
    L.fd3d_upload_fault_sw.argtypes = [H] + [_FP] * 6
    L.fd3d_upload_fault_fvw.argtypes = [H] + [_DP] * 11
    L.fd3d_init_state.argtypes = [H]
    L.fd3d_step.argtypes = [H, ctypes.c_int, _FP, _FP]
    L.fd3d_sync.argtypes = [H]
    L.fd3d_step_begin.argtypes = [H, ctypes.c_int]
    L.fd3d_step_wait.argtypes = [H, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(_FP), ctypes.POINTER(_FP)]
    L.fd3d_steps_done.argtypes = [H]
    L.fd3d_download_final.argtypes = [H, _FP]
    L.fd3d_download_field.argtypes = [H, ctypes.c_int, _FP]
    L.fd3d_upload_field.argtypes = [H, ctypes.c_int, _FP]
    L.fd3d_set_profiling.argtypes = [H, ctypes.c_int]
    L.fd3d_set_tuning.argtypes = [H] + [ctypes.c_int] * 5
    L.fd3d_get_tuning.argtypes = [H, ctypes.POINTER(ctypes.c_int)]
    L.fd3d_kernel_times.argtypes = [H, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), _DP,
                                    ctypes.POINTER(ctypes.c_long)]
    L.fd3d_step_bytes.restype = ctypes.c_double
    L.fd3d_step_bytes.argtypes = [H, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]
    L.fd3d_last_error.restype = ctypes.c_char_p
    L.fd3d_last_error.argtypes = [H]
    L.fd3d_destroy.argtypes = [H]
    L.fd3d_read_outputs.argtypes = [H, ctypes.c_int, _FP, _FP]
    L.fd3d_set_source_bins.argtypes = [H, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_int]
    L.fd3d_download_source_bins.argtypes = [H, _FP, _FP, _FP]
    L.fd3d_device_memory.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    B = ctypes.c_void_p
    L.fd3d_batch_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(B)]
    L.fd3d_batch_create_static.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(B)]
    L.fd3d_batch_add.argtypes = [B, H]
    L.fd3d_batch_remove.argtypes = [B, H]
    L.fd3d_batch_size.argtypes = [B]
    L.fd3d_batch_step.argtypes = [B, ctypes.c_int]
    L.fd3d_batch_sync.argtypes = [B]
    L.fd3d_batch_info.argtypes = [B, ctypes.POINTER(ctypes.c_int)]
    L.fd3d_batch_set_fused_items.argtypes = [B, ctypes.c_int]
    L.fd3d_batch_elapsed.restype = ctypes.c_double
    L.fd3d_batch_elapsed.argtypes = [B, ctypes.POINTER(ctypes.c_int)]
    L.fd3d_batch_last_error.restype = ctypes.c_char_p
    L.fd3d_batch_last_error.argtypes = [B]
    L.fd3d_batch_destroy.argtypes = [B]
    if path is None:
        _lib = L
    return L


def _f32(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float32))


def _f64(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _fp(a: np.ndarray):
    return a.ctypes.data_as(_FP)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_DP)


def make_config(case: C.Case, fvw_scalars: Optional[Dict[str, float]] = None, newton_max: int = 0,
                ring_steps: int = 0) -> Tuple[Config, np.ndarray]:
    sta = np.asarray(case.stations, dtype=np.int32).reshape(-1) if case.stations else np.zeros(3, np.int32)
    cfg = Config()
    cfg.nxt, cfg.nyt, cfg.nzt, cfg.nabc = case.nxt, case.nyt, case.nzt, case.nabc
    cfg.fspace = int(case.fspace)
    fvw = case.friction == "fvw"
    cfg.friction = FD3D_FVW if fvw else FD3D_SW
    cfg.nucleation = int(fvw)
    cfg.nstations = len(case.stations)
    cfg.stations = sta.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    cfg.dh, cfg.dt, cfg.damp_s = case.dh, case.dt, case.damp_s
    cfg.pml_vp, cfg.pml_fact = case.pml_vp, case.pml_fact
    if fvw and fvw_scalars:
        for name in C.FVW_SCALARS:
            setattr(cfg, name, float(fvw_scalars[name]))
    cfg.newton_max = newton_max
    cfg.ring_steps = ring_steps
    return cfg, sta


def footprint(case: C.Case) -> int:
    cfg, _sta = make_config(case)
    return int(load().fd3d_footprint(ctypes.byref(cfg)))


def device_memory() -> Tuple[int, int]:
    """(free, total) bytes of device memory."""
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    rc = load().fd3d_device_memory(ctypes.byref(f), ctypes.byref(t))
    if rc != 0:
        raise FD3DError(rc, "hipMemGetInfo failed")
    return int(f.value), int(t.value)


class Solver:
    """One rupture problem resident on the GPU (the replacement of fd3d()).

    Mirrors the reference's flow: the caller fills the module arrays
    (medium and fault) as fd3d_init()/forwardspecial*() would, then the
    time loop runs on the device; per-step fault slabs and station samples
    come back as numpy arrays shaped like the reference's .res streams.
    """

    def __init__(self, case: C.Case, fault: Optional[Dict] = None, fvw_scalars=None, medium=None,
                 newton_max: int = 0, ring_steps: int = 0, init: bool = True):
        import time
        clock = time.perf_counter
        t0 = clock()
        self.L = load()
        self.case = case
        fvw = case.friction == "fvw"
        if fvw and fault is None:
            fault, fvw_scalars = C.fault_fvw(case)
        self.cfg, self._sta = make_config(case, fvw_scalars, newton_max, ring_steps)
        self.ring_steps = ring_steps if ring_steps > 0 else 64   # the library's default ring
        self.h = ctypes.c_void_p()
        t1 = clock()
        rc = self.L.fd3d_create(ctypes.byref(self.cfg), ctypes.byref(self.h))
        if rc != 0:
            self.h = None
            raise FD3DError(rc, "fd3d_create failed")
        t2 = clock()
        try:
            if medium is None:
                medium = C.material(case, np.float32)
            lam, mu, d = (_f32(x) for x in medium)
            self._chk(self.L.fd3d_upload_medium(self.h, _fp(lam), _fp(mu), _fp(d)))
            t3 = clock()
            if fvw:
                arrs = [_f64(fault[n]) for n in C.FVW_FIELDS]
                self._chk(self.L.fd3d_upload_fault_fvw(self.h, *[_dp(a) for a in arrs]))
            else:
                if fault is None:
                    fault = C.fault_sw(case, np.float32)
                arrs = [_f32(fault[n]) for n in C.SW_FIELDS]
                self._chk(self.L.fd3d_upload_fault_sw(self.h, *[_fp(a) for a in arrs]))
            t4 = clock()
            if init:
                self.init()
        except Exception:
            self.close()   # release the device memory of a half-built problem
            raise
        # host seconds of each set-up phase
        self.setup_times = {"inputs": t1 - t0, "create": t2 - t1, "medium": t3 - t2, "fault": t4 - t3,
                            "init": clock() - t4}

    def _chk(self, rc: int):
        if rc != 0:
            msg = self.L.fd3d_last_error(self.h).decode() if self.h else ""
            raise FD3DError(rc, msg)

    def init(self):
        self._chk(self.L.fd3d_init_state(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.L.fd3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------------ run
    def step(self, nsteps: int, outputs: bool = True) -> Dict[str, np.ndarray]:
        """Advance nsteps; returns per-step slabs (nsteps, nxtT, nztT) and stations."""
        case = self.case
        nxT, nzT = case.fault_shape
        nsta = len(case.stations)
        res: Dict[str, np.ndarray] = {}
        if not outputs:
            self._chk(self.L.fd3d_step(self.h, nsteps, None, None))
            return res
        out = np.empty((nsteps, NOUT, nzT, nxT), dtype=np.float32)
        seis = np.empty((nsteps, max(nsta, 1), 3), dtype=np.float32)
        self._chk(self.L.fd3d_step(self.h, nsteps, _fp(out), _fp(seis) if nsta else None))
        for o, name in enumerate(OUT_NAMES):
            res[name] = out[:, o].transpose(0, 2, 1)
        res["seis"] = seis[:, :nsta, :]
        return res

    def step_begin(self, nsteps: int):
        """Advance nsteps (<= ring_steps/2) and queue the copy of their
        outputs to pinned host memory; returns at once (fd3d_step_begin)."""
        self._chk(self.L.fd3d_step_begin(self.h, nsteps))

    def step_wait(self, copy: bool = True) -> Dict[str, np.ndarray]:
        """Outputs of the oldest chunk in flight (fd3d_step_wait), shaped as
        step() returns them.  copy=False returns views of the handle's pinned
        buffer, valid until the step_begin after the next."""
        n = ctypes.c_int()
        po, ps = _FP(), _FP()
        self._chk(self.L.fd3d_step_wait(self.h, ctypes.byref(n), ctypes.byref(po), ctypes.byref(ps)))
        case = self.case
        nxT, nzT = case.fault_shape
        nsta = len(case.stations)
        out = np.ctypeslib.as_array(po, shape=(n.value, NOUT, nzT, nxT))
        seis = np.ctypeslib.as_array(ps, shape=(n.value, max(nsta, 1), 3))
        if copy:
            out, seis = out.copy(), seis.copy()
        res = {name: out[:, o].transpose(0, 2, 1) for o, name in enumerate(OUT_NAMES)}
        res["seis"] = seis[:, :nsta, :]
        return res

    def stream(self, nsteps: int, chunk: int = 32, copy: bool = True):
        """Advance nsteps, yielding each chunk's outputs while the GPU
        computes the next chunk (the overlapped output stream).  With
        copy=False the yielded arrays are views of the pinned buffer, valid
        until the generator resumes."""
        chunk = max(1, min(chunk, self.ring_steps // 2))
        left = nsteps
        inflight = 0
        while left > 0 or inflight:
            # keep two chunks in flight: the one handed out last is done with
            # once the generator resumes, so its buffer can take the next
            while left > 0 and inflight < 2:
                n = min(chunk, left)
                self.step_begin(n)
                left -= n
                inflight += 1
            yield self.step_wait(copy=copy)
            inflight -= 1

    def read_outputs(self, nsteps: int) -> Dict[str, np.ndarray]:
        """Slabs and stations of the last nsteps steps from the device ring
        (for problems advanced by a Batch)."""
        case = self.case
        nxT, nzT = case.fault_shape
        nsta = len(case.stations)
        out = np.empty((nsteps, NOUT, nzT, nxT), dtype=np.float32)
        seis = np.empty((nsteps, max(nsta, 1), 3), dtype=np.float32)
        self._chk(self.L.fd3d_read_outputs(self.h, nsteps, _fp(out), _fp(seis) if nsta else None))
        res = {name: out[:, o].transpose(0, 2, 1) for o, name in enumerate(OUT_NAMES)}
        res["seis"] = seis[:, :nsta, :]
        return res

    # ------------------------------------------------------ inversion bins
    def set_source_bins(self, NL: int = 0, NW: int = 0, dL: float = 0.0, dW: float = 0.0, dtseis: float = 1.0,
                        nbins: Optional[int] = None):
        """Accumulate MSRX/MSRZ (NL x NW patches of dL x dW metres) and the
        moment rate in bins of dtseis seconds, as the reference's time loop
        does for the inversion (fd3d_theo.f90:560-578).  Re-initialises the
        state (the bins start from zero with the wavefield)."""
        self._bins = (NL, NW, nbins or self.case.ntfd)
        self._chk(self.L.fd3d_set_source_bins(self.h, NL, NW, dL, dW, dtseis, self._bins[2]))
        self.init()

    def source_bins(self) -> Dict[str, np.ndarray]:
        NL, NW, nb = self._bins
        msrx = np.zeros(max(NL * NW * nb, 1), np.float32)
        msrz = np.zeros(max(NL * NW * nb, 1), np.float32)
        mr = np.zeros(nb, np.float32)
        self._chk(self.L.fd3d_download_source_bins(self.h, _fp(msrx), _fp(msrz), _fp(mr)))
        return {"MSRX": msrx[:NL * NW * nb].reshape(NL * NW, nb), "MSRZ": msrz[:NL * NW * nb].reshape(NL * NW, nb),
                "MomentRate": mr}

    def steps_done(self) -> int:
        return int(self.L.fd3d_steps_done(self.h))

    def sync(self):
        self._chk(self.L.fd3d_sync(self.h))

    def final(self) -> Dict[str, np.ndarray]:
        c = self.case
        buf = np.empty(7 * c.nxt * c.nzt, dtype=np.float32)
        self._chk(self.L.fd3d_download_final(self.h, _fp(buf)))
        buf = buf.reshape(7, c.nzt, c.nxt).transpose(0, 2, 1)
        return {n: buf[q].copy() for q, n in enumerate(FINAL_NAMES)}

    def field(self, name: str) -> np.ndarray:
        c = self.case
        buf = np.empty(c.cells, dtype=np.float32)
        self._chk(self.L.fd3d_download_field(self.h, FIELDS.index(name), _fp(buf)))
        return buf.reshape((c.nxt, c.nyt, c.nzt), order="F")

    def set_field(self, name: str, arr: np.ndarray):
        a = _f32(arr)
        self._chk(self.L.fd3d_upload_field(self.h, FIELDS.index(name), _fp(a)))

    # ---------------------------------------------------------------- tuning
    def set_tuning(self, mode: int, tz_vel: int = 16, ny_vel: int = 64, tz_str: int = 16, ny_str: int = 64):
        self._chk(self.L.fd3d_set_tuning(self.h, mode, tz_vel, ny_vel, tz_str, ny_str))

    def get_tuning(self) -> Tuple[int, int, int, int, int]:
        buf = (ctypes.c_int * 5)()
        self._chk(self.L.fd3d_get_tuning(self.h, buf))
        return tuple(buf)

    # ------------------------------------------------------------ profiling
    def set_profiling(self, on: bool):
        self._chk(self.L.fd3d_set_profiling(self.h, int(on)))

    def kernel_times(self) -> Dict[str, Tuple[float, int]]:
        n = 32
        names = (ctypes.c_char_p * n)()
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_long * n)()
        k = self.L.fd3d_kernel_times(self.h, n, names, ms, cnt)
        if k < 0:
            self._chk(k)
        return {names[q].decode(): (ms[q], cnt[q]) for q in range(k)}

    def step_bytes(self) -> Tuple[float, int, int]:
        b = ctypes.c_long()
        p = ctypes.c_long()
        v = self.L.fd3d_step_bytes(self.h, ctypes.byref(b), ctypes.byref(p))
        return v, b.value, p.value


PERSISTENT, STAGED = 0, 1


class Batch:
    """Many resident problems advanced together (fd3d_batch_*): one
    persistent launch per time step (or one launch per stage), replayed
    from a captured HIP graph.  Members are Solver objects; their outputs
    are read back with Solver.read_outputs / Solver.final."""

    def __init__(self, max_problems: int, tz: int = 16, mode: int = 1, static: bool = False):
        """static: the batch keeps its tables in the library's fixed device
        array and allocates no device memory (fd3d_batch_create_static; staged
        launches, at most 64 problems, one such batch per GPU at a time)."""
        self.L = load()
        self.b = ctypes.c_void_p()
        if static:
            rc = self.L.fd3d_batch_create_static(max_problems, tz, ctypes.byref(self.b))
        else:
            rc = self.L.fd3d_batch_create(max_problems, tz, mode, ctypes.byref(self.b))
        if rc != 0:
            self.b = None
            raise FD3DError(rc, "fd3d_batch_create_static failed" if static else "fd3d_batch_create failed")
        self.members = []

    def _chk(self, rc: int):
        if rc != 0:
            raise FD3DError(rc, self.L.fd3d_batch_last_error(self.b).decode() if self.b else "")

    def add(self, s: Solver):
        self._chk(self.L.fd3d_batch_add(self.b, s.h))
        self.members.append(s)

    def remove(self, s: Solver):
        self._chk(self.L.fd3d_batch_remove(self.b, s.h))
        self.members.remove(s)

    def __len__(self):
        return int(self.L.fd3d_batch_size(self.b))

    def step(self, nsteps: int):
        self._chk(self.L.fd3d_batch_step(self.b, nsteps))

    def sync(self):
        self._chk(self.L.fd3d_batch_sync(self.b))

    def set_fused_items(self, min_items_per_cu: int):
        """Least fused work items per CU at which members asking for the
        fused interior sweep run it in the batch (0: always)."""
        self._chk(self.L.fd3d_batch_set_fused_items(self.b, min_items_per_cu))

    def info(self) -> Dict[str, int]:
        buf = (ctypes.c_int * 4)()
        self._chk(self.L.fd3d_batch_info(self.b, buf))
        return {"workers": buf[0], "items": buf[1], "segments": buf[2], "launches": buf[3]}

    def elapsed(self) -> Tuple[float, int]:
        n = ctypes.c_int()
        ms = self.L.fd3d_batch_elapsed(self.b, ctypes.byref(n))
        return ms, n.value

    def close(self):
        if getattr(self, "b", None):
            self.L.fd3d_batch_destroy(self.b)
            self.b = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def gpu_available() -> bool:
    """True when a HIP device is visible (checked without initialising torch)."""
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        return False
    n = ctypes.c_int(0)
    try:
        rc = hip.hipGetDeviceCount(ctypes.byref(n))
    except Exception:
        return False
    return rc == 0 and n.value > 0
