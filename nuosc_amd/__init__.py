"""nuosc_amd -- MI355X (gfx950) implementation of the nuosc RK4 time-stepper core.

Python binding (ctypes) of the C ABI in include/nuosc_hip.h.  ``NuOsc`` mirrors
the reference's ``NuOsc`` class (nuosc_class.h:99-191): same constructor
arguments, ``set_mu``/``set_pmo``/``set_renorm``, ``fillInitValue``,
``step_rk4``, ``calRHS``, ``analysis``, ``renormalize`` and snapshot slices.

There is no CPU fallback: the compute path is libnuosc_hip.so, and importing
the binding raises if that library is missing (build it with
``python -m nuosc_amd.build``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnuosc_hip.so")
CLI_PATH = os.path.join(HERE, "bin", "nuosc")
HEADER = os.path.join(os.path.dirname(HERE), "include", "nuosc_hip.h")

FIELDS = ("ee", "mm", "emr", "emi", "bee", "bmm", "bemr", "bemi")
# 3 flavours: the 8 fields above, then these 10 (field indices 8..17)
FIELDS3 = FIELDS + ("tt", "mtr", "mti", "ter", "tei", "btt", "bmtr", "bmti", "bter", "btei")
DBL_MAX = np.finfo(np.float64).max

NUOSC_OK, NUOSC_EINVAL, NUOSC_EHIP, NUOSC_ENOMEM, NUOSC_EUNSTABLE, NUOSC_ENODEV, NUOSC_EAGAIN = range(7)
PREC_FP64, PREC_FP32 = 64, 32
VGRID_VORONOI, VGRID_EQUAL = 0, 1
BC_PERIODIC, BC_OPEN_Z = 0, 1
# launch forms of the batched stage kernels (NUOSC_BATCH_*)
BATCH_GRID, BATCH_PERSISTENT = 0, 1
# advection schemes (NUOSC_ADV_*): the reference's default and its ADVEC_* variants
ADV_CFD4, ADV_CFD2, ADV_UPWIND2, ADV_LOPSIDED, ADV_MUSCL, ADV_OFF = range(6)


class NuOscError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nuosc error {code}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [
        ("nn", C.c_int), ("gx", C.c_int), ("gz", C.c_int),
        ("x0", C.c_double), ("x1", C.c_double), ("z0", C.c_double), ("z1", C.c_double),
        ("dx", C.c_double), ("dz", C.c_double),
        ("cfl", C.c_double), ("ko", C.c_double), ("mu", C.c_double), ("pmo", C.c_double),
        ("renorm", C.c_int), ("vgrid", C.c_int), ("precision", C.c_int), ("bc", C.c_int),
        ("device", C.c_int), ("use_graph", C.c_int), ("zc", C.c_double),
        ("ko_order", C.c_int), ("adv", C.c_int), ("nflavor", C.c_int),
    ]


class Request(C.Structure):
    _fields_ = [
        ("cfg", Config), ("ipt", C.c_int),
        ("alpha", C.c_double), ("eps0", C.c_double), ("sigma", C.c_double), ("lnue", C.c_double),
        ("lnueb", C.c_double), ("lnuex", C.c_double), ("lnuebx", C.c_double),
        ("seed", C.c_uint64), ("nsteps", C.c_int), ("ana_every", C.c_int),
    ]


class Result(C.Structure):
    _fields_ = [
        ("id", C.c_longlong), ("status", C.c_int), ("nsteps", C.c_int),
        ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
        ("device_bytes", C.c_longlong), ("n_samples", C.c_int),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("queued", C.c_int), ("active", C.c_int), ("finished", C.c_longlong), ("resident_bytes", C.c_longlong),
        ("updates", C.c_double), ("busy_s", C.c_double), ("now_s", C.c_double),
    ]


class Analysis(C.Structure):
    _fields_ = [
        ("phy_time", C.c_double), ("maxdP", C.c_double), ("surv", C.c_double), ("survb", C.c_double),
        ("avgP", C.c_double), ("avgPb", C.c_double), ("aM0", C.c_double), ("Lex", C.c_double),
        ("ELNe", C.c_double), ("rho_ex2", C.c_double), ("nonfinite", C.c_double), ("status", C.c_int),
    ]

    COLUMNS = ("phy_time", "maxdP", "surv", "survb", "avgP", "avgPb", "aM0", "Lex", "ELNe")

    def row(self):
        """The nine analysis.dat columns (nuosc_ana.cpp:126-129)."""
        return np.array([getattr(self, k) for k in self.COLUMNS])


_lib = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p

# exported symbol -> (argtypes, restype)
SIGNATURES = {
    "nuosc_abi_version": ([], C.c_int),
    "nuosc_build_info": ([], C.c_char_p),
    "nuosc_vgrid": ([C.c_int, C.c_int, _dp, _dp, _dp, _dp], C.c_int),
    "nuosc_create": ([C.POINTER(Config), C.POINTER(_vp)], C.c_int),
    "nuosc_destroy": ([_vp], C.c_int),
    "nuosc_last_error": ([_vp], C.c_char_p),
    "nuosc_set_mu": ([_vp, C.c_double], C.c_int),
    "nuosc_set_pmo": ([_vp, C.c_double], C.c_int),
    "nuosc_set_renorm": ([_vp, C.c_int], C.c_int),
    "nuosc_get_dims": ([_vp, _ip, _ip, _ip, C.POINTER(C.c_longlong)], C.c_int),
    "nuosc_get_scalars": ([_vp, _dp, _dp, _dp, _dp], C.c_int),
    "nuosc_set_phy_time": ([_vp, C.c_double], C.c_int),
    "nuosc_get_vgrid": ([_vp, _dp, _dp, _dp, _dp], C.c_int),
    "nuosc_get_coords": ([_vp, _dp, _dp], C.c_int),
    "nuosc_get_g0": ([_vp, _dp, _dp], C.c_int),
    "nuosc_get_device_bytes": ([_vp, C.POINTER(C.c_longlong)], C.c_int),
    "nuosc_fill_init": ([_vp, C.c_int] + [C.c_double] * 7 + [C.c_uint64], C.c_int),
    "nuosc_step_rk4": ([_vp, C.c_int], C.c_int),
    "nuosc_step_srk3": ([_vp, C.c_int], C.c_int),
    "nuosc_synchronize": ([_vp], C.c_int),
    "nuosc_calrhs": ([_vp, _dp, _dp], C.c_int),
    "nuosc_analysis": ([_vp, C.POINTER(Analysis)], C.c_int),
    "nuosc_renormalize": ([_vp], C.c_int),
    "nuosc_download_field": ([_vp, C.c_int, _dp], C.c_int),
    "nuosc_download_field_ghosted": ([_vp, C.c_int, _dp], C.c_int),
    "nuosc_upload_field": ([_vp, C.c_int, _dp], C.c_int),
    "nuosc_state_updated": ([_vp], C.c_int),
    "nuosc_p3_slice": ([_vp, _ip, C.c_int, _ip, C.c_int, _dp], C.c_int),
    "nuosc_p3_slice_noalloc": ([_vp, _ip, C.c_int, _ip, C.c_int, _dp], C.c_int),
    "nuosc_last_step_ms": ([_vp, _dp], C.c_int),
    "nuosc_profile_stages": ([_vp, C.c_int, _dp], C.c_int),
    "nuosc_get_plan": ([_vp, _ip, _ip], C.c_int),
    "nuosc_set_zlen": ([_vp, C.c_int], C.c_int),
    "nuosc_autotune": ([_vp, C.c_int, _ip, _dp], C.c_int),
    "nuosc_batch_create": ([C.c_int, C.c_int, C.POINTER(_vp)], C.c_int),
    "nuosc_batch_destroy": ([_vp], C.c_int),
    "nuosc_batch_add": ([_vp, _vp], C.c_int),
    "nuosc_batch_remove": ([_vp, _vp], C.c_int),
    "nuosc_batch_size": ([_vp], C.c_int),
    "nuosc_batch_step": ([_vp, C.c_int], C.c_int),
    "nuosc_batch_last_error": ([_vp], C.c_char_p),
    "nuosc_batch_set_zlen": ([_vp, C.c_int], C.c_int),
    "nuosc_batch_autotune": ([_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)], C.c_int),
    "nuosc_batch_set_launch": ([_vp, C.c_int], C.c_int),
    "nuosc_batch_get_launch": ([_vp], C.c_int),
    "nuosc_estimate_device_bytes": ([C.POINTER(Config), C.POINTER(C.c_longlong)], C.c_int),
    "nuosc_server_create": ([C.c_int, C.c_int, C.c_longlong, C.c_int, C.POINTER(_vp)], C.c_int),
    "nuosc_server_submit": ([_vp, C.POINTER(Request), C.POINTER(C.c_longlong)], C.c_int),
    "nuosc_server_take_result": ([_vp, C.c_longlong, C.c_double, C.POINTER(Result), _dp, C.c_int], C.c_int),
    "nuosc_server_stats": ([_vp, C.POINTER(ServerStats)], C.c_int),
    "nuosc_server_destroy": ([_vp], C.c_int),
    "nuosc_server_set_age_bound": ([_vp, C.c_double], C.c_int),
}


def lib():
    """Load libnuosc_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m nuosc_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def vgrid(nn: int, mode: int = VGRID_VORONOI):
    """Velocity grid (vx, vy, vz, vw) as the library builds it (host only)."""
    n = 10 * nn * nn + 2
    out = [np.zeros(n) for _ in range(4)]
    got = lib().nuosc_vgrid(nn, mode, *[_ptr(a) for a in out])
    if got != n:
        raise NuOscError(NUOSC_EINVAL, "bad nn or mode")
    return tuple(out)


class NuOsc:
    """GPU NuOsc.  Constructor follows NuOsc(px,pz,nv,nphi,gx,gz,x0,x1,z0,z1,dx,dz,CFL,ko)
    (nuosc_class.h:131-133) without px, pz, nphi; ``nv`` is the icosahedron level."""

    def __init__(self, nv=4, gx=2, gz=2, x0=-1.0, x1=1.0, z0=-10.0, z1=10.0, dx=0.1, dz=0.1,
                 CFL=0.4, ko=0.0, mu=1.0, pmo=0.1, renorm=False, precision=PREC_FP64,
                 vgrid_mode=VGRID_VORONOI, device=0, use_graph=True, bc=0, zc=0.0, ko_order=3, adv=ADV_CFD4,
                 nflavor=2):
        """bc: BC_PERIODIC (the reference's boundary) or BC_OPEN_Z (open z
        ends with the ipt-0 inflow pulse, x periodic); zc: centre of the ipt-0
        profile (0 in the reference); ko_order (3 or 5) and adv (ADV_*): the
        discretisation (the reference's KO_ORD_3 / ADVEC_* build options);
        nflavor 2 (8 fields, FIELDS) or 3 (18 fields, FIELDS3: the
        reference's N_FLAVOR_3 build option)."""
        L = lib()
        cfg = Config(nn=nv, gx=gx, gz=gz, x0=x0, x1=x1, z0=z0, z1=z1, dx=dx, dz=dz, cfl=CFL, ko=ko,
                     mu=mu, pmo=pmo, renorm=int(renorm), vgrid=vgrid_mode, precision=precision, bc=bc,
                     device=device, use_graph=int(use_graph), zc=zc, ko_order=ko_order, adv=adv, nflavor=nflavor)
        h = _vp()
        rc = L.nuosc_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise NuOscError(rc, (L.nuosc_last_error(None) or b"").decode())
        self._h = h
        self.precision = precision
        nx, nz, nvv, lp = C.c_int(), C.c_int(), C.c_int(), C.c_longlong()
        L.nuosc_get_dims(h, C.byref(nx), C.byref(nz), C.byref(nvv), C.byref(lp))
        self.nx, self.nz, self.nv, self.lpts = nx.value, nz.value, nvv.value, lp.value
        self.gx, self.gz = gx, gz
        self.nvar = 18 if nflavor == 3 else 8

    # -- plumbing ---------------------------------------------------------
    def _check(self, rc):
        if rc != 0:
            raise NuOscError(rc, (lib().nuosc_last_error(self._h) or b"").decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().nuosc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def shape(self):
        return (self.nx, self.nz, self.nv)

    # -- reference setters (nuosc_class.h:148-159) ---------------------------
    def set_mu(self, mu):
        self._check(lib().nuosc_set_mu(self._h, mu))

    def set_pmo(self, pmo):
        self._check(lib().nuosc_set_pmo(self._h, pmo))

    def set_renorm(self, renorm):
        self._check(lib().nuosc_set_renorm(self._h, int(renorm)))

    # -- scalars ----------------------------------------------------------
    def _scalars(self):
        dt, ds, t = C.c_double(), C.c_double(), C.c_double()
        n0 = (C.c_double * 2)()
        self._check(lib().nuosc_get_scalars(self._h, C.byref(dt), C.byref(ds), C.byref(t), n0))
        return dt.value, ds.value, t.value, (n0[0], n0[1])

    @property
    def dt(self):
        return self._scalars()[0]

    @property
    def ds_L(self):
        return self._scalars()[1]

    @property
    def phy_time(self):
        return self._scalars()[2]

    def set_phy_time(self, t):
        self._check(lib().nuosc_set_phy_time(self._h, float(t)))

    @property
    def n_nue0(self):
        return self._scalars()[3]

    def vgrid(self):
        out = [np.zeros(self.nv) for _ in range(4)]
        self._check(lib().nuosc_get_vgrid(self._h, *[_ptr(a) for a in out]))
        return tuple(out)

    def coords(self):
        X, Z = np.zeros(self.nx), np.zeros(self.nz)
        self._check(lib().nuosc_get_coords(self._h, _ptr(X), _ptr(Z)))
        return X, Z

    def g0(self):
        a, b = np.zeros(self.nv), np.zeros(self.nv)
        self._check(lib().nuosc_get_g0(self._h, _ptr(a), _ptr(b)))
        return a, b

    def device_bytes(self):
        b = C.c_longlong()
        self._check(lib().nuosc_get_device_bytes(self._h, C.byref(b)))
        return b.value

    # -- the path ---------------------------------------------------------
    def fillInitValue(self, ipt=0, alpha=0.9, eps0=0.1, sigma=100.0, lnue=0.6, lnueb=0.53,
                      lnuex=DBL_MAX, lnuebx=DBL_MAX, seed=0):
        self._check(lib().nuosc_fill_init(self._h, int(ipt), alpha, eps0, sigma, lnue, lnueb,
                                          lnuex, lnuebx, int(seed)))

    def step_rk4(self, nsteps=1):
        self._check(lib().nuosc_step_rk4(self._h, int(nsteps)))

    def step_srk3(self, nsteps=1):
        """SSP-RK3 steps (NuOsc::step_srk3, with boundary fills per stage)."""
        self._check(lib().nuosc_step_srk3(self._h, int(nsteps)))

    def synchronize(self):
        self._check(lib().nuosc_synchronize(self._h))

    def last_step_ms(self):
        ms = C.c_double()
        self._check(lib().nuosc_last_step_ms(self._h, C.byref(ms)))
        return ms.value

    def profile_stages(self, nsteps=2):
        """Average device time (ms) of each of the four stage kernels."""
        ms = np.zeros(4)
        self._check(lib().nuosc_profile_stages(self._h, int(nsteps), _ptr(ms)))
        return ms

    def plan(self):
        z, w = C.c_int(), C.c_int()
        self._check(lib().nuosc_get_plan(self._h, C.byref(z), C.byref(w)))
        return z.value, w.value

    def set_zlen(self, zlen):
        self._check(lib().nuosc_set_zlen(self._h, int(zlen)))

    def autotune(self, trials=2):
        """Pick the z-segment length of the stage kernel by timing it on this
        grid (at a step boundary; the state is not modified)."""
        z, ms = C.c_int(), C.c_double()
        self._check(lib().nuosc_autotune(self._h, int(trials), C.byref(z), C.byref(ms)))
        return z.value, ms.value

    def calRHS(self, state):
        """RHS of an interior state array (nvar, nx, nz, nv) -> (nvar, nx, nz, nv)."""
        a = np.ascontiguousarray(state, dtype=np.float64)
        assert a.shape == (self.nvar, self.nx, self.nz, self.nv)
        out = np.zeros_like(a)
        self._check(lib().nuosc_calrhs(self._h, _ptr(a), _ptr(out)))
        return out

    def analysis(self, check=False):
        a = Analysis()
        rc = lib().nuosc_analysis(self._h, C.byref(a))
        if rc not in (0, NUOSC_EUNSTABLE) or (check and rc):
            self._check(rc)
        return a

    def renormalize(self):
        self._check(lib().nuosc_renormalize(self._h))

    # -- state transfer ---------------------------------------------------
    def get_field(self, f):
        out = np.zeros(self.nx * self.nz * self.nv)
        self._check(lib().nuosc_download_field(self._h, int(f), _ptr(out)))
        return out.reshape(self.nx, self.nz, self.nv)

    def get_state(self):
        return np.stack([self.get_field(f) for f in range(self.nvar)])

    def get_field_ghosted(self, f):
        shp = (self.nx + 2 * self.gx, self.nz + 2 * self.gz, self.nv)
        out = np.zeros(int(np.prod(shp)))
        self._check(lib().nuosc_download_field_ghosted(self._h, int(f), _ptr(out)))
        return out.reshape(shp)

    def set_state(self, state):
        a = np.asarray(state, dtype=np.float64)
        assert a.shape == (self.nvar, self.nx, self.nz, self.nv)
        for f in range(self.nvar):
            self._check(lib().nuosc_upload_field(self._h, f, _ptr(np.ascontiguousarray(a[f]).reshape(-1))))
        self._check(lib().nuosc_state_updated(self._h))

    def p3_slice(self, xidx, vidx, noalloc=False):
        """P3 at x indices xidx and direction indices vidx: (sx, nz, sv).
        noalloc: through nuosc_p3_slice_noalloc (no device allocation; the
        slice size is bounded)."""
        xi = np.ascontiguousarray(xidx, dtype=np.int32)
        vi = np.ascontiguousarray(vidx, dtype=np.int32)
        out = np.zeros(len(xi) * self.nz * len(vi))
        fn = lib().nuosc_p3_slice_noalloc if noalloc else lib().nuosc_p3_slice
        self._check(fn(self._h, xi.ctypes.data_as(_ip), len(xi), vi.ctypes.data_as(_ip), len(vi), _ptr(out)))
        return out.reshape(len(xi), self.nz, len(vi))


class Batch:
    """Many independent NuOsc runs advanced together, one launch per RK4
    stage for all of them (nuosc_batch_* in include/nuosc_hip.h)."""

    def __init__(self, device=0, precision=PREC_FP64):
        h = _vp()
        rc = lib().nuosc_batch_create(device, precision, C.byref(h))
        if rc != 0:
            raise NuOscError(rc, (lib().nuosc_last_error(None) or b"").decode())
        self._h = h
        self.members = []

    def _check(self, rc):
        if rc != 0:
            raise NuOscError(rc, (lib().nuosc_batch_last_error(self._h) or b"").decode())

    def add(self, sim):
        self._check(lib().nuosc_batch_add(self._h, sim._h))
        self.members.append(sim)

    def remove(self, sim):
        if getattr(sim, "_h", None) is None:   # closed: the library removed it from the batch
            if sim in self.members:
                self.members.remove(sim)
            return
        self._check(lib().nuosc_batch_remove(self._h, sim._h))
        self.members.remove(sim)

    def __len__(self):
        return lib().nuosc_batch_size(self._h)

    def step(self, nsteps=1):
        self._check(lib().nuosc_batch_step(self._h, int(nsteps)))

    def set_zlen(self, zlen):
        """z-segment length shared by all members (0: each member's own plan)."""
        self._check(lib().nuosc_batch_set_zlen(self._h, int(zlen)))

    def autotune(self, trials=2):
        """Pick the fastest shared z-segment length and launch form; returns
        (zlen, stage-1 ms); the form chosen is `launch`."""
        z, ms = C.c_int(), C.c_double()
        self._check(lib().nuosc_batch_autotune(self._h, int(trials), C.byref(z), C.byref(ms)))
        return z.value, ms.value

    def set_launch(self, launch):
        """BATCH_GRID (one block per logical block) or BATCH_PERSISTENT
        (resident blocks drawing logical blocks from per-XCD work queues)."""
        self._check(lib().nuosc_batch_set_launch(self._h, int(launch)))

    @property
    def launch(self):
        return lib().nuosc_batch_get_launch(self._h)

    def close(self):
        if getattr(self, "_h", None):
            lib().nuosc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def estimate_device_bytes(nv=4, x0=-1.0, x1=1.0, z0=-10.0, z1=10.0, dx=0.1, dz=0.1, precision=PREC_FP64, nflavor=2,
                          **_):
    cfg = Config(nn=nv, gx=2, gz=2, x0=x0, x1=x1, z0=z0, z1=z1, dx=dx, dz=dz, precision=precision, nflavor=nflavor)
    b = C.c_longlong()
    rc = lib().nuosc_estimate_device_bytes(C.byref(cfg), C.byref(b))
    if rc != 0:
        raise NuOscError(rc, "bad configuration")
    return b.value


class Server:
    """Continuous-batching run server (nuosc_server_* in include/nuosc_hip.h).

    submit() queues a run and returns its id; results() yields finished runs
    as (Result, series) with series rows (t, Pee, Peebar, maxdP, |M0|)."""

    def __init__(self, device=0, precision=PREC_FP64, hbm_budget_bytes=200 << 30, max_tick_steps=10):
        h = _vp()
        rc = lib().nuosc_server_create(device, precision, int(hbm_budget_bytes), int(max_tick_steps), C.byref(h))
        if rc != 0:
            raise NuOscError(rc, "server creation failed")
        self._h = h

    def submit(self, nsteps, ana_every=10, ipt=5, alpha=0.9, eps0=1e-6, sigma=1.0, lnue=0.6, lnueb=0.5,
               lnuex=DBL_MAX, lnuebx=DBL_MAX, seed=0, nv=4, x0=-1.0, x1=1.0, z0=-1.0, z1=1.0, dx=0.1, dz=0.1,
               CFL=0.4, ko=0.1, mu=1.0, pmo=0.0, nflavor=2):
        cfg = Config(nn=nv, gx=2, gz=2, x0=x0, x1=x1, z0=z0, z1=z1, dx=dx, dz=dz, cfl=CFL, ko=ko, mu=mu, pmo=pmo,
                     renorm=0, vgrid=0, precision=0, bc=0, device=0, use_graph=0, nflavor=nflavor)
        req = Request(cfg=cfg, ipt=ipt, alpha=alpha, eps0=eps0, sigma=sigma, lnue=lnue, lnueb=lnueb, lnuex=lnuex,
                      lnuebx=lnuebx, seed=seed, nsteps=nsteps, ana_every=ana_every)
        rid = C.c_longlong()
        rc = lib().nuosc_server_submit(self._h, C.byref(req), C.byref(rid))
        if rc != 0:
            raise NuOscError(rc, "submit rejected")
        return rid.value

    def take(self, rid=-1, timeout=60.0, max_samples=100000):
        info = Result()
        buf = np.zeros(5 * max_samples)
        rc = lib().nuosc_server_take_result(self._h, int(rid), float(timeout), C.byref(info), _ptr(buf), max_samples)
        if rc == NUOSC_EAGAIN:
            return None
        if rc != 0:
            raise NuOscError(rc, "take_result failed")
        n = min(info.n_samples, max_samples)
        return info, buf[:5 * n].reshape(n, 5).copy()

    def set_age_bound(self, seconds):
        """Seconds the oldest queued run may be passed over by younger runs
        that fit (0: strict arrival order)."""
        rc = lib().nuosc_server_set_age_bound(self._h, float(seconds))
        if rc != 0:
            raise NuOscError(rc, "bad age bound")

    def stats(self):
        st = ServerStats()
        lib().nuosc_server_stats(self._h, C.byref(st))
        return st

    def close(self):
        if getattr(self, "_h", None):
            lib().nuosc_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
