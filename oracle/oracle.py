"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.  The product (nuosc_amd, libnuosc_hip.so)
never does.  See nuosc_oracle.h for what the oracle restates and how each part
is pinned.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_VGRID = os.path.join(HERE, "_ref", "ref_vgrid")

FIELDS = ("ee", "mm", "emr", "emi", "bee", "bmm", "bemr", "bemi")
# 3 flavours: the 8 fields above, then these 10
FIELDS3 = FIELDS + ("tt", "mtr", "mti", "ter", "tei", "btt", "bmtr", "bmti", "bter", "btei")
NVAR_MAX = 18
DBL_MAX = np.finfo(np.float64).max

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class _Sim(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("nz", C.c_int), ("nv", C.c_int), ("gx", C.c_int), ("gz", C.c_int),
        ("x0", C.c_double), ("x1", C.c_double), ("z0", C.c_double), ("z1", C.c_double),
        ("dx", C.c_double), ("dz", C.c_double),
        ("X", C.POINTER(C.c_double)), ("Z", C.POINTER(C.c_double)),
        ("vx", C.POINTER(C.c_double)), ("vy", C.POINTER(C.c_double)),
        ("vz", C.POINTER(C.c_double)), ("vw", C.POINTER(C.c_double)),
        ("lpts", C.c_long),
        ("dt", C.c_double), ("ds_L", C.c_double), ("cfl", C.c_double), ("ko", C.c_double),
        ("mu", C.c_double), ("pmo", C.c_double), ("ct", C.c_double), ("st", C.c_double),
        ("renorm", C.c_int),
        ("phy_time", C.c_double),
        ("n_nue0", C.c_double * 2),
        ("G0", C.POINTER(C.c_double)), ("G0b", C.POINTER(C.c_double)),
        ("stat", C.POINTER(C.c_double) * NVAR_MAX), ("rhs", C.POINTER(C.c_double) * NVAR_MAX),
        ("pre", C.POINTER(C.c_double) * NVAR_MAX), ("cor", C.POINTER(C.c_double) * NVAR_MAX),
        ("analysis_lines", C.c_int),
        ("bc", C.c_int), ("zc", C.c_double), ("inflow_on", C.c_int),
        ("inflow_eps0", C.c_double), ("inflow_sigma", C.c_double),
        ("ko_order", C.c_int), ("adv", C.c_int),
        ("nvar", C.c_int), ("hv", C.c_double * 9),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.orc_vgrid_icosahedron.argtypes = [C.c_int, dp, dp, dp, dp]
        L.orc_vgrid_icosahedron.restype = C.c_int
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_double] * 8 + [C.c_int]
        L.orc_create.restype = C.POINTER(_Sim)
        L.orc_destroy.argtypes = [C.POINTER(_Sim)]
        L.orc_set_physics.argtypes = [C.POINTER(_Sim), C.c_double, C.c_double, C.c_int]
        L.orc_fill_init.argtypes = [C.POINTER(_Sim), C.c_int] + [C.c_double] * 7 + [C.c_uint64]
        L.orc_fill_init.restype = C.c_int
        L.orc_seed_uniforms.argtypes = [C.c_uint64, C.c_long, dp, dp]
        L.orc_periodic_ghosts.argtypes = [C.POINTER(_Sim), dp * NVAR_MAX]
        L.orc_set_boundary.argtypes = [C.POINTER(_Sim), C.c_int, C.c_double]
        L.orc_set_boundary.restype = C.c_int
        L.orc_set_scheme.argtypes = [C.POINTER(_Sim), C.c_int, C.c_int]
        L.orc_set_scheme.restype = C.c_int
        L.orc_set_flavors.argtypes = [C.POINTER(_Sim), C.c_int]
        L.orc_set_flavors.restype = C.c_int
        L.orc_vacuum_h3.argtypes = [dp]
        L.orc_open_ghosts.argtypes = [C.POINTER(_Sim), dp * NVAR_MAX, C.c_double]
        L.orc_fill_ghosts.argtypes = [C.POINTER(_Sim), dp * NVAR_MAX, C.c_double]
        L.orc_calrhs.argtypes = [C.POINTER(_Sim), dp * NVAR_MAX, dp * NVAR_MAX]
        L.orc_step_rk4.argtypes = [C.POINTER(_Sim)]
        L.orc_step_srk3.argtypes = [C.POINTER(_Sim)]
        L.orc_renormalize.argtypes = [C.POINTER(_Sim), dp * NVAR_MAX]
        L.orc_analysis.argtypes = [C.POINTER(_Sim), dp]
        L.orc_analysis.restype = C.c_int
        L.orc_format_analysis_line.argtypes = [C.POINTER(_Sim), dp, C.c_char_p, C.c_int]
        L.orc_format_analysis_line.restype = C.c_int
        L.orc_get_field.argtypes = [C.POINTER(_Sim), C.c_int, C.c_int, dp]
        L.orc_set_field.argtypes = [C.POINTER(_Sim), C.c_int, C.c_int, dp]
        L.orc_field_ptr.argtypes = [C.POINTER(_Sim), C.c_int, C.c_int]
        L.orc_field_ptr.restype = dp
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def vgrid(nn: int):
    """(vx, vy, vz, vw) of the icosahedral Voronoi grid at refinement nn."""
    n = 10 * nn * nn + 2
    out = [np.zeros(n) for _ in range(4)]
    got = lib().orc_vgrid_icosahedron(nn, *[_dp(a) for a in out])
    assert got == n
    return tuple(out)


def seed_uniforms(seed: int, lin: int):
    ur, ui = C.c_double(), C.c_double()
    lib().orc_seed_uniforms(seed, lin, C.byref(ur), C.byref(ui))
    return ur.value, ui.value


def vacuum_h3():
    """The oracle's 3-flavour vacuum Hamiltonian {hee, hmm, htt, hemr, hemi,
    hmtr, hmti, hter, htei}."""
    out = np.zeros(9)
    lib().orc_vacuum_h3(_dp(out))
    return out


def ref_vgrid(nn: int):
    """Velocity grid printed by the reference's own icosahedron.h (oracle/_ref),
    or None when that binary was not built."""
    if not os.path.exists(REF_VGRID):
        return None
    txt = subprocess.run([REF_VGRID, str(nn)], check=True, capture_output=True, text=True).stdout
    rows = [[float.fromhex(t) for t in line.split()] for line in txt.strip().splitlines()]
    a = np.array(rows)
    return a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy(), a[:, 3].copy()


class OracleSim:
    """The oracle's NuOsc: reference layout and operation order, FP64."""

    BUF_STAT, BUF_RHS, BUF_PRE, BUF_COR = 0, 1, 2, 3
    BC_PERIODIC, BC_OPEN_Z = 0, 1
    # advection schemes (ORC_ADV_*)
    ADV_CFD4, ADV_CFD2, ADV_UPWIND2, ADV_LOPSIDED, ADV_MUSCL, ADV_OFF = range(6)

    def __init__(self, nn=4, gx=2, gz=2, x0=-3.2, x1=3.2, z0=-3.2, z1=3.2, dx=0.1, dz=0.1,
                 cfl=0.4, ko=0.1, mu=1.0, pmo=0.0, renorm=0, vgrid_mode=0, bc=0, zc=0.0, ko_order=3, adv=0,
                 nflavor=2):
        L = lib()
        p = L.orc_create(nn, gx, gz, x0, x1, z0, z1, dx, dz, cfl, ko, vgrid_mode)
        if not p:
            raise ValueError("orc_create rejected the configuration")
        self._p = p
        L.orc_set_physics(p, mu, pmo, renorm)
        if L.orc_set_boundary(p, bc, zc) != 0:
            self.close()
            raise ValueError("orc_set_boundary rejected the configuration")
        if L.orc_set_scheme(p, ko_order, adv) != 0:
            self.close()
            raise ValueError("orc_set_scheme rejected the configuration (ghost width below the stencil radius?)")
        if L.orc_set_flavors(p, nflavor) != 0:
            self.close()
            raise ValueError("orc_set_flavors rejected the configuration")

    def close(self):
        if self._p:
            lib().orc_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def s(self):
        return self._p.contents

    @property
    def nvar(self):
        return self.s.nvar

    @property
    def shape(self):
        s = self.s
        return (s.nx, s.nz, s.nv)

    def vgrid(self):
        s = self.s
        n = s.nv
        return tuple(np.ctypeslib.as_array(getattr(s, k), (n,)).copy() for k in ("vx", "vy", "vz", "vw"))

    @property
    def G0(self):
        return np.ctypeslib.as_array(self.s.G0, (self.s.nv,)).copy()

    @property
    def G0b(self):
        return np.ctypeslib.as_array(self.s.G0b, (self.s.nv,)).copy()

    @property
    def phy_time(self):
        return self.s.phy_time

    @property
    def n_nue0(self):
        return (self.s.n_nue0[0], self.s.n_nue0[1])

    def fill_init(self, ipt=0, alpha=0.9, eps0=0.1, sigma=100.0, lnue=0.6, lnueb=0.53,
                  lnuex=DBL_MAX, lnuebx=DBL_MAX, seed=0):
        rc = lib().orc_fill_init(self._p, ipt, alpha, eps0, sigma, lnue, lnueb, lnuex, lnuebx, seed)
        if rc != 0:
            raise ValueError(f"ipt={ipt} not supported by the oracle")

    def step(self, n=1):
        for _ in range(n):
            lib().orc_step_rk4(self._p)

    def step_srk3(self, n=1):
        for _ in range(n):
            lib().orc_step_srk3(self._p)

    def analysis(self):
        out = np.zeros(9)
        status = lib().orc_analysis(self._p, _dp(out))
        return out, status

    def analysis_line(self, out):
        buf = C.create_string_buffer(1024)
        n = lib().orc_format_analysis_line(self._p, _dp(np.ascontiguousarray(out, dtype=np.float64)), buf, 1024)
        return buf.value[:n].decode()

    def get_state(self, buf=0):
        nx, nz, nv = self.shape
        out = np.zeros((self.nvar, nx, nz, nv))
        for f in range(self.nvar):
            tmp = np.zeros(nx * nz * nv)
            lib().orc_get_field(self._p, buf, f, _dp(tmp))
            out[f] = tmp.reshape(nx, nz, nv)
        return out

    def set_state(self, arr, buf=0):
        nx, nz, nv = self.shape
        arr = np.asarray(arr, dtype=np.float64)
        assert arr.shape == (self.nvar, nx, nz, nv)
        for f in range(self.nvar):
            lib().orc_set_field(self._p, buf, f, _dp(np.ascontiguousarray(arr[f]).reshape(-1)))

    def raw_buffer(self, buf, f):
        """Ghost-inclusive buffer (lpts doubles) as a numpy view."""
        ptr = lib().orc_field_ptr(self._p, buf, f)
        return np.ctypeslib.as_array(ptr, (self.s.lpts,))

    def _ptrs(self, buf):
        arr = (C.POINTER(C.c_double) * NVAR_MAX)()
        for f in range(self.nvar):
            arr[f] = lib().orc_field_ptr(self._p, buf, f)
        return arr

    def fill_ghosts(self, buf=0, t=None):
        """Boundary fill of the configured mode (open z: at time t, default
        the current physical time)."""
        lib().orc_fill_ghosts(self._p, self._ptrs(buf), self.phy_time if t is None else t)

    def ghosted(self, buf=0):
        """(nvar, nx+2gx, nz+2gz, nv) view of a buffer including its ghost band."""
        s = self.s
        return np.stack([self.raw_buffer(buf, f).reshape(s.nx + 2 * s.gx, s.nz + 2 * s.gz, s.nv)
                         for f in range(self.nvar)])

    def calrhs(self, state):
        """RHS of an interior state (nvar,nx,nz,nv): boundary fill at the current
        time, then calRHS."""
        self.set_state(state, buf=self.BUF_PRE)
        self.fill_ghosts(self.BUF_PRE)
        lib().orc_calrhs(self._p, self._ptrs(self.BUF_COR), self._ptrs(self.BUF_PRE))
        return self.get_state(self.BUF_COR)

    def renormalize(self):
        lib().orc_renormalize(self._p, self._ptrs(self.BUF_STAT))
