"""ctypes binding of the FP64 CPU oracle (oracle/liblbmc_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product never uses it.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# LBMC_ORACLE_LIB selects another build (the planted-mistake builds of
# tests/test_oracle_pins_bite.py)
LIB_PATH = os.environ.get("LBMC_ORACLE_LIB") or os.path.join(HERE, "liblbmc_oracle.so")

D2Q9, D3Q15 = 0, 1
BGK, NTPOR, MRT, MRTPOR = 0, 1, 2, 3
FIX_D3Q15_MINV, FIX_D3Q15_MEQ, FIX_D3Q15_ZH = 0x1, 0x2, 0x4


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run make -C oracle")
    lib = ct.CDLL(LIB_PATH)
    P = ct.c_void_p
    D = ct.POINTER(ct.c_double)
    I = ct.POINTER(ct.c_int)
    DP = ct.POINTER(D)
    sig = {
        "orc_lattice_tables": (ct.c_int, [ct.c_int, ct.c_uint, I, D, I, D, D]),
        "orc_create": (P, [ct.c_int, I, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_double, D,
                           ct.c_double, ct.c_uint]),
        "orc_destroy": (None, [P]),
        "orc_set_fields": (ct.c_int, [P, D, I, I, DP, DP, D]),
        "orc_init_static": (None, [P]),
        "orc_init_equilibrium": (None, [P, D, DP]),
        "orc_step": (None, [P, ct.c_longlong, ct.c_int]),
        "orc_run": (None, [P, ct.c_longlong, ct.c_int, ct.c_int]),
        "orc_get_f": (None, [P, DP]),
        "orc_set_f": (None, [P, DP]),
        "orc_get_macros": (None, [P, D, DP]),
        "orc_node_types": (ct.c_int, [P, ct.POINTER(ct.c_uint8), ct.POINTER(ct.c_uint16), I]),
        "orc_links": (None, [P, ct.POINTER(ct.c_uint16)]),
        "orc_kinetic_energy": (ct.c_double, [P]),
        "orc_threads": (ct.c_int, []),
        "orc_set_threads": (None, [ct.c_int]),
        "orc_node_collide": (None, [P, ct.c_int, D, D, D, D, ct.c_double, D]),
        "orc_node_micro": (None, [P, ct.c_int, D, D, D, D]),
        "orc_node_macro": (None, [P, ct.c_int, D, D, ct.c_double, D]),
        "orc_equilibrium": (None, [P, ct.c_double, D, D]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = _load()


def _d(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


def _dp(arrays, n):
    arr = (ct.POINTER(ct.c_double) * n)()
    for k in range(n):
        a = arrays[k] if arrays is not None and k < len(arrays) else None
        arr[k] = _d(a) if a is not None else ct.POINTER(ct.c_double)()
    return ct.cast(arr, ct.POINTER(ct.POINTER(ct.c_double)))


def tables(lattice, quirks=0):
    """(e[D,Q], w[Q], opp[Q], M[Q,Q], Minv[Q,Q]) of the oracle."""
    q = 9 if lattice == D2Q9 else 15
    dim = 2 if lattice == D2Q9 else 3
    e = np.zeros(3 * q, dtype=np.int32)
    w = np.zeros(q)
    opp = np.zeros(q, dtype=np.int32)
    M = np.zeros(q * q)
    Minv = np.zeros(q * q)
    _lib.orc_lattice_tables(lattice, quirks, e.ctypes.data_as(ct.POINTER(ct.c_int)), _d(w),
                            opp.ctypes.data_as(ct.POINTER(ct.c_int)), _d(M), _d(Minv))
    return e.reshape(3, q)[:dim].copy(), w, opp, M.reshape(q, q), Minv.reshape(q, q)


def set_threads(n):
    _lib.orc_set_threads(int(n))


def threads():
    return int(_lib.orc_threads())


class Oracle:
    """FP64 restatement of one LBM-C lattice (same constructor as lbmc.Lattice)."""

    def __init__(self, lattice=D3Q15, length=(16, 16, 16), collision_type=BGK, forcing=False,
                 macro_bc=False, micro_bc=False, tau=1.0, tau_mrt=None, c_smag=0.0, quirks=0,
                 **_ignored):
        self.lattice = lattice
        self.dim = 2 if lattice == D2Q9 else 3
        self.q = 9 if lattice == D2Q9 else 15
        L = list(length) + [1] * (3 - len(length))
        self.length = tuple(int(v) for v in L[: self.dim])
        self.n = int(np.prod(self.length))
        self.tau = float(tau)
        Li = (ct.c_int * 3)(*[int(v) for v in L[:3]])
        tm = np.zeros(15)
        if tau_mrt is not None:
            tm[: len(tau_mrt)] = tau_mrt
        self._tm = tm
        self._h = _lib.orc_create(lattice, Li, collision_type, int(bool(forcing)),
                                  int(bool(macro_bc)), int(bool(micro_bc)), float(tau), _d(tm),
                                  float(c_smag), quirks)
        if not self._h:
            raise ValueError("oracle: invalid descriptor")

    def close(self):
        if self._h:
            _lib.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _f64(self, a):
        if a is None:
            return None
        a = np.ascontiguousarray(np.asarray(a, dtype=np.float64)).reshape(-1)
        assert a.size == self.n
        return a

    def _i32(self, a):
        if a is None:
            return None
        a = np.ascontiguousarray(np.asarray(a, dtype=np.int32)).reshape(-1)
        assert a.size == self.n
        return a

    def set_fields(self, geometry=None, macro_bc=None, micro_bc=None, force=None, u=None,
                   rho=None):
        g = self._f64(geometry)
        ma = self._i32(macro_bc)
        mi = self._i32(micro_bc)
        fo = [self._f64(v) for v in force] if force is not None else None
        uu = [self._f64(v) for v in u] if u is not None else None
        r = self._f64(rho)
        self._keep = (g, ma, mi, fo, uu, r)
        st = _lib.orc_set_fields(
            self._h, _d(g) if g is not None else None,
            ma.ctypes.data_as(ct.POINTER(ct.c_int)) if ma is not None else None,
            mi.ctypes.data_as(ct.POINTER(ct.c_int)) if mi is not None else None,
            _dp(fo, 3) if fo is not None else None, _dp(uu, 3) if uu is not None else None,
            _d(r) if r is not None else None)
        if st != 0:
            raise ValueError("oracle: invalid field codes")

    def init_static(self):
        _lib.orc_init_static(self._h)

    def init_equilibrium(self, rho, u):
        r = self._f64(rho)
        uu = [self._f64(v) for v in u]
        _lib.orc_init_equilibrium(self._h, _d(r), _dp(uu, 3))

    def step(self, t, store_macros=True):
        _lib.orc_step(self._h, int(t), int(bool(store_macros)))

    def run(self, t0, nsteps, store_every=1):
        _lib.orc_run(self._h, int(t0), int(nsteps), int(store_every))

    def get_f(self):
        f = [np.empty(self.n) for _ in range(self.q)]
        _lib.orc_get_f(self._h, _dp(f, 15))
        return np.stack(f)

    def set_f(self, f):
        f = [self._f64(v) for v in f]
        _lib.orc_set_f(self._h, _dp(f, 15))

    def get_macros(self):
        rho = np.empty(self.n)
        u = [np.empty(self.n) for _ in range(self.dim)]
        _lib.orc_get_macros(self._h, _d(rho), _dp(u, 3))
        return rho, np.stack(u)

    def node_types(self):
        mask = np.empty(self.n, dtype=np.uint8)
        table = np.zeros(256, dtype=np.uint16)
        nc = ct.c_int()
        st = _lib.orc_node_types(self._h, mask.ctypes.data_as(ct.POINTER(ct.c_uint8)),
                                 table.ctypes.data_as(ct.POINTER(ct.c_uint16)), ct.byref(nc))
        if st != 0:
            raise ValueError("oracle: more than 255 classes")
        return mask, table[: nc.value].copy()

    def links(self):
        out = np.empty(self.n, dtype=np.uint16)
        _lib.orc_links(self._h, out.ctypes.data_as(ct.POINTER(ct.c_uint16)))
        return out

    def kinetic_energy(self):
        return float(_lib.orc_kinetic_energy(self._h))

    # -- single-node operators
    def collide(self, code, f, rho, u, F=None, B=0.0, tau=None):
        """Apply collision code 0..8 to one node; returns (f, rho, u, tau)."""
        f = np.array(f, dtype=np.float64)
        r = ct.c_double(rho)
        uu = np.zeros(3)
        uu[: self.dim] = u
        FF = np.zeros(3)
        if F is not None:
            FF[: self.dim] = F
        t = ct.c_double(self.tau if tau is None else tau)
        _lib.orc_node_collide(self._h, code, _d(f), ct.byref(r), _d(uu), _d(FF), float(B),
                              ct.byref(t))
        return f, r.value, uu[: self.dim].copy(), t.value

    def micro(self, code, f, rho, u, nb=None):
        f = np.array(f, dtype=np.float64)
        r = ct.c_double(rho)
        uu = np.zeros(3)
        uu[: self.dim] = u
        nbv = np.zeros(15) if nb is None else np.array(nb, dtype=np.float64)
        _lib.orc_node_micro(self._h, code, _d(f), ct.byref(r), _d(uu), _d(nbv))
        return f, r.value, uu[: self.dim].copy()

    def macro(self, code, rho, u, prho, pu):
        r = ct.c_double(rho)
        uu = np.zeros(3)
        uu[: self.dim] = u
        pp = np.zeros(3)
        pp[: self.dim] = pu
        _lib.orc_node_macro(self._h, code, ct.byref(r), _d(uu), float(prho), _d(pp))
        return r.value, uu[: self.dim].copy()

    def equilibrium(self, rho, u):
        uu = np.zeros(3)
        uu[: self.dim] = u
        feq = np.zeros(self.q)
        _lib.orc_equilibrium(self._h, float(rho), _d(uu), _d(feq))
        return feq
