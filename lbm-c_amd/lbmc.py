"""ctypes binding of the C ABI in include/lbmc.h (liblbmc.so, built in-tree).

The binding mirrors the reference's host-side vocabulary: a lattice is
created from the run-time constants the reference keeps in DomainConstant
(LBM-C/data_types.cuh:23-43), fed the per-node fields ModelBuilder loads
(LBM-C/model_builder.cu:173-225) and stepped with the reference's time index
(LBM-C/index.cu:102-108).  Arrays are numpy arrays with N = Lx*Ly(*Lz)
entries in the reference's order (x fastest); a (Lz, Ly, Lx) or (Ly, Lx)
C-ordered array is that order.

There is no fallback: if liblbmc.so is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbmc.so")

# enums of include/lbmc.h
D2Q9, D3Q15 = 0, 1
BGK, NTPOR, MRT, MRTPOR = 0, 1, 2, 3
FP64, FP32, FP16S, BF16S = 0, 1, 2, 3
FIX_D3Q15_MINV, FIX_D3Q15_MEQ, FIX_D3Q15_ZH = 0x1, 0x2, 0x4
STORAGE_NAMES = {FP64: "f64", FP32: "f32", FP16S: "f16s", BF16S: "bf16s"}
SLOT_BYTES = {FP64: 8, FP32: 4, FP16S: 2, BF16S: 2}

STATUS = {0: "LBMC_OK", 1: "LBMC_ERR_INVALID", 2: "LBMC_ERR_NOMEM", 3: "LBMC_ERR_DEVICE",
          4: "LBMC_ERR_STATE", 5: "LBMC_ERR_UNSUPPORTED"}

EXPORTS = [
    "lbmc_create", "lbmc_destroy", "lbmc_last_error", "lbmc_set_fields", "lbmc_get_node_types",
    "lbmc_get_links", "lbmc_init_static", "lbmc_init_equilibrium", "lbmc_step", "lbmc_run",
    "lbmc_time_steps", "lbmc_get_macros", "lbmc_get_node", "lbmc_get_f", "lbmc_set_f",
    "lbmc_get_f_noalloc",
    "lbmc_kinetic_energy", "lbmc_sync", "lbmc_info", "lbmc_copy_bandwidth",
    "lbmc_mlp_create", "lbmc_mlp_destroy", "lbmc_mlp_last_error", "lbmc_mlp_set_params",
    "lbmc_mlp_get_params", "lbmc_mlp_set_scaling", "lbmc_mlp_forward", "lbmc_mlp_set_eval_fp16",
    "lbmc_mlp_loss_grad",
    "lbmc_mlp_set_data", "lbmc_mlp_train", "lbmc_set_closure", "lbmc_set_closure_interval",
    "lbmc_set_closure_shift",
    "lbmc_batch_create", "lbmc_batch_destroy", "lbmc_batch_last_error", "lbmc_batch_run",
    "lbmc_batch_time_steps", "lbmc_batch_autotune", "lbmc_batch_set_shape", "lbmc_batch_info",
]

MLP_NPARAM = 4737


class LbmcError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


class Desc(ct.Structure):
    _fields_ = [
        ("lattice", ct.c_int),
        ("length", ct.c_int * 3),
        ("collision_type", ct.c_int),
        ("forcing", ct.c_int),
        ("macro_bc", ct.c_int),
        ("micro_bc", ct.c_int),
        ("tau", ct.c_double),
        ("tau_mrt", ct.c_double * 15),
        ("c_smag", ct.c_double),
        ("storage", ct.c_int),
        ("quirks", ct.c_uint),
        ("device", ct.c_int),
    ]


class MlpScaling(ct.Structure):
    _fields_ = [("mu", ct.c_float * 7), ("isd", ct.c_float * 7), ("out_a", ct.c_float),
                ("out_b", ct.c_float)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "(make -C lbm-c_amd)")
    lib = ct.CDLL(LIB_PATH)
    P = ct.c_void_p
    dpp = ct.POINTER(ct.POINTER(ct.c_double))
    fp = ct.POINTER(ct.c_float)
    sig = {
        "lbmc_create": (ct.c_int, [ct.POINTER(Desc), ct.POINTER(P)]),
        "lbmc_destroy": (ct.c_int, [P]),
        "lbmc_last_error": (ct.c_char_p, [P]),
        "lbmc_set_fields": (ct.c_int, [P, ct.POINTER(ct.c_double), ct.POINTER(ct.c_int),
                                       ct.POINTER(ct.c_int), dpp, dpp, ct.POINTER(ct.c_double)]),
        "lbmc_get_node_types": (ct.c_int, [P, ct.POINTER(ct.c_uint8), ct.POINTER(ct.c_uint16),
                                           ct.POINTER(ct.c_int)]),
        "lbmc_get_links": (ct.c_int, [P, ct.POINTER(ct.c_uint16)]),
        "lbmc_init_static": (ct.c_int, [P]),
        "lbmc_init_equilibrium": (ct.c_int, [P, ct.POINTER(ct.c_double), dpp]),
        "lbmc_step": (ct.c_int, [P, ct.c_longlong, ct.c_int]),
        "lbmc_run": (ct.c_int, [P, ct.c_longlong, ct.c_int, ct.c_int]),
        "lbmc_time_steps": (ct.c_int, [P, ct.c_longlong, ct.c_int, ct.c_int,
                                       ct.POINTER(ct.c_float)]),
        "lbmc_get_macros": (ct.c_int, [P, ct.POINTER(ct.c_double), dpp]),
        "lbmc_get_node": (ct.c_int, [P, ct.c_longlong, ct.POINTER(ct.c_double),
                                     ct.POINTER(ct.c_double)]),
        "lbmc_get_f": (ct.c_int, [P, dpp]),
        "lbmc_set_f": (ct.c_int, [P, dpp]),
        "lbmc_get_f_noalloc": (ct.c_int, [P, dpp]),
        "lbmc_kinetic_energy": (ct.c_int, [P, ct.POINTER(ct.c_double)]),
        "lbmc_sync": (ct.c_int, [P]),
        "lbmc_info": (ct.c_int, [P, ct.POINTER(ct.c_longlong), ct.POINTER(ct.c_int),
                                 ct.POINTER(ct.c_int)]),
        "lbmc_copy_bandwidth": (ct.c_int, [ct.c_int, ct.c_size_t, ct.c_int,
                                           ct.POINTER(ct.c_float)]),
        "lbmc_mlp_create": (ct.c_int, [ct.c_int, fp, ct.POINTER(MlpScaling), ct.POINTER(P)]),
        "lbmc_mlp_destroy": (None, [P]),
        "lbmc_mlp_last_error": (ct.c_char_p, [P]),
        "lbmc_mlp_set_params": (ct.c_int, [P, fp]),
        "lbmc_mlp_get_params": (ct.c_int, [P, fp]),
        "lbmc_mlp_set_scaling": (ct.c_int, [P, ct.POINTER(MlpScaling)]),
        "lbmc_mlp_forward": (ct.c_int, [P, fp, ct.c_longlong, fp]),
        "lbmc_mlp_set_eval_fp16": (ct.c_int, [P, ct.c_int]),
        "lbmc_mlp_loss_grad": (ct.c_int, [P, fp, fp, ct.c_longlong, fp, ct.POINTER(ct.c_double)]),
        "lbmc_mlp_set_data": (ct.c_int, [P, fp, fp, ct.c_longlong]),
        "lbmc_mlp_train": (ct.c_int, [P, ct.c_int, ct.c_int, ct.c_float, ct.c_ulonglong,
                                      ct.POINTER(ct.c_double)]),
        "lbmc_set_closure": (ct.c_int, [P, P, ct.c_double, ct.c_double]),
        "lbmc_set_closure_interval": (ct.c_int, [P, ct.c_int]),
        "lbmc_set_closure_shift": (ct.c_int, [P, ct.c_double]),
        "lbmc_batch_create": (ct.c_int, [ct.POINTER(P), ct.c_int, ct.POINTER(P)]),
        "lbmc_batch_destroy": (ct.c_int, [P]),
        "lbmc_batch_last_error": (ct.c_char_p, [P]),
        "lbmc_batch_run": (ct.c_int, [P, ct.c_longlong, ct.c_int, ct.c_int]),
        "lbmc_batch_time_steps": (ct.c_int, [P, ct.c_longlong, ct.c_int, ct.c_int, fp]),
        "lbmc_batch_autotune": (ct.c_int, [P, ct.c_longlong, ct.c_int, ct.POINTER(ct.c_int), fp]),
        "lbmc_batch_set_shape": (ct.c_int, [P, ct.c_int, ct.c_int, ct.c_int]),
        "lbmc_batch_info": (ct.c_int, [P, ct.POINTER(ct.c_longlong), ct.POINTER(ct.c_int),
                                       ct.POINTER(ct.c_int), ct.POINTER(ct.c_int), ct.POINTER(ct.c_int)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = _load()


def lib():
    return _lib


def _dptr(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


def _ptr_array(arrays, n):
    """double*[n] from a list of float64 arrays (None -> NULL)."""
    arr = (ct.POINTER(ct.c_double) * n)()
    for k in range(n):
        a = arrays[k] if arrays is not None and k < len(arrays) else None
        arr[k] = _dptr(a) if a is not None else ct.POINTER(ct.c_double)()
    return ct.cast(arr, ct.POINTER(ct.POINTER(ct.c_double)))


def _f64(a, n):
    if a is None:
        return None
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64)).reshape(-1)
    if a.size != n:
        raise ValueError(f"array has {a.size} entries, the lattice has {n} nodes")
    return a


def _i32(a, n):
    if a is None:
        return None
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32)).reshape(-1)
    if a.size != n:
        raise ValueError(f"array has {a.size} entries, the lattice has {n} nodes")
    return a


class Lattice:
    """One lattice on one GPU: the device-side state of an LBM-C run."""

    def __init__(self, lattice=D3Q15, length=(16, 16, 16), collision_type=BGK, forcing=False,
                 macro_bc=False, micro_bc=False, tau=1.0, tau_mrt=None, c_smag=0.0,
                 storage=FP32, quirks=0, device=0):
        self.lattice = lattice
        self.dim = 2 if lattice == D2Q9 else 3
        self.q = 9 if lattice == D2Q9 else 15
        L = list(length) + [1] * (3 - len(length))
        self.length = tuple(int(v) for v in L[: self.dim])
        self.n = int(np.prod(self.length))
        self.storage = storage
        d = Desc()
        d.lattice = lattice
        for k in range(3):
            d.length[k] = int(L[k]) if k < self.dim else 1
        d.collision_type = collision_type
        d.forcing = int(bool(forcing))
        d.macro_bc = int(bool(macro_bc))
        d.micro_bc = int(bool(micro_bc))
        d.tau = float(tau)
        if tau_mrt is not None:
            for k, v in enumerate(tau_mrt):
                d.tau_mrt[k] = float(v)
        d.c_smag = float(c_smag)
        d.storage = storage
        d.quirks = quirks
        d.device = device
        self.desc = d
        self._h = ct.c_void_p()
        st = _lib.lbmc_create(ct.byref(d), ct.byref(self._h))
        if st != 0:
            raise LbmcError(st, _lib.lbmc_last_error(None).decode())

    # -- plumbing
    def _check(self, st):
        if st != 0:
            raise LbmcError(st, _lib.lbmc_last_error(self._h).decode())

    def close(self):
        if self._h:
            _lib.lbmc_destroy(self._h)
            self._h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- setup
    def set_fields(self, geometry=None, macro_bc=None, micro_bc=None, force=None, u=None,
                   rho=None):
        n = self.n
        g = _f64(geometry, n)
        ma = _i32(macro_bc, n)
        mi = _i32(micro_bc, n)
        fo = [_f64(v, n) for v in force] if force is not None else None
        uu = [_f64(v, n) for v in u] if u is not None else None
        r = _f64(rho, n)
        self._keep = (g, ma, mi, fo, uu, r)
        self._check(_lib.lbmc_set_fields(
            self._h,
            _dptr(g) if g is not None else None,
            ma.ctypes.data_as(ct.POINTER(ct.c_int)) if ma is not None else None,
            mi.ctypes.data_as(ct.POINTER(ct.c_int)) if mi is not None else None,
            _ptr_array(fo, 3) if fo is not None else None,
            _ptr_array(uu, 3) if uu is not None else None,
            _dptr(r) if r is not None else None))

    def init_static(self):
        self._check(_lib.lbmc_init_static(self._h))

    def init_equilibrium(self, rho, u):
        n = self.n
        r = _f64(rho, n)
        uu = [_f64(v, n) for v in u]
        self._check(_lib.lbmc_init_equilibrium(self._h, _dptr(r), _ptr_array(uu, 3)))

    # -- stepping
    def step(self, t, store_macros=True):
        self._check(_lib.lbmc_step(self._h, int(t), int(bool(store_macros))))

    def run(self, t0, nsteps, store_every=1):
        self._check(_lib.lbmc_run(self._h, int(t0), int(nsteps), int(store_every)))

    def time_steps(self, t0, nsteps, store_every=0):
        ms = ct.c_float()
        self._check(_lib.lbmc_time_steps(self._h, int(t0), int(nsteps), int(store_every),
                                         ct.byref(ms)))
        return float(ms.value)

    def sync(self):
        self._check(_lib.lbmc_sync(self._h))

    # -- read-back
    def get_macros(self):
        rho = np.empty(self.n)
        u = [np.empty(self.n) for _ in range(self.dim)]
        self._check(_lib.lbmc_get_macros(self._h, _dptr(rho), _ptr_array(u, 3)))
        return rho, np.stack(u)

    def get_node(self, idx):
        rho = ct.c_double()
        u = (ct.c_double * 3)()
        self._check(_lib.lbmc_get_node(self._h, int(idx), ct.byref(rho), u))
        return rho.value, np.array(u[: self.dim])

    def get_f(self):
        f = [np.empty(self.n) for _ in range(self.q)]
        self._check(_lib.lbmc_get_f(self._h, _ptr_array(f, 15)))
        return np.stack(f)

    def get_f_noalloc(self):
        """get_f through the library's fixed device scratch (no allocation)."""
        f = [np.empty(self.n) for _ in range(self.q)]
        self._check(_lib.lbmc_get_f_noalloc(self._h, _ptr_array(f, 15)))
        return np.stack(f)

    def set_f(self, f):
        f = [_f64(v, self.n) for v in f]
        self._check(_lib.lbmc_set_f(self._h, _ptr_array(f, 15)))

    def node_types(self):
        mask = np.empty(self.n, dtype=np.uint8)
        table = np.zeros(256, dtype=np.uint16)
        nc = ct.c_int()
        self._check(_lib.lbmc_get_node_types(
            self._h, mask.ctypes.data_as(ct.POINTER(ct.c_uint8)),
            table.ctypes.data_as(ct.POINTER(ct.c_uint16)), ct.byref(nc)))
        return mask, table[: nc.value].copy()

    def links(self):
        out = np.empty(self.n, dtype=np.uint16)
        self._check(_lib.lbmc_get_links(self._h, out.ctypes.data_as(ct.POINTER(ct.c_uint16))))
        return out

    def kinetic_energy(self):
        e = ct.c_double()
        self._check(_lib.lbmc_kinetic_energy(self._h, ct.byref(e)))
        return e.value

    # -- learned closure
    def set_closure(self, model, tau_lo=0.5005, tau_hi=2.0):
        """Relax plain fluid nodes with the model's tau (Smagorinsky outside
        [tau_lo, tau_hi]); model=None turns the closure off."""
        self._closure_model = model
        h = model._h if model is not None else None
        self._check(_lib.lbmc_set_closure(self._h, h, float(tau_lo), float(tau_hi)))

    def set_closure_interval(self, every):
        """Evaluate the closure network every `every` calls (reusing each
        node's recorded tau in between)."""
        self._check(_lib.lbmc_set_closure_interval(self._h, int(every)))

    def set_closure_shift(self, dtau):
        """Add dtau to the network's tau' on this lattice (before the band check)."""
        self._check(_lib.lbmc_set_closure_shift(self._h, float(dtau)))


def copy_bandwidth(nbytes, reps=20, device=0):
    """Time a plain device copy of nbytes (read nbytes + write nbytes) reps times; ms total."""
    ms = ct.c_float()
    st = _lib.lbmc_copy_bandwidth(int(device), ct.c_size_t(int(nbytes)), int(reps), ct.byref(ms))
    if st != 0:
        raise LbmcError(st, _lib.lbmc_last_error(None).decode())
    return float(ms.value)


def _fptr(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_float))


def _f32(a, shape_tail=()):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


class Mlp:
    """The closure network on one device: FP32 master weights, Adam state and
    the operand images the kernels read (bf16; an fp16 evaluation image for
    set_eval_fp16).

    params: LBMC_MLP_NPARAM floats (w1[64][8] with the bias in column 7,
    w2[64][64], b2[64], w3[64], b3).  Inputs are normalised as
    (feature - mu) * isd; the closure's tau is out_a * y + out_b."""

    def __init__(self, params, mu=None, isd=None, out_a=1.0, out_b=0.0, device=0):
        p = _f32(params).reshape(-1)
        if p.size != MLP_NPARAM:
            raise ValueError(f"{p.size} parameters, expected {MLP_NPARAM}")
        self.device = device
        self._h = ct.c_void_p()
        sc = self._scaling(mu, isd, out_a, out_b)
        st = _lib.lbmc_mlp_create(int(device), _fptr(p), ct.byref(sc), ct.byref(self._h))
        if st != 0:
            raise LbmcError(st, _lib.lbmc_mlp_last_error(None).decode())

    @staticmethod
    def _scaling(mu, isd, out_a, out_b):
        sc = MlpScaling()
        mu = np.zeros(7) if mu is None else np.asarray(mu, dtype=np.float64)
        isd = np.ones(7) if isd is None else np.asarray(isd, dtype=np.float64)
        for k in range(7):
            sc.mu[k] = float(mu[k])
            sc.isd[k] = float(isd[k])
        sc.out_a = float(out_a)
        sc.out_b = float(out_b)
        return sc

    def _check(self, st):
        if st != 0:
            raise LbmcError(st, _lib.lbmc_mlp_last_error(self._h).decode())

    def close(self):
        if self._h:
            _lib.lbmc_mlp_destroy(self._h)
            self._h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_scaling(self, mu=None, isd=None, out_a=1.0, out_b=0.0):
        sc = self._scaling(mu, isd, out_a, out_b)
        self._check(_lib.lbmc_mlp_set_scaling(self._h, ct.byref(sc)))

    def set_params(self, params):
        p = _f32(params).reshape(-1)
        if p.size != MLP_NPARAM:
            raise ValueError(f"{p.size} parameters, expected {MLP_NPARAM}")
        self._check(_lib.lbmc_mlp_set_params(self._h, _fptr(p)))

    def get_params(self):
        p = np.empty(MLP_NPARAM, dtype=np.float32)
        self._check(_lib.lbmc_mlp_get_params(self._h, _fptr(p)))
        return p

    def set_eval_fp16(self, on=True):
        """Evaluate with fp16 operands instead of bf16 (forward, and the
        closure steps of lattices that attach the model afterwards)."""
        self._check(_lib.lbmc_mlp_set_eval_fp16(self._h, 1 if on else 0))

    def forward(self, feat):
        x = _f32(feat).reshape(-1, 7)
        y = np.empty(x.shape[0], dtype=np.float32)
        self._check(_lib.lbmc_mlp_forward(self._h, _fptr(x), x.shape[0], _fptr(y)))
        return y

    def loss_grad(self, feat, target):
        x = _f32(feat).reshape(-1, 7)
        t = _f32(target).reshape(-1)
        g = np.empty(MLP_NPARAM, dtype=np.float32)
        loss = ct.c_double()
        self._check(_lib.lbmc_mlp_loss_grad(self._h, _fptr(x), _fptr(t), x.shape[0], _fptr(g),
                                            ct.byref(loss)))
        return loss.value, g

    def set_data(self, feat, target):
        x = _f32(feat).reshape(-1, 7)
        t = _f32(target).reshape(-1)
        if t.size != x.shape[0]:
            raise ValueError("one target per sample")
        self._check(_lib.lbmc_mlp_set_data(self._h, _fptr(x), _fptr(t), x.shape[0]))

    def train(self, epochs, batch=16384, lr=1e-3, seed=0):
        out = np.zeros(max(epochs, 1))
        self._check(_lib.lbmc_mlp_train(self._h, int(epochs), int(batch), float(lr), int(seed),
                                        out.ctypes.data_as(ct.POINTER(ct.c_double))))
        return out[:epochs]


class Batch:
    """Several Lattice objects (same lattice type, storage and device) stepped
    together by one persistent launch per step.  The members keep their own
    read-back methods; they must stay open while the batch is used."""

    def __init__(self, members):
        self.members = list(members)
        arr = (ct.c_void_p * len(self.members))(*[m._h.value for m in self.members])
        self._h = ct.c_void_p()
        st = _lib.lbmc_batch_create(arr, len(self.members), ct.byref(self._h))
        if st != 0:
            raise LbmcError(st, _lib.lbmc_batch_last_error(None).decode())

    def _check(self, st):
        if st != 0:
            raise LbmcError(st, _lib.lbmc_batch_last_error(self._h).decode())

    def close(self):
        if self._h:
            _lib.lbmc_batch_destroy(self._h)
            self._h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, t0, nsteps, store_every=1):
        self._check(_lib.lbmc_batch_run(self._h, int(t0), int(nsteps), int(store_every)))

    def time_steps(self, t0, nsteps, store_every=0):
        ms = ct.c_float()
        self._check(_lib.lbmc_batch_time_steps(self._h, int(t0), int(nsteps), int(store_every),
                                               ct.byref(ms)))
        return float(ms.value)

    def autotune(self, t0, steps_per_trial=20):
        g = ct.c_int()
        ms = ct.c_float()
        self._check(_lib.lbmc_batch_autotune(self._h, int(t0), int(steps_per_trial), ct.byref(g),
                                             ct.byref(ms)))
        return g.value, float(ms.value)

    def set_shape(self, grid, chunk_nodes=256, blocks_per_cu=0):
        self._check(_lib.lbmc_batch_set_shape(self._h, int(grid), int(chunk_nodes), int(blocks_per_cu)))

    def info(self):
        n = ct.c_longlong()
        g = ct.c_int()
        c = ct.c_int()
        cn = ct.c_int()
        bpc = ct.c_int()
        self._check(_lib.lbmc_batch_info(self._h, ct.byref(n), ct.byref(g), ct.byref(c), ct.byref(cn),
                                         ct.byref(bpc)))
        return {"nodes": n.value, "grid": g.value, "chunks": c.value, "chunk_nodes": cn.value,
                "blocks_per_cu": bpc.value}

    def shape(self):
        """(grid, chunk_nodes, blocks_per_cu): the launch shape, e.g. after autotune."""
        i = self.info()
        return i["grid"], i["chunk_nodes"], i["blocks_per_cu"]
