"""ctypes binding of libshrt_hip.so (include/shrt.h) for tests, bench and tools.

Mirrors the reference's source-level interface (time_integral.h, fft.h, shear.h):
`Shrt.initialize()` is initialize(), `Shrt.time_advance(n)` is time_advance
called n times, `xtok` / `ktox` are fft.h's transforms, `shearing_field` is the
remap of shear.cu.  Arrays are numpy, in the reference layouts:
spectral (nkx, nky) complex, physical (nx, ny) real.

The library must be built (make -C shearing-rt_amd); there is no fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SHRT_LIB", os.path.join(HERE, "libshrt_hip.so"))
REPO = os.path.dirname(HERE)
HEADER = os.path.join(REPO, "include", "shrt.h")

FP32, FP64 = 4, 8
OMEGA, PHI, RHO = 0, 1, 2
P_OMEGA, P_PHI, P_RHO, P_VX, P_VY, P_BRACKET = 0, 1, 2, 3, 4, 5
RING = {"aomg": 0, "domg": 1, "arho": 2, "drho": 3}

_lib = None


class ShrtError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("Lx", ctypes.c_double), ("Ly", ctypes.c_double),
        ("delt", ctypes.c_double), ("cfl_num", ctypes.c_double),
        ("nu", ctypes.c_double), ("kappa", ctypes.c_double),
        ("sigma", ctypes.c_double), ("g", ctypes.c_double),
        ("rho0", ctypes.c_double), ("rho0_prime", ctypes.c_double),
        ("rho_eps1", ctypes.c_double),
        ("noise", ctypes.c_int), ("linear", ctypes.c_int),
        ("precision", ctypes.c_int), ("graph_steps", ctypes.c_int),
    ]


class SeededIC(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("kmin", ctypes.c_double), ("kmax", ctypes.c_double),
                ("amp", ctypes.c_double), ("expo", ctypes.c_double)]


class Time(ctypes.Structure):
    _fields_ = [("istep", ctypes.c_int64), ("time", ctypes.c_double), ("delt", ctypes.c_double),
                ("remap_count", ctypes.c_int), ("halvings", ctypes.c_int),
                ("last_remap_step", ctypes.c_int64)]


class Diag(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("enstrophy", ctypes.c_double), ("rho_var", ctypes.c_double)]


class Tuning(ctypes.Structure):
    _fields_ = [("p1", ctypes.c_int), ("p2", ctypes.c_int), ("p3", ctypes.c_int), ("ms", ctypes.c_double * 8),
                ("fuse", ctypes.c_int), ("step_ms", ctypes.c_double * 2)]


class Halving(ctypes.Structure):
    _fields_ = [("istep", ctypes.c_int64), ("time", ctypes.c_double), ("value", ctypes.c_double),
                ("delt", ctypes.c_double), ("which", ctypes.c_int), ("count", ctypes.c_int),
                ("counts", ctypes.c_int * 4), ("values", ctypes.c_double * 4), ("delt_before", ctypes.c_double)]


# Every entry point of include/shrt.h with its ctypes signature.
_VP = ctypes.c_void_p
_SIGS = {
    "shrt_create": ([ctypes.POINTER(Params), ctypes.POINTER(_VP)], ctypes.c_int),
    "shrt_destroy": ([_VP], ctypes.c_int),
    "shrt_initialize": ([_VP], ctypes.c_int),
    "shrt_initialize_seeded": ([_VP, ctypes.POINTER(SeededIC)], ctypes.c_int),
    "shrt_step": ([_VP, ctypes.c_int], ctypes.c_int),
    "shrt_step_async": ([_VP, ctypes.c_int], ctypes.c_int),
    "shrt_synchronize": ([_VP], ctypes.c_int),
    "shrt_get_spectral": ([_VP, ctypes.c_int, _VP], ctypes.c_int),
    "shrt_set_spectral": ([_VP, ctypes.c_int, _VP], ctypes.c_int),
    "shrt_get_spectral_unremapped": ([_VP, ctypes.c_int, _VP], ctypes.c_int),
    "shrt_get_spectral_modes": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, _VP], ctypes.c_int),
    "shrt_xtok_noalloc": ([_VP, _VP, _VP], ctypes.c_int),
    "shrt_get_buffer": ([_VP, ctypes.c_int, ctypes.c_int, _VP], ctypes.c_int),
    "shrt_set_buffer": ([_VP, ctypes.c_int, ctypes.c_int, _VP], ctypes.c_int),
    "shrt_get_physical": ([_VP, ctypes.c_int, ctypes.c_int, _VP], ctypes.c_int),
    "shrt_xtok": ([_VP, _VP, _VP], ctypes.c_int),
    "shrt_ktox": ([_VP, _VP, _VP], ctypes.c_int),
    "shrt_ktox_shear": ([_VP, _VP, _VP], ctypes.c_int),
    "shrt_autotune": ([_VP, ctypes.c_int, ctypes.POINTER(Tuning)], ctypes.c_int),
    "shrt_set_variants": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "shrt_get_variants": ([_VP, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                           ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "shrt_set_fusion": ([_VP, ctypes.c_int], ctypes.c_int),
    "shrt_set_update_precision": ([_VP, ctypes.c_int], ctypes.c_int),
    "shrt_ensemble_create": ([ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int,
                              ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "shrt_ensemble_destroy": ([_VP], ctypes.c_int),
    "shrt_ensemble_step": ([_VP, ctypes.c_int], ctypes.c_int),
    "shrt_ensemble_synchronize": ([_VP], ctypes.c_int),
    "shrt_ensemble_time_steps": ([_VP, ctypes.c_int, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "shrt_closure_create": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                             ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "shrt_closure_destroy": ([_VP], ctypes.c_int),
    "shrt_closure_info": ([_VP, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "shrt_closure_get_params": ([_VP, _VP], ctypes.c_int),
    "shrt_closure_set_params": ([_VP, _VP], ctypes.c_int),
    "shrt_closure_forward": ([_VP, _VP, ctypes.c_int, _VP], ctypes.c_int),
    "shrt_closure_gradients": ([_VP, _VP, _VP, ctypes.c_int, ctypes.POINTER(ctypes.c_double), _VP], ctypes.c_int),
    "shrt_closure_set_data": ([_VP, _VP, _VP, ctypes.c_int], ctypes.c_int),
    "shrt_closure_train": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint64, _VP], ctypes.c_int),
    "shrt_closure_correct": ([_VP, _VP, _VP, ctypes.c_double, _VP, ctypes.c_double, ctypes.c_double,
                              ctypes.c_int, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "shrt_shearing_field": ([_VP, _VP, _VP, _VP], ctypes.c_int),
    "shrt_get_time": ([_VP, ctypes.POINTER(Time)], ctypes.c_int),
    "shrt_set_time": ([_VP, ctypes.c_int64, ctypes.c_double, ctypes.c_double], ctypes.c_int),
    "shrt_get_shift": ([_VP, _VP, _VP], ctypes.c_int),
    "shrt_set_shift": ([_VP, _VP], ctypes.c_int),
    "shrt_diagnostics": ([_VP, ctypes.POINTER(Diag)], ctypes.c_int),
    "shrt_get_halvings": ([_VP, ctypes.POINTER(Halving), ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
                          ctypes.c_int),
    "shrt_time_steps": ([_VP, ctypes.c_int, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "shrt_kernel_times": ([_VP, ctypes.c_int, _VP], ctypes.c_int),
    "shrt_fused_kernel_time": ([_VP, ctypes.c_int, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "shrt_copy_bench": ([ctypes.c_longlong, ctypes.c_int, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "shrt_stream_copy_bench": ([ctypes.c_longlong, ctypes.c_int, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "shrt_sizes": ([ctypes.c_int, ctypes.c_int, _VP], ctypes.c_int),
    "shrt_supported": ([ctypes.c_int, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "shrt_device_count": ([], ctypes.c_int),
    "shrt_last_error": ([], ctypes.c_char_p),
    "shrt_version": ([], ctypes.c_char_p),
}


def lib():
    """Load libshrt_hip.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ShrtError(f"{LIB_PATH} is missing: build it with `make -C shearing-rt_amd`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _check(st):
    if st != 0:
        raise ShrtError(f"status {st}: {lib().shrt_last_error().decode()}")


def sizes(nx, ny):
    out = np.zeros(8, dtype=np.int32)
    _check(lib().shrt_sizes(nx, ny, out.ctypes.data))
    return dict(zip(["nkx", "nky", "nkxh", "nkxh2", "nkxpad", "ncy", "M", "I"], map(int, out)))


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class Shrt:
    """One simulation context (one instance of the reference's global state)."""

    DEFAULTS = dict(nx=512, ny=1024, Lx=np.pi, Ly=np.pi, delt=1e-3, cfl_num=0.1, nu=1e-3,
                    kappa=1e-5, sigma=1.0, g=1.0, rho0=1.0, rho0_prime=1.0, rho_eps1=1e-2,
                    noise=1, linear=0, precision=FP64, graph_steps=0)

    def __init__(self, **kw):
        p = dict(self.DEFAULTS)
        p.update(kw)
        self.params = p
        ints = ("nx", "ny", "noise", "linear", "precision", "graph_steps")
        self._p = Params(**{k: (int(v) if k in ints else float(v)) for k, v in p.items()})
        h = ctypes.c_void_p()
        _check(lib().shrt_create(ctypes.byref(self._p), ctypes.byref(h)))
        self.h = h
        self.nx, self.ny = p["nx"], p["ny"]
        self.s = sizes(self.nx, self.ny)
        self.real = np.float32 if p["precision"] == FP32 else np.float64
        self.cplx = np.complex64 if p["precision"] == FP32 else np.complex128

    def close(self):
        if getattr(self, "h", None):
            lib().shrt_destroy(self.h)
            self.h = None

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
    def shape_k(self):
        return (self.s["nkx"], self.s["nky"])

    # --- lifecycle ---------------------------------------------------------
    def initialize(self):
        _check(lib().shrt_initialize(self.h))

    def initialize_seeded(self, seed, kmin, kmax, amp, expo):
        ic = SeededIC(seed, kmin, kmax, amp, expo)
        _check(lib().shrt_initialize_seeded(self.h, ctypes.byref(ic)))

    def time_advance(self, n=1):
        _check(lib().shrt_step(self.h, int(n)))

    def step_async(self, n):
        _check(lib().shrt_step_async(self.h, int(n)))

    def synchronize(self):
        _check(lib().shrt_synchronize(self.h))

    # --- state access ------------------------------------------------------
    def spectral(self, field):
        out = np.zeros(self.shape_k, dtype=self.cplx)
        _check(lib().shrt_get_spectral(self.h, field, _ptr(out)))
        return out

    def spectral_modes(self, field, nx_c, ny_c=None):
        """The field on the retained modes of a coarser grid (a spectral filter on the device)."""
        sc = sizes(nx_c, ny_c or nx_c)
        out = np.zeros((sc["nkx"], sc["nky"]), dtype=self.cplx)
        _check(lib().shrt_get_spectral_modes(self.h, field, nx_c, ny_c or nx_c, _ptr(out)))
        return out

    def spectral_unremapped(self, field):
        """The field as stored, before a pending remap (k_data_bef at a remap step)."""
        out = np.zeros(self.shape_k, dtype=self.cplx)
        _check(lib().shrt_get_spectral_unremapped(self.h, field, _ptr(out)))
        return out

    def set_spectral(self, field, arr):
        a = np.ascontiguousarray(arr, dtype=self.cplx).reshape(self.shape_k)
        _check(lib().shrt_set_spectral(self.h, field, _ptr(a)))

    def buffer(self, ring, slot=0):
        out = np.zeros(self.shape_k, dtype=self.cplx)
        _check(lib().shrt_get_buffer(self.h, RING[ring], slot, _ptr(out)))
        return out

    def set_buffer(self, ring, slot, arr):
        a = np.ascontiguousarray(arr, dtype=self.cplx).reshape(self.shape_k)
        _check(lib().shrt_set_buffer(self.h, RING[ring], slot, _ptr(a)))

    def physical(self, pfield, lab_frame=0):
        out = np.zeros((self.nx, self.ny), dtype=self.real)
        _check(lib().shrt_get_physical(self.h, pfield, lab_frame, _ptr(out)))
        return out

    def xtok(self, phys):
        a = np.ascontiguousarray(phys, dtype=self.real).reshape(self.nx, self.ny)
        out = np.zeros(self.shape_k, dtype=self.cplx)
        _check(lib().shrt_xtok(self.h, _ptr(a), _ptr(out)))
        return out

    def ktox(self, spec):
        a = np.ascontiguousarray(spec, dtype=self.cplx).reshape(self.shape_k)
        out = np.zeros((self.nx, self.ny), dtype=self.real)
        _check(lib().shrt_ktox(self.h, _ptr(a), _ptr(out)))
        return out

    def ktox_shear(self, spec):
        a = np.ascontiguousarray(spec, dtype=self.cplx).reshape(self.shape_k)
        out = np.zeros((self.nx, self.ny), dtype=self.real)
        _check(lib().shrt_ktox_shear(self.h, _ptr(a), _ptr(out)))
        return out

    def shearing_field(self, jump, spec):
        j = np.ascontiguousarray(jump, dtype=np.int32)
        a = np.ascontiguousarray(spec, dtype=self.cplx).reshape(self.shape_k)
        out = np.zeros_like(a)
        _check(lib().shrt_shearing_field(self.h, _ptr(j), _ptr(a), _ptr(out)))
        return out

    def state(self):
        t = Time()
        _check(lib().shrt_get_time(self.h, ctypes.byref(t)))
        return dict(istep=t.istep, time=t.time, delt=t.delt, remap_count=t.remap_count,
                    halvings=t.halvings, last_remap_step=t.last_remap_step)

    def set_time(self, istep, time, delt):
        _check(lib().shrt_set_time(self.h, int(istep), float(time), float(delt)))

    def shift(self):
        s = np.zeros(self.s["nkx"])
        j = np.zeros(self.s["nkx"], dtype=np.int32)
        _check(lib().shrt_get_shift(self.h, _ptr(s), _ptr(j)))
        return s, j

    def set_shift(self, s):
        a = np.ascontiguousarray(s, dtype=np.float64)
        _check(lib().shrt_set_shift(self.h, _ptr(a)))

    def diagnostics(self):
        d = Diag()
        _check(lib().shrt_diagnostics(self.h, ctypes.byref(d)))
        return dict(energy=d.energy, enstrophy=d.enstrophy, rho_var=d.rho_var)

    def time_steps(self, n):
        """Milliseconds (HIP events on the context stream) for n steps."""
        ms = ctypes.c_double()
        _check(lib().shrt_time_steps(self.h, int(n), ctypes.byref(ms)))
        return ms.value

    def kernel_times(self, n):
        """Average milliseconds per launch of the three step kernels over n steps."""
        out = np.zeros(3)
        _check(lib().shrt_kernel_times(self.h, int(n), _ptr(out)))
        return out

    def fused_kernel_time(self, n):
        """Average milliseconds per launch of the fused P3+P1 kernel over n launches."""
        ms = ctypes.c_double()
        _check(lib().shrt_fused_kernel_time(self.h, int(n), ctypes.byref(ms)))
        return ms.value

    def autotune(self, reps=20):
        t = Tuning()
        _check(lib().shrt_autotune(self.h, int(reps), ctypes.byref(t)))
        return dict(p1=t.p1, p2=t.p2, p3=t.p3, ms=list(t.ms), fuse=t.fuse, step_ms=list(t.step_ms))

    def set_variants(self, p1, p2, p3):
        _check(lib().shrt_set_variants(self.h, int(p1), int(p2), int(p3)))

    def variants(self):
        """The pass variants and fusion flag in use: dict(p1, p2, p3, fuse)."""
        v = [ctypes.c_int() for _ in range(4)]
        _check(lib().shrt_get_variants(self.h, *[ctypes.byref(x) for x in v]))
        return dict(p1=v[0].value, p2=v[1].value, p3=v[2].value, fuse=v[3].value)

    def set_update_precision(self, wide):
        """FP32 contexts: state update in FP64 with stochastic rounding (True, default) or FP32 arithmetic."""
        _check(lib().shrt_set_update_precision(self.h, 1 if wide else 0))

    def set_fusion(self, on):
        """Run captured batches with P3 of step n and P1 of step n+1 fused (default off)."""
        _check(lib().shrt_set_fusion(self.h, 1 if on else 0))

    def halvings(self, cap=256):
        arr = (Halving * cap)()
        n = ctypes.c_int()
        _check(lib().shrt_get_halvings(self.h, arr, cap, ctypes.byref(n)))
        return [dict(istep=arr[i].istep, time=arr[i].time, value=arr[i].value, delt=arr[i].delt,
                     which=arr[i].which, count=arr[i].count, counts=list(arr[i].counts),
                     values=list(arr[i].values), delt_before=arr[i].delt_before) for i in range(n.value)]


class Ensemble:
    """Independent contexts advanced together, one captured graph per batch of
    `steps_per_graph` steps (shrt_ensemble_*)."""

    def __init__(self, members, steps_per_graph=8):
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        h = ctypes.c_void_p()
        _check(lib().shrt_ensemble_create(arr, len(self.members), int(steps_per_graph), ctypes.byref(h)))
        self.h = h

    def step(self, n):
        _check(lib().shrt_ensemble_step(self.h, int(n)))
        _check(lib().shrt_ensemble_synchronize(self.h))

    def time_steps(self, n):
        ms = ctypes.c_double()
        _check(lib().shrt_ensemble_time_steps(self.h, int(n), ctypes.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            lib().shrt_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Closure:
    """The learned spectral-convolution closure (shrt_closure_*): inputs
    [batch][4][n*n] (omega, rho, phi, shear rate), outputs [batch][2][n*n]."""

    def __init__(self, n, kmax, width=32, layers=4, max_batch=8, seed=1):
        h = ctypes.c_void_p()
        _check(lib().shrt_closure_create(int(n), int(kmax), int(width), int(layers), int(max_batch), int(seed),
                                         ctypes.byref(h)))
        self.h = h
        self.n, self.kmax, self.width, self.layers, self.max_batch = n, kmax, width, layers, max_batch
        npar, nm = ctypes.c_longlong(), ctypes.c_int()
        _check(lib().shrt_closure_info(self.h, ctypes.byref(npar), ctypes.byref(nm)))
        self.nparams, self.nmodes = npar.value, nm.value

    def params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        _check(lib().shrt_closure_get_params(self.h, _ptr(p)))
        return p

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        assert a.size == self.nparams
        _check(lib().shrt_closure_set_params(self.h, _ptr(a)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        B = x.shape[0]
        assert x.shape == (B, 4, self.n * self.n)
        out = np.zeros((B, 2, self.n * self.n), dtype=np.float32)
        _check(lib().shrt_closure_forward(self.h, _ptr(x), B, _ptr(out)))
        return out

    def gradients(self, x, target):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = ctypes.c_double()
        _check(lib().shrt_closure_gradients(self.h, _ptr(x), _ptr(t), x.shape[0], ctypes.byref(loss), _ptr(g)))
        return loss.value, g

    def set_data(self, x, target):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        assert x.shape[0] == t.shape[0]
        _check(lib().shrt_closure_set_data(self.h, _ptr(x), _ptr(t), x.shape[0]))

    def train(self, iters, batch, lr, seed=0):
        losses = np.zeros(int(iters))
        _check(lib().shrt_closure_train(self.h, int(iters), int(batch), float(lr), int(seed), _ptr(losses)))
        return losses

    def correct(self, ctx, scale_in, sigma, scale_out, dt_tot, guard, kcut=0):
        """Correct the solver context `ctx` (a Shrt) in place on the device;
        returns False when the fallback left it unchanged."""
        si = np.ascontiguousarray(scale_in, dtype=np.float64)
        so = np.ascontiguousarray(scale_out, dtype=np.float64)
        assert si.shape == (3,) and so.shape == (2,)
        applied = ctypes.c_int()
        _check(lib().shrt_closure_correct(self.h, ctx.h, _ptr(si), float(sigma), _ptr(so), float(dt_tot),
                                          float(guard), int(kcut), ctypes.byref(applied)))
        return bool(applied.value)

    def close(self):
        if getattr(self, "h", None):
            lib().shrt_closure_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def header_symbols():
    """Function names declared in include/shrt.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(shrt_[a-z_0-9]+)\s*\(", txt)))
