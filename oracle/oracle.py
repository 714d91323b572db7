"""ctypes wrapper of the FP64 CPU oracle (oracle/shrt_oracle.cpp).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() and by
bench.py's CPU-baseline leg, always as the checker, never as the measured path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_shrt.so")
_lib = None


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("Lx", ctypes.c_double), ("Ly", ctypes.c_double),
        ("delt", ctypes.c_double), ("cfl_num", ctypes.c_double),
        ("nu", ctypes.c_double), ("kappa", ctypes.c_double),
        ("sigma", ctypes.c_double), ("g", ctypes.c_double),
        ("rho0", ctypes.c_double), ("rho0_prime", ctypes.c_double),
        ("rho_eps1", ctypes.c_double),
        ("noise", ctypes.c_int), ("linear", ctypes.c_int),
    ]


# Defaults of the reference's input_data (file_access.cu:137-163).
DEFAULTS = dict(nx=512, ny=1024, Lx=np.pi, Ly=np.pi, delt=1e-3, cfl_num=0.1,
                nu=1e-3, kappa=1e-5, sigma=1.0, g=1.0, rho0=1.0, rho0_prime=1.0,
                rho_eps1=1e-2, noise=1, linear=0)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        vp = ctypes.c_void_p
        L.orc_sizes.argtypes = [ctypes.c_int, ctypes.c_int, ip]
        L.orc_fft.argtypes = [ctypes.c_int, ctypes.c_int, dp]
        L.orc_create.argtypes = [ctypes.POINTER(OrcParams)]
        L.orc_create.restype = vp
        L.orc_destroy.argtypes = [vp]
        L.orc_set_threads.argtypes = [vp, ctypes.c_int]
        L.orc_initialize_reference.argtypes = [vp]
        L.orc_seeded_spectrum.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_uint64, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]
        L.orc_set_spectral.argtypes = [vp, ctypes.c_int, dp]
        L.orc_get_buffer.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.orc_set_buffer.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.orc_get_physical.argtypes = [vp, ctypes.c_int, dp]
        L.orc_get_kdata_bef.argtypes = [vp, ctypes.c_int, dp]
        L.orc_step.argtypes = [vp, ctypes.c_int]
        L.orc_get_state.argtypes = [vp, dp]
        L.orc_set_time.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_get_shift.argtypes = [vp, dp, ip]
        L.orc_set_shift.argtypes = [vp, dp]
        L.orc_xtok.argtypes = [vp, dp, dp]
        L.orc_ktox.argtypes = [vp, dp, dp]
        L.orc_ktox_shear.argtypes = [vp, dp, dp]
        L.orc_shearing_field.argtypes = [vp, ip, dp, dp]
        L.orc_poisson_bracket.argtypes = [vp, dp, dp]
        L.orc_check_cfl.argtypes = [vp]
        L.orc_diagnostics.argtypes = [vp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def sizes(nx, ny):
    """fft.cu:91-96 -> dict(nkx, nky, nkxh, nkxh2, nkxpad, ncy, M, I)."""
    out = np.zeros(8, dtype=np.int32)
    lib().orc_sizes(nx, ny, _ip(out))
    keys = ["nkx", "nky", "nkxh", "nkxh2", "nkxpad", "ncy", "M", "I"]
    return {k: int(v) for k, v in zip(keys, out)}


def fft(x, direction=-1):
    """Unnormalised complex FFT of a 1-D complex array (direction -1 = e^{-i})."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().orc_fft(a.size, direction, _dp(a.view(np.float64)))
    return a


def seeded_spectrum(nx, ny, Lx, Ly, seed, kmin, kmax, amp, expo):
    s = sizes(nx, ny)
    out = np.zeros(s["M"], dtype=np.complex128)
    lib().orc_seeded_spectrum(nx, ny, Lx, Ly, seed, kmin, kmax, amp, expo,
                              _dp(out.view(np.float64)))
    return out.reshape(s["nkx"], s["nky"])


class Oracle:
    """One instance of the reference algorithm in FP64 (reference layouts)."""

    RING = {"aomg": 0, "domg": 1, "arho": 2, "drho": 3, "aphi": 4}

    def __init__(self, threads=1, **kw):
        p = dict(DEFAULTS)
        p.update(kw)
        self.params = p
        self._p = OrcParams(**{k: (int(v) if k in ("nx", "ny", "noise", "linear") else float(v))
                               for k, v in p.items()})
        self.h = lib().orc_create(ctypes.byref(self._p))
        self.nx, self.ny = p["nx"], p["ny"]
        self.s = sizes(self.nx, self.ny)
        lib().orc_set_threads(self.h, threads)

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape_k(self):
        return (self.s["nkx"], self.s["nky"])

    def initialize_reference(self):
        lib().orc_initialize_reference(self.h)

    def set_spectral(self, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.complex128).reshape(-1)
        assert a.size == self.s["M"]
        lib().orc_set_spectral(self.h, {"omega": 0, "rho": 1, "phi": 2}[field],
                               _dp(a.view(np.float64)))

    def buffer(self, ring, slot=0):
        out = np.zeros(self.s["M"], dtype=np.complex128)
        lib().orc_get_buffer(self.h, self.RING[ring], slot, _dp(out.view(np.float64)))
        return out.reshape(self.shape_k)

    def set_buffer(self, ring, slot, arr):
        a = np.ascontiguousarray(arr, dtype=np.complex128).reshape(-1)
        lib().orc_set_buffer(self.h, self.RING[ring], slot, _dp(a.view(np.float64)))

    def kdata_bef(self, field):
        """omega / phi / rho as update_shear handed them to k_data_bef at the last remap."""
        out = np.zeros(self.s["M"], dtype=np.complex128)
        lib().orc_get_kdata_bef(self.h, {"omega": 0, "phi": 1, "rho": 2}[field], _dp(out.view(np.float64)))
        return out.reshape(self.shape_k)

    def physical(self, which):
        out = np.zeros(self.nx * self.ny, dtype=np.float64)
        lib().orc_get_physical(self.h, {"vx": 0, "vy": 1, "rtmp": 2}[which], _dp(out))
        return out.reshape(self.nx, self.ny)

    def step(self, n=1):
        lib().orc_step(self.h, n)

    def state(self):
        out = np.zeros(6)
        lib().orc_get_state(self.h, _dp(out))
        return dict(istep=int(out[0]), time=out[1], delt=out[2], remap_count=int(out[3]),
                    last_remap_step=int(out[4]), halvings=int(out[5]))

    def set_time(self, istep, time, delt):
        lib().orc_set_time(self.h, int(istep), float(time), float(delt))

    def shift(self):
        s = np.zeros(self.s["nkx"])
        j = np.zeros(self.s["nkx"], dtype=np.int32)
        lib().orc_get_shift(self.h, _dp(s), _ip(j))
        return s, j

    def set_shift(self, s):
        a = np.ascontiguousarray(s, dtype=np.float64)
        lib().orc_set_shift(self.h, _dp(a))

    def xtok(self, phys):
        a = np.ascontiguousarray(phys, dtype=np.float64).reshape(-1)
        out = np.zeros(self.s["M"], dtype=np.complex128)
        lib().orc_xtok(self.h, _dp(a), _dp(out.view(np.float64)))
        return out.reshape(self.shape_k)

    def ktox(self, spec):
        a = np.ascontiguousarray(spec, dtype=np.complex128).reshape(-1)
        out = np.zeros(self.nx * self.ny)
        lib().orc_ktox(self.h, _dp(a.view(np.float64)), _dp(out))
        return out.reshape(self.nx, self.ny)

    def ktox_shear(self, spec):
        a = np.ascontiguousarray(spec, dtype=np.complex128).reshape(-1)
        out = np.zeros(self.nx * self.ny)
        lib().orc_ktox_shear(self.h, _dp(a.view(np.float64)), _dp(out))
        return out.reshape(self.nx, self.ny)

    def shearing_field(self, jump, spec):
        j = np.ascontiguousarray(jump, dtype=np.int32)
        a = np.ascontiguousarray(spec, dtype=np.complex128).reshape(-1)
        out = np.zeros_like(a)
        lib().orc_shearing_field(self.h, _ip(j), _dp(a.view(np.float64)),
                                 _dp(out.view(np.float64)))
        return out.reshape(self.shape_k)

    def poisson_bracket(self, spec):
        a = np.ascontiguousarray(spec, dtype=np.complex128).reshape(-1)
        out = np.zeros(self.nx * self.ny)
        lib().orc_poisson_bracket(self.h, _dp(a.view(np.float64)), _dp(out))
        return out.reshape(self.nx, self.ny)

    def check_cfl(self):
        lib().orc_check_cfl(self.h)

    def diagnostics(self):
        out = np.zeros(3)
        lib().orc_diagnostics(self.h, _dp(out))
        return dict(energy=out[0], enstrophy=out[1], rho_var=out[2])

    def k_tables(self):
        """kx, ky exactly as fourier.cu:123-128 builds them."""
        nkx, nky, nkxh = self.s["nkx"], self.s["nky"], self.s["nkxh"]
        kx = np.zeros(nkx)
        for i in range(nkxh):
            kx[i] = i * 2 * np.pi / self.params["Lx"]
        for i in range(nkxh - 1):
            kx[nkx - i - 1] = -kx[i + 1]
        ky = np.array([i * 2 * np.pi / self.params["Ly"] for i in range(nky)])
        return kx, ky
