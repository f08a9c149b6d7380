"""ctypes binding of libpach.so (include/pach.h).

The library is the only compute path: if it is missing or cannot be loaded
this module raises, and nothing falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libpach.so")

PACH_OK = 0
ERRORS = {-1: "EINVAL", -2: "EHIP", -3: "ESTATE", -4: "ENOMEM", -5: "ECANCEL", -6: "ERANGE"}
PACH_ERANGE = -6
MODE_STOCHASTIC, MODE_SPECULAR = 0, 1
PREC_FP32, PREC_FP64, PREC_MIXED = 0, 1, 2
DIR_GEODESIC, DIR_RANDOM, DIR_SUPPLIED = 0, 1, 2
NPLANES = 11
PLANES = {"energy": 0, "pressure": 1, "pos_x": 2, "pos_y": 3, "pos_z": 4, "neg_x": 5,
          "neg_y": 6, "neg_z": 7, "spec": 8, "diffuse": 9, "spec_scat": 10}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Params(C.Structure):
    _fields_ = [("cutoff_m", C.c_double), ("oct_lo", C.c_int), ("oct_hi", C.c_int),
                ("is_order", C.c_int), ("mode", C.c_int), ("precision", C.c_int),
                ("async_", C.c_int), ("ray_count", C.c_int64), ("seed", C.c_uint64),
                ("p1_max_bounces", C.c_int), ("max_bounces", C.c_int), ("batch_rays", C.c_int64)]


class Progress(C.Structure):
    _fields_ = [("rays_done", C.c_int64), ("rays_total", C.c_int64), ("octave_rays", C.c_int64),
                ("running", C.c_int), ("status", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("bounces_broad", C.c_int64), ("bounces_late", C.c_int64),
                ("bounces_verylate", C.c_int64), ("octave_rays", C.c_int64),
                ("spawned", C.c_int64), ("misses", C.c_int64), ("deposits", C.c_int64),
                ("p1_cap_hits", C.c_int64), ("guard_hits", C.c_int64), ("spills", C.c_int64),
                ("resumes", C.c_int64), ("lost", C.c_double), ("e_init", C.c_double),
                ("h_oct", C.c_int), ("rays_done", C.c_int64), ("ms_phase1", C.c_double),
                ("ms_octave", C.c_double), ("launches_phase1", C.c_int64),
                ("launches_octave", C.c_int64), ("ms_trace", C.c_double),
                ("loop_steps", C.c_int64), ("voxels", C.c_int64), ("tri_tests", C.c_int64),
                ("shade_passes", C.c_int64), ("lanes_shaded", C.c_int64), ("revolutions", C.c_int64),
                ("conv_receiver_visible", C.c_int), ("conv_receiver_occluded", C.c_int), ("converged", C.c_int),
                ("exact_rewalks", C.c_int64), ("exact_walks", C.c_int64), ("ms_exact", C.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["bounces"] = d["bounces_broad"] + d["bounces_late"] + d["bounces_verylate"]
        return d


class Tuning(C.Structure):
    _fields_ = [("grid", C.c_int * 3), ("block", C.c_int), ("occupancy", C.c_int), ("shade_queue", C.c_int),
                ("walk_flat", C.c_int), ("walk_steps", C.c_int), ("shade_min", C.c_int), ("refill_min", C.c_int),
                ("dump_min", C.c_int), ("trials", C.c_int), ("rate_start", C.c_double),
                ("rate_best", C.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["grid"] = list(self.grid)
        return d


class IsStats(C.Structure):
    _fields_ = [("checks", C.c_int64), ("paths", C.c_int64), ("launches", C.c_int64), ("ms_kernel", C.c_double),
                ("ms_total", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FilterStats(C.Structure):
    _fields_ = [("channels", C.c_int64), ("samples", C.c_int64), ("launches", C.c_int64),
                ("ms_kernel", C.c_double), ("ms_total", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


CARRIER_LEN = 16384  # ETCToFilter's carrier length


# pach_is_path (one valid image-source path), as a numpy record
IS_MAX_ORDER = 8
IS_PATH = np.dtype([("receiver", "<i4"), ("order", "<i4"), ("seq", "<i4", (IS_MAX_ORDER,)),
                    ("poly", "<i4", (IS_MAX_ORDER,)), ("pts", "<f8", (IS_MAX_ORDER + 2, 3)),
                    ("length", "<f8"), ("time_s", "<f8"), ("energy", "<f8", (8,))])
assert IS_PATH.itemsize == 392


# knobs of pach_tune that an autotuning result sets
TUNED_KNOBS = ("block", "occupancy", "shade_queue", "walk_flat", "walk_steps", "shade_min", "refill_min",
               "dump_min")


# (name, restype, argtypes) of every entry point declared in include/pach.h
SIGNATURES = [
    ("pach_abi_version", C.c_int, []),
    ("pach_create", C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    ("pach_destroy", None, [C.c_void_p]),
    ("pach_last_error", C.c_char_p, [C.c_void_p]),
    ("pach_uniform_medium", C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, _dp,
                                      _dp, _dp]),
    ("pach_scene_upload", C.c_int, [C.c_void_p, _dp, C.c_int, _ip, C.c_int, _dp, _dp, _dp,
                                    C.c_int, C.c_int, C.c_int]),
    ("pach_grid_build", C.c_int, [_dp, C.c_int, C.POINTER(C.c_int), _dp, _dp, C.POINTER(C.c_uint32),
                                  C.POINTER(C.c_uint32), C.POINTER(C.c_int64), C.POINTER(C.c_uint8)]),
    ("pach_grid_face_skip", C.c_int, [_dp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_uint32)]),
    ("pach_grid_octant_cubes", C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_uint8), C.POINTER(C.c_uint8)]),
    ("pach_grid_octant_boxes", C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_uint8), C.POINTER(C.c_uint8)]),
    ("pach_source_set", C.c_int, [C.c_void_p, _dp, _dp, C.c_int, _dp, C.c_int64]),
    ("pach_receivers_set", C.c_int, [C.c_void_p, _dp, C.c_int, C.c_double, _dp, C.c_double,
                                     C.c_double, C.c_int, C.c_int]),
    ("pach_rooms_set", C.c_int, [C.c_void_p, C.c_int, _dp, _dp, C.POINTER(C.c_int64), _ip, _ip]),
    ("pach_map_receivers_set", C.c_int, [C.c_void_p, _dp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int]),
    ("pach_map_receivers_clear", C.c_int, [C.c_void_p]),
    ("pach_map_histogram", C.c_int, [C.c_void_p, C.c_int, _dp, C.c_int]),
    ("pach_tune", C.c_int, [C.c_void_p, C.c_char_p, C.c_int64]),
    ("pach_exact_histograms", C.c_int, [C.c_void_p, C.c_int]),
    ("pach_direct_sound", C.c_int, [_dp, C.c_int, _ip, _dp, C.c_int, _dp, _dp, _dp, C.c_int, _dp, C.c_double,
                                    _dp, _dp, _ip]),
    ("pach_convergence_state_size", C.c_int64, [C.c_int, C.c_int, C.c_int]),
    ("pach_convergence_step", C.c_int, [C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_int, C.c_double, _dp,
                                        C.POINTER(C.c_int)]),
    ("pach_image_sources", C.c_int, [C.c_void_p, _ip, C.c_int, C.c_int, C.c_void_p, C.c_int64,
                                     C.POINTER(C.c_int64)]),
    ("pach_image_sources_stats", C.c_int, [C.c_void_p, C.POINTER(IsStats)]),
    ("pach_etc_to_filter", C.c_int, [C.c_void_p, _dp, C.c_int, C.c_int, _dp, C.c_int, _dp]),
    ("pach_receiver_filters", C.c_int, [C.c_void_p, _dp, _dp]),
    ("pach_etc_to_filter_noalloc", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, _dp, C.c_int, C.c_void_p]),
    ("pach_filter_stats_get", C.c_int, [C.c_void_p, C.POINTER(FilterStats)]),
    ("pach_autotune", C.c_int, [C.c_void_p, C.POINTER(Params), C.c_int64, C.POINTER(Tuning)]),
    ("pach_trace", C.c_int, [C.c_void_p, C.POINTER(Params)]),
    ("pach_poll", C.c_int, [C.c_void_p, C.POINTER(Progress)]),
    ("pach_wait", C.c_int, [C.c_void_p]),
    ("pach_cancel", C.c_int, [C.c_void_p]),
    ("pach_histogram", C.c_int, [C.c_void_p, C.c_int, _dp, C.c_int]),
    ("pach_stats_get", C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    ("pach_shoot", C.c_int, [C.c_void_p, _dp, _dp, C.c_int64, C.c_int, _ip, _dp, C.POINTER(C.c_int64)]),
    ("pach_percent_lost", C.c_int, [C.c_void_p, _dp]),
    ("pach_schroeder_integral", C.c_int, [_dp, C.c_int, _dp]),
    ("pach_t_x", C.c_int, [_dp, C.c_int, C.c_int, C.c_int, _dp]),
    ("pach_early_decay_time", C.c_int, [_dp, C.c_int, C.c_int, _dp]),
    ("pach_clarity", C.c_int, [_dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, _dp]),
]

class DnConfig(C.Structure):
    _fields_ = [("bins", C.c_int), ("patch", C.c_int), ("width", C.c_int), ("blocks", C.c_int),
                ("n_desc", C.c_int), ("max_batch", C.c_int), ("beta1", C.c_float), ("beta2", C.c_float),
                ("eps", C.c_float), ("weight_decay", C.c_float)]


_fp = C.POINTER(C.c_float)
_u16p = C.POINTER(C.c_uint16)

# (name, restype, argtypes) of every entry point declared in include/pach_denoise.h
DN_SIGNATURES = [
    ("pach_dn_create", C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(DnConfig)]),
    ("pach_dn_destroy", None, [C.c_void_p]),
    ("pach_dn_last_error", C.c_char_p, [C.c_void_p]),
    ("pach_dn_param_count", C.c_int64, [C.c_void_p]),
    ("pach_dn_set_params", C.c_int, [C.c_void_p, _fp]),
    ("pach_dn_get_params", C.c_int, [C.c_void_p, _fp]),
    ("pach_dn_forward", C.c_int, [C.c_void_p, _fp, _fp, C.c_int, _fp]),
    ("pach_dn_loss_grad", C.c_int, [C.c_void_p, _fp, _fp, _fp, C.c_int, _dp, _fp]),
    ("pach_dn_set_loss_weights", C.c_int, [C.c_void_p, _fp]),
    ("pach_dn_dataset", C.c_int, [C.c_void_p, _fp, _fp, _fp, C.c_int64]),
    ("pach_dn_train_step", C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.c_int, C.c_float, _dp]),
    ("pach_dn_gemm", C.c_int, [C.c_void_p, _u16p, _u16p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _fp]),
]

_lib = None


class PachError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def lib():
    """Load libpach.so (raises if it is absent: there is no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libpach.so not built at {LIB_PATH}; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES + DN_SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _chk(rc, ctx=None):
    if rc != PACH_OK:
        msg = lib().pach_last_error(ctx).decode() if ctx else ""
        raise PachError(rc, msg)


def uniform_medium(temp_c=20.0, hr=50.0, pa_hpa=1013.25, air_choice=0, edge_correction=False):
    """(attenuation dB/m [8], speed of sound m/s, density kg/m^3) of Uniform_Medium."""
    at = np.zeros(8)
    c = C.c_double()
    rho = C.c_double()
    _chk(lib().pach_uniform_medium(air_choice, temp_c, hr, pa_hpa, int(edge_correction), _p(at),
                                   C.byref(c), C.byref(rho)))
    return at, c.value, rho.value


def direct_sound(room, power=None, atten=None, c=None):
    """pach_direct_sound for a rooms.Room: (energy [R, 8], time_s [R], audible [R])."""
    if atten is None or c is None:
        at, cc, _ = uniform_medium()
        atten = at if atten is None else atten
        c = cc if c is None else c
    v = np.ascontiguousarray(room.verts, dtype=np.float64)
    sf = np.ascontiguousarray(room.surf, dtype=np.int32)
    tr = np.ascontiguousarray(room.trans, dtype=np.float64)
    src = np.ascontiguousarray(room.source, dtype=np.float64)
    pw = np.ascontiguousarray(np.ones(8) if power is None else power, dtype=np.float64)
    ctr = np.ascontiguousarray(room.receivers, dtype=np.float64).reshape(-1, 3)
    R = ctr.shape[0]
    at = np.ascontiguousarray(atten, dtype=np.float64)
    energy = np.zeros((R, 8))
    time_s = np.zeros(R)
    aud = np.zeros(R, dtype=np.int32)
    _chk(lib().pach_direct_sound(_p(v), v.shape[0], sf.ctypes.data_as(_ip), _p(tr), tr.shape[0], _p(src), _p(pw),
                                 _p(ctr), R, _p(at), float(c), _p(energy), _p(time_s), aud.ctypes.data_as(_ip)))
    return energy, time_s, aud.astype(bool)


CONV_SPECMIN, CONV_MINIMUM = 0, 1


class ConvergenceCheck:
    """One receiver's convergence check (pach_convergence_step) driven from the host."""

    def __init__(self, kind, n, fs=1000, direct_s=0.0):
        self.kind, self.n, self.fs, self.direct_s = kind, n, fs, direct_s
        size = lib().pach_convergence_state_size(kind, n, fs)
        if size < 0:
            raise PachError(-1, "bad convergence check")
        self.state = np.zeros(size)

    def step(self, energy, spec=None, diff=None, rogue=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None
                for a in (energy, spec, diff, rogue)]
        out = C.c_int()
        _chk(lib().pach_convergence_step(self.kind, *[None if a is None else _p(a) for a in arrs], self.n, self.fs,
                                         self.direct_s, _p(self.state), C.byref(out)))
        return bool(out.value)


def schroeder_integral(etc):
    e = np.ascontiguousarray(etc, dtype=np.float64)
    out = np.zeros_like(e)
    _chk(lib().pach_schroeder_integral(_p(e), e.size, _p(out)))
    return out


def t_x(sch, decay_depth=30, sample_rate=1000):
    s = np.ascontiguousarray(sch, dtype=np.float64)
    out = C.c_double()
    _chk(lib().pach_t_x(_p(s), s.size, decay_depth, sample_rate, C.byref(out)))
    return out.value


def early_decay_time(sch, sample_rate=1000):
    s = np.ascontiguousarray(sch, dtype=np.float64)
    out = C.c_double()
    _chk(lib().pach_early_decay_time(_p(s), s.size, sample_rate, C.byref(out)))
    return out.value


def clarity(etc, sample_rate=1000.0, seconds=0.08, start_time=0.0, pressure=False):
    e = np.ascontiguousarray(etc, dtype=np.float64)
    out = C.c_double()
    _chk(lib().pach_clarity(_p(e), e.size, sample_rate, seconds, start_time, int(pressure),
                            C.byref(out)))
    return out.value


class Tracer:
    """One libpach context on one GPU (a HIP stream of its own)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _chk(lib().pach_create(C.byref(h), device))
        self.h = h
        self.R = 0
        self.sample_ct = 0

    def close(self):
        if getattr(self, "h", None):
            lib().pach_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        _chk(rc, self.h)

    def upload_scene(self, verts, surf, alpha, scat, trans=None, grid=(0, 0, 0)):
        v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1, 3, 3)
        s = np.ascontiguousarray(surf, dtype=np.int32)
        a = np.ascontiguousarray(alpha, dtype=np.float64).reshape(-1, 8)
        sc = np.ascontiguousarray(scat, dtype=np.float64).reshape(-1, 8)
        tr = None if trans is None else np.ascontiguousarray(trans, dtype=np.float64).reshape(-1, 8)
        g = grid if hasattr(grid, "__len__") else (grid, grid, grid)
        self._chk(lib().pach_scene_upload(self.h, _p(v), v.shape[0], s.ctypes.data_as(_ip),
                                          a.shape[0], _p(a), _p(sc),
                                          _p(tr) if tr is not None else None, *map(int, g)))

    def set_source(self, origin, power_w, dir_mode=DIR_GEODESIC, dirs=None):
        o = np.ascontiguousarray(origin, dtype=np.float64)
        pw = np.ascontiguousarray(power_w, dtype=np.float64)
        if dirs is not None:
            d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
            self._chk(lib().pach_source_set(self.h, _p(o), _p(pw), dir_mode, _p(d), d.shape[0]))
        else:
            self._chk(lib().pach_source_set(self.h, _p(o), _p(pw), dir_mode, None, 0))

    def set_receivers(self, centers, radius, atten, c, rho_c, sample_rate, sample_ct):
        ctr = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
        at = np.ascontiguousarray(atten, dtype=np.float64)
        self._chk(lib().pach_receivers_set(self.h, _p(ctr), ctr.shape[0], radius, _p(at), c, rho_c,
                                           sample_rate, sample_ct))
        self.R = ctr.shape[0]
        self.sample_ct = sample_ct

    @staticmethod
    def _params(rays, cutoff_m, seed=1, mode=MODE_STOCHASTIC, precision=PREC_FP32, oct_lo=0, oct_hi=7,
                is_order=-1, p1_max_bounces=128, max_bounces=100000, batch_rays=0, async_=False):
        p = Params()
        p.cutoff_m = cutoff_m
        p.oct_lo, p.oct_hi, p.is_order = oct_lo, oct_hi, is_order
        p.mode, p.precision, p.async_ = mode, precision, int(async_)
        p.ray_count, p.seed = rays, seed
        p.p1_max_bounces, p.max_bounces, p.batch_rays = p1_max_bounces, max_bounces, batch_rays
        return p

    def autotune(self, *, rays, cutoff_m, sample_rays=0, **kw):
        """pach_autotune: times sample traces and applies the fastest settings."""
        p = self._params(rays, cutoff_m, **kw)
        t = Tuning()
        self._chk(lib().pach_autotune(self.h, C.byref(p), int(sample_rays), C.byref(t)))
        return t.as_dict()

    def apply_tuning(self, tuning):
        """Re-applies an autotune result (the grid through the next scene upload)."""
        for k in TUNED_KNOBS:
            self.tune(k, tuning[k])

    def trace(self, *, rays, cutoff_m, **kw):
        p = self._params(rays, cutoff_m, **kw)
        rc = lib().pach_trace(self.h, C.byref(p))
        if rc == -5:
            return rc
        self._chk(rc)
        return rc

    def set_rooms(self, src, power, rays, surf_count, rec_count):
        s = np.ascontiguousarray(src, dtype=np.float64).reshape(-1, 3)
        pw = np.ascontiguousarray(power, dtype=np.float64).reshape(-1, 8)
        ry = np.ascontiguousarray(rays, dtype=np.int64)
        sc = np.ascontiguousarray(surf_count, dtype=np.int32)
        rc = np.ascontiguousarray(rec_count, dtype=np.int32)
        self._chk(lib().pach_rooms_set(self.h, s.shape[0], _p(s), _p(pw), ry.ctypes.data_as(C.POINTER(C.c_int64)),
                                       sc.ctypes.data_as(_ip), rc.ctypes.data_as(_ip)))

    def clear_rooms(self):
        self._chk(lib().pach_rooms_set(self.h, 0, None, None, None, None, None))

    def tune(self, key, value):
        self._chk(lib().pach_tune(self.h, key.encode(), int(value)))

    def exact_histograms(self, on=True):
        """Order-independent (bit-reproducible) histogram accumulation for later traces."""
        self._chk(lib().pach_exact_histograms(self.h, 1 if on else 0))

    def poll(self):
        pr = Progress()
        self._chk(lib().pach_poll(self.h, C.byref(pr)))
        return {k: getattr(pr, k) for k, _ in pr._fields_}

    def wait(self):
        return lib().pach_wait(self.h)

    def cancel(self):
        self._chk(lib().pach_cancel(self.h))

    def image_sources(self, max_order, object_of_surface=None, n_objects=0):
        """Valid image-source paths (pach_image_sources) of the uploaded scene,
        source and receiver centres: an IS_PATH array sorted by receiver, then
        in the reference's depth-first sequence order."""
        if object_of_surface is not None:
            obj = np.ascontiguousarray(object_of_surface, dtype=np.int32)
            optr = obj.ctypes.data_as(_ip)
            n_objects = int(n_objects) if n_objects else int(obj.max()) + 1
        else:
            optr = None
        n = C.c_int64()
        cap = 4096
        while True:
            out = np.zeros(cap, dtype=IS_PATH)
            rc = lib().pach_image_sources(self.h, optr, n_objects, int(max_order), out.ctypes.data, cap,
                                          C.byref(n))
            if rc == PACH_ERANGE:
                cap = int(n.value)
                continue
            self._chk(rc)
            return out[:n.value].copy()

    def etc_to_filter(self, prms, swl, sample_frequency):
        """Pach_SP.ETCToFilter of prms [n_ch][8][T] (or [8][T]): [n_ch][T + 16383]."""
        pr = np.ascontiguousarray(prms, dtype=np.float64)
        single = pr.ndim == 2
        pr = pr.reshape(-1, 8, pr.shape[-1])
        sw = np.ascontiguousarray(swl, dtype=np.float64)
        out = np.zeros((pr.shape[0], pr.shape[2] + CARRIER_LEN - 1))
        self._chk(lib().pach_etc_to_filter(self.h, _p(pr), pr.shape[0], pr.shape[2], _p(sw), int(sample_frequency),
                                           _p(out)))
        return out[0] if single else out

    def etc_to_filter_noalloc(self, d_prms, n_ch, T, swl, sample_frequency, d_out):
        """pach_etc_to_filter_noalloc on device buffers (addresses as integers):
        d_prms [n_ch][8][T], d_out [n_ch][T + 16383] doubles."""
        sw = np.ascontiguousarray(swl, dtype=np.float64)
        self._chk(lib().pach_etc_to_filter_noalloc(self.h, C.c_void_p(int(d_prms)), int(n_ch), int(T), _p(sw),
                                                   int(sample_frequency), C.c_void_p(int(d_out))))

    def receiver_filters(self, swl=None):
        """Create_Filter(SWL, Rho_C) of every receiver of the last trace: [R][7][SampleCT + 12287]."""
        sw = np.ascontiguousarray(np.full(8, 120.0) if swl is None else swl, dtype=np.float64)
        out = np.zeros((self.R, 7, self.sample_ct + CARRIER_LEN - 1 - 4096))
        self._chk(lib().pach_receiver_filters(self.h, _p(sw), _p(out)))
        return out

    def filter_stats(self):
        st = FilterStats()
        self._chk(lib().pach_filter_stats_get(self.h, C.byref(st)))
        return st.as_dict()

    def image_source_stats(self):
        st = IsStats()
        self._chk(lib().pach_image_sources_stats(self.h, C.byref(st)))
        return st.as_dict()

    def set_map(self, centers, increment, directional=True, sample_ct=1000, time1pt=False):
        """Mapping receivers (pach_map_receivers_set) for the next traces;
        time1pt: single-sample histograms (sample_ct 1)."""
        ctr = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
        if time1pt:
            sample_ct = 1
        self._chk(lib().pach_map_receivers_set(self.h, _p(ctr), ctr.shape[0], float(increment), int(directional),
                                               int(sample_ct), int(bool(time1pt))))
        self.map_n, self.map_sct = ctr.shape[0], int(sample_ct)

    def clear_map(self):
        self._chk(lib().pach_map_receivers_clear(self.h))
        self.map_n = 0

    def map_histogram(self, plane=0, scaled=True):
        out = np.zeros((self.map_n, 8, self.map_sct))
        self._chk(lib().pach_map_histogram(self.h, plane, _p(out), int(scaled)))
        return out

    def map_planes(self, scaled=False):
        return np.stack([self.map_histogram(k, scaled) for k in range(8)])

    def histogram(self, plane=0, scaled=True):
        out = np.zeros((self.R, 8, self.sample_ct))
        self._chk(lib().pach_histogram(self.h, plane, _p(out), int(scaled)))
        return out

    def planes(self, scaled=False):
        return np.stack([self.histogram(k, scaled) for k in range(NPLANES)])

    def stats(self):
        st = Stats()
        self._chk(lib().pach_stats_get(self.h, C.byref(st)))
        return st.as_dict()

    def shoot(self, origins, dirs, precision=PREC_FP64):
        """Polygon_Scene.shoot for each ray: (triangle [n] (-1: miss), distance [n], exact re-walks)."""
        o = np.ascontiguousarray(origins, dtype=np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        n = o.shape[0]
        tri = np.zeros(n, dtype=np.int32)
        t = np.zeros(n)
        ne = C.c_int64()
        self._chk(lib().pach_shoot(self.h, _p(o), _p(d), n, int(precision), tri.ctypes.data_as(_ip), _p(t),
                                   C.byref(ne)))
        return tri, t, ne.value

    def percent_lost(self):
        out = C.c_double()
        self._chk(lib().pach_percent_lost(self.h, C.byref(out)))
        return out.value


def trace_room(tr: Tracer, room, *, rays, cutoff_s, seed=1, mode=MODE_STOCHASTIC,
               precision=PREC_FP32, sample_rate=1000, grid=(0, 0, 0), dir_mode=DIR_GEODESIC,
               dirs=None, power_w=None, medium=None, **kw):
    """Upload a rooms.Room and trace it; returns the context's stats."""
    at, c, rho = medium if medium is not None else uniform_medium()
    tr.upload_scene(room.verts, room.surf, room.alpha, room.scat, room.trans, grid)
    tr.set_source(room.source, np.ones(8) if power_w is None else power_w, dir_mode, dirs)
    sct = int(np.floor(cutoff_s * 1000 * sample_rate / 1000))
    tr.set_receivers(room.receivers, room.radius, at, c, c * rho, sample_rate, sct)
    tr.trace(rays=rays, cutoff_m=cutoff_s * c, seed=seed, mode=mode, precision=precision, **kw)
    return tr.stats()


def image_sources_room(tr: Tracer, room, max_order, *, medium=None, power_w=None, grid=(0, 0, 0),
                       object_of_surface=None, n_objects=0):
    """Upload a rooms.Room and return its valid image-source paths (IS_PATH array)."""
    at, c, rho = medium if medium is not None else uniform_medium()
    tr.upload_scene(room.verts, room.surf, room.alpha, room.scat, room.trans, grid)
    tr.set_source(room.source, np.ones(8) if power_w is None else power_w, DIR_GEODESIC)
    tr.set_receivers(room.receivers, room.radius, at, c, c * rho, 1000, 1000)
    return tr.image_sources(max_order, object_of_surface, n_objects)


def trace_batch(tr: Tracer, packed, rays, *, cutoff_s, seed=1, mode=MODE_STOCHASTIC,
                precision=PREC_FP32, sample_rate=1000, grid=(0, 0, 0), power_w=None, medium=None, **kw):
    """Trace rooms.PackedRooms in one launch sequence; rays[i] source rays in room i.
    Returns the context's stats; per-room histograms come from batch_histograms."""
    at, c, rho = medium if medium is not None else uniform_medium()
    room = packed.room
    tr.upload_scene(room.verts, room.surf, room.alpha, room.scat, room.trans, grid)
    n = len(rays)
    pw = np.ones((n, 8)) if power_w is None else np.broadcast_to(power_w, (n, 8))
    tr.set_source(packed.sources[0], pw[0], DIR_GEODESIC)
    sct = int(np.floor(cutoff_s * 1000 * sample_rate / 1000))
    tr.set_receivers(room.receivers, room.radius, at, c, c * rho, sample_rate, sct)
    tr.set_rooms(packed.sources, pw, rays, packed.surf_count, packed.rec_count)
    try:
        tr.trace(rays=0, cutoff_m=cutoff_s * c, seed=seed, mode=mode, precision=precision, **kw)
        return tr.stats()
    finally:
        tr.clear_rooms()


def batch_histograms(tr: Tracer, packed, plane=0, scaled=True):
    """List of per-room histograms [rec_count_i, 8, sample_ct] of the last batch."""
    H = tr.histogram(plane, scaled)
    out, r0 = [], 0
    for k in packed.rec_count:
        out.append(H[r0:r0 + int(k)])
        r0 += int(k)
    return out
