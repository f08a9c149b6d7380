"""ctypes front end of the CPU FP64 oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

NPLANES = 11
PLANE_NAMES = ("E", "P", "px", "py", "pz", "nx", "ny", "nz", "spec", "diff", "spec_scat")
MODE_STOCHASTIC, MODE_SPECULAR = 0, 1
DIR_GEODESIC, DIR_RANDOM, DIR_SUPPLIED = 0, 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Params(C.Structure):
    _fields_ = [("cutoff_m", C.c_double), ("oct_lo", C.c_int), ("oct_hi", C.c_int),
                ("is_order", C.c_int), ("ray_begin", C.c_int64), ("ray_end", C.c_int64),
                ("seed", C.c_uint64), ("mode", C.c_int), ("dir_mode", C.c_int),
                ("dirs", _dp), ("n_dirs", C.c_int64), ("p1_max_bounces", C.c_int),
                ("max_bounces", C.c_int), ("use_grid", C.c_int)]


class Setup(C.Structure):
    _fields_ = [("src", C.c_double * 3), ("power", C.c_double * 8), ("R", C.c_int),
                ("centers", _dp), ("radius", C.c_double), ("c", C.c_double),
                ("atten", C.c_double * 8), ("rho_c", C.c_double), ("sample_rate", C.c_int),
                ("sample_ct", C.c_int), ("map", C.c_void_p), ("map_planes", C.POINTER(_dp))]


class Counters(C.Structure):
    _fields_ = [("bounces_broad", C.c_int64), ("bounces_late", C.c_int64),
                ("bounces_verylate", C.c_int64), ("spawned", C.c_int64),
                ("misses", C.c_int64), ("deposits", C.c_int64), ("octave_rays", C.c_int64),
                ("p1_cap_hits", C.c_int64), ("guard_hits", C.c_int64), ("lost", C.c_double),
                ("e_init", C.c_double), ("h_oct", C.c_int), ("absorbed", C.c_double),
                ("remaining", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RayOut(C.Structure):
    _fields_ = [("d", C.c_double * 3), ("I", C.c_double), ("mask", C.c_int), ("band", C.c_int)]

    def as_tuple(self):
        return np.array(self.d[:]), self.I, self.mask, self.band


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.or_philox4x32_10.argtypes = [C.POINTER(C.c_uint32)] * 3
        L.or_mix64.restype = C.c_uint64
        L.or_mix64.argtypes = [C.c_uint64]
        L.or_child_uid.restype = C.c_uint64
        L.or_child_uid.argtypes = [C.c_uint64, C.c_uint64]
        L.or_attenuation.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, _dp]
        L.or_density.restype = C.c_double
        L.or_density.argtypes = [C.c_double] * 3
        L.or_sound_speed.restype = C.c_double
        L.or_sound_speed.argtypes = [C.c_double]
        L.or_scene_create.restype = C.c_void_p
        L.or_scene_create.argtypes = [_dp, C.c_int, _ip, C.c_int, _dp, _dp, _dp,
                                      C.c_int, C.c_int, C.c_int]
        L.or_scene_destroy.argtypes = [C.c_void_p]
        L.or_shoot_brute.restype = C.c_int
        L.or_shoot_brute.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.or_shoot_brute_n.argtypes = [C.c_void_p, _dp, _dp, C.c_int64, _ip, _dp]
        L.or_shoot_grid.restype = C.c_int
        L.or_shoot_grid.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.or_surface_area.restype = C.c_double
        L.or_surface_area.argtypes = [C.c_void_p, C.c_int]
        L.or_geodesic_faces.argtypes = [_dp]
        L.or_source_direction.argtypes = [C.c_int, C.c_uint64, C.c_int64, _dp, C.c_int64, _dp]
        L.or_trace.restype = C.c_int
        L.or_trace.argtypes = [C.c_void_p, C.POINTER(Setup), C.POINTER(Params),
                               C.POINTER(_dp), C.POINTER(Counters)]
        L.or_trace_mt.restype = C.c_int
        L.or_trace_mt.argtypes = [C.c_void_p, C.POINTER(Setup), C.POINTER(Params),
                                  C.POINTER(_dp), C.POINTER(Counters), C.c_int]
        L.or_h_oct.restype = C.c_int
        L.or_h_oct.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int]
        for name in ("or_schroeder",):
            getattr(L, name).argtypes = [_dp, C.c_int, _dp]
        L.or_log10data.argtypes = [_dp, C.c_int, C.c_double, _dp]
        L.or_t_x.restype = C.c_double
        L.or_t_x.argtypes = [_dp, C.c_int, C.c_int, C.c_int]
        L.or_edt.restype = C.c_double
        L.or_edt.argtypes = [_dp, C.c_int, C.c_int]
        L.or_clarity.restype = C.c_double
        L.or_clarity.argtypes = [_dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int]
        L.or_receiver_check.argtypes = [C.POINTER(Setup), C.POINTER(_dp), _dp, _dp, _dp,
                                        C.c_double, C.c_int, C.c_double, C.c_int,
                                        C.POINTER(C.c_int64)]
        L.or_sample_diffuse.argtypes = [_dp, C.c_double, C.c_double, _dp]
        L.or_sincos_turn.argtypes = [C.c_double, _dp, _dp]
        L.or_sabine.argtypes = [C.c_void_p, C.c_double, _dp, _dp]
        L.or_eyring.argtypes = [C.c_void_p, C.c_double, _dp, _dp]
        L.or_image_sources.restype = C.c_int64
        L.or_image_sources.argtypes = [C.c_void_p, _ip, C.c_int, C.c_int, _dp, _dp, _dp, C.c_int, _dp,
                                       C.c_double, C.c_double, C.c_void_p, C.c_int64]
        L.or_map_create.restype = C.c_void_p
        L.or_map_create.argtypes = [_dp, C.c_int, C.c_double, _dp, C.c_double, C.c_double, _dp, C.c_int, C.c_int]
        L.or_map_destroy.argtypes = [C.c_void_p]
        L.or_map_set_time1pt.argtypes = [C.c_void_p, C.c_int]
        L.or_map_grid.restype = C.c_int
        L.or_map_grid.argtypes = [C.c_void_p, _ip, _dp, _dp]
        L.or_map_voxel.restype = C.c_int
        L.or_map_voxel.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _ip, C.c_int]
        L.or_map_check_octave.argtypes = [C.c_void_p, C.POINTER(_dp), _dp, _dp, _dp, C.c_double, C.c_int,
                                          C.c_double, C.POINTER(C.c_int64)]
        L.or_map_check_broadband.argtypes = [C.c_void_p, C.POINTER(_dp), _dp, _dp, _dp, _dp, C.c_double,
                                             C.POINTER(C.c_int64)]
        L.or_lambert_bounds.restype = C.c_int
        L.or_lambert_bounds.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp]
        _u32p = C.POINTER(C.c_uint32)
        L.or_scatter_late_once.restype = C.c_int
        L.or_scatter_late_once.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_int, C.c_int,
                                           _u32p, C.POINTER(RayOut), C.POINTER(RayOut)]
        L.or_scatter_verylate_once.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_int, C.c_int,
                                               _u32p, C.POINTER(RayOut)]
        L.or_scatter_early_once.restype = C.c_int
        L.or_scatter_early_once.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _u32p, _dp, _dp,
                                            C.POINTER(RayOut), C.c_int]
        L.or_image_sources_count_mt.restype = C.c_int64
        L.or_image_sources_count_mt.argtypes = [C.c_void_p, _ip, C.c_int, C.c_int, _dp, _dp, _dp, C.c_int, _dp,
                                                C.c_double, C.c_double, C.c_int, C.c_int, C.c_int]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return list(o)


def attenuation(air_choice=0, temp_c=20.0, hr=50.0, pa_hpa=1013.25, edge=False):
    out = np.zeros(8)
    lib().or_attenuation(air_choice, temp_c, hr, pa_hpa, int(edge), _p(out))
    return out


def density(temp_c=20.0, hr=50.0, pa_hpa=1013.25):
    return lib().or_density(temp_c, hr, pa_hpa)


def sound_speed(temp_c=20.0):
    return lib().or_sound_speed(temp_c)


@dataclass
class Medium:
    c: float
    atten: np.ndarray
    rho: float

    @property
    def rho_c(self):
        return self.c * self.rho


def standard_medium(temp_c=20.0, hr=50.0, pa_hpa=1013.25, air_choice=0):
    return Medium(c=sound_speed(temp_c), atten=attenuation(air_choice, temp_c, hr, pa_hpa),
                  rho=density(temp_c, hr, pa_hpa))


IS_MAX_ORDER = 8
IS_PATH = np.dtype([("rec", "<i4"), ("order", "<i4"), ("seq", "<i4", (IS_MAX_ORDER,)),
                    ("poly", "<i4", (IS_MAX_ORDER,)), ("pts", "<f8", (IS_MAX_ORDER + 2, 3)),
                    ("length", "<f8"), ("time", "<f8"), ("energy", "<f8", (8,))])


class Scene:
    """Oracle scene (triangles + per-polygon materials + its own voxel grid)."""

    def __init__(self, room, grid=(16, 16, 16)):
        self._v = np.ascontiguousarray(room.verts, dtype=np.float64)
        self._s = np.ascontiguousarray(room.surf, dtype=np.int32)
        self._a = np.ascontiguousarray(room.alpha, dtype=np.float64)
        self._sc = np.ascontiguousarray(room.scat, dtype=np.float64)
        self._t = np.ascontiguousarray(room.trans, dtype=np.float64)
        g = grid if hasattr(grid, "__len__") else (grid, grid, grid)
        self.h = lib().or_scene_create(_p(self._v), self._v.shape[0],
                                       self._s.ctypes.data_as(_ip), self._a.shape[0],
                                       _p(self._a), _p(self._sc), _p(self._t), *map(int, g))
        if not self.h:
            raise ValueError("oracle scene creation failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_scene_destroy(self.h)
            self.h = None

    def shoot(self, o, d, grid=True):
        o = np.ascontiguousarray(o, dtype=np.float64)
        d = np.ascontiguousarray(d, dtype=np.float64)
        t = np.zeros(1)
        f = lib().or_shoot_grid if grid else lib().or_shoot_brute
        tri = f(self.h, _p(o), _p(d), _p(t))
        return tri, (float(t[0]) if tri >= 0 else None)

    def shoot_brute_many(self, origins, dirs):
        """or_shoot_brute for every ray: (triangle [n] (-1: miss), distance [n])."""
        o = np.ascontiguousarray(origins, dtype=np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        n = o.shape[0]
        tri = np.zeros(n, dtype=np.int32)
        t = np.zeros(n)
        lib().or_shoot_brute_n(self.h, _p(o), _p(d), n, tri.ctypes.data_as(_ip), _p(t))
        return tri, t

    def surface_area(self, s):
        return lib().or_surface_area(self.h, s)

    def h_oct(self, power, oct_lo=0, oct_hi=7):
        pw = np.ascontiguousarray(power, dtype=np.float64)
        return lib().or_h_oct(self.h, _p(pw), oct_lo, oct_hi)

    def image_sources(self, max_order, src, power, receivers, medium, obj_of_surf=None, n_obj=None):
        """Valid specular paths of every sequence of 1..max_order objects
        (ImageSourceData, planar objects): a structured array of IS_PATH, in
        the reference's depth-first sequence order, receivers inner."""
        src = np.ascontiguousarray(src, dtype=np.float64)
        pw = np.ascontiguousarray(power, dtype=np.float64)
        rec = np.ascontiguousarray(receivers, dtype=np.float64).reshape(-1, 3)
        at = np.ascontiguousarray(medium.atten, dtype=np.float64)
        S = self._a.shape[0]
        if obj_of_surf is None:
            obj_ptr, n_obj = None, S
        else:
            obj = np.ascontiguousarray(obj_of_surf, dtype=np.int32)
            obj_ptr, n_obj = obj.ctypes.data_as(_ip), int(n_obj if n_obj is not None else obj.max() + 1)
        args = (self.h, obj_ptr, n_obj, int(max_order), _p(src), _p(pw), _p(rec), rec.shape[0], _p(at),
                float(medium.c), float(medium.rho_c))
        n = lib().or_image_sources(*args, None, 0)
        if n < 0:
            raise ValueError("image sources: bad arguments")
        out = np.zeros(n, dtype=IS_PATH)
        lib().or_image_sources(*args, out.ctypes.data, n)
        return out

    def image_source_count(self, max_order, src, power, receivers, medium, q0_begin, q0_end, threads=1):
        """Valid paths of the sequences whose first object is in [q0_begin, q0_end) (polygons as objects)."""
        src = np.ascontiguousarray(src, dtype=np.float64)
        pw = np.ascontiguousarray(power, dtype=np.float64)
        rec = np.ascontiguousarray(receivers, dtype=np.float64).reshape(-1, 3)
        at = np.ascontiguousarray(medium.atten, dtype=np.float64)
        n = lib().or_image_sources_count_mt(self.h, None, self._a.shape[0], int(max_order), _p(src), _p(pw),
                                            _p(rec), rec.shape[0], _p(at), float(medium.c), float(medium.rho_c),
                                            int(q0_begin), int(q0_end), int(threads))
        if n < 0:
            raise ValueError("image sources: bad arguments")
        return int(n)

    def lambert_bounds(self, surf, band):
        """(lo, hi) of Lambert_Scattering for polygon surf, band (Classes_Materials.cs:4463-4485)."""
        lo, hi = C.c_double(), C.c_double()
        if lib().or_lambert_bounds(self.h, surf, band, C.byref(lo), C.byref(hi)) != 0:
            raise ValueError("bad polygon or band")
        return lo.value, hi.value

    def scatter_late(self, sf, band, d, N, I, w, mask=0, transmit=False):
        """One Scatter_Late event; returns (parent, child or None) as (d, I, mask, band)."""
        d, N = (np.ascontiguousarray(v, dtype=np.float64) for v in (d, N))
        ww = (C.c_uint32 * 4)(*[int(x) for x in w])
        par, ch = RayOut(), RayOut()
        n = lib().or_scatter_late_once(self.h, sf, band, _p(d), _p(N), float(I), int(mask), int(transmit), ww,
                                       C.byref(par), C.byref(ch))
        return par.as_tuple(), (ch.as_tuple() if n else None)

    def scatter_verylate(self, sf, band, d, N, I, w, mask=0, transmit=False):
        d, N = (np.ascontiguousarray(v, dtype=np.float64) for v in (d, N))
        ww = (C.c_uint32 * 4)(*[int(x) for x in w])
        out = RayOut()
        lib().or_scatter_verylate_once(self.h, sf, band, _p(d), _p(N), float(I), int(mask), int(transmit), ww,
                                       C.byref(out))
        return out.as_tuple()

    def scatter_early(self, sf, d, N, E, w):
        """One Scatter_Early event of a broadband ray with energies E (after Absorb):
        returns (outgoing direction, remaining energies, children in enqueue order)."""
        d, N, E = (np.ascontiguousarray(v, dtype=np.float64) for v in (d, N, E))
        ww = (C.c_uint32 * 16)(*[int(x) for x in w])
        d_out, E_out = np.zeros(3), np.zeros(8)
        ch = (RayOut * 24)()
        n = lib().or_scatter_early_once(self.h, sf, _p(d), _p(N), _p(E), ww, _p(d_out), _p(E_out), ch, 24)
        return d_out, E_out, [ch[i].as_tuple() for i in range(n)]

    def sabine(self, volume, atten):
        out = np.zeros(8)
        a = np.ascontiguousarray(atten, dtype=np.float64)
        lib().or_sabine(self.h, volume, _p(a), _p(out))
        return out

    def eyring(self, volume, atten):
        out = np.zeros(8)
        a = np.ascontiguousarray(atten, dtype=np.float64)
        lib().or_eyring(self.h, volume, _p(a), _p(out))
        return out


def source_direction(dir_mode, seed, src_id, dirs=None):
    d = np.zeros(3)
    if dirs is not None:
        dd = np.ascontiguousarray(dirs, dtype=np.float64)
        lib().or_source_direction(dir_mode, seed, src_id, _p(dd), dd.shape[0], _p(d))
    else:
        lib().or_source_direction(dir_mode, seed, src_id, None, 0, _p(d))
    return d


def geodesic_faces():
    out = np.zeros((80, 3, 3))
    lib().or_geodesic_faces(_p(out))
    return out


class Map:
    """Mapping receivers (or_map): spheres of diameter `increment` at `centers`."""

    def __init__(self, centers, increment, src, medium, directional=True, sample_ct=1000, time1pt=False):
        self.centers = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
        s = np.ascontiguousarray(src, dtype=np.float64)
        at = np.ascontiguousarray(medium.atten, dtype=np.float64)
        self.M, self.sample_ct = self.centers.shape[0], 1 if time1pt else int(sample_ct)
        self.h = lib().or_map_create(_p(self.centers), self.M, float(increment), _p(s), float(medium.c),
                                     float(medium.rho_c), _p(at), int(directional), self.sample_ct)
        if not self.h:
            raise ValueError("oracle map creation failed")
        if time1pt:
            lib().or_map_set_time1pt(self.h, 1)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_map_destroy(self.h)
            self.h = None

    def grid(self):
        ct = (C.c_int * 3)()
        mn, vd = np.zeros(3), np.zeros(3)
        lib().or_map_grid(self.h, ct, _p(mn), _p(vd))
        return tuple(ct), mn, vd

    def voxel(self, x, y, z):
        buf = (C.c_int * self.M)()
        n = lib().or_map_voxel(self.h, x, y, z, buf, self.M)
        return list(buf[:n])

    def planes(self):
        return np.zeros((8, self.M, 8, self.sample_ct))

    @staticmethod
    def _ptrs(pl):
        ptrs = (_dp * 8)()
        for k in range(8):
            ptrs[k] = _p(pl[k])
        return ptrs

    def check_octave(self, pl, o, d, end, I, band, t_sum):
        dep = C.c_int64()
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (o, d, end)]
        lib().or_map_check_octave(self.h, self._ptrs(pl), _p(a[0]), _p(a[1]), _p(a[2]), float(I), int(band),
                                  float(t_sum), C.byref(dep))
        return dep.value

    def check_broadband(self, pl, o, d, end, E, t_sum):
        dep = C.c_int64()
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (o, d, end, E)]
        lib().or_map_check_broadband(self.h, self._ptrs(pl), _p(a[0]), _p(a[1]), _p(a[2]), _p(a[3]),
                                     float(t_sum), C.byref(dep))
        return dep.value


def trace(room, *, rays, cutoff_s, seed=1, mode=MODE_STOCHASTIC, dir_mode=DIR_GEODESIC,
          dirs=None, oct_lo=0, oct_hi=7, is_order=-1, sample_rate=1000, sample_ct=None,
          power=None, medium=None, grid=(16, 16, 16), use_grid=True, ray_begin=0,
          p1_max_bounces=128, max_bounces=100000, threads=1, planes=NPLANES, scene=None, map_rx=None):
    """Trace ``rays`` source rays; returns (planes[11, R, 8, sample_ct], counters dict).
    With ``map_rx`` (a Map) the receivers are the map's and the planes returned
    are the map's [8, M, 8, map sample_ct].

    Histograms are raw sums (not divided by the ray count), i.e. what the
    receivers hold before Combine_ThreadLocal_Results' Scale."""
    med = medium or standard_medium()
    if sample_ct is None:
        sample_ct = int(np.floor(cutoff_s * 1000 * sample_rate / 1000))
    sc = scene or Scene(room, grid)
    R = room.receivers.shape[0]
    su = Setup()
    su.src[:] = list(map(float, room.source))
    pw = np.ones(8) if power is None else np.asarray(power, dtype=np.float64)
    su.power[:] = list(pw)
    centers = np.ascontiguousarray(room.receivers, dtype=np.float64)
    su.R = R
    su.centers = _p(centers)
    su.radius = room.radius
    su.c = med.c
    su.atten[:] = list(med.atten)
    su.rho_c = med.rho_c
    su.sample_rate = sample_rate
    su.sample_ct = sample_ct
    p = Params()
    p.cutoff_m = cutoff_s * med.c
    p.oct_lo, p.oct_hi, p.is_order = oct_lo, oct_hi, is_order
    p.ray_begin, p.ray_end = ray_begin, ray_begin + rays
    p.seed = seed
    p.mode, p.dir_mode = mode, dir_mode
    dd = None
    if dirs is not None:
        dd = np.ascontiguousarray(dirs, dtype=np.float64)
        p.dirs = _p(dd)
        p.n_dirs = dd.shape[0]
    p.p1_max_bounces, p.max_bounces, p.use_grid = p1_max_bounces, max_bounces, int(use_grid)
    map_out = None
    if map_rx is not None:
        map_out = map_rx.planes()
        map_ptrs = Map._ptrs(map_out)
        su.map = map_rx.h
        su.map_planes = map_ptrs
    out = np.zeros((NPLANES, R, 8, sample_ct))
    ptrs = (_dp * NPLANES)()
    for k in range(NPLANES):
        ptrs[k] = _p(out[k]) if k < planes else None
    cnt = Counters()
    if threads > 1:
        rc = lib().or_trace_mt(sc.h, C.byref(su), C.byref(p), ptrs, C.byref(cnt), threads)
    else:
        rc = lib().or_trace(sc.h, C.byref(su), C.byref(p), ptrs, C.byref(cnt))
    if rc != 0:
        raise RuntimeError(f"or_trace failed: {rc}")
    if map_out is not None:
        return map_out, cnt.as_dict()
    return out, cnt.as_dict()


def scale(planes, n_rays):
    """Combine_ThreadLocal_Results -> Receiver_Bank.Scale (Classes_Receivers.cs:1774-1786,
    :1088-1102): energy, direction and analysis planes / N, pressure / sqrt(N)."""
    out = planes / float(n_rays)
    out[1] = planes[1] / np.sqrt(float(n_rays))
    return out


def schroeder(etc):
    e = np.ascontiguousarray(etc, dtype=np.float64)
    out = np.zeros_like(e)
    lib().or_schroeder(_p(e), e.size, _p(out))
    return out


def t_x(sch, depth=30, fs=1000):
    s = np.ascontiguousarray(sch, dtype=np.float64)
    return lib().or_t_x(_p(s), s.size, depth, fs)


def edt(sch, fs=1000):
    s = np.ascontiguousarray(sch, dtype=np.float64)
    return lib().or_edt(_p(s), s.size, fs)


def clarity(etc, fs=1000.0, seconds=0.08, start=0.0, pressure=False):
    e = np.ascontiguousarray(etc, dtype=np.float64)
    return lib().or_clarity(_p(e), e.size, fs, seconds, start, int(pressure))


def log10data(x, limit_db=-60.0):
    a = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(a)
    lib().or_log10data(_p(a), a.size, limit_db, _p(out))
    return out


def make_setup(receivers, radius, medium=None, sample_rate=1000, sample_ct=2000, src=(0, 0, 0),
               power=None):
    """Setup record for single-call helpers; keeps its arrays alive on the object."""
    med = medium or standard_medium()
    su = Setup()
    su.src[:] = list(map(float, src))
    su.power[:] = list(np.ones(8) if power is None else power)
    su._centers = np.ascontiguousarray(receivers, dtype=np.float64).reshape(-1, 3)
    su.R = su._centers.shape[0]
    su.centers = _p(su._centers)
    su.radius = radius
    su.c = med.c
    su.atten[:] = list(med.atten)
    su.rho_c = med.rho_c
    su.sample_rate = sample_rate
    su.sample_ct = sample_ct
    return su


def receiver_check(su, o, d, end, I, band, t_sum=0.0, ana=0):
    """One CheckRay over every receiver; returns (planes[11, R, 8, n], deposits)."""
    out = np.zeros((NPLANES, su.R, 8, su.sample_ct))
    ptrs = (_dp * NPLANES)(*[_p(out[k]) for k in range(NPLANES)])
    o, d, end = (np.ascontiguousarray(v, dtype=np.float64) for v in (o, d, end))
    dep = C.c_int64()
    lib().or_receiver_check(C.byref(su), ptrs, _p(o), _p(d), _p(end), I, band, t_sum, ana,
                            C.byref(dep))
    return out, dep.value


def sample_diffuse(N, u_az, u_zen):
    n = np.ascontiguousarray(N, dtype=np.float64)
    out = np.zeros(3)
    lib().or_sample_diffuse(_p(n), u_az, u_zen, _p(out))
    return out


def sincos_turn(u):
    s, c = C.c_double(), C.c_double()
    lib().or_sincos_turn(u, C.byref(s), C.byref(c))
    return s.value, c.value
