"""Host-side checks of libpach.so that need no GPU: symbol exports, argument
validation, the medium and decay-curve entry points, and the voxel grid the
scene upload builds."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pyoracle as O
from pachydermacoustic_universal_amd import _abi as A
from pachydermacoustic_universal_amd import rooms

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for h in sorted(os.listdir(inc)):
        if h.endswith(".h"):
            text = re.sub(r"/\*.*?\*/", "", open(os.path.join(inc, h)).read(), flags=re.S)
            names |= set(re.findall(r"\b(pach_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    lib = A.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    # and the ctypes table covers every declaration with a signature
    assert set(n for n, _, _ in A.SIGNATURES + A.DN_SIGNATURES) == set(names)


def test_abi_version():
    assert A.lib().pach_abi_version() == 1


def test_null_arguments_are_rejected():
    lib = A.lib()
    assert lib.pach_create(None, 0) == -1
    assert lib.pach_trace(None, None) == -1
    assert lib.pach_poll(None, None) == -1
    assert lib.pach_histogram(None, 0, None, 0) == -1
    assert lib.pach_schroeder_integral(None, 3, None) == -1
    out = np.zeros(8)
    assert lib.pach_uniform_medium(7, 20, 50, 1013.25, 0, out.ctypes.data_as(C.POINTER(C.c_double)),
                                   None, None) == -1


def gpu_count():
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        return 0
    n = C.c_int(0)
    return n.value if hip.hipGetDeviceCount(C.byref(n)) == 0 else 0


def test_create_without_gpu_fails_cleanly():
    if gpu_count() > 0:
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert A.lib().pach_create(C.byref(h), 0) == -2
    assert not h.value


def test_uniform_medium_matches_oracle():
    for air, T, hr, pa, edge in [(0, 20, 50, 1013.25, 0), (0, 30, 70, 1000.0, 0), (1, 25, 40, 1013.25, 0),
                                 (2, 20, 60, 1013.25, 1), (2, 10, 80, 1013.25, 0)]:
        at, c, rho = A.uniform_medium(T, hr, pa, air, edge)
        np.testing.assert_allclose(at, O.attenuation(air, T, hr, pa, edge), rtol=1e-12, atol=0)
        assert c == O.sound_speed(T)
        assert abs(rho - O.density(T, hr, pa)) < 1e-14


def test_decay_math_matches_oracle():
    rng = np.random.default_rng(3)
    for _ in range(5):
        n = 3000
        etc = np.exp(-rng.uniform(5, 30) * np.arange(n) / 1000) * rng.uniform(0.5, 1.5, n)
        sch = A.schroeder_integral(etc)
        np.testing.assert_allclose(sch, O.schroeder(etc), rtol=1e-13, atol=1e-16)
        assert abs(A.t_x(sch, 30, 1000) - O.t_x(sch, 30, 1000)) < 1e-12
        assert abs(A.t_x(sch, 20, 1000) - O.t_x(sch, 20, 1000)) < 1e-12
        assert abs(A.early_decay_time(sch, 1000) - O.edt(sch, 1000)) < 1e-12
        assert abs(A.clarity(etc, 1000.0, 0.08, 0.0) - O.clarity(etc, 1000.0, 0.08, 0.0)) < 1e-12
        assert abs(A.clarity(etc, 1000.0, 0.05, 0.013) - O.clarity(etc, 1000.0, 0.05, 0.013)) < 1e-12


def build_grid(verts, dims=(0, 0, 0)):
    lib = A.lib()
    v = np.ascontiguousarray(verts, dtype=np.float64)
    d = (C.c_int * 3)(*dims)
    org = np.zeros(3)
    cell = np.zeros(3)
    nl = C.c_int64(0)
    dp = C.POINTER(C.c_double)
    u32 = C.POINTER(C.c_uint32)
    assert lib.pach_grid_build(v.ctypes.data_as(dp), v.shape[0], d, org.ctypes.data_as(dp),
                               cell.ctypes.data_as(dp), None, None, C.byref(nl), None) == 0
    G = tuple(d)
    ncell = G[0] * G[1] * G[2]
    ranges = np.zeros((ncell, 2), np.uint32)
    tris = np.zeros(max(nl.value, 1), np.uint32)
    dist = np.zeros(ncell, np.uint8)
    assert lib.pach_grid_build(v.ctypes.data_as(dp), v.shape[0], d, org.ctypes.data_as(dp),
                               cell.ctypes.data_as(dp), ranges.ctypes.data_as(u32),
                               tris.ctypes.data_as(u32), C.byref(nl),
                               dist.ctypes.data_as(C.POINTER(C.c_uint8))) == 0
    return G, org, cell, ranges, tris[:nl.value], dist


@pytest.mark.parametrize("seed,dims,p", [(0, (9, 7, 5), 0.08), (1, (6, 11, 4), 0.2), (2, (5, 5, 5), 0.0)])
def test_octant_cubes_are_the_largest_empty_cubes(seed, dims, p):
    """oct[v][o] is the largest K whose cube (corner v, K voxels along each
    axis in o's directions, voxels outside the grid empty) holds no occupied
    voxel, checked by brute force; 0 on occupied voxels; 255 when the cube
    never meets one."""
    rng = np.random.default_rng(seed)
    gx, gy, gz = dims
    occ = rng.random((gz, gy, gx)) < p
    dist = np.where(occ, 0, 1).astype(np.uint8).ravel()
    oct_ = np.zeros(dist.size * 8, np.uint8)
    d = (C.c_int * 3)(*dims)
    assert A.lib().pach_grid_octant_cubes(d, dist.ctypes.data_as(C.POINTER(C.c_uint8)),
                                          oct_.ctypes.data_as(C.POINTER(C.c_uint8))) == 0
    oct_ = oct_.reshape(gz, gy, gx, 8)

    def empty(x, y, z, K, s):
        xs = [x + s[0] * i for i in range(K)]
        ys = [y + s[1] * i for i in range(K)]
        zs = [z + s[2] * i for i in range(K)]
        return not any(occ[c, b, a] for c in zs if 0 <= c < gz for b in ys if 0 <= b < gy
                       for a in xs if 0 <= a < gx)
    for z in range(gz):
        for y in range(gy):
            for x in range(gx):
                for o in range(8):
                    s = (-1 if o & 1 else 1, -1 if o & 2 else 1, -1 if o & 4 else 1)
                    K = int(oct_[z, y, x, o])
                    if occ[z, y, x]:
                        assert K == 0
                        continue
                    assert K >= 1 and empty(x, y, z, K, s)
                    if K < 255:
                        assert not empty(x, y, z, K + 1, s), (x, y, z, o, K)
                    else:
                        assert empty(x, y, z, max(dims) + 1, s)


@pytest.mark.parametrize("seed,dims,p", [(0, (9, 7, 5), 0.08), (1, (6, 11, 4), 0.2), (2, (5, 5, 5), 0.0),
                                         (3, (12, 4, 6), 0.05)])
def test_octant_boxes_are_the_longest_empty_boxes(seed, dims, p):
    """For each empty voxel and octant: E_a, the longest empty box K long on
    the two axes other than a (K = the octant cube's edge), by brute force;
    the box's first axis A is the a with the largest E_a (lowest on a tie),
    its second B the longer of the other two (the one after A on a tie), and
    its B length the longest empty one with A at E_A and the third axis at K
    -- the box empty and one voxel more on A or B not.  A slip in the axis or
    octant order, a wrong tie rule or a box one voxel too long fails here."""
    rng = np.random.default_rng(seed)
    gx, gy, gz = dims
    occ = rng.random((gz, gy, gx)) < p
    dist = np.where(occ, 0, 1).astype(np.uint8).ravel()
    u8 = C.POINTER(C.c_uint8)
    d = (C.c_int * 3)(*dims)
    oct_ = np.zeros(dist.size * 8, np.uint8)
    ext = np.zeros(dist.size * 24, np.uint8)
    assert A.lib().pach_grid_octant_cubes(d, dist.ctypes.data_as(u8), oct_.ctypes.data_as(u8)) == 0
    assert A.lib().pach_grid_octant_boxes(d, dist.ctypes.data_as(u8), ext.ctypes.data_as(u8)) == 0
    oct_ = oct_.reshape(gz, gy, gx, 8)
    ext = ext.reshape(gz, gy, gx, 3, 8)
    big = max(dims) + 1

    def empty(x, y, z, n, s):
        xs = [x + s[0] * i for i in range(min(n[0], big))]
        ys = [y + s[1] * i for i in range(min(n[1], big))]
        zs = [z + s[2] * i for i in range(min(n[2], big))]
        return not any(occ[c, b, a] for c in zs if 0 <= c < gz for b in ys if 0 <= b < gy
                       for a in xs if 0 <= a < gx)

    def longest(x, y, z, s, n, a, K):
        m = list(n)
        L = K
        while L < 255:
            m[a] = L + 1
            if not empty(x, y, z, m, s):
                return L
            if L + 1 >= big:
                return 255
            L += 1
        return 255
    longer = 0
    for z in range(gz):
        for y in range(gy):
            for x in range(gx):
                for o in range(8):
                    ea, eb, ax = (int(v) for v in ext[z, y, x, :, o])
                    if occ[z, y, x]:
                        assert (ea, eb, ax) == (0, 0, 0)
                        continue
                    s = (-1 if o & 1 else 1, -1 if o & 2 else 1, -1 if o & 4 else 1)
                    K = int(oct_[z, y, x, o])
                    E = [longest(x, y, z, s, [K, K, K], a, K) for a in range(3)]
                    A_ = max(range(3), key=lambda a: (E[a], -a))
                    a1, a2 = (A_ + 1) % 3, (A_ + 2) % 3
                    B_ = a1 if E[a1] >= E[a2] else a2
                    assert ax == A_ | B_ << 2, (x, y, z, o, E, ax)
                    assert ea == E[A_], (x, y, z, o, E, ea)
                    n = [K, K, K]
                    n[A_] = E[A_]
                    assert eb == longest(x, y, z, s, n, B_, K), (x, y, z, o, eb)
                    n[B_] = eb
                    assert empty(x, y, z, n, s)
                    longer += (ea > K) + (eb > K)
    if p > 0:
        assert longer > 0  # boxes do reach past the cubes


@pytest.mark.parametrize("room_fn,dims", [(rooms.shoebox, (0, 0, 0)), (rooms.fan_hall, (32, 32, 32)),
                                          (rooms.fan_hall, (7, 5, 3))])
def test_face_skip_masks_mark_entries_listed_across_each_face(room_fn, dims):
    """Bit i of skip[voxel][f] is set exactly when entry i (< 32) of the
    voxel's list is also in the list of its neighbour across face f (-x, +x,
    -y, +y, -z, +z): the walk, stepping in through f, tested it there."""
    room = room_fn()
    G, org, cell, ranges, tris, dist = build_grid(room.verts, dims)
    v = np.ascontiguousarray(room.verts, dtype=np.float64)
    d = (C.c_int * 3)(*dims)
    ncell = G[0] * G[1] * G[2]
    skip = np.zeros((ncell, 6), np.uint32)
    assert A.lib().pach_grid_face_skip(v.ctypes.data_as(C.POINTER(C.c_double)), v.shape[0], d,
                                       skip.ctypes.data_as(C.POINTER(C.c_uint32))) == 0
    assert tuple(d) == G
    lists = [tris[a:a + n] for a, n in ranges]
    n_set = 0
    for z in range(G[2]):
        for y in range(G[1]):
            for x in range(G[0]):
                c = (z * G[1] + y) * G[0] + x
                for f in range(6):
                    q = [x, y, z]
                    q[f // 2] += 1 if f % 2 else -1
                    want = 0
                    if 0 <= q[f // 2] < G[f // 2]:
                        nb = set(lists[(q[2] * G[1] + q[1]) * G[0] + q[0]].tolist())
                        for i, t in enumerate(lists[c][:32]):
                            if int(t) in nb:
                                want |= 1 << i
                    assert int(skip[c, f]) == want, (x, y, z, f)
                    n_set += bin(want).count("1")
    assert n_set > 0


@pytest.mark.parametrize("room_fn,dims", [(rooms.shoebox, (0, 0, 0)), (rooms.fan_hall, (32, 32, 32)),
                                          (rooms.fan_hall, (0, 0, 0))])
def test_grid_lists_cover_every_triangle_point(room_fn, dims):
    room = room_fn()
    G, org, cell, ranges, tris, dist = build_grid(room.verts, dims)
    if dims[0]:
        assert G == dims
    lists = [set(tris[a:a + n].tolist()) for a, n in ranges]
    rng = np.random.default_rng(0)
    for t, tv in enumerate(room.verts):
        # random barycentric points plus the vertices and edge midpoints
        u = rng.random((64, 2))
        u = np.where(u.sum(1, keepdims=True) > 1, 1 - u, u)
        pts = tv[0] + u[:, :1] * (tv[1] - tv[0]) + u[:, 1:] * (tv[2] - tv[0])
        pts = np.vstack([pts, tv, (tv + np.roll(tv, 1, 0)) / 2])
        idx = np.clip(np.floor((pts - org) / cell).astype(int), 0, np.array(G) - 1)
        for x, y, z in idx:
            assert t in lists[(z * G[1] + y) * G[0] + x]


def test_grid_distance_field_is_linf_distance():
    room = rooms.fan_hall()
    G, org, cell, ranges, tris, dist = build_grid(room.verts, (12, 14, 8))
    occ = np.argwhere((ranges[:, 1] > 0).reshape(G[2], G[1], G[0]))
    cells = np.argwhere(np.ones((G[2], G[1], G[0]), bool))
    # brute-force Chebyshev distance to the nearest occupied voxel
    d = np.abs(cells[:, None, :] - occ[None, :, :]).max(2).min(1)
    np.testing.assert_array_equal(dist, np.minimum(d, 255))


def test_grid_rejects_bad_input():
    v = np.zeros((1, 3, 3))
    v[0, 0, 0] = np.nan
    d = (C.c_int * 3)(0, 0, 0)
    nl = C.c_int64(0)
    assert A.lib().pach_grid_build(v.ctypes.data_as(C.POINTER(C.c_double)), 1, d, None, None, None,
                                   None, C.byref(nl), None) == -1


@pytest.mark.parametrize("r,scale", [(0.25, 40.0), (0.05, 150.0)])
def test_map_entry_screen_never_rejects_a_line_the_fp64_test_accepts(r, scale):
    """The mapping receivers' FP32 screen (trace_impl.h map_check) uses an FP32
    copy Cf of each centre and widens its threshold by 2 (2 r e + e^2), e =
    sqrt(3) 2^-24 max|C| (runtime.cpp pach_map_receivers_set).  On lines that
    pass within r of a centre by 1e-9 .. 1e-4 of r (where the FP64 test's
    discriminant is barely >= 0) the screen must accept every one the FP64
    test accepts; without the widening it would reject some of them."""
    rng = np.random.default_rng(11)
    n = 400_000
    f32 = np.float32
    o = rng.uniform(-scale, scale, (n, 3)).astype(f32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(f32)
    # a unit normal to d, and the centre at distance rho from the line
    a = np.cross(d.astype(np.float64), rng.normal(size=(n, 3)))
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    rho = r * (1 - 10.0 ** rng.uniform(-9, -4, n))
    t = rng.uniform(-3, 3, n)
    C = o.astype(np.float64) + t[:, None] * d.astype(np.float64) + rho[:, None] * a
    # the FP64 test of map_check: b^2 - c >= 0
    m = o.astype(np.float64) - C
    b = np.einsum("ij,ij->i", m, d.astype(np.float64))
    c = np.einsum("ij,ij->i", m, m) - r * r
    hit = b * b - c >= 0
    assert hit.sum() > n // 4
    e = np.sqrt(3.0) * np.ldexp(np.abs(C).max(), -24)
    scr = f32(2 * (2 * r * e + e * e))
    # the FP32 screen, operation by operation in FP32
    fx = o - C.astype(f32)
    fb = (fx[:, 0] * d[:, 0] + fx[:, 1] * d[:, 1]) + fx[:, 2] * d[:, 2]
    fm = (fx[:, 0] * fx[:, 0] + fx[:, 1] * fx[:, 1]) + fx[:, 2] * fx[:, 2]
    lhs = fb * fb - (fm - f32(r * r))
    base = -f32(1e-5) * fm - f32(1e-12)
    assert not np.any(hit & (lhs < base - scr))
    assert np.any(hit & (lhs < base)), "the widening is not exercised"
