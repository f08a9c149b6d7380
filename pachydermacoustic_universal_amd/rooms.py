"""Seeded synthetic rooms for the stochastic tracer.

Each generator returns a :class:`Room`: a triangulated, watertight polygon
scene with one material entry per polygon (the layout ``Polygon_Scene.Construct``
builds, Classes_Scenes.cs:447-630: absorption, scattering and transmission are
stored per polygon at :484-487), plus the source origin and receiver centres
of the workload.  Polygons are quads or triangles; every polygon is split into
triangles that keep the polygon's index as their surface id, so ``Surf_ID``
semantics (material lookup per polygon) are unchanged.

The medium of every workload is 20 C, 50 % RH, 1013.25 hPa (ISO 9613-1,
Air_Choice 0); the source is omni at 120 dB per octave band (1 W).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

NBANDS = 8


@dataclass
class Room:
    name: str
    verts: np.ndarray          # [T, 3, 3] float64 triangle vertices
    surf: np.ndarray           # [T] int32 polygon (surface) index of each triangle
    alpha: np.ndarray          # [S, 8] absorption per polygon and octave band
    scat: np.ndarray           # [S, 8] scattering coefficient
    trans: np.ndarray          # [S, 8] transmission coefficient
    source: np.ndarray         # [3]
    receivers: np.ndarray      # [R, 3]
    radius: float = 0.5
    volume: float = 0.0
    meta: dict = field(default_factory=dict)

    @property
    def n_tris(self) -> int:
        return int(self.verts.shape[0])

    @property
    def n_surfaces(self) -> int:
        return int(self.alpha.shape[0])

    def bbox(self):
        v = self.verts.reshape(-1, 3)
        return v.min(0), v.max(0)


def _tri_volume(verts: np.ndarray) -> float:
    """Signed-volume sum (divergence theorem) of a closed triangle mesh."""
    a, b, c = verts[:, 0], verts[:, 1], verts[:, 2]
    return float(abs(np.einsum("ij,ij->i", a, np.cross(b, c)).sum()) / 6.0)


def _quads_to_tris(quads: np.ndarray):
    """[Q, 4, 3] quads -> [2Q, 3, 3] triangles and the quad index of each."""
    t0 = quads[:, [0, 1, 2]]
    t1 = quads[:, [0, 2, 3]]
    tris = np.stack([t0, t1], axis=1).reshape(-1, 3, 3)
    surf = np.repeat(np.arange(quads.shape[0], dtype=np.int32), 2)
    return tris, surf


def shoebox(lx=10.0, ly=8.0, lz=4.0, alpha=0.10, scat=0.20, trans=0.0,
            source=(2.0, 3.0, 1.5), n_receivers=3, radius=0.5, seed=1,
            receivers=None, name="shoebox") -> Room:
    """Axis-aligned box as 6 quad polygons (12 triangles), normals inward.

    The plain-method small workload: 10 x 8 x 4 m, uniform absorption 0.10,
    scattering 0.20, 3 receivers of radius 0.5 m drawn uniformly from the box
    shrunk by 1 m (seed 1).
    """
    x0, y0, z0, x1, y1, z1 = 0.0, 0.0, 0.0, lx, ly, lz
    P = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]])
    # vertex order gives inward normals (e1 x e2 points into the room)
    faces = np.array([[0, 1, 2, 3],   # floor  z=0, normal +z
                      [4, 7, 6, 5],   # ceiling z=lz, normal -z
                      [0, 4, 5, 1],   # y=0, normal +y
                      [3, 2, 6, 7],   # y=ly, normal -y
                      [0, 3, 7, 4],   # x=0, normal +x
                      [1, 5, 6, 2]])  # x=lx, normal -x
    tris, surf = _quads_to_tris(P[faces])
    S = faces.shape[0]
    A = np.broadcast_to(np.asarray(alpha, dtype=np.float64), (S, NBANDS)).copy()
    Sc = np.broadcast_to(np.asarray(scat, dtype=np.float64), (S, NBANDS)).copy()
    Tr = np.broadcast_to(np.asarray(trans, dtype=np.float64), (S, NBANDS)).copy()
    if receivers is None:
        rng = np.random.default_rng(seed)
        lo = np.array([x0 + 1, y0 + 1, z0 + 1])
        hi = np.array([x1 - 1, y1 - 1, z1 - 1])
        receivers = lo + (hi - lo) * rng.random((n_receivers, 3))
    return Room(name=name, verts=tris, surf=surf, alpha=A, scat=Sc, trans=Tr,
                source=np.asarray(source, dtype=np.float64),
                receivers=np.asarray(receivers, dtype=np.float64).reshape(-1, 3),
                radius=float(radius), volume=lx * ly * lz,
                meta={"dims": (lx, ly, lz)})


def _deformed_box(mapping, nu, nv, nw, jitter, rng):
    """Closed, watertight quad mesh of the unit cube's surface pushed through
    ``mapping(u, v, w) -> xyz``; every shared vertex is jittered once, so
    neighbouring quads keep common vertices.  Returns [Q, 4, 3] quads whose
    normals point into the room."""
    N = (nu, nv, nw)
    verts = {}

    def vid(i, j, k):
        key = (i, j, k)
        if key not in verts:
            p = np.asarray(mapping(i / nu, j / nv, k / nw), dtype=np.float64)
            verts[key] = p + (rng.uniform(-jitter, jitter, 3) if jitter > 0 else 0.0)
        return verts[key]

    quads = []
    # each face: fixed axis a at value 0 or N[a]; (b, c) spans the face.
    for a in range(3):
        b, c = [x for x in range(3) if x != a]
        for side in (0, 1):
            fixed = 0 if side == 0 else N[a]
            for ib in range(N[b]):
                for ic in range(N[c]):
                    corners = []
                    for db, dc in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        idx = [0, 0, 0]
                        idx[a] = fixed
                        idx[b] = ib + db
                        idx[c] = ic + dc
                        corners.append(vid(*idx))
                    # orientation: (b, c) order gives normal along +a for a cyclic
                    # (a, b, c); flip so normals point inward.
                    cyclic = (b - a) % 3 == 1
                    inward_positive = side == 0
                    if cyclic != inward_positive:
                        corners = corners[::-1]
                    quads.append(corners)
    return np.asarray(quads, dtype=np.float64)


def fan_hall(seed=2, nu=14, nv=18, nw=8, height=12.0, front_half=8.0, back_half=18.0,
             depth=30.0, jitter=0.15, n_receivers=16, radius=0.5,
             source=(0.0, 4.0, 1.5), name="fan_hall") -> Room:
    """Seeded polyhedral hall: a fan-shaped plan (16 m stage wall splaying to a
    36 m back wall over 30 m) extruded to 12 m, every wall, floor and ceiling
    subdivided into quads and vertex-jittered (+-0.15 m) to ~2,000 triangles.
    Each quad is one polygon with absorption U(0.02, 0.40) and scattering
    U(0.05, 0.70) per band; 16 receivers 1.2 m over the audience area."""
    rng = np.random.default_rng(seed)

    def mapping(u, v, w):
        half = front_half + (back_half - front_half) * v
        return (-half + 2 * half * u, depth * v, height * w)

    quads = _deformed_box(mapping, nu, nv, nw, jitter, rng)
    tris, surf = _quads_to_tris(quads)
    S = quads.shape[0]
    alpha = rng.uniform(0.02, 0.40, (S, NBANDS))
    scat = rng.uniform(0.05, 0.70, (S, NBANDS))
    trans = np.zeros((S, NBANDS))
    recs = []
    for _ in range(n_receivers):
        u = rng.uniform(0.15, 0.85)
        v = rng.uniform(0.35, 0.95)
        x, y, _z = mapping(u, v, 0.0)
        recs.append((x, y, 1.2))
    return Room(name=name, verts=tris, surf=surf, alpha=alpha, scat=scat, trans=trans,
                source=np.asarray(source, dtype=np.float64),
                receivers=np.asarray(recs, dtype=np.float64), radius=float(radius),
                volume=_tri_volume(tris), meta={"grid": (nu, nv, nw), "jitter": jitter})


def hall_variant(seed, n_receivers=16, radius=0.5) -> Room:
    """The polyhedral-hall generator with seeded proportions: height U(7, 16) m,
    stage half-width U(5, 10) m splaying to a back half-width U(+2, 22) m over a
    depth U(18, 40) m, source on the stage; jitter, materials and receivers as
    fan_hall (the denoiser's training rooms)."""
    rng = np.random.default_rng(1_000_003 + seed)
    height = rng.uniform(7, 16)
    front = rng.uniform(5, 10)
    back = rng.uniform(front + 2, 22)
    depth = rng.uniform(18, 40)
    src = (rng.uniform(-0.5, 0.5) * front, rng.uniform(2.0, 0.2 * depth), 1.5)
    return fan_hall(seed=seed, height=height, front_half=front, back_half=back, depth=depth,
                    n_receivers=n_receivers, radius=radius, source=src, name=f"hall_variant{seed}")


def _box_tris(lo, hi, bottom=True):
    """Axis-aligned box [lo, hi] as triangles with outward normals (the room
    side of an obstacle); without its bottom face when bottom is False."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    P = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]])
    faces = [[0, 4, 5, 1], [3, 2, 6, 7], [0, 3, 7, 4], [1, 5, 6, 2], [4, 7, 6, 5]]
    if bottom:
        faces.append([0, 1, 2, 3])
    # the shoebox order above gives inward normals; reversed: outward
    return _quads_to_tris(P[np.array(faces)][:, ::-1])[0]


def concert_hall(seed=5, n_receivers=64, radius=0.5, name="concert_hall") -> Room:
    """Procedural concert hall of ~50,000 triangles: a 28 x 44 x 18 m shell
    subdivided into 0.5 m quads (walls and ceiling vertex-jittered +-5 cm, the
    floor flat), 180 seat-block boxes (4 x 0.8 x 0.45 m, open to the floor) in
    six columns over the audience area, and four balcony slabs (3 x 30 x
    0.8 m, subdivided) along the side walls at 5 and 9 m.  Each quad or box
    face is a polygon with absorption U(0.02, 0.40) and scattering
    U(0.05, 0.70) per band; the source is on the stage, and 64 receivers sit
    on a jittered 8 x 8 grid 1.25 m over the floor of the audience area."""
    rng = np.random.default_rng(seed)
    W, Lh, H = 28.0, 44.0, 18.0

    def shell_map(u, v, w):
        return (-W / 2 + W * u, Lh * v, H * w)

    quads = _deformed_box(shell_map, 56, 88, 36, 0.05, rng)
    floor = np.abs(quads[..., 2]) < 0.08
    quads[..., 2][floor] = 0.0
    parts = [quads]
    # seat blocks: 6 columns x 30 rows, 1 m pitch, aisles between columns
    boxes = []
    for col in range(6):
        x0 = -12.5 + col * 4.2
        for row in range(30):
            y0 = 12.0 + row * 0.95
            boxes.append(_box_tris((x0, y0, 0.0), (x0 + 4.0, y0 + 0.8, 0.45), bottom=False))
    # balcony slabs, 0.2 m clear of the side walls, as subdivided closed boxes
    slabs = []
    for side in (-1, 1):
        for z in (5.0, 9.0):
            xa = side * (W / 2 - 0.2)
            xb = side * (W / 2 - 3.2)
            lo = (min(xa, xb), 10.0, z)
            hi = (max(xa, xb), 40.0, z + 0.8)

            def slab_map(u, v, w, lo=lo, hi=hi):
                return (lo[0] + (hi[0] - lo[0]) * u, lo[1] + (hi[1] - lo[1]) * v, lo[2] + (hi[2] - lo[2]) * w)

            q = _deformed_box(slab_map, 6, 60, 2, 0.0, rng)
            slabs.append(q[:, ::-1])  # outward normals: the room side of the slab
    shell_tris, shell_surf = _quads_to_tris(parts[0])
    tris = [shell_tris]
    surf = [shell_surf]
    nsurf = int(shell_surf.max()) + 1
    for q in slabs:
        t, sf = _quads_to_tris(q)
        tris.append(t)
        surf.append(sf + nsurf)
        nsurf += q.shape[0]
    for b in boxes:
        tris.append(b)
        surf.append(nsurf + np.arange(b.shape[0], dtype=np.int32) // 2)
        nsurf += b.shape[0] // 2
    tris = np.concatenate(tris)
    surf = np.concatenate(surf).astype(np.int32)
    alpha = rng.uniform(0.02, 0.40, (nsurf, NBANDS))
    scat = rng.uniform(0.05, 0.70, (nsurf, NBANDS))
    gx, gy = np.meshgrid(np.linspace(-10.5, 10.5, 8), np.linspace(14.0, 38.0, 8))
    recs = np.stack([gx.ravel() + rng.uniform(-0.4, 0.4, 64), gy.ravel() + rng.uniform(-0.4, 0.4, 64),
                     np.full(64, 1.25)], axis=1)[:n_receivers]
    return Room(name=name, verts=tris, surf=surf, alpha=alpha, scat=scat, trans=np.zeros((nsurf, NBANDS)),
                source=np.array([0.0, 4.0, 1.5]), receivers=recs, radius=float(radius),
                volume=_tri_volume(tris), meta={"seed": seed})


def concert_variant(seed, n_receivers=64, radius=0.5) -> Room:
    """A concert hall of the concert_hall kind with seeded proportions (the
    denoiser's training rooms of that class; concert_hall itself is never
    one of them): a W x L x H shell with W U(20, 34) m, L U(32, 52) m,
    H U(11, 22) m in ~1 m quads (walls and ceiling jittered +-5 cm, the floor
    flat), 4-7 columns of seat-block boxes (0.8 m deep, 0.40-0.55 m high,
    open to the floor) at a row pitch U(0.9, 1.05) m over the audience area,
    0-2 levels of subdivided balcony slabs along the side walls, per-polygon
    absorption U(0.02, 0.40) and scattering U(0.05, 0.70) per band, the
    source on the stage and receivers on a jittered 8 x 8 grid 1.25 m over the
    audience floor (~8,000-20,000 triangles)."""
    rng = np.random.default_rng(2_000_003 + seed)
    W, Lh, H = rng.uniform(20, 34), rng.uniform(32, 52), rng.uniform(11, 22)

    def shell_map(u, v, w):
        return (-W / 2 + W * u, Lh * v, H * w)

    quads = _deformed_box(shell_map, int(round(W)), int(round(Lh)), int(round(H)), 0.05, rng)
    floor = np.abs(quads[..., 2]) < 0.08
    quads[..., 2][floor] = 0.0
    tris, surf = _quads_to_tris(quads)
    tris, surf = [tris], [surf]
    nsurf = int(surf[0].max()) + 1
    # balconies: 0-2 levels, 0.2 m clear of the side walls
    levels = [z for z in (5.0, 9.0)[:int(rng.integers(0, 3))] if z + 4.0 < H]
    bal_depth = rng.uniform(2.0, 3.5)
    for z in levels:
        for side in (-1, 1):
            xa, xb = side * (W / 2 - 0.2), side * (W / 2 - 0.2 - bal_depth)
            lo = (min(xa, xb), 0.22 * Lh, z)
            hi = (max(xa, xb), 0.9 * Lh, z + 0.8)

            def slab_map(u, v, w, lo=lo, hi=hi):
                return (lo[0] + (hi[0] - lo[0]) * u, lo[1] + (hi[1] - lo[1]) * v, lo[2] + (hi[2] - lo[2]) * w)

            q = _deformed_box(slab_map, 3, int(round(0.68 * Lh)), 1, 0.0, rng)[:, ::-1]
            t, sf = _quads_to_tris(q)
            tris.append(t)
            surf.append(sf + nsurf)
            nsurf += q.shape[0]
    # seat blocks in columns with 0.2 m aisles
    ncol = int(rng.integers(4, 8))
    x_lo, x_hi = -W / 2 + 1.5, W / 2 - 1.5
    bw = (x_hi - x_lo - 0.2 * (ncol - 1)) / ncol
    pitch = rng.uniform(0.9, 1.05)
    bh = rng.uniform(0.40, 0.55)
    y0, y1 = 0.27 * Lh, 0.87 * Lh
    for col in range(ncol):
        x0 = x_lo + col * (bw + 0.2)
        for row in range(int((y1 - y0) // pitch)):
            yb = y0 + row * pitch
            b = _box_tris((x0, yb, 0.0), (x0 + bw, yb + 0.8, bh), bottom=False)
            tris.append(b)
            surf.append(nsurf + np.arange(b.shape[0], dtype=np.int32) // 2)
            nsurf += b.shape[0] // 2
    tris = np.concatenate(tris)
    surf = np.concatenate(surf).astype(np.int32)
    alpha = rng.uniform(0.02, 0.40, (nsurf, NBANDS))
    scat = rng.uniform(0.05, 0.70, (nsurf, NBANDS))
    gx, gy = np.meshgrid(np.linspace(-0.375 * W, 0.375 * W, 8), np.linspace(0.3 * Lh, 0.86 * Lh, 8))
    recs = np.stack([gx.ravel() + rng.uniform(-0.4, 0.4, 64), gy.ravel() + rng.uniform(-0.4, 0.4, 64),
                     np.full(64, 1.25)], axis=1)[:n_receivers]
    src = np.array([rng.uniform(-3.0, 3.0), rng.uniform(2.0, 0.12 * Lh + 2.0), 1.5])
    return Room(name=f"concert_variant{seed}", verts=tris, surf=surf, alpha=alpha, scat=scat,
                trans=np.zeros((nsurf, NBANDS)), source=src, receivers=recs, radius=float(radius),
                volume=_tri_volume(tris), meta={"dims": (W, Lh, H), "columns": ncol, "balconies": len(levels)})


def small_workload() -> Room:
    """Plain method, small: the 12-triangle shoebox."""
    return shoebox()


def hall_workload() -> Room:
    """Seeded polyhedral hall (~2,000 triangles, 16 receivers)."""
    return fan_hall()


def convex_room(seed, n_tris, size=(3.0, 30.0), n_receivers=4, radius=0.5, name="convex"):
    """Random convex polyhedron with ~n_tris triangles (hull of n_tris/2 + 2
    points on a random ellipsoid, sides in `size`), one polygon per triangle,
    normals into the room; source and receivers at seeded interior points."""
    from scipy.spatial import ConvexHull

    rng = np.random.default_rng(seed)
    n_pts = max(n_tris // 2 + 2, 8)
    axes = rng.uniform(size[0], size[1], 3) / 2
    p = rng.normal(size=(n_pts, 3))
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    p *= axes
    hull = ConvexHull(p)
    tris = p[hull.simplices]
    c = p.mean(0)
    # orient inward: normal must point toward the centroid
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    flip = np.einsum("ij,ij->i", n, c - tris[:, 0]) < 0
    tris[flip] = tris[flip][:, ::-1]
    T = tris.shape[0]
    alpha = rng.uniform(0.02, 0.40, (T, NBANDS))
    scat = rng.uniform(0.05, 0.70, (T, NBANDS))

    def interior():
        # a vertex pulled 30-60 % of the way toward the centroid stays inside
        v = p[rng.integers(n_pts)]
        return c + (v - c) * rng.uniform(0.3, 0.6)

    src = interior()
    recs = np.array([interior() for _ in range(n_receivers)])
    return Room(name=name, verts=tris, surf=np.arange(T, dtype=np.int32), alpha=alpha, scat=scat,
                trans=np.zeros((T, NBANDS)), source=src, receivers=recs, radius=float(radius),
                volume=float(hull.volume), meta={"axes": axes})


def batch_rooms(seed=3, n=256, n_receivers=4, radius=0.5):
    """The batched workload: n seeded rooms, each a shoebox with sides U(3, 30)
    m or a random convex polyhedron with 12-500 triangles, random per-band
    absorption U(0.02, 0.40) and scattering U(0.05, 0.70) per polygon, one source
    and 4 receivers; ray counts log-uniform in [1e4, 1e6]."""
    rng = np.random.default_rng(seed)
    out, rays = [], []
    for i in range(n):
        if rng.random() < 0.5:
            L = rng.uniform(3, 30, 3)
            a = rng.uniform(0.02, 0.40, (6, NBANDS))
            s = rng.uniform(0.05, 0.70, (6, NBANDS))
            lo = np.minimum(0.6, L * 0.2)
            src = lo + (L - 2 * lo) * rng.random(3)
            recs = lo + (L - 2 * lo) * rng.random((n_receivers, 3))
            r = shoebox(*L, alpha=a, scat=s, source=src, receivers=recs, radius=min(radius, float(lo.min()) * 0.5),
                        name=f"batch{i}_box")
            r.radius = radius
        else:
            r = convex_room(int(rng.integers(1 << 31)), int(rng.integers(12, 501)), n_receivers=n_receivers,
                            radius=radius, name=f"batch{i}_convex")
        out.append(r)
        rays.append(int(round(10 ** rng.uniform(4, 6))))
    return out, rays


@dataclass
class PackedRooms:
    room: Room                 # all rooms side by side in one scene
    offsets: np.ndarray        # [n, 3] translation of each room
    tri_count: np.ndarray      # [n]
    surf_count: np.ndarray     # [n]
    rec_count: np.ndarray      # [n]
    sources: np.ndarray        # [n, 3] (translated)

    def member(self, i):
        """Room i in packed coordinates (the geometry the batch traces)."""
        t0, t1 = int(self.tri_count[:i].sum()), int(self.tri_count[:i + 1].sum())
        s0, s1 = int(self.surf_count[:i].sum()), int(self.surf_count[:i + 1].sum())
        r0, r1 = int(self.rec_count[:i].sum()), int(self.rec_count[:i + 1].sum())
        R = self.room
        return Room(name=f"member{i}", verts=R.verts[t0:t1], surf=R.surf[t0:t1] - s0, alpha=R.alpha[s0:s1],
                    scat=R.scat[s0:s1], trans=R.trans[s0:s1], source=self.sources[i],
                    receivers=R.receivers[r0:r1], radius=R.radius)


def pack_rooms(rooms_list, gap=4.0):
    """Places rooms on a square lattice with `gap` metres between bounding
    boxes, so no ray can pass from one room into another; polygons, receivers
    and sources are concatenated room by room."""
    n = len(rooms_list)
    side = int(np.ceil(np.sqrt(n)))
    ext = np.array([r.bbox()[1] - r.bbox()[0] for r in rooms_list])
    cell = ext.max(0) + gap
    offs, verts, surf, al, sc, tr, recs, srcs = [], [], [], [], [], [], [], []
    sbase = 0
    for i, r in enumerate(rooms_list):
        lo = r.bbox()[0]
        o = np.array([(i % side) * cell[0], (i // side) * cell[1], 0.0]) - lo
        offs.append(o)
        verts.append(r.verts + o)
        surf.append(r.surf + sbase)
        sbase += r.n_surfaces
        al.append(r.alpha)
        sc.append(r.scat)
        tr.append(r.trans)
        recs.append(r.receivers + o)
        srcs.append(r.source + o)
    radius = rooms_list[0].radius
    packed = Room(name=f"packed{n}", verts=np.concatenate(verts), surf=np.concatenate(surf).astype(np.int32),
                  alpha=np.concatenate(al), scat=np.concatenate(sc), trans=np.concatenate(tr),
                  source=srcs[0], receivers=np.concatenate(recs), radius=radius)
    return PackedRooms(room=packed, offsets=np.array(offs),
                       tri_count=np.array([r.n_tris for r in rooms_list]),
                       surf_count=np.array([r.n_surfaces for r in rooms_list]),
                       rec_count=np.array([len(r.receivers) for r in rooms_list]),
                       sources=np.array(srcs))
